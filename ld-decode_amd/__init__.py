"""MI355X-native RF demodulation + time-base correction for LaserDisc RF.

Drop-in GPU path for ld-decode's ``RFDecode.demodblock`` (block FM
demodulation with overlap-save stitching) and ``Field.downscale`` (the
sinc-LUT field resampler).  All sample arithmetic runs in hand-written HIP
kernels for gfx950 in ``libldgpu.so``, called through a C ABI (ctypes);
PyTorch-ROCm only owns device buffers and streams.

The directory name contains a hyphen, so it is imported under the module name
``lddecode_amd`` (see ``load_package``).
"""

__all__ = ["load_package"]


def load_package():
    import sys
    return sys.modules[__name__]
