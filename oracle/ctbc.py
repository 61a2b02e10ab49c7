"""ctypes binding of the oracle's C resampler loop (oracle/tbc_loop.c).

Test infrastructure: used by the CPU tests (cross-check against
oracle/tbc.py:scale_field) and by the benchmark's CPU baseline."""

import ctypes
import pathlib
import subprocess

import numpy as np

from . import tbc as T

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "liboracle.so"
_lib = None


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", str(_HERE), "liboracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built (run oracle.ctbc.build())")
        L = ctypes.CDLL(str(LIB_PATH))
        p = ctypes.c_void_p
        L.ora_scale_field.argtypes = [p, ctypes.c_int64, p, p, p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, p]
        L.ora_scale_field.restype = ctypes.c_int
        _lib = L
    return _lib


def scale_field(buf, pixel_locs, wowfactors, lut, lineoffset, outwidth, nout,
                smoothing=0.0, threshold=15, coord_f32=True):
    """Same contract and result as oracle.tbc.scale_field (utils.py:127-183)."""
    buf = np.ascontiguousarray(buf, dtype=np.float32)
    locs = np.ascontiguousarray(pixel_locs, dtype=np.float64)
    la = np.ascontiguousarray(T.level_adjusts(wowfactors, outwidth, smoothing, threshold), dtype=np.float64)
    lut = np.ascontiguousarray(lut, dtype=np.float32)
    start = outwidth * (lineoffset + 1)
    if start + nout > len(locs):
        raise ValueError("pixel_locs shorter than the requested output")
    out = np.empty(nout, dtype=np.float64)
    lib().ora_scale_field(buf.ctypes.data, len(buf), locs.ctypes.data, la.ctypes.data, lut.ctypes.data,
                          start, nout, int(bool(coord_f32)), out.ctypes.data)
    return out
