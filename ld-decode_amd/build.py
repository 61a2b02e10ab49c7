"""Build recipe for libldgpu.so (hipcc, gfx950).

    python -m lddecode_amd.build        (or import and call build())

Compiles every HIP translation unit under csrc/ into one shared library next
to this file.  All device code is built with packed-FP32 instructions
disabled.
"""

import os
import pathlib
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libldgpu.so"
SOURCES = ["demod.hip", "tbc.hip", "lines.hip", "pulses.hip", "input.hip", "dropouts.hip", "audio.hip", "locate.cpp"]
ARCH = os.environ.get("LDGPU_ARCH", "gfx950")


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return str(pathlib.Path(rocm) / "bin" / "hipcc")


def build(verbose=False):
    """Compile csrc/*.hip into libldgpu.so (always recompiles)."""
    objs = []
    flags = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]
    for src in SOURCES:
        obj = CSRC / (pathlib.Path(src).stem + ".o")
        cmd = [hipcc(), *flags, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
