"""Write tests/golden/sinc_lut_rows.npz: sampled rows of the reference's own
data file lddecode/sinc_lut.npz (65537 x 16 float32), read with numpy's safe
loader (no pickle).  Only data is kept: the row indices and their values.

Run in a container that has the reference checked out at /root/reference:
    python tests/golden/make_sinc_lut_rows.py
"""

import pathlib

import numpy as np

REF = pathlib.Path("/root/reference/lddecode/sinc_lut.npz")
OUT = pathlib.Path(__file__).with_name("sinc_lut_rows.npz")


def main():
    with np.load(REF, allow_pickle=False) as z:
        lut = z["downscale_sinc_lut"]
    rng = np.random.default_rng(20240605)
    rows = np.unique(np.concatenate([
        [0, 1, 2, 4095, 4096, 16384, 32767, 32768, 32769, 49152, 65534, 65535, 65536],
        rng.integers(0, 65537, 243),
    ]))
    np.savez_compressed(OUT, rows=rows.astype(np.int64), values=lut[rows],
                        shape=np.array(lut.shape, dtype=np.int64))
    print(f"wrote {OUT} ({len(rows)} rows)")


if __name__ == "__main__":
    main()
