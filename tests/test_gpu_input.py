"""GPU parity of the packed 10-bit input: the unpack kernel against the
oracle's unpack_data_4_40 (bit-exact), and the demodulator reading packed
bytes against the same capture as int16.  The two read identical sample
values; the packed kernel is a separate instantiation whose floating-point
contraction the compiler may schedule differently, so its outputs agree
with the int16 kernel's within 4 units in the last place of each channel's
magnitude (about 0.5 ppm), at every sample alignment of the first block."""

import numpy as np
import pytest

from lddecode_amd import synth
from oracle import packed as PK

pytestmark = pytest.mark.gpu


def test_unpack_kernel_matches_oracle(gpu_rf):
    import torch
    rf = gpu_rf("NTSC", audio=False)
    rng = np.random.default_rng(7)
    codes = rng.integers(0, 1024, size=40_003)
    buf = PK.pack_4_40(codes)
    dev = torch.from_numpy(buf).cuda()
    want = ((codes.astype(np.int32) - 512) << 6).astype(np.int16)
    for first, n in ((0, 40_000), (1, 1), (2, 7), (3, 39_990), (1001, 12_345)):
        got = rf.unpack_10bit(dev, first, n).cpu().numpy()
        np.testing.assert_array_equal(got, want[first:first + n])
        if n % 4 == 0:
            np.testing.assert_array_equal(got, PK.load_packed(buf, first, n))


@pytest.mark.parametrize("first", [0, 1, 2, 3])
def test_demod_packed_equals_int16(gpu_rf, first):
    import torch
    rf = gpu_rf("NTSC", audio=True)
    cap = synth.make_capture("NTSC", 6 * 31712 + 32768 + 8, seed=5, snr_db=40.0)
    packed = torch.from_numpy(synth.pack_10bit(cap.data)).cuda()
    ints = torch.from_numpy(cap.data).cuda()
    a = rf.demod_blocks(ints, first=first, nblocks=5)
    b = rf.demod_blocks(packed, first=first, nblocks=5)
    def close(x, y, name):
        x, y = x.cpu().numpy(), y.cpu().numpy()
        ulp = np.spacing(np.float32(np.abs(x).max()))
        d = np.abs(x.astype(np.float64) - y)
        assert d.max() <= 4 * ulp, (name, d.max(), ulp)

    for k in rf.video_fields:
        close(a["video"][k], b["video"][k], k)
    close(a["rfhpf"], b["rfhpf"], "rfhpf")
    for k in ("audio_left", "audio_right"):
        close(a["audio"][k], b["audio"][k], k)
