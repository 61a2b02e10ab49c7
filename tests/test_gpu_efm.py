"""GPU parity of the EFM front-end (demodblock's digital-audio output,
core.py:847-851): the int16 EFM output of k_demod_block_efm against the
oracle's FP64 np.int16(np.clip(ifft(X * Fefm).real)) on the same blocks,
within 1 code (FP32 transforms before the truncation) on a small fraction
of samples, with the clip reached; and the video outputs of the EFM kernel
equal to the plain kernel's.
"""

import numpy as np
import pytest

from lddecode_amd import synth
from oracle import efm as E

pytestmark = pytest.mark.gpu

_RF = {}


def _rf_efm(system):
    from lddecode_amd.rfdecode import RFDecode
    if system not in _RF:
        _RF[system] = RFDecode(system=system, decode_analog_audio=0, decode_digital_audio=True, device=0)
    return _RF[system]


def _capture(system, nblocks=6, seed=5):
    cap = synth.make_capture(system, 31712 * (nblocks - 1) + 32768, seed=seed, snr_db=40.0, audio=False)
    x = cap.data.astype(np.int32)
    # an EFM-band signal: a strong 800 kHz tone (reaching the int16 clip)
    # plus band noise
    rng = np.random.default_rng(seed)
    t = np.arange(len(x))
    x = x // 4 + (9000 * np.cos(2 * np.pi * 0.8e6 * t / 40e6)).astype(np.int32) + (rng.normal(size=len(x)) * 800).astype(np.int32)
    return np.clip(x, -32768, 32767).astype(np.int16)


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_efm_output_matches_oracle(gpu_rf, system):
    import torch
    rf = _rf_efm(system)
    data = _capture(system)
    nb = 6
    out = rf.demod_blocks(torch.from_numpy(data).cuda(), 0, nb)
    got = out["efm"].cpu().numpy()
    assert got.dtype == np.int16
    filt = E.efm_filter(rf.freq_hz, rf.blocklen)
    S = rf.block_stride
    diffs = []
    for b in range(nb):
        X = np.fft.fft(data[b * S:b * S + rf.blocklen].astype(np.float64))
        want = E.efm_block(X, filt)
        d = got[b * S:(b + 1) * S].astype(np.int32) - want.astype(np.int32)
        diffs.append(d)
    d = np.concatenate(diffs)
    assert np.abs(d).max() <= 1
    assert (d == 0).mean() >= 0.97
    assert (np.abs(got.astype(np.int32)) >= 32767).any()
    # the plain kernel's video outputs are unchanged by the EFM phase
    plain = gpu_rf(system, audio=False)
    ref = plain.demod_blocks(torch.from_numpy(data).cuda(), 0, nb)
    for k in ("demod", "demod_05", "rfhpf"):
        a = (ref["video"][k] if k != "rfhpf" else ref["rfhpf"]).cpu().numpy()
        b = (out["video"][k] if k != "rfhpf" else out["rfhpf"]).cpu().numpy()
        np.testing.assert_allclose(b, a, rtol=0, atol=4 * np.spacing(np.abs(a).max()))


def test_demodblock_returns_efm(gpu_rf):
    """RFDecode.demodblock (one uncut block) returns rv['efm'] as the
    reference's demodblock does."""
    import torch
    rf = _rf_efm("NTSC")
    data = _capture("NTSC", nblocks=1)
    rv = rf.demodblock(data=torch.from_numpy(data[:rf.blocklen]).cuda(), cut=False)
    assert "efm" in rv and rv["efm"].numel() == rf.blocklen
    want = E.efm_block(np.fft.fft(data[:rf.blocklen].astype(np.float64)), E.efm_filter(rf.freq_hz, rf.blocklen),
                       cut=False)
    d = rv["efm"].cpu().numpy().astype(np.int32) - want
    assert np.abs(d).max() <= 1 and (d == 0).mean() >= 0.97
