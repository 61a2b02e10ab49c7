"""8192-, 16384- and 65536-sample blocks (RFDecode(blocklen=...), the
reference's blocklen constructor argument, core.py:284-288): the single-level
4096- and 8192-point kernels and the two-level transform kernel (65536) against the FP64
oracle built at the same block length (its filters are freqz samples at
that many bins), in FP32 and with the split-BF16 MFMA radix-64 stages,
single and stitched blocks, and the batched decode at that length against
the per-capture decode.

Tolerances as the 32768-sample parity tests: video max |err| <= 25 Hz and
1e-4 relative (to max(|value|, 1 MHz)); rfhpf <= 0.05 codes; audio stage 1
<= 10 Hz.  MFMA variants: 1e-4 relative on every channel.
"""

import numpy as np
import pytest

from lddecode_amd import _lib, synth
from oracle import dsp as O

pytestmark = pytest.mark.gpu

LENGTHS = [8192, 16384, 65536]
_RF, _ORF = {}, {}


def _rf(system, n):
    if (system, n) not in _RF:
        from lddecode_amd.rfdecode import RFDecode
        _RF[system, n] = RFDecode(system=system, blocklen=n, decode_analog_audio=44100, device=0)
    return _RF[system, n]


def _orf(system, n):
    if (system, n) not in _ORF:
        _ORF[system, n] = O.OracleRF(system, blocklen=n, decode_analog_audio=True)
    return _ORF[system, n]


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


def _check(got, ref, fields, tol=25.0, rel=1e-4):
    for k in fields:
        g, r = _np(got[k]), ref[k]
        assert g.shape == r.shape, k
        err = np.abs(g - r)
        assert err.max() <= tol, (k, err.max())
        assert (err / np.maximum(np.abs(r), 1e6)).max() <= rel, k


@pytest.mark.parametrize("n", LENGTHS)
@pytest.mark.parametrize("system,cut", [("NTSC", True), ("NTSC", False), ("PAL", True)])
def test_block_matches_oracle(torch_cuda, n, system, cut):
    rf, orf = _rf(system, n), _orf(system, n)
    assert rf.block_stride == n - 1056
    cap = synth.make_capture(system, n, seed=7, snr_db=40.0)
    got = rf.demodblock(data=cap.data, mtf_level=0, cut=cut)
    ref = orf.demodblock(data=cap.data.astype(np.float64), mtf_level=0, cut=cut)
    _check(got["video"], ref["video"], rf.video_fields)
    d = np.abs(_np(got["rfhpf"]) - ref["rfhpf"])
    assert d.max() <= 0.05, d.max()
    for k in ("audio_left", "audio_right"):
        g, r = _np(got["audio"][k]), ref["audio"][k]
        assert g.shape == r.shape
        assert np.abs(g - r).max() <= 10.0, (k, np.abs(g - r).max())


@pytest.mark.parametrize("variant", [_lib.LD_FFT_MFMA_BF16X3, _lib.LD_FFT_MFMA_BF16X6])
@pytest.mark.parametrize("n", LENGTHS)
@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_block_mfma_matches_oracle(torch_cuda, n, system, variant):
    rf, orf = _rf(system, n), _orf(system, n)
    cap = synth.make_capture(system, n, seed=8, snr_db=40.0)
    ref = orf.demodblock(data=cap.data.astype(np.float64), mtf_level=0, cut=True)
    try:
        rf.set_fft_variant(variant)
        got = rf.demodblock(data=cap.data, mtf_level=0, cut=True)
    finally:
        rf.set_fft_variant(_lib.LD_FFT_FP32)
    for k in rf.video_fields:
        e = np.abs(_np(got["video"][k]) - ref["video"][k])
        print(f"{system} {n} MFMA variant {variant} {k}: max {e.max():.2f} Hz, rms {np.sqrt(np.mean(e ** 2)):.3f}", flush=True)
    _check(got["video"], ref["video"], rf.video_fields, tol=1e9, rel=1e-4)


@pytest.mark.parametrize("n", LENGTHS)
@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_stitched_blocks_match_oracle(torch_cuda, n, system):
    rf, orf = _rf(system, n), _orf(system, n)
    S = rf.block_stride
    cap = synth.make_capture(system, 2 * S + n, seed=21, snr_db=40.0)
    got = rf.demod_blocks(cap.data, 0, 3)
    ref, start = O.demod_span(orf, cap.data.astype(np.float64), 0, 3 * S - 1)
    assert start == 0
    _check(got["video"], {k: ref[k] for k in rf.video_fields}, rf.video_fields)
    assert np.abs(_np(got["rfhpf"]) - ref["rfhpf"]).max() <= 0.05
    # every block's outputs are those of the block alone
    one = rf.demod_blocks(cap.data, S, 1)
    for k in rf.video_fields:
        np.testing.assert_array_equal(_np(got["video"][k])[S:2 * S], _np(one["video"][k]))


@pytest.mark.parametrize("n", LENGTHS)
def test_batch_equals_per_capture_decode(torch_cuda, n):
    import torch
    from lddecode_amd import batch as batchmod
    from lddecode_amd import capture
    plan = [("NTSC", 3_200_000), ("PAL", 3_600_000)]
    b = synth.make_batch(plan, "cuda", seed=12, ntemplates=1)
    rfs = {"NTSC": _rf("NTSC", n), "PAL": _rf("PAL", n)}
    got = {}

    def on_wave(info):
        torch.cuda.synchronize()
        for s in ("NTSC", "PAL"):
            if s not in info:
                continue
            d = info[s]
            W, lines = rfs[s].SysParams["outlinelen"], rfs[s].SysParams["frame_lines"] // 2 + 1
            for f in d["fields"]:
                b0, nb, slot0 = d["lanes"][f["lane"]]
                start = f["demod_offset"] + (b0 - slot0) * rfs[s].block_stride
                row = f["out_offset"] // (W * lines)
                got[(f["lane"], start)] = d["out_u16"][row * W * lines:(row + 1) * W * lines].cpu().numpy()

    dec = batchmod.BatchDecoder(rfs, b.data, b.systems, b.offsets, b.lengths, wave_blocks=20 * 65536 // n, min_fields=1,
                                graph=True, on_wave=on_wave)
    n = dec.run()
    torch.cuda.synchronize()
    assert dec.waves >= 2 and n >= 4
    for k in range(len(plan)):
        cd = capture.CaptureDecoder(rfs[b.systems[k]], b.capture(k), locate=True, max_fields=16)
        out = cd.step().clone()
        torch.cuda.synchronize()
        W, lines = cd.W, cd.linesout
        want = {f["demod_offset"]: out[j * W * lines:(j + 1) * W * lines].cpu().numpy()
                for j, f in enumerate(cd.fields)}
        mine = {s: v for (lane, s), v in got.items() if lane == k}
        assert sorted(mine) == sorted(want)
        for s, w in want.items():
            np.testing.assert_array_equal(mine[s], w)
