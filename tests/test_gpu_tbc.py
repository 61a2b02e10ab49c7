"""GPU parity of the hsync refinement and the field TBC against the oracle.

Tolerances: line locations |err| <= 1e-3 samples with identical bad-line
flags; TBC uint16 codes bit-exact when both sides resample the same demod
buffer; end to end (GPU FP32 demod vs FP64 oracle demod) every code within
+-1 LSB and >= 99.9% of codes exact or +-1, in both the
float32-coordinate (reference) and FP64-coordinate modes.
"""

import numpy as np
import pytest

from lddecode_amd import _lib, capture, synth
from lddecode_amd import field as fieldmod
from oracle import dsp as O
from oracle import tbc as T
from oracle.params import out_scale

pytestmark = pytest.mark.gpu

LL_TOL = 1e-3


def _np(t):
    return t.detach().cpu().numpy()


def _oracle_field(orf, demod, ll, lineoffset, coord_f32, smoothing=0.0, lut=None):
    sp = orf.SysParams
    W, lines = sp["outlinelen"], sp["frame_lines"] // 2 + 1
    locs, wow = T.computewow_scaled(ll, orf.linelen, W, lines, lineoffset)
    lut = T.build_kaiser_lut() if lut is None else lut
    ds = T.scale_field(np.asarray(demod, np.float32), locs, wow, lut, lineoffset, W, lines * W,
                       smoothing=smoothing, coord_f32=coord_f32)
    dp = orf.DecoderParams
    u16 = T.hz_to_output_array(ds.astype(np.float32), dp["ire0"], dp["hz_ire"], sp["outputZero"], dp["vsync_ire"],
                               out_scale(orf.system, dp["vsync_ire"]))
    return ds, u16


@pytest.mark.parametrize("system,seed", [("NTSC", 2), ("PAL", 3)])
def test_field_decode_matches_oracle(gpu_rf, oracle_rf, system, seed):
    """Configs C2/C3: one field decoded from stitched blocks; hsync refine
    and TBC against the oracle run on the oracle's own FP64 demod."""
    import torch
    rf, orf = gpu_rf(system, audio=False), oracle_rf(system, audio=False)
    cap = synth.make_capture(system, 1_700_000 if system == "NTSC" else 1_900_000, seed=seed,
                             snr_db=40.0 if system == "NTSC" else 36.0)
    fields, ll1 = capture.truth_fields(rf, cap, 1, first_field=1)
    fd = fields[0]
    S = rf.block_stride
    nb = 29 if system == "NTSC" else 33
    org = fd["demod_offset"]
    x = torch.from_numpy(cap.data).cuda()
    got = rf.demod_blocks(x, org, nb, audio=False)
    ref, _ = O.demod_span(orf, cap.data.astype(np.float64), org, nb * S - 1)
    # hsync refinement: GPU on GPU demod_05 vs oracle on oracle demod_05
    ll2g, badg = fieldmod.refine_linelocs_hsync(rf, got["video"]["demod_05"], ll1)
    ll2o, bado = T.refine_linelocs_hsync(ref["demod_05"], ll1, np.zeros(len(ll1), bool), system, 40.0,
                                         orf.iretohz, orf.DecoderParams["vsync_ire"])
    badg = _np(badg)
    np.testing.assert_array_equal(badg, bado)
    assert np.abs(_np(ll2g) - ll2o).max() <= LL_TOL
    assert (~bado).sum() > 0.9 * len(bado)
    lineoffset = 0 if system == "NTSC" else 2
    for coord_f32 in (True, False):
        f = fieldmod.Field(rf, got, linelocs=ll2o, lineoffset=lineoffset, coord_f32=coord_f32)
        u16 = _np(f.downscale(final=True)[0]).astype(np.int64)
        _, u16o = _oracle_field(orf, ref["demod"], ll2o, lineoffset, coord_f32)
        d = u16 - u16o.astype(np.int64)
        assert np.abs(d).max() <= 1, np.abs(d).max()
        assert np.mean(np.abs(d) <= 1) >= 0.999
        # same demod buffer on both sides: bit-exact codes
        same = {"video": {"demod": torch.from_numpy(ref["demod"].astype(np.float32)).cuda()}}
        f2 = fieldmod.Field(rf, same, linelocs=ll2o, lineoffset=lineoffset, coord_f32=coord_f32)
        np.testing.assert_array_equal(_np(f2.downscale(final=True)[0]).astype(np.uint16), u16o)


def test_tbc_smoothing_and_float_output(gpu_rf, oracle_rf):
    """wow_level_adjust_smoothing > 0 (utils.py:142-149): closed-form EMA on
    the GPU vs the oracle's sequential recurrence; float32 dsout within 1 Hz."""
    import torch
    rf, orf = gpu_rf("NTSC", audio=False), oracle_rf("NTSC", audio=False)
    rng = np.random.default_rng(12)
    demod = (8.1e6 + rng.normal(0, 3e5, 800_000)).astype(np.float32)
    ll = 5000 + np.arange(273) * 2542.3 + rng.uniform(-2, 2, 273)
    ll[40] += 30     # an outlier line (clamped to the median wow)
    for smoothing in (0.0, 0.5):
        ds_o, u16_o = _oracle_field(orf, demod, ll, 0, True, smoothing=smoothing)
        f = fieldmod.Field(rf, {"video": {"demod": torch.from_numpy(demod).cuda()}}, linelocs=ll,
                           wow_level_adjust_smoothing=smoothing)
        ds = _np(f.downscale(final=False)[0]).astype(np.float64)
        assert np.abs(ds - ds_o).max() <= 1.0
        u16 = _np(f.downscale(final=True)[0]).astype(np.int64)
        assert np.abs(u16 - u16_o.astype(np.int64)).max() <= 1


def test_tbc_asymmetric_lut(gpu_rf, oracle_rf):
    """A sinc table without the mirror symmetry (rows above 32768 perturbed)
    is read in full: float32 output equals the oracle's with the same table,
    bit for bit, in both coordinate modes.  The symmetric table is restored
    afterwards."""
    import torch
    rf, orf = gpu_rf("NTSC", audio=False), oracle_rf("NTSC", audio=False)
    rng = np.random.default_rng(21)
    demod = (8.1e6 + rng.normal(0, 3e5, 800_000)).astype(np.float32)
    ll = 5000 + np.arange(273) * 2542.3 + rng.uniform(-2, 2, 273)
    lut = T.build_kaiser_lut()
    lut[32769::7] *= np.float32(1.0 + 2.0 ** -8)
    try:
        fieldmod.upload_sinc_lut(rf, lut)
        for coord_f32 in (True, False):
            ds_o, _ = _oracle_field(orf, demod, ll, 0, coord_f32, lut=lut)
            f = fieldmod.Field(rf, {"video": {"demod": torch.from_numpy(demod).cuda()}}, linelocs=ll,
                               coord_f32=coord_f32)
            ds = _np(f.downscale(final=False)[0])
            np.testing.assert_array_equal(ds, ds_o.astype(np.float32))
    finally:
        fieldmod.upload_sinc_lut(rf)


def test_tbc_distorted_lines(gpu_rf, oracle_rf):
    """Line locations that stretch some lines past a chunk's LDS stage (the
    taps of those chunks come from global memory) and one line that runs
    backwards: float32 output equals the oracle's bit for bit in both
    coordinate modes."""
    import torch
    rf, orf = gpu_rf("NTSC", audio=False), oracle_rf("NTSC", audio=False)
    rng = np.random.default_rng(33)
    demod = (8.1e6 + rng.normal(0, 3e5, 800_000)).astype(np.float32)
    ll = 5000 + np.arange(273) * 2542.3 + rng.uniform(-2, 2, 273)
    ll[100:] += 3000          # line 99: ~6.1 samples per pixel (1560 per 256-pixel chunk)
    ll[180:] += 15000         # line 179: ~7 times the usual length
    ll[150] = ll[149] - 50.0  # line 149 runs backwards, line 150 twice as long
    for coord_f32 in (True, False):
        ds_o, _ = _oracle_field(orf, demod, ll, 0, coord_f32)
        f = fieldmod.Field(rf, {"video": {"demod": torch.from_numpy(demod).cuda()}}, linelocs=ll,
                           coord_f32=coord_f32)
        ds = _np(f.downscale(final=False)[0])
        np.testing.assert_array_equal(ds, ds_o.astype(np.float32))


def test_batched_fields_equal_single(gpu_rf):
    """Three fields in one tbc_fields launch equal three single-field calls."""
    import torch
    rf = gpu_rf("PAL", audio=False)
    rng = np.random.default_rng(13)
    demod = torch.from_numpy((7.1e6 + rng.normal(0, 2e5, 3_000_000)).astype(np.float32)).cuda()
    fields, lls, singles = [], [], []
    W, lines = 1135, 313
    for f in range(3):
        ll = 3000 + np.arange(326) * 2560.1 + rng.uniform(-2, 2, 326)
        off = f * 900_000
        fields.append(dict(demod_offset=off, demod_len=900_000, lineloc_offset=326 * f, nlines=326,
                           lineoffset=2 + (f % 2), linesout=lines, outwidth=W, inlinelen=2560.0,
                           out_offset=f * lines * W))
        lls.append(ll)
        one = dict(fields[-1], demod_offset=0, lineloc_offset=0, out_offset=0)
        _, u = fieldmod.tbc_fields(rf, demod[off:off + 900_000].contiguous(), [one],
                                   torch.from_numpy(ll).cuda())
        singles.append(_np(u))
    _, u16 = fieldmod.tbc_fields(rf, demod, fields, torch.from_numpy(np.concatenate(lls)).cuda())
    np.testing.assert_array_equal(_np(u16), np.concatenate(singles))


def test_tbc_fields_rejects_bad_geometry(gpu_rf):
    """ld_tbc_fields checks every field before it launches anything: a line
    offset outside 0..1024, output widths outside 1..2048, too few line
    locations and a demod span past the buffer are errors with their codes."""
    import torch
    rf = gpu_rf("PAL", audio=False)
    demod = torch.zeros(900_000, dtype=torch.float32, device="cuda")
    ll = torch.from_numpy(3000 + np.arange(326) * 2560.0).cuda()
    good = dict(demod_offset=0, demod_len=900_000, lineloc_offset=0, nlines=326, lineoffset=2, linesout=313,
                outwidth=1135, inlinelen=2560.0, out_offset=0)
    fieldmod.tbc_fields(rf, demod, [good], ll)
    for bad, code in ((dict(lineoffset=1025), _lib.LD_EINVAL), (dict(lineoffset=-1), _lib.LD_EINVAL),
                      (dict(outwidth=2049), _lib.LD_EINVAL), (dict(outwidth=0), _lib.LD_EINVAL),
                      (dict(nlines=1), _lib.LD_EINVAL), (dict(demod_len=900_001), _lib.LD_ERANGE)):
        with pytest.raises(_lib.LdGpuError) as e:
            fieldmod.tbc_fields(rf, demod, [dict(good, **bad)], ll)
        assert e.value.code == code, bad
    torch.cuda.synchronize()


def test_capture_decoder_end_to_end(gpu_rf):
    """A 6-field NTSC capture through CaptureDecoder (demod + hsync + TBC in
    three launches groups): every field's sync tip and blanking land at the
    output codes the level model predicts (size-independent property)."""
    import torch
    rf = gpu_rf("NTSC", audio=False)
    cap = synth.make_capture("NTSC", 8 * 667_334, seed=31, snr_db=46.0)
    fields, ll1 = capture.truth_fields(rf, cap, 6, first_field=1)
    dec = capture.CaptureDecoder(rf, torch.from_numpy(cap.data).cuda(), fields, ll1)
    out = _np(dec.step()).reshape(6, 263, 910).astype(np.int64)
    torch.cuda.synchronize()
    bad = _np(dec.linebad)
    assert bad.mean() < 0.1
    osc = out_scale("NTSC", -40)                 # output codes per IRE
    blank = 1024 + 40 * osc                      # 0 IRE
    # 4fsc pixels from the line's sync edge: sync tip 0.7-3.5 us, back porch
    # (after the burst, before active video) 8.2-8.9 us, on mid-field lines
    tip = out[:, 40:200, 10:50]
    porch = out[:, 40:200, 117:128]
    assert abs(np.median(tip) - 1024) < 3 * osc
    assert abs(np.median(porch) - blank) < 3 * osc
