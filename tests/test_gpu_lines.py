"""GPU parity of the line-location refinements (NTSC colour burst, PAL
pilot, fix_badlines, subcarrier shift) against the oracle.

Tolerances: on the same demod buffers the GPU reproduces the oracle's line
locations within 1e-9 samples (FP64 on both sides; the PAL circular mean's
sin/cos/atan2 may differ in the last bit), with identical bad-line flags,
phase-adjust medians and field phase IDs.  From the GPU's own FP32 demod the
locations agree within 0.02 samples on lines both sides call good.
"""

import types

import numpy as np
import pytest

from lddecode_amd import capture, synth
from lddecode_amd import field as fieldmod
from oracle import dsp as O
from oracle import lines as L
from oracle import tbc as T

pytestmark = pytest.mark.gpu


def _batch(system, nf, seed):
    orf = O.OracleRF(system, decode_analog_audio=False)
    rf = types.SimpleNamespace(system=system, blockcut=1024, SysParams=orf.SysParams, block_stride=31712,
                               blocklen=32768)
    fl = 667_334 if system == "NTSC" else 800_000
    cap = synth.make_capture(system, (nf + 2) * fl + 100_000, seed=seed, snr_db=40.0, jitter_ns=30)
    fields, ll1 = capture.truth_fields(rf, cap, nf, first_field=1)
    org0 = fields[0]["demod_offset"]
    nbt = (fields[-1]["demod_offset"] - org0) // 31712 + (29 if system == "NTSC" else 33)
    ref, _ = O.demod_span(orf, cap.data.astype(np.float64), org0, nbt * 31712 - 1)
    video = {k: np.ascontiguousarray(v.astype(np.float32)) for k, v in ref.items()}
    return orf, cap, fields, ll1, org0, nbt, video


def _oracle_lines(orf, video, fields, ll1, org0, system):
    p = L.LineRefineParams(types.SimpleNamespace(SysParams=orf.SysParams, DecoderParams=orf.DecoderParams,
                                                 system=system, freq=40.0, linelen=orf.linelen))
    nl = fields[0]["nlines"]
    ll2s, bads, lf, out_ll, out_bad, meds, fpids = [], [], [], [], [], [], []
    prev = 0.0
    for f, fd in enumerate(fields):
        off = fd["demod_offset"] - org0
        l1 = ll1[f * nl:(f + 1) * nl]
        ll2, bad = T.refine_linelocs_hsync(video["demod_05"][off:], l1, np.zeros(nl, bool), system, 40.0,
                                           orf.iretohz, orf.DecoderParams["vsync_ire"])
        first = (f % 2) == 0
        lc = 263 if first else 262
        badc = L.compute_deriv_error(ll2, bad)
        if system == "NTSC":
            ll, b, fpid, med = L.ntsc_field_lines(p, video["demod"][off:], video["demod_burst"][off:], ll2, badc,
                                                  lc, first, prev)
            prev = med
            meds.append(med)
            fpids.append(fpid)
        else:
            ll, b = L.pal_field_lines(p, video["demod_pilot"][off:], ll2, badc)
        ll2s.append(ll2)
        bads.append(bad)
        out_ll.append(ll)
        out_bad.append(b)
        lf.append(dict(demod_offset=off, demod_len=len(video["demod"]) - off, lineloc_offset=f * nl, nlines=nl,
                       linecount=lc, first_field=int(first)))
    return (np.concatenate(ll2s), np.concatenate(bads), lf, np.concatenate(out_ll), np.concatenate(out_bad),
            meds, fpids)


@pytest.mark.parametrize("system,nf", [("NTSC", 3), ("PAL", 2)])
def test_refine_same_input_matches_oracle(gpu_rf, system, nf):
    """Same demod buffers and hsync results on both sides; NTSC with three
    fields so each field's burst search starts from the previous median."""
    import torch
    rf = gpu_rf(system, audio=False)
    orf, cap, fields, ll1, org0, nbt, video = _batch(system, nf, seed=2)
    ll2, bad, lf, exp_ll, exp_bad, meds, fpids = _oracle_lines(orf, video, fields, ll1, org0, system)
    dev = {k: torch.from_numpy(v).cuda() for k, v in video.items()}
    ll, b, phase, fpid = fieldmod.refine_lines_fields(rf, dev, lf, torch.from_numpy(ll2).cuda(),
                                                      torch.from_numpy(bad.astype(np.uint8)).cuda())
    ll, b = ll.cpu().numpy(), b.cpu().numpy().astype(bool)
    np.testing.assert_array_equal(b, exp_bad)
    assert np.abs(ll - exp_ll).max() <= 1e-9
    if system == "NTSC":
        np.testing.assert_allclose(phase.cpu().numpy(), meds, rtol=0, atol=1e-12)
        assert fpid.cpu().numpy().tolist() == fpids


def test_refine_from_gpu_demod(gpu_rf):
    """The whole NTSC line path on the GPU (FP32 demod, FP64 hsync, burst
    refine) against the oracle's FP64 path."""
    import torch
    rf = gpu_rf("NTSC", audio=False)
    orf, cap, fields, ll1, org0, nbt, video = _batch("NTSC", 2, seed=7)
    ll2o, bado, lf, exp_ll, exp_bad, _, _ = _oracle_lines(orf, video, fields, ll1, org0, "NTSC")
    x = torch.from_numpy(cap.data).cuda()
    got = rf.demod_blocks(x, org0, nbt, audio=False)
    hs = [dict(demod_offset=d["demod_offset"], demod_len=d["demod_len"], lineloc_offset=d["lineloc_offset"],
               nlines=d["nlines"]) for d in lf]
    ll2, bad = fieldmod.refine_linelocs_hsync_fields(rf, got["video"]["demod_05"], hs, torch.from_numpy(ll1).cuda())
    ll, b, _, _ = fieldmod.refine_lines_fields(rf, got["video"], lf, ll2, bad)
    ll, b = ll.cpu().numpy(), b.cpu().numpy().astype(bool)
    both = ~b & ~exp_bad
    assert both.mean() > 0.7
    assert np.abs(ll - exp_ll)[both].max() <= 0.02
    assert (b == exp_bad).mean() >= 0.97


def test_field_process_lines(gpu_rf):
    """Field.process_lines on one PAL field equals the batched entry point."""
    import torch
    rf = gpu_rf("PAL", audio=False)
    orf, cap, fields, ll1, org0, nbt, video = _batch("PAL", 1, seed=3)
    ll2, bad, lf, exp_ll, exp_bad, _, _ = _oracle_lines(orf, video, fields, ll1, org0, "PAL")
    dev = {"video": {k: torch.from_numpy(v).cuda() for k, v in video.items()}}
    f = fieldmod.Field(rf, dev, linelocs=ll2, linebad=bad.astype(np.uint8))
    ll = f.process_lines(linecount=263, first_field=True).cpu().numpy()
    assert np.abs(ll - exp_ll).max() <= 1e-9
    np.testing.assert_array_equal(f.linebad.cpu().numpy(), exp_bad)
