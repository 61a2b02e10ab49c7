"""The line-location oracle (oracle/linelocs.py) against closed forms and
synthetic ground truth.

The reference's line location has no golden vectors, and its module cannot be
imported here (lddecode.core needs numba at import time), so these checks pin
the restatement instead:
  * findpulses on hand-built level sequences (index-0 and unterminated runs
    dropped, the 5000-sample limit, equality counts as "in pulse");
  * an ideal sync raster (pulses at exact half-line times, no filtering):
    every line start is the exact first pulse sample, first / second field
    alternate as the raster was built, the walk steps one field at a time,
    and a missing hsync pulse is filled by the closed-form interpolation and
    flagged bad;
  * demodulated synthetic captures (FP64 demod): line starts sit a constant
    edge delay after the true line starts, with the right field parity.
"""

import numpy as np
import pytest

from lddecode_amd import synth
from lddecode_amd.params import system_tables
from oracle import dsp as O
from oracle import linelocs as LL


def _params(system):
    return LL.LocParams.from_rf(O.OracleRF(system, decode_analog_audio=False))


# ---------------------------------------------------------------- findpulses
def test_findpulses_closed_forms():
    hi = 5.0
    x = np.full(40, 10.0, np.float32)
    x[0:3] = 0.0          # starts at index 0: dropped
    x[5:9] = 5.0          # equal to the threshold counts: start 5, len 4
    x[12] = 1.0           # one sample: start 12, len 1
    x[20:25] = -3.0       # start 20, len 5
    x[37:] = 0.0          # never ends: dropped
    s, n = LL.findpulses(x, hi)
    assert s.tolist() == [5, 12, 20] and n.tolist() == [4, 1, 5]
    y = np.full(12000, 10.0, np.float32)
    y[100:5100] = 0.0     # 5000 samples: kept
    y[6000:11001] = 0.0   # 5001 samples: dropped
    s, n = LL.findpulses(y, hi)
    assert s.tolist() == [100] and n.tolist() == [5000]
    # float32 data against a float64 threshold: 5.0000001 rounds to 5.0 in
    # float32, so a float32 comparison would call this sample a pulse
    z = np.full(10, 10.0, np.float32)
    z[4] = np.float32(5.0)
    s, _ = LL.findpulses(z, 4.9999999)
    assert s.tolist() == []


def test_numba_median_and_mean():
    a = np.array([3.0, 1.0, 2.0, 10.0], np.float32)
    assert LL.numba_median(a) == 2.5
    assert LL.numba_median(a[:3]) == 2.0
    r = np.random.default_rng(0).normal(size=1001).astype(np.float32)
    assert LL.np_mean_f32(r) == np.mean(r)


# ------------------------------------------------------------ ideal raster
def _ideal(system, nfields, drop=()):
    """Sync-only demod_05 in Hz over nfields + 2 fields, built from the
    synthetic raster with no jitter: sync pulses at the sync level, the rest
    at blanking.  drop: half-line indices whose pulse is removed."""
    sysp, _ = system_tables(system)
    hpf = 525 if system == "NTSC" else 625
    n = int((nfields + 2) * hpf * sysp["line_period"] * 20) + 1000
    ire, _, hstart = synth.ire_raster(system, sysp, n, 40e6, 0.0, np.random.default_rng(0))
    sync = ire <= sysp["vsync_ire"] + 0.5
    for h in drop:
        a = int(np.ceil(hstart[h]))
        sync[a:a + 400] = False
    x = np.where(sync, sysp["ire0"] + sysp["hz_ire"] * sysp["vsync_ire"], sysp["ire0"]).astype(np.float32)
    return x, hstart


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_ideal_raster_walk(system):
    p = _params(system)
    x, hstart = _ideal(system, 4)
    out = LL.walk_fields(p, x, 5, 0, lambda k, res, org: res.linelocs1, demod=x)
    valid = [r for r in out if r["valid"]]
    assert len(valid) >= 3
    hpf = 525 if system == "NTSC" else 625
    fields_seen = []
    for r in valid:
        res = r["res"]
        org = r["startloc"]
        # which raster field is this?  line 0 sits on field f's line-0 half-line
        f = int(round((res.linelocs1[0] + org) / (hpf * p.line_period * 20)))
        want = np.ceil(hstart[synth.field_line0(system, f) + 2 * np.arange(p.proclines)]) - org
        np.testing.assert_array_equal(res.linelocs1, want)
        assert not res.linebad.any()
        assert res.isFirstField == (f % 2 == 0)
        assert res.linecount == (263 if f % 2 == 0 else 262)
        fields_seen.append(f)
    assert np.all(np.diff(fields_seen) == 1), fields_seen


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_ideal_raster_vblank_and_timings(system):
    p = _params(system)
    x, hstart = _ideal(system, 2)
    starts, lens = LL.findpulses(x, p.iretohz(-20))
    LT, valid = LL.refinepulses(p, starts, lens)
    # pulse lengths are exact sample counts: hsync 4.7 us = 188 samples
    assert LT["hsync_median"] == 188
    kinds = np.array([v[0] for v in valid])
    # the raster's second field (field 1) vblank: numPulses of each kind
    seg = kinds[(np.array([v[1][0] for v in valid]) > hstart[525 if system == "NTSC" else 625] - 100)]
    first_blank = np.flatnonzero(seg > 0)[0]
    run = seg[first_blank:first_blank + 3 * p.numPulses]
    assert run.tolist() == ([LL.EQPL1] * p.numPulses + [LL.VSYNC] * p.numPulses + [LL.EQPL2] * p.numPulses)
    assert LL.compute_line_len(p, valid) == pytest.approx(p.line_period * 40, abs=0.01)


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_missing_hsync_is_interpolated(system):
    """Remove the hsync pulse of line 100 of field 2: its location is the
    average of its neighbours' (closed form of the gap fill) and it is
    flagged bad; every other line is unchanged."""
    p = _params(system)
    h = synth.field_line0(system, 2) + 2 * 100
    x0, hstart = _ideal(system, 3)
    x1, _ = _ideal(system, 3, drop=(h,))
    def first_field2(x):
        out = LL.walk_fields(p, x, 5, 0, lambda k, res, org: res.linelocs1, demod=x)
        for r in out:
            if r["valid"] and r["res"].isFirstField:
                return r
        raise AssertionError("no first field")
    r0, r1 = first_field2(x0), first_field2(x1)
    a, b = r0["res"].linelocs1, r1["res"].linelocs1
    assert r0["startloc"] == r1["startloc"]
    assert b[100] == (a[99] + a[101]) / 2
    assert r1["res"].linebad[100] and r1["res"].linebad.sum() == 1
    np.testing.assert_array_equal(np.delete(a, 100), np.delete(b, 100))


def test_process_vblank_parity_and_line0():
    """processVBlank's line 0 on an NTSC pulse list built by hand: the
    first field's line 0 is 4 lines before its first vsync pulse and the
    second field's 3.5 lines (firstFieldH = (0.5, 1), distfroml1 = 3)."""
    p = _params("NTSC")
    L = p.linelen
    for first, gap in ((True, 2), (False, 1)):
        pulses = [(1000 + k * L, 188) for k in range(20)]             # hsync lines
        h0 = pulses[-1][0] + gap * L / 2                               # first eq pulse
        for k in range(18):
            n = 1084 if 6 <= k < 12 else 92
            pulses.append((int(round(h0 + k * L / 2)), n))
        last = pulses[-1][0]
        pulses += [(int(round(last + L / 2 + k * L)), 188) for k in range(1, 30)]
        s = np.array([q[0] for q in pulses]); n = np.array([q[1] for q in pulses])
        _, valid = LL.refinepulses(p, s, n)
        line0, isfirst, fb, conf, _ = LL.process_vblank(p, valid, 0)
        vs = int(round(h0 + 6 * L / 2))
        assert isfirst == first and conf == 100
        assert line0 == int(vs - (4.0 if first else 3.5) * L)


# ------------------------------------------------ demodulated synthetic RF
@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_walk_on_demodulated_capture(system):
    orf = O.OracleRF(system, decode_analog_audio=False)
    p = LL.LocParams.from_rf(orf)
    fl = 667_334 if system == "NTSC" else 800_000
    cap = synth.make_capture(system, 5 * fl, seed=3, snr_db=40.0, jitter_ns=50, audio=False)
    nbt = (len(cap.data) - 32768) // 31712 + 1
    ref, _ = O.demod_span(orf, cap.data.astype(np.float64), 0, nbt * 31712 - 1)
    d05 = ref["demod_05"].astype(np.float32)
    out = LL.walk_fields(p, d05, 5, 0, lambda k, res, org: res.linelocs1 + 13.0, demod=ref["demod"].astype(np.float32))
    valid = [r for r in out if r["valid"]]
    assert len(valid) == 3
    hpf = 525 if system == "NTSC" else 625
    hs = cap.halfline_starts
    for r in valid:
        res = r["res"]
        raw = res.linelocs1 + r["startloc"] + 1024
        f = int(round(raw[0] / (hpf * p.line_period * 20)))
        idx = synth.field_line0(system, f) + 2 * np.arange(p.proclines)
        d = raw - hs[idx]
        # the pulse edge of demod_05 trails the true start by its filter delay
        assert 9.5 < d.min() and d.max() < 14.5, (d.min(), d.max())
        assert res.isFirstField == (f % 2 == 0)
        assert not res.linebad.any()
