"""libldgpu's host-side line location (ld_locate_field / ld_locate_fields,
no GPU needed) against the oracle on identical pulse lists.

The pulse lists come from the oracle's FP64 demodulation of synthetic
captures, so both sides see exactly the same integers; everything the
reference decides from them (pulse classes, vertical blanking, line 0,
field parity, the line table and its gap fill, the next field's offset)
must then agree exactly.  Damaged pulse trains (pulses removed, lengths
disturbed) drive the fallback paths of the state machine and processVBlank.
"""

import types

import numpy as np
import pytest

from lddecode_amd import _lib, locate, synth
from oracle import dsp as O
from oracle import linelocs as LL
from oracle import tbc as T


def _rf(system):
    orf = O.OracleRF(system, decode_analog_audio=False)
    return types.SimpleNamespace(lib=_lib.load(), system=system, SysParams=orf.SysParams,
                                 DecoderParams=orf.DecoderParams, freq=orf.freq, linelen=orf.linelen,
                                 blocklen=32768, blockcut=1024, block_stride=31712, hsync_tolerance=0.4), orf


_CACHE = {}


def _capture(system, nfields=5, seed=3):
    key = (system, nfields, seed)
    if key not in _CACHE:
        orf = O.OracleRF(system, decode_analog_audio=False)
        fl = 667_334 if system == "NTSC" else 800_000
        cap = synth.make_capture(system, nfields * fl, seed=seed, snr_db=40.0, jitter_ns=50, audio=False)
        nbt = (len(cap.data) - 32768) // 31712 + 1
        ref, _ = O.demod_span(orf, cap.data.astype(np.float64), 0, nbt * 31712 - 1)
        _CACHE[key] = (ref["demod_05"].astype(np.float32), ref["demod"].astype(np.float32))
    return _CACHE[key]


def _edges_means(p, d05):
    """ld_find_pulses / ld_pulse_means' outputs, computed in numpy."""
    below = d05.astype(np.float64) <= p.iretohz(-20)
    prev = np.concatenate([[False], below[:-1]])
    edges = np.flatnonzero(below != prev).astype(np.int64)
    s, e = edges[0:len(edges) - 1:2], edges[1::2]
    s = s[:len(e)]
    means = np.full(len(s), np.nan, np.float32)
    for k in range(len(s)):
        n = e[k] - s[k]
        if n > p.usectoinpx(10) and n <= 5000:
            means[k] = LL.np_mean_f32(d05[int(s[k] + p.freq):int(s[k] + n - p.freq)])
    return edges, means


def _loc_equal(c, o, startloc):
    """One C walk entry against one oracle field."""
    res = o["res"]
    assert c.startloc == o["startloc"] == startloc
    assert c.start == pytest.approx(o["start"], abs=0, rel=0)
    assert (c.status == _lib.LD_LOC_OK) == o["valid"]
    if not o["valid"]:
        assert c.offset == res.nextfieldoffset if res.nextfieldoffset is not None else True
        return
    assert bool(c.is_first_field) == res.isFirstField
    assert c.linecount == res.linecount
    assert c.sync_confidence == res.sync_confidence
    assert c.line0loc == res.line0loc
    assert c.meanlinelen == res.meanlinelen
    assert c.vblank_next == (res.vblank_next if res.vblank_next is not None else -1)
    assert c.nextfieldoffset == res.nextfieldoffset


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_walk_matches_oracle(system):
    rf, orf = _rf(system)
    p = LL.LocParams.from_rf(orf)
    d05, dm = _capture(system)
    edges, means = _edges_means(p, d05)

    def finish(k, res, org):
        ll2, _ = T.refine_linelocs_hsync(d05[org:], res.linelocs1, res.linebad, system, 40.0, orf.iretohz,
                                         orf.DecoderParams["vsync_ire"])
        return ll2

    ora = LL.walk_fields(p, d05, 5, 0, finish, demod=dm)
    facts = []
    for o in ora:
        f = _lib.FieldFacts()
        if o["valid"]:
            llc, meds = LL.field_facts(p, dm[o["startloc"]:o["startloc"] + o["nblocks"] * 31712], o["final"],
                                       o["res"].isFirstField)
            f.ll_linecount = llc
            f.medians[:] = meds
            f.known = 1
        facts.append(f)
    locs, ll, bad, _, _ = locate.walk(rf, edges, means, len(d05), 0.0, None, facts, max_fields=5)
    assert len(locs) == len(ora)
    for k, (c, o) in enumerate(zip(locs, ora)):
        _loc_equal(c, o, o["startloc"])
        if o["valid"]:
            np.testing.assert_array_equal(ll[k], o["res"].linelocs1)
            np.testing.assert_array_equal(bad[k].astype(bool), o["res"].linebad)
    assert sum(o["valid"] for o in ora) == 3
    # the walk with provisional facts reaches the same fields here
    locs2, ll2, _, _, _ = locate.walk(rf, edges, means, len(d05), 0.0, None, None, max_fields=5)
    assert [c.startloc for c in locs2] == [c.startloc for c in locs]
    np.testing.assert_array_equal(ll2, ll)


def test_walk_many_equals_sequential_walks():
    """ld_locate_fields_many over several jobs on several threads (NTSC and
    PAL, with and without facts, different start points) returns exactly
    what each ld_locate_fields call returns alone."""
    calls = []
    for system in ("NTSC", "PAL"):
        rf, orf = _rf(system)
        p = LL.LocParams.from_rf(orf)
        d05, _ = _capture(system)
        edges, means = _edges_means(p, d05)
        for start, facts in ((0.0, None), (float(len(d05) // 3), None), (0.0, [_lib.FieldFacts()] * 2)):
            calls.append(dict(rf=rf, edges=edges, means=means, total=len(d05), start=start, facts=facts,
                              max_fields=4))
    want = [locate.walk(**c) for c in calls]
    for threads in (1, 3, 16):
        got = locate.walk_many(calls, threads=threads)
        assert len(got) == len(want)
        for (l1, a1, b1, p1, v1), (l2, a2, b2, p2, v2) in zip(want, got):
            assert [bytes(c) for c in l1] == [bytes(c) for c in l2]
            np.testing.assert_array_equal(a1, a2)
            np.testing.assert_array_equal(b1, b2)
            np.testing.assert_array_equal(v1, v2)
            assert bytes(p1) == bytes(p2)
    assert locate.walk_many([]) == []


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_prev_field_dependence(system):
    """getLine0's inputs from the previous field: a final location moved by
    a fraction of a line changes the median of three (line0loc follows it
    when it is the middle value), and a skip score below 50 lifts
    processVBlank's limit.  The C walk follows the oracle in both."""
    rf, orf = _rf(system)
    p = LL.LocParams.from_rf(orf)
    d05, dm = _capture(system)
    edges, means = _edges_means(p, d05)
    for shift, quiet in ((0.3 * p.linelen, False), (-0.35 * p.linelen, True), (3.0, False)):
        def finish(k, res, org):
            return res.linelocs1 + shift

        ora = LL.walk_fields(p, d05, 5, 0, finish, demod=None if not quiet else np.zeros_like(dm) + p.ire0)
        facts = []
        for o in ora:
            f = _lib.FieldFacts()
            if o["valid"]:
                lc = o["res"].linecount if system == "NTSC" else (312 if o["res"].isFirstField else 313)
                f.ll_linecount = o["final"][lc]
                f.medians[:] = [p.ire0] * 8 if quiet else [p.iretohz(p.vsync_ire)] * 8
                f.known = 1
            facts.append(f)
        locs, ll, bad, _, _ = locate.walk(rf, edges, means, len(d05), 0.0, None, facts, max_fields=5)
        assert len(locs) == len(ora)
        for k, (c, o) in enumerate(zip(locs, ora)):
            _loc_equal(c, o, o["startloc"])
            if o["valid"]:
                np.testing.assert_array_equal(ll[k], o["res"].linelocs1)


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_damaged_pulse_trains(system):
    """compute_linelocs on pulse lists with pulses removed or lengths
    disturbed (seeded): C and oracle agree on every outcome, including
    failed fields and the gap fill."""
    rf, orf = _rf(system)
    p = LL.LocParams.from_rf(orf)
    d05, _ = _capture(system)
    S = 31712
    ora = LL.walk_fields(p, d05, 3, 0, lambda k, res, org: res.linelocs1)
    start = next(o["start"] for o in ora if o["valid"])
    _, startloc, nb = LL.field_window(p, start)
    buf = d05[startloc:startloc + nb * S]
    starts, lens = LL.findpulses(buf, p.iretohz(-20))
    rng = np.random.default_rng(11)
    outcomes = set()
    for trial in range(40):
        keep = rng.random(len(starts)) > rng.choice([0.0, 0.02, 0.1, 0.3])
        if trial % 3 == 0:      # drop a run of pulses (a dropout across the vblank, sometimes)
            a = rng.integers(0, len(starts) - 30)
            keep[a:a + rng.integers(5, 30)] = False
        s = starts[keep].copy()
        n = lens[keep].copy()
        if trial % 4 == 1:
            n = n + rng.integers(-30, 30, size=len(n))
            n = np.maximum(n, 1)
        prev = None
        if trial % 5 == 2:
            prev = LL.PrevField(True, startloc - 21 * S, np.zeros(400) + 1234.5, 263, bool(trial % 2), 90,
                                int(rng.choice([0, 50, 100])))
        res = LL.FieldLines(None, None, None)
        try:
            LT, valid = LL.refinepulses(p, s, n)
            ml = LL.compute_line_len(p, valid)
            res = _oracle_from_pulses(p, s, n, prev, startloc)
        except (ValueError, TypeError, IndexError):
            res = None
        cprev = None
        if prev is not None:
            cprev = _lib.PrevField(valid=1, is_first_field=int(prev.isFirstField), linecount=prev.linecount,
                                   sync_confidence=prev.sync_confidence, startloc=prev.startloc,
                                   ll_linecount=prev.linelocs[prev.linecount],
                                   skip_score=prev.skip_score, fields_written=3)
        if res is None:
            with pytest.raises(_lib.LdGpuError):
                locate.locate_one(rf, s, n, cprev, startloc, 3)
            outcomes.add("raise")
            continue
        c, ll, bad = locate.locate_one(rf, s, n, cprev, startloc, 3)
        assert (c.status == _lib.LD_LOC_OK) == (res.linelocs1 is not None), trial
        assert c.nextfieldoffset == res.nextfieldoffset, trial
        if res.linelocs1 is not None:
            np.testing.assert_array_equal(ll, res.linelocs1)
            np.testing.assert_array_equal(bad.astype(bool), res.linebad)
            assert bool(c.is_first_field) == res.isFirstField and c.sync_confidence == res.sync_confidence
            outcomes.add("ok" if not res.linebad.any() else "filled")
        else:
            outcomes.add("failed")
    assert {"ok", "filled", "failed"} <= outcomes, outcomes


def _oracle_from_pulses(p, starts, lens, prev, startloc):
    """compute_linelocs after getpulses, from a given pulse list."""
    orig = LL.getpulses
    try:
        LL.getpulses = lambda p_, d_, fw_=1: (starts, lens, None)
        return LL.compute_linelocs(p, None, prev, startloc, 3)
    finally:
        LL.getpulses = orig


def test_skip_score_closed_forms():
    rf, orf = _rf("NTSC")
    p = LL.LocParams.from_rf(orf)
    P = locate.locate_params(rf)
    sync, black, white = p.iretohz(p.vsync_ire), p.iretohz(0), p.iretohz(100)
    for meds, want in (([sync] * 2 + [black] * 6, 100), ([sync] + [black] * 4 + [white] * 3, 50),
                       ([black] * 5 + [white] * 3, 25), ([white] * 8, 0)):
        f = _lib.FieldFacts()
        f.medians[:] = meds
        assert locate.skip_scores(rf, [f], P)[0] == want == LL.skip_score_from(p, meds)


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_recheck_settles_to_sequential_walk(system):
    """The batched decoder's scheme on the host: walk on provisional
    previous-field values, measure each field's facts from its line table
    (here final = linelocs1 + a shift, standing in for the GPU alignment),
    re-check, and walk again from the first changed field until the check
    passes.  The settled records equal the oracle's sequential walk, line 0
    included, for shifts that leave the line tables alone and for ones that
    move line 0 across a pulse."""
    rf, orf = _rf(system)
    p = LL.LocParams.from_rf(orf)
    d05, _ = _capture(system)
    edges, means = _edges_means(p, d05)
    for shift in (1.7, 0.3 * p.linelen, -0.45 * p.linelen):
        def facts_for(locs, ll):
            out = []
            for k, c in enumerate(locs):
                f = _lib.FieldFacts()
                if c.status == _lib.LD_LOC_OK:
                    lc = c.linecount if system == "NTSC" else (312 if c.is_first_field else 313)
                    f.ll_linecount = ll[k][lc] + shift
                    f.medians[:] = [p.iretohz(p.vsync_ire)] * 8
                    f.known = 1
                out.append(f)
            return out

        locs, ll, bad, _, vp = locate.walk(rf, edges, means, len(d05), 0.0, None, None, max_fields=5)
        for _ in range(6):
            facts = facts_for(locs, ll)
            first, checked = locate.recheck(rf, locs, ll, bad, vp, facts)
            if first < 0:
                locs = checked
                break
            keep = facts[:first] + [_lib.FieldFacts() for _ in facts[first:]]
            locs, ll, bad, _, vp = locate.walk(rf, edges, means, len(d05), 0.0, None, keep, max_fields=5)
        else:
            raise AssertionError("did not settle")
        ora = LL.walk_fields(p, d05, 5, 0, lambda k, res, org: res.linelocs1 + shift)
        assert len(ora) == len(locs)
        for k, (c, o) in enumerate(zip(locs, ora)):
            _loc_equal(c, o, o["startloc"])
            if o["valid"]:
                np.testing.assert_array_equal(ll[k], o["res"].linelocs1)
