"""Known answers, derived by hand from the reference's code, for the
reference-specific decisions of line location (A10), the NTSC burst / PAL
pilot alignment (A12) and computedelays -- the parts no golden vector or
closed form pins otherwise.  Each input is built so that the answer can be
worked out on paper from the cited lines (the derivations are in the
comments), and each test names the plausible mistake it catches.  The
line-0 cases also run through the library's native walk (ld_locate_field,
host code), which must reach the same answers.

NTSC constants used below (40 MHz): rf.linelen L = 2542 samples, H = L / 2,
usectoinpx(x) = x * 40; get_timings (core.py:1761-1810) with every hsync
188 samples long gives LT hsync (168, 208), eq (72, 112), vsync
(542, 1124); numPulses 6, firstFieldH (0.5, 1), field_lines (263, 262).
"""

import types

import numpy as np
import pytest

from lddecode_amd import _lib, locate
from oracle import dsp as O
from oracle import lines as LN
from oracle import linelocs as LL

L, H = 2542, 1271
HS, EQ, VS = 188, 92, 800          # hsync 4.7 us, eq 2.3 us, vsync 20 us pulse lengths
HSYNC, EQPL1, VSYNC, EQPL2 = LL.HSYNC, LL.EQPL1, LL.VSYNC, LL.EQPL2


def _p(system="NTSC"):
    return LL.LocParams.from_rf(O.OracleRF(system, decode_analog_audio=False))


def _lt(p):
    LT = LL.get_timings(p, [HS] * 20)
    return LT


def test_timings_by_hand():
    """get_timings (core.py:1761-1810): median hsync 188, +-usectoinpx(0.5);
    eq = usectoinpx(2.3 -+ 0.5); vsync = (usectoinpx(27.1 / 2), usectoinpx(28.1))."""
    LT = _lt(_p())
    assert LT["hsync"] == pytest.approx((168.0, 208.0))
    assert LT["eq"] == pytest.approx((72.0, 112.0))
    assert LT["vsync"] == pytest.approx((542.0, 1124.0))


# ------------------------------------------------------------ state machine
def _vblank_train():
    """Two hsyncs, then 6 eq, 6 vsync, 6 eq at half-line spacing, then an
    hsync half a line after the last eq."""
    return ([(0, HS), (L, HS)] + [(2 * L + k * H, EQ) for k in range(6)] +
            [(5 * L + k * H, VS) for k in range(6)] + [(8 * L + k * H, EQ) for k in range(6)] + [(11 * L, HS)])


def _machine(pulses):
    p = _p()
    done, valid = LL.run_vblank_state_machine(p, pulses, _lt(p))
    return done, [v[0] for v in valid], [v[1][0] for v in valid], [v[2] for v in valid]


def test_state_machine_full_vblank():
    """core.py:1826-1911: HSYNC -> EQPL1 -> VSYNC -> EQPL2 -> HSYNC (done).
    Quality (core.py:1812-1823): hsync to hsync 1.0 line in (0.9, 1.1);
    hsync -> eq 1.0 in (0.4, 1.1); eq / vsync spacing 0.5 in (0.4, 0.6); the
    closing hsync 0.5 line after the last eq in (0.4, 1.1); the first pulse
    has no predecessor: False."""
    done, types_, starts, good = _machine(_vblank_train())
    assert done is True
    assert types_ == [HSYNC] * 2 + [EQPL1] * 6 + [VSYNC] * 6 + [EQPL2] * 6 + [HSYNC]
    assert starts[-1] == 11 * L
    assert good == [False] + [True] * 20


def test_state_machine_rejects_transition_before_state_end():
    """A vsync-length pulse at 4.7 L, before the first eq's state_end =
    2 L + (3 - 0.1) L = 4.9 L (core.py:1889-1897), is dropped; the train
    then reads as without it.  Mistake caught: no state_end rejection (the
    early pulse would open VSYNC and appear in the list)."""
    tr = _vblank_train()
    tr.insert(8, (int(4.7 * L), VS))
    done, types_, starts, _ = _machine(tr)
    assert done is True
    assert int(4.7 * L) not in starts
    assert types_ == [HSYNC] * 2 + [EQPL1] * 6 + [VSYNC] * 6 + [EQPL2] * 6 + [HSYNC]


@pytest.mark.parametrize("at,accepted", [(12456, True), (12455, False)])
def test_state_end_uses_state_length_minus_a_tenth(at, accepted):
    """state_end = start + (state_length - 0.1) * inlinelen = 5080 + 2.9 L =
    12455.8: a vsync at 12456 opens VSYNC, one at 12455 is dropped.  Mistake
    caught: state_end without the -0.1 (12710 would drop both)."""
    done, types_, _, _ = _machine([(0, HS), (L, HS), (2 * L, EQ), (at, VS)])
    assert done is False
    assert types_ == [HSYNC, HSYNC, EQPL1] + ([VSYNC] if accepted else [])


def test_vsync_to_hsync_waits_for_state_end():
    """In VSYNC an hsync-length pulse is taken only after state_end = 5 L +
    2.9 L = 7.9 L (core.py:1880-1881): one at 7.7 L is ignored.  One at 8 L
    is taken (HSYNC; state_end stays 7.9 L, no state_length pending), so the
    eq pulses after it open EQPL1 again at 8.5 L (state_end 8.5 L + 2.9 L =
    11.4 L) and the hsync at 11 L, a transition before 11.4 L, is dropped:
    the machine ends not done."""
    tr = _vblank_train()
    tr.insert(14, (int(7.7 * L), HS))
    done, types_, starts, _ = _machine(tr)
    assert done is True and int(7.7 * L) not in starts
    assert types_ == [HSYNC] * 2 + [EQPL1] * 6 + [VSYNC] * 6 + [EQPL2] * 6 + [HSYNC]
    tr = _vblank_train()
    tr[14] = (8 * L, HS)                    # the first post-vsync eq becomes an hsync
    done, types_, starts, _ = _machine(tr)
    assert done is False
    assert types_ == [HSYNC] * 2 + [EQPL1] * 6 + [VSYNC] * 6 + [HSYNC] + [EQPL1] * 5
    assert 11 * L not in starts


def test_eqpl2_early_hsync_ends_the_machine_unappended():
    """EQPL2 -> hsync sets done before the state_end check (core.py:1884-1887);
    an hsync at 10.5 L (state_end 8 L + 2.9 L = 10.9 L) is then dropped and
    the machine returns done with the list ending in EQPL2.  Mistake caught:
    'done only when appended' (it would go on to the hsync at 11 L)."""
    tr = _vblank_train()
    tr[19] = (int(10.5 * L), HS)            # the sixth post-vsync eq is hsync-length
    done, types_, starts, _ = _machine(tr)
    assert done is True
    assert types_ == [HSYNC] * 2 + [EQPL1] * 6 + [VSYNC] * 6 + [EQPL2] * 5
    assert 11 * L not in starts


def test_quality_flags_by_spacing():
    """core.py:1812-1823: hsync at 1.2 L after an hsync: 1.2 outside (0.9, 1.1)
    -> False; the eq 0.8 L later: inside (0.4, 1.1) -> True."""
    _, types_, _, good = _machine([(0, HS), (int(1.2 * L), HS), (2 * L, EQ)])
    assert types_ == [HSYNC, HSYNC, EQPL1] and good == [False, False, True]


# ------------------------------------------------------------ processVBlank
def _field(lead, off, first=True, shift=(0, 0, 0), nvs=6, neq1=6, neq2=6, tail=10, spacing=None):
    """Classified pulses (type, (start, len), good) of `lead` hsyncs from
    line `off`, a vblank (EQPL1, VSYNC, EQPL2 groups at half-line spacing,
    each group's starts shifted by `shift`), then `tail` hsyncs.  first:
    the vblank starts one line after the last lead hsync (first-field
    timing), else 1.5 lines."""
    v = [(HSYNC, ((off + k) * L, HS), True) for k in range(lead)]
    t = (off + lead + (0 if first else 0.5)) * L
    sp = spacing or [0] * 18
    for k in range(neq1):
        v.append((EQPL1, (int(t + k * H) + shift[0] + sp[k], EQ), True))
    t += neq1 * H
    for k in range(nvs):
        v.append((VSYNC, (int(t + k * H) + shift[1] + sp[6 + k % 6], VS), True))
    t += nvs * H
    for k in range(neq2):
        v.append((EQPL2, (int(t + k * H) + shift[2] + sp[12 + k % 6], EQ), True))
    t += neq2 * H
    v += [(HSYNC, (int(t + k * L), HS), True) for k in range(tail)]   # half a line after the last eq
    return v


def test_process_vblank_first_field():
    """core.py:1988-2056.  Blank range: firstvsync = 16 (10 lead hsyncs, 6
    eq), firstblank 10, lastblank 27.  The VSYNC group (checked first) sits
    at j = 6 with 6 matches in the next 10 -> firstloc 13 L; distfroml1 =
    (2 - 1) * 6 / 2 = 3; dist = (13 L - 9 L) / L = 4 -> hdist 8 -> even ==
    (firstFieldH[1] != 1) = False -> first field; eqgap firstFieldH[1] = 1;
    line0 = 13 L - 4 L = 9 L = 22878, confidence 100."""
    r = LL.process_vblank(_p(), _field(10, 0), 0)
    assert r == (22878, True, 10, 100, False)


def test_process_vblank_second_field():
    """The vblank half a line later: firstloc 13.5 L, dist 4.5 -> hdist 9,
    odd -> second field; eqgap firstFieldH[0] = 0.5 -> line0 = 13.5 L -
    3.5 L = 10 L = 25420.  Mistake caught: parity or eqgap inverted."""
    r = LL.process_vblank(_p(), _field(10, 0, first=False), 0)
    assert r == (25420, False, 10, 100, False)


def test_process_vblank_takes_vsync_group_first():
    """VSYNC is tried before EQPL1 / EQPL2 (core.py:2013): with the vsync
    group 3 samples late its line0 (9 L + 3 = 22881) wins over the eq
    groups' 22878.  Mistake caught: EQPL1 first."""
    r = LL.process_vblank(_p(), _field(10, 0, shift=(0, 3, 0)), 0)
    assert r == (22881, True, 10, 100, False)


def test_process_vblank_group_needs_exactly_numpulses():
    """A 7-pulse vsync run (the first post eq read as vsync) fails
    ptmatch[j:j+10].sum() == 6 (core.py:2019-2020) and breaks out; the
    EQPL1 group (2 samples early) is used: firstloc 10 L - 2, distfroml1 0,
    dist ~ 1 -> hdist 2, first field, line0 = 10 L - 2 - L = 22876.
    Mistake caught: no '+4' run check (the vsync group would give 22878)."""
    v = _field(10, 0, shift=(-2, 0, 0))
    v[22] = (VSYNC, v[22][1], True)
    r = LL.process_vblank(_p(), v, 0)
    assert r == (22876, True, 10, 100, False)


def test_process_vblank_pal_vsync_flip():
    """PAL (L 2560, numPulses 5, firstFieldH (1, 0.5)): vsync group at
    12.5 L, presync 9 L -> dist 3.5, hdist 7, odd == (0.5 != 1) -> True,
    then distfroml1 = 2.5 is a half-line count -> flipped to False
    (core.py:2046-2047); eqgap firstFieldH[0] = 1 -> line0 = 12.5 L - 3.5 L
    = 9 L = 23040.  Mistake caught: the PAL flip (9.5 L, first field)."""
    LP, HP = 2560, 1280
    v = [(HSYNC, (k * LP, HS), True) for k in range(10)]
    v += [(EQPL1, (10 * LP + k * HP, EQ), True) for k in range(5)]
    v += [(VSYNC, (int(12.5 * LP) + k * HP, VS), True) for k in range(5)]
    v += [(EQPL2, (15 * LP + k * HP, EQ), True) for k in range(5)]
    v += [(HSYNC, (int(17.5 * LP) + k * LP, HS), True) for k in range(10)]
    r = LL.process_vblank(_p("PAL"), v, 0)
    assert r == (23040, False, 10, 100, False)


def test_process_vblank_fallback_half_confidence():
    """Every group's spacing broken (second differences of +-20 > 0.2 * 40):
    the fallback (core.py:2058-2088) checks the four edge pulses' quality
    and NTSC |gap1 + gap2| = |L + (0.5 L - 10)| in (1.4 L, 1.6 L);
    gap1 / L = 1 in (0.95, 1.05) -> first field; line0 = the presync start
    9 L = 22878 with confidence 50.  Mistake caught: confidence 100."""
    jig = [0, 10] * 9
    r = LL.process_vblank(_p(), _field(10, 0, spacing=jig), 0)
    assert r == (22878, True, 10, 50, False)


def test_process_vblank_fallback_failure_zeroes_confidence():
    """Same, but the presync 0.3 L late: |gap1 + gap2| = 1.2 L - 10 is out
    of range -> (None, None, None, 0) and the field's sync_confidence set
    to 0 (core.py:2081-2083)."""
    v = _field(10, 0, spacing=[0, 10] * 9)
    v[9] = (HSYNC, (int(9.3 * L), HS), True)
    r = LL.process_vblank(_p(), v, 0)
    assert r == (None, None, None, 0, True)


# ------------------------------------------------------------ getLine0
def _two_fields(b_shift=6, b_first=False, gap_lines=263):
    """Field A (first-field vblank after 10 hsyncs, line 0 at 9 L = 22878),
    hsyncs every line from 19 L, then field B's vblank half a line after
    hsync `19 + gap_lines - 10` (second-field timing), shifted by b_shift."""
    a = _field(10, 0, tail=0)
    last = 19 + gap_lines - 10                      # line of B's presync hsync
    a += [(HSYNC, (k * L, HS), True) for k in range(19, last + 1)]
    t = (last + 0.5) * L + b_shift
    b = [(EQPL1, (int(t + k * H), EQ), True) for k in range(6)]
    b += [(VSYNC, (int(t + 3 * L + k * H), VS), True) for k in range(6)]
    b += [(EQPL2, (int(t + 6 * L + k * H), EQ), True) for k in range(6)]
    b += [(HSYNC, (int(t + 9 * L + k * L), HS), True) for k in range(10)]
    return a + b


def _prev(ll0, first, conf=100, skip=0, startloc=0, linecount=263):
    ll = np.zeros(400)
    ll[linecount] = ll0
    return LL.PrevField(True, startloc, ll, linecount, first, conf, skip)


def test_line0_median_of_three():
    """core.py:2157-2220.  Local: 22878 (first field).  Next: field B's
    presync at 272 L = 691424, its vsync group at 275.5 L + 6 = 700327:
    dist 3.502 -> hdist 7 -> second field; line0 = 700327 - 3.5 L = 691430
    = vblank_next; isFirstField_next = not False; fieldlen = 263 L = 668546
    -> 22884.  Prev: its linelocs[263] = 122880 with startloc 100000 later
    -> 22880.  All three: median 22880 (mean would be 22880.67, local alone
    22878); votes local True + next True + (not prev.first) -> True;
    confidence 100."""
    p = _p()
    r = LL.get_line0(p, _two_fields(), float(L), _prev(122880, True), 100000)
    assert r == (22880.0, 691430, True, 100)


def test_line0_vote_negates_prev():
    """Field A a second field (line0 10 L = 25420, local False), field B too
    (next = not False = True); the previous field a second field ->
    isFirstField_prev = True (core.py:2203): two of three -> True.
    Mistake caught: prev's flag not negated (then one of three -> False)."""
    v = _field(10, 0, first=False, tail=0)
    v += [(HSYNC, (int((k + 0.5) * L), HS), True) for k in range(19, 273)]
    t = 273 * L
    v += [(EQPL1, (int(t + k * H), EQ), True) for k in range(6)]
    v += [(VSYNC, (int(t + 3 * L + k * H), VS), True) for k in range(6)]
    v += [(EQPL2, (int(t + 6 * L + k * H), EQ), True) for k in range(6)]
    v += [(HSYNC, (int(t + 9 * L + k * L), HS), True) for k in range(10)]
    line0, _, first, conf = LL.get_line0(_p(), v, float(L), _prev(25000, False), 0)
    assert first is True and conf == 100
    assert line0 == 25000.0            # median of 25420, 272.5 L - 263 L = 24149, 25000


def test_line0_local_only_caps_confidence_at_90():
    """No previous field: local (confidence 100 > 50) -> min(100, 90) = 90
    (core.py:2224-2226), vblank_next still reported."""
    assert LL.get_line0(_p(), _two_fields(), float(L), None, 0) == (22878, 691430, True, 90)


@pytest.mark.parametrize("conf_prev,want", [(100, 90), (60, 50), (15, 10)])
def test_line0_from_prev_confidence(conf_prev, want):
    """No vblank in the buffer: line 0 from the previous field with
    confidence max(conf_prev - 10, 10) (core.py:2227-2231, where
    np.max(conf_prev - 10, 0) reduces over axis 0 of a scalar: the value
    itself).  Mistake caught: max(conf_prev - 10, 0) without the floor of
    10 (5 for conf_prev 15)."""
    v = [(HSYNC, (k * L, HS), True) for k in range(300)]
    r = LL.get_line0(_p(), v, float(L), _prev(22880, False, conf=conf_prev), 0)
    assert r == (22880.0, None, True, want)


def test_line0_half_confidence_local_needs_prev():
    """A fallback vblank (confidence 50) is not enough alone: 'conf_local >
    50' fails, so with no previous field and no next vblank getLine0 finds
    nothing (core.py:2224-2236); with a previous field its line 0 wins.
    Mistake caught: '>= 50'."""
    v = _field(10, 0, spacing=[0, 10] * 9, tail=300)
    assert LL.get_line0(_p(), v, float(L), None, 0) == (None, None, None, 100)
    r = LL.get_line0(_p(), v, float(L), _prev(22900, True, conf=80), 0)
    assert r == (22900.0, None, False, 70)


def test_line0_limit_after_skip():
    """A previous field whose skip_check scored >= 50 limits the local
    search to the first 100 pulses (core.py:2166-2172): a vblank after 120
    hsyncs is then not seen and the previous field's line 0 is used."""
    v = _field(120, 0, tail=300)
    near = LL.get_line0(_p(), v, float(L), _prev(300000, True, conf=100, skip=0), 0)
    far = LL.get_line0(_p(), v, float(L), _prev(300000, True, conf=100, skip=50), 0)
    assert near == (119 * L, None, True, 90)
    assert far == (300000.0, None, False, 90)


def test_line0_negative_next_sets_confidence_10():
    """Field B only 110 lines after A: vblank_next = 118 L = 299956, line 0
    from it 299956 - 263 L < 0 -> sync_confidence 10 (core.py:2190-2191),
    which also caps the local answer."""
    r = LL.get_line0(_p(), _two_fields(b_shift=0, gap_lines=109), float(L), None, 0)
    assert r == (22878, 299956, True, 10)


def _native(pulses, prev, startloc):
    """ld_locate_field on the raw pulse list (host code of libldgpu)."""
    orf = O.OracleRF("NTSC", decode_analog_audio=False)
    rf = types.SimpleNamespace(lib=_lib.load(), system="NTSC", SysParams=orf.SysParams,
                               DecoderParams=orf.DecoderParams, freq=orf.freq, linelen=orf.linelen,
                               blocklen=32768, blockcut=1024, block_stride=31712, hsync_tolerance=0.4)
    s = np.array([q[1][0] for q in pulses], dtype=np.int64)
    n = np.array([q[1][1] for q in pulses], dtype=np.int64)
    cprev = None
    if prev is not None:
        cprev = _lib.PrevField(valid=1, is_first_field=int(prev.isFirstField), linecount=prev.linecount,
                               sync_confidence=prev.sync_confidence, startloc=prev.startloc,
                               ll_linecount=prev.linelocs[prev.linecount], skip_score=prev.skip_score,
                               fields_written=3)
    c, _, _ = locate.locate_one(rf, s, n, cprev, startloc, 3)
    return c


@pytest.mark.parametrize("case", ["median", "local", "negative_next"])
def test_native_walk_reaches_the_same_line0(case):
    """The same pulse trains as raw (start, length) lists through the
    library's classification and getLine0 (ld_locate_field)."""
    if case == "median":
        c = _native(_two_fields(), _prev(122880, True), 100000)
        assert (c.line0loc, c.vblank_next, c.is_first_field, c.sync_confidence) == (22880.0, 691430, 1, 100)
    elif case == "local":
        c = _native(_two_fields(), None, 0)
        assert (c.line0loc, c.vblank_next, c.is_first_field, c.sync_confidence) == (22878.0, 691430, 1, 90)
    else:
        c = _native(_two_fields(b_shift=0, gap_lines=109), None, 0)
        assert (c.line0loc, c.vblank_next, c.is_first_field, c.sync_confidence) == (22878.0, 299956, 1, 10)


# ------------------------------------------------------------ bursts
def test_clb_findbursts_median_over_all_sixteen_slots():
    """utils.py:1339-1368 on burstarea [1, 1, -1, -1, 1, 1, -1, -1, 1 ...]
    (threshold 0.5): crossings at 1.5, 3.5, 5.5, 7.5 (rising at 3.5, 7.5),
    then no crossing in 17 samples -> stop, 4 found; the other 12 slots
    stay 0.  bstart 10, s_rem 0.25, zcburstdiv 4: cycles 2.8125, 3.3125,
    3.8125, 4.3125 and 2.4375 for the 12 empty slots; residuals round(c) - c
    = 0.1875, -0.3125, 0.1875, -0.3125 and -0.4375 x 12 -> median of all 16
    = -0.4375 (of the 4 found only: -0.0625).  Rising count over 16 slots:
    isrising xor odd(round) = [T, F, F, T, F x 12] -> 2.  Second pass from
    -0.4375: residuals -0.375, 0.125, -0.375, 0.125, 0 x 12 -> median 0,
    phase stays; rounds 2, 3, 3, 4, 2 -> xor [F, F, T, T, ...] -> 2."""
    area = np.array([1, 1, -1, -1, 1, 1, -1, -1] + [1] * 12, dtype=np.float32)
    n, pa, rc = LN.clb_findbursts(area, 0.5, 10, 0.25, 4.0, 0.0)
    assert (n, rc) == (4, 2) and pa == pytest.approx(-0.4375, abs=1e-12)
    n, pa2, rc = LN.clb_findbursts(area, 0.5, 10, 0.25, 4.0, pa)
    assert (n, rc) == (4, 2) and pa2 == pytest.approx(-0.4375, abs=1e-12)


def test_burst_outlier_window_and_last_valid_carry(monkeypatch):
    """refine_linelocs_burst's application step (core.py:3416-3440), with
    compute_burst_offsets' per-line values given (adjs_new = a * fsc / 40 so
    that the sample adjustments are a; evenly spaced lines: lfreq = 40).
    Base a = 1.0; line 9: 1.4; line 10: -1.2 (median distance 2.2: outlier
    -> takes line 9's 1.4, the last applied value, not the median); line 20:
    2.99 (distance 1.99: applied); line 25: 3.01 (distance 2.01: outlier ->
    line 24's 1.0); line 30 missing (flagged bad, not in adjs -> carries
    line 29's 1.0); line 0: -2.0 (outlier; nothing applied yet -> the
    median 1.0).  Field phase ID of (first field, field14) = (True, False)
    -> 3.  Mistakes caught: outliers given the median, |a| <= 2 instead of
    the distance to the median (line 20 dropped, line 10 applied)."""
    p = LN.LineRefineParams(O.OracleRF("NTSC", decode_analog_audio=False))
    a = {l: 1.0 for l in range(266)}
    a[9], a[10], a[20], a[25], a[0] = 1.4, -1.2, 2.99, 3.01, -2.0
    del a[30]
    scale = p.fsc_mhz / 40.0
    monkeypatch.setattr(LN, "compute_burst_offsets",
                        lambda *args: (False, {l: v * scale for l, v in a.items()}, 0.0))
    ll2 = 5000.0 + np.arange(300, dtype=np.float64) * L
    bad = np.zeros(300, dtype=bool)
    ll3, bad3, fpid, _ = LN.refine_linelocs_burst(p, None, None, ll2, bad, 263, True, 0.0)
    d = ll3[:266] - ll2[:266]
    want = np.full(266, 1.0)
    want[9], want[10], want[20], want[25], want[0], want[30] = 1.4, 1.4, 2.99, 1.0, 1.0, 1.0
    np.testing.assert_allclose(d, want, atol=1e-9)
    assert bad3[30] and not bad3[29] and fpid == 3


# ------------------------------------------------------------ PAL pilot
def test_pilot_circular_mean_across_the_wrap():
    """refine_linelocs_pilot (core.py:3054-3093), lines every 2560 samples at
    ...+0.5 (lsoffset 0.5), plen = 40 / 3.75 / 2 = 16/3.  Each line's 241
    samples are 0.25 except a peak 1.0 at pk, 0.5, -0.5 after it: calczc
    from the peak -> pk + 1.5, zc = (pk + 1) * 3 / 16.  Lines 0-160 pk 4
    (zc 0.9375), 161-321 pk 10 (2.0625), 322 pk 15 (3.0): fractions 0.9375,
    0.0625, 0 -> circular mean 0 (the arithmetic mean would be 0.4985);
    phase_distance wraps 0.9375 to -0.0625 -> shifts -1/3, +1/3, 0 samples.
    Mistakes caught: an arithmetic mean (+-2.34 samples), no wrap (+5)."""
    p = LN.LineRefineParams(O.OracleRF("PAL", decode_analog_audio=False))
    n = 323
    ll = 1000.5 + np.arange(n + 3, dtype=np.float64) * 2560
    pk = np.array([4] * 161 + [10] * 161 + [15])
    buf = np.full(int(ll[-1]) + 4000, 0.25, dtype=np.float32)
    for l in range(n):
        b = int(ll[l]) + pk[l]
        buf[b:b + 3] = (1.0, 0.5, -0.5)
    out = LN.refine_linelocs_pilot(p, buf, ll)
    shift = np.array([-1.0] * 161 + [1.0] * 161 + [0.0]) / 3.0
    np.testing.assert_allclose(out[:n] - ll[:n], shift, atol=1e-9)


# ------------------------------------------------------------ computedelays
@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_computedelays_against_a_linear_model(system):
    """computedelays (core.py:952-1046) measures the demodulated fake line's
    sync and white edges.  Independent route: the same edges through the
    linear chain Fvideo_lpf * Femp * FVideo (FM modulation and demodulation
    taken as identity), plus the RF filter's group delay at the carrier of
    each level (RFVideo's phase slope) and the discriminator's half sample,
    agree within 0.75 samples; sync (lower carrier) is later than white by
    about the RF group-delay difference.  Mistakes caught: FVideo05 in place
    of FVideo (+28 samples), the emphasis left out (+2.7 / +4.3), the
    start offset not subtracted."""
    import scipy.fft as sfft
    o = O.OracleRF(system, decode_analog_audio=False)
    F, N = o.Filters, o.blocklen
    vs = o.DecoderParams["vsync_ire"]
    fake = np.full(N, o.iretohz(0))
    fake[1500:1500 + int(4.7 * o.freq)] = o.iretohz(vs)
    fake[3000:3500] = o.iretohz(100)
    lin = sfft.ifft(sfft.fft(fake) * F["Fvideo_lpf"] * F["Femp"] * F["FVideo"]).real
    ph = np.unwrap(np.angle(F["RFVideo"][:N // 2]))

    def gd(hz):
        k = int(round(hz / o.freq_hz * N))
        return -(ph[k + 1] - ph[k - 1]) / (2 * 2 * np.pi / N)

    sync = O.calczc(lin, 1500, o.iretohz(vs / 2), count=512) - 1500 + gd(o.iretohz(vs / 2)) + 0.5
    white = O.calczc(lin, 3000, o.iretohz(50), count=512) - 3000 + gd(o.iretohz(50)) + 0.5
    assert abs(o.delays["video_sync"] - sync) < 0.75, (o.delays, sync)
    assert abs(o.delays["video_white"] - white) < 0.75, (o.delays, white)
    assert 0 < o.delays["video_sync"] - o.delays["video_white"] < 0.1
    assert o.delays["video_rot"] == int(np.round(o.delays["video_rot"])) and 1 <= o.delays["video_rot"] <= 6
