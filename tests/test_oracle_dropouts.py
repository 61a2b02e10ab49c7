"""The dropout-detection oracle (oracle/dropouts.py) pinned by closed forms,
and libldgpu's native error-list code (ld_dropout_list, host only) against
it.

Pins: numba's std restated two ways (scalar loops and in-order cumulative
sums) must agree bit for bit and sit at the float64 std; build_errlist and
dropout_errlist_to_tbc on hand-worked cases (merging within 20 samples, the
1.7x padding and its freq*12 cap, the -8/+4 widening of every run but the
last, errors before the first line ignored, multi-line splitting with
round-half-even ends); the sample mask on a demodulated synthetic field
with injected carrier losses (each loss is flagged, a clean field is not).
"""

import types

import numpy as np
import pytest

from lddecode_amd import _lib
from lddecode_amd import dropouts as DR
from oracle import dropouts as D


def test_std_forms_agree():
    rng = np.random.default_rng(1)
    for n, scale, off in ((1, 1.0, 0.0), (7, 3.0, 0.5), (5000, 3.0, 0.2), (20000, 1e4, 3e5)):
        x = (rng.normal(size=n) * scale + off).astype(np.float32)
        a, b = D.nb_std_f32_loop(x), D.nb_std_f32(x)
        assert a == b
        assert abs(a - float(np.std(x.astype(np.float64)))) <= 1e-6 * max(a, 1e-30) + 1e-12


def test_build_errlist_closed_forms():
    ll = np.array([50.0, 3000.0])
    # a run 100, 101, 105 (pads 1.7 then 8.5), then a lone error at 200
    got = D.build_errlist(np.array([100, 101, 105, 200]), ll, 0, 40)
    assert got == [(92, 112.5), (200, 200)]
    # a long run: the padding reaches its cap freq*12 = 480, so the run
    # ends at 580; 610 starts the last run, which gets no -8/+4
    e = np.concatenate([np.arange(100, 601, 10), np.arange(610, 701, 10)])
    got = D.build_errlist(e, ll, 0, 40)
    assert len(got) == 2
    assert got[0] == (92, 584.0)
    assert got[1][0] == 610 and abs(got[1][1] - (610 + 90 * 1.7)) < 1e-9
    # errors before linelocs[lineoffset] do not start runs
    assert D.build_errlist(np.array([10, 20, 100]), ll, 0, 40) == [(100, 100)]


def test_errlist_to_tbc_closed_forms():
    ll = 1000.0 + 2000.0 * np.arange(11)
    # one short dropout on line 1: start int(45.5) = 45, end round(91.0)
    assert D.errlist_to_tbc([(3100.0, 3200.0)], ll, 0, 8, 910) == [(2, 45, 91)]
    # end round(4800/2000*910) = 2184 > 910: split over three lines
    got = D.errlist_to_tbc([(5500.0, 9800.0)], ll, 0, 8, 910)
    assert got == [(3, 227, 910), (4, 0, 910), (5, 0, 364)]
    # a start exactly on a line start belongs to the line before (inrange is
    # inclusive at both ends): 1137.5 rounds half to even to 1138, split as
    # (1, 910, 910) + (2, 0, 228); likewise 682.5 -> 682 with 420 pixels
    assert D.errlist_to_tbc([(3000.0, 3500.0)], ll, 0, 8, 910) == [(1, 910, 910), (2, 0, 228)]
    assert D.errlist_to_tbc([(3000.0, 4250.0)], ll, 0, 8, 420) == [(1, 420, 420), (2, 0, 262)]
    # exact halves (2048-sample lines, 512 pixels): 2.5 -> 2, 3.5 -> 4
    l2 = 1024.0 + 2048.0 * np.arange(11)
    assert D.errlist_to_tbc([(3080.0, 3082.0)], l2, 0, 8, 512) == [(2, 2, 2)]
    assert D.errlist_to_tbc([(3080.0, 3086.0)], l2, 0, 8, 512) == [(2, 2, 4)]
    # a leading run before linelocs[lineoffset] is dropped; lineoffset 2
    got = D.errlist_to_tbc([(100.0, 150.0), (5100.0, 5200.0)], ll, 2, 5, 910)
    assert got == [(1, 45, 91)]
    lines, starts, ends = D.dropout_detect(np.isin(np.arange(12000), [5100, 5101, 5105]), ll, 2, 5, 910, 40)
    # the last (here only) run keeps its bounds: start int(100/2000*910)
    assert lines == [0] and starts == [45] and ends == [round(108.5 / 2000 * 910)]


def _host_rf(system="NTSC"):
    return types.SimpleNamespace(lib=_lib.load(), SysParams={"outlinelen": 910 if system == "NTSC" else 1135},
                                 freq=40.0)


def _oracle_lists(errmap, ll, lineoffset, linecount, W, freq):
    lines, starts, ends = [], [], []
    if len(errmap) > 0 and errmap[-1] > ll[lineoffset]:
        for r in D.errlist_to_tbc(D.build_errlist(errmap, ll, lineoffset, freq), ll, lineoffset, linecount, W):
            lines.append(r[0] - 1)
            starts.append(int(r[1]))
            ends.append(int(r[2]))
    return lines, starts, ends


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_native_error_list_matches_oracle(system):
    rf = _host_rf(system)
    W = rf.SysParams["outlinelen"]
    rng = np.random.default_rng(3 if system == "NTSC" else 4)
    nl, lo, lc = (273, 0, 263) if system == "NTSC" else (326, 2, 312)
    for trial in range(40):
        ll = 700.0 + np.cumsum(rng.normal(2541.0, 0.3, size=nl))
        n = int(ll[-1]) + 3000
        kind = trial % 5
        if kind == 0:
            err = np.zeros(0, np.int64)
        elif kind == 1:                     # only errors before the first line
            err = np.sort(rng.choice(int(ll[lo]), size=30, replace=False)).astype(np.int64)
        else:
            runs = []
            for _ in range(rng.integers(1, 60)):
                a = int(rng.integers(0, n - 8000))
                ln = int(rng.choice([1, 3, 25, 400, 2600, 7000]))
                step = int(rng.choice([1, 5, 19, 21]))
                runs.append(np.arange(a, a + ln, step))
            if kind == 3:                   # sparse noise errors
                runs.append(rng.choice(n, size=500, replace=False))
            err = np.unique(np.concatenate(runs)).astype(np.int64)
        want = _oracle_lists(err, ll, lo, lc, W, rf.freq)
        got = DR.error_list(rf, err, ll, lo, lc)
        assert got == want, (trial, kind)
    # more entries than the first output capacity
    ll = 700.0 + 2541.0 * np.arange(nl)
    err = np.concatenate([np.arange(int(x) + 100, int(x) + 104) for x in ll[lo:lo + lc]] +
                         [np.arange(int(x) + 900, int(x) + 903) for x in ll[lo:lo + lc]])
    err = np.sort(err).astype(np.int64)
    want = _oracle_lists(err, ll, lo, lc, W, rf.freq)
    assert len(want[0]) > 256
    assert DR.error_list(rf, err, ll, lo, lc) == want


def test_native_error_list_rejects_bad_arguments():
    rf = _host_rf()
    ll = np.arange(10, dtype=np.float64) * 2541
    with pytest.raises(_lib.LdGpuError):
        DR.error_list(rf, np.array([5], np.int64), ll, 0, 10)    # lineoffset + linecount + 1 > nlines


def test_mask_flags_injected_losses(oracle_rf):
    """The oracle's mask on the FP64 demodulation of an NTSC field with the
    carrier lost over three spans: each span (delayed by the demodulation)
    is flagged, nothing is flagged on the clean copy, and the grouped list
    puts each loss on its line."""
    from lddecode_amd import synth
    from oracle import dsp as O
    from oracle import linelocs as LL
    orf = oracle_rf("NTSC", audio=False)
    p = LL.LocParams.from_rf(orf)
    n = 4 * 667_334
    spans = [(800_000, 40), (1_000_000, 3000), (1_200_000, 7)]
    out = {}
    for tag, drops in (("clean", ()), ("lossy", spans)):
        cap = synth.make_capture("NTSC", n, seed=11, snr_db=40.0, jitter_ns=50, audio=False, dropouts=drops)
        nbt = (len(cap.data) - 32768) // 31712 + 1
        ref, _ = O.demod_span(orf, cap.data.astype(np.float64), 0, nbt * 31712 - 1)
        vv = {k: v.astype(np.float32) for k, v in ref.items()}
        out[tag] = (cap, vv, vv["rfhpf"])
    cap, v, _ = out["clean"]
    ora = LL.walk_fields(p, v["demod_05"], 3, 0, lambda k, res, org: res.linelocs1, demod=v["demod"])
    o = next(o for o in ora if o["valid"])
    res = o["res"]
    sl = slice(o["startloc"], o["startloc"] + o["nblocks"] * 31712)
    results = {}
    for tag in ("clean", "lossy"):
        _, vv, rr = out[tag]
        iserr = D.dropout_detect_demod(p, rr[sl], vv["demod"][sl], vv["demod_raw"][sl], vv["demod_05"][sl],
                                       res.linelocs1, res.LT["hsync"][1], res.isFirstField, 0, res.linecount,
                                       orf.delays["video_rot"])
        results[tag] = iserr
    assert not results["clean"].any()
    pos = np.flatnonzero(results["lossy"]) + o["startloc"] + 1024
    for a0, ln in spans:
        hit = pos[(pos >= a0 - 100) & (pos <= a0 + ln + 200)]
        assert len(hit) > 0, (a0, ln)
    assert len(pos) < 20_000
    lines, starts, ends = D.dropout_detect(results["lossy"], res.linelocs1, 0, res.linecount, 910, 40.0)
    for a0, ln in spans:
        l = int(np.searchsorted(res.linelocs1, a0 - o["startloc"] - 1024, side="right")) - 1
        assert any(abs(x - l) <= 1 for x in lines), (a0, l, lines)
    assert all(0 <= a <= b <= 910 for a, b in zip(starts, ends))
