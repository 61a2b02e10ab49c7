"""GPU parity of line location: the pulse edges, sync levels, range means
and skip-check medians against numpy / the oracle bit for bit, and the whole
located decode (GPU pulses + host walk + verification) against the oracle's
sequential walk over the same demodulated buffers.
"""

import ctypes

import numpy as np
import pytest

from lddecode_amd import _lib, capture, locate, synth
from oracle import dsp as O
from oracle import linelocs as LL

pytestmark = pytest.mark.gpu


def _edges_np(x, high):
    below = x.astype(np.float64) <= high
    prev = np.concatenate([[False], below[:-1]])
    return np.flatnonzero(below != prev).astype(np.int64)


def _means_np(x, edges, min_len, trim):
    s, e = edges[0:len(edges) - 1:2], edges[1::2]
    s = s[:len(e)]
    out = np.full(len(s), np.nan, np.float32)
    for k in range(len(s)):
        n = e[k] - s[k]
        if n > min_len and n <= 5000:
            a, b = int(s[k] + trim), int(s[k] + n - trim)
            if b > a:
                out[k] = LL.np_mean_f32(x[a:b])
    return out


def _pulses_gpu(rf, x, high, cap=None):
    import torch
    xt = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()
    pf = locate.PulseFinder(rf, len(x), cap_edges=cap)
    pf.launch(xt, len(x), high=high)
    e, m = pf.result()
    return e.copy(), m.copy()


_CAP = {}


def _demod(system, nfields=5, seed=3):
    key = (system, nfields, seed)
    if key not in _CAP:
        orf = O.OracleRF(system, decode_analog_audio=False)
        fl = 667_334 if system == "NTSC" else 800_000
        cap = synth.make_capture(system, nfields * fl, seed=seed, snr_db=40.0, jitter_ns=50, audio=False)
        nbt = (len(cap.data) - 32768) // 31712 + 1
        ref, _ = O.demod_span(orf, cap.data.astype(np.float64), 0, nbt * 31712 - 1)
        _CAP[key] = (cap, {k: v.astype(np.float32) for k, v in ref.items()})
    return _CAP[key]


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_pulse_edges_and_levels_exact(gpu_rf, system):
    rf = gpu_rf(system, audio=False)
    _, v = _demod(system)
    d05 = v["demod_05"]
    high = locate.pulse_threshold(rf)
    e, m = _pulses_gpu(rf, d05, high)
    want = _edges_np(d05, high)
    np.testing.assert_array_equal(e, want)
    wm = _means_np(d05, want, rf.freq * 10.0, rf.freq)
    np.testing.assert_array_equal(np.isnan(m), np.isnan(wm))
    np.testing.assert_array_equal(m[~np.isnan(m)], wm[~np.isnan(wm)])
    assert (~np.isnan(m)).sum() >= 3 * 6      # every field's vertical sync pulses


def test_pulse_edges_adversarial(gpu_rf):
    """Thresholds met exactly, dense noise edges, a buffer starting inside a
    pulse, lengths that are not multiples of a tile, and a capacity too small
    for the first pass (the finder grows and searches again)."""
    rf = gpu_rf("NTSC", audio=False)
    rng = np.random.default_rng(5)
    for n in (5, 17, 4095, 4096, 4097, 100_003):
        x = rng.normal(size=n).astype(np.float32)
        x[::7] = np.float32(0.25)
        x[0] = -1.0
        for high in (0.25, 0.0, float(np.float32(0.25)) + 1e-12):
            e, _ = _pulses_gpu(rf, x, high, cap=64)
            np.testing.assert_array_equal(e, _edges_np(x, high))


def _pack_bits(below):
    """Paired in-pulse words, 2 * ceil(n / 64) int32 of them: bit j < 16 of
    word w = below[32 w + 2 j], bit 16 + j = below[32 w + 2 j + 1]."""
    n = len(below)
    nw = 2 * -(-n // 64)
    b = np.zeros(nw * 32, np.uint8)
    b[:n] = below
    b = b.reshape(-1, 32)
    b = np.concatenate([b[:, 0::2], b[:, 1::2]], axis=1)     # bit order within the word
    return np.packbits(b[:, ::-1], axis=1).view(">u4").astype(np.uint32).view(np.int32).ravel()


def _pulses_bits_gpu(rf, words, n):
    import torch
    wt = torch.from_numpy(np.ascontiguousarray(words)).cuda()
    cap = n + 2
    edges = torch.empty(cap, dtype=torch.int64, device="cuda")
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    ws = torch.empty(rf.lib.ld_pulse_bits_workspace_bytes(rf.handle, n), dtype=torch.uint8, device="cuda")
    _lib.check(rf.lib.ld_find_pulses_bits(rf.handle, wt.data_ptr(), n, edges.data_ptr(), cap, count.data_ptr(),
                                          ws.data_ptr(), ws.numel(), None))
    c = int(count.cpu()[0])
    return edges[:c].cpu().numpy()


def test_pulse_bits_edges(gpu_rf):
    """ld_find_pulses_bits against the same edges from the samples: word and
    tile boundaries, a buffer starting inside a pulse, a run reaching the
    end, lengths around 32 / 64 / a tile; bits past n are ignored."""
    rf = gpu_rf("NTSC", audio=False)
    rng = np.random.default_rng(11)
    for n in (1, 5, 31, 32, 33, 63, 64, 65, 16383, 16384, 16385, 100_003):
        for density in (0.05, 0.5):
            below = rng.random(n) < density
            below[0] = True
            if n > 40:
                below[31:33] = (False, True)      # an edge on a word boundary
                below[-5:] = True                 # a run reaching the end
            words = _pack_bits(below)
            junk = np.zeros(len(words) * 32, bool)
            junk[n:] = rng.random(len(junk) - n) < 0.5
            words |= _pack_bits(junk)[:len(words)]
            x = np.where(below, -1.0, 1.0).astype(np.float32)
            np.testing.assert_array_equal(_pulses_bits_gpu(rf, words, n), _edges_np(x, 0.0), err_msg=str(n))


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_demod_in_pulse_bits(gpu_rf, system):
    """The demod's fused in-pulse bits equal demod_05 <= threshold (in
    double) for every stitched sample, and the bit search gives the sample
    search's edges."""
    import torch
    rf = gpu_rf(system, audio=False)
    fl = 667_334 if system == "NTSC" else 800_000
    cap = synth.make_capture(system, 3 * fl, seed=4, snr_db=40.0, jitter_ns=50, audio=False)
    dec = capture.CaptureDecoder(rf, torch.from_numpy(cap.data).cuda(), locate=True, max_fields=4)
    dec.demod()
    torch.cuda.synchronize()
    d05 = dec.buffers["demod_05"].cpu().numpy()
    hi = locate.pulse_threshold(rf)
    words = dec.buffers["demod_05_below"].cpu().numpy()
    nw = -(-len(d05) // 32)                   # words the demod writes (the rest is padding)
    np.testing.assert_array_equal(words[:nw], _pack_bits(d05.astype(np.float64) <= hi)[:nw])
    e_bits = _pulses_bits_gpu(rf, words, len(d05))
    np.testing.assert_array_equal(e_bits, _edges_np(d05, hi))
    e_float, _ = _pulses_gpu(rf, d05, hi)
    np.testing.assert_array_equal(e_bits, e_float)


def test_pulse_search_many_buffers(gpu_rf):
    """ld_pulse_search over an NTSC and a PAL buffer at once (one launch per
    step for both): each buffer's pinned host lists equal the numpy edges and
    np.mean levels of its own samples, and its count; a third job whose
    capacity is below its edge count exports exactly the first cap edges."""
    import torch
    bufs, finders, searches = [], [], []
    for system, seed in (("NTSC", 4), ("PAL", 5), ("NTSC", 6)):
        rf = gpu_rf(system, audio=False)
        fl = 667_334 if system == "NTSC" else 800_000
        cap = synth.make_capture(system, 2 * fl, seed=seed, snr_db=40.0, jitter_ns=50, audio=False)
        dec = capture.CaptureDecoder(rf, torch.from_numpy(cap.data).cuda(), locate=True, max_fields=4)
        dec.demod()
        d05 = dec.buffers["demod_05"]
        n = d05.numel()
        pf = locate.PulseFinder(rf, n, cap_edges=1000 if len(finders) == 2 else None)
        bufs.append((rf, d05.cpu().numpy()))
        finders.append(pf)
        searches.append((pf, d05, n, dec.buffers["demod_05_below"]))
    ws = torch.empty(locate.pulse_search_workspace(searches), dtype=torch.uint8, device="cuda")
    locate.pulse_search(searches, ws, torch.cuda.current_stream())
    torch.cuda.synchronize()
    for k, ((rf, x), pf) in enumerate(zip(bufs, finders)):
        want = _edges_np(x, locate.pulse_threshold(rf))
        cnt = int(pf.h_count[0])
        assert cnt == len(want), f"job {k}"
        ne = min(cnt, pf.cap)
        np.testing.assert_array_equal(pf.h_edges.numpy()[:ne], want[:ne], err_msg=f"job {k}")
        lvl = _means_np(x, want[:ne], 10.0 * ((rf.freq / rf.linelen) * rf.linelen), rf.freq)
        np.testing.assert_array_equal(pf.h_means.numpy()[:ne // 2], lvl[:ne // 2], err_msg=f"job {k}")
    assert finders[2].cap < int(finders[2].h_count[0])


def test_pulse_means_long_pulses(gpu_rf):
    """Pulse levels whose trimmed bodies straddle the 1280 samples the mean
    kernels stage in LDS (longer ones are read in place) up to the 5000-sample
    limit (findpulses' longest): numpy's float32 pairwise mean of each body,
    bit for bit, from the single-buffer finder and from ld_pulse_search."""
    import torch
    rf = gpu_rf("NTSC", audio=False)
    trim = int(rf.freq)
    rng = np.random.default_rng(11)
    lengths = [500, 1000, 1279 + 2 * trim, 1280 + 2 * trim, 1281 + 2 * trim, 1409 + 2 * trim, 2000, 4096,
               4999, 5000, 5001, 7000]
    gap = 3000
    n = gap + sum(L + gap for L in lengths) + 123
    x = (1.0 + 0.05 * rng.standard_normal(n)).astype(np.float32)
    pos = gap
    for L in lengths:
        x[pos:pos + L] = (0.1 + 0.05 * rng.standard_normal(L)).astype(np.float32)
        pos += L + gap
    high = 0.5
    want = _edges_np(x, high)
    assert len(want) == 2 * len(lengths)
    lvl = _means_np(x, want, 10.0 * ((rf.freq / rf.linelen) * rf.linelen), rf.freq)
    assert np.isnan(lvl[-2:]).all() and not np.isnan(lvl[:-2]).any()    # 5001 and 7000: past the limit
    e, m = _pulses_gpu(rf, x, high)
    np.testing.assert_array_equal(e, want)
    np.testing.assert_array_equal(np.isnan(m), np.isnan(lvl))
    np.testing.assert_array_equal(m[~np.isnan(m)], lvl[~np.isnan(lvl)])
    # the fused multi-buffer search, from the in-pulse bits of the same samples
    xt = torch.from_numpy(x).cuda()
    bits = torch.from_numpy(_pack_bits(x.astype(np.float64) <= high)).cuda()
    pf = locate.PulseFinder(rf, n)
    searches = [(pf, xt, n, bits)]
    ws = torch.empty(locate.pulse_search_workspace(searches), dtype=torch.uint8, device="cuda")
    locate.pulse_search(searches, ws, torch.cuda.current_stream())
    torch.cuda.synchronize()
    assert int(pf.h_count[0]) == len(want)
    np.testing.assert_array_equal(pf.h_edges.numpy()[:len(want)], want)
    got = pf.h_means.numpy()[:len(lengths)]
    np.testing.assert_array_equal(np.isnan(got), np.isnan(lvl))
    np.testing.assert_array_equal(got[~np.isnan(got)], lvl[~np.isnan(lvl)])


def test_range_means_exact(gpu_rf):
    import torch
    rf = gpu_rf("NTSC", audio=False)
    rng = np.random.default_rng(9)
    x = (rng.normal(size=200_000) * 1e5 + 7.6e6).astype(np.float32)
    b = rng.integers(0, 190_000, size=40)
    ln = rng.integers(0, 20000, size=40)
    ln[:5] = (0, 1, 129, 8192, 16385)
    end = b + ln
    xt = torch.from_numpy(x).cuda()
    out = torch.empty(40, dtype=torch.float32, device="cuda")
    bt, et = torch.from_numpy(b.astype(np.int64)).cuda(), torch.from_numpy(end.astype(np.int64)).cuda()
    _lib.check(rf.lib.ld_range_means(rf.handle, xt.data_ptr(), len(x), bt.data_ptr(), et.data_ptr(), 40,
                                     out.data_ptr(), None))
    got = out.cpu().numpy()
    for k in range(40):
        want = LL.np_mean_f32(x[b[k]:end[k]]) if end[k] > b[k] else np.float32(np.nan)
        assert (np.isnan(got[k]) and np.isnan(want)) or got[k] == want, (k, got[k], want)


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_located_decode_matches_oracle_walk(gpu_rf, system):
    """CaptureDecoder(locate=True) against the oracle's sequential walk over
    the GPU's own demod buffers, each oracle field finished with the GPU's
    final line locations for it: the same fields, parity, confidence, line 0,
    next-field offsets and linelocs1, bit for bit; and the located lines sit
    at the synthetic truth."""
    import torch
    rf = gpu_rf(system, audio=False)
    cap, _ = _demod(system)
    dec = capture.CaptureDecoder(rf, torch.from_numpy(cap.data).cuda(), locate=True, max_fields=8)
    dec.step()
    torch.cuda.synchronize()
    locs, ll1, bad1 = dec.walk_state[:3]
    d05 = dec.buffers["demod_05"].cpu().numpy()
    dm = dec.buffers["demod"].cpu().numpy()
    final = dec.linelocs.cpu().numpy()
    L = locate.proclines(rf)
    valid = [k for k, c in enumerate(locs) if c.status == _lib.LD_LOC_OK]
    assert len(valid) >= 3
    fin = {k: final[j * L:(j + 1) * L] for j, k in enumerate(valid)}
    p = LL.LocParams.from_rf(O.OracleRF(system, decode_analog_audio=False))
    ora = LL.walk_fields(p, d05, len(locs), 0, lambda k, res, org: fin[k], demod=dm)
    assert len(ora) == len(locs)
    for k, (c, o) in enumerate(zip(locs, ora)):
        assert c.startloc == o["startloc"]
        assert (c.status == _lib.LD_LOC_OK) == o["valid"]
        if o["valid"]:
            r = o["res"]
            assert bool(c.is_first_field) == r.isFirstField and c.sync_confidence == r.sync_confidence
            assert c.line0loc == r.line0loc and c.nextfieldoffset == r.nextfieldoffset
            np.testing.assert_array_equal(ll1[k], r.linelocs1)
    # truth: even-numbered raster fields are first fields; line starts trail
    # the true starts by demod_05's edge delay
    hpf = 525 if system == "NTSC" else 625
    hs = cap.halfline_starts
    for j, k in enumerate(valid):
        c = locs[k]
        raw = ll1[k] + c.startloc + 1024
        f = int(round(raw[0] / (hpf * rf.SysParams["line_period"] * 20)))
        idx = synth.field_line0(system, f) + 2 * np.arange(L)
        d = raw - hs[idx]
        assert 9.5 < d.min() and d.max() < 14.5
        assert bool(c.is_first_field) == (f % 2 == 0)


def test_pipelined_located_equals_sequential(gpu_rf):
    """PipelinedDecoder(locate=True) over four passes gives the sequential
    CaptureDecoder's fields and output codes (the last three passes each
    still hold their output: three buffer sets)."""
    import torch
    rf = gpu_rf("NTSC", audio=False)
    cap, _ = _demod("NTSC")
    x = torch.from_numpy(cap.data).cuda()
    seq = capture.CaptureDecoder(rf, x, locate=True, max_fields=8)
    want = seq.step().clone()
    torch.cuda.synchronize()
    nf = len(seq.fields)
    pipe = capture.PipelinedDecoder(rf, x, locate=True, max_fields=8)
    outs = [pipe.step() for _ in range(4)]
    assert outs[0] is None and outs[1] is None
    outs = outs[2:] + pipe.join()
    torch.cuda.synchronize()
    assert len(outs) == 4
    for d in outs[-3:]:
        assert [f["demod_offset"] for f in d.fields] == [f["demod_offset"] for f in seq.fields]
        n = nf * seq.linesout * seq.W
        np.testing.assert_array_equal(d.out_u16[:n].cpu().numpy(), want[:n].cpu().numpy())


def test_pipelined_alternating_mtf_matches_sequential(gpu_rf):
    """Consecutive pipelined passes at alternating MTF levels (the reference
    changes the level per field, core.py:3841-3852): each pass's RFVideo
    table is uploaded while the previous passes may still run, and every
    returned set must equal a sequential decode at its own level, checked
    before the next step() call reuses a set."""
    import torch
    rf = gpu_rf("NTSC", audio=False)
    cap, _ = _demod("NTSC")
    x = torch.from_numpy(cap.data).cuda()
    levels = (0.0, 1.0)
    want, plan = {}, {}
    for lv in levels:
        seq = capture.CaptureDecoder(rf, x, locate=True, max_fields=8)
        want[lv] = seq.step(mtf_level=lv).clone()
        torch.cuda.synchronize()
        plan[lv] = [f["demod_offset"] for f in seq.fields]
    n = len(plan[0.0]) * seq.linesout * seq.W
    assert not torch.equal(want[0.0][:n], want[1.0][:n]), "the MTF level must change the output"
    pipe = capture.PipelinedDecoder(rf, x, locate=True, max_fields=8)
    checked = 0
    for k in range(7):
        d = pipe.step(mtf_level=levels[k % 2])
        if d is None:
            continue
        torch.cuda.synchronize()
        lv = levels[(k - 2) % 2]          # step k returns pass k - 2
        assert [f["demod_offset"] for f in d.fields] == plan[lv]
        np.testing.assert_array_equal(d.out_u16[:n].cpu().numpy(), want[lv][:n].cpu().numpy())
        checked += 1
    rest = pipe.join()
    torch.cuda.synchronize()
    for j, d in enumerate(rest):
        lv = levels[(7 - 2 + j) % 2]
        np.testing.assert_array_equal(d.out_u16[:n].cpu().numpy(), want[lv][:n].cpu().numpy())
        checked += 1
    assert checked == 7


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_field_compute_linelocs(gpu_rf, system):
    """Field.compute_linelocs on one field buffer (the reference's per-field
    call) against the oracle's compute_linelocs on the same buffer, without
    and with a previous field (its final line locations and buffers)."""
    import torch
    from lddecode_amd import field as fieldmod
    rf = gpu_rf(system, audio=False)
    _, v = _demod(system)
    p = LL.LocParams.from_rf(O.OracleRF(system, decode_analog_audio=False))
    ora = LL.walk_fields(p, v["demod_05"], 4, 0, lambda k, res, org: res.linelocs1 + 2.5, demod=v["demod"])
    valid = [o for o in ora if o["valid"]]
    a, b = valid[0], valid[1]
    S = 31712

    def buf(o):
        sl = slice(o["startloc"], o["startloc"] + o["nblocks"] * S)
        return {"video": {k: torch.from_numpy(np.ascontiguousarray(v[k][sl])).cuda() for k in ("demod", "demod_05")}}

    fa = fieldmod.Field(rf, buf(a))
    ll, bad, off = fa.compute_linelocs(None, fields_written=1, startloc=a["startloc"])
    ra = LL.compute_linelocs(p, v["demod_05"][a["startloc"]:a["startloc"] + a["nblocks"] * S], None, a["startloc"], 1)
    np.testing.assert_array_equal(ll, ra.linelocs1)
    np.testing.assert_array_equal(bad, ra.linebad)
    assert off == ra.nextfieldoffset and fa.isFirstField == ra.isFirstField
    fa.linelocs = ll + 2.5                       # stands in for its aligned positions
    fb = fieldmod.Field(rf, buf(b))
    ll, bad, off = fb.compute_linelocs(fa, fields_written=1, startloc=b["startloc"])
    np.testing.assert_array_equal(ll, b["res"].linelocs1)
    assert fb.line0loc == b["res"].line0loc and fb.sync_confidence == b["res"].sync_confidence
    assert off == b["res"].nextfieldoffset


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_field_facts_match_oracle(gpu_rf, system):
    """ld_field_facts_measure (skip_check's eight line medians and the final
    location of line linecount) against the oracle's field_facts on the same
    buffers, bit for bit: first and second fields, repeated samples (ties in
    the order statistics), and fields too short for some of the lines
    (NaN medians, NaN linecount location)."""
    import torch
    rf = gpu_rf(system, audio=False)
    p = LL.LocParams.from_rf(rf)
    length = p.usectoinpx(p.line_period)
    oc = p.outlinecount
    lo_of = (lambda first: (2 if first else 3)) if system == "PAL" else (lambda first: 0)
    lc_of = (lambda first: 312 if first else 313) if system == "PAL" else (lambda first: 263 if first else 262)
    rng = np.random.default_rng(21)
    ire0, hz_ire = rf.DecoderParams["ire0"], rf.DecoderParams["hz_ire"]
    specs = [(p.proclines, 1), (p.proclines, 0), (oc + 4, 1), (oc + 12, 0)]
    demods, lls, arr = [], [], (_lib.LineField * len(specs))()
    off = loff = 0
    for f, (nlines, first) in enumerate(specs):
        span = int((nlines + 3) * p.linelen)
        d = (ire0 + hz_ire * rng.uniform(-45.0, 105.0, span)).astype(np.float32)
        d[::7] = d[3]                                  # ties
        final = 100.0 + np.arange(nlines) * (p.linelen + 0.37) + rng.uniform(-3.0, 3.0, nlines)
        r = arr[f]
        r.demod_offset, r.demod_len, r.lineloc_offset, r.nlines = off, span, loff, nlines
        r.linecount, r.first_field, r.chain = 0, first, 0
        demods.append(d)
        lls.append(final)
        off += span
        loff += nlines
    dt = torch.from_numpy(np.concatenate(demods)).cuda()
    lt = torch.from_numpy(np.concatenate(lls)).cuda()
    n = len(specs)
    facts = torch.empty(n * locate.FACTS_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    ws = torch.empty(rf.lib.ld_field_facts_workspace_bytes(rf.handle, n), dtype=torch.uint8, device="cuda")
    _lib.check(rf.lib.ld_field_facts_measure(rf.handle, dt.data_ptr(), dt.numel(), arr, n, lt.data_ptr(), lt.numel(),
                                             1 if system == "PAL" else 0, oc, float(length), facts.data_ptr(),
                                             ws.data_ptr(), ws.numel(), None))
    got = facts.cpu().numpy().view(locate.FACTS_DTYPE)
    for f, (nlines, first) in enumerate(specs):
        d, final = demods[f], lls[f]
        want = []
        for j in range(8):
            line = oc + j + lo_of(first)
            b = final[line] if line < nlines else None
            want.append(LL.numba_median(d[int(b):int(b + length + 1)]) if b is not None else np.nan)
        lc = lc_of(first)
        want_llc = final[lc] if lc < nlines else np.nan
        if nlines == p.proclines:                     # the oracle's own function, where it applies
            llc_o, meds_o = LL.field_facts(p, d, final, bool(first))
            assert llc_o == want_llc and meds_o == want
        np.testing.assert_array_equal(got["medians"][f], np.array(want), err_msg=f"field {f}")
        np.testing.assert_array_equal(got["ll_linecount"][f], want_llc, err_msg=f"field {f}")
        assert got["known"][f] == 1
        if nlines == oc + 4:
            assert np.isnan(got["medians"][f][-1]) and not np.isnan(got["medians"][f][0])
