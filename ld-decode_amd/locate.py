"""Line location: the GPU finds sync pulses, their levels and the skip-check
medians; libldgpu's host code runs the pulse state machine, the
vertical-blanking analysis, line 0 and the field walk.

Mirrors Field.getpulses / compute_linelocs (lddecode/core.py:2242-2481) and
the offsets of LDdecode.decodefield / readfield (core.py:4018-4191).

A field's line 0 can depend on the previous field's *final* line locations
and on its skip_check (core.py:2163-2207), which only exist after that
field's hsync and burst / pilot alignment.  The batched walk therefore runs
ahead on provisional values (the previous field's linelocs1 at its linecount
plus an offset, skip score 100), the GPU aligns every field, and the walk is
repeated with the measured values (:class:`FieldFacts`); fields whose result
changes are located and aligned again.  The result is the reference's
sequential answer.
"""

import ctypes
import types

import numpy as np

from . import _lib

NTSC_LINES = 273   # proclines: outlinecount + 10
PAL_LINES = 326    # outlinecount + 10 + 3


def proclines(rf):
    return rf.SysParams["frame_lines"] // 2 + 1 + 10 + (3 if rf.system == "PAL" else 0)


def readlen(rf):
    """LDdecode.readlen (core.py:3625-3631)."""
    if rf.system == "PAL":
        return rf.linelen * 400
    return ((rf.linelen * 350) // 16384) * 16384


FIELDLOC_DTYPE = np.dtype(_lib.FieldLoc)
FACTS_DTYPE = np.dtype(_lib.FieldFacts)


def locate_params(rf, prov_delta=0.0):
    sp, dp = rf.SysParams, rf.DecoderParams
    p = _lib.LocateParams()
    p.freq = float(rf.freq)
    p.linelen = float(rf.linelen)
    p.ire0, p.hz_ire, p.vsync_ire = float(dp["ire0"]), float(dp["hz_ire"]), float(dp["vsync_ire"])
    p.hsync_us, p.eq_us, p.vsync_us = float(sp["hsyncPulseUS"]), float(sp["eqPulseUS"]), float(sp["vsyncPulseUS"])
    p.line_period_us = float(sp["line_period"])
    p.first_field_h[0], p.first_field_h[1] = (float(v) for v in sp["firstFieldH"])
    p.hsync_tolerance = float(getattr(rf, "hsync_tolerance", 0.4))
    p.prov_delta = float(prov_delta)
    p.pal = 1 if rf.system == "PAL" else 0
    p.num_pulses = int(sp["numPulses"])
    p.field_lines[0], p.field_lines[1] = (int(v) for v in sp["field_lines"])
    p.frame_lines = int(sp["frame_lines"])
    p.blocklen = int(rf.blocklen)
    p.blockcut = int(rf.blockcut)
    p.block_stride = int(rf.block_stride)
    p.readlen = int(readlen(rf))
    return p


def locate_params_from(base, prov_delta):
    """A copy of `base` (locate_params of the same decoder) with prov_delta
    set: the per-lane form of locate_params without reading the decoder's
    tables again."""
    p = _lib.LocateParams.from_buffer_copy(base)
    p.prov_delta = float(prov_delta)
    return p


def pulse_threshold(rf):
    """getpulses' first-pass level: iretohz(-20) (core.py:2245-2248)."""
    return rf.DecoderParams["ire0"] + rf.DecoderParams["hz_ire"] * -20.0


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class PulseFinder:
    """Device buffers and pinned host copies for the pulse pass over one
    stitched demod_05 buffer of up to `n` samples."""

    def __init__(self, rf, n, cap_edges=None):
        torch = rf.torch
        self.rf, self.torch, self.n = rf, torch, int(n)
        self.cap = int(cap_edges or max(4096, n // 400))
        self.cap -= self.cap % 2
        dev = rf.device
        self.edges = torch.empty(self.cap, dtype=torch.int64, device=dev)
        self.count = torch.zeros(1, dtype=torch.int64, device=dev)
        self.means = torch.empty(self.cap // 2, dtype=torch.float32, device=dev)
        self.ws = torch.empty(rf.lib.ld_pulse_workspace_bytes(rf.handle, self.n), dtype=torch.uint8, device=dev)
        self.h_edges = torch.empty(self.cap, dtype=torch.int64, pin_memory=True)
        self.h_count = torch.empty(1, dtype=torch.int64, pin_memory=True)
        self.h_means = torch.empty(self.cap // 2, dtype=torch.float32, pin_memory=True)
        self.done = None

    def launch(self, demod05, n=None, stream=None, high=None, bits=None, record=True):
        """Enqueue the pulse search, the sync levels and their copies to the
        host on `stream` (default: current).  bits: (int32 words, threshold)
        of the demod's in-pulse bits over demod05; used when that threshold
        is this search's.  record=False: no completion event (the caller
        tracks the stream, e.g. under graph capture)."""
        torch, rf = self.torch, self.rf
        n = int(n if n is not None else demod05.numel())
        assert n <= self.n and demod05.dtype == torch.float32 and demod05.is_contiguous()
        st = stream if stream is not None else torch.cuda.current_stream(rf.device)
        sp = ctypes.c_void_p(st.cuda_stream)
        hi = pulse_threshold(rf) if high is None else high
        if bits is not None and float(bits[1]) == float(hi):
            _lib.check(rf.lib.ld_find_pulses_bits(rf.handle, bits[0].data_ptr(), n, self.edges.data_ptr(), self.cap,
                                                  self.count.data_ptr(), self.ws.data_ptr(), self.ws.numel(), sp))
        else:
            _lib.check(rf.lib.ld_find_pulses(rf.handle, demod05.data_ptr(), n, float(hi), self.edges.data_ptr(),
                                             self.cap, self.count.data_ptr(), self.ws.data_ptr(), self.ws.numel(),
                                             sp))
        _lib.check(rf.lib.ld_pulse_means(rf.handle, demod05.data_ptr(), n, self.edges.data_ptr(),
                                         self.count.data_ptr(), self.cap // 2,
                                         10.0 * ((rf.freq / rf.linelen) * rf.linelen),    # usectoinpx(10)
                                         float(rf.freq), self.means.data_ptr(), sp))
        # the edges found (not the whole capacity) and their levels, written
        # by the GPU into the pinned host buffers
        _lib.check(rf.lib.ld_pulse_export(rf.handle, self.edges.data_ptr(), self.count.data_ptr(),
                                          self.means.data_ptr(), self.cap, self.h_edges.data_ptr(),
                                          self.h_means.data_ptr(), self.h_count.data_ptr(), sp))
        if record:
            self.done = torch.cuda.Event()
            self.done.record(st)
        self._args = (demod05, n, st, hi)

    def result(self):
        """Wait for the pass; (edges int64, means float32) as numpy arrays."""
        self.done.synchronize()
        cnt = int(self.h_count[0])
        if cnt > self.cap:          # more edges than room: grow and search again
            demod05, n, st, hi = self._args
            self.__init__(self.rf, self.n, cap_edges=cnt + 1024)
            self.launch(demod05, n, st, hi)
            return self.result()
        return self.h_edges.numpy()[:cnt], self.h_means.numpy()[:cnt // 2]


def pulse_search_jobs(searches):
    """ld_pulse_job entries for [(PulseFinder, demod05 tensor, samples,
    in-pulse bit words), ...]: each finder's device and pinned lists."""
    jobs = (_lib.PulseJob * len(searches))()
    for j, (pf, d05, n, bits) in zip(jobs, searches):
        rf = pf.rf
        assert n <= pf.n and d05.dtype == pf.torch.float32 and d05.is_contiguous()
        j.bits, j.n, j.x = bits.data_ptr(), int(n), d05.data_ptr()
        j.edges, j.count, j.means, j.cap = pf.edges.data_ptr(), pf.count.data_ptr(), pf.means.data_ptr(), pf.cap
        j.min_len = 10.0 * ((rf.freq / rf.linelen) * rf.linelen)      # usectoinpx(10)
        j.trim = float(rf.freq)
        j.host_edges, j.host_means, j.host_count = (pf.h_edges.data_ptr(), pf.h_means.data_ptr(),
                                                    pf.h_count.data_ptr())
    return jobs


def pulse_search(searches, workspace, stream):
    """The pulse searches of several buffers in one ld_pulse_search call on
    `stream` (each PulseFinder's result() then reads its lists once the
    stream's work has completed)."""
    jobs = pulse_search_jobs(searches)
    rf = searches[0][0].rf
    _lib.check(rf.lib.ld_pulse_search(rf.handle, jobs, len(searches), workspace.data_ptr(), workspace.numel(),
                                      ctypes.c_void_p(stream.cuda_stream)))


def pulse_search_workspace(searches):
    """Device bytes ld_pulse_search needs for these buffers."""
    jobs = pulse_search_jobs(searches)
    rf = searches[0][0].rf
    return rf.lib.ld_pulse_search_workspace_bytes(rf.handle, jobs, len(searches))


def _facts_array(facts, n):
    """n FieldFacts: the given ones first, the rest unknown."""
    if facts is None:
        return None
    if isinstance(facts, np.ndarray):
        a = np.zeros(max(n, 1), FACTS_DTYPE)
        k = min(len(facts), n)
        a[:k] = facts[:k]
        return (_lib.FieldFacts * max(n, 1)).from_buffer(a)
    arr = (_lib.FieldFacts * max(n, 1))()
    for i, f in enumerate(list(facts)[:n]):
        arr[i] = f
    return arr


MAX_VPULSES = 2048     # classified pulses kept per field for the re-check
VPULSE_DTYPE = np.dtype([("start", np.int64), ("len", np.int32), ("kind", np.int16), ("good", np.int16)])


class WalkBuffers:
    """Host arrays a walk writes into: field records, linelocs1 / linebad
    rows and the classified pulses kept for the re-check."""

    def __init__(self, rf, cap):
        L = proclines(rf)
        self.cap = int(cap)
        self.out = (_lib.FieldLoc * self.cap)()
        self.ll = np.empty((self.cap, L), dtype=np.float64)
        self.bad = np.empty((self.cap, L), dtype=np.uint8)
        self.vp = np.empty((self.cap, MAX_VPULSES), dtype=VPULSE_DTYPE)
        self.vp[:] = 0          # touch the pages once
        self.out_np = np.frombuffer(self.out, dtype=FIELDLOC_DTYPE)    # a view of out


def _walk_args(rf, edges, means, total, start, prev=None, facts=None, max_fields=64, params=None, bufs=None, at=0):
    """The arguments of one ld_locate_fields call (kept alive with it)."""
    P = params or locate_params(rf)
    L = proclines(rf)
    m = max(max_fields, 1)
    if bufs is None:
        bufs = WalkBuffers(rf, m)
        at = 0
    if at + max_fields > bufs.cap:
        raise ValueError("walk buffers too small")
    fsz = ctypes.sizeof(_lib.FieldLoc)
    return types.SimpleNamespace(
        rf=rf, P=P, L=L, max_fields=int(max_fields), total=int(total), start=float(start), prev=prev, bufs=bufs,
        out=ctypes.cast(ctypes.addressof(bufs.out) + at * fsz, ctypes.POINTER(_lib.FieldLoc)),
        ll=bufs.ll[at:], bad=bufs.bad[at:], vp=bufs.vp[at:], po=_lib.PrevField(),
        edges=np.ascontiguousarray(edges, dtype=np.int64), means=np.ascontiguousarray(means, dtype=np.float32),
        fa=_facts_array(facts, m))


def _walk_result(a, n):
    """walk()'s results from the buffers of a finished call of n fields."""
    out, ll, bad = a.out, a.ll, a.bad
    for k in range(n):
        if out[k].status != _lib.LD_LOC_OK:
            ll[k] = np.nan
            bad[k] = 0
    return [_lib.FieldLoc.from_buffer_copy(out[k]) for k in range(n)], ll[:n], bad[:n], a.po, a.vp[:n]


def walk(rf, edges, means, total, start, prev=None, facts=None, max_fields=64, params=None, bufs=None, at=0):
    """ld_locate_fields: returns (locs, linelocs1, linebad, prev_out,
    vpulses) with locs a list of _lib.FieldLoc, linelocs1 / linebad arrays
    of shape (fields, proclines) and vpulses the fields' classified pulses
    (fields, MAX_VPULSES).  With bufs (WalkBuffers), the results are written
    from row `at` on and the arrays returned are views of those rows."""
    a = _walk_args(rf, edges, means, total, start, prev, facts, max_fields, params, bufs, at)
    nf = ctypes.c_int(0)
    _lib.check(rf.lib.ld_locate_fields(ctypes.byref(a.P), _ptr(a.edges), len(a.edges), _ptr(a.means), a.total,
                                       a.start, ctypes.byref(prev) if prev is not None else None, a.fa,
                                       a.max_fields, a.out, _ptr(a.ll), _ptr(a.bad), a.L, _ptr(a.vp), MAX_VPULSES,
                                       ctypes.byref(a.po), ctypes.byref(nf)))
    return _walk_result(a, nf.value)


class PendingWalk:
    """Several independent walk() calls (dicts of its keyword arguments)
    running in one ld_locate_fields_many_start call on up to `threads` host
    threads.  The calls' arguments stay referenced here until wait()."""

    def __init__(self, calls, threads=8):
        self.args = [_walk_args(**c) for c in calls]
        self.jobs = jobs = (_lib.WalkJob * max(len(self.args), 1))()
        for j, a in zip(jobs, self.args):
            j.p = ctypes.addressof(a.P)
            j.edges, j.nedges = _ptr(a.edges), len(a.edges)
            j.means, j.total, j.start = _ptr(a.means), a.total, a.start
            j.prev = ctypes.addressof(a.prev) if a.prev is not None else None
            j.facts = ctypes.addressof(a.fa) if a.fa is not None else None
            j.out = ctypes.cast(a.out, ctypes.c_void_p).value
            j.linelocs1, j.linebad, j.vpulses = _ptr(a.ll), _ptr(a.bad), _ptr(a.vp)
            j.prev_out = ctypes.addressof(a.po)
            j.max_fields, j.max_lines, j.max_vpulses = a.max_fields, a.L, MAX_VPULSES
        self.run = None
        if self.args:
            run = ctypes.c_void_p()
            _lib.check(self.args[0].rf.lib.ld_locate_fields_many_start(jobs, len(self.args), int(threads),
                                                                      ctypes.byref(run)))
            self.run = run

    def wait(self):
        """Join the walks; returns [(walk arguments, fields visited)] in call order."""
        if self.run is not None:
            run, self.run = self.run, None
            _lib.check(self.args[0].rf.lib.ld_locate_fields_many_wait(run))
        return [(a, j.nfields) for a, j in zip(self.args, self.jobs)]

    def __del__(self):
        if getattr(self, "run", None) is not None:   # never leave the threads running on freed arguments
            self.args[0].rf.lib.ld_locate_fields_many_wait(self.run)


def walk_many(calls, threads=8):
    """Several independent walk() calls (dicts of its keyword arguments) in
    one ld_locate_fields_many call on up to `threads` host threads; returns
    each call's walk() result, in order."""
    if not calls:
        return []
    return [_walk_result(a, n) for a, n in PendingWalk(calls, threads).wait()]


def recheck(rf, locs, linelocs1, linebad, vpulses, facts, prev=None, params=None):
    """ld_locate_recheck: (first changed field or -1, the records updated
    to the measured line-0 decisions).  linelocs1 / linebad are checked in
    place."""
    P = params or locate_params(rf)
    L = proclines(rf)
    n = len(locs)
    if n == 0:
        return -1, []
    arr = (_lib.FieldLoc * n)(*locs)
    assert linelocs1.shape == (n, L) and linelocs1.flags.c_contiguous and linebad.flags.c_contiguous
    vp = np.ascontiguousarray(vpulses)
    fa = _facts_array(facts, n)
    first = ctypes.c_int(-1)
    _lib.check(rf.lib.ld_locate_recheck(ctypes.byref(P), arr, _ptr(linelocs1), _ptr(linebad), L, _ptr(vp),
                                        vp.shape[1], n, ctypes.byref(prev) if prev is not None else None, fa,
                                        ctypes.byref(first)))
    return first.value, list(arr)


def recheck_records(rf, locs, linelocs1, linebad, vpulses, facts, prev, params):
    """recheck() on the walk's own records: locs a contiguous FIELDLOC_DTYPE
    array (updated in place), facts a FACTS_DTYPE array of the same length.
    Returns the first changed field or -1."""
    n = len(locs)
    if n == 0:
        return -1
    L = linelocs1.shape[1]
    assert locs.flags.c_contiguous and linelocs1.flags.c_contiguous and linebad.flags.c_contiguous
    assert vpulses.flags.c_contiguous and facts.flags.c_contiguous and len(facts) == n
    first = ctypes.c_int(-1)
    _lib.check(rf.lib.ld_locate_recheck(
        ctypes.byref(params), ctypes.cast(locs.ctypes.data, ctypes.POINTER(_lib.FieldLoc)), _ptr(linelocs1),
        _ptr(linebad), L, _ptr(vpulses), vpulses.shape[1], n, ctypes.byref(prev) if prev is not None else None,
        ctypes.cast(facts.ctypes.data, ctypes.POINTER(_lib.FieldFacts)), ctypes.byref(first)))
    return first.value


def window_pulses(rf, edges, startloc, buflen):
    """The field buffer's pulse list as getpulses returns it (starts
    relative to the buffer): runs that start after its first sample, end
    inside it and last at most 5000 samples."""
    s, e = edges[0:len(edges) - 1:2], edges[1::2]
    s = s[:len(e)]
    keep = (s > startloc) & (e < startloc + buflen) & ((e - s) <= 5000)
    return s[keep] - startloc, (e - s)[keep]


def locate_one(rf, starts, lens, prev=None, startloc=0, fields_written=1, params=None):
    """ld_locate_field: compute_linelocs from one field's pulse list."""
    P = params or locate_params(rf)
    L = proclines(rf)
    out = _lib.FieldLoc()
    ll = np.empty(L, dtype=np.float64)
    bad = np.empty(L, dtype=np.uint8)
    starts = np.ascontiguousarray(starts, dtype=np.int64)
    lens = np.ascontiguousarray(lens, dtype=np.int64)
    _lib.check(rf.lib.ld_locate_field(ctypes.byref(P), _ptr(starts), _ptr(lens), len(starts),
                                      ctypes.byref(prev) if prev is not None else None, int(startloc),
                                      int(fields_written), ctypes.byref(out), _ptr(ll), _ptr(bad), L))
    return out, ll, bad


def _f32_median(v):
    return np.float32(np.median(np.asarray(v, dtype=np.float32)))


def relevel_pulses(rf, demod05, edges, means, loc, stream=None, origin=0):
    """getpulses' second pass for a field whose vsync level is off
    (core.py:2292-2322): black levels from the equalising pulses around the
    vertical sync (GPU means), the threshold halfway between black and sync,
    and findpulses again over the field buffer (GPU).  Returns the field's
    (starts, lens).  origin: the walk coordinate of demod05[0]."""
    torch = rf.torch
    ws, n = int(loc.startloc), int(loc.buflen)
    s, e = edges[0:len(edges) - 1:2], edges[1::2]
    s = s[:len(e)]
    k = np.flatnonzero((s > ws) & (e < ws + n) & ((e - s) <= 5000))
    st, ln = s[k] - ws, (e - s)[k]
    minlength = 10.0 * ((rf.freq / rf.linelen) * rf.linelen)      # usectoinpx(10)
    vl = np.flatnonzero(ln > minlength)
    synclevel = _f32_median(means[k[vl]])
    r = [i for i in list(range(vl[0] - 5, vl[0])) + list(range(vl[-1] + 1, vl[-1] + 6)) if 0 <= i < len(st)]
    r = [i for i in r if rf.freq * 0.75 <= ln[i] <= rf.freq * 2.5]
    begin = np.array([int(st[i] + rf.freq * 5) + ws - origin for i in r], dtype=np.int64)
    end = np.array([int(st[i] + rf.freq * 20) + ws - origin for i in r], dtype=np.int64)
    dev = rf.device
    cur = stream if stream is not None else torch.cuda.current_stream(dev)
    out = torch.empty(max(len(r), 1), dtype=torch.float32, device=dev)
    if len(r):
        b = torch.from_numpy(begin).to(dev)
        en = torch.from_numpy(end).to(dev)
        _lib.check(rf.lib.ld_range_means(rf.handle, demod05.data_ptr(), demod05.numel(), b.data_ptr(),
                                         en.data_ptr(), len(r), out.data_ptr(), ctypes.c_void_p(cur.cuda_stream)))
    cur.synchronize()
    blacklevel = _f32_median(out[:len(r)].cpu().numpy()) if len(r) else np.float32(np.nan)
    high = np.float32((blacklevel + synclevel) / np.float32(2))
    pf = PulseFinder(rf, n)
    pf.launch(demod05[ws - origin:ws - origin + n], n, cur, float(high))
    e2, _ = pf.result()
    s2, t2 = e2[0:len(e2) - 1:2], e2[1::2]
    s2 = s2[:len(t2)]
    keep = (s2 != 0) & ((t2 - s2) <= 5000)
    return s2[keep], (t2 - s2)[keep]


def recalibrated_ire0(rf, demod05, loc):
    """getpulses' first-field retry (core.py:2250-2257): ire0 = the 15th
    percentile of the field buffer (numpy's linear method: the two order
    statistics around 0.15 * (n - 1), interpolated in float64)."""
    torch = rf.torch
    ws, n = int(loc.startloc), int(loc.buflen)
    x = demod05[ws:ws + n]
    pos = 0.15 * (n - 1)
    lo = int(np.floor(pos))
    vals = torch.sort(x).values[lo:lo + 2].double().cpu().numpy()
    frac = pos - lo
    return float(vals[0] + (vals[1] - vals[0]) * frac) if len(vals) > 1 else float(vals[0])


def skip_scores(rf, facts, params=None):
    P = params or locate_params(rf)
    out = []
    for f in facts:
        m = (ctypes.c_double * 8)(*f.medians)
        out.append(rf.lib.ld_skip_score(ctypes.byref(P), m))
    return out


class FieldWalker:
    """The decodefield walk over one stitched demod_05 buffer, resolving
    the getpulses retries on the GPU.  walk() visits fields until max_fields
    of them are valid (or the buffer ends) and returns every field visited
    (valid or not) with its linelocs1 / linebad."""

    def __init__(self, rf, total, max_fields=64, finder=True):
        self.rf = rf
        self.total = int(total)
        self.max_fields = max_fields
        # finder=False: the caller supplies the pulses (walk(edges_means=...));
        # a search buffer is made only if a first-field ire0 recalibration
        # needs a new search
        self._finder = PulseFinder(rf, total) if finder else None
        self.prov_delta = 0.0
        self.bufs = WalkBuffers(rf, max_fields + 16)
        self.allow_recalibration = True

    @property
    def finder(self):
        if self._finder is None:
            self._finder = PulseFinder(self.rf, self.total)
        return self._finder

    def pulses(self, demod05, stream=None, bits=None):
        self.finder.launch(demod05, self.total, stream, bits=bits)

    def first_call(self, start=0.0, prev=None, facts=None, edges_means=None, params=None):
        """The walk() keyword arguments of this walker's first native call
        for walk(start, prev, facts, edges_means=...), so that a batch of
        walkers can make those calls together (walk_many) and then each
        continue with walk(..., first=its result)."""
        edges, means = self.finder.result() if edges_means is None else edges_means
        if self.max_fields > self.bufs.cap:
            self.bufs = WalkBuffers(self.rf, 2 * self.bufs.cap + self.max_fields)
        return dict(rf=self.rf, edges=edges, means=means, total=self.total, start=start, prev=prev, facts=facts,
                    max_fields=self.max_fields, params=params or locate_params(self.rf, self.prov_delta),
                    bufs=self.bufs, at=0)

    def walk_records(self, demod05, start, prev, facts, first_args, nfields, stream=None, edges_means=None,
                     origin=0):
        """walk() continued from its first native call, made elsewhere (a
        PendingWalk of first_call()'s arguments: `first_args` and the
        `nfields` it visited).  Returns (records, linelocs1, linebad,
        prev_out, next start, vpulses) with records a FIELDLOC_DTYPE view of
        this walker's buffers (valid, like the arrays, until its next walk).
        A walk the first call completes takes no per-field Python work."""
        B = first_args.bufs
        nf = int(nfields)
        recs = B.out_np[:nf]
        if nf == 0:
            return recs, B.ll[:0], B.bad[:0], prev, start, B.vp[:0]
        status = recs["status"]
        tail = int(status[-1])
        ok = status == _lib.LD_LOC_OK
        if tail not in (_lib.LD_LOC_NEED_LEVELS, _lib.LD_LOC_NEED_IRE0) and \
                (nf < self.max_fields or int(np.count_nonzero(ok)) >= self.max_fields):
            ll, bad = B.ll[:nf], B.bad[:nf]
            ll[~ok] = np.nan
            bad[~ok] = 0
            return recs, ll, bad, first_args.po, float(recs["start"][-1] + recs["offset"][-1]), B.vp[:nf]
        locs, ll, bad, pv, nxt, vp = self.walk(demod05, start, prev, facts, stream, edges_means, origin,
                                               first=_walk_result(first_args, nf))
        out = self.bufs.out
        for k, c in enumerate(locs):
            out[k] = c
        return self.bufs.out_np[:len(locs)], ll, bad, pv, nxt, vp

    def walk(self, demod05, start=0.0, prev=None, facts=None, stream=None, edges_means=None, origin=0, first=None):
        """Returns (locs, linelocs1, linebad, prev_out, next start,
        vpulses); the arrays are views of this walker's buffers, valid until
        its next walk.  edges_means: the buffer's pulses (edge pairs and
        levels, numpy, in walk coordinates) when the caller searched them
        itself; origin: the walk coordinate of demod05[0] (a window of a
        longer capture); first: the result of the native call first_call()
        describes, already made."""
        rf = self.rf
        edges, means = self.finder.result() if edges_means is None else edges_means
        P = locate_params(rf, self.prov_delta)
        B = self.bufs
        locs = []
        nvalid = 0
        while nvalid < self.max_fields:
            n = len(locs)
            if first is not None:
                got, _, _, po, _ = first
                first = None
            else:
                if n + (self.max_fields - nvalid) > B.cap:
                    old, B = B, WalkBuffers(rf, 2 * B.cap + self.max_fields)
                    B.ll[:n], B.bad[:n], B.vp[:n] = old.ll[:n], old.bad[:n], old.vp[:n]
                    self.bufs = B
                fk = None if facts is None else facts[n:]
                requested = self.max_fields - nvalid
                got, _, _, po, _ = walk(rf, edges, means, self.total, start, prev, fk, requested, P, B, n)
            requested = self.max_fields - nvalid
            if not got:
                break
            tail = got[-1]
            if tail.status in (_lib.LD_LOC_NEED_LEVELS, _lib.LD_LOC_NEED_IRE0):
                locs += got[:-1]
                nvalid += sum(c.status == _lib.LD_LOC_OK for c in got[:-1])
                written = po.fields_written
                pv = po if len(got) > 1 else prev
                if tail.status == _lib.LD_LOC_NEED_IRE0:
                    if not self.allow_recalibration:
                        raise RuntimeError("first-field ire0 recalibration needed where the decoder's "
                                           "parameters are shared (batched captures)")
                    rf.DecoderParams["ire0"] = recalibrated_ire0(rf, demod05, tail)
                    P = locate_params(rf, self.prov_delta)
                    self.finder.launch(demod05, self.total, stream)
                    edges, means = self.finder.result()
                    start = tail.start
                    prev = pv
                    continue
                st, ln = relevel_pulses(rf, demod05, edges, means, tail, stream, origin)
                loc, l1, b1 = locate_one(rf, st, ln, pv if (pv is not None and pv.valid) else None,
                                         tail.startloc, written, P)
                loc.start, loc.readloc, loc.startloc, loc.buflen = tail.start, tail.readloc, tail.startloc, tail.buflen
                loc.offset = (loc.nextfieldoffset - (loc.readloc - loc.startloc)
                              if loc.status == _lib.LD_LOC_OK else loc.nextfieldoffset)
                k = len(locs)
                B.ll[k], B.bad[k] = l1, b1
                B.vp[k] = 0
                locs.append(loc)
                nvalid += loc.status == _lib.LD_LOC_OK
                prev = _prev_from(loc, l1, rf, P, written, None if facts is None else
                                  (facts[k] if k < len(facts) else None))
                start = loc.start + loc.offset
                continue
            locs += got
            nvalid += sum(c.status == _lib.LD_LOC_OK for c in got)
            prev = po
            start = got[-1].start + got[-1].offset
            if len(got) < requested:
                break       # the next field's buffer passes the end
        n = len(locs)
        return locs, B.ll[:n], B.bad[:n], prev, start, B.vp[:n]


def _prev_from(loc, ll1, rf, P, written, fact):
    """The state a resolved field leaves (what ld_locate_fields keeps)."""
    if isinstance(fact, np.void):       # a FACTS_DTYPE record
        fact = _lib.FieldFacts.from_buffer_copy(fact.tobytes())
    pv = _lib.PrevField()
    pv.fields_written = written + (1 if loc.status == _lib.LD_LOC_OK else 0)
    if loc.status != _lib.LD_LOC_OK:
        return pv
    lc = (312 if loc.is_first_field else 313) if rf.system == "PAL" else loc.linecount
    pv.valid, pv.is_first_field, pv.linecount = 1, loc.is_first_field, lc
    pv.sync_confidence, pv.startloc = loc.sync_confidence, loc.startloc
    if fact is not None and fact.known:
        pv.ll_linecount = fact.ll_linecount
        pv.skip_score = skip_scores(rf, [fact], P)[0]
    else:
        pv.ll_linecount = ll1[lc] + P.prov_delta
        pv.skip_score = 100
    return pv
