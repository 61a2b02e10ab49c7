"""Whole-capture GPU decode: the batched demod + hsync + TBC path.

A capture resident in device memory is demodulated block by block into
stitched buffers (DemodCache.read, core.py:1344-1413, without the LRU or
threads: every block is demodulated exactly once in one set of launches).
Fields are then described by their buffer origin and line locations, as
``LDdecode.readfield`` / ``Field`` see them (core.py:4100-4199): every field's
hsync refinement runs in one launch and every field's TBC in one more.
"""

import ctypes

import numpy as np

from . import _lib
from . import audio as audiomod
from . import dropouts as dropoutmod
from . import field as fieldmod
from . import locate as locatemod


def field_origin(rf, begin):
    """Stitched-buffer index of a field read starting at raw sample `begin`:
    DemodCache.read returns blocks from begin // S, so the field's buffer
    starts at raw (begin // S) * S + blockcut, stitched index (begin // S) * S
    (core.py:1371, :1409)."""
    S = rf.block_stride
    return (begin // S) * S


class CaptureDecoder:
    """Pre-allocated demod + line location + hsync + TBC over one capture.

    data: int16 (or float32) device tensor holding the capture.

    With fields / linelocs1 given, the field plan is fixed: fields is a list
    of dicts with demod_offset (stitched index of the field buffer origin),
    demod_len, nlines and, for the burst / pilot alignment, linecount and
    first_field; linelocs1 holds each field's line locations relative to its
    origin, concatenated (float64, host or device).

    With locate=True the fields are found in the demodulated capture the way
    LDdecode.readfield walks it (core.py:4100-4191, Field.compute_linelocs
    core.py:2325-2481): sync pulses on the GPU, the field walk on the host,
    up to max_fields valid fields per pass, starting at raw sample `start`.
    The walk uses provisional values for what each field reads from the one
    before; after the alignment the measured values are checked against them
    and changed fields are located and aligned again (see locate.py).

    lines=True runs the burst (NTSC) / pilot (PAL) alignment, fix_badlines
    and the NTSC subcarrier shift between the hsync refine and the TBC, as
    FieldNTSC/FieldPAL.process do; prev_phase_adjust seeds the first NTSC
    field's burst search.
    """

    def __init__(self, rf, data, fields=None, linelocs1=None, lineoffset=0, coord_f32=True, smoothing=0.0,
                 audio=False, lines=True, prev_phase_adjust=0.0, locate=False, max_fields=64, start=0.0,
                 audio_rate=None, first_field_number=0):
        torch = rf.torch
        self.rf = rf
        self.torch = torch
        self.data = data
        N, S = rf.blocklen, rf.block_stride
        # samples in the capture (uint8: the packed 10-bit format, 4 per 5 bytes)
        nsamp = data.numel() // 5 * 4 if data.dtype == torch.uint8 else data.numel()
        self.nblocks = (nsamp - N) // S + 1
        self.total = self.nblocks * S
        dev = rf.device
        names = list(rf.video_fields) + ["rfhpf"]
        self.buffers = {n: torch.empty(self.total, dtype=torch.float32, device=dev) for n in names}
        if rf.decode_digital_audio:
            self.buffers["efm"] = torch.empty(self.total, dtype=torch.int16, device=dev)
        if locate and S % 32 == 0 and (rf.blockcut + int(rf.Filters["F05_offset"])) % 32 == 0:
            # demod_05's in-pulse bits, written by the demod for the pulse search
            # (whole 32-sample words per block; otherwise the search reads the samples)
            self.buffers["demod_05_below"] = torch.empty(2 * -(-self.total // 64), dtype=torch.int32, device=dev)
        self._below_high = None
        self.audio = bool(audio)
        if self.audio:
            na = rf.filterbank.audio["left"]["nbins"]
            alen = na - rf.blockcut // (N // na) - (-(-32 // (N // na)))
            for n in ("audio_left", "audio_right"):
                self.buffers[n] = torch.empty(self.nblocks * alen, dtype=torch.float32, device=dev)
        self.W = rf.SysParams["outlinelen"]
        self.linesout = rf.SysParams["frame_lines"] // 2 + 1
        self.default_lineoffset = lineoffset
        self.locating = bool(locate)
        self.start = float(start)
        if self.locating:
            self.max_fields = int(max_fields)
            self.L = locatemod.proclines(rf)
            max_lines = self.max_fields * self.L
        else:
            self.max_fields = len(fields)
            max_lines = sum(int(f["nlines"]) for f in fields)
        self.max_lines = max_lines
        self.linelocs1 = torch.empty(max_lines, dtype=torch.float64, device=dev)
        self.linebad1 = torch.zeros(max_lines, dtype=torch.uint8, device=dev)
        self.linelocs2 = torch.empty_like(self.linelocs1)
        self.linebad = torch.empty(max_lines, dtype=torch.uint8, device=dev)
        self.lines_on = bool(lines)
        self.prev_phase_adjust = prev_phase_adjust
        if self.lines_on:
            self.linelocs = torch.empty_like(self.linelocs1)
            self.linebad_final = torch.empty(max_lines, dtype=torch.uint8, device=dev)
            self.phase_adjust = torch.empty(self.max_fields, dtype=torch.float64, device=dev)
            self.field_phase_id = torch.empty(self.max_fields, dtype=torch.int32, device=dev)
            self.ln_ws = torch.empty(max(1, rf.lib.ld_refine_workspace_bytes(rf.handle, self.max_fields, max_lines)),
                                     dtype=torch.uint8, device=dev)
        else:
            self.linelocs = self.linelocs2
        self.out_u16 = torch.empty(self.max_fields * self.linesout * self.W, dtype=torch.uint16, device=dev)
        self.hs_ws = torch.empty(max(1, rf.lib.ld_hsync_workspace_bytes(rf.handle, self.max_fields)),
                                 dtype=torch.uint8, device=dev)
        self.tbc_ws = torch.empty(rf.lib.ld_tbc_workspace_bytes(rf.handle, self.max_fields),
                                  dtype=torch.uint8, device=dev)
        rf._workspace(self.nblocks)
        self.coord_f32 = coord_f32
        self.smoothing = smoothing
        self.fields = []
        self.dropout_det = None
        self.audio_out = None
        # audio_rate (e.g. 44100): every pass also runs audio stage 2 and the
        # output-rate audio after the TBC; the fields of a pass are numbered
        # from first_field_number for the first tick's offset
        self.audio_rate = audio_rate if audio else None
        self.first_field_number = int(first_field_number)
        if self.locating:
            fsz = ctypes.sizeof(_lib.FieldFacts)
            self.walker = locatemod.FieldWalker(rf, self.total, self.max_fields)
            self.facts_dev = torch.empty(self.max_fields * fsz, dtype=torch.uint8, device=dev)
            self.facts_host = torch.empty(self.max_fields * fsz, dtype=torch.uint8, pin_memory=True)
            self.facts_ws = torch.empty(rf.lib.ld_field_facts_workspace_bytes(rf.handle, self.max_fields),
                                        dtype=torch.uint8, device=dev)
            self.ll_host = torch.empty(max_lines, dtype=torch.float64, pin_memory=True)
            self.bad_host = torch.empty(max_lines, dtype=torch.uint8, pin_memory=True)
            self.walk_state = None
            self.facts_done = None
            self.relocated = 0
        else:
            self.set_fields(fields, linelocs1)

    # -------------------------------------------------------- field plan
    def set_fields(self, fields, linelocs1, linebad1=None):
        """Fix the pass's field list and copy its linelocs1 (and input bad
        flags) into the device buffers."""
        torch = self.torch
        rf = self.rf
        out, off, lo = fieldmod.FieldPlan(), 0, 0
        for f in fields:
            fd = dict(f)
            fd.setdefault("lineoffset", self.default_lineoffset)
            fd.setdefault("linesout", self.linesout)
            fd.setdefault("outwidth", self.W)
            fd.setdefault("inlinelen", float(rf.linelen))
            fd.setdefault("first_field", 1)
            fd.setdefault("linecount", 263 if fd["first_field"] else 262)
            fd["lineloc_offset"] = lo
            fd["out_offset"] = off
            fd["demod_len"] = min(fd.get("demod_len", self.total), self.total - fd["demod_offset"])
            lo += fd["nlines"]
            off += fd["linesout"] * fd["outwidth"]
            out.append(fd)
        if lo > self.max_lines or len(out) > self.max_fields:
            raise ValueError("more fields or lines than this decoder was built for")
        self.fields = out
        self.npixels = off
        if linelocs1 is not None:
            ll = linelocs1 if isinstance(linelocs1, torch.Tensor) else torch.from_numpy(
                np.ascontiguousarray(linelocs1, dtype=np.float64))
            assert ll.numel() == lo, "line count mismatch"
            self.linelocs1[:lo].copy_(ll.to(torch.float64), non_blocking=True)
        if linebad1 is not None:
            b = linebad1 if isinstance(linebad1, torch.Tensor) else torch.from_numpy(
                np.ascontiguousarray(linebad1, dtype=np.uint8))
            self.linebad1[:lo].copy_(b.to(torch.uint8), non_blocking=True)

    # ------------------------------------------------------------ stages
    def demod(self, mtf_level=0):
        below = self.buffers.get("demod_05_below")
        self._below_high = locatemod.pulse_threshold(self.rf) if below is not None else None
        return self.rf.demod_blocks(self.data, 0, self.nblocks, mtf_level=mtf_level, cut=True, audio=self.audio,
                                    out=self.buffers, below=below, below_high=self._below_high)

    def pulses(self):
        """Enqueue the sync-pulse pass over demod_05 and its copy to the host
        (from the demod's in-pulse bits while their threshold holds)."""
        below = self.buffers.get("demod_05_below")
        bits = (below, self._below_high) if below is not None and self._below_high is not None else None
        self.walker.pulses(self.buffers["demod_05"], bits=bits)

    def locate(self, facts=None, prev=None):
        """Host: walk the fields (waits for the pulse pass), then copy the
        valid fields' linelocs1 / bad flags to the device (current stream)."""
        rf = self.rf
        locs, ll1, bad1, prev_out, nxt, vp = self.walker.walk(self.buffers["demod_05"], self.start, prev, facts)
        self.walk_state = (locs, ll1, bad1, prev_out, nxt, vp)
        valid = [k for k, c in enumerate(locs) if c.status == _lib.LD_LOC_OK]
        L = self.L
        fields = []
        for k in valid:
            c = locs[k]
            first = bool(c.is_first_field)
            lo = (2 if first else 3) if rf.system == "PAL" else self.default_lineoffset
            fields.append(dict(demod_offset=int(c.startloc), demod_len=int(c.buflen), nlines=L,
                               linecount=int(c.linecount), first_field=int(first), lineoffset=lo,
                               hsync_max=float(c.hsync_max)))
        n = len(valid) * L
        if valid:
            ll = ll1[valid]
            bad = fieldmod.compute_deriv_error_rows(ll, bad1[valid].astype(bool))
            self.ll_host[:n].copy_(self.torch.from_numpy(ll.reshape(-1)))
            self.bad_host[:n].copy_(self.torch.from_numpy(bad.reshape(-1).astype(np.uint8)))
            self.set_fields(fields, self.ll_host[:n], self.bad_host[:n])
        else:
            self.set_fields([], None)
        return locs

    def refine(self):
        if not self.fields:
            return None
        n = sum(f["nlines"] for f in self.fields)
        bad_in = self.linebad1 if self.locating else None
        return fieldmod.refine_linelocs_hsync_fields(
            self.rf, self.buffers["demod_05"], self.fields, self.linelocs1[:n] if self.locating else self.linelocs1,
            bad_in[:n] if bad_in is not None else None, out=(self.linelocs2, self.linebad), workspace=self.hs_ws)

    def lines(self):
        """Burst / pilot alignment + fix_badlines (+ NTSC shift) of every field."""
        if not self.lines_on or not self.fields:
            return self.linelocs
        fieldmod.refine_lines_fields(self.rf, self.buffers, self.fields, self.linelocs2, self.linebad,
                                     self.prev_phase_adjust,
                                     out=(self.linelocs, self.linebad_final, self.phase_adjust, self.field_phase_id),
                                     workspace=self.ln_ws)
        return self.linelocs

    def tbc(self):
        if not self.fields:
            return self.out_u16
        return fieldmod.tbc_fields(self.rf, self.buffers["demod"], self.fields, self.linelocs,
                                   coord_f32=self.coord_f32, smoothing=self.smoothing, out_u16=self.out_u16,
                                   final=True, workspace=self.tbc_ws)

    def measure_facts(self):
        """Enqueue the per-field facts (final location at linecount, skip
        medians) and their copy to the host."""
        torch, rf = self.torch, self.rf
        if not self.fields:
            self.facts_done = None
            return
        st = torch.cuda.current_stream(rf.device)
        arr = fieldmod._line_fields(self.fields)
        length = rf.SysParams["line_period"] * ((rf.freq / rf.linelen) * rf.linelen)
        _lib.check(rf.lib.ld_field_facts_measure(
            rf.handle, self.buffers["demod"].data_ptr(), self.total, arr, len(self.fields),
            self.linelocs.data_ptr(), self.linelocs.numel(), 1 if rf.system == "PAL" else 0, self.linesout,
            float(length), self.facts_dev.data_ptr(), self.facts_ws.data_ptr(), self.facts_ws.numel(),
            ctypes.c_void_p(st.cuda_stream)))
        nb = len(self.fields) * ctypes.sizeof(_lib.FieldFacts)
        self.facts_host[:nb].copy_(self.facts_dev[:nb], non_blocking=True)
        self.facts_done = torch.cuda.Event()
        self.facts_done.record(st)

    def verify(self, mtf_level=None, max_rounds=4):
        """Wait for the measured facts and walk again with them.  If any
        field's location changes, adopt the new walk and re-run the hsync,
        alignment, TBC and facts (current stream), until the walk is stable.
        Returns the number of extra rounds."""
        rounds = 0
        while self.facts_done is not None:
            self.facts_done.synchronize()
            locs, ll1, bad1, prev_out, nxt, vp = self.walk_state
            measured = self._facts_list()
            facts = []
            j = 0
            for c in locs:
                if c.status == _lib.LD_LOC_OK:
                    facts.append(measured[j])
                    j += 1
                else:
                    facts.append(_lib.FieldFacts())
            self._tune_prov_delta(locs, ll1, measured)
            P = locatemod.locate_params(self.rf, self.walker.prov_delta)
            first, checked = locatemod.recheck(self.rf, locs, ll1, bad1, vp, facts, None, P)
            if first < 0:
                # every field keeps its line table with the measured values;
                # the records now carry the sequential walk's line 0
                self.walk_state = (checked, ll1, bad1, prev_out, nxt, vp)
                return rounds
            if rounds >= max_rounds:
                raise RuntimeError("line location did not settle")
            rounds += 1
            self.relocated += 1
            # walk again: measured values up to the first changed field, the
            # rest provisional until the next check
            self.locate(facts=facts[:first] + [_lib.FieldFacts() for _ in facts[first:]])
            self.refine()
            self.lines()
            self.tbc()
            self.audio_stage()
            self.measure_facts()
        return rounds

    def audio_output(self, freq=44100, timeoffsets=None, stream=None):
        """Enqueue analogue audio stage 2 (RFDecode.audio_phase2,
        core.py:884-950) for every field read of the pass and its int16
        output at `freq` (downscale_audio, core.py:1538-1578) from the
        fields' final line locations; audio_results() returns it.
        timeoffsets: each field's first tick offset (audio.time_offset)."""
        if not self.audio:
            raise ValueError("this decoder was built without analogue audio")
        if self.audio_out is None:
            self.audio_out = audiomod.AudioOutput(self.rf, self.max_fields)
        self.audio_out.stage2(self.buffers["audio_left"], self.buffers["audio_right"], self.fields, stream)
        n = sum(f["nlines"] for f in self.fields)
        self.audio_out.downscale(self.fields, self.linelocs[:max(n, 1)], freq, timeoffsets, stream)

    def audio_stage(self, stream=None):
        """audio_output at audio_rate for the pass's fields (when set)."""
        if self.audio_rate is None or not self.fields:
            return
        sp = self.rf.SysParams
        offs = [audiomod.time_offset(sp, self.first_field_number + k, bool(f["first_field"]), self.audio_rate)
                for k, f in enumerate(self.fields)]
        self.audio_output(self.audio_rate, offs, stream)

    def audio_results(self):
        """Per field: (int16 pairs, next time offset, failed)."""
        return self.audio_out.results() if self.audio_out is not None else []

    def dropouts(self, stream=None):
        """Enqueue dropout detection (Field.dropout_detect, core.py:2978-2996)
        for every field of the pass, over its buffers and final line
        locations; dropout_lists() returns the results."""
        if self.dropout_det is None:
            self.dropout_det = dropoutmod.DropoutDetector(self.rf)
        n = sum(f["nlines"] for f in self.fields)
        self.dropout_det.detect(self.buffers, self.fields, self.linelocs[:n], stream=stream)

    def dropout_lists(self):
        """Per field of the pass, the reference's (lines, starts, ends)."""
        return self.dropout_det.lists() if self.dropout_det is not None else []

    def _facts_list(self):
        n = len(self.fields)
        raw = self.facts_host[:n * ctypes.sizeof(_lib.FieldFacts)].numpy()
        arr = (_lib.FieldFacts * n).from_buffer_copy(raw.tobytes())
        return list(arr)

    def _tune_prov_delta(self, locs, ll1, measured):
        rf = self.rf
        d = []
        j = 0
        for k, c in enumerate(locs):
            if c.status != _lib.LD_LOC_OK:
                continue
            lc = (312 if c.is_first_field else 313) if rf.system == "PAL" else c.linecount
            f = measured[j]
            j += 1
            if f.known and np.isfinite(f.ll_linecount):
                d.append(f.ll_linecount - ll1[k][lc])
        if d:
            self.walker.prov_delta = float(np.median(d))

    def step(self, mtf_level=0):
        """One pass of the hot path over the whole capture."""
        self.demod(mtf_level)
        if self.locating:
            self.pulses()
            self.locate()
        self.refine()
        self.lines()
        self.tbc()
        self.audio_stage()
        if self.locating:
            self.measure_facts()
            self.verify()
        return self.out_u16


def truth_fields(rf, cap, nfields, first_field=0, margin_lines=2, delay=13.0):
    """Field plan from a synthetic capture's ground truth: line l of field f
    starts at half-line synth.field_line0(system, f) + 2*l (line 0 is the
    last line start before the field's equalising pulses; even-numbered
    fields are first fields); the field
    buffer begins at the block holding the start of line -margin_lines;
    linelocs1 = rounded true starts + the typical filter delay of demod_05's
    sync edge, relative to the buffer origin."""
    from .synth import field_line0
    proclines = rf.SysParams["frame_lines"] // 2 + 1 + 10 + (3 if rf.system == "PAL" else 0)
    hs = cap.halfline_starts
    fields, lls = [], []
    for f in range(first_field, first_field + nfields):
        idx = field_line0(rf.system, f) + 2 * np.arange(proclines)
        if idx[0] < 2 * margin_lines or idx[-1] >= len(hs):
            raise ValueError(f"field {f} does not fit the capture")
        raw = hs[idx]
        begin = int(hs[idx[0] - 2 * margin_lines]) - rf.blockcut
        org = field_origin(rf, max(begin, 0))
        is_first = (f % 2) == 0
        fields.append(dict(demod_offset=org, nlines=proclines, first_field=int(is_first),
                           linecount=263 if is_first else 262))
        lls.append(np.round(raw - (org + rf.blockcut) + delay))
    return fields, np.concatenate(lls)


_STREAMS = {}


def _pipeline_streams(torch, device):
    """The two default-priority streams every PipelinedDecoder on a device
    shares.  The runtime spreads streams over a few hardware queues in
    creation order; streams made later (another decoder's) can share one
    queue and then run one after the other, so the pair is made once."""
    key = torch.device(device).index or 0
    if key not in _STREAMS:
        _STREAMS[key] = (torch.cuda.Stream(device=device), torch.cuda.Stream(device=device))
    return _STREAMS[key]


class PipelinedDecoder:
    """CaptureDecoder buffer sets on two streams, so that consecutive passes
    over a capture (or consecutive batches) overlap.  Both streams have the
    default priority.  Arguments as CaptureDecoder.

    Fixed field plan (two sets): pass k's demod and hsync refine run on
    stream A while pass k-1's line alignment and TBC run on stream B;
    `step()` returns the uint16 output of the pass it queued (valid once
    stream B reaches it).

    locate=True (three sets): step k queues pass k's demod and sync-pulse
    search on stream A, walks pass k-1's fields on the host (its pulses are
    ready by then) and queues its hsync refine, alignment, TBC and facts on
    stream B, and finally verifies pass k-2's walk against its measured
    facts (re-running that pass in the rare case a field moves).  The host
    walk thus overlaps the GPU work of both streams.  `step()` returns the
    CaptureDecoder set holding pass k-2 (None on the first two calls);
    `join()` finishes the remaining passes and returns their sets."""

    def __init__(self, rf, data, fields=None, linelocs1=None, **kw):
        torch = rf.torch
        self.rf = rf
        self.torch = torch
        self.locating = bool(kw.get("locate", False))
        nsets = 3 if self.locating else 2
        self.sets = [CaptureDecoder(rf, data, fields, linelocs1, **kw) for _ in range(nsets)]
        self.nblocks = self.sets[0].nblocks
        self.stream_a, self.stream_b = _pipeline_streams(torch, rf.device)
        self.free = [None] * nsets  # event: the set's previous pass is finished (and verified)
        self.k = 0
        self.pending_b = None       # (set, ready event): pass waiting for its walk and stage B
        self.pending_v = None       # set whose walk waits for verification

    def step(self, mtf_level=0):
        if self.locating:
            return self._step_locating(mtf_level)
        torch = self.torch
        i = self.k % 2
        d = self.sets[i]
        cur = torch.cuda.current_stream(self.rf.device)
        with torch.cuda.stream(self.stream_a):
            self.stream_a.wait_stream(cur)
            if self.free[i] is not None:
                self.stream_a.wait_event(self.free[i])
            d.demod(mtf_level)
            d.refine()
            ready = torch.cuda.Event()
            ready.record(self.stream_a)
        with torch.cuda.stream(self.stream_b):
            self.stream_b.wait_event(ready)
            d.lines()
            d.tbc()
            d.audio_stage()
            done = torch.cuda.Event()
            done.record(self.stream_b)
        self.free[i] = done
        self.k += 1
        return d.out_u16

    def _step_locating(self, mtf_level):
        torch = self.torch
        i = self.k % 3
        d = self.sets[i]
        cur = torch.cuda.current_stream(self.rf.device)
        with torch.cuda.stream(self.stream_a):
            self.stream_a.wait_stream(cur)
            if self.free[i] is not None:
                self.stream_a.wait_event(self.free[i])
            d.demod(mtf_level)
            d.pulses()
            ready = torch.cuda.Event()
            ready.record(self.stream_a)
        nxt_v = self._stage_b()
        out = self._verify()
        self.pending_v = nxt_v
        self.pending_b = (i, ready)
        self.k += 1
        return out

    def _stage_b(self):
        """Walk the pending pass's fields and queue its stage B."""
        if self.pending_b is None:
            return None
        torch = self.torch
        j, ready = self.pending_b
        p = self.sets[j]
        with torch.cuda.stream(self.stream_b):
            self.stream_b.wait_event(ready)
            p.locate()
            p.refine()
            p.lines()
            p.tbc()
            p.audio_stage()
            p.measure_facts()
        self.pending_b = None
        return j

    def _verify(self):
        if self.pending_v is None:
            return None
        torch = self.torch
        j = self.pending_v
        p = self.sets[j]
        with torch.cuda.stream(self.stream_b):
            rounds = p.verify()
            if rounds == 0 and p.facts_done is not None:
                # nothing was re-run: the set is free once its own stage B
                # (ending with the facts copy) is; an event recorded now
                # would also wait for the next pass's stage B, queued since
                done = p.facts_done
            else:
                done = torch.cuda.Event()
                done.record(self.stream_b)
        self.free[j] = done
        self.pending_v = None
        return p

    def join(self):
        """Finish every queued pass and make torch's current stream wait for
        them.  Returns the sets of the passes it finished, in order (locate
        mode; an empty list otherwise)."""
        done = []
        if self.locating:
            nxt_v = self._stage_b()
            last = self._verify()
            if last is not None:
                done.append(last)
            self.pending_v = nxt_v
            last = self._verify()
            if last is not None:
                done.append(last)
        cur = self.torch.cuda.current_stream(self.rf.device)
        cur.wait_stream(self.stream_a)
        cur.wait_stream(self.stream_b)
        return done
