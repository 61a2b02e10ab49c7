"""Whole-capture GPU decode: the batched demod + hsync + TBC path.

A capture resident in device memory is demodulated block by block into
stitched buffers (DemodCache.read, core.py:1344-1413, without the LRU or
threads: every block is demodulated exactly once in one set of launches).
Fields are then described by their buffer origin and line locations, as
``LDdecode.readfield`` / ``Field`` see them (core.py:4100-4199): every field's
hsync refinement runs in one launch and every field's TBC in one more.
"""

import numpy as np

from . import field as fieldmod


def field_origin(rf, begin):
    """Stitched-buffer index of a field read starting at raw sample `begin`:
    DemodCache.read returns blocks from begin // S, so the field's buffer
    starts at raw (begin // S) * S + blockcut, stitched index (begin // S) * S
    (core.py:1371, :1409)."""
    S = rf.block_stride
    return (begin // S) * S


class CaptureDecoder:
    """Pre-allocated demod + hsync + TBC of `fields` over one capture.

    data: int16 (or float32) device tensor holding the capture.
    fields: list of dicts with demod_offset (stitched index of the field
    buffer origin), demod_len, nlines and, for the burst / pilot alignment,
    linecount and first_field; linelocs1: per-field line locations relative
    to each origin, concatenated (float64, host or device).  lines=True runs
    the burst (NTSC) / pilot (PAL) alignment, fix_badlines and the NTSC
    subcarrier shift between the hsync refine and the TBC, as
    FieldNTSC/FieldPAL.process do; prev_phase_adjust seeds the first NTSC
    field's burst search.
    """

    def __init__(self, rf, data, fields, linelocs1, lineoffset=0, coord_f32=True, smoothing=0.0, audio=False,
                 lines=True, prev_phase_adjust=0.0):
        torch = rf.torch
        self.rf = rf
        self.data = data
        N, S = rf.blocklen, rf.block_stride
        self.nblocks = (data.numel() - N) // S + 1
        self.total = self.nblocks * S
        dev = rf.device
        names = list(rf.video_fields) + ["rfhpf"]
        self.buffers = {n: torch.empty(self.total, dtype=torch.float32, device=dev) for n in names}
        self.audio = bool(audio)
        if self.audio:
            na = rf.filterbank.audio["left"]["nbins"]
            alen = na - rf.blockcut // (N // na) - (-(-32 // (N // na)))
            for n in ("audio_left", "audio_right"):
                self.buffers[n] = torch.empty(self.nblocks * alen, dtype=torch.float32, device=dev)
        W = rf.SysParams["outlinelen"]
        linesout = rf.SysParams["frame_lines"] // 2 + 1
        self.fields = []
        off = 0
        lo = 0
        for f in fields:
            fd = dict(f)
            fd.setdefault("lineoffset", lineoffset)
            fd.setdefault("linesout", linesout)
            fd.setdefault("outwidth", W)
            fd.setdefault("inlinelen", float(rf.linelen))
            fd["lineloc_offset"] = lo
            fd["out_offset"] = off
            fd["demod_len"] = min(fd.get("demod_len", self.total), self.total - fd["demod_offset"])
            lo += fd["nlines"]
            off += fd["linesout"] * fd["outwidth"]
            self.fields.append(fd)
        self.npixels = off
        ll = linelocs1 if isinstance(linelocs1, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(linelocs1, dtype=np.float64))
        self.linelocs1 = ll.to(dev, torch.float64).contiguous()
        assert self.linelocs1.numel() == lo, "line count mismatch"
        self.linelocs2 = torch.empty_like(self.linelocs1)
        self.linebad = torch.empty(lo, dtype=torch.uint8, device=dev)
        self.lines_on = bool(lines)
        self.prev_phase_adjust = prev_phase_adjust
        if self.lines_on:
            for fd in self.fields:
                fd.setdefault("first_field", 1)
                fd.setdefault("linecount", 263 if fd["first_field"] else 262)
            self.linelocs = torch.empty_like(self.linelocs1)
            self.linebad_final = torch.empty(lo, dtype=torch.uint8, device=dev)
            self.phase_adjust = torch.empty(len(self.fields), dtype=torch.float64, device=dev)
            self.field_phase_id = torch.empty(len(self.fields), dtype=torch.int32, device=dev)
            self.ln_ws = torch.empty(max(1, rf.lib.ld_refine_workspace_bytes(rf.handle, len(self.fields), lo)),
                                     dtype=torch.uint8, device=dev)
        else:
            self.linelocs = self.linelocs2
        self.out_u16 = torch.empty(self.npixels, dtype=torch.uint16, device=dev)
        self.hs_ws = torch.empty(max(1, rf.lib.ld_hsync_workspace_bytes(rf.handle, len(self.fields))),
                                 dtype=torch.uint8, device=dev)
        self.tbc_ws = torch.empty(rf.lib.ld_tbc_workspace_bytes(rf.handle, len(self.fields)),
                                  dtype=torch.uint8, device=dev)
        rf._workspace(self.nblocks)
        self.coord_f32 = coord_f32
        self.smoothing = smoothing

    def demod(self, mtf_level=0):
        return self.rf.demod_blocks(self.data, 0, self.nblocks, mtf_level=mtf_level, cut=True, audio=self.audio,
                                    out=self.buffers)

    def refine(self):
        return fieldmod.refine_linelocs_hsync_fields(
            self.rf, self.buffers["demod_05"], self.fields, self.linelocs1, out=(self.linelocs2, self.linebad),
            workspace=self.hs_ws)

    def lines(self):
        """Burst / pilot alignment + fix_badlines (+ NTSC shift) of every field."""
        if not self.lines_on:
            return self.linelocs
        fieldmod.refine_lines_fields(self.rf, self.buffers, self.fields, self.linelocs2, self.linebad,
                                     self.prev_phase_adjust,
                                     out=(self.linelocs, self.linebad_final, self.phase_adjust, self.field_phase_id),
                                     workspace=self.ln_ws)
        return self.linelocs

    def tbc(self):
        return fieldmod.tbc_fields(self.rf, self.buffers["demod"], self.fields, self.linelocs,
                                   coord_f32=self.coord_f32, smoothing=self.smoothing, out_u16=self.out_u16,
                                   final=True, workspace=self.tbc_ws)

    def step(self, mtf_level=0):
        """One pass of the hot path over the whole capture (no host sync)."""
        self.demod(mtf_level)
        self.refine()
        self.lines()
        self.tbc()
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


class PipelinedDecoder:
    """Two CaptureDecoder buffer sets on two streams, so that consecutive
    passes over a capture (or consecutive batches) overlap: pass k's demod
    and hsync refine run on stream A while pass k-1's line alignment and
    TBC run on stream B.  The NTSC burst alignment walks the fields in
    order in one workgroup; overlapped this way it occupies one CU while
    the next demod fills the others.  Both streams have the default
    priority.  Arguments as CaptureDecoder; `step()` returns the uint16
    output of the pass it queued (valid once stream B reaches it)."""

    def __init__(self, rf, data, fields, linelocs1, **kw):
        torch = rf.torch
        self.rf = rf
        self.torch = torch
        self.sets = [CaptureDecoder(rf, data, fields, linelocs1, **kw) for _ in range(2)]
        self.nblocks = self.sets[0].nblocks
        self.stream_a = torch.cuda.Stream(device=rf.device)
        self.stream_b = torch.cuda.Stream(device=rf.device)
        self.free = [None, None]    # event: the set's previous TBC has finished
        self.k = 0

    def step(self, mtf_level=0):
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
            done = torch.cuda.Event()
            done.record(self.stream_b)
        self.free[i] = done
        self.k += 1
        return d.out_u16

    def join(self):
        """Make torch's current stream wait for everything queued so far."""
        cur = self.torch.cuda.current_stream(self.rf.device)
        cur.wait_stream(self.stream_a)
        cur.wait_stream(self.stream_b)
