"""Analogue audio stage 2 and the output-rate audio on the GPU.

RFDecode.audio_phase2 (core.py:884-950) filters each field read's stage-1
audio (one blocklen-point FFT filter per channel after removing clip peaks);
downscale_audio (core.py:1423-1578) turns it into interleaved int16 pairs at
44.1 kHz (or a multiple of the line rate) from the field's line locations.
Both run batched over the fields of a pass: ``AudioOutput.stage2`` and
``AudioOutput.downscale`` enqueue one launch each, ``results`` copies the
int16 pairs back.
"""

import ctypes
import math

import numpy as np

from . import _lib


def ticks(line_period, linecount, timeoffset=0.0, freq=44100):
    """The tick grid of _downscale_audio_compute_locs_and_swow
    (core.py:1443-1456): (first tick time, rate in Hz, tick count, frame
    time).  A negative freq is a multiple of the line rate and starts at 0;
    the count is numba's arange length ceil((stop - start) / step)."""
    if freq < 0:
        timeoffset = 0
        freq = (1000000 / line_period) * -freq
    frametime = linecount / (1000000 / line_period)
    soundgap = 1 / freq
    n = int(max(math.ceil((frametime + (soundgap / 2) - timeoffset) / soundgap), 0))
    return timeoffset, freq, n, frametime


def time_offset(sysp, field_number, is_first, freq=44100):
    """Field.downscale's first-tick offset (core.py:2644-2670) for output
    field `field_number` (the count of fields written before it, which the
    reference derives from lastfieldwritten and the read position)."""
    if freq < 16000:
        return 0
    linecount = sum(sysp["field_lines"]) * (field_number // 2)
    if not is_first:
        linecount += sysp["field_lines"][0]
    samples_per_line = (sysp["line_period"] / 1000000) / (1 / freq)
    count = linecount * samples_per_line
    frac = count - np.floor(count)
    return (1 - frac) * (1 / freq) if frac > .5 else -frac * (1 / freq)


class AudioOutput:
    """Device buffers for the stage-2 audio and the int16 output of up to
    max_fields field reads per pass."""

    def __init__(self, rf, max_fields=64):
        if not rf.decode_analog_audio:
            raise ValueError("analogue audio is off in this RFDecode")
        self.rf = rf
        self.torch = rf.torch
        self.max_fields = int(max_fields)
        self.fdiv = int(rf.Filters["audio_fdiv"])
        self.a2 = None
        self.out = None
        self.failed = self.torch.zeros(self.max_fields, dtype=self.torch.int32, device=rf.device)
        self.ws = None
        self._plan = None

    def _stream(self, stream):
        return stream if stream is not None else self.torch.cuda.current_stream(self.rf.device).cuda_stream

    def _workspace(self, need):
        if self.ws is None or self.ws.numel() < need:
            self.ws = self.torch.empty(need, dtype=self.torch.uint8, device=self.rf.device)
        return self.ws

    def stage2(self, audio_left, audio_right, fields, stream=None):
        """audio_phase2 of every field read: fields are dicts with
        demod_offset / demod_len (stitched video samples, whole blocks); the
        field's stage-1 audio is the same blocks' audio.  Returns the
        float64 (left, right) device tensors holding each field's stage-2
        audio from its 'a2_offset' (set in the dicts)."""
        torch, rf = self.torch, self.rf
        nf = len(fields)
        if nf > self.max_fields:
            raise ValueError("more fields than this AudioOutput was built for")
        arr = (_lib.AudioField * nf)()
        off = 0
        for k, f in enumerate(fields):
            if f["demod_offset"] % self.fdiv or f["demod_len"] % self.fdiv:
                raise ValueError("field buffers must start and end on audio samples (whole blocks)")
            n = f["demod_len"] // self.fdiv
            arr[k].audio_offset, arr[k].audio_len, arr[k].a2_offset = f["demod_offset"] // self.fdiv, n, off
            f["a2_offset"], f["audio_len"] = off, n
            off += n
        if self.a2 is None or self.a2[0].numel() < off:
            self.a2 = tuple(torch.empty(max(off, 1), dtype=torch.float64, device=rf.device) for _ in range(2))
        ws = self._workspace(max(rf.lib.ld_audio2_workspace_bytes(rf.handle, arr, nf), 1))
        if nf:
            _lib.check(rf.lib.ld_audio2_fields(
                rf.handle, audio_left.data_ptr(), audio_right.data_ptr(), min(audio_left.numel(), audio_right.numel()),
                arr, nf, self.a2[0].data_ptr(), self.a2[1].data_ptr(), self.a2[0].numel(), ws.data_ptr(),
                ws.numel(), ctypes.c_void_p(self._stream(stream))))
        self._fields = fields
        return self.a2

    def downscale(self, fields, linelocs, freq=44100, timeoffsets=None, stream=None):
        """downscale_audio for every field after stage2(): fields carry
        lineloc_offset, nlines and linecount (their final line locations in
        the float64 device tensor linelocs); timeoffsets: per field first
        tick offset (seconds, default 0).  Enqueues the int16 output;
        results() returns it."""
        torch, rf = self.torch, self.rf
        nf = len(fields)
        sp = rf.SysParams
        arr = (_lib.AudioField * nf)()
        plan = []
        off = 0
        rate = None
        for k, f in enumerate(fields):
            t0 = 0.0 if timeoffsets is None else float(timeoffsets[k])
            start, rate, n, frametime = ticks(sp["line_period"], int(f["linecount"]), t0, freq)
            a = arr[k]
            a.audio_offset, a.audio_len, a.a2_offset = 0, f["audio_len"], f["a2_offset"]
            a.lineloc_offset, a.nlines, a.linecount = f["lineloc_offset"], f["nlines"], f["linecount"]
            a.timeoffset, a.nout, a.out_offset = start, max(n - 1, 0), off
            plan.append((off, max(n - 1, 0), start + (n - 1) * (1 / rate) - frametime))
            off += max(n - 1, 0)
        if self.out is None or self.out.numel() < 2 * max(off, 1):
            self.out = torch.empty(2 * max(off, 1), dtype=torch.int16, device=rf.device)
        p = _lib.AudioOutParams(float(sp["line_period"]), float(rf.linelen), float(rate or 44100),
                                float(self.fdiv), float(sp["audio_lfreq"]), float(sp["audio_rfreq"]))
        ll = linelocs if linelocs.dtype == torch.float64 else linelocs.to(torch.float64)
        ws = self._workspace(max(ctypes.sizeof(_lib.AudioField) * nf, 1))
        if nf:
            _lib.check(rf.lib.ld_audio_downscale(
                rf.handle, self.a2[0].data_ptr(), self.a2[1].data_ptr(), self.a2[0].numel(), arr, nf,
                ll.data_ptr(), ll.numel(), p, self.out.data_ptr(), self.out.numel() // 2, self.failed.data_ptr(),
                ws.data_ptr(), ws.numel(), ctypes.c_void_p(self._stream(stream))))
        self._plan = plan
        self.done = torch.cuda.Event()
        self.done.record(torch.cuda.ExternalStream(self._stream(stream)))

    def results(self):
        """Per field: (int16 array of interleaved left/right pairs, the next
        field's time offset as downscale_audio returns it, failed)."""
        if not self._plan:
            return []
        self.done.synchronize()
        out = self.out.cpu().numpy()
        failed = self.failed[:len(self._plan)].cpu().numpy()
        return [(out[2 * o:2 * (o + n)].copy(), nxt, bool(failed[k])) for k, (o, n, nxt) in enumerate(self._plan)]
