"""Dropout detection on the GPU (Field.dropout_detect, core.py:2808-2996).

Per field: the error samples of its buffer (rfhpf beyond three standard
deviations, demod_raw above half the sample rate, demod / demod_05 outside
their valid levels with the lower sync-area limit over each line's hsync
and the vertical-sync lines) are found and compacted on the GPU for a whole
batch of fields at once; the grouping into dropout runs and their TBC line
coordinates (build_errlist, dropout_errlist_to_tbc) run natively on the host
over each field's error positions.  The result per field is the reference's
(lines, starts, ends) triple of lists.
"""

import ctypes

import numpy as np

from . import _lib


def dropout_params(rf):
    """The float32 levels of dropout_detect_demod (core.py:2840-2866)."""
    pal = rf.system == "PAL"
    f32 = np.float32
    vs = rf.SysParams["vsync_ire"]
    P = _lib.DropoutParams()
    P.freq_hz_half = float(rf.freq_hz_half)
    P.vmin = f32(rf.iretohz(-70 if pal else -50))
    P.vmax = f32(rf.iretohz(150 if pal else 160))
    P.vmin05 = f32(rf.iretohz(-30))
    P.vmax05 = f32(rf.iretohz(115))
    P.sync_min = f32(rf.iretohz(vs - 60 if pal else vs - 35))
    P.sync_min05 = f32(rf.iretohz(vs - 10))
    P.video_rot = int(rf.delays["video_rot"]) if rf.delays else 0
    P.pal = int(pal)
    return P


def default_hsync_max(rf):
    """LT['hsync'][1] of get_timings (core.py:1761-1810) at the nominal
    hsync length, for a field plan whose pulses were not measured."""
    spl = (rf.freq / rf.linelen) * rf.linelen
    return spl * rf.SysParams["hsyncPulseUS"] + spl * 0.5


def _fields_array(rf, fields):
    arr = (_lib.DropoutField * len(fields))()
    for k, f in enumerate(fields):
        a = arr[k]
        a.demod_offset, a.demod_len = int(f["demod_offset"]), int(f["demod_len"])
        a.lineloc_offset, a.nlines = int(f["lineloc_offset"]), int(f["nlines"])
        a.lineoffset, a.linecount = int(f.get("lineoffset", 0)), int(f["linecount"])
        a.first_field = int(bool(f.get("first_field", 1)))
        a.hsync_max = float(f.get("hsync_max", default_hsync_max(rf)))
    return arr


class DropoutDetector:
    """Reusable buffers for the batched dropout pass.

    detect() enqueues the error search over a batch of fields on the current
    stream (or `stream`); lists() waits for it and returns, per field, the
    reference's (lines, starts, ends).  The position buffer grows (and the
    search is repeated) when a batch has more error samples than it holds."""

    def __init__(self, rf, cap=1 << 20):
        self.rf = rf
        torch = rf.torch
        self.torch = torch
        self.cap = int(cap)
        self.positions = torch.empty(self.cap, dtype=torch.int64, device=rf.device)
        self.starts_dev = None
        self.starts_host = None
        self.ws = None
        self._args = None

    def _stream(self, stream):
        if stream is not None:
            return stream
        return self.torch.cuda.current_stream(self.rf.device).cuda_stream

    def detect(self, buffers, fields, linelocs, stream=None):
        """buffers: the stitched device channels (demod, demod_raw, demod_05,
        rfhpf); fields: dicts with demod_offset, demod_len, lineloc_offset,
        nlines, lineoffset, linecount, first_field and hsync_max; linelocs:
        the fields' final line locations (float64 device tensor)."""
        torch, rf = self.torch, self.rf
        nf = len(fields)
        if nf == 0:
            self._args = None
            return
        arr = _fields_array(rf, fields)
        need = rf.lib.ld_dropout_workspace_bytes(rf.handle, arr, nf)
        if self.ws is None or self.ws.numel() < need:
            self.ws = torch.empty(need, dtype=torch.uint8, device=rf.device)
        if self.starts_dev is None or self.starts_dev.numel() < nf + 1:
            self.starts_dev = torch.empty(nf + 1, dtype=torch.int64, device=rf.device)
            self.starts_host = torch.empty(nf + 1, dtype=torch.int64, pin_memory=True)
        total = buffers["demod"].numel()
        for n in ("demod", "demod_raw", "demod_05", "rfhpf"):
            if buffers[n].dtype != torch.float32 or buffers[n].numel() < total:
                raise ValueError(f"dropout detection: channel {n} must be float32 of {total} samples")
        ll = linelocs if linelocs.dtype == torch.float64 else linelocs.to(torch.float64)
        P = dropout_params(rf)
        self._args = (buffers, arr, nf, ll.contiguous(), P, fields)
        self._launch(stream)

    def _launch(self, stream):
        rf = self.rf
        buffers, arr, nf, ll, P, _ = self._args
        st = self._stream(stream)
        _lib.check(rf.lib.ld_dropout_detect(
            rf.handle, buffers["rfhpf"].data_ptr(), buffers["demod"].data_ptr(), buffers["demod_raw"].data_ptr(),
            buffers["demod_05"].data_ptr(), buffers["demod"].numel(), arr, nf, ll.data_ptr(), ll.numel(), P,
            self.positions.data_ptr(), self.cap, self.starts_dev.data_ptr(), self.ws.data_ptr(), self.ws.numel(),
            ctypes.c_void_p(st)))
        self.starts_host[:nf + 1].copy_(self.starts_dev[:nf + 1], non_blocking=True)
        self._event = self.torch.cuda.Event()
        self._event.record(self.torch.cuda.ExternalStream(st) if stream is not None
                           else self.torch.cuda.current_stream(rf.device))
        self._stream_used = stream

    def positions_per_field(self):
        """Wait for detect(); per field the sorted error sample positions
        (buffer-relative int64 numpy arrays)."""
        if self._args is None:
            return []
        nf = self._args[2]
        self._event.synchronize()
        fs = self.starts_host[:nf + 1].numpy().copy()
        if fs[-1] > self.cap:
            self.cap = int(fs[-1]) + (int(fs[-1]) >> 2) + 1024
            self.positions = self.torch.empty(self.cap, dtype=self.torch.int64, device=self.rf.device)
            self._launch(self._stream_used)
            self._event.synchronize()
            fs = self.starts_host[:nf + 1].numpy().copy()
        pos = self.positions[:int(fs[-1])].cpu().numpy()
        return [pos[fs[k]:fs[k + 1]] for k in range(nf)]

    def lists(self):
        """Wait for detect(); per field the reference's (lines, starts, ends)."""
        if self._args is None:
            return []
        per = self.positions_per_field()
        _, _, nf, ll, _, fields = self._args
        ll_host = ll.cpu().numpy()
        rf = self.rf
        out = []
        for k, f in enumerate(fields):
            o = int(f["lineloc_offset"])
            lines = ll_host[o:o + int(f["nlines"])]
            out.append(error_list(rf, per[k], lines, int(f.get("lineoffset", 0)), int(f["linecount"])))
        return out


def error_list(rf, errmap, linelocs, lineoffset, linecount):
    """build_errlist + dropout_errlist_to_tbc (core.py:2890-2996) in native
    host code: the (lines, starts, ends) of one field from its sorted error
    positions and final line locations."""
    errmap = np.ascontiguousarray(errmap, dtype=np.int64)
    ll = np.ascontiguousarray(linelocs, dtype=np.float64)
    W = int(rf.SysParams["outlinelen"])
    cap = 256
    while True:
        lines = np.empty(cap, np.int32)
        starts = np.empty(cap, np.int32)
        ends = np.empty(cap, np.int32)
        n = ctypes.c_int()
        _lib.check(rf.lib.ld_dropout_list(
            errmap.ctypes.data, len(errmap), ll.ctypes.data, len(ll), lineoffset, linecount, W, float(rf.freq),
            lines.ctypes.data, starts.ctypes.data, ends.ctypes.data, cap, ctypes.byref(n)))
        if n.value <= cap:
            k = n.value
            return lines[:k].tolist(), starts[:k].tolist(), ends[:k].tolist()
        cap = n.value
