"""GPU field resampler (TBC) and FP64 hsync refinement: the host mirror of
ld-decode's ``Field`` video path (core.py:1587-2722).

``Field`` keeps the reference's attribute names (``linelocs``, ``lineoffset``,
``inlinelen``, ``outlinelen``, ``outlinecount``, ``out_scale``) and
``downscale(...) -> (dsout, dsaudio, efmout)``; ``tbc_fields`` is the batched
form that resamples many fields in one launch.  Line locations come from the
caller or from ``refine_linelocs_hsync`` (GPU, FP64).
"""

import ctypes

import numpy as np
from scipy.special import i0

from . import _lib
from .params import out_scale as _out_scale

SINC_TAPS = 16
SINC_PHASES = 1 << 16


def build_sinc_lut(beta=5.0, taps=SINC_TAPS, phases=SINC_PHASES):
    """(phases + 1) x taps float32 Kaiser-windowed sinc table (utils.py:73-124):
    row p holds the weights for fractional phase p/phases of taps at distances
    (taps/2 - 1 - t) + p/phases, each row normalised to unit sum (float32
    weights divided by their float32 sum); the last row repeats row phases-1."""
    half = taps // 2
    dist = np.arange(half - 1, -half - 1, -1, dtype=np.float64)[None, :] + \
        (np.arange(phases, dtype=np.float64) / phases)[:, None]
    r = dist / half
    window = np.where(np.abs(r) <= 1.0, i0(beta * np.sqrt(np.clip(1.0 - r * r, 0.0, None))) / i0(beta), 0.0)
    w = np.sinc(dist) * window
    lut = np.empty((phases + 1, taps), dtype=np.float32)
    lut[:phases] = w.astype(np.float32) / w.sum(axis=1, keepdims=True).astype(np.float32)
    lut[phases] = lut[phases - 1]
    return lut


def upload_sinc_lut(rf, lut=None):
    lut = build_sinc_lut() if lut is None else np.ascontiguousarray(lut, dtype=np.float32)
    ptr = lut.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    _lib.check(rf.lib.ld_set_sinc_lut(rf.handle, ptr, lut.shape[0], lut.shape[1]))
    rf.downscale_sinc_lut = lut


def level_params(rf, system=None):
    """hz_to_output constants (core.py:1715-1734, utils.py:1061-1075)."""
    dp, sp = rf.DecoderParams, rf.SysParams
    lp = _lib.LevelParams()
    lp.ire0 = dp["ire0"]
    lp.hz_ire = dp["hz_ire"]
    lp.output_zero = sp["outputZero"]
    lp.vsync_ire = dp["vsync_ire"]
    lp.out_scale = _out_scale(system or rf.system, dp["vsync_ire"])
    return lp


def tbc_fields(rf, demod, fields, linelocs, coord_f32=True, smoothing=0.0, threshold=15,
               out_f32=None, out_u16=None, final=True, stream=None, workspace=None):
    """Resample several fields in one launch.

    demod: float32 device tensor; fields: list of dicts with keys
    demod_offset, demod_len, lineloc_offset, nlines, lineoffset, linesout,
    outwidth, inlinelen, out_offset; linelocs: float64 device tensor.
    Returns (out_f32, out_u16) device tensors (one may be None)."""
    torch = rf.torch
    if getattr(rf, "downscale_sinc_lut", None) is None:
        upload_sinc_lut(rf)
    nf = len(fields)
    arr = _field_array(fields)
    total = getattr(fields, "total_out", None)
    if total is None:
        total = max(f["out_offset"] + f["linesout"] * f["outwidth"] for f in fields)
    if final and out_u16 is None:
        out_u16 = torch.empty(total, dtype=torch.uint16, device=rf.device)
    if not final and out_f32 is None:
        out_f32 = torch.empty(total, dtype=torch.float32, device=rf.device)
    ws_bytes = rf.lib.ld_tbc_workspace_bytes(rf.handle, nf)
    ws = workspace if workspace is not None else torch.empty(ws_bytes, dtype=torch.uint8, device=rf.device)
    lp = level_params(rf)
    out_len = (out_u16 if out_u16 is not None else out_f32).numel()
    st = stream if stream is not None else rf._stream()
    _lib.check(rf.lib.ld_tbc_fields(
        rf.handle, demod.data_ptr(), demod.numel(), arr, nf, linelocs.data_ptr(), linelocs.numel(),
        1 if coord_f32 else 0, float(smoothing), float(threshold), ctypes.byref(lp),
        out_f32.data_ptr() if out_f32 is not None else None,
        out_u16.data_ptr() if out_u16 is not None else None, out_len,
        ws.data_ptr(), ws.numel(), ctypes.c_void_p(st)))
    return out_f32, out_u16


class FieldPlan(list):
    """A pass's field dicts (a list) that builds each C struct array of
    them once: the batched calls of one pass share it."""

    def struct_array(self, ctype, build):
        cache = self.__dict__.setdefault("_arrays", {})
        if ctype not in cache:
            cache[ctype] = build(self)
        return cache[ctype]


_DTYPES = {}


def _columns_array(ctype, n, cols):
    dt = _DTYPES.get(ctype)
    if dt is None:
        dt = _DTYPES[ctype] = np.dtype(ctype)
    a = np.zeros(max(n, 1), dt)
    for k, _ in ctype._fields_:
        if k in cols:
            a[k][:n] = cols[k]
    return (ctype * n).from_buffer(a)


class ColumnPlan:
    """A pass's fields held as columns (numpy arrays of equal length): the C
    struct arrays are built from the columns at once, and the per-field
    dicts only when something iterates or indexes the plan."""

    def __init__(self, cols):
        self.cols = {k: np.asarray(v) for k, v in cols.items()}
        keys = list(self.cols)
        self.n = len(self.cols[keys[0]]) if keys else 0
        self._dicts = None
        self._arrays = {_lib.TbcField: _columns_array(_lib.TbcField, self.n, self.cols),
                        _lib.LineField: _columns_array(_lib.LineField, self.n, self.cols)}
        self.total_out = None
        if self.n:
            c = self.cols
            self.total_out = int(np.max(c["out_offset"] + c["linesout"].astype(np.int64) * c["outwidth"]))

    def struct_array(self, ctype, build):
        if ctype not in self._arrays:
            self._arrays[ctype] = build(self)
        return self._arrays[ctype]

    def _rows(self):
        if self._dicts is None:
            keys = list(self.cols)
            self._dicts = [dict(zip(keys, vals)) for vals in zip(*(self.cols[k].tolist() for k in keys))]
        return self._dicts

    def __len__(self):
        return self.n

    def __iter__(self):
        return iter(self._rows())

    def __getitem__(self, i):
        return self._rows()[i]


def plan_from_columns(cols):
    """A FieldPlan from per-key columns (equal-length sequences): the field
    dicts plus the TBC and line struct arrays built column-wise at once."""
    keys = list(cols)
    n = len(cols[keys[0]]) if keys else 0
    plan = FieldPlan(dict(zip(keys, vals)) for vals in zip(*(cols[k] for k in keys)))
    arrays = plan.__dict__.setdefault("_arrays", {})
    arrays[_lib.TbcField] = _columns_array(_lib.TbcField, n, cols)
    arrays[_lib.LineField] = _columns_array(_lib.LineField, n, cols)
    if n:
        plan.total_out = int(np.max(np.asarray(cols["out_offset"]) +
                                    np.asarray(cols["linesout"]) * np.asarray(cols["outwidth"])))
    return plan


def _struct_array(ctype, fields, keys, default=None):
    """(ctype * n) over a numpy structured array of the same layout, filled
    column by column (numpy reads the struct layout from ctypes)."""
    dt = _DTYPES.get(ctype)
    if dt is None:
        dt = _DTYPES[ctype] = np.dtype(ctype)
    n = len(fields)
    a = np.zeros(max(n, 1), dt)
    for k in keys:
        a[k][:n] = [f[k] if default is None else f.get(k, default) for f in fields]
    return (ctype * n).from_buffer(a)


def _build_tbc(fields):
    return _struct_array(_lib.TbcField, fields, [k for k, _ in _lib.TbcField._fields_], default=0)


def _build_lines(fields):
    a = _struct_array(_lib.LineField, fields,
                      ("demod_offset", "demod_len", "lineloc_offset", "nlines", "linecount", "first_field"))
    if fields and "chain" in fields[0]:
        for i, f in enumerate(fields):
            a[i].chain = int(f["chain"])
    return a


def _field_array(fields):
    if hasattr(fields, "struct_array"):
        return fields.struct_array(_lib.TbcField, _build_tbc)
    return _build_tbc(fields)


def refine_linelocs_hsync_fields(rf, demod_05, fields, linelocs1, linebad=None, stream=None, out=None,
                                 workspace=None):
    """Field.refine_linelocs_hsync (core.py:2484-2551) for many fields in one
    launch, FP64 positions.  fields: dicts with demod_offset, demod_len,
    lineloc_offset, nlines; linelocs1: float64 device tensor (per-field
    relative positions).  Returns (linelocs2 float64, linebad uint8) device tensors."""
    torch = rf.torch
    n = linelocs1.numel()
    bad = None
    if linebad is not None:
        bad = linebad.to(device=rf.device, dtype=torch.uint8).contiguous()
    if out is not None:
        ll2, bad2 = out
    else:
        ll2 = torch.empty_like(linelocs1)
        bad2 = torch.empty(n, dtype=torch.uint8, device=rf.device)
    arr = _field_array(fields)
    ws = workspace
    if ws is None:
        ws = torch.empty(max(1, rf.lib.ld_hsync_workspace_bytes(rf.handle, len(fields))), dtype=torch.uint8,
                         device=rf.device)
    dp = rf.DecoderParams
    st = stream if stream is not None else rf._stream()
    _lib.check(rf.lib.ld_refine_hsync(
        rf.handle, demod_05.data_ptr(), demod_05.numel(), arr, len(fields), linelocs1.data_ptr(),
        bad.data_ptr() if bad is not None else None, n, float(rf.freq), rf.iretohz(dp["vsync_ire"] / 2),
        rf.iretohz(-55), rf.iretohz(30), 1 if rf.system == "PAL" else 0, ll2.data_ptr(), bad2.data_ptr(),
        ws.data_ptr(), ws.numel(), ctypes.c_void_p(st)))
    return ll2, bad2


def refine_linelocs_hsync(rf, demod_05, linelocs1, linebad=None, stream=None):
    """Single-field form: linelocs1 (host or device) index demod_05 directly."""
    torch = rf.torch
    if not isinstance(linelocs1, torch.Tensor):
        linelocs1 = torch.from_numpy(np.ascontiguousarray(linelocs1, dtype=np.float64))
    ll1 = linelocs1.to(device=rf.device, dtype=torch.float64).contiguous()
    if linebad is not None and not isinstance(linebad, torch.Tensor):
        linebad = torch.from_numpy(np.ascontiguousarray(linebad, dtype=np.uint8))
    fd = dict(demod_offset=0, demod_len=demod_05.numel(), lineloc_offset=0, nlines=ll1.numel())
    ll2, bad2 = refine_linelocs_hsync_fields(rf, demod_05, [fd], ll1, linebad, stream)
    return ll2, bad2.bool()


def compute_deriv_error(linelocs, baserr):
    """core.py:2553-2564 (host; a few hundred values)."""
    ll = np.asarray(linelocs, dtype=np.float64)
    dd = np.abs(np.diff(np.diff(ll))) > 4
    d1 = np.zeros(len(ll), bool)
    d1[1:-1] = dd
    d2 = np.zeros(len(ll), bool)
    d2[2:] = dd
    return np.asarray(baserr, bool) | d1 | d2


def compute_deriv_error_rows(linelocs, baserr):
    """Field.compute_deriv_error (core.py:2553-2564) for each row of a
    (fields, lines) array."""
    ll = np.asarray(linelocs, dtype=np.float64)
    d = np.abs(np.diff(np.diff(ll, axis=1), axis=1)) > 4
    out = np.array(baserr, dtype=bool, copy=True)
    out[:, 1:-1] |= d
    out[:, 2:] |= d
    return out


def line_params(rf):
    """ld_line_params of an RFDecode: rf.freq, rf.linelen, fsc/pilot MHz, hz_ire."""
    sp = rf.SysParams
    return _lib.LineParams(float(rf.freq), float(rf.linelen), float(sp["fsc_mhz"]),
                           float(sp.get("pilot_mhz", sp["fsc_mhz"])), float(rf.DecoderParams["hz_ire"]))


def _line_fields(fields):
    if hasattr(fields, "struct_array"):
        return fields.struct_array(_lib.LineField, _build_lines)
    return _build_lines(fields)


def refine_lines_fields(rf, video, fields, linelocs2, linebad, prev_phase_adjust=0.0, stream=None, out=None,
                        workspace=None, chain_seeds=None):
    """Burst (NTSC) or pilot (PAL) line alignment + fix_badlines (+ NTSC
    subcarrier shift) for a batch of fields (FieldNTSC.process
    core.py:3472-3495, FieldPAL.process core.py:3212-3237).

    video: dict of DEVICE float32 channels (``demod``/``demod_burst`` for
    NTSC, ``demod_pilot`` for PAL); fields: dicts with demod_offset,
    demod_len, lineloc_offset, nlines, linecount (263/262 for a first/second
    field, as compute_linelocs sets it; the pilot pass does not read it) and
    first_field;
    linelocs2 / linebad: the hsync refine's outputs for the whole batch.
    out: optional preallocated (linelocs, linebad, phase, fpid) device
    tensors; workspace: optional device buffer of ld_refine_workspace_bytes.
    chain_seeds (NTSC): the fields form independent burst chains (their
    "chain" keys, consecutive runs numbered from 0, e.g. one per capture),
    chain c starting from chain_seeds[c] instead of prev_phase_adjust.
    Returns (linelocs float64, linebad uint8, phase_adjust_median per field or
    None, fieldPhaseID per field or None) as device tensors."""
    torch = rf.torch
    n = linelocs2.numel()
    ll2 = linelocs2.to(rf.device, torch.float64).contiguous()
    bad = linebad.to(rf.device, torch.uint8).contiguous()
    if out is not None:
        ll, bad_out, phase, fpid = out
    else:
        ll = torch.empty(n, dtype=torch.float64, device=rf.device)
        bad_out = torch.empty(n, dtype=torch.uint8, device=rf.device)
        phase = torch.empty(len(fields), dtype=torch.float64, device=rf.device)
        fpid = torch.empty(len(fields), dtype=torch.int32, device=rf.device)
    wsb = rf.lib.ld_refine_workspace_bytes(rf.handle, len(fields), n)
    ws = workspace if workspace is not None else torch.empty(max(wsb, 1), dtype=torch.uint8, device=rf.device)
    if ws.numel() < wsb:
        raise ValueError(f"refine workspace of {ws.numel()} bytes < {wsb}")
    lf = _line_fields(fields)
    lp = line_params(rf)
    st = rf._stream() if stream is None else stream
    if rf.system == "NTSC":
        demod, burst = video["demod"], video["demod_burst"]
        if chain_seeds is None:
            _lib.check(rf.lib.ld_refine_burst_ntsc(
                rf.handle, demod.data_ptr(), burst.data_ptr(), demod.numel(), lf, len(fields), ctypes.byref(lp),
                ll2.data_ptr(), bad.data_ptr(), n, float(prev_phase_adjust), ll.data_ptr(), bad_out.data_ptr(),
                phase.data_ptr(), fpid.data_ptr(), ws.data_ptr(), ws.numel(), st))
        else:
            seeds = (ctypes.c_double * len(chain_seeds))(*[float(v) for v in chain_seeds])
            _lib.check(rf.lib.ld_refine_burst_ntsc_chains(
                rf.handle, demod.data_ptr(), burst.data_ptr(), demod.numel(), lf, len(fields), ctypes.byref(lp),
                ll2.data_ptr(), bad.data_ptr(), n, seeds, len(chain_seeds), ll.data_ptr(), bad_out.data_ptr(),
                phase.data_ptr(), fpid.data_ptr(), ws.data_ptr(), ws.numel(), st))
        return ll, bad_out, phase, fpid
    pilot = video["demod_pilot"]
    _lib.check(rf.lib.ld_refine_pilot_pal(
        rf.handle, pilot.data_ptr(), pilot.numel(), lf, len(fields), ctypes.byref(lp), ll2.data_ptr(),
        bad.data_ptr(), n, ll.data_ptr(), bad_out.data_ptr(), ws.data_ptr(), ws.numel(), st))
    return ll, bad_out, None, None


class Field:
    """One field's TBC state (the video half of core.py:1587-2722).

    decode: the dict returned by ``RFDecode.demod_blocks`` (stitched buffers).
    linelocs: line locations in that buffer's sample index (FP64).
    """

    def __init__(self, rf, decode, linelocs=None, lineoffset=0, linebad=None,
                 wow_level_adjust_smoothing=0, wow_interpolation_method="linear", coord_f32=True):
        if wow_interpolation_method != "linear":
            raise NotImplementedError("only the default linear wow interpolation is on the GPU path")
        self.rf = rf
        self.data = decode
        self.inlinelen = rf.linelen
        self.outlinelen = rf.SysParams["outlinelen"]
        self.outlinecount = (rf.SysParams["frame_lines"] // 2) + 1
        self.lineoffset = lineoffset
        self.linelocs = linelocs
        self.linebad = linebad
        self.wow_level_adjust_smoothing = wow_level_adjust_smoothing
        self.coord_f32 = coord_f32
        self.out_scale = _out_scale(rf.system, rf.DecoderParams["vsync_ire"])
        self.dspicture = None

    def refine_linelocs_hsync(self, linelocs1=None, linebad=None):
        ll1 = self.linelocs if linelocs1 is None else linelocs1
        bad = self.linebad if linebad is None else linebad
        ll2, bad2 = refine_linelocs_hsync(self.rf, self.data["video"]["demod_05"], ll1, bad)
        self.linelocs2 = ll2
        self.linebad = bad2
        return ll2

    def process_lines(self, linecount, first_field, prev_phase_adjust=0.0):
        """The rest of FieldNTSC/FieldPAL.process after the hsync refine:
        burst (NTSC) or pilot (PAL) alignment, fix_badlines and the NTSC
        subcarrier shift.  Sets and returns self.linelocs; sets linebad,
        and for NTSC phase_adjust_median and fieldPhaseID."""
        torch = self.rf.torch
        ll2 = self.linelocs2 if getattr(self, "linelocs2", None) is not None else self.linelocs
        if not isinstance(ll2, torch.Tensor):
            ll2 = torch.as_tensor(np.asarray(ll2, dtype=np.float64))
        ll2 = ll2.to(self.rf.device, torch.float64)
        bad = self.linebad
        if bad is None:
            bad = torch.zeros(ll2.numel(), dtype=torch.uint8)
        elif not isinstance(bad, torch.Tensor):
            bad = torch.as_tensor(np.asarray(bad, dtype=np.uint8))
        video = self.data["video"]
        n = video["demod"].numel()
        fd = dict(demod_offset=0, demod_len=n, lineloc_offset=0, nlines=ll2.numel(), linecount=linecount,
                  first_field=int(bool(first_field)))
        ll, bad2, phase, fpid = refine_lines_fields(self.rf, video, [fd], ll2, bad, prev_phase_adjust)
        self.linelocs = ll
        self.linebad = bad2.bool()
        if phase is not None:
            self.phase_adjust_median = float(phase[0].item())
            self.fieldPhaseID = int(fpid[0].item())
        return ll

    def compute_linelocs(self, prevfield=None, fields_written=1, startloc=0):
        """Field.compute_linelocs (core.py:2325-2481) over this field's
        buffer: the GPU finds the sync pulses of demod_05 and their levels
        (getpulses, with its re-threshold and ire0 retries), the native state
        machine places the lines.  prevfield: a located Field before this one
        (its final `linelocs`, `linecount`, `isFirstField`, `sync_confidence`,
        `startloc` and buffers are read, as getLine0 reads them) or None.
        startloc: this buffer's stitched origin (for frameoffset).  Returns
        (linelocs1, linebad, nextfieldoffset) like the reference (None, None,
        offset when no field was found) and sets isFirstField, linecount,
        sync_confidence, vblank_next, valid."""
        from . import locate as L
        rf = self.rf
        torch = rf.torch
        d05 = self.data["video"]["demod_05"]
        n = d05.numel()
        self.startloc = int(startloc)
        pf = L.PulseFinder(rf, n)
        pf.launch(d05, n)
        edges, means = pf.result()
        loc0 = _lib.FieldLoc()
        loc0.startloc, loc0.buflen = 0, n

        def buffer_pulses(edges, means):
            # findpulses on the buffer: runs that end inside it, not starting
            # at sample 0, at most 5000 samples long
            s_all, e_all = edges[0:len(edges) - 1:2], edges[1::2]
            s_all = s_all[:len(e_all)]
            keep = (s_all != 0) & ((e_all - s_all) <= 5000)
            return s_all[keep], (e_all - s_all)[keep], means[:len(e_all)][keep]

        starts, lens, lmeans = buffer_pulses(edges, means)
        prev = None
        if prevfield is not None and getattr(prevfield, "valid", False):
            prev = prevfield._as_prev(fields_written)
        P = L.locate_params(rf)
        if len(starts) == 0 and not fields_written:
            rf.DecoderParams["ire0"] = L.recalibrated_ire0(rf, d05, loc0)
            P = L.locate_params(rf)
            pf.launch(d05, n)
            edges, means = pf.result()
            starts, lens, lmeans = buffer_pulses(edges, means)
        if len(starts):
            long = lens > 10.0 * ((rf.freq / rf.linelen) * rf.linelen)    # usectoinpx(10)
            if not long.any():
                starts = None       # getpulses returns None: no vertical sync seen
            else:
                m = lmeans[long].astype(np.float32)
                level = np.float32(np.median(m))
                dp = rf.DecoderParams
                ire = np.float32((level - np.float32(dp["ire0"])) / np.float32(dp["hz_ire"]))
                if not abs(ire - np.float32(dp["vsync_ire"])) < 5:
                    starts, lens = L.relevel_pulses(rf, d05, edges, means, loc0)
        if starts is None:
            off = rf.linelen * 200 if fields_written else int(rf.freq * 1e6)
            self.valid = False
            self.nextfieldoffset = off
            return None, None, off
        loc, ll1, bad = L.locate_one(rf, starts, lens, prev, self.startloc, fields_written, P)
        self.isFirstField = bool(loc.is_first_field)
        self.linecount = int(loc.linecount)
        self.sync_confidence = int(loc.sync_confidence)
        self.vblank_next = None if loc.vblank_next < 0 else int(loc.vblank_next)
        self.line0loc = loc.line0loc
        self.hsync_max = loc.hsync_max
        self.nextfieldoffset = loc.nextfieldoffset
        self.valid = loc.status == _lib.LD_LOC_OK
        if not self.valid:
            return None, None, loc.nextfieldoffset
        self.linelocs1 = ll1
        self.linebad = bad.astype(bool)
        return ll1, self.linebad, loc.nextfieldoffset

    def dropout_detect(self, linecount=None, first_field=None):
        """Field.dropout_detect (core.py:2978-2996) on this field's buffers
        (decode['video'] demod / demod_raw / demod_05 and decode['rfhpf'])
        with its final `linelocs`: returns (lines, starts, ends) as lists.
        linecount / first_field default to what compute_linelocs found;
        the hsync length is the one its get_timings measured (hsync_max)."""
        from . import dropouts as D
        rf = self.rf
        torch = rf.torch
        ll = self.linelocs
        if not isinstance(ll, torch.Tensor):
            ll = torch.as_tensor(np.asarray(ll, dtype=np.float64))
        ll = ll.to(rf.device, torch.float64).contiguous()
        video = self.data["video"]
        bufs = {n: video[n] for n in ("demod", "demod_raw", "demod_05")}
        bufs["rfhpf"] = self.data["rfhpf"]
        first = self.isFirstField if first_field is None else first_field
        fd = dict(demod_offset=0, demod_len=bufs["demod"].numel(), lineloc_offset=0, nlines=ll.numel(),
                  lineoffset=self.lineoffset, linecount=self.linecount if linecount is None else linecount,
                  first_field=int(bool(first)))
        if getattr(self, "hsync_max", None) is not None:
            fd["hsync_max"] = self.hsync_max
        det = D.DropoutDetector(rf, cap=1 << 16)
        det.detect(bufs, [fd], ll)
        return det.lists()[0]

    def _as_prev(self, fields_written):
        """The state getLine0 reads of this (located, aligned) field."""
        from . import locate as L
        rf = self.rf
        torch = rf.torch
        pal = rf.system == "PAL"
        lc = (312 if self.isFirstField else 313) if pal else self.linecount
        ll = self.linelocs
        if not isinstance(ll, torch.Tensor):
            ll = torch.as_tensor(np.asarray(ll, dtype=np.float64))
        ll = ll.to(rf.device, torch.float64).contiguous()
        demod = self.data["video"]["demod"]
        fd = dict(demod_offset=0, demod_len=demod.numel(), lineloc_offset=0, nlines=ll.numel(),
                  linecount=self.linecount, first_field=int(bool(self.isFirstField)))
        facts = torch.empty(ctypes.sizeof(_lib.FieldFacts), dtype=torch.uint8, device=rf.device)
        ws = torch.empty(rf.lib.ld_field_facts_workspace_bytes(rf.handle, 1), dtype=torch.uint8, device=rf.device)
        length = rf.SysParams["line_period"] * ((rf.freq / rf.linelen) * rf.linelen)
        _lib.check(rf.lib.ld_field_facts_measure(
            rf.handle, demod.data_ptr(), demod.numel(), _line_fields([fd]), 1, ll.data_ptr(), ll.numel(),
            1 if pal else 0, rf.SysParams["frame_lines"] // 2 + 1, float(length), facts.data_ptr(),
            ws.data_ptr(), ws.numel(), ctypes.c_void_p(rf._stream())))
        f = _lib.FieldFacts.from_buffer_copy(facts.cpu().numpy().tobytes())
        pv = _lib.PrevField()
        pv.valid, pv.is_first_field, pv.linecount = 1, int(bool(self.isFirstField)), lc
        pv.sync_confidence, pv.startloc = int(self.sync_confidence), int(getattr(self, "startloc", 0))
        pv.ll_linecount = f.ll_linecount
        pv.skip_score = L.skip_scores(rf, [f])[0]
        pv.fields_written = fields_written
        return pv

    def downscale_audio(self, freq=44100, lineinfo=None, lastfieldwritten=None, timeoffset=None):
        """downscale_audio (core.py:1538-1578) of this field: interleaved
        int16 pairs at `freq` (negative: a multiple of the line rate).
        decode['audio'] holds the field read's stage-1 audio (float32, as
        RFDecode.demod_blocks returns it; stage 2 runs first) or its stage-2
        audio (float64).  The first tick's offset follows Field.downscale
        (core.py:2644-2670) from lastfieldwritten = (field number, readloc)
        and this field's readloc, unless timeoffset is given.  Sets
        audio_next_offset; returns the int16 array."""
        from . import audio as A
        rf = self.rf
        torch = rf.torch
        au = self.data["audio"]
        left, right = au["audio_left"], au["audio_right"]
        out = A.AudioOutput(rf, 1)
        n = min(left.numel(), right.numel())
        fd = dict(demod_offset=0, demod_len=n * out.fdiv)
        if left.dtype == torch.float64:
            out.a2 = (left.contiguous(), right.contiguous())
            fd.update(a2_offset=0, audio_len=n)
        else:
            out.stage2(left.contiguous(), right.contiguous(), [fd])
        ll = self.linelocs if lineinfo is None else lineinfo
        if not isinstance(ll, torch.Tensor):
            ll = torch.as_tensor(np.asarray(ll, dtype=np.float64))
        ll = ll.to(rf.device, torch.float64).contiguous()
        if timeoffset is None:
            timeoffset = 0.0
            if lastfieldwritten and freq >= 16000:
                per_field = rf.freq_hz / rf.SysParams["FPS"] / 2
                gap = (getattr(self, "readloc", 0) - lastfieldwritten[1]) / per_field
                number = int(np.round(lastfieldwritten[0] + gap))
                timeoffset = A.time_offset(rf.SysParams, number, bool(getattr(self, "isFirstField", True)), freq)
        fd.update(lineloc_offset=0, nlines=ll.numel(), linecount=int(self.linecount))
        out.downscale([fd], ll, freq, [timeoffset])
        pcm, nxt, failed = out.results()[0]
        self.audio_next_offset = nxt
        self.audio_failed = failed
        return pcm

    def downscale(self, lineinfo=None, linesout=None, outwidth=None, channel="demod", audio=0,
                  final=False, lastfieldwritten=None):
        """core.py:2626-2722: returns (dsout, dsaudio, efmout); dsout is the
        float32 Hz field, or the uint16 codes when final=True; dsaudio the
        int16 audio pairs at rate `audio` (None when audio is 0); efmout the
        EFM samples from line 1 to line linecount + 1 when the RFDecode
        decodes digital audio (decode['efm'], as demod_blocks returns it)."""
        dsaudio = None
        if audio and self.rf.decode_analog_audio:
            dsaudio = self.downscale_audio(audio, lineinfo, lastfieldwritten)
        torch = self.rf.torch
        ll = self.linelocs if lineinfo is None else lineinfo
        if not isinstance(ll, torch.Tensor):
            ll = torch.as_tensor(np.asarray(ll, dtype=np.float64))
        ll = ll.to(self.rf.device, torch.float64).contiguous()
        outwidth = self.outlinelen if outwidth is None else outwidth
        linesout = self.outlinecount if linesout is None else linesout
        src = self.data["video"][channel]
        fd = dict(demod_offset=0, demod_len=src.numel(), lineloc_offset=0, nlines=ll.numel(),
                  lineoffset=self.lineoffset, linesout=linesout, outwidth=outwidth,
                  inlinelen=float(self.inlinelen), out_offset=0)
        f32, u16 = tbc_fields(self.rf, src, [fd], ll, coord_f32=self.coord_f32,
                              smoothing=self.wow_level_adjust_smoothing, final=final)
        dsout = u16 if final else f32
        if final:
            self.dspicture = dsout
        efmout = None
        if self.rf.decode_digital_audio and "efm" in self.data and getattr(self, "linecount", None):
            # the field's active span of the EFM output (core.py:2704-2707)
            llh = ll.cpu().numpy()
            efmout = self.data["efm"][int(llh[1]):int(llh[self.linecount + 1])]
        self.efmout = efmout
        return dsout, dsaudio, efmout
