"""ctypes binding of libldgpu.so (declarations: include/ldgpu.h).

The library is loaded from this package directory.  There is no fallback: if
it is missing, every entry point raises :class:`LdGpuUnavailable`.
"""

import ctypes
import os
import pathlib
import threading

LIB_PATH = pathlib.Path(__file__).resolve().with_name("libldgpu.so")

LD_OK, LD_EINVAL, LD_ERANGE, LD_EHIP, LD_ESTATE = 0, -1, -2, -3, -4
LD_FILT_RFHPF, LD_FILT_RFVIDEO, LD_FILT_VIDEO, LD_FILT_VIDEO05, LD_FILT_BURST, LD_FILT_PILOT, LD_FILT_EFM = range(7)
LD_IN_INT16, LD_IN_FLOAT32, LD_IN_PACKED10 = 0, 1, 2
LD_FFT_FP32, LD_FFT_MFMA_BF16X3, LD_FFT_MFMA_BF16X6 = 0, 1, 2
ABI_VERSION = 1

# every symbol include/ldgpu.h declares
EXPORTS = (
    "ld_abi_version", "ld_last_error", "ld_create", "ld_destroy", "ld_set_filter", "ld_set_audio",
    "ld_set_demod_params", "ld_demod_workspace_bytes", "ld_demod_blocks", "ld_set_sinc_lut",
    "ld_tbc_workspace_bytes", "ld_tbc_fields", "ld_hsync_workspace_bytes", "ld_refine_hsync", "ld_set_notch",
    "ld_refine_workspace_bytes", "ld_refine_burst_ntsc", "ld_refine_pilot_pal",
    "ld_pulse_workspace_bytes", "ld_find_pulses", "ld_pulse_bits_workspace_bytes", "ld_find_pulses_bits",
    "ld_pulse_means", "ld_range_means",
    "ld_field_facts_workspace_bytes", "ld_field_facts_measure", "ld_locate_field", "ld_locate_fields",
    "ld_skip_score", "ld_locate_recheck", "ld_unpack_10bit",
    "ld_dropout_workspace_bytes", "ld_dropout_detect", "ld_dropout_list",
    "ld_set_audio2", "ld_audio2_workspace_bytes", "ld_audio2_fields", "ld_audio_downscale",
    "ld_demod_noalloc_max_bytes", "ld_demod_blocks_noalloc", "ld_refine_burst_ntsc_chains",
    "ld_demod_batch", "ld_set_fft_variant", "ld_fft_forward", "ld_locate_fields_many",
    "ld_locate_fields_many_start", "ld_locate_fields_many_wait", "ld_pulse_export",
    "ld_pulse_search_workspace_bytes", "ld_pulse_search",
)
LD_LOC_OK, LD_LOC_FAILED, LD_LOC_NEED_LEVELS, LD_LOC_NEED_IRE0 = 0, 1, 2, 3


class LdGpuUnavailable(RuntimeError):
    """libldgpu.so is missing or could not be loaded."""


class LdGpuError(RuntimeError):
    """A libldgpu call returned an error code."""

    def __init__(self, code, message):
        super().__init__(f"libldgpu error {code}: {message}")
        self.code = code


class DemodOutputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "rfhpf", "demod", "demod_raw", "demod_05", "demod_burst", "demod_pilot",
        "audio_left", "audio_right")] + [("video_stride", ctypes.c_int64), ("audio_stride", ctypes.c_int64),
                                         ("efm", ctypes.c_void_p), ("demod_05_below", ctypes.c_void_p),
                                         ("below_high", ctypes.c_double)]


class BatchBlock(ctypes.Structure):
    _fields_ = [("first", ctypes.c_int64), ("sys", ctypes.c_int32), ("slot", ctypes.c_int32)]


class TbcField(ctypes.Structure):
    _fields_ = [
        ("demod_offset", ctypes.c_int64),
        ("demod_len", ctypes.c_int64),
        ("lineloc_offset", ctypes.c_int64),
        ("nlines", ctypes.c_int32),
        ("lineoffset", ctypes.c_int32),
        ("linesout", ctypes.c_int32),
        ("outwidth", ctypes.c_int32),
        ("inlinelen", ctypes.c_double),
        ("out_offset", ctypes.c_int64),
    ]


class LevelParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("ire0", "hz_ire", "output_zero", "vsync_ire", "out_scale")]


class LineField(ctypes.Structure):
    _fields_ = [
        ("demod_offset", ctypes.c_int64),
        ("demod_len", ctypes.c_int64),
        ("lineloc_offset", ctypes.c_int64),
        ("nlines", ctypes.c_int32),
        ("linecount", ctypes.c_int32),
        ("first_field", ctypes.c_int32),
        ("chain", ctypes.c_int32),
    ]


class LineParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("freq_mhz", "linelen", "fsc_mhz", "pilot_mhz", "hz_ire")]


class FieldFacts(ctypes.Structure):
    _fields_ = [("ll_linecount", ctypes.c_double), ("medians", ctypes.c_double * 8), ("known", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class LocateParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "freq", "linelen", "ire0", "hz_ire", "vsync_ire", "hsync_us", "eq_us", "vsync_us", "line_period_us")] + [
        ("first_field_h", ctypes.c_double * 2), ("hsync_tolerance", ctypes.c_double),
        ("prov_delta", ctypes.c_double), ("pal", ctypes.c_int32), ("num_pulses", ctypes.c_int32),
        ("field_lines", ctypes.c_int32 * 2), ("frame_lines", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("blocklen", ctypes.c_int64), ("blockcut", ctypes.c_int64), ("block_stride", ctypes.c_int64),
        ("readlen", ctypes.c_int64)]


class PrevField(ctypes.Structure):
    _fields_ = [("valid", ctypes.c_int32), ("is_first_field", ctypes.c_int32), ("linecount", ctypes.c_int32),
                ("sync_confidence", ctypes.c_int32), ("startloc", ctypes.c_int64), ("ll_linecount", ctypes.c_double),
                ("skip_score", ctypes.c_int32), ("fields_written", ctypes.c_int32)]


class FieldLoc(ctypes.Structure):
    _fields_ = [("start", ctypes.c_double), ("readloc", ctypes.c_int64), ("startloc", ctypes.c_int64),
                ("buflen", ctypes.c_int64), ("line0loc", ctypes.c_double), ("meanlinelen", ctypes.c_double),
                ("nextfieldoffset", ctypes.c_double), ("offset", ctypes.c_double), ("vblank_next", ctypes.c_int64)] + [
        (n, ctypes.c_int32) for n in ("status", "is_first_field", "linecount", "sync_confidence", "skipdetected",
                                      "npulses", "nlines", "analyzed", "nvalid")] + [
        ("last_start", ctypes.c_int64), ("vb_line0", ctypes.c_double * 2), ("vb_firstblank", ctypes.c_int64 * 2), ("vb_conf", ctypes.c_int32 * 2),
        ("vb_flags", ctypes.c_int32 * 2), ("reserved", ctypes.c_int32), ("hsync_max", ctypes.c_double)]


class VPulse(ctypes.Structure):
    _fields_ = [("start", ctypes.c_int64), ("len", ctypes.c_int32), ("kind", ctypes.c_int16), ("good", ctypes.c_int16)]


class PulseJob(ctypes.Structure):
    """ld_pulse_job: one buffer of ld_pulse_search."""
    _fields_ = [("bits", ctypes.c_void_p), ("n", ctypes.c_int64), ("x", ctypes.c_void_p),
                ("edges", ctypes.c_void_p), ("count", ctypes.c_void_p), ("means", ctypes.c_void_p),
                ("cap", ctypes.c_int64), ("min_len", ctypes.c_double), ("trim", ctypes.c_double),
                ("host_edges", ctypes.c_void_p), ("host_means", ctypes.c_void_p), ("host_count", ctypes.c_void_p)]


class WalkJob(ctypes.Structure):
    """ld_walk_job: one ld_locate_fields call of ld_locate_fields_many."""
    _fields_ = [("p", ctypes.c_void_p), ("edges", ctypes.c_void_p), ("nedges", ctypes.c_int64),
                ("means", ctypes.c_void_p), ("total", ctypes.c_int64), ("start", ctypes.c_double),
                ("prev", ctypes.c_void_p), ("facts", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("linelocs1", ctypes.c_void_p), ("linebad", ctypes.c_void_p), ("vpulses", ctypes.c_void_p),
                ("prev_out", ctypes.c_void_p)] + [
        (n, ctypes.c_int32) for n in ("max_fields", "max_lines", "max_vpulses", "nfields", "status", "reserved")]


class DropoutField(ctypes.Structure):
    _fields_ = [("demod_offset", ctypes.c_int64), ("demod_len", ctypes.c_int64), ("lineloc_offset", ctypes.c_int64),
                ("nlines", ctypes.c_int32), ("lineoffset", ctypes.c_int32), ("linecount", ctypes.c_int32),
                ("first_field", ctypes.c_int32), ("hsync_max", ctypes.c_double)]


class DropoutParams(ctypes.Structure):
    _fields_ = [("freq_hz_half", ctypes.c_double)] + [
        (n, ctypes.c_float) for n in ("vmin", "vmax", "vmin05", "vmax05", "sync_min", "sync_min05")] + [
        ("video_rot", ctypes.c_int32), ("pal", ctypes.c_int32)]


class AudioField(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("audio_offset", "audio_len", "a2_offset", "lineloc_offset",
                                              "out_offset")] + [
        ("timeoffset", ctypes.c_double)] + [(n, ctypes.c_int32) for n in ("nlines", "linecount", "nout",
                                                                          "reserved")]


class AudioOutParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("line_period", "linelen", "freq", "scale", "lfreq", "rfreq")]


_lib = None
_lock = threading.Lock()


def _declare(lib):
    c_int, c_i64, c_vp, c_dbl, c_sz = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_double, ctypes.c_size_t
    P = ctypes.POINTER
    sig = {
        "ld_abi_version": (c_int, []),
        "ld_last_error": (ctypes.c_char_p, []),
        "ld_create": (c_int, [c_int, c_int, P(c_vp)]),
        "ld_destroy": (None, [c_vp]),
        "ld_set_filter": (c_int, [c_vp, c_int, P(c_dbl), c_int, c_dbl, c_dbl, c_int]),
        "ld_set_audio": (c_int, [c_vp, c_int, P(c_dbl), c_int, c_int, c_dbl, c_dbl]),
        "ld_set_demod_params": (c_int, [c_vp, c_dbl, c_dbl, c_dbl, c_dbl, c_int]),
        "ld_set_notch": (c_int, [c_vp, c_int, c_int]),
        "ld_demod_workspace_bytes": (c_sz, [c_vp, c_int]),
        "ld_demod_blocks": (c_int, [c_vp, c_vp, c_int, c_i64, c_i64, c_i64, c_int, c_int,
                                    P(DemodOutputs), c_vp, c_sz, c_vp]),
        "ld_demod_noalloc_max_bytes": (c_sz, []),
        "ld_set_fft_variant": (c_int, [c_vp, c_int]),
        "ld_fft_forward": (c_int, [c_vp, c_vp, c_vp, c_int, c_int, c_vp]),
        "ld_demod_batch": (c_int, [P(c_vp), c_vp, c_int, c_i64, c_vp, c_int, P(ctypes.c_int32), c_int,
                                   P(DemodOutputs), c_vp, c_sz, c_vp]),
        "ld_demod_blocks_noalloc": (c_int, [c_vp, c_vp, c_int, c_i64, c_i64, c_i64, c_int, c_int,
                                            P(DemodOutputs), c_vp]),
        "ld_set_sinc_lut": (c_int, [c_vp, P(ctypes.c_float), c_int, c_int]),
        "ld_tbc_workspace_bytes": (c_sz, [c_vp, c_int]),
        "ld_tbc_fields": (c_int, [c_vp, c_vp, c_i64, P(TbcField), c_int, c_vp, c_i64, c_int, c_dbl, c_dbl,
                                  P(LevelParams), c_vp, c_vp, c_i64, c_vp, c_sz, c_vp]),
        "ld_hsync_workspace_bytes": (c_sz, [c_vp, c_int]),
        "ld_refine_hsync": (c_int, [c_vp, c_vp, c_i64, P(TbcField), c_int, c_vp, c_vp, c_i64, c_dbl, c_dbl, c_dbl,
                                    c_dbl, c_int, c_vp, c_vp, c_vp, c_sz, c_vp]),
        "ld_refine_workspace_bytes": (c_sz, [c_vp, c_int, c_i64]),
        "ld_refine_burst_ntsc": (c_int, [c_vp, c_vp, c_vp, c_i64, P(LineField), c_int, P(LineParams), c_vp, c_vp,
                                         c_i64, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
        "ld_refine_burst_ntsc_chains": (c_int, [c_vp, c_vp, c_vp, c_i64, P(LineField), c_int, P(LineParams), c_vp,
                                                c_vp, c_i64, P(c_dbl), c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz,
                                                c_vp]),
        "ld_refine_pilot_pal": (c_int, [c_vp, c_vp, c_i64, P(LineField), c_int, P(LineParams), c_vp, c_vp, c_i64,
                                        c_vp, c_vp, c_vp, c_sz, c_vp]),
        "ld_pulse_workspace_bytes": (c_sz, [c_vp, c_i64]),
        "ld_find_pulses": (c_int, [c_vp, c_vp, c_i64, c_dbl, c_vp, c_i64, c_vp, c_vp, c_sz, c_vp]),
        "ld_pulse_bits_workspace_bytes": (c_sz, [c_vp, c_i64]),
        "ld_find_pulses_bits": (c_int, [c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_sz, c_vp]),
        "ld_pulse_means": (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_dbl, c_dbl, c_vp, c_vp]),
        "ld_range_means": (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp, c_int, c_vp, c_vp]),
        "ld_pulse_export": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
        "ld_pulse_search_workspace_bytes": (c_sz, [c_vp, P(PulseJob), c_int]),
        "ld_pulse_search": (c_int, [c_vp, P(PulseJob), c_int, c_vp, c_sz, c_vp]),
        "ld_field_facts_workspace_bytes": (c_sz, [c_vp, c_int]),
        "ld_field_facts_measure": (c_int, [c_vp, c_vp, c_i64, P(LineField), c_int, c_vp, c_i64, c_int, c_int, c_dbl,
                                           c_vp, c_vp, c_sz, c_vp]),
        "ld_locate_field": (c_int, [P(LocateParams), c_vp, c_vp, c_int, P(PrevField), c_i64, c_int, P(FieldLoc),
                                    c_vp, c_vp, c_int]),
        "ld_locate_fields": (c_int, [P(LocateParams), c_vp, c_i64, c_vp, c_i64, c_dbl, P(PrevField),
                                     P(FieldFacts), c_int, P(FieldLoc), c_vp, c_vp, c_int, c_vp, c_int,
                                     P(PrevField), P(c_int)]),
        "ld_locate_fields_many": (c_int, [P(WalkJob), c_int, c_int]),
        "ld_locate_fields_many_start": (c_int, [P(WalkJob), c_int, c_int, P(c_vp)]),
        "ld_locate_fields_many_wait": (c_int, [c_vp]),
        "ld_skip_score": (c_int, [P(LocateParams), P(c_dbl)]),
        "ld_unpack_10bit": (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp]),
        "ld_locate_recheck": (c_int, [P(LocateParams), P(FieldLoc), c_vp, c_vp, c_int, c_vp, c_int, c_int,
                                      P(PrevField), P(FieldFacts), P(c_int)]),
        "ld_set_audio2": (c_int, [c_vp, c_int, P(c_dbl), c_int, c_dbl]),
        "ld_audio2_workspace_bytes": (c_sz, [c_vp, P(AudioField), c_int]),
        "ld_audio2_fields": (c_int, [c_vp, c_vp, c_vp, c_i64, P(AudioField), c_int, c_vp, c_vp, c_i64, c_vp, c_sz,
                                     c_vp]),
        "ld_audio_downscale": (c_int, [c_vp, c_vp, c_vp, c_i64, P(AudioField), c_int, c_vp, c_i64,
                                       P(AudioOutParams), c_vp, c_i64, c_vp, c_vp, c_sz, c_vp]),
        "ld_dropout_workspace_bytes": (c_sz, [c_vp, P(DropoutField), c_int]),
        "ld_dropout_detect": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, P(DropoutField), c_int, c_vp, c_i64,
                                      P(DropoutParams), c_vp, c_i64, c_vp, c_vp, c_sz, c_vp]),
        "ld_dropout_list": (c_int, [c_vp, c_i64, c_vp, c_int, c_int, c_int, c_int, c_dbl, c_vp, c_vp, c_vp, c_int,
                                    P(c_int)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load(path=None):
    """Load (once) and return the ctypes library."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = pathlib.Path(path) if path else LIB_PATH
        if not p.exists():
            raise LdGpuUnavailable(
                f"{p} not found: build it with `python -m lddecode_amd.build` "
                "(there is no CPU fallback)")
        try:
            lib = ctypes.CDLL(os.fspath(p))
        except OSError as e:
            raise LdGpuUnavailable(f"cannot load {p}: {e}") from e
        _declare(lib)
        if lib.ld_abi_version() != ABI_VERSION:
            raise LdGpuUnavailable(f"{p}: ABI version {lib.ld_abi_version()} != {ABI_VERSION}")
        if path is None:
            _lib = lib
        return lib


def check(rc):
    """Raise LdGpuError for a non-zero return code."""
    if rc != LD_OK:
        msg = load().ld_last_error()
        raise LdGpuError(rc, msg.decode() if msg else "")
    return rc
