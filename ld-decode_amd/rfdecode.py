"""GPU RFDecode: the host mirror of ld-decode's ``RFDecode`` (core.py:253-1046).

``RFDecode`` keeps the reference's constructor arguments, parameter tables,
``Filters`` names, ``demodblock(data, mtf_level, fftdata, cut)`` call and
returned keys, but every per-sample operation runs in libldgpu's HIP kernels.
``demod_blocks`` is the batched form used for whole captures: it demodulates
many overlapping blocks in one set of launches and writes the cut outputs
straight into stitched buffers, which is what ``DemodCache.read`` assembles
(core.py:1344-1413).

Buffers are torch tensors on the HIP device; torch is only the allocator.
"""

import copy
import ctypes

import numpy as np

from . import _lib
from .field import upload_sinc_lut
from .filters import FilterBank, audio2_half, efm_filter, hermitian_half
from .hostdsp import calczc, genwave
from .params import BLOCKSIZE, system_tables

VIDEO_FIELDS_NTSC = ("demod", "demod_raw", "demod_05", "demod_burst")
VIDEO_FIELDS_PAL = VIDEO_FIELDS_NTSC + ("demod_pilot",)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.LdGpuUnavailable("no HIP device visible to torch (there is no CPU fallback)")
    return torch


def _device(torch, index):
    if isinstance(index, torch.device):
        return index
    return torch.device("cuda", torch.cuda.current_device() if index is None else index)


class Channels(dict):
    """Named float32 device arrays of equal length: the GPU counterpart of the
    reference's ``np.rec.array`` outputs (``rv['video']['demod']`` works the
    same way).  ``to_numpy()`` returns the host recarray."""

    @property
    def names(self):
        return tuple(self.keys())

    def __len__(self):
        return len(next(iter(self.values()))) if dict.__len__(self) else 0

    def to_numpy(self):
        arrs = [v.detach().cpu().numpy() for v in self.values()]
        return np.rec.array(arrs, names=list(self.keys()))


class RFDecode:
    """GPU RF decoder for one system (NTSC or PAL) and block length."""

    def __init__(self, inputfreq=40, system="NTSC", blocklen=BLOCKSIZE, decode_digital_audio=False,
                 decode_analog_audio=0, has_analog_audio=True, extra_options=None,
                 decoder_params_override=None, device=None):
        extra_options = extra_options or {}
        decoder_params_override = decoder_params_override or {}
        torch = _torch()
        self.torch = torch
        self.device = _device(torch, device)
        # core.py:328-329
        self.NTSC_ColorNotchFilter = bool(extra_options.get("NTSC_ColorNotchFilter", False))
        self.PAL_V4300D_NotchFilter = bool(extra_options.get("PAL_V4300D_NotchFilter", False))
        # core.py:319-403
        self.blocklen = blocklen
        self.blockcut = 1024
        self.blockcut_end = 0
        self.system = system
        self.setupcount = 0
        self.freq = inputfreq
        self.freq_half = inputfreq / 2
        self.freq_hz = inputfreq * 1000000
        self.freq_hz_half = self.freq_hz / 2
        self.mtf_mult = extra_options.get("MTF_level", 1.0)
        self.mtf_offset = extra_options.get("MTF_offset", 0)
        self.SysParams, self.DecoderParams = system_tables(system, extra_options.get("lowband", False))
        for k in ("ire0", "hz_ire", "vsync_ire"):
            self.DecoderParams[k] = self.SysParams[k]
        for k, v in decoder_params_override.items():
            self.DecoderParams[k] = v
        self.SysParams["analog_audio"] = has_analog_audio
        self.SysParams["AC3"] = extra_options.get("AC3", False)
        if self.SysParams["AC3"]:
            self.SysParams["audio_rfreq"] = self.SysParams["audio_rfreq_AC3"]
        fw = extra_options.get("audio_filterwidth", 0)
        if fw is not None and fw > 0:
            self.DecoderParams["audio_filterwidth"] = fw
        deemp = list(self.DecoderParams["video_deemp"])
        d_low, d_high = extra_options.get("deemp_coeff", (0, 0))
        if d_low > 0:
            deemp[1] = 1 / (d_low * 1000000)
        if d_high > 0:
            deemp[0] = 1 / (d_high * 1000000)
        self.DecoderParams["video_deemp"] = deemp
        self.DecoderParams["video_deemp_strength"] = extra_options.get("deemp_str", 1.0)
        self.linelen = int(np.round(self.freq_hz / (1000000.0 / self.SysParams["line_period"])))
        self.samplesperline = self.freq / self.linelen
        self.hsync_tolerance = 0.4
        self.decode_digital_audio = decode_digital_audio
        if decode_digital_audio and blocklen != 32768:
            # the EFM equaliser is built for 32768 bins (core.py:465)
            raise ValueError("EFM output needs blocklen 32768")
        self.decode_analog_audio = decode_analog_audio
        self.delays = None

        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self.lib.ld_create(self.device.index or 0, blocklen, ctypes.byref(h)))
        self.handle = h
        self.fft_variant = _lib.LD_FFT_FP32
        if self.system == "PAL" and self.PAL_V4300D_NotchFilter:
            # the spur window of core.py:785-788, applied per block in the kernel
            lo, hi = int(blocklen * (8.42 / self.freq)), int(1 + (blocklen * (8.6 / self.freq)))
            _lib.check(self.lib.ld_set_notch(h, lo, hi))
        self._ws = None
        self._mtf_uploaded = None
        self.downscale_sinc_lut = None
        upload_sinc_lut(self)
        self.computefilters()
        self.blockcut_end = 32  # F05_offset (core.py:402, :659)

    def set_fft_variant(self, variant):
        """The transform variant of later demodulations: _lib.LD_FFT_FP32
        (default), _lib.LD_FFT_MFMA_BF16X3 (radix-64 stages as split-BF16
        matrix products on the matrix cores, hi + lo terms) or
        _lib.LD_FFT_MFMA_BF16X6 (hi + mid + lo terms, FP32-level error)."""
        _lib.check(self.lib.ld_set_fft_variant(self.handle, int(variant)))
        self.fft_variant = int(variant)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.ld_destroy(h)
            self.handle = None

    # ------------------------------------------------------------ tables
    @property
    def video_fields(self):
        return VIDEO_FIELDS_PAL if self.system == "PAL" else VIDEO_FIELDS_NTSC

    @property
    def block_stride(self):
        """Usable samples per block (core.py:1073)."""
        return self.blocklen - self.blockcut - 32

    def iretohz(self, ire, spec=False):
        p = self.SysParams if spec else self.DecoderParams
        return p["ire0"] + p["hz_ire"] * ire

    def hztoire(self, hz, spec=False):
        p = self.SysParams if spec else self.DecoderParams
        return (hz - p["ire0"]) / p["hz_ire"]

    def _upload(self, fid, spec, scale, dc=0.0, shift=0):
        half = np.ascontiguousarray(spec[: self.blocklen // 2 + 1], dtype=np.complex128)
        ptr = half.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        _lib.check(self.lib.ld_set_filter(self.handle, fid, ptr, len(half), scale, float(np.real(dc)), shift))

    def _upload_rfvideo(self, level):
        """RFVideo * MTF**level (core.py:800-803), scaled by 1/N for the inverse
        transform; the complex power is folded on the host per level."""
        if self._mtf_uploaded == level:
            return
        spec = self.Filters["RFVideo"]
        if level != 0:
            spec = spec * self.Filters["MTF"] ** level
        self._upload(_lib.LD_FILT_RFVIDEO, spec, 1.0 / self.blocklen)
        self._mtf_uploaded = level

    def computefilters(self):
        """core.py:405-420: (re)build every table, upload, recalibrate delays."""
        self.setupcount += 1
        bank = FilterBank(self.system, self.blocklen, self.freq_hz, self.SysParams, self.DecoderParams,
                          analog_audio=bool(self.decode_analog_audio), color_notch=self.NTSC_ColorNotchFilter)
        self.filterbank = bank
        self.Filters = dict(bank.filters)
        self.Filters["F05_offset"] = bank.offset05
        self.Filters["FVideoBurst_offset"] = bank.offset_burst
        M = self.blocklen // 2
        c2r = 1.0 / M
        F = self.Filters
        self._upload(_lib.LD_FILT_RFHPF, F["Frfhpf"], c2r)
        self._upload(_lib.LD_FILT_VIDEO, F["FVideo"], c2r, F["FVideo"][0], 0)
        self._upload(_lib.LD_FILT_VIDEO05, F["FVideo05"], c2r, F["FVideo05"][0], bank.offset05)
        self._upload(_lib.LD_FILT_BURST, F["FVideoBurst"], c2r, F["FVideoBurst"][0], bank.offset_burst)
        if self.system == "PAL":
            self._upload(_lib.LD_FILT_PILOT, F["FVideoPilot"], c2r, F["FVideoPilot"][0], 0)
        if self.decode_digital_audio:
            # core.py:416-418; the GPU applies the part ifft(X * Fefm).real sees
            F["Fefm"] = efm_filter(self.freq_hz, self.blocklen)
            self._upload(_lib.LD_FILT_EFM, hermitian_half(F["Fefm"]), c2r)
        self._mtf_uploaded = None
        self._upload_rfvideo(0)
        if self.decode_analog_audio:
            self.audio = {}
            for ci, ch in enumerate(("left", "right")):
                a = bank.audio[ch]
                self.audio[ch] = a
                t = np.ascontiguousarray(a["filt1"], dtype=np.complex128)
                ptr = t.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                _lib.check(self.lib.ld_set_audio(self.handle, ci, ptr, len(t), a["lowbin"], a["a1_freq"],
                                                 a["low_freq"]))
                h2 = np.ascontiguousarray(audio2_half(a["audio2_filter"]), dtype=np.complex128)
                p2 = h2.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                _lib.check(self.lib.ld_set_audio2(self.handle, ci, p2, len(h2), float(a["center"])))
            self.Filters["audio_fdiv"] = self.blocklen // bank.audio["left"]["nbins"]
        rot = self.delays["video_rot"] if self.delays else 0
        self._set_params(rot)
        self.computedelays()

    def _set_params(self, rot):
        # the clipped demod (core.py:809) minus the blanking frequency keeps the
        # second FFT's input near zero mean
        _lib.check(self.lib.ld_set_demod_params(self.handle, float(self.freq_hz), 1500000.0,
                                                self.freq_hz * 0.75, float(self.DecoderParams["ire0"]), int(rot)))

    def computedelays(self, mtf_level=0):
        """core.py:952-1046: demodulate a synthetic line on the GPU and measure
        the filter delays; video_rot sets the rfhpf cut."""
        DP, SP = self.DecoderParams, self.SysParams
        fake = np.full(self.blocklen, self.iretohz(0), dtype=np.double)
        synclen = int(4.7 * self.freq)
        fake[1500:1500 + synclen] = self.iretohz(DP["vsync_ire"])
        fake[2000:2000 + synclen] = self.iretohz(DP["vsync_ire"])
        porch_end = 2000 + synclen + int(0.6 * self.freq)
        burst_end = porch_end + int(1.2 * self.freq)
        fake[porch_end:burst_end] += (genwave(np.full(burst_end - porch_end, SP["fsc_mhz"]), self.freq / 2)
                                      * DP["hz_ire"] * 20)
        fake[3000:3500] = self.iretohz(100)
        fake[4500:5000] = self.iretohz(100)
        fake[4200:5500] += genwave(np.full(1300, SP["fsc_mhz"]), self.freq / 2) * DP["hz_ire"] * 20
        fake[2000:2000 + synclen] = self.iretohz(DP["vsync_ire"]) + (
            genwave(np.full(synclen, SP["fsc_mhz"]), self.freq / 2) * DP["hz_ire"] * DP["vsync_ire"])
        F = self.Filters
        emp = np.fft.ifft(np.fft.fft(fake) * F["Fvideo_lpf"] * F["Femp"]).real
        sig = genwave(emp, self.freq_hz / 2) * 4096 + 8192
        sig[6000:6005] = 0
        dec = self.demodblock(data=sig.astype(np.float32), mtf_level=mtf_level, cut=False, _raw_level=True)
        vd = dec["video"]["demod"].cpu().numpy()
        vraw = dec["video"]["demod_raw"].cpu().numpy()
        self.delays = {
            "video_sync": calczc(vd, 1500, self.iretohz(DP["vsync_ire"] / 2), count=512) - 1500,
            "video_white": calczc(vd, 3000, self.iretohz(50), count=512) - 3000,
            "video_rot": int(np.round(calczc(vd, 6000, self.iretohz(-10), count=512) - 6000)),
        }
        self.limits = {"sync": (np.min(vraw[1400:2800]), np.max(vraw[1400:2800])),
                       "viewable": (np.min(vraw[2900:6000]), np.max(vraw[2900:6000]))}
        self._set_params(self.delays["video_rot"])
        return dec

    # ------------------------------------------------------------ demod
    def _level(self, mtf_level):
        """core.py:747-752"""
        return (mtf_level * self.mtf_mult + self.mtf_offset) * self.DecoderParams["MTF_basemult"]

    def _stream(self):
        """hipStream_t of torch's current stream on this device."""
        return self.torch.cuda.current_stream(self.device).cuda_stream

    def _workspace(self, nblocks):
        need = self.lib.ld_demod_workspace_bytes(self.handle, nblocks)
        if self._ws is None or self._ws.numel() < need:
            self._ws = self.torch.empty(need, dtype=self.torch.uint8, device=self.device)
        return self._ws

    def _as_device_input(self, data):
        """(device tensor, LD_IN_* format, samples).  uint8 data is the packed
        10-bit capture format (4 samples in 5 bytes, utils.py:356-390)."""
        torch = self.torch
        if isinstance(data, torch.Tensor):
            t = data if data.device == self.device else data.to(self.device)
        else:
            a = np.asarray(data)
            if a.dtype not in (np.int16, np.float32, np.uint8):
                a = a.astype(np.float32)
            t = torch.from_numpy(np.ascontiguousarray(a)).to(self.device)
        if t.dtype == torch.uint8:
            return t.contiguous(), _lib.LD_IN_PACKED10, t.numel() // 5 * 4
        if t.dtype == torch.int16:
            return t.contiguous(), _lib.LD_IN_INT16, t.numel()
        if t.dtype != torch.float32:
            t = t.float()
        return t.contiguous(), _lib.LD_IN_FLOAT32, t.numel()

    def unpack_10bit(self, packed, first=0, n=None, stream=None):
        """load_packed_data_4_40's unpacking (utils.py:371-410) on the GPU:
        samples [first, first + n) of a packed capture as int16 (code - 512)
        << 6."""
        torch = self.torch
        x, dt, total = self._as_device_input(packed)
        if dt != _lib.LD_IN_PACKED10:
            raise TypeError("unpack_10bit takes packed bytes (uint8)")
        n = total - first if n is None else n
        out = torch.empty(n, dtype=torch.int16, device=self.device)
        st = stream if stream is not None else self._stream()
        _lib.check(self.lib.ld_unpack_10bit(self.handle, x.data_ptr(), x.numel(), int(first), int(n),
                                            out.data_ptr(), ctypes.c_void_p(st)))
        return out

    def demod_blocks(self, data, first=0, nblocks=None, mtf_level=0, cut=True, step=None,
                     audio=None, channels=None, out=None, stream=None, _raw_level=False, below=None,
                     below_high=None):
        """Demodulate blocks b = 0..nblocks-1 reading data[first + b*step : + blocklen].

        With cut=True and the default step (blocklen - 1056) the outputs are the
        stitched DemodCache arrays: element i is raw sample first + 1024 + i.
        below: int32 device words receiving demod_05's in-pulse bits against
        below_high (the input of ld_find_pulses_bits), 2 * ceil(n / 64)
        words for n stitched samples.
        Returns {'video': Channels, 'rfhpf': tensor, 'audio': Channels}."""
        torch = self.torch
        x, dt, nsamp = self._as_device_input(data)
        N = self.blocklen
        S = self.block_stride
        step = S if step is None else step
        if nblocks is None:
            nblocks = (nsamp - first - N) // step + 1
        if nblocks <= 0:
            raise ValueError("not enough input for one block")
        self._upload_rfvideo(mtf_level if _raw_level else self._level(mtf_level))
        vlen = S if cut else N
        if audio is None:
            audio = bool(self.decode_analog_audio)
        nb_a = self.filterbank.audio["left"]["nbins"] if audio else 0
        fdiv = N // nb_a if audio else 1
        alen = 0
        if audio:
            # cut [blockcut//fdiv : -blockcut_end//fdiv] (core.py:872-875)
            alen = (nb_a - self.blockcut // fdiv - (-(-32 // fdiv))) if cut else nb_a
        names = list(self.video_fields) + ["rfhpf"] + (["efm"] if self.decode_digital_audio else [])
        if channels is not None:
            names = [n for n in names if n in channels]
        res = {}
        o = _lib.DemodOutputs()
        for n in names:
            odt = torch.int16 if n == "efm" else torch.float32
            if out is not None and n in out:
                t = out[n]
                if t.numel() < nblocks * (S if n == "rfhpf" else vlen) or t.dtype != odt:
                    raise ValueError(f"output buffer {n} too small or of the wrong type")
            else:
                t = torch.empty(nblocks * (S if n == "rfhpf" else vlen), dtype=odt, device=self.device)
            res[n] = t
            setattr(o, n, t.data_ptr())
        if audio:
            for n in ("audio_left", "audio_right"):
                if out is not None and n in out:
                    t = out[n]
                else:
                    t = torch.empty(nblocks * alen, dtype=torch.float32, device=self.device)
                res[n] = t
                setattr(o, n, t.data_ptr())
        if "rfhpf" in res and not cut and nblocks > 1:
            raise ValueError("uncut output with several blocks: call per block")
        o.video_stride = vlen
        o.audio_stride = alen
        if below is not None and "demod_05" in res:
            if below.dtype != torch.int32 or below.numel() < 2 * -(-(nblocks * vlen) // 64):
                raise ValueError("in-pulse bit buffer too small or not int32")
            o.demod_05_below = below.data_ptr()
            o.below_high = float(below_high)
        ws = self._workspace(nblocks)
        st = stream if stream is not None else self._stream()
        _lib.check(self.lib.ld_demod_blocks(self.handle, x.data_ptr(), dt, nsamp, first, step, nblocks,
                                            1 if cut else 0, ctypes.byref(o), ws.data_ptr(), ws.numel(),
                                            ctypes.c_void_p(st)))
        rv = {"video": Channels((n, res[n]) for n in self.video_fields if n in res)}
        if "rfhpf" in res:
            rv["rfhpf"] = res["rfhpf"]
        if "efm" in res:
            rv["efm"] = res["efm"]
        if audio:
            rv["audio"] = Channels((n, res[n]) for n in ("audio_left", "audio_right"))
        rv["setupcount"] = self.setupcount
        return rv

    def audio_phase2(self, field_audio):
        """RFDecode.audio_phase2 (core.py:917-950) for one field read's
        stage-1 audio (Channels or a dict of float32 device tensors, the
        stitched audio of whole blocks): clip removal and the stage-2 filter
        per channel, returned as Channels of float64 tensors (a read longer
        than a block holds float32-rounded values, as the reference's array
        does in that case)."""
        from . import audio as A
        left, right = field_audio["audio_left"], field_audio["audio_right"]
        n = min(left.numel(), right.numel())
        out = A.AudioOutput(self, 1)
        out.stage2(left.contiguous(), right.contiguous(), [dict(demod_offset=0, demod_len=n * out.fdiv)])
        return Channels((("audio_left", out.a2[0][:n]), ("audio_right", out.a2[1][:n])))

    def demodblock(self, data=None, mtf_level=0, fftdata=None, cut=False, _raw_level=False):
        """RFDecode.demodblock (core.py:747-879) on one block.

        The GPU recomputes the block spectrum from `data` (reading 2 bytes per
        sample is cheaper than storing 8); `fftdata` alone is not accepted."""
        if data is None:
            if fftdata is not None:
                raise NotImplementedError("demodblock on the GPU needs the raw block (data=)")
            raise Exception("demodblock called without raw or FFT data")
        n = len(data)
        if n < self.blocklen:
            raise ValueError(f"block of {n} samples < blocklen {self.blocklen}")
        return self.demod_blocks(data, 0, 1, mtf_level=mtf_level, cut=cut, _raw_level=_raw_level)
