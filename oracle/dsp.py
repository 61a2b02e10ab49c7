"""FP64 restatement of ld-decode's RF demodulation (test oracle only).

Follows ``lddecode/core.py`` and ``lddecode/utils.py``; each function cites the
lines it restates.  All arithmetic is float64 / complex128 (numpy, scipy.fft,
scipy.signal), the same libraries the reference calls.

How each part is pinned (tests/test_oracle_*.py):
  * filter tables: scipy.signal is the reference's own design library; closed
    forms pin the composition (Butterworth |H|^2 = 1/(1+(tan(w/2)/tan(wc/2))^2n),
    Hilbert mask 1/2/0, firwin unit DC gain, de-emphasis DC gain 1,
    group-delay equaliser unit magnitude and conjugate symmetry).
  * FFT stages: brute-force DFT on small N and Parseval.
  * unwrap_hilbert: closed form (FM tone at f -> exactly f; [0, fs) wrap).
  * demodblock: a constant-frequency carrier demodulates to its frequency on
    every branch with unit DC gain; roll/cut index algebra by construction
    tests (an impulse through the pipeline lands where the reference's
    roll/cut put it).
  * stitching: equals one long overlap-save (closed-form index map).
"""

import math

import numpy as np
import scipy.fft as sfft
import scipy.signal as sps

from .params import filter_params, sys_params

TAU = 2 * np.pi
BLOCKSIZE = 32 * 1024  # core.py:66


# --------------------------------------------------------------------------
# utils.py helpers
# --------------------------------------------------------------------------

def polar2z(r, theta):
    """utils.py:763"""
    return r * np.exp(1j * theta)


def emphasis_iir(t1, t2, fs):
    """utils.py:766-778 -- bilinear first-order (de)emphasis, zero at t1, pole at t2."""
    w1 = 2 * fs * np.tan((1 / t1) / (2 * fs))
    w2 = 2 * fs * np.tan((1 / t2) / (2 * fs))
    tf_b, tf_a = sps.zpk2tf([-w1], [-w2], w2 / w1)
    return sps.bilinear(tf_b, tf_a, fs)


def filtfft(filt, blocklen):
    """utils.py:782-783 -- whole-circle freqz at blocklen bins."""
    return sps.freqz(filt[0], filt[1], blocklen, whole=1)[1]


def build_hilbert(fft_size):
    """utils.py:882-890 -- analytic-signal mask: 1 at DC and N/2, 2 between, 0 above."""
    if fft_size % 2:
        raise Exception("build_hilbert: must have even fft_size")
    out = np.zeros(fft_size)
    out[0] = out[fft_size // 2] = 1
    out[1:fft_size // 2] = 2
    return out


def unwrap_hilbert(hilbert, freq_hz):
    """utils.py:898-926 -- conjugate-product FM discriminator, Hz in [0, freq_hz)."""
    hilbert = np.asarray(hilbert)
    out = np.empty(len(hilbert), dtype=np.float64)
    out[0] = 0.0
    prod = hilbert[1:] * np.conj(hilbert[:-1])
    d = np.arctan2(prod.imag, prod.real)
    out[1:] = np.where(d < 0.0, d + TAU, d)
    return out * (freq_hz / TAU)


def nb_round(x):
    """utils.py:1027-1029 (int(np.round(x)), half to even)."""
    return int(np.round(x))


def fft_determine_slices(center, min_bandwidth, freq_hz, bins_in):
    """utils.py:929-950"""
    binwidth = freq_hz / bins_in
    cbin = nb_round(center / binwidth)
    bbins = nb_round(min_bandwidth / binwidth)
    nbins = 2 * (2 ** math.ceil(math.log2(bbins * 2)))
    lowbin = cbin - (nbins // 4)
    cut_freq = binwidth * nbins
    return lowbin, nbins, cut_freq


def fft_do_slice(fdomain, lowbin, nbins, blocklen):
    """utils.py:953-962"""
    h = nbins // 2
    return np.concatenate([fdomain[lowbin:lowbin + h],
                           fdomain[blocklen - lowbin - h:blocklen - lowbin]])


def genwave(rate, freq, initialphase=0.0):
    """utils.py:1009-1023 -- FM synthesis by phase accumulation (scalar loop
    restated in closed form: out[i] = sin(phi_i), phi_i = phi_0 + pi*sum(rate[:i])/freq,
    wrapped).  The wrap does not change sin(), so the cumulative sum is exact
    up to float64 rounding of the running phase."""
    rate = np.asarray(rate, dtype=np.float64)
    inc = np.pi * (rate / freq)
    phase = initialphase + np.concatenate([[0.0], np.cumsum(inc[:-1])])
    return np.sin(phase)


def calczc(data, start_offset, target, edge=0, count=16, reverse=False):
    """utils.py:795-854 -- first threshold crossing after start_offset with a
    linear-interpolated sub-sample position (None if not found)."""
    if reverse:
        rev = calczc(np.asarray(data)[int(start_offset)::-1], 0, target, edge, count)
        return None if rev is None else start_offset - rev
    s = int(start_offset)
    icount = int(count + 1)
    if edge == 0:
        edge = 1 if data[s] < target else -1
    seg = np.asarray(data[s:s + icount], dtype=np.float64)
    loc = None
    if edge == 1:
        hit = np.nonzero((seg[:-1] < target) & (seg[1:] >= target))[0]
    else:
        hit = np.nonzero((seg[:-1] > target) & (seg[1:] <= target))[0]
    if len(hit):
        loc = int(hit[0]) + 1
    if loc is None:
        return None
    x = s + loc
    a = float(data[x - 1]) - target
    b = float(data[x]) - target
    y = -a / (-a + b) if (b - a) != 0 else 0.0
    return x - 1 + y


# --------------------------------------------------------------------------
# RFDecode restatement (core.py:253-1046)
# --------------------------------------------------------------------------

class OracleRF:
    """FP64 restatement of ``RFDecode`` (core.py:284-1046) for one system."""

    def __init__(self, system="NTSC", inputfreq=40, blocklen=BLOCKSIZE,
                 decode_analog_audio=True, has_analog_audio=True,
                 mtf_mult=1.0, mtf_offset=0.0, deemp_str=1.0,
                 decoder_params_override=None, pal_v4300d_notch=False, ntsc_color_notch=False):
        # core.py:319-403
        self.system = system
        self.PAL_V4300D_NotchFilter = pal_v4300d_notch     # extra_options, core.py:329
        self.NTSC_ColorNotchFilter = ntsc_color_notch      # extra_options, core.py:328
        self.blocklen = blocklen
        self.blockcut = 1024
        self.blockcut_end = 0
        self.freq = inputfreq
        self.freq_half = inputfreq / 2
        self.freq_hz = inputfreq * 1000000
        self.freq_hz_half = self.freq_hz / 2
        self.mtf_mult = mtf_mult
        self.mtf_offset = mtf_offset
        self.SysParams = sys_params(system)
        self.DecoderParams = filter_params(system)
        for k in ("ire0", "hz_ire", "vsync_ire"):
            self.DecoderParams[k] = self.SysParams[k]
        for k, v in (decoder_params_override or {}).items():
            self.DecoderParams[k] = v
        self.SysParams["analog_audio"] = has_analog_audio
        self.DecoderParams["video_deemp"] = list(self.DecoderParams["video_deemp"])
        self.DecoderParams["video_deemp_strength"] = deemp_str
        self.linelen = int(np.round(self.freq_hz / (1000000.0 / self.SysParams["line_period"])))
        self.decode_analog_audio = decode_analog_audio
        self.computefilters()
        self.blockcut_end = self.Filters["F05_offset"]

    # -- helpers core.py:739-745, :467-478
    def iretohz(self, ire):
        return self.DecoderParams["ire0"] + self.DecoderParams["hz_ire"] * ire

    def hztoire(self, hz):
        return (hz - self.DecoderParams["ire0"]) / self.DecoderParams["hz_ire"]

    def notchrange(self, f, notchwidth, hz=False):
        d = self.freq_hz_half if hz else self.freq_half
        return [(f - notchwidth) / d, (f + notchwidth) / d]

    def computefilters(self):
        """core.py:405-420"""
        self.computevideofilters()
        if self.decode_analog_audio:
            self.computeaudiofilters()
        self.computedelays()

    def build_groupdelay_equalizer(self, lpf_fft):
        """core.py:480-542 -- unit-magnitude all-pass whose group delay is the
        IEC target minus the LPF's own group delay."""
        N = self.blocklen
        fs = self.freq_hz
        binfreq = np.abs(np.fft.fftfreq(N, 1.0 / fs))
        if self.system == "PAL":
            gd_f = np.array([0.0, 0.5e6, 2.0e6, 3.0e6, 4.0e6, 4.4336e6, 4.8e6, 5.5e6])
            gd_t = np.array([0.0, 0.0, 10e-9, 35e-9, 85e-9, 135e-9, 200e-9, 200e-9])
        else:
            gd_f = np.array([0.0, 0.5e6, 2.0e6, 3.0e6, 3.58e6, 4.0e6, 4.2e6, 4.8e6])
            gd_t = np.array([0.0, 0.0, 15e-9, 45e-9, 80e-9, 135e-9, 200e-9, 200e-9])
        target = np.interp(binfreq, gd_f, gd_t)
        phase = np.unwrap(np.angle(lpf_fft))
        lpf_gd = -np.gradient(phase) / (2 * np.pi * (fs / N))
        i05 = np.argmin(np.abs(binfreq - 0.5e6))
        residual = target - (lpf_gd - lpf_gd[i05])
        lpf_freq = self.DecoderParams["video_lpf_freq"]
        t0, t1 = lpf_freq + 0.3e6, lpf_freq + 1.3e6
        taper = np.clip((t1 - binfreq) / (t1 - t0), 0.0, 1.0)
        residual = residual * taper
        residual[binfreq < 0.4e6] = 0.0
        half = N // 2
        dphi = -2 * np.pi * np.cumsum(residual[:half + 1]) * (fs / N)
        eq = np.ones(N, dtype=complex)
        eq[:half + 1] = np.exp(1j * dphi)
        eq[half + 1:] = np.conj(eq[1:half][::-1])
        eq[0] = 1.0
        eq[half] = 1.0
        return eq

    def computevideofilters(self):
        """core.py:544-679"""
        F = self.Filters = {}
        SP, DP, N = self.SysParams, self.DecoderParams, self.blocklen
        F["Frfhpf"] = filtfft(sps.butter(1, 10 / self.freq_half, btype="highpass"), N)
        lo = DP["MTF_freq"] / self.freq_half
        hi = (self.freq_half + (self.freq_half - DP["MTF_freq"])) / self.freq_half
        to_z = lambda pole: polar2z(DP["MTF_poledist"], np.pi * pole)
        F["MTF"] = filtfft(sps.zpk2tf([], [to_z(lo), to_z(hi)], 1), N)
        if "video_bpf_low_order" in DP:
            hp = sps.butter(DP["video_bpf_low_order"], DP["video_bpf_low"] / self.freq_hz_half, btype="highpass")
            lp = sps.butter(DP["video_bpf_high_order"], DP["video_bpf_high"] / self.freq_hz_half, btype="lowpass")
            F["RFVideo"] = filtfft(hp, N) * filtfft(lp, N)
        else:
            bp = sps.butter(DP["video_bpf_order"],
                            [DP["video_bpf_low"] / self.freq_hz_half, DP["video_bpf_high"] / self.freq_hz_half],
                            btype="bandpass")
            F["RFVideo"] = filtfft(bp, N)
        if SP["analog_audio"] and self.system == "NTSC":
            cl = sps.butter(DP["audio_notchorder"], self.notchrange(SP["audio_lfreq"], DP["audio_notchwidth"], True), btype="bandstop")
            cr = sps.butter(DP["audio_notchorder"], self.notchrange(SP["audio_rfreq"], DP["audio_notchwidth"], True), btype="bandstop")
            F["Fcutl"] = filtfft(cl, N)
            F["Fcutr"] = filtfft(cr, N)
            F["RFVideo"] = F["RFVideo"] * F["Fcutl"] * F["Fcutr"]
        F["hilbert"] = build_hilbert(N)
        F["RFVideo"] = F["RFVideo"] * F["hilbert"]
        F["Fvideo_lpf"] = filtfft(sps.butter(DP["video_lpf_order"], DP["video_lpf_freq"] / self.freq_hz_half, "low"), N)
        if self.system == "NTSC" and self.NTSC_ColorNotchFilter:   # core.py:625-631
            notch = sps.butter(3, [DP["video_lpf_freq"] / 1000000 / self.freq_half, 5.0 / self.freq_half], "bandstop")
            F["Fvideo_lpf"] = F["Fvideo_lpf"] * filtfft(notch, N)
        d1, d2 = DP["video_deemp"]
        F["Fdeemp"] = filtfft(emphasis_iir(d1, d2, self.freq_hz), N)
        F["Femp"] = filtfft(emphasis_iir(d2, d1, self.freq_hz), N)
        F["FVideo"] = F["Fvideo_lpf"] * (F["Fdeemp"] ** DP["video_deemp_strength"])
        F["FVideoGD"] = self.build_groupdelay_equalizer(F["Fvideo_lpf"])
        F["FVideo"] = F["FVideo"] * F["FVideoGD"]
        F05 = sps.firwin(65, [0.5 / self.freq_half], pass_zero=True)
        F["F05_offset"] = 32
        F["FVideo05"] = F["Fvideo_lpf"] * F["Fdeemp"] * filtfft((F05, [1.0]), N)
        Fb = sps.firwin(81, self.notchrange(SP["fsc_mhz"], 0.2), pass_zero=False)
        F["FVideoBurst_offset"] = 40
        F["Fburst"] = filtfft((Fb, [1.0]), N)
        F["FVideoBurst"] = F["Fvideo_lpf"] * F["Fdeemp"] * F["Fburst"]
        if self.system == "PAL":
            F["Fpilot"] = filtfft(sps.butter(1, self.notchrange(SP["pilot_mhz"], 0.1), btype="bandpass"), N)
            F["FVideoPilot"] = F["Fvideo_lpf"] * F["Fdeemp"] * F["Fpilot"]

    def computeaudiofilters(self):
        """core.py:681-737 (stage-1 filters; stage 2 is outside the path)."""
        SP, DP = self.SysParams, self.DecoderParams
        self.audio = {}
        for ch, cf in (("left", SP["audio_lfreq"]), ("right", SP["audio_rfreq"])):
            fir = sps.firwin(DP["audio_filterorder"], self.notchrange(cf, DP["audio_filterwidth"], True), pass_zero=False)
            a1 = filtfft([fir, 1.0], self.blocklen)
            lowbin, nbins, a1_freq = fft_determine_slices(cf, 200000, self.freq_hz, self.blocklen)
            a = {"lowbin": lowbin, "nbins": nbins, "a1_freq": a1_freq,
                 "low_freq": self.freq_hz * (lowbin / self.blocklen)}
            a["filt1"] = fft_do_slice(a1, lowbin, nbins, self.blocklen) * build_hilbert(nbins)
            self.audio[ch] = a
            self.Filters["audio_fdiv"] = self.blocklen // nbins

    def mtf_level_total(self, mtf_level):
        """core.py:747-752"""
        return (mtf_level * self.mtf_mult + self.mtf_offset) * self.DecoderParams["MTF_basemult"]

    def demodblock(self, data=None, mtf_level=0, fftdata=None, cut=False):
        """core.py:747-879 (demodblock -> demodblock_cpu), FP64 outputs."""
        return self.demodblock_cpu(data, self.mtf_level_total(mtf_level), fftdata, cut)

    def demodblock_cpu(self, data=None, mtf_level=0, fftdata=None, cut=False):
        """core.py:755-879.  Returns float64 arrays; the reference casts each
        output to float32 when packing its recarray (:821-845)."""
        N, F = self.blocklen, self.Filters
        if fftdata is not None:
            X = fftdata
        elif data is not None:
            X = sfft.fft(np.asarray(data[:N], dtype=np.float64))
        else:
            raise Exception("demodblock called without raw or FFT data")
        rv = {}
        rot = self.delays["video_rot"] if getattr(self, "delays", None) is not None else 0
        hpf = sfft.ifft(X * F["Frfhpf"]).real
        rv["rfhpf"] = hpf[self.blockcut - rot: -self.blockcut_end - rot]
        if self.system == "PAL" and self.PAL_V4300D_NotchFilter:
            X = v4300d_notch(X, N, self.freq)
        Xf = X * F["RFVideo"]
        if mtf_level != 0:
            Xf = Xf * F["MTF"] ** mtf_level
        z = sfft.ifft(Xf)
        demod = unwrap_hilbert(z, self.freq_hz)
        D = sfft.fft(np.clip(demod, 1500000, self.freq_hz * 0.75))
        video = {
            "demod": sfft.ifft(D * F["FVideo"]).real,
            "demod_raw": demod,
            "demod_05": np.roll(sfft.ifft(D * F["FVideo05"]).real, -F["F05_offset"]),
            "demod_burst": np.roll(sfft.ifft(D * F["FVideoBurst"]).real, -F["FVideoBurst_offset"]),
        }
        if self.system == "PAL":
            video["demod_pilot"] = sfft.ifft(D * F["FVideoPilot"]).real
        if cut:
            video = {k: v[self.blockcut:-self.blockcut_end] for k, v in video.items()}
        rv["video"] = video
        if self.decode_analog_audio:
            aud = {}
            for ch in ("left", "right"):
                a = self.audio[ch]
                a1 = sfft.ifft(fft_do_slice(X, a["lowbin"], a["nbins"], N) * a["filt1"])
                aud["audio_" + ch] = unwrap_hilbert(a1, a["a1_freq"]) + a["low_freq"]
            fdiv = N // self.audio["left"]["nbins"]
            if cut:
                aud = {k: v[self.blockcut // fdiv: -self.blockcut_end // fdiv] for k, v in aud.items()}
            rv["audio"] = aud
        rv["spectrum"] = X
        return rv

    def computedelays(self, mtf_level=0):
        """core.py:952-1046 -- demodulate a synthetic line and measure filter delays."""
        fake = np.zeros(self.blocklen, dtype=np.double)
        fake[:] = self.iretohz(0)
        synclen_full = int(4.7 * self.freq)
        fake[1500:1500 + synclen_full] = self.iretohz(self.DecoderParams["vsync_ire"])
        fake[2000:2000 + synclen_full] = self.iretohz(self.DecoderParams["vsync_ire"])
        porch_end = 2000 + synclen_full + int(0.6 * self.freq)
        burst_end = porch_end + int(1.2 * self.freq)
        rate = np.full(burst_end - porch_end, self.SysParams["fsc_mhz"], dtype=np.double)
        fake[porch_end:burst_end] += genwave(rate, self.freq / 2) * self.DecoderParams["hz_ire"] * 20
        fake[3000:3500] = self.iretohz(100)
        fake[4500:5000] = self.iretohz(100)
        rate = np.full(5500 - 4200, self.SysParams["fsc_mhz"], dtype=np.double)
        fake[4200:5500] += genwave(rate, self.freq / 2) * self.DecoderParams["hz_ire"] * 20
        rate = np.full(synclen_full, self.SysParams["fsc_mhz"], dtype=np.double)
        fake[2000:2000 + synclen_full] = self.iretohz(self.DecoderParams["vsync_ire"]) + (
            genwave(rate, self.freq / 2) * self.DecoderParams["hz_ire"] * self.DecoderParams["vsync_ire"])
        F = self.Filters
        emp = sfft.ifft(sfft.fft(fake) * F["Fvideo_lpf"] * F["Femp"]).real
        sig = genwave(emp, self.freq_hz / 2)
        sig *= 4096
        sig += 8192
        sig[6000:6005] = 0
        self.fake_signal = sig
        dec = self.demodblock_cpu(sig, mtf_level=mtf_level)
        vd = dec["video"]["demod"]
        vraw = dec["video"]["demod_raw"]
        self.delays = {
            "video_sync": calczc(vd, 1500, self.iretohz(self.DecoderParams["vsync_ire"] / 2), count=512) - 1500,
            "video_white": calczc(vd, 3000, self.iretohz(50), count=512) - 3000,
        }
        self.delays["video_rot"] = int(np.round(calczc(vd, 6000, self.iretohz(-10), count=512) - 6000))
        self.limits = {"sync": (np.min(vraw[1400:2800]), np.max(vraw[1400:2800])),
                       "viewable": (np.min(vraw[2900:6000]), np.max(vraw[2900:6000]))}
        return dec


def v4300d_notch_window(blocklen, freq_mhz):
    """core.py:785-788: the spectrum bins searched for the LD-V4300D spur."""
    return int(blocklen * (8.42 / freq_mhz)), int(1 + (blocklen * (8.6 / freq_mhz)))


def v4300d_notch(X, blocklen, freq_mhz):
    """core.py:774-798 on a copy of the full spectrum X: bins of the window
    whose magnitude exceeds mean + 3 std (population) of the window are
    zeroed with both neighbours, and so are their mirror bins."""
    X = np.array(X, copy=True)
    lo, hi = v4300d_notch_window(blocklen, freq_mhz)
    mag = np.sqrt(X[lo:hi].real ** 2 + X[lo:hi].imag ** 2)
    m = np.mean(mag) + np.std(mag) * 3
    for i in np.where(mag > m)[0]:
        for k in (i - 1 + lo, i + lo, i + 1 + lo):
            X[k] = 0
            X[blocklen - k] = 0
    return X


# --------------------------------------------------------------------------
# DemodCache stitching (core.py:1073, :1204-1413)
# --------------------------------------------------------------------------

def block_stride(rf):
    """core.py:1073 -- usable samples per block."""
    return rf.blocklen - (rf.blockcut + rf.blockcut_end)


def demod_span(rf, data, begin, length, mtf_level=0):
    """DemodCache.read (core.py:1344-1413) without the threading/LRU: demodulate
    blocks begin//S .. (begin+length)//S (block b reads raw [b*S, b*S+N)) with
    cut=True and concatenate.  Returns (dict of arrays, startloc) where array
    index i corresponds to raw sample startloc + blockcut + i.
    Returns None when a block would run past the end of data (EOF, :1256-1266)."""
    S = block_stride(rf)
    N = rf.blocklen
    end = begin + length
    out = {}
    for b in range(begin // S, end // S + 1):
        raw = data[b * S: b * S + N]
        if len(raw) < N:
            return None
        r = rf.demodblock(data=raw, mtf_level=mtf_level, cut=True)
        parts = dict(r["video"])
        parts["rfhpf"] = r["rfhpf"]
        if "audio" in r:
            parts.update(r["audio"])
        for k, v in parts.items():
            out.setdefault(k, []).append(v)
    return {k: np.concatenate(v) for k, v in out.items()}, (begin // S) * S
