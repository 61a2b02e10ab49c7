"""Host-side FP64 filter design for the GPU demodulator.

Every filter is a complex frequency response sampled at the ``blocklen`` bins
of the block FFT, exactly as the reference builds them with
``scipy.signal.freqz(b, a, blocklen, whole=True)`` (``lddecode/utils.py:782``).
Because the GPU transforms real blocks, only bins ``0..blocklen/2`` are
uploaded; for a real-coefficient filter the other half is the conjugate mirror.

Design formulas follow ``RFDecode.computevideofilters`` (core.py:544-679),
``build_groupdelay_equalizer`` (core.py:480-542), ``computeaudiofilters``
(core.py:681-737) and the helpers ``emphasis_iir`` (utils.py:766-778),
``build_hilbert`` (utils.py:882-890) and ``fft_determine_slices``
(utils.py:929-950).  scipy.signal is the reference's own design library and is
used here on the host only; no sample data ever passes through it.
"""

import math

import numpy as np
import scipy.signal as sps


def freq_response(b, a, n):
    """Whole-circle response of b/a at n bins (utils.py:782-783)."""
    return sps.freqz(b, a, n, whole=True)[1]


def emphasis(t_zero, t_pole, fs):
    """First-order (de)emphasis: zero at 1/t_zero, pole at 1/t_pole rad/s,
    bilinear-mapped with prewarping (utils.py:766-778)."""
    wz = 2 * fs * np.tan((1 / t_zero) / (2 * fs))
    wp = 2 * fs * np.tan((1 / t_pole) / (2 * fs))
    b, a = sps.zpk2tf([-wz], [-wp], wp / wz)
    return sps.bilinear(b, a, fs)


def analytic_mask(n):
    """Weights that keep DC and Nyquist once and double the positive bins
    (utils.py:882-890)."""
    if n % 2:
        raise Exception("build_hilbert: must have even fft_size")
    m = np.zeros(n)
    m[0] = m[n // 2] = 1
    m[1:n // 2] = 2
    return m


def audio_bins(center, min_bw, freq_hz, nbins_in):
    """(lowbin, nbins, decimated rate) covering center +- min_bw (utils.py:929-950)."""
    bw = freq_hz / nbins_in
    cbin = int(np.round(center / bw))
    half = int(np.round(min_bw / bw))
    nbins = 2 * (2 ** math.ceil(math.log2(half * 2)))
    return cbin - nbins // 4, nbins, bw * nbins


def audio2_filter(a1_freq, n):
    """Stage-2 audio filter at the stage-1 rate a1_freq (core.py:727-734): the
    lowest-order Butterworth low-pass with at most 1 dB ripple to 20 kHz and
    9 dB attenuation from 24 kHz, times the 75 us / 5.3 us de-emphasis; whole
    circle at n bins."""
    order, wn = sps.buttord(20000 / (a1_freq / 2), 24000 / (a1_freq / 2), 1, 9)
    return freq_response(*sps.butter(order, wn), n) * freq_response(*emphasis(5.3e-6, 75e-6, a1_freq), n)


def hermitian_half(h):
    """The half spectrum (bins 0..n/2) that a real inverse transform applies
    for np.fft.ifft(X * h).real with X a real signal's spectrum: the
    Hermitian part (h[k] + conj(h[n-k])) / 2.  For a one-sided table (the
    EFM equaliser) that is half of it; the stage-2 audio low-pass in b/a
    form is conjugate-symmetric only to ~1e-4."""
    n = len(h)
    k = np.arange(n // 2 + 1)
    return 0.5 * (h[k] + np.conj(h[(-k) % n]))


audio2_half = hermitian_half


def supergauss(x, freq, order=1, centerfreq=0):
    """Super-Gaussian window (utils.py:870-877)."""
    return np.exp(-2 * np.power((2 * (x - centerfreq) * (math.log(2.0) / 2.0) ** (1 / (2 * order))) / freq,
                                2 * order))


def efm_filter(freq_hz, n):
    """The EFM equaliser Fefm (core.py:420-465): cubic interpolation of the
    11 amplitude / phase knots from 0 to 1.9 MHz onto the bins below 1.9 MHz
    (one-sided), times 8 and a 60th-order super-Gaussian 20 kHz - 1.6 MHz
    band-pass that the reference builds for 32768 bins."""
    import scipy.interpolate as spi
    knots = np.linspace(0.0e6, 1.9e6, num=11)
    amp = np.array([0.0, 0.215, 0.41, 0.73, 0.98, 1.03, 0.99, 0.81, 0.59, 0.42, 0.0])
    phase = [p * 1.25 for p in (0.0, -0.92, -1.03, -1.11, -1.2, -1.2, -1.2, -1.2, -1.05, -0.95, -0.8)]
    per_bin = freq_hz / n
    nz = int(knots[-1] / per_bin) + 1
    f = np.arange(nz) * per_bin
    a = spi.interp1d(knots, amp, kind="cubic")(f)
    p = spi.interp1d(knots, phase, kind="cubic")(f)
    coeffs = np.zeros(n, dtype=complex)
    coeffs[:nz] = a * (np.cos(p) + (complex(0, -1) * np.sin(p)))
    bl = 32768
    sg = supergauss(np.linspace(0, 20000000, bl // 2 + 1), 1600000 - 20000, 60, (1600000 + 20000) / 2.0)[:-1]
    return coeffs * 8 * np.concatenate([sg, np.flip(sg)])


def gd_equaliser(system, lpf, fs, lpf_freq):
    """Unit-magnitude all-pass whose group delay is the IEC 60856/60857 target
    curve minus the video LPF's own group delay (core.py:480-542)."""
    n = len(lpf)
    f = np.abs(np.fft.fftfreq(n, 1.0 / fs))
    if system == "PAL":
        knots_f = [0.0, 0.5e6, 2.0e6, 3.0e6, 4.0e6, 4.4336e6, 4.8e6, 5.5e6]
        knots_t = [0.0, 0.0, 10e-9, 35e-9, 85e-9, 135e-9, 200e-9, 200e-9]
    else:
        knots_f = [0.0, 0.5e6, 2.0e6, 3.0e6, 3.58e6, 4.0e6, 4.2e6, 4.8e6]
        knots_t = [0.0, 0.0, 15e-9, 45e-9, 80e-9, 135e-9, 200e-9, 200e-9]
    target = np.interp(f, np.array(knots_f), np.array(knots_t))
    gd = -np.gradient(np.unwrap(np.angle(lpf))) / (2 * np.pi * (fs / n))
    ref_bin = np.argmin(np.abs(f - 0.5e6))
    resid = target - (gd - gd[ref_bin])
    lo, hi = lpf_freq + 0.3e6, lpf_freq + 1.3e6
    resid = resid * np.clip((hi - f) / (hi - lo), 0.0, 1.0)
    resid[f < 0.4e6] = 0.0
    h = n // 2
    phi = -2 * np.pi * np.cumsum(resid[:h + 1]) * (fs / n)
    eq = np.ones(n, dtype=complex)
    eq[:h + 1] = np.exp(1j * phi)
    eq[h + 1:] = np.conj(eq[1:h][::-1])
    eq[0] = 1.0
    eq[h] = 1.0
    return eq


class FilterBank:
    """All frequency-domain filters of one (system, blocklen, parameters) set.

    Attributes are full-length complex128 arrays named as in the reference's
    ``RFDecode.Filters`` dict (``Frfhpf``, ``RFVideo`` (analytic mask folded in),
    ``MTF``, ``Fvideo_lpf``, ``Fdeemp``, ``Femp``, ``FVideo``, ``FVideo05``,
    ``FVideoBurst``, ``FVideoPilot``) plus per-channel audio stage-1 data.
    """

    def __init__(self, system, blocklen, freq_hz, sysp, decp, analog_audio=True, color_notch=False):
        self.system = system
        self.blocklen = n = blocklen
        self.freq_hz = fs = freq_hz
        nyq_mhz = fs / 2e6
        nyq = fs / 2
        F = self.filters = {}
        F["Frfhpf"] = freq_response(*sps.butter(1, 10 / nyq_mhz, btype="highpass"), n)
        # MTF: two poles mirrored about fs/4 (core.py:557-567)
        p_lo = decp["MTF_freq"] / nyq_mhz
        p_hi = (nyq_mhz + (nyq_mhz - decp["MTF_freq"])) / nyq_mhz
        poles = [decp["MTF_poledist"] * np.exp(1j * np.pi * p) for p in (p_lo, p_hi)]
        F["MTF"] = freq_response(*sps.zpk2tf([], poles, 1), n)
        if "video_bpf_low_order" in decp:
            hp = sps.butter(decp["video_bpf_low_order"], decp["video_bpf_low"] / nyq, btype="highpass")
            lp = sps.butter(decp["video_bpf_high_order"], decp["video_bpf_high"] / nyq, btype="lowpass")
            rfv = freq_response(*hp, n) * freq_response(*lp, n)
        else:
            bp = sps.butter(decp["video_bpf_order"],
                            [decp["video_bpf_low"] / nyq, decp["video_bpf_high"] / nyq], btype="bandpass")
            rfv = freq_response(*bp, n)
        if sysp["analog_audio"] and system == "NTSC":
            for key in ("audio_lfreq", "audio_rfreq"):
                c, w = sysp[key], decp["audio_notchwidth"]
                notch = sps.butter(decp["audio_notchorder"], [(c - w) / nyq, (c + w) / nyq], btype="bandstop")
                rfv = rfv * freq_response(*notch, n)
        F["RFVideo"] = rfv * analytic_mask(n)
        F["Fvideo_lpf"] = freq_response(*sps.butter(decp["video_lpf_order"], decp["video_lpf_freq"] / nyq, "low"), n)
        if system == "NTSC" and color_notch:
            # NTSC_ColorNotchFilter (core.py:625-631): 3rd-order band-stop from
            # the video LPF corner to 5 MHz, folded into every post-demod filter
            stop = sps.butter(3, [decp["video_lpf_freq"] / 1e6 / nyq_mhz, 5.0 / nyq_mhz], "bandstop")
            F["Fvideo_lpf"] = F["Fvideo_lpf"] * freq_response(*stop, n)
        t1, t2 = decp["video_deemp"]
        F["Fdeemp"] = freq_response(*emphasis(t1, t2, fs), n)
        F["Femp"] = freq_response(*emphasis(t2, t1, fs), n)
        strength = decp.get("video_deemp_strength", 1.0)
        F["FVideo"] = (F["Fvideo_lpf"] * (F["Fdeemp"] ** strength)
                       * gd_equaliser(system, F["Fvideo_lpf"], fs, decp["video_lpf_freq"]))
        lpf_deemp = F["Fvideo_lpf"] * F["Fdeemp"]
        fir05 = sps.firwin(65, [0.5 / nyq_mhz], pass_zero=True)
        F["FVideo05"] = lpf_deemp * freq_response(fir05, [1.0], n)
        fsc = sysp["fsc_mhz"]
        firb = sps.firwin(81, [(fsc - 0.2) / nyq_mhz, (fsc + 0.2) / nyq_mhz], pass_zero=False)
        F["FVideoBurst"] = lpf_deemp * freq_response(firb, [1.0], n)
        if system == "PAL":
            pil = sysp["pilot_mhz"]
            bp = sps.butter(1, [(pil - 0.1) / nyq_mhz, (pil + 0.1) / nyq_mhz], btype="bandpass")
            F["FVideoPilot"] = lpf_deemp * freq_response(*bp, n)
        # sample shifts applied to the 0.5 MHz and burst outputs (core.py:659, :667)
        self.offset05 = 32
        self.offset_burst = 40

        # analogue audio stage 1 (core.py:681-737)
        self.audio = {}
        if analog_audio:
            for ch, key in (("left", "audio_lfreq"), ("right", "audio_rfreq")):
                c, w = sysp[key], decp["audio_filterwidth"]
                fir = sps.firwin(decp["audio_filterorder"], [(c - w) / nyq, (c + w) / nyq], pass_zero=False)
                resp = freq_response(fir, 1.0, n)
                lowbin, nb, rate = audio_bins(c, 200000, fs, n)
                h = nb // 2
                sl = np.concatenate([resp[lowbin:lowbin + h], resp[n - lowbin - h:n - lowbin]])
                self.audio[ch] = {"lowbin": lowbin, "nbins": nb, "a1_freq": rate,
                                  "low_freq": fs * (lowbin / n),
                                  "filt1": sl * analytic_mask(nb),
                                  "audio2_filter": audio2_filter(rate, n),
                                  "center": sysp[key]}

    def half(self, name):
        """Bins 0..blocklen/2 of a filter."""
        return self.filters[name][: self.blocklen // 2 + 1]
