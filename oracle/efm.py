"""FP64 restatement of ld-decode's EFM (digital audio) front-end: the
equaliser filter of RFDecode.computeefmfilter (core.py:420-465, with
gen_bpf_supergauss / supergauss utils.py:859-877) and the per-block output
of demodblock (core.py:847-851).  Test infrastructure only.

Kept from the reference: the 11 equaliser knots (0 to 1.9 MHz, amplitude
and 1.25 x phase) interpolated with scipy's cubic interp1d onto the bins
below 1.9 MHz, scaled by 8 and by a 60th-order super-Gaussian band-pass
built for 32768 bins whatever the block length (core.py:465); the output
is np.int16(np.clip(ifft(X * Fefm).real, -32768, 32767)) cut like the
video outputs.
"""

import math

import numpy as np
import scipy.interpolate as spi


def supergauss(x, freq, order=1, centerfreq=0):
    """utils.py:870-877."""
    return np.exp(-2 * np.power((2 * (x - centerfreq) * (math.log(2.0) / 2.0) ** (1 / (2 * order))) / freq,
                                2 * order))


def gen_bpf_supergauss(freq_low, freq_high, order, nyquist_hz, block_len):
    """utils.py:859-867."""
    sg = supergauss(np.linspace(0, nyquist_hz, block_len // 2 + 1), freq_high - freq_low, order,
                    (freq_high + freq_low) / 2.0)[:-1]
    return np.concatenate([sg, np.flip(sg)])


KNOT_FREQS = np.linspace(0.0e6, 1.9e6, num=11)
KNOT_AMP = np.array([0.0, 0.215, 0.41, 0.73, 0.98, 1.03, 0.99, 0.81, 0.59, 0.42, 0.0])
KNOT_PHASE = np.array([0.0, -0.92, -1.03, -1.11, -1.2, -1.2, -1.2, -1.2, -1.05, -0.95, -0.8]) * 1.25


def efm_filter(freq_hz, blocklen):
    """Filters['Fefm'] (core.py:420-465)."""
    freq_per_bin = freq_hz / blocklen
    coeffs = np.zeros(blocklen, dtype=complex)
    a_interp = spi.interp1d(KNOT_FREQS, KNOT_AMP, kind="cubic")
    p_interp = spi.interp1d(KNOT_FREQS, list(KNOT_PHASE), kind="cubic")
    nonzero_bins = int(KNOT_FREQS[-1] / freq_per_bin) + 1
    bin_freqs = np.arange(nonzero_bins) * freq_per_bin
    amp, phase = a_interp(bin_freqs), p_interp(bin_freqs)
    coeffs[:nonzero_bins] = amp * (np.cos(phase) + (complex(0, -1) * np.sin(phase)))
    return coeffs * 8 * gen_bpf_supergauss(20000, 1600000, 60, 20000000, 32768)


def efm_block(X, filt, blockcut=1024, blockcut_end=32, cut=True):
    """demodblock's EFM output (core.py:847-851) from the block spectrum."""
    out = np.fft.ifft(X * filt)
    if cut:
        out = out[blockcut:-blockcut_end]
    return np.int16(np.clip(out.real, -32768, 32767))
