"""FP64 restatement of ld-decode's field resampler (TBC) and hsync refinement
(test oracle only).

Pinned by (tests/test_oracle_tbc.py):
  * build_kaiser_lut: rows of the reference's own data file
    ``lddecode/sinc_lut.npz`` (fixture ``tests/golden/sinc_lut_rows.npz``),
    plus closed forms (rows sum to 1, phase 0 is the unit impulse at tap 7).
  * computewow_scaled: is the reference's own scipy call (core.py:2613);
    the closed form the GPU kernel uses (lerp on segment p // outlinelen,
    slope of the right-hand segment at knots) is checked against it.
  * scale_field: brute force at small sizes against a direct per-pixel
    evaluation of the Kaiser-windowed sinc, and the identity that median/MAD
    over per-pixel wow equals median/MAD over per-line slopes.
  * hz_to_output_array: closed form (1 LSB = out_scale/hz_ire Hz).
  * refine_linelocs_hsync: generator ground truth (line starts known to
    1e-9 samples) on synthetic fields.
"""

import numpy as np
import scipy.interpolate as spi
from scipy.special import i0

from .dsp import calczc

SINC_TAPS = 16        # utils.py:76
SINC_PHASES = 2 ** 16  # utils.py:77
KAISER_BETA = 5.0     # utils.py:75 (commented default); matches sinc_lut.npz


def build_kaiser_lut(beta=KAISER_BETA, taps=SINC_TAPS, phases=SINC_PHASES):
    """utils.py:98-124 (the commented generator of sinc_lut.npz).  Row p holds
    the normalised Kaiser-windowed sinc weights for fractional phase p/phases;
    tap t weighs sample int(coord) - (taps/2 - 1) + t, at distance
    (taps/2 - 1 - t) + phase.  Row ``phases`` repeats row ``phases-1``."""
    a = taps // 2
    offsets = np.arange(a - 1, -a - 1, -1, dtype=np.float64)
    phase = np.arange(phases, dtype=np.float64) / phases
    x = offsets[None, :] + phase[:, None]
    r = x / a
    win = np.where(np.abs(r) <= 1.0, i0(beta * np.sqrt(np.clip(1.0 - r * r, 0.0, None))) / i0(beta), 0.0)
    w = np.sinc(x) * win
    # the reference stores each weight as float32, sums the unrounded float64
    # weights, and divides the float32 array by that Python-float sum, which
    # numpy evaluates in float32 (weak scalar): reproduced bit-exactly
    w32 = w.astype(np.float32) / w.sum(axis=1, keepdims=True).astype(np.float32)
    table = np.empty((phases + 1, taps), dtype=np.float32)
    table[:phases] = w32
    table[phases] = table[phases - 1]
    return table


def computewow_scaled(linelocs, inlinelen, outlinelen, outlinecount, lineoffset):
    """core.py:2591-2623 with wow_interpolation_method='linear': the reference's
    own call, scipy.interpolate.make_interp_spline(k=1) through
    (i*inlinelen, linelocs[i]), evaluated (value and first derivative) at
    x_p = p*inlinelen/outlinelen for p < (outlinecount + lineoffset + 1)*outlinelen.
    Returns (interpolated_pixel_locs, wowfactors), float64."""
    ll = np.asarray(linelocs, dtype=np.float64)
    expected = np.array([i * inlinelen for i in range(len(ll))], dtype=np.float64)
    spl = spi.make_interp_spline(expected, ll, k=1, bc_type=None, check_finite=False)
    outscale = inlinelen / outlinelen
    x = np.arange(outlinecount * outlinelen + (lineoffset + 1) * outlinelen) * outscale
    return spl(x), spl(x, 1)


def level_adjusts(wowfactors, outwidth, smoothing=0.0, threshold=15):
    """utils.py:131-149 -- outlier-clamped wow (|w-median| > threshold*MAD ->
    median), then an optional per-pixel exponential moving average."""
    w = np.asarray(wowfactors, dtype=np.float64)
    med = np.median(w)
    mad = np.median(np.abs(w - med))
    thr = threshold * mad if mad > 0 else 0.001
    la = np.where(np.abs(w - med) > thr, med, w)
    if smoothing > 0:
        alpha = 1 / (smoothing * outwidth)
        la = la.copy()
        for i in range(1, len(la)):
            la[i] = alpha * la[i] + (1 - alpha) * la[i - 1]
    return la


def scale_field(buf, pixel_locs, wowfactors, lut, lineoffset, outwidth, nout,
                smoothing=0.0, threshold=15, coord_f32=True):
    """utils.py:127-183 -- 16-tap LUT resampler.  Returns float64 dsout[nout].

    coord_f32=True reproduces the reference's float32 cast of the absolute
    source coordinate (utils.py:160); False keeps it in float64."""
    buf = np.asarray(buf, dtype=np.float32)
    la = level_adjusts(wowfactors, outwidth, smoothing, threshold)
    start = outwidth * (lineoffset + 1)
    idx = np.arange(start, start + nout)
    if coord_f32:
        coord = pixel_locs[idx].astype(np.float32)
        ci = coord.astype(np.int64)
        frac = coord - ci.astype(np.float32)                      # float32, exact
        ph = frac * np.float32(SINC_PHASES)                       # float32, exact
        p0 = ph.astype(np.int64)
        alpha = (ph - p0.astype(np.float32)).astype(np.float32)
    else:
        coord = pixel_locs[idx]
        ci = np.floor(coord).astype(np.int64)
        frac = coord - ci
        ph = frac * SINC_PHASES
        p0 = ph.astype(np.int64)
        alpha = (ph - p0).astype(np.float32)
    w0 = lut[p0].astype(np.float32)
    w1 = lut[p0 + 1].astype(np.float32)
    # ws + alpha*(we - ws) in float32 with one rounding of the multiply-add
    # (fastmath contracts it to an FMA); float64 holds alpha*d + ws exactly
    d = (w1 - w0).astype(np.float32)
    w = (w0.astype(np.float64) + alpha[:, None].astype(np.float64) * d.astype(np.float64)).astype(np.float32)
    taps = ci[:, None] - (SINC_TAPS // 2 - 1) + np.arange(SINC_TAPS)[None, :]
    # buf[start+t] * weight is a float32 product (float32 operands in the
    # reference's numba loop, utils.py:179-181); the running sum is float64,
    # accumulated tap by tap
    prod = (buf[taps] * w).astype(np.float32).astype(np.float64)
    res = np.zeros(len(idx))
    for t in range(SINC_TAPS):
        res += prod[:, t]
    return la[idx] * res


def hz_to_output_array(x, ire0, hz_ire, output_zero, vsync_ire, out_scale):
    """utils.py:1061-1075 -- Hz -> uint16 code, round half up, clipped."""
    scale = out_scale / hz_ire
    offset = output_zero - vsync_ire * out_scale - ire0 * scale
    v = np.asarray(x, dtype=np.float64) * scale + offset + 0.5
    return np.clip(v, 0, 65535).astype(np.uint16)


def compute_deriv_error(linelocs, baserr):
    """core.py:2553-2564"""
    ll = np.asarray(linelocs, dtype=np.float64)
    dd = np.abs(np.diff(np.diff(ll))) > 4
    d1 = np.zeros(len(ll), bool)
    d1[1:-1] = dd
    d2 = np.zeros(len(ll), bool)
    d2[2:] = dd
    return np.asarray(baserr, bool) | d1 | d2


def refine_linelocs_hsync(demod_05, linelocs1, linebad, system, freq, iretohz, vsync_ire):
    """core.py:2484-2551 -- per line, sub-sample hsync leading edge on demod_05.

    iretohz is the DecoderParams IRE->Hz map.  Returns (linelocs2, linebad)."""
    d = np.asarray(demod_05)
    ll1 = np.asarray(linelocs1, dtype=np.float64)
    bad = np.array(linebad, dtype=bool)
    ll2 = ll1.copy()
    for i in range(len(ll1)):
        if 3 <= i <= 6 or (system == "PAL" and 1 <= i <= 2):
            bad[i] = True
            continue
        start = ll1[i] - freq
        zc = calczc(d, start, iretohz(vsync_ire / 2), count=freq * 2)
        if zc is not None and not bad[i]:
            ll2[i] = zc
            area = d[int(zc - freq * 0.75):int(zc + freq * 8)]
            if np.min(area) < iretohz(-55) or np.max(area) > iretohz(30):
                bad[i] = True
                ll2[i] = ll1[i]
            else:
                porch = _median_f32(d[int(zc + freq * 8):int(zc + freq * 9)])
                sync = _median_f32(d[int(zc + freq * 1):int(zc + freq * 2.5)])
                zc2 = calczc(d, start, (porch + sync) / 2, count=400)
                if zc2 is not None and np.abs(zc2 - zc) < (freq / 2):
                    ll2[i] = zc2
                else:
                    bad[i] = True
        else:
            bad[i] = True
        if bad[i]:
            ll2[i] = ll1[i]
    return ll2, bad


def _median_f32(a):
    """nb_median (utils.py:1233-1235) over a float32 slice: numba's median
    averages the two middle values as 0.5*(a+b) with a+b in float32."""
    s = np.sort(np.asarray(a, dtype=np.float32))
    n = len(s)
    if n % 2:
        return float(s[n // 2])
    return 0.5 * float(np.float32(s[n // 2 - 1] + s[n // 2]))
