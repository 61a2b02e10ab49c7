"""Pins for the FP64 oracle (CPU): reference data files, closed forms and
library routines the reference itself calls.  No test here retypes the
oracle's formula; each compares against something independent of it."""

import numpy as np
import pytest
import scipy.interpolate as spi
import scipy.signal as sps

from oracle import dsp as O
from oracle import tbc as T
from oracle.params import filter_params, out_scale, sys_params

GOLDEN = __import__("pathlib").Path(__file__).with_name("golden")


# ------------------------------------------------------------------ params
def test_system_geometry_closed_forms():
    ntsc, pal = sys_params("NTSC"), sys_params("PAL")
    assert ntsc["line_period"] == pytest.approx(63.5555555, abs=1e-6)   # 227.5 / (315/88)
    assert ntsc["outlinelen"] == 910 and pal["outlinelen"] == 1135       # 4 fsc per line
    assert pal["vsync_ire"] == pytest.approx(-300 / 7)
    assert out_scale("NTSC", -40) == pytest.approx((0xC800 - 0x0400) / 140)
    assert out_scale("PAL", pal["vsync_ire"]) == pytest.approx((0xD300 - 0x0100) / (100 + 300 / 7))
    rf = O.OracleRF("NTSC", decode_analog_audio=False)
    assert rf.linelen == 2542
    assert O.OracleRF("PAL", decode_analog_audio=False).linelen == 2560


# ------------------------------------------------------------------ LUT
def test_sinc_lut_matches_reference_data_file():
    """tests/golden/sinc_lut_rows.npz holds rows of lddecode/sinc_lut.npz."""
    g = np.load(GOLDEN / "sinc_lut_rows.npz")
    lut = T.build_kaiser_lut()
    assert tuple(g["shape"]) == lut.shape
    np.testing.assert_array_equal(lut[g["rows"]], g["values"])


def test_sinc_lut_closed_forms():
    lut = T.build_kaiser_lut()
    assert lut[0, 7] == 1.0 and np.all(np.abs(np.delete(lut[0], 7)) < 1e-15)
    np.testing.assert_allclose(lut[:-1].astype(np.float64).sum(axis=1), 1.0, atol=2e-6)
    # half phase is symmetric about the tap pair (7, 8)
    np.testing.assert_allclose(lut[32768], lut[32768][::-1], atol=2e-7)


# ------------------------------------------------------------------ filters
@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_filter_closed_forms(system):
    rf = O.OracleRF(system)
    F, N = rf.Filters, rf.blocklen
    w = 2 * np.pi * np.arange(N // 2 + 1) / N
    dp = filter_params(system)
    # Butterworth low-pass through the bilinear map: |H|^2 = 1/(1+(tan(w/2)/tan(wc/2))^2n)
    wc = np.pi * dp["video_lpf_freq"] / (rf.freq_hz / 2)
    mag2 = 1 / (1 + (np.tan(w / 2) / np.tan(wc / 2)) ** (2 * dp["video_lpf_order"]))
    np.testing.assert_allclose(np.abs(F["Fvideo_lpf"][: N // 2 + 1]) ** 2, mag2, atol=1e-9)
    # emphasis and de-emphasis are exact inverses, unit gain at DC
    np.testing.assert_allclose(F["Fdeemp"] * F["Femp"], 1.0, atol=1e-12)
    assert abs(F["Fdeemp"][0] - 1) < 1e-12
    # analytic mask: nothing above Nyquist, DC and Nyquist kept once
    assert np.all(F["RFVideo"][N // 2 + 1:] == 0)
    # group-delay equaliser is unit-magnitude and conjugate-symmetric
    eq = F["FVideoGD"]
    np.testing.assert_allclose(np.abs(eq), 1.0, atol=1e-12)
    np.testing.assert_allclose(eq[1:N // 2], np.conj(eq[N - 1:N // 2:-1]), atol=1e-12)
    # windowed-FIR low-pass is normalised to unit DC gain (firwin scale=True)
    assert F["FVideo05"][0] == pytest.approx(1.0, abs=1e-9)
    assert F["FVideo"][0] == pytest.approx(1.0, abs=1e-9)
    # band-pass FIR normalised at its centre frequency fsc
    kc = int(round(rf.SysParams["fsc_mhz"] * 1e6 / rf.freq_hz * N))
    assert abs(F["Fburst"][kc]) == pytest.approx(1.0, abs=2e-3)
    # RF high-pass: first-order Butterworth, |H|^2 = 1/(1+(tan(wc/2)/tan(w/2))^2)
    wc = np.pi * 10 / 20
    k = np.arange(1, N // 2)
    mag2 = 1 / (1 + (np.tan(wc / 2) / np.tan(w[k] / 2)) ** 2)
    np.testing.assert_allclose(np.abs(F["Frfhpf"][k]) ** 2, mag2, atol=1e-9)


def test_audio_slices_match_reference_values():
    """Bin slices computed with fft_determine_slices (utils.py:929-950)."""
    a = O.OracleRF("NTSC").audio
    assert (a["left"]["lowbin"], a["right"]["lowbin"]) == (1629, 2048)
    p = O.OracleRF("PAL").audio
    assert (p["left"]["lowbin"], p["right"]["lowbin"]) == (304, 618)
    for ch in (a["left"], a["right"], p["left"], p["right"]):
        assert ch["nbins"] == 1024 and ch["a1_freq"] == 1.25e6


# ------------------------------------------------------------------ DSP
def test_fft_matches_brute_force_dft():
    rng = np.random.default_rng(0)
    x = rng.normal(size=64)
    n = np.arange(64)
    dft = np.exp(-2j * np.pi * np.outer(n, n) / 64) @ x
    import scipy.fft as sfft
    np.testing.assert_allclose(sfft.fft(x), dft, atol=1e-12)


def test_unwrap_hilbert_known_answer():
    fs = 40e6
    n = np.arange(4096)
    for f in (7.6e6, 9.3e6, 19.9e6, -3e6):
        z = np.exp(2j * np.pi * f * n / fs)
        out = O.unwrap_hilbert(z, fs)
        assert out[0] == 0.0
        np.testing.assert_allclose(out[1:], f % fs, atol=1e-4)


def test_calczc_linear_ramp():
    d = np.arange(100, dtype=np.float64) * 2.0          # crosses 37.3 at x = 18.65
    assert O.calczc(d, 5, 37.3, count=40) == pytest.approx(18.65)
    assert O.calczc(d, 5, 1000.0, count=40) is None
    assert O.calczc(d[::-1].copy(), 5, 37.3, count=200) == pytest.approx(99 - 18.65)


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_demodblock_constant_carrier(system):
    """A constant FM carrier demodulates to its frequency on every branch
    (unit DC gain), burst/pilot to freq*H(0); rfhpf of a pure tone stays a tone."""
    rf = O.OracleRF(system)
    # a bin-centred carrier near 50 IRE: periodic in the block, so the
    # circular (overlap-save) filtering has no edge transient
    f = round(rf.iretohz(50) * rf.blocklen / rf.freq_hz) * rf.freq_hz / rf.blocklen
    n = np.arange(rf.blocklen)
    x = 8192 * np.cos(2 * np.pi * f * n / rf.freq_hz)
    r = rf.demodblock(data=x, cut=True)
    v = r["video"]
    np.testing.assert_allclose(v["demod_raw"], f, rtol=1e-9)
    # demod_raw[0] = 0 by definition (utils.py:918) and clips to 1.5 MHz; its
    # filtered response wraps onto the block end, so compare the interior
    inner = slice(0, -600)
    np.testing.assert_allclose(v["demod"][inner], f, rtol=1e-6)
    np.testing.assert_allclose(v["demod_05"][inner], f, rtol=1e-6)
    np.testing.assert_allclose(v["demod_burst"][inner], f * rf.Filters["FVideoBurst"][0].real, atol=5.0)
    assert len(v["demod"]) == rf.blocklen - 1024 - 32


def test_roll_and_cut_index_algebra():
    """cut=True outputs are exactly samples [1024, N-32) of the uncut ones
    (core.py:839-845), for every channel including the rolled ones."""
    rf = O.OracleRF("NTSC")
    n = np.arange(rf.blocklen)
    f = np.where(n < 16000, rf.iretohz(0), rf.iretohz(100))
    phase = np.concatenate([[0.0], np.cumsum(2 * np.pi * f[:-1] / rf.freq_hz)])
    x = 8192 * np.cos(phase)
    cut = rf.demodblock(data=x, cut=True)["video"]
    full = rf.demodblock(data=x, cut=False)["video"]
    for k in ("demod", "demod_05", "demod_burst", "demod_raw"):
        np.testing.assert_array_equal(cut[k], full[k][1024:-32])


def test_stitching_equals_per_block():
    rf = O.OracleRF("NTSC", decode_analog_audio=False)
    rng = np.random.default_rng(2)
    S = O.block_stride(rf)
    data = rng.integers(-2000, 2000, size=3 * S + rf.blocklen).astype(np.float64)
    out, start = O.demod_span(rf, data, 10, 2 * S)
    assert start == 0
    for b in range(3):
        r = rf.demodblock(data=data[b * S: b * S + rf.blocklen], cut=True)
        np.testing.assert_array_equal(out["demod"][b * S:(b + 1) * S], r["video"]["demod"])
    assert O.demod_span(rf, data[: 2 * S], 0, 2 * S) is None    # EOF


# ------------------------------------------------------------------ TBC
def _random_linelocs(rng, n, linelen=2542.0, jitter=2.0):
    return 3000 + np.arange(n) * (linelen + 0.2222) + rng.uniform(-jitter, jitter, n)


@pytest.mark.parametrize("inl,W,lines,lo", [(2542, 910, 263, 0), (2560, 1135, 313, 2), (2560, 1135, 313, 3)])
def test_computewow_is_segment_lerp(inl, W, lines, lo):
    """The k=1 spline equals the lerp on segment j = p // outlinelen, with the
    derivative taken on the right-hand segment at every line start."""
    rng = np.random.default_rng(3)
    ll = _random_linelocs(rng, lines + lo + 10, linelen=inl)
    locs, wow = T.computewow_scaled(ll, inl, W, lines, lo)
    p = np.arange(len(locs))
    j = p // W
    slope = np.diff(ll) / inl
    np.testing.assert_allclose(locs, ll[j] + (p * (inl / W) - j * inl) * slope[j], atol=1e-9, rtol=0)
    np.testing.assert_allclose(wow, slope[j], atol=1e-12, rtol=0)


def test_level_adjust_median_identity():
    """median/MAD over per-pixel wow (outwidth repeats per line) equal the
    per-line statistics, and outliers are replaced by the median."""
    rng = np.random.default_rng(4)
    slopes = 1 + rng.normal(0, 1e-4, 264)
    slopes[[10, 100]] = [1.5, 0.2]
    wow = np.repeat(slopes, 910)
    la = T.level_adjusts(wow, 910)
    med = np.median(slopes)
    assert np.median(wow) == med
    assert la[10 * 910] == med and la[100 * 910] == med
    assert la[11 * 910] == slopes[11]


def test_scale_field_against_direct_kaiser_sinc():
    """FP64-coordinate resampling equals direct evaluation of the normalised
    Kaiser-windowed sinc at each pixel's fractional position (the LUT's
    linear phase interpolation error is far below 1e-5 of the signal)."""
    from scipy.special import i0
    rng = np.random.default_rng(5)
    buf = rng.normal(0, 1e5, 4000).astype(np.float32) + 8e6
    ll = 100 + np.arange(4) * 1000.37
    locs, wow = T.computewow_scaled(ll, 1000, 400, 1, 0)
    lut = T.build_kaiser_lut()
    out = T.scale_field(buf, locs, np.ones_like(wow), lut, 0, 400, 400, coord_f32=False)
    for q in rng.integers(0, 400, 20):
        c = locs[400 + q]
        ci = int(c)
        d = np.arange(7, -9, -1) + (c - ci)
        win = np.where(np.abs(d / 8) <= 1, i0(5 * np.sqrt(np.clip(1 - (d / 8) ** 2, 0, None))) / i0(5), 0)
        wgt = np.sinc(d) * win
        wgt /= wgt.sum()
        direct = np.dot(buf[ci - 7: ci + 9].astype(np.float64), wgt)
        assert out[q] == pytest.approx(direct, rel=2e-7)


@pytest.mark.parametrize("coord_f32,smoothing", [(True, 0.0), (False, 0.0), (True, 0.5)])
def test_c_resampler_equals_numpy_oracle(coord_f32, smoothing):
    """The oracle's two resampler implementations (numpy, and the compiled C
    loop the CPU baseline runs) agree bit for bit on an NTSC field with an
    outlier line, in both coordinate modes and with level smoothing."""
    from oracle import ctbc
    ctbc.build()
    rng = np.random.default_rng(9)
    buf = (8.1e6 + rng.normal(0, 3e5, 800_000)).astype(np.float32)
    ll = 5000 + np.arange(273) * 2542.3 + rng.uniform(-2, 2, 273)
    ll[40] += 30
    locs, wow = T.computewow_scaled(ll, 2542.2, 910, 263, 0)
    lut = T.build_kaiser_lut()
    a = T.scale_field(buf, locs, wow, lut, 0, 910, 263 * 910, smoothing=smoothing, coord_f32=coord_f32)
    b = ctbc.scale_field(buf, locs, wow, lut, 0, 910, 263 * 910, smoothing=smoothing, coord_f32=coord_f32)
    np.testing.assert_array_equal(a, b)


def test_v4300d_notch_window():
    """core.py:785-788 at 40 MSPS, N = 32768: int(6897.664) and int(1 + 7045.12)."""
    assert O.v4300d_notch_window(32768, 40) == (6897, 7046)
    assert O.v4300d_notch_window(65536, 40) == (13795, 14091)


def test_v4300d_notch_closed_forms():
    """core.py:774-798 on constructed spectra: a window of equal magnitudes
    (std 0) zeroes nothing; one 100x spur among 148 unit bins (mean 1.66,
    std 8.06: threshold 25.9) zeroes exactly that bin, both neighbours and
    their three mirror bins; a spur on the window's first bin zeroes the bin
    just below the window too."""
    N = 32768
    lo, hi = O.v4300d_notch_window(N, 40)
    rng = np.random.default_rng(3)
    X = np.exp(2j * np.pi * rng.random(N))
    np.testing.assert_array_equal(O.v4300d_notch(X, N, 40), X)
    for spur in (lo + 70, lo):
        Y = X.copy()
        Y[spur] *= 100
        Z = O.v4300d_notch(Y, N, 40)
        zero = {spur - 1, spur, spur + 1} | {N - k for k in (spur - 1, spur, spur + 1)}
        assert set(np.flatnonzero(Z == 0)) == zero
        keep = np.ones(N, bool)
        keep[list(zero)] = False
        np.testing.assert_array_equal(Z[keep], Y[keep])


def test_ntsc_color_notch_filter():
    """NTSC_ColorNotchFilter (core.py:625-631): a 3rd-order Butterworth band
    stop from the video LPF corner to 5 MHz multiplies Fvideo_lpf: unity at
    DC, a null inside the band, and every post-demod filter changes by the
    same factor."""
    a = O.OracleRF("NTSC", decode_analog_audio=False)
    b = O.OracleRF("NTSC", decode_analog_audio=False, ntsc_color_notch=True)
    N = a.blocklen
    r = b.Filters["Fvideo_lpf"][: N // 2] / a.Filters["Fvideo_lpf"][: N // 2]
    f = np.arange(N // 2) * 40e6 / N
    lo = a.DecoderParams["video_lpf_freq"]
    assert abs(r[0] - 1) < 1e-12
    assert np.abs(r[(f > lo) & (f < 5e6)]).min() < 1e-3
    assert np.abs(np.abs(r[f < 1e6]) - 1).max() < 1e-2
    for k in ("FVideo05", "FVideoBurst"):
        q = b.Filters[k][: N // 2] / a.Filters[k][: N // 2]
        ok = np.abs(a.Filters[k][: N // 2]) > 1e-9
        np.testing.assert_allclose(q[ok], r[ok], rtol=1e-9)


def test_hz_to_output_closed_form():
    sp = sys_params("NTSC")
    osc = out_scale("NTSC", -40)
    ire = np.array([-40.0, 0.0, 7.5, 100.0])
    hz = sp["ire0"] + sp["hz_ire"] * ire
    codes = T.hz_to_output_array(hz, sp["ire0"], sp["hz_ire"], 1024, -40, osc)
    expect = np.floor(1024 + (ire + 40) * osc + 0.5).astype(np.uint16)
    np.testing.assert_array_equal(codes, expect)
    assert codes[0] == 1024 and codes[-1] == 0xC800
    assert T.hz_to_output_array(np.array([-1e9, 1e12]), sp["ire0"], sp["hz_ire"], 1024, -40, osc).tolist() == [0, 65535]


def test_compute_deriv_error():
    ll = np.arange(10) * 100.0
    ll[5] += 10
    e = T.compute_deriv_error(ll, np.zeros(10, bool))
    assert e.tolist() == [False, False, False, False, True, True, True, True, False, False]
