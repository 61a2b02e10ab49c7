"""The analogue audio stage-2 / output-rate oracle (oracle/audio.py) pinned
by the reference's own known answers and closed forms, and the package's
host-side tick arithmetic against it.

Pins:
  * tests/test_audio_rate.py of the reference: a 525-line NTSC frame with
    evenly spaced lines gives exactly 1470 ticks at -2.8 x the line rate,
    1471 at 44.1 kHz, unit wow throughout; -2.8 resolves to 44055.944... Hz.
  * the stage-2 filter: gain 1 at DC (to the precision of its b/a form), conjugate-symmetric (a real
    filter), the package's table equal to the oracle's; a tone on an exact
    FFT bin over a whole block comes out scaled and shifted by H at that bin.
  * findpeaks' comparison with the last value, the +-8 clip mask in both
    channels (identity filter), and the overlap-save branch's assembly,
    whose last segment is one sample late in the reference.
  * dsa_rescale_and_clip, the in-order mean, the first-tick offset.
"""

import numpy as np
import pytest
import scipy.signal as sps

from lddecode_amd import audio as A
from lddecode_amd import filters as F
from oracle import audio as OA
from oracle import dsp as O

def _orf(system="NTSC"):
    return O.OracleRF(system, decode_analog_audio=True)


def test_reference_audio_rate_answers():
    orf = _orf()
    lp = orf.SysParams["line_period"]
    lineinfo = np.arange(0, (525 + 4) * 910, 910, dtype=np.double)
    _, swow, arange, _ = OA.compute_locs_and_swow(lineinfo, lp, 910, 525, 0.0, -2.8, 32)
    assert len(arange) - 1 == 1470
    np.testing.assert_allclose(swow, 1.0, atol=1e-9)
    _, _, arange, _ = OA.compute_locs_and_swow(lineinfo, lp, 910, 525, 0.0, 44100, 32)
    assert len(arange) - 1 == 1471
    assert (1_000_000 / lp) * 2.8 == 44055.944055944055
    # the package's tick arithmetic gives the same counts
    assert A.ticks(lp, 525, 0.0, -2.8)[2] - 1 == 1470
    assert A.ticks(lp, 525, 0.0, 44100)[2] - 1 == 1471
    for t0 in (0.0, -1.1e-5, 1.7e-5):
        for lc in (262, 263, 312, 313):
            _, _, ar, ft = OA.compute_locs_and_swow(lineinfo, lp, 910, lc, t0, 44100, 32)
            s, rate, n, ft2 = A.ticks(lp, lc, t0, 44100)
            assert n == len(ar) and ft2 == ft and s == t0


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_stage2_filter_closed_forms(system):
    orf = _orf(system)
    for ch in ("left", "right"):
        h = OA.audio2_filter(orf, ch)
        n = len(h)
        # DC: 1 for both factors in exact arithmetic; the order-10 low-pass in
        # the reference's b/a form evaluates to 0.99993 there (its
        # denominator's coefficients, up to 177 in size, cancel to 1.5e-10)
        a1 = orf.audio[ch]["a1_freq"]
        order, wn = sps.buttord(20000 / (a1 / 2), 24000 / (a1 / 2), 1, 9)
        assert order == (10 if system == "NTSC" else order)
        e = O.emphasis_iir(5.3e-6, 75e-6, a1)
        assert abs(e[0].sum() / e[1].sum() - 1.0) < 1e-12
        assert abs(h[0] - 1.0) < 1e-3 and abs(h[0].imag) < 1e-12
        # freqz of this ill-conditioned b/a pair is conjugate-symmetric only
        # to ~1e-4; ifft(X H).real applies its Hermitian part, the table the
        # GPU takes: (H[k] + conj(H[N - k])) / 2 over the half spectrum
        asym = np.abs(h[1:n // 2] - np.conj(h[n - 1:n // 2:-1])).max()
        assert asym < 2e-4
        x = np.random.default_rng(1).normal(size=n)
        X = np.fft.fft(x)
        heff = 0.5 * (h[:n // 2 + 1] + np.conj(h[(-np.arange(n // 2 + 1)) % n]))
        np.testing.assert_allclose(np.fft.ifft(X * h).real, np.fft.irfft(X[:n // 2 + 1] * heff, n), atol=1e-12)
        np.testing.assert_allclose(F.audio2_half(h), heff, atol=0)
        pkg = F.audio2_filter(orf.audio[ch]["a1_freq"], n)
        np.testing.assert_allclose(pkg, h, atol=1e-14)
        # |H| falls below the 24 kHz stop band and is ~1 well inside 20 kHz
        k = lambda f: int(round(f / a1 * n))
        assert abs(abs(h[k(1000)]) - 1) < 0.1 and abs(h[k(30000)]) < 0.36


def test_stage2_tone_on_a_bin():
    """A block-long tone on bin k (no zero padding): the output is the tone
    times |H[k]| and shifted by arg H[k], plus the carrier."""
    orf = _orf()
    N = orf.blocklen
    sp = orf.SysParams
    n = np.arange(N)
    k = 300
    amp = 40000.0
    left = (sp["audio_lfreq"] + amp * np.cos(2 * np.pi * k * n / N)).astype(np.float32)
    right = (sp["audio_rfreq"] + amp * np.sin(2 * np.pi * k * n / N)).astype(np.float32)
    out = OA.audio_phase2(orf, {"audio_left": left, "audio_right": right})
    for name, ch, x, c in (("audio_left", "left", left, sp["audio_lfreq"]),
                           ("audio_right", "right", right, sp["audio_rfreq"])):
        h = OA.audio2_filter(orf, ch)
        raw = x.astype(np.float32) - np.float32(c)
        X = np.fft.fft(raw.astype(np.float64))
        want = np.fft.ifft(X * h).real + c
        np.testing.assert_allclose(out[name], want, atol=1e-6)
        # closed form of the same: tone amplitude |H[k]| (raw is the float32 tone)
        ref = c + amp * abs(h[k]) * (np.cos if name == "audio_left" else np.sin)(2 * np.pi * k * n / N + np.angle(h[k]))
        assert np.max(np.abs(out[name] - ref)) < 0.05 * amp * 1e-3 + 0.2
        assert out[name].dtype == np.float64


def test_findpeaks_and_clip_mask():
    a = np.zeros(20, np.float32)
    a[5], a[6], a[7] = 600000, 700000, 650000
    assert OA.findpeaks(a, 500000) == [4]
    b = a.copy()
    b[-1] = 800000            # the comparison is with the last value
    assert OA.findpeaks(b, 500000) == []
    b[-1] = 499999            # below the threshold: zeroed before comparing
    assert OA.findpeaks(b, 500000) == [4]
    # the clip mask: identity filters, a left spike zeroes [l-8, l+8) of both
    orf = _orf()
    sp = orf.SysParams
    N = orf.blocklen
    ident = {"left": np.ones(N, complex), "right": np.ones(N, complex)}
    n = 1000
    left = np.full(n, sp["audio_lfreq"] + 1000, np.float32)
    right = np.full(n, sp["audio_rfreq"] - 2000, np.float32)
    left[300], left[301] = sp["audio_lfreq"] + 600000, sp["audio_lfreq"] + 900000   # peak at i = 300 -> l = 299
    out = OA.runfilter_audio_phase2(orf, {"audio_left": left, "audio_right": right}, 0, ident)
    for name, x, c in (("audio_left", left, sp["audio_lfreq"]), ("audio_right", right, sp["audio_rfreq"])):
        raw = x.astype(np.float32) - np.float32(c)
        z = np.zeros(n, bool)
        z[291:307] = True
        np.testing.assert_allclose(out[name][z], c, atol=1e-6)
        np.testing.assert_allclose(out[name][~z], raw[~z].astype(np.float64) + c, atol=1e-6)


def test_overlap_save_branch_assembly():
    """Reads longer than a block: with identity filters every position is
    the input, except that the last segment is one sample late."""
    orf = _orf()
    sp = orf.SysParams
    N = orf.blocklen
    ident = {"left": np.ones(N, complex), "right": np.ones(N, complex)}
    n = 2 * N + 777
    rng = np.random.default_rng(2)
    left = (sp["audio_lfreq"] + rng.normal(size=n) * 1000).astype(np.float32)
    right = (sp["audio_rfreq"] + rng.normal(size=n) * 1000).astype(np.float32)
    out = OA.audio_phase2(orf, {"audio_left": left, "audio_right": right}, ident)
    tail = N - 512
    np.testing.assert_allclose(out["audio_left"][:n - tail], left[:n - tail], atol=0.07)
    np.testing.assert_allclose(out["audio_left"][n - tail:], left[n - tail - 1:n - 1], atol=0.07)
    assert out["audio_left"].dtype == np.float32


def test_timing_form_of_the_output():
    """downscale_audio_np (the CPU baseline's numpy form) gives the loop
    form's pairs, next offset and failure flag on a wobbly field."""
    orf = _orf()
    sp = orf.SysParams
    rng = np.random.default_rng(0)
    n = 29 * 991
    aud = {"audio_left": sp["audio_lfreq"] + rng.normal(size=n) * 30000,
           "audio_right": sp["audio_rfreq"] + rng.normal(size=n) * 30000}
    ll = 1000 + np.cumsum(rng.normal(2541, 0.5, size=273))
    ll[120] += 80                      # a jump that the 1.5% wow clamp holds
    for t0 in (-1e-5, 0.0, 2e-5):
        a = OA.downscale_audio(aud, ll, sp, orf.linelen, 32, 263, t0, 44100)
        b = OA.downscale_audio_np(aud, ll, sp, orf.linelen, 32, 263, t0, 44100)
        np.testing.assert_array_equal(a[0], b[0])
        assert a[1] == b[1] and a[2] == b[2]


def test_rescale_mean_and_offsets():
    assert OA.dsa_rescale_and_clip(0.0) == 0
    assert OA.dsa_rescale_and_clip(1e9) == 32766 and OA.dsa_rescale_and_clip(-1e9) == -32766
    assert OA.dsa_rescale_and_clip(371081.0 / 32767.0 * 100.0) == 100
    assert OA.nb_mean([1.0, 2.0, 4.0]) == 7.0 / 3.0
    # to_output: two ticks over known means, swow 1
    left = np.array([10.0, 20.0, 30.0, 40.0, 50.0]) + 100.0
    right = np.array([1.0, 1.0, 1.0, 1.0, 1.0]) + 200.0
    out, failed = OA.to_output(np.array([0.0, 1.0, 2.0]), np.array([0.0, 2.0, 4.0]), np.ones(3), left, right,
                               100.0, 200.0)
    assert not failed
    assert list(out) == [-OA.dsa_rescale_and_clip(15.0), -OA.dsa_rescale_and_clip(1.0),
                         -OA.dsa_rescale_and_clip(35.0), -OA.dsa_rescale_and_clip(1.0)]
    _, failed = OA.to_output(np.array([0.0, 1.0]), np.array([0.0, 5.0]), np.ones(2), left, right, 0, 0)
    assert failed              # the tick's end reaches the end of the audio
    sp = _orf().SysParams
    for fn, first in ((0, True), (1, False), (2, True), (17, False), (1000, True)):
        want = OA.audio_offset(sp, 40e6, 5.0, first, 44100, (fn, 5.0))
        assert A.time_offset(sp, fn, first, 44100) == want
    assert A.time_offset(sp, 3, False, 8000) == 0
    # field 1 (second field): 263 lines of 2.8027... ticks, offset -frac / 44100
    cnt = 263 * (sp["line_period"] / 1e6) * 44100
    assert abs(A.time_offset(sp, 1, False, 44100) + (cnt - np.floor(cnt)) / 44100) < 1e-15
