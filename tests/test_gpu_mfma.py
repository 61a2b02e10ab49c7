"""The split-BF16 MFMA transform variants (radix-64 stages of the 16384-point
complex transform as bf16 matrix products with FP32 accumulation: hi + lo
terms, LD_FFT_MFMA_BF16X3, or hi + mid + lo, LD_FFT_MFMA_BF16X6) against
numpy's FFT in float64 and, inside the demodulator, against the FP64 oracle.
Bars: the FP32 transform within 1e-6 relative RMS; the two-term MFMA
transform within 3e-5 relative RMS (two stages at about 2^-17 each, FP32
accumulation of 64 terms); the three-term one within 1e-6, as FP32;
demodulated video within 1e-4 relative of the FP64 oracle, and the
three-term variant's demod error within twice the FP32 path's."""

import numpy as np
import pytest

from lddecode_amd import _lib, synth

pytestmark = pytest.mark.gpu

M = 16384


def _fft(rf, torch, x, variant):
    xin = torch.from_numpy(x.astype(np.complex64)).to(rf.device)
    out = torch.empty_like(xin)
    _lib.check(rf.lib.ld_fft_forward(rf.handle, xin.data_ptr(), out.data_ptr(), x.size // M, variant, None))
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.complex128)


def _rel_rms(got, want):
    return float(np.sqrt(np.mean(np.abs(got - want) ** 2) / np.mean(np.abs(want) ** 2)))


@pytest.mark.parametrize("variant,bar", [(_lib.LD_FFT_FP32, 1e-6), (_lib.LD_FFT_MFMA_BF16X3, 3e-5),
                                         (_lib.LD_FFT_MFMA_BF16X6, 1e-6)])
def test_fft_variants_match_numpy(gpu_rf, torch_cuda, variant, bar):
    rf = gpu_rf("NTSC", audio=False)
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(4 * M) + 1j * rng.standard_normal(4 * M))
    x[M:2 * M] *= 1000.0                      # scale invariance
    got = _fft(rf, torch_cuda, x, variant)
    want = np.fft.fft(x.reshape(4, M), axis=1).reshape(-1)
    err = _rel_rms(got, want)
    print(f"variant {variant}: relative RMS error {err:.3e}", flush=True)
    assert err < bar


@pytest.mark.parametrize("variant,tol", [(_lib.LD_FFT_MFMA_BF16X3, 2e-4), (_lib.LD_FFT_MFMA_BF16X6, 2e-6)])
def test_mfma_fft_known_answers(gpu_rf, torch_cuda, variant, tol):
    rf = gpu_rf("NTSC", audio=False)
    n = np.arange(M)
    # a complex tone on bin 1234, a DC block and an impulse at 77
    x = np.concatenate([np.exp(2j * np.pi * 1234 * n / M), np.ones(M), (n == 77).astype(complex),
                        np.exp(-2j * np.pi * 5 * n / M) * 3.0])
    got = _fft(rf, torch_cuda, x, variant).reshape(4, M)
    tone = np.zeros(M, complex)
    tone[1234] = M
    np.testing.assert_allclose(got[0], tone, atol=M * tol)
    dc = np.zeros(M, complex)
    dc[0] = M
    np.testing.assert_allclose(got[1], dc, atol=M * tol)
    np.testing.assert_allclose(got[2], np.exp(-2j * np.pi * 77 * n / M), atol=tol)
    neg = np.zeros(M, complex)
    neg[M - 5] = 3 * M
    np.testing.assert_allclose(got[3], neg, atol=3 * M * tol)


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_mfma_demod_matches_oracle(gpu_rf, torch_cuda, oracle_rf, system):
    """One block demodulated with the MFMA transforms: every video channel
    within 1e-4 relative of the FP64 oracle; the error is also reported
    next to the FP32 path's."""
    rf = gpu_rf(system, audio=True)
    cap = synth.make_capture(system, 32768, seed=1, snr_db=40.0)
    ref = oracle_rf(system, audio=True).demodblock(data=cap.data.astype(np.float64), mtf_level=0, cut=True)
    res = {}
    try:
        for variant in (_lib.LD_FFT_FP32, _lib.LD_FFT_MFMA_BF16X3, _lib.LD_FFT_MFMA_BF16X6):
            rf.set_fft_variant(variant)
            got = rf.demodblock(data=cap.data, mtf_level=0, cut=True)
            res[variant] = {k: got["video"][k].cpu().numpy().astype(np.float64) for k in rf.video_fields}
    finally:
        rf.set_fft_variant(_lib.LD_FFT_FP32)
    for k in rf.video_fields:
        e32 = np.abs(res[0][k] - ref["video"][k])
        e16 = np.abs(res[1][k] - ref["video"][k])
        e6 = np.abs(res[2][k] - ref["video"][k])
        print(f"{system} {k}: FP32 max {e32.max():.2f} Hz rms {np.sqrt(np.mean(e32**2)):.3f}; "
              f"MFMA x3 max {e16.max():.2f} Hz rms {np.sqrt(np.mean(e16**2)):.3f}; "
              f"MFMA x6 max {e6.max():.2f} Hz rms {np.sqrt(np.mean(e6**2)):.3f}", flush=True)
        # relative to the value, or to 1 MHz for the channels that swing
        # around zero (demod_burst), as the FP32 parity tests measure it
        assert (e16 / np.maximum(np.abs(ref["video"][k]), 1e6)).max() <= 1e-4, k
        assert (e6 / np.maximum(np.abs(ref["video"][k]), 1e6)).max() <= 1e-4, k
        assert np.sqrt(np.mean(e6 ** 2)) <= 2.0 * np.sqrt(np.mean(e32 ** 2)) + 0.05, k
