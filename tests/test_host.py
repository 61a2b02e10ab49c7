"""CPU tests of the host side: parameter and filter tables, the synthetic
generator, the LUT builder, and the C ABI library (load, exports, argument
checks that return before any device call)."""

import ctypes
import re

import numpy as np
import pytest

from lddecode_amd import _lib, filters, params, synth
from lddecode_amd import field as fieldmod
from lddecode_amd.hostdsp import calczc, genwave
from oracle import dsp as O
from oracle import tbc as T
from oracle.params import filter_params, sys_params

ROOT = __import__("pathlib").Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_param_tables_match_oracle(system):
    sp, dp = params.system_tables(system)
    osp, odp = sys_params(system), filter_params(system)
    # bit-equal: the tables feed the audio tick grid and the field lengths
    for k, v in osp.items():
        np.testing.assert_array_equal(np.asarray(sp[k], float), np.asarray(v, float), err_msg=k)
    for k, v in odp.items():
        assert dp[k] == v, k


def test_ntsc_line_period_is_the_references_expression():
    # core.py:119 forms 1 / (fsc / 227.5); 227.5 / fsc differs in the last bit
    sp, _ = params.system_tables("NTSC")
    assert sp["line_period"] == 1 / ((315.0 / 88.0) / 227.5) == 63.55555555555555


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_filterbank_matches_oracle(system):
    orf = O.OracleRF(system)
    sp, dp = params.system_tables(system)
    for k in ("ire0", "hz_ire", "vsync_ire"):
        dp[k] = sp[k]
    bank = filters.FilterBank(system, 32768, 40e6, sp, dp)
    for name in ("Frfhpf", "MTF", "RFVideo", "Fvideo_lpf", "Fdeemp", "Femp", "FVideo", "FVideo05", "FVideoBurst"):
        np.testing.assert_allclose(bank.filters[name], orf.Filters[name], rtol=0, atol=1e-12, err_msg=name)
    if system == "PAL":
        np.testing.assert_allclose(bank.filters["FVideoPilot"], orf.Filters["FVideoPilot"], atol=1e-12)
    for ch in ("left", "right"):
        a, b = bank.audio[ch], orf.audio[ch]
        assert (a["lowbin"], a["nbins"], a["a1_freq"], a["low_freq"]) == (b["lowbin"], b["nbins"], b["a1_freq"], b["low_freq"])
        np.testing.assert_allclose(a["filt1"], b["filt1"], atol=1e-12)


def test_filterbank_color_notch_matches_oracle():
    orf = O.OracleRF("NTSC", ntsc_color_notch=True)
    sp, dp = params.system_tables("NTSC")
    for k in ("ire0", "hz_ire", "vsync_ire"):
        dp[k] = sp[k]
    bank = filters.FilterBank("NTSC", 32768, 40e6, sp, dp, color_notch=True)
    for name in ("Fvideo_lpf", "FVideo", "FVideo05", "FVideoBurst"):
        np.testing.assert_allclose(bank.filters[name], orf.Filters[name], rtol=0, atol=1e-12, err_msg=name)


def test_sinc_lut_builder_matches_oracle():
    np.testing.assert_array_equal(fieldmod.build_sinc_lut(), T.build_kaiser_lut())


def test_host_helpers():
    d = np.arange(100, dtype=np.float64) * 2.0
    assert calczc(d, 5, 37.3, count=40) == pytest.approx(O.calczc(d, 5, 37.3, count=40))
    r = np.full(1000, 7.5e6)
    np.testing.assert_allclose(genwave(r, 20e6), O.genwave(r, 20e6), atol=1e-9)


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_synthetic_capture(system):
    cap = synth.make_capture(system, 400_000, seed=7, snr_db=40.0)
    assert cap.data.dtype == np.int16 and len(cap.data) == 400_000
    codes = (cap.data.astype(np.int32) >> 6) + 512
    assert np.all((cap.data.astype(np.int32) & 63) == 0)          # 10-bit codes << 6
    assert codes.min() >= 0 and codes.max() <= 1023
    H = cap.sysp["line_period"] * 40
    dl = np.diff(cap.halfline_starts)
    assert np.all(np.abs(dl - H / 2) <= 4.0 + 1e-9)                 # +-50 ns jitter per line
    again = synth.make_capture(system, 400_000, seed=7, snr_db=40.0)
    np.testing.assert_array_equal(cap.data, again.data)             # seeded


# ------------------------------------------------------------------ C ABI
def _declared_symbols():
    text = (ROOT / "include" / "ldgpu.h").read_text()
    return sorted(set(re.findall(r"\b(ld_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared_symbols()
    assert set(declared) == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.ld_abi_version() == _lib.ABI_VERSION


def test_abi_argument_errors_without_device():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.ld_create(0, 1000, ctypes.byref(h)) == _lib.LD_EINVAL
    assert b"blocklen" in lib.ld_last_error()
    assert lib.ld_create(0, 32768, None) == _lib.LD_EINVAL
    out = _lib.DemodOutputs()
    assert lib.ld_demod_blocks(None, None, 0, 0, 0, 0, 1, 1, ctypes.byref(out), None, 0, None) == _lib.LD_EINVAL
    assert lib.ld_demod_workspace_bytes(None, 4) == 0
    with pytest.raises(_lib.LdGpuError):
        _lib.check(lib.ld_set_filter(None, 0, None, 0, 1.0, 0.0, 0))
    assert lib.ld_set_notch(None, 6897, 7046) == _lib.LD_EINVAL


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_lib.LdGpuUnavailable):
        _lib.load(tmp_path / "nope.so")


def test_batch_plan_is_seeded_log_uniform():
    plan = synth.batch_plan(seed=4)
    assert plan == synth.batch_plan(seed=4) and plan != synth.batch_plan(seed=5)
    secs = np.array([n for _, n in plan]) / 40e6
    assert len(plan) == 64 and secs.min() >= 1.0 and secs.max() <= 30.0
    # log-uniform on [1, 30]: the mean of log(length) is (log 1 + log 30) / 2
    assert abs(np.log(secs).mean() - np.log(30) / 2) < 0.3
    npal = sum(s == "PAL" for s, _ in plan)
    assert 20 <= npal <= 44


def test_templates_are_whole_colour_sequences():
    for system, fsc_key in (("NTSC", "fsc_mhz"), ("PAL", "fsc_mhz")):
        n = synth.template_samples(system)
        sp, _ = params.system_tables(system)
        cycles = n / 40e6 * sp[fsc_key] * 1e6
        assert abs(cycles - round(cycles)) < 1e-6, system
        lines = n / (sp["line_period"] * 40)
        assert abs(lines * 2 - round(lines * 2)) < 1e-9


def test_make_batch_on_cpu_is_deterministic():
    import torch
    plan = [("NTSC", 300_000), ("PAL", 200_000)]
    tm = {"NTSC": [synth.field_template("NTSC", 1)[:100_000]] * 1, "PAL": [synth.field_template("PAL", 2)[:80_000]]}
    a = synth.make_batch(plan, "cpu", seed=3, templates=tm, chunk=65536)
    b = synth.make_batch(plan, "cpu", seed=3, templates=tm, chunk=131072)
    x = a.data.numpy()
    assert a.offsets == [0, 300_000] and a.lengths == [300_000, 200_000]
    assert np.all(x % 64 == 0) and x.min() >= -32768 and x.max() <= 511 * 64
    # chunking does not change the signal (phase carried across chunks); the
    # noise draws follow the chunks, so compare with noise off
    a0 = synth.make_batch(plan, "cpu", seed=3, templates=tm, chunk=65536, snr_db=None).data.numpy()
    b0 = synth.make_batch(plan, "cpu", seed=3, templates=tm, chunk=131072, snr_db=None).data.numpy()
    assert np.mean(a0 == b0) > 0.999
    assert np.std(x) > 1e4 and b.data.shape == a.data.shape


def test_demod_batch_argument_checks_without_device():
    lib = _lib.load()
    rc = lib.ld_demod_batch(None, None, 0, 0, None, 1, None, 1, None, None, 0, None)
    assert rc == _lib.LD_EINVAL
