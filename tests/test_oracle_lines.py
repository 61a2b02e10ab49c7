"""Pins for the oracle's line-location refinements (CPU): synthetic ground
truth (the capture generator's burst and pilot follow absolute time while
every line start carries timing jitter, so a correct alignment puts every
refined line on the unjittered grid up to one common offset), known answers
for the burst zero-crossing phase and closed forms for bad-line repair and
the circular mean."""

import types

import numpy as np
import pytest

from lddecode_amd import capture, synth
from lddecode_amd.synth import field_line0
from oracle import dsp as O
from oracle import lines as L
from oracle import tbc as T


def _params(orf):
    return L.LineRefineParams(types.SimpleNamespace(SysParams=orf.SysParams, DecoderParams=orf.DecoderParams,
                                                    system=orf.system, freq=40.0, linelen=orf.linelen))


def _field_case(system, seed, jitter_ns):
    """One field of a synthetic capture through the oracle: demod, hsync
    refine, and the unjittered start of every line relative to the field's
    buffer."""
    orf = O.OracleRF(system, decode_analog_audio=False)
    rf = types.SimpleNamespace(system=system, blockcut=1024, SysParams=orf.SysParams, block_stride=31712,
                               blocklen=32768)
    fl = 667_334 if system == "NTSC" else 800_000
    cap = synth.make_capture(system, 3 * fl + 100_000, seed=seed, snr_db=40.0, jitter_ns=jitter_ns)
    fields, ll1 = capture.truth_fields(rf, cap, 1, first_field=1)
    org = fields[0]["demod_offset"]
    nb = 29 if system == "NTSC" else 33
    ref, _ = O.demod_span(orf, cap.data.astype(np.float64), org, nb * 31712 - 1)
    video = {k: v.astype(np.float32) for k, v in ref.items()}
    ll2, bad = T.refine_linelocs_hsync(video["demod_05"], ll1, np.zeros(len(ll1), bool), system, 40.0,
                                       orf.iretohz, orf.DecoderParams["vsync_ire"])
    bad = L.compute_deriv_error(ll2, bad)
    idx = field_line0(system, 1) + 2 * np.arange(len(ll1))
    nominal = idx * (orf.SysParams["line_period"] * 40 / 2) - (org + 1024)
    return orf, video, ll2, bad, nominal


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_refinement_recovers_unjittered_grid(system):
    """After burst (NTSC) / pilot (PAL) alignment every good line sits at the
    same offset from its unjittered start: the spread left by the hsync
    refine (uniform +-1.2 sample jitter) shrinks by more than 4x."""
    orf, video, ll2, bad, nominal = _field_case(system, seed=2, jitter_ns=30)
    p = _params(orf)
    if system == "NTSC":
        # field 1 is a second field; its burst search starts from the phase
        # median a previous field leaves behind (compute_burst_offsets uses
        # prevfield.phase_adjust_median), here from a first pass over itself
        _, _, _, med0 = L.ntsc_field_lines(p, video["demod"], video["demod_burst"], ll2, bad, 262, False, 0.0)
        ll, badf, fpid, med = L.ntsc_field_lines(p, video["demod"], video["demod_burst"], ll2, bad, 262, False,
                                                 med0)
        n = L.NTSC_BURST_LINES
        assert fpid in (1, 2, 3, 4) and np.isfinite(med)
    else:
        ll, badf = L.pal_field_lines(p, video["demod_pilot"], ll2, bad)
        n = L.PAL_PILOT_LINES
    good = ~badf[:n]
    good[:12] = False
    good[n - 20:] = False     # the next field's blanking: no burst / pilot there
    before = (ll2 - nominal)[:n][good]
    after = (ll - nominal)[:n][good]
    assert good.sum() > 0.6 * n
    assert after.std() < before.std() / 4, (after.std(), before.std())
    assert np.percentile(np.abs(after - np.median(after)), 95) < 0.5


def test_clb_findbursts_known_phase():
    """A clean burst whose zero crossings sit at phi + k*D (D = half a
    subcarrier period in samples): two passes from phase_adjust 0 settle on
    round(phi/D) - phi/D, and the caller's adjustment (-phase_adjust) is the
    fractional crossing phase phi/D - round(phi/D)."""
    D = 5.587
    for phi in (0.3 * D, 0.8 * D, 1.45 * D):
        n = np.arange(120)
        burst = np.sin(np.pi * (n - phi) / D).astype(np.float32)
        thr = np.float32(np.sqrt(np.mean(np.square(burst - np.mean(burst)))))
        pa = 0.0
        for _ in range(2):
            cnt, pa, rising = L.clb_findbursts(burst, thr, 0, 0.0, D, pa)
        assert cnt == 16
        assert -pa == pytest.approx(phi / D - np.round(phi / D), abs=2e-3)


def test_fix_badlines_interpolates_between_good_neighbours():
    ll = 100.0 + np.arange(40) * 2542.25
    bad = np.zeros(40, bool)
    broken = ll.copy()
    broken[[10, 11, 12, 30]] += [7.0, -3.0, 9.0, 5.0]
    bad[[10, 11, 12, 30]] = True
    fixed, bad2 = L.fix_badlines("NTSC", broken, bad)
    np.testing.assert_allclose(fixed, ll, rtol=0, atol=1e-9)
    # NTSC never repairs from line 0 (firstcheck 1); PAL does
    b0 = ll.copy()
    b0[1] += 8
    bad0 = np.zeros(40, bool)
    bad0[1] = True
    assert L.fix_badlines("NTSC", b0, bad0)[0][1] == b0[1]
    assert L.fix_badlines("PAL", b0, bad0)[0][1] == pytest.approx(ll[1], abs=1e-9)
    # NaN locations fall back to the backup
    bn = ll.copy()
    bn[20] = np.nan
    out, _ = L.fix_badlines("NTSC", bn, np.zeros(40, bool), ll)
    assert out[20] == ll[20]


def test_angular_mean_and_phase_distance():
    assert L.angular_mean([0.95, 1.05]) == pytest.approx(1.0, abs=1e-12) or \
        L.angular_mean([0.95, 1.05]) == pytest.approx(0.0, abs=1e-12)
    assert L.angular_mean([0.2, 0.3, 5.25]) == pytest.approx(0.25, abs=1e-9)
    assert L.angular_mean([0.7, 0.8]) == pytest.approx(0.75, abs=1e-12)
    assert L.phase_distance(3.9, 0.1) == pytest.approx(-0.2)
    assert L.phase_distance(0.05, 0.95) == pytest.approx(0.1)
    assert L.phase_distance(0.5, 0.25) == pytest.approx(0.25)
