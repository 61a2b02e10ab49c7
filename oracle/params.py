"""System and filter parameter tables (oracle copy).

Restates the constant tables of ``lddecode/core.py:62-251`` (SysParams_NTSC,
SysParams_PAL, FilterParams_NTSC, FilterParams_PAL and the derived entries
computed at ``core.py:106-170``).  The numbers are the reference's (a table of
constants has to be alike); the code that derives the dependent entries is
written here from the formulas.

Pinned by: closed forms in ``tests/test_oracle_params.py`` (NTSC line period
227.5/fsc = 63.5556 us, outlinelen 910, PAL outlinelen 1135, out_scale values).
"""

import numpy as np


def _round_half_even(x):
    # nb_round (utils.py:1027-1029) is int(np.round(x)): numpy rounds half to even
    return int(np.round(x))


def calclinelen(sysp, mult, mhz):
    """core.py:111-115 -- integer samples per line at mult*mhz."""
    if isinstance(mhz, str):
        mhz = sysp[mhz]
    return _round_half_even(sysp["line_period"] * mhz * mult)


def sys_params(system):
    """core.py:70-103 (NTSC), :121-154 (PAL), derived :106-118, :156-160, :170."""
    if system == "NTSC":
        fsc = 315.0 / np.double(88.0)
        sp = {
            "fsc_mhz": fsc,
            "pilot_mhz": fsc,
            "frame_lines": 525,
            "field_lines": (263, 262),
            "ire0": 8100000,
            "hz_ire": 1700000 / 140.0,
            "vsync_ire": -40,
            "analog_audio": True,
            "audio_lfreq": (1000000 * 315 / 88 / 227.5) * 146.25,
            "audio_rfreq": (1000000 * 315 / 88 / 227.5) * 178.75,
            "colorBurstUS": (5.3, 7.8),
            "firstFieldH": (0.5, 1),
            "numPulses": 6,
            "hsyncPulseUS": 4.7,
            "eqPulseUS": 2.3,
            "vsyncPulseUS": 27.1,
            "outputZero": 1024,
            "fieldPhases": 4,
        }
        sp["line_period"] = 1 / (sp["fsc_mhz"] / np.double(227.5))
        sp["activeVideoUS"] = (9.45, sp["line_period"] - 1.0)
        sp["FPS"] = 1000000 / (525 * sp["line_period"])
        sp["outlinelen"] = calclinelen(sp, 4, "fsc_mhz")
        sp["outfreq"] = 4 * sp["fsc_mhz"]
        return sp
    if system == "PAL":
        sp = {
            "FPS": 25,
            "fsc_mhz": ((1 / 64) * 283.75) + (25 / 1000000),
            "pilot_mhz": 3.75,
            "frame_lines": 625,
            "field_lines": (312, 313),
            "line_period": 64,
            "ire0": 7100000,
            "hz_ire": 800000 / 100.0,
            "analog_audio": True,
            "audio_lfreq": (1000000 / 64) * 43.75,
            "audio_rfreq": (1000000 / 64) * 68.25,
            "colorBurstUS": (5.6, 7.85),
            "activeVideoUS": (10.5, 64 - 1.5),
            "firstFieldH": (1, 0.5),
            "numPulses": 5,
            "hsyncPulseUS": 4.7,
            "eqPulseUS": 2.35,
            "vsyncPulseUS": 27.3,
            "outputZero": 256,
            "fieldPhases": 8,
        }
        sp["outlinelen"] = calclinelen(sp, 4, "fsc_mhz")
        sp["outlinelen_pilot"] = calclinelen(sp, 4, "pilot_mhz")
        sp["outfreq"] = 4 * sp["fsc_mhz"]
        sp["vsync_ire"] = -0.3 * (100 / 0.7)
        return sp
    raise ValueError(f"unknown system {system!r}")


def filter_params(system):
    """core.py:172-204 (NTSC), :211-245 (PAL)."""
    if system == "NTSC":
        return {
            "audio_notchwidth": 200000,
            "audio_notchorder": 2,
            "video_deemp": (120e-9, 320e-9),
            "video_bpf_low": 3700000,
            "video_bpf_low_order": 2,
            "video_bpf_high": 13800000,
            "video_bpf_high_order": 3,
            "video_bpf_order": 4,
            "video_lpf_freq": 4500000,
            "video_lpf_order": 6,
            "MTF_basemult": 0.4,
            "MTF_poledist": 0.9,
            "MTF_freq": 12.2,
            "video_hpf_freq": 10000000,
            "video_hpf_order": 4,
            "audio_filterwidth": 150000,
            "audio_filterorder": 512,
        }
    if system == "PAL":
        return {
            "audio_notchwidth": 200000,
            "audio_notchorder": 2,
            "video_deemp": (100e-9, 400e-9),
            "video_bpf_low": 2300000,
            "video_bpf_low_order": 2,
            "video_bpf_high": 14000000,
            "video_bpf_high_order": 3,
            "video_bpf_order": 2,
            "video_lpf_freq": 5800000,
            "video_lpf_order": 7,
            "MTF_basemult": 1.0,
            "MTF_poledist": 0.70,
            "MTF_freq": 10,
            "video_hpf_freq": 10000000,
            "video_hpf_order": 4,
            "audio_filterwidth": 100000,
            "audio_filterorder": 900,
        }
    raise ValueError(f"unknown system {system!r}")


def out_scale(system, vsync_ire):
    """FieldNTSC.process core.py:3475-3477, FieldPAL.process core.py:3215-3217."""
    if system == "NTSC":
        return np.double(0xC800 - 0x0400) / (100 - vsync_ire)
    return np.double(0xD300 - 0x0100) / (100 - vsync_ire)
