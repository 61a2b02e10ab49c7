"""NTSC/PAL system constants and decoder filter parameters.

Mirrors the constant tables of ``lddecode/core.py:62-251`` (SysParams_NTSC,
SysParams_PAL, FilterParams_NTSC/PAL and the ``*_lowband`` variants) so the
host API accepts the same ``decoder_params_override`` keys.
"""

import copy

import numpy as np

BLOCKSIZE = 32 * 1024  # core.py:66


def _lines_to_samples(sp, mult, mhz_key):
    """core.py:111-115 (calclinelen): samples per line at mult * sp[mhz_key] MHz."""
    return int(np.round(sp["line_period"] * sp[mhz_key] * mult))


_NTSC_FSC = 315.0 / 88.0

SYS_NTSC = {
    "fsc_mhz": _NTSC_FSC,
    "pilot_mhz": _NTSC_FSC,
    "frame_lines": 525,
    "field_lines": (263, 262),
    "ire0": 8100000,
    "hz_ire": 1700000 / 140.0,
    "vsync_ire": -40,
    "analog_audio": True,
    "audio_lfreq": (1000000 * 315 / 88 / 227.5) * 146.25,
    "audio_rfreq": (1000000 * 315 / 88 / 227.5) * 178.75,
    "audio_rfreq_AC3": 2880000,
    "colorBurstUS": (5.3, 7.8),
    "firstFieldH": (0.5, 1),
    "numPulses": 6,
    "hsyncPulseUS": 4.7,
    "eqPulseUS": 2.3,
    "vsyncPulseUS": 27.1,
    "outputZero": 1024,
    "fieldPhases": 4,
    # 227.5 colour cycles per line, formed as the reference forms it
    # (core.py:119): 1 / (fsc / 227.5) is one ulp below 227.5 / fsc
    "line_period": 1 / (_NTSC_FSC / 227.5),
}
SYS_NTSC["activeVideoUS"] = (9.45, SYS_NTSC["line_period"] - 1.0)
SYS_NTSC["FPS"] = 1000000 / (525 * SYS_NTSC["line_period"])
SYS_NTSC["outlinelen"] = _lines_to_samples(SYS_NTSC, 4, "fsc_mhz")
SYS_NTSC["outfreq"] = 4 * SYS_NTSC["fsc_mhz"]

SYS_PAL = {
    "FPS": 25,
    "fsc_mhz": (283.75 / 64) + (25 / 1000000),
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
    "vsync_ire": -0.3 * (100 / 0.7),
}
SYS_PAL["outlinelen"] = _lines_to_samples(SYS_PAL, 4, "fsc_mhz")
SYS_PAL["outlinelen_pilot"] = _lines_to_samples(SYS_PAL, 4, "pilot_mhz")
SYS_PAL["outfreq"] = 4 * SYS_PAL["fsc_mhz"]

FILT_NTSC = {
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
FILT_NTSC_LOWBAND = dict(FILT_NTSC, video_bpf_low=3800000, video_bpf_high=12500000,
                         video_lpf_freq=4200000)

FILT_PAL = {
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
FILT_PAL_LOWBAND = dict(FILT_PAL, video_bpf_low=3200000, video_bpf_high=13000000,
                        video_bpf_order=2, video_lpf_freq=4800000)


def system_tables(system, lowband=False):
    """Deep copies of (SysParams, DecoderParams) for a system, as core.py:343-357."""
    if system == "NTSC":
        return copy.deepcopy(SYS_NTSC), copy.deepcopy(FILT_NTSC_LOWBAND if lowband else FILT_NTSC)
    if system == "PAL":
        return copy.deepcopy(SYS_PAL), copy.deepcopy(FILT_PAL_LOWBAND if lowband else FILT_PAL)
    raise ValueError(f"system must be 'NTSC' or 'PAL', not {system!r}")


def out_scale(system, vsync_ire):
    """Output code units per IRE: core.py:3475-3477 (NTSC), :3215-3217 (PAL)."""
    top = (0xC800 - 0x0400) if system == "NTSC" else (0xD300 - 0x0100)
    return np.double(top) / (100 - vsync_ire)
