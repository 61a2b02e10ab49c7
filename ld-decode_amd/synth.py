"""Seeded synthetic LaserDisc RF captures (NTSC / PAL), FP64.

The signal model follows what the reference itself uses to exercise its
decoder: an IRE raster with equalising / vertical / horizontal sync pulses
(``tests/test_ac3rf.py:123-151``), IRE -> Hz with the system's ``ire0`` /
``hz_ire`` (``core.py:739-741``), pre-emphasis by ``Fvideo_lpf * Femp``
(``core.py:1007-1016``, ``scripts/bench_rf.py:58-61``) and phase-accumulated FM
(``genwave``, ``utils.py:1009-1023``).  On top of that: 75% colour bars with a
colour burst, the two analogue-audio FM carriers, per-line timing jitter,
AWGN, and 10-bit quantisation stored as ``(code - 512) << 6`` int16
(``utils.py:371-390``).

The generator records the true (jittered) start of every line's sync pulse,
in samples, as ground truth for line-location and TBC checks.
"""

import numpy as np
import scipy.fft as sfft
import scipy.signal as sps

from .filters import emphasis
from .params import system_tables

FS = 40e6

# 75% colour bars: (luma IRE above black, chroma peak IRE, chroma phase deg)
_BARS = [(100, 0, 0), (89, 31, 167), (72, 44, 283), (62, 41, 241),
         (48, 41, 61), (38, 44, 103), (21, 31, 347), (0, 0, 0)]


class Capture:
    """A synthetic capture: ``data`` int16 RF, ``line_starts`` float64 sample
    positions of every sync-pulse leading edge (full lines only), and the
    system parameters used."""

    def __init__(self, data, line_starts, halfline_starts, system, sysp, fs, snr_db):
        self.data = data
        self.line_starts = line_starts
        self.halfline_starts = halfline_starts
        self.system = system
        self.sysp = sysp
        self.fs = fs
        self.snr_db = snr_db


def _field_layout(system):
    """(halflines per field, eq pulses, vsync pulses, first video halfline of
    an even-numbered field; odd-numbered fields start their video one
    half-line later)."""
    if system == "NTSC":
        return 525, 6, 6, 18
    return 625, 5, 5, 15


def field_line0(system, field):
    """Global half-line index at which line 0 of field `field` begins: the
    last line start before the field's first equalising pulse (lines run on
    whole-line steps from there).  Even-numbered fields are the reference's
    first fields (their line 0 sits 2 half-lines before the pulses for NTSC,
    1 for PAL; core.py:2039-2054 with firstFieldH), odd-numbered fields the
    second ones."""
    hpf, _, _, first = _field_layout(system)
    first_q = first + (field % 2)
    return field * hpf - (2 if first_q % 2 == 0 else 1)


def ire_raster(system, sysp, nsamples, fs, jitter_samples, rng, start_halfline=0, pilot_ire=0.0):
    """IRE value of every sample plus the line-start table.

    Half-line h (global index) starts at h * H/2 * fs plus the jitter of its
    full line.  Within a field (hpf half-lines), half-lines [0, neq) carry
    equalising pulses, [neq, neq+nvs) vertical sync pulses, [neq+nvs, 3*neq)
    equalising pulses, then every other half-line from the field's first
    video half-line (first, or first + 1 in odd-numbered fields, so that the
    line grid runs on across the interlaced field boundary) starts a video
    line with an hsync pulse, burst and colour bars.  pilot_ire > 0 adds the PAL disc
    pilot (pilot_mhz, continuous phase) of that peak amplitude on every
    hsync pulse."""
    hpf, neq, nvs, first = _field_layout(system)
    H = sysp["line_period"] * 1e-6 * fs          # samples per line
    hh = H / 2
    nhalf = int(np.ceil(nsamples / hh)) + 4
    hidx = np.arange(nhalf) + start_halfline
    jit = rng.uniform(-jitter_samples, jitter_samples, size=nhalf // 2 + 4)
    hstart = (hidx - start_halfline) * hh + jit[(hidx - start_halfline) // 2]
    n = np.arange(nsamples, dtype=np.float64)
    h_of_n = np.searchsorted(hstart, n, side="right") - 1
    h_of_n = np.clip(h_of_n, 0, nhalf - 1)
    t_us = (n - hstart[h_of_n]) / fs * 1e6      # time since the half-line start
    hg = h_of_n + start_halfline
    hf = hg % hpf                                 # half-line within field
    first_n = first + (hg // hpf) % 2             # first video half-line
    sync = sysp["vsync_ire"]
    setup = 7.5 if system == "NTSC" else 0.0
    ire = np.zeros(nsamples)
    eq = (hf < neq) | ((hf >= neq + nvs) & (hf < 3 * neq))
    vs = (hf >= neq) & (hf < neq + nvs)
    ire[eq & (t_us < sysp["eqPulseUS"])] = sync
    ire[vs & (t_us < sysp["vsyncPulseUS"])] = sync
    video = hf >= first_n
    second_half = ((hf - first_n) % 2) == 1
    # time since the line's hsync for video half-lines
    tl = t_us + np.where(second_half, sysp["line_period"] / 2, 0.0)
    vline = video & ~second_half
    ire[vline & (t_us < sysp["hsyncPulseUS"])] = sync
    fsc = sysp["fsc_mhz"]
    t_abs_us = n / fs * 1e6
    if pilot_ire > 0:
        tip = vline & (t_us < sysp["hsyncPulseUS"])
        ire[tip] += pilot_ire * np.sin(2 * np.pi * sysp["pilot_mhz"] * t_abs_us[tip])
    b0, b1 = sysp["colorBurstUS"]
    burst = video & (tl >= b0) & (tl < b1)
    burst_phase = np.pi if system == "NTSC" else 3 * np.pi / 4
    ire[burst] = 20 * np.sin(2 * np.pi * fsc * t_abs_us[burst] + burst_phase)
    a0, a1 = sysp["activeVideoUS"]
    act = video & (tl >= a0) & (tl < a1)
    bar = np.clip(((tl[act] - a0) / (a1 - a0) * 8).astype(int), 0, 7)
    lum = np.array([b[0] for b in _BARS], float) * 0.75
    cam = np.array([b[1] for b in _BARS], float)
    cph = np.deg2rad(np.array([b[2] for b in _BARS], float))
    ire[act] = (setup + lum[bar] * (1 - setup / 100)
                + cam[bar] * np.sin(2 * np.pi * fsc * t_abs_us[act] + cph[bar]))
    # full-line starts: video half-lines at even offset from the field's first
    hf_all = hidx % hpf
    first_all = first + (hidx // hpf) % 2
    is_line = (hf_all >= first_all) & (((hf_all - first_all) % 2) == 0)
    is_line |= hf_all < 3 * neq  # vblank half-lines also begin with a pulse
    line_starts = hstart[is_line & (hstart < nsamples)]
    return ire, line_starts, hstart[hstart < nsamples]


def make_capture(system="NTSC", nsamples=700_000, seed=1, snr_db=40.0, jitter_ns=50.0,
                 audio=True, fs=FS, start_halfline=0, pilot_ire=None, dropouts=()):
    """Generate one capture.  Returns :class:`Capture`.  pilot_ire: peak
    amplitude of the PAL sync-tip pilot (default 10 IRE for PAL, none for
    NTSC).  dropouts: (start, length) sample spans where the RF carrier is
    lost (the signal drops to the noise floor), as on a damaged disc."""
    rng = np.random.default_rng(seed)
    sysp, decp = system_tables(system)
    jitter = jitter_ns * 1e-9 * fs
    if pilot_ire is None:
        pilot_ire = 10.0 if system == "PAL" else 0.0
    ire, line_starts, hstarts = ire_raster(system, sysp, nsamples, fs, jitter, rng, start_halfline, pilot_ire)
    hz = sysp["ire0"] + sysp["hz_ire"] * ire
    # pre-emphasis: Fvideo_lpf * Femp applied in the frequency domain over the
    # whole capture (the reference applies the same product per block)
    lpf_b, lpf_a = sps.butter(decp["video_lpf_order"], decp["video_lpf_freq"] / (fs / 2), "low")
    t1, t2 = decp["video_deemp"]
    emp_b, emp_a = emphasis(t2, t1, fs)
    nf = sfft.next_fast_len(nsamples, real=True)
    w = np.linspace(0, np.pi, nf // 2 + 1)
    H = sps.freqz(lpf_b, lpf_a, worN=w)[1] * sps.freqz(emp_b, emp_a, worN=w)[1]
    mean = hz.mean()
    spec = sfft.rfft(hz - mean, nf) * H
    hz = sfft.irfft(spec, nf)[:nsamples] + mean * H[0].real
    # phase-accumulated FM: phase_i = 2*pi*sum_{m<i} f_m / fs (genwave model)
    phase = np.empty(nsamples)
    phase[0] = 0.0
    np.cumsum(hz[:-1] * (2 * np.pi / fs), out=phase[1:])
    sig = np.sin(phase)
    del phase
    if audio:
        t = np.arange(nsamples) / fs
        if system == "NTSC":
            carriers = [(sysp["audio_lfreq"], 1000.0, 100e3), (sysp["audio_rfreq"], 3000.0, 100e3)]
        else:
            carriers = [(sysp["audio_lfreq"], 1000.0, 50e3), (sysp["audio_rfreq"], 3000.0, 50e3)]
        for fc, fa, dev in carriers:
            sig += 0.1 * np.sin(2 * np.pi * fc * t - (dev / fa) * np.cos(2 * np.pi * fa * t))
        del t
    for a0, n in dropouts:
        sig[a0:a0 + n] = 0.0
    if snr_db is not None:
        sigma = np.sqrt(0.5 / 10 ** (snr_db / 10))
        sig += rng.normal(0.0, sigma, nsamples)
    code = np.clip(np.round(512 + 511 * sig / np.max(np.abs(sig))), 0, 1023).astype(np.int32)
    data = ((code - 512) << 6).astype(np.int16)
    return Capture(data, line_starts, hstarts, system, sysp, fs, snr_db)


def pack_10bit(samples):
    """An int16 capture ((code - 512) << 6) in the packed 10-bit format of
    the Domesday Duplicator: four 10-bit codes in five bytes, most
    significant bits first (the layout utils.py:356-390 unpacks); a partial
    last group is padded with code 0."""
    c = (np.asarray(samples, dtype=np.int32) >> 6) + 512
    c = np.concatenate([c, np.zeros((-len(c)) % 4, np.int32)]).astype(np.uint32).reshape(-1, 4)
    out = np.empty((len(c), 5), dtype=np.uint8)
    out[:, 0] = c[:, 0] >> 2
    out[:, 1] = ((c[:, 0] & 0x03) << 6) | (c[:, 1] >> 4)
    out[:, 2] = ((c[:, 1] & 0x0F) << 4) | (c[:, 2] >> 6)
    out[:, 3] = ((c[:, 2] & 0x3F) << 2) | (c[:, 3] >> 8)
    out[:, 4] = c[:, 3] & 0xFF
    return out.reshape(-1)


# ---------------------------------------------------------------- batches
# Fields per template: whole colour sequences spanning an integer number of
# samples and subcarrier (and PAL pilot) cycles at 40 MSPS, so that templates
# tile without a phase step: 12 NTSC fields = 6300 half-lines of 1271.1 samples
# = 8,008,000 samples (716,625 fsc cycles); 8 PAL fields = 6,400,000 samples.
TEMPLATE_FIELDS = {"NTSC": 12, "PAL": 8}


def template_samples(system, fs=FS):
    sysp, _ = system_tables(system)
    hpf, _, _, _ = _field_layout(system)
    n = TEMPLATE_FIELDS[system] * hpf / 2 * sysp["line_period"] * 1e-6 * fs
    if abs(n - round(n)) > 1e-6:
        raise ValueError(f"{system} template is not a whole number of samples at {fs} Hz")
    return int(round(n))


def batch_plan(seed=4, ncaptures=64, min_s=1.0, max_s=30.0, pal_fraction=0.5, fs=FS):
    """The batched workload: capture lengths log-uniform in [min_s, max_s]
    seconds and systems NTSC / PAL with probability 1 - pal_fraction /
    pal_fraction, from `seed`.  Returns a list of (system, samples)."""
    rng = np.random.default_rng(seed)
    secs = np.exp(rng.uniform(np.log(min_s), np.log(max_s), ncaptures))
    pal = rng.random(ncaptures) < pal_fraction
    return [("PAL" if p else "NTSC", int(round(s * fs))) for s, p in zip(secs, pal)]


def field_template(system, seed, jitter_ns=50.0, fs=FS):
    """One tileable template: the pre-emphasised instantaneous frequency
    (float64 Hz) of TEMPLATE_FIELDS[system] fields starting at the first
    equalising pulse of an even-numbered field, with per-line jitter drawn
    from `seed`.  The pre-emphasis is applied circularly over the template,
    which is what it is when the template repeats."""
    rng = np.random.default_rng(seed)
    sysp, decp = system_tables(system)
    n = template_samples(system, fs)
    pilot = 10.0 if system == "PAL" else 0.0
    ire, _, _ = ire_raster(system, sysp, n, fs, jitter_ns * 1e-9 * fs, rng, 0, pilot)
    hz = sysp["ire0"] + sysp["hz_ire"] * ire
    del ire
    lpf_b, lpf_a = sps.butter(decp["video_lpf_order"], decp["video_lpf_freq"] / (fs / 2), "low")
    t1, t2 = decp["video_deemp"]
    emp_b, emp_a = emphasis(t2, t1, fs)
    w = np.arange(n // 2 + 1) * (2 * np.pi / n)
    H = sps.freqz(lpf_b, lpf_a, worN=w)[1] * sps.freqz(emp_b, emp_a, worN=w)[1]
    mean = hz.mean()
    out = sfft.irfft(sfft.rfft(hz - mean) * H, n) + mean * H[0].real
    return out


class Batch:
    """Captures of a batch back to back in one int16 tensor: capture k is
    data[offsets[k] : offsets[k] + lengths[k]], of system systems[k]."""

    def __init__(self, data, systems, offsets, lengths, fs, snr_db):
        self.data = data
        self.systems = list(systems)
        self.offsets = [int(o) for o in offsets]
        self.lengths = [int(n) for n in lengths]
        self.fs = fs
        self.snr_db = snr_db

    def capture(self, k):
        return self.data[self.offsets[k]: self.offsets[k] + self.lengths[k]]


def make_batch(plan, device, seed=4, snr_db=40.0, ntemplates=2, jitter_ns=50.0, fs=FS, chunk=1 << 25,
               templates=None):
    """Generate every capture of `plan` (batch_plan) on `device` with torch:
    each capture tiles its system's templates (field_template; a seeded
    random template per repeat, from a seeded random phase of the first),
    then the FM carrier is phase-accumulated in float64 over the whole
    capture, the two analogue-audio FM carriers (as make_capture) and AWGN
    are added and the result is quantised to 10-bit codes stored as
    (code - 512) << 6.  Work proceeds in chunks of `chunk` samples."""
    import torch
    rng = np.random.default_rng(seed)
    if templates is None:
        templates = {}
        for s in sorted({p[0] for p in plan}):
            templates[s] = [field_template(s, seed * 1000 + 17 * t + (1 if s == "PAL" else 0), jitter_ns, fs)
                            for t in range(ntemplates)]
    tdev = {s: [torch.from_numpy(t).to(device) for t in ts] for s, ts in templates.items()}
    total = sum(n for _, n in plan)
    data = torch.empty(total, dtype=torch.int16, device=device)
    offsets, o = [], 0
    for _, n in plan:
        offsets.append(o)
        o += n
    sigma = None if snr_db is None else float(np.sqrt(0.5 / 10 ** (snr_db / 10)))
    peak = 1.2 + (6 * sigma if sigma else 0.0)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    two_pi_fs = 2 * np.pi / fs
    for k, (system, n) in enumerate(plan):
        ts = tdev[system]
        T = ts[0].numel()
        reps = (n + T - 1) // T + 2
        order = rng.integers(0, len(ts), size=reps)
        phase0 = int(rng.integers(0, T))          # where in the first template the capture begins
        sysp, _ = system_tables(system)
        dev_hz = 100e3 if system == "NTSC" else 50e3
        carriers = [(sysp["audio_lfreq"], 1000.0, dev_hz), (sysp["audio_rfreq"], 3000.0, dev_hz)]
        carry = 0.0
        for c0 in range(0, n, chunk):
            m = min(chunk, n - c0)
            hz = torch.empty(m, dtype=torch.float64, device=device)
            p, filled = phase0 + c0, 0
            while filled < m:
                q, r = divmod(p, T)
                take = min(T - r, m - filled)
                hz[filled:filled + take] = ts[order[q]][r:r + take]
                filled += take
                p += take
            inc = hz.mul_(two_pi_fs)
            ph = torch.cumsum(inc, 0)
            ph.sub_(inc).add_(carry)                     # exclusive sum: phase[0] = carry
            carry = float(ph[-1] + inc[-1]) % (2 * np.pi)
            del inc, hz
            sig = torch.sin(ph)
            t = torch.arange(c0, c0 + m, dtype=torch.float64, device=device).div_(fs)
            for fc, fa, dv in carriers:
                sig.add_(torch.sin(t * (2 * np.pi * fc) - (dv / fa) * torch.cos(t * (2 * np.pi * fa))), alpha=0.1)
            del t, ph
            if sigma:
                sig.add_(torch.randn(m, dtype=torch.float64, device=device, generator=gen), alpha=sigma)
            code = torch.round(sig.mul_(511.0 / peak).add_(512.0)).clamp_(0, 1023).to(torch.int32)
            data[offsets[k] + c0: offsets[k] + c0 + m] = ((code - 512) << 6).to(torch.int16)
            del sig, code
    return Batch(data, [p[0] for p in plan], offsets, [p[1] for p in plan], fs, snr_db)
