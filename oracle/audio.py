"""FP64 restatement of ld-decode's analogue audio stage 2 and its 44.1 kHz
output: the per-field-read FFT filter with clip removal
(RFDecode.runfilter_audio_phase2 / audio_phase2, core.py:884-950, called
from DemodCache.read core.py:1405-1407), the stage-2 filter tables
(core.py:727-737), and downscale_audio (core.py:1423-1580) with the
output-rate bookkeeping of Field.downscale (core.py:2644-2670).  Test
infrastructure only.

Conventions kept from the reference:
  * Stage-1 audio is float32 (the demodulator's recarray); `raw -= center`
    stays float32, the zero-padded FFT input is complex128, the filter
    product and inverse are float64, and the single-block result is
    returned as float64 (the multi-block result is written into a float32
    array of the input's dtype).
  * findpeaks compares each value with the LAST value of the thresholded
    copy (utils.py:1207-1217) and reports loc - 1.
  * _downscale_audio_compute_locs_and_swow and _downscale_audio_to_output
    run under numba: np.arange there is start + i * step, np.mean of float64
    is an in-order sum divided by the count (parity of these two details is
    unpinned: numba is not installed; the reference's tests/test_audio_rate.py
    pins the sample counts and unit wow).
"""

import numpy as np
import scipy.signal as sps

from . import dsp as O


def audio2_filter(orf, channel):
    """Stage-2 filter of a channel (core.py:727-734): 20 kHz Butterworth
    low-pass times the 75 us / 5.3 us de-emphasis, whole circle, blocklen
    bins, at the stage-1 output rate a1_freq."""
    a1 = orf.audio[channel]["a1_freq"]
    N, Wn = sps.buttord(20000 / (a1 / 2), 24000 / (a1 / 2), 1, 9)
    lpf = O.filtfft(sps.butter(N, Wn), orf.blocklen)
    deemp = O.filtfft(O.emphasis_iir(5.3e-6, 75e-6, a1), orf.blocklen)
    return lpf * deemp


def findpeaks(array, low=0):
    """utils.py:1207-1217."""
    a2 = np.array(array, copy=True)
    a2[np.where(a2 < low)] = 0
    return [int(loc) - 1 for loc in np.where(np.logical_and(a2[:-1] > a2[-1], a2[1:] > a2[:-1]))[0]]


def runfilter_audio_phase2(orf, frame_audio, start, filters=None):
    """core.py:884-915: {'audio_left', 'audio_right'} float64 arrays."""
    sp = orf.SysParams
    filters = filters or {ch: audio2_filter(orf, ch) for ch in ("left", "right")}
    out = {}
    clips = None
    for name, center, ch in (("audio_left", sp["audio_lfreq"], "left"), ("audio_right", sp["audio_rfreq"], "right")):
        raw = np.array(frame_audio[name][start:start + orf.blocklen], dtype=np.float32, copy=True)
        raw -= center
        if name == "audio_left":
            clips = findpeaks(raw, 500000)
        for l in clips:
            raw[max(0, l - 8):min(l + 8, len(raw))] = 0
        filt = filters[ch]
        if len(raw) < len(filt):
            a2_in = np.zeros_like(filt)
            a2_in[:len(raw)] = raw
        else:
            a2_in = raw
        out[name] = np.fft.ifft(np.fft.fft(a2_in) * filt).real[:len(raw)] + center
    return out


def audio_phase2(orf, field_audio, filters=None):
    """core.py:917-950 over one field read's stage-1 audio (dict of float32
    arrays).  Returns a dict of arrays (float64 in the single-block case)."""
    filters = filters or {ch: audio2_filter(orf, ch) for ch in ("left", "right")}
    n = len(field_audio["audio_left"])
    tmp = runfilter_audio_phase2(orf, field_audio, 0, filters)
    if len(tmp["audio_left"]) >= n:
        return {k: v[:n] for k, v in tmp.items()}
    out = {k: np.zeros(n, dtype=np.float32) for k in tmp}
    for k in out:
        out[k][:len(tmp[k])] = tmp[k]
    askip = 512
    sjump = orf.blocklen - askip
    ostart = len(tmp["audio_left"])
    for sample in range(sjump, n - sjump, sjump):
        tmp = runfilter_audio_phase2(orf, field_audio, sample, filters)
        m = len(tmp["audio_left"])
        for k in out:
            out[k][ostart:ostart + m - askip] = tmp[k][askip:]
        ostart += m - askip
    tmp = runfilter_audio_phase2(orf, field_audio, n - orf.blocklen - 1, filters)
    m = len(tmp["audio_left"])
    for k in out:
        out[k][n - (m - askip):] = tmp[k][askip:]
    return out


def nb_arange(start, stop, step):
    """numba's np.arange(start, stop, step): ceil((stop - start) / step)
    items, start + i * step."""
    n = int(max(np.ceil((stop - start) / step), 0))
    return np.array([start + i * step for i in range(n)], dtype=np.float64)


def compute_locs_and_swow(lineinfo, line_period, linelen, linecount, timeoffset, freq, scale):
    """_downscale_audio_compute_locs_and_swow (core.py:1423-1486)."""
    if freq < 0:
        timeoffset = 0
        freq = (1000000 / line_period) * -freq
    frametime = linecount / (1000000 / line_period)
    soundgap = 1 / freq
    arange = nb_arange(timeoffset, frametime + (soundgap / 2), soundgap)
    locs = np.zeros(len(arange))
    swow = np.zeros(len(arange))
    for i, t in enumerate(arange):
        linenum = ((t * 1000000) / line_period) + 1
        intlinenum = int(linenum)
        if linenum < 0:
            lineloc_cur = int(lineinfo[0] + (linelen * linenum))
            lineloc_next = lineloc_cur + linelen
        elif len(lineinfo) > linenum + 2:
            lineloc_cur, lineloc_next = lineinfo[intlinenum], lineinfo[intlinenum + 1]
        else:
            lineloc_cur = lineinfo[-2]
            lineloc_next = lineloc_cur + linelen
        sampleloc = lineloc_cur
        sampleloc += (lineloc_next - lineloc_cur) * (linenum - np.floor(linenum))
        swow[i] = (lineloc_next - lineloc_cur) / linelen
        swow[i] = ((swow[i] - 1)) + 1
        if i and np.abs(swow[i] - swow[i - 1]) > 0.015:
            swow[i] = swow[i - 1]
        locs[i] = sampleloc / scale
    return locs, swow, arange, frametime


def nb_mean(x):
    """numba's np.mean of a float64 array: in-order sum / count."""
    s = 0.0
    for v in np.asarray(x, dtype=np.float64):
        s += v
    return s / len(x)


def dsa_rescale_and_clip(x):
    """utils.py:1328-1332."""
    value = int(np.round(x * 32767.0 / 371081.0))
    return min(max(value, -32766), 32766)


def to_output(arange, locs, swow, left, right, lfreq, rfreq):
    """_downscale_audio_to_output (core.py:1489-1535): (int16 interleaved
    output, failed)."""
    out = np.zeros(2 * (len(arange) - 1), dtype=np.int16)
    failed = False
    for i in range(len(arange) - 1):
        start, end = int(locs[i]), int(locs[i + 1])
        if end > start and end < len(left):
            ol = (nb_mean(left[start:end]) * swow[i]) - lfreq
            orr = (nb_mean(right[start:end]) * swow[i]) - rfreq
            out[2 * i] = -dsa_rescale_and_clip(ol)
            out[2 * i + 1] = -dsa_rescale_and_clip(orr)
        else:
            failed = True
    return out, failed


def downscale_audio(audio, lineinfo, sysp, linelen, scale, linecount, timeoffset=0, freq=44100):
    """downscale_audio (core.py:1538-1578): (int16 output, next time
    offset, failed)."""
    locs, swow, arange, frametime = compute_locs_and_swow(lineinfo, sysp["line_period"], linelen, linecount,
                                                          timeoffset, freq, scale)
    out, failed = to_output(arange, locs, swow, audio["audio_left"], audio["audio_right"], sysp["audio_lfreq"],
                            sysp["audio_rfreq"])
    return out, arange[-1] - frametime, failed


def downscale_audio_np(audio, lineinfo, sysp, linelen, scale, linecount, timeoffset=0, freq=44100):
    """downscale_audio with numpy array operations (the means from running
    sums, so in a different order): the CPU baseline's
    timing form of the numba-compiled reference, equal to downscale_audio
    except where a mean lies within rounding of an int16 step."""
    lp = sysp["line_period"]
    if freq < 0:
        timeoffset = 0
        freq = (1000000 / lp) * -freq
    frametime = linecount / (1000000 / lp)
    soundgap = 1 / freq
    arange = nb_arange(timeoffset, frametime + (soundgap / 2), soundgap)
    li = np.asarray(lineinfo, dtype=np.float64)
    linenum = ((arange * 1000000) / lp) + 1
    il = linenum.astype(np.int64)
    inside = (linenum >= 0) & (len(li) > linenum + 2)
    cur = np.where(inside, li[np.clip(il, 0, len(li) - 2)], li[-2])
    nxt = np.where(inside, li[np.clip(il + 1, 0, len(li) - 1)], li[-2] + linelen)
    neg = linenum < 0
    cur = np.where(neg, np.trunc(li[0] + linelen * linenum), cur)
    nxt = np.where(neg, cur + linelen, nxt)
    locs = (cur + (nxt - cur) * (linenum - np.floor(linenum))) / scale
    w = ((nxt - cur) / linelen - 1) + 1
    for i in range(1, len(w)):
        if abs(w[i] - w[i - 1]) > 0.015:
            w[i] = w[i - 1]
    s = locs[:-1].astype(np.int64)
    e = locs[1:].astype(np.int64)
    left, right = np.asarray(audio["audio_left"]), np.asarray(audio["audio_right"])
    ok = (e > s) & (e < len(left)) & (s >= 0)
    out = np.zeros(2 * (len(arange) - 1), dtype=np.int16)
    if ok.any():
        cnt = (e - s)[ok]
        cl = np.concatenate([[0.0], np.cumsum(left, dtype=np.float64)])
        cr = np.concatenate([[0.0], np.cumsum(right, dtype=np.float64)])
        ml = cl[e[ok]] - cl[s[ok]]
        mr = cr[e[ok]] - cr[s[ok]]
        vl = (ml / cnt) * w[:-1][ok] - sysp["audio_lfreq"]
        vr = (mr / cnt) * w[:-1][ok] - sysp["audio_rfreq"]
        sc = lambda v: -np.clip(np.round(v * 32767.0 / 371081.0), -32766, 32766).astype(np.int16)
        idx = np.flatnonzero(ok)
        out[2 * idx], out[2 * idx + 1] = sc(vl), sc(vr)
    return out, arange[-1] - frametime, not ok.all()


def audio_offset(sysp, freq_hz, readloc, is_first, freq, lastfieldwritten):
    """Field.downscale's audio time offset (core.py:2644-2670) for output
    rate `freq` given lastfieldwritten = (field number, readloc)."""
    if not (lastfieldwritten and freq >= 16000):
        return 0
    rf_samples_per_field = freq_hz / sysp["FPS"] / 2
    read_gap = (readloc - lastfieldwritten[1]) / rf_samples_per_field
    field_number = int(np.round(lastfieldwritten[0] + read_gap))
    linecount = sum(sysp["field_lines"]) * (field_number // 2)
    if not is_first:
        linecount += sysp["field_lines"][0]
    samples_per_line = (sysp["line_period"] / 1000000) / (1 / freq)
    count = linecount * samples_per_line
    frac = count - np.floor(count)
    return (1 - frac) * (1 / freq) if frac > .5 else -frac * (1 / freq)
