"""FP64 restatement of the line-location refinements that follow the hsync
refine: NTSC colour-burst alignment, PAL pilot alignment, bad-line repair
and the NTSC subcarrier shift (core.py:2566-2589, :2999-3047, :3054-3093,
:3212-3237, :3361-3495; utils.py:1288-1368).  Test infrastructure only.

Conventions kept from the reference, as written there:
  * Field.lineoffset is 0 while these run (FieldPAL sets 2/3 afterwards,
    core.py:3230-3231) and Field.linecount is 263/262 for a first/second
    field (core.py:2340) in both systems; get_linelen switches to a backward
    difference at line >= linecount + lineoffset (core.py:1668-1669).
  * compute_line_bursts reads self.linelocs, which is linelocs2 at that point
    (core.py:1652), for its line frequency.
  * clb_findbursts takes its median and rising count over all 16 slots of
    its zero-crossing arrays, unfilled slots included (zeros).
  * burst areas are float32 (the demod buffers' dtype); nb_mean accumulates
    a float32 sum sequentially and divides in float64 (numba's np.mean), and
    subtracting that Python float from a float32 array rounds it to float32
    first (NumPy 2 scalar promotion).  Zero-crossing positions are stored in
    a float32 array (utils.py:1354); everything else is float64.
"""

import numpy as np

from .dsp import calczc

NTSC_BURST_LINES = 266      # compute_burst_offsets / refine_linelocs_burst line range
PAL_PILOT_LINES = 323       # refine_linelocs_pilot line range
FSC_PHASE_DEG = 117.25      # FieldNTSC.process subcarrier phase offset


def nb_mean_f32(a):
    """nb_mean (utils.py:1242-1244) of a float32 array under numba: a float32
    accumulator over the elements in order, divided by the count in float64."""
    c = np.float32(0.0)
    for v in np.asarray(a, dtype=np.float32):
        c = np.float32(c + v)
    return float(c) / len(a)


def get_linelen(linelocs, line, linecount, lineoffset, linelen):
    """Field.get_linelen (core.py:1656-1679)."""
    ll = linelocs
    if line >= linecount + lineoffset:
        length = (ll[line] - ll[line - 1]) / 1
    elif line > 0:
        length = (ll[line + 1] - ll[line - 1]) / 2
    else:
        length = (ll[line + 1] - ll[line]) / 1
    if length <= 0:
        return linelen
    return length


def clb_findbursts(burstarea, threshold, bstart, s_rem, zcburstdiv, phase_adjust):
    """utils.py:1339-1368 with i = 0 and endburstarea = len - 1.  Returns
    (zc_count, phase_adjust, rising_count)."""
    isrising = np.zeros(16, dtype=bool)
    zcs = np.zeros(16, dtype=np.float32)
    zc_count = 0
    rising_count = 0
    j = 0
    end = len(burstarea) - 1
    while j < end and zc_count < 16:
        if np.abs(burstarea[j]) > threshold:
            zc = calczc(burstarea, j, 0)
            if zc is not None:
                isrising[zc_count] = burstarea[j] < 0
                zcs[zc_count] = zc
                zc_count += 1
                j = int(zc) + 1
            else:
                break
        else:
            j += 1
    if zc_count:
        zc_cycles = ((bstart + zcs.astype(np.float64) - s_rem) / zcburstdiv) + phase_adjust
        zc_rounds = (zc_cycles + 0.5).astype(np.int32)
        phase_adjust += np.median(zc_rounds - zc_cycles)
        rising_count = int(np.sum(np.bitwise_xor(isrising, (zc_rounds % 2) != 0)))
    return zc_count, phase_adjust, rising_count


class LineRefineParams:
    """The RFDecode/SysParams constants the refinements read."""

    def __init__(self, rf):
        sp = rf.SysParams
        self.system = rf.system
        self.freq = rf.freq                     # MHz
        self.linelen = rf.linelen               # nominal samples per line (int)
        self.samplesperline = rf.freq / rf.linelen
        self.fsc_mhz = sp["fsc_mhz"]
        self.pilot_mhz = sp.get("pilot_mhz", sp["fsc_mhz"])
        self.hz_ire = rf.DecoderParams["hz_ire"]


def compute_line_bursts(p, demod, demod_burst, linelocs, line, linecount, prev_phaseadjust, lineoffset=0):
    """Field.compute_line_bursts (core.py:2999-3047).  Returns (rising,
    phase_adjust) or (None, None)."""
    line = line + lineoffset
    s = int(linelocs[line])
    s_rem = linelocs[line] - s
    lfreq = p.samplesperline * get_linelen(linelocs, line, linecount, lineoffset, p.linelen)
    fsc_mhz_inv = 1 / p.fsc_mhz
    bstart = int(21 * fsc_mhz_inv * lfreq)
    bend = int(28 * fsc_mhz_inv * lfreq)
    burstarea = np.asarray(demod_burst[s + bstart: s + bend], dtype=np.float32)
    if len(burstarea) == 0:
        return None, None
    burstarea = burstarea - np.float32(nb_mean_f32(burstarea))
    threshold = np.sqrt(np.mean(np.square(burstarea - np.mean(burstarea))))
    bd = np.asarray(demod[s + bstart: s + bend], dtype=np.float32)
    bd = bd - np.float32(nb_mean_f32(bd))
    if float(np.max(np.abs(bd))) > 30 * p.hz_ire:
        return None, None
    zcburstdiv = (lfreq * fsc_mhz_inv) / 2
    phase_adjust = -prev_phaseadjust
    for _ in range(2):
        zc_count, phase_adjust, rising_count = clb_findbursts(burstarea, threshold, bstart, s_rem, zcburstdiv,
                                                              phase_adjust)
    rising = rising_count > (zc_count / 2)
    return rising, -phase_adjust


def compute_burst_offsets(p, demod, demod_burst, linelocs, linecount, prev_phaseadjust):
    """FieldNTSC.compute_burst_offsets (core.py:3361-3388) for a field whose
    own phase_adjust_median is still 0 (so the previous field's value seeds
    every line).  Returns (field14, adjs, phase_adjust_median)."""
    rising_sum = 0
    adjs = {}
    for l in range(NTSC_BURST_LINES):
        rising, phase_adjust = compute_line_bursts(p, demod, demod_burst, linelocs, l, linecount, prev_phaseadjust)
        if rising is None:
            continue
        adjs[l] = phase_adjust / 2
        even_line = not (l % 2)
        rising_sum += 1 if (even_line and rising) else 0
    field14 = rising_sum > (len(adjs.keys()) // 4)
    median = np.median([adjs[a] for a in adjs]) * 2 if adjs else np.nan
    return field14, adjs, median


def compute_deriv_error(linelocs, baserr):
    """Field.compute_deriv_error (core.py:2553-2564)."""
    ll = np.asarray(linelocs, dtype=np.float64)
    dd = np.abs(np.diff(np.diff(ll))) > 4
    d1 = np.zeros(len(ll), bool)
    d1[1:-1] = dd
    d2 = np.zeros(len(ll), bool)
    d2[2:] = dd
    return np.asarray(baserr, bool) | d1 | d2


def fix_badlines(system, linelocs_in, linebad, linelocs_backup_in=None, lineoffset=0):
    """Field.fix_badlines (core.py:2566-2589).  Returns (linelocs, linebad)."""
    linebad = compute_deriv_error(linelocs_in, linebad)
    linelocs = np.array(linelocs_in, dtype=np.float64, copy=True)
    if linelocs_backup_in is not None:
        backup = np.array(linelocs_backup_in, dtype=np.float64, copy=True)
        nan = np.isnan(linelocs)
        linelocs[nan] = backup[nan]
    for l in np.where(linebad)[0]:
        prevgood = l - 1
        while prevgood >= 0 and linebad[prevgood]:
            prevgood -= 1
        nextgood = l + 1
        while nextgood < len(linelocs) and linebad[nextgood]:
            nextgood += 1
        firstcheck = 0 if system == "PAL" else 1
        if prevgood >= firstcheck and nextgood < (len(linelocs) + lineoffset):
            gap = (linelocs[nextgood] - linelocs[prevgood]) / (nextgood - prevgood)
            linelocs[l] = (gap * (l - prevgood)) + linelocs[prevgood]
    return linelocs, linebad


def refine_linelocs_burst(p, demod, demod_burst, linelocs2, linebad, linecount, first_field, prev_phaseadjust):
    """FieldNTSC.refine_linelocs_burst (core.py:3391-3440).  Returns
    (linelocs3, linebad, field_phase_id, phase_adjust_median)."""
    linelocs_adj = np.array(linelocs2, dtype=np.float64, copy=True)
    linebad = np.array(linebad, dtype=bool, copy=True)
    field14, adjs_new, median = compute_burst_offsets(p, demod, demod_burst, linelocs2, linecount,
                                                      prev_phaseadjust)
    for l in range(1, NTSC_BURST_LINES):
        if l not in adjs_new:
            linebad[l] = True
    adjs = {}
    for l in range(NTSC_BURST_LINES):
        if not (np.isnan(linelocs_adj[l]) or linebad[l]):
            lfreq = p.samplesperline * get_linelen(linelocs2, l, linecount, 0, p.linelen)
            if l in adjs_new:
                adjs[l] = adjs_new[l] * lfreq * (1 / p.fsc_mhz)
    if len(adjs.keys()):
        adjs_median = np.median([adjs[a] for a in adjs])
        lastvalid_adj = adjs_median
        for l in range(NTSC_BURST_LINES):
            if l in adjs and (adjs[l] - adjs_median >= -2) and (adjs[l] - adjs_median <= 2):
                linelocs_adj[l] += adjs[l]
                lastvalid_adj = adjs[l]
            else:
                linelocs_adj[l] += lastvalid_adj
        map4 = {(True, True): 1, (False, False): 2, (True, False): 3, (False, True): 4}
        field_phase_id = map4[(bool(first_field), bool(field14))]
    else:
        field_phase_id = 1
    return linelocs_adj, linebad, field_phase_id, median


def ntsc_field_lines(p, demod, demod_burst, linelocs2, linebad, linecount, first_field, prev_phaseadjust):
    """FieldNTSC.process after the hsync refine (core.py:3490-3495): burst
    refine, fix_badlines with linelocs2 as backup, subcarrier shift.
    Returns (linelocs, linebad, field_phase_id, phase_adjust_median)."""
    ll3, bad, fpid, median = refine_linelocs_burst(p, demod, demod_burst, linelocs2, linebad, linecount,
                                                   first_field, prev_phaseadjust)
    ll4, bad = fix_badlines("NTSC", ll3, bad, linelocs2)
    shift = (FSC_PHASE_DEG / 360) / p.fsc_mhz * p.freq
    return np.array(ll4) - shift, bad, fpid, median


def angular_mean(zcs):
    """angular_mean_helper (utils.py:1288-1302) and the mean angle of
    refine_linelocs_pilot (core.py:3081-3084), in cycles [0, 1)."""
    x = np.asarray(zcs, dtype=np.float64)
    x2 = x - x.astype(np.int32)
    angles = np.exp(1j * x2 * np.pi * 2)
    am = np.angle(np.mean(angles)) / (np.pi * 2)
    return 1 + am if am < 0 else am


def phase_distance(x, c):
    """utils.py:1305-1315."""
    d = (x - np.floor(x)) - c
    if d < -0.5:
        d += 1
    elif d > 0.5:
        d -= 1
    return d


def refine_linelocs_pilot(p, demod_pilot, linelocs):
    """FieldPAL.refine_linelocs_pilot (core.py:3054-3093) with lineoffset 0."""
    linelocs = np.array(linelocs, dtype=np.float64, copy=True)
    plen = {}
    zcs = []
    length = 6 * (p.samplesperline * p.linelen)      # usectoinpx(6), line=None
    for l in range(PAL_PILOT_LINES):
        adjfreq = p.freq
        if l > 1:
            adjfreq /= (linelocs[l] - linelocs[l - 1]) / p.linelen
        plen[l] = (adjfreq / p.pilot_mhz) / 2
        begin = linelocs[l] + 0.0
        ls = slice(int(begin), int(begin + length + 1))
        lsoffset = linelocs[l] - ls.start
        pilots = np.asarray(demod_pilot[ls], dtype=np.float32)
        peakloc = int(np.argmax(np.abs(pilots)))
        zc_base = calczc(pilots, peakloc, 0)
        if zc_base is not None:
            zc = (zc_base - lsoffset) / plen[l]
        else:
            zc = zcs[-1] if len(zcs) else 0
        zcs.append(zc)
    am = angular_mean(zcs)
    for l in range(PAL_PILOT_LINES):
        linelocs[l] += (phase_distance(zcs[l], am) * plen[l]) * 1
    return linelocs


def pal_field_lines(p, demod_pilot, linelocs2, linebad):
    """FieldPAL.process after the hsync refine (core.py:3221-3226): two pilot
    passes, then fix_badlines.  Returns (linelocs, linebad)."""
    ll3 = refine_linelocs_pilot(p, demod_pilot, linelocs2)
    ll3a = refine_linelocs_pilot(p, demod_pilot, ll3)
    return fix_badlines("PAL", ll3a, linebad)
