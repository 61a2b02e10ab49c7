"""FP64 restatement of ld-decode's line location: sync-pulse detection,
pulse classification, vertical-blanking analysis, line 0 and the per-line
start positions (linelocs1) that the hsync refine starts from, plus the
field-to-field walk of LDdecode.decodefield.  Test infrastructure only.

Reference: lddecode/core.py:1761-2481 (Field.get_timings .. compute_linelocs),
core.py:2122-2155 (skip_check), core.py:4018-4098 (decodefield offsets),
utils.py:1101-1157 (findpulses).

Conventions kept from the reference, as written there:
  * Field.usectoinpx(x) with no line is x * (freq / linelen) * linelen
    (samplesperline * get_linelen(), core.py:1682-1685; linelen is the
    integer rf.linelen, core.py:386-387).
  * findpulses compares float32 samples against a float64 threshold
    (numba promotes), keeps runs of samples <= high of length 0..5000 that
    do not start at index 0 and that end inside the buffer
    (utils.py:1101-1141); its low threshold is ignored.
  * getpulses averages with NumPy (float32 pairwise sums), takes a float32
    median, and NumPy 2 keeps float32 in the arithmetic with Python floats.
  * skip_check's nb_median is numba's: the middle value, or for an even
    count the float32 sum of the two middle values halved in float64.
  * get_timings falls back to SysParams['hsyncPulseUS'] (microseconds, not
    samples) when no pulse has an hsync-like length (core.py:1781-1784).
  * getLine0's ``np.max(conf_prev - 10, 0)`` is conf_prev - 10 (the 0 is an
    axis), so the fallback confidence is max(conf_prev - 10, 10).
"""

from dataclasses import dataclass, field as dc_field
from typing import List, Optional

import numpy as np

HSYNC, EQPL1, VSYNC, EQPL2 = range(4)        # core.py:1584
MAX_SYNCLEN = 5000                           # findpulses_numba_raw default


def inrange(a, mi, ma):
    """utils.py:787-788."""
    return (a >= mi) & (a <= ma)


def nb_round(m):
    """utils.py:1238-1239 (round half to even)."""
    return int(np.round(m))


# ---------------------------------------------------------------- pulses
def findpulses(sync_ref, high):
    """findpulses_numba_raw (utils.py:1101-1141): (starts, lengths) int64."""
    x = np.asarray(sync_ref, dtype=np.float32).astype(np.float64)
    below = x <= float(high)
    if len(below) == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    prev = np.concatenate([[False], below[:-1]])
    starts = np.flatnonzero(below & ~prev)
    ends = np.flatnonzero(~below & prev)
    starts = starts[:len(ends)]                 # a trailing pulse never ends
    lens = ends - starts
    keep = (lens >= 0) & (lens <= MAX_SYNCLEN) & (starts != 0)
    return starts[keep].astype(np.int64), lens[keep].astype(np.int64)


def np_mean_f32(a):
    """np.mean of a float32 array: NumPy's pairwise float32 sum / count."""
    a = np.asarray(a, dtype=np.float32)
    return np.float32(np.add.reduce(a, dtype=np.float32) / np.float32(len(a))) if len(a) else np.float32(np.nan)


def numba_median(a):
    """nb_median (utils.py:1233-1235) under numba, returned as float64."""
    s = np.sort(np.asarray(a, dtype=np.float32))
    n = len(s)
    if n == 0:
        return float("nan")
    if n % 2:
        return float(s[n // 2])
    return float(np.float32(s[n // 2 - 1] + s[n // 2])) / 2


@dataclass
class LocParams:
    """The RFDecode / SysParams / DecoderParams values line location reads."""
    system: str
    freq: float                 # MHz
    linelen: int                # rf.linelen (rounded)
    ire0: float
    hz_ire: float
    vsync_ire: float
    hsyncPulseUS: float
    eqPulseUS: float
    vsyncPulseUS: float
    numPulses: int
    firstFieldH: tuple
    field_lines: tuple
    frame_lines: int
    line_period: float
    blocklen: int = 32768
    blockcut: int = 1024
    block_stride: int = 31712
    hsync_tolerance: float = 0.4

    @classmethod
    def from_rf(cls, rf):
        sp, dp = rf.SysParams, rf.DecoderParams
        return cls(system=rf.system, freq=rf.freq, linelen=rf.linelen, ire0=dp["ire0"], hz_ire=dp["hz_ire"],
                   vsync_ire=dp["vsync_ire"], hsyncPulseUS=sp["hsyncPulseUS"], eqPulseUS=sp["eqPulseUS"],
                   vsyncPulseUS=sp["vsyncPulseUS"], numPulses=sp["numPulses"], firstFieldH=tuple(sp["firstFieldH"]),
                   field_lines=tuple(sp["field_lines"]), frame_lines=sp["frame_lines"],
                   line_period=sp["line_period"], blocklen=getattr(rf, "blocklen", 32768),
                   block_stride=getattr(rf, "blocklen", 32768) - 1024 - 32)

    @property
    def freq_hz_half(self):
        return self.freq * 1000000 / 2

    @property
    def linefreq(self):
        return (self.freq / self.linelen) * self.linelen

    def usectoinpx(self, x):
        return x * self.linefreq

    def iretohz(self, ire):
        return self.ire0 + self.hz_ire * ire

    def hztoire(self, hz):
        return (hz - self.ire0) / self.hz_ire

    @property
    def outlinecount(self):
        return self.frame_lines // 2 + 1

    @property
    def proclines(self):
        """compute_linelocs: outlinecount + lineoffset (0) + 10, +3 for PAL."""
        return self.outlinecount + 10 + (3 if self.system == "PAL" else 0)

    @property
    def readlen(self):
        """LDdecode.readlen (core.py:3625-3631)."""
        if self.system == "PAL":
            return self.linelen * 400
        return ((self.linelen * 350) // 16384) * 16384


def getpulses(p, demod_05, fields_written=1):
    """Field.getpulses (core.py:2242-2322).  Returns (starts, lens, new_ire0)
    where new_ire0 is the recalibrated rf ire0 of the first-field retry (or
    None); starts is None when no vertical sync pulse was seen."""
    d = np.asarray(demod_05, dtype=np.float32)
    hi = p.iretohz(-20)
    starts, lens = findpulses(d, hi)
    new_ire0 = None
    if len(starts) == 0:
        if fields_written:
            return starts, lens, None
        new_ire0 = float(np.percentile(d, 15))
        q = LocParams(**{**p.__dict__, "ire0": new_ire0})
        starts, lens = findpulses(d, q.iretohz(-20))
        if len(starts) == 0:
            return starts, lens, new_ire0
        p = q
    minlength = p.usectoinpx(10)
    vsync_locs, vsync_means = [], []
    for i, (s, n) in enumerate(zip(starts, lens)):
        if n > minlength:
            vsync_locs.append(i)
            vsync_means.append(np_mean_f32(d[int(s + p.freq):int(s + n - p.freq)]))
    if not vsync_means:
        return None, None, new_ire0
    synclevel = np.float32(np.median(np.array(vsync_means, dtype=np.float32)))
    if abs(np.float32((synclevel - np.float32(p.ire0)) / np.float32(p.hz_ire)) - np.float32(p.vsync_ire)) < 5:
        return starts, lens, new_ire0
    black_means = []
    r = list(range(vsync_locs[0] - 5, vsync_locs[0])) + list(range(vsync_locs[-1] + 1, vsync_locs[-1] + 6))
    for i in r:
        if i < 0 or i >= len(starts):
            continue
        if inrange(lens[i], p.freq * 0.75, p.freq * 2.5):
            black_means.append(np_mean_f32(d[int(starts[i] + p.freq * 5):int(starts[i] + p.freq * 20)]))
    blacklevel = np.float32(np.median(np.array(black_means, dtype=np.float32)))
    high = np.float32((blacklevel + synclevel) / np.float32(2))
    s2, l2 = findpulses(d, float(high))
    return s2, l2, new_ire0


# ------------------------------------------------------- classification
def get_timings(p, lens):
    """Field.get_timings (core.py:1761-1810): LT ranges in samples."""
    hsync_typical = p.usectoinpx(p.hsyncPulseUS)
    cmin = p.usectoinpx(p.hsyncPulseUS - 1.75)
    cmax = p.usectoinpx(p.hsyncPulseUS + 2)
    hl = [n for n in lens if inrange(n, cmin, cmax)]
    med = float(np.median(hl)) if hl else p.hsyncPulseUS
    LT = {"hsync_median": med}
    LT["hsync"] = (med + p.usectoinpx(-0.5), med + p.usectoinpx(0.5))
    off = med - hsync_typical
    LT["hsync_offset"] = off
    LT["eq"] = (p.usectoinpx(p.eqPulseUS - 0.5) + off, p.usectoinpx(p.eqPulseUS + 0.5) + off)
    LT["vsync"] = (p.usectoinpx(p.vsyncPulseUS * 0.5) + off, p.usectoinpx(p.vsyncPulseUS + 1) + off)
    return LT


def pulse_qualitycheck(p, prev, cur):
    """core.py:1812-1823; prev/cur are (type, (start, len), ...)."""
    if prev[0] > 0 and cur[0] > 0:
        lo, hi = 0.4, 0.6
    elif prev[0] == 0 and cur[0] == 0:
        lo, hi = 0.9, 1.1
    else:
        lo, hi = 0.4, 1.1
    return bool(inrange((cur[1][0] - prev[1][0]) / p.linelen, lo, hi))


def run_vblank_state_machine(p, pulses, LT):
    """core.py:1826-1911.  pulses: list of (start, len).  Returns (done,
    validpulses) with validpulses entries (type, (start, len), good)."""
    done = False
    valid = []
    state_end = 0
    state_length = None
    half = p.numPulses / 2
    for pl in pulses:
        spulse = None
        state = valid[-1][0] if valid else -1
        n = pl[1]
        if state == -1:
            if inrange(n, *LT["hsync"]):
                spulse = (HSYNC, pl)
        elif state == HSYNC:
            if inrange(n, *LT["hsync"]):
                spulse = (HSYNC, pl)
            elif inrange(n, *LT["eq"]):
                spulse = (EQPL1, pl)
                state_length = half
            elif inrange(n, *LT["vsync"]):
                spulse = (VSYNC, pl)
        elif state == EQPL1:
            if inrange(n, *LT["eq"]):
                spulse = (EQPL1, pl)
            elif inrange(n, *LT["vsync"]):
                spulse = (VSYNC, pl)
                state_length = half
            elif inrange(n, *LT["hsync"]):
                spulse = (HSYNC, pl)
        elif state == VSYNC:
            if inrange(n, *LT["eq"]):
                spulse = (EQPL2, pl)
                state_length = half
            elif inrange(n, *LT["vsync"]):
                spulse = (VSYNC, pl)
            elif pl[0] > state_end and inrange(n, *LT["hsync"]):
                spulse = (HSYNC, pl)
        elif state == EQPL2:
            if inrange(n, *LT["eq"]):
                spulse = (EQPL2, pl)
            elif inrange(n, *LT["hsync"]):
                spulse = (HSYNC, pl)
                done = True
        if spulse is not None and spulse[0] != state:
            if spulse[1][0] < state_end:
                spulse = None
            elif state_length:
                state_end = spulse[1][0] + (state_length - 0.1) * p.linelen
                state_length = None
        if spulse is not None:
            good = pulse_qualitycheck(p, valid[-1], spulse) if valid else False
            valid.append((spulse[0], spulse[1], good))
        if done:
            return done, valid
    return done, valid


def refinepulses(p, starts, lens):
    """Field.refinepulses (core.py:1914-1949).  Returns (LT, validpulses)."""
    LT = get_timings(p, lens)
    raw = [(int(s), int(n)) for s, n in zip(starts, lens)]
    i = 0
    valid = []
    while i < len(raw):
        cur = raw[i]
        if inrange(cur[1], *LT["hsync"]):
            good = pulse_qualitycheck(p, valid[-1], (0, cur)) if valid else False
            valid.append((HSYNC, cur, good))
            i += 1
        elif i > 2 and inrange(cur[1], *LT["eq"]) and (len(valid) and valid[-1][0] == HSYNC):
            done, vb = run_vblank_state_machine(p, raw[i - 2:i + 24], LT)
            if done:
                valid.extend(vb[2:])
                i += len(vb) - 2
            else:
                i += 1
        else:
            i += 1
    return LT, valid


def get_blank_range(valid, start=0):
    """Field.getBlankRange (core.py:1952-1977)."""
    vt = np.array([v[0] for v in valid])
    vs = np.flatnonzero(vt[start:] == VSYNC)
    firstvsync = vs[0] + start if len(vs) else None
    if firstvsync is None or firstvsync < 10:
        return None, None
    for newstart in range(firstvsync - 10, firstvsync - 4):
        bl = np.flatnonzero(vt[newstart:] > 0)
        if len(bl) == 0:
            continue
        firstblank = bl[0] + newstart
        hl = np.flatnonzero(vt[firstblank:] == 0)
        if len(hl) == 0:
            continue
        lastblank = hl[0] + firstblank - 1
        if (lastblank - firstblank) > 12:
            return int(firstblank), int(lastblank)
    return None, None


def process_vblank(p, valid, start, limit=None):
    """Field.processVBlank (core.py:1988-2088).  Returns (line0loc,
    isfirstfield, firstblank, confidence, zeroed): zeroed is True on the
    fallback that also sets the field's sync_confidence to 0 (core.py:2082)."""
    firstblank, lastblank = get_blank_range(valid, start)
    lastvalid = len(valid) if limit is None else start + limit
    if firstblank is None or firstblank > lastvalid:
        return None, None, None, None, False
    loc_presync = valid[firstblank - 1][1][0]
    pt = np.array([v[0] for v in valid[firstblank:]])
    pstart = np.array([v[1][0] for v in valid[firstblank:]])
    plen = np.array([v[1][1] for v in valid[firstblank:]])
    npul = p.numPulses
    for i in (VSYNC, EQPL1, EQPL2):
        ptmatch = pt == i
        grouploc = None
        for j in range(0, lastblank - firstblank):
            if ptmatch[j:j + npul].all():
                if ptmatch[j:j + npul + 4].sum() != npul:
                    break
                gaps = np.diff(np.diff(pstart[j:j + npul]))
                lengths = np.diff(plen[j:j + npul])
                if np.max(gaps) < (p.freq * 0.2) and np.max(lengths) < (p.freq * 0.2):
                    grouploc = j
                    break
        if grouploc is None:
            continue
        firstloc = valid[firstblank + grouploc][1][0]
        distfroml1 = ((i - 1) * npul) * 0.5
        dist = (firstloc - loc_presync) / p.linelen
        hdist = nb_round(dist * 2)
        isfirst = (hdist % 2) == (p.firstFieldH[1] != 1)
        if (distfroml1 * 2) % 2:
            isfirst = not isfirst
        eqgap = p.firstFieldH[int(isfirst)]
        line0 = firstloc - ((eqgap + distfroml1) * p.linelen)
        return int(line0), bool(isfirst), firstblank, 100, False
    if valid[firstblank - 1][2] and valid[firstblank][2] and valid[lastblank][2] and valid[lastblank + 1][2]:
        gap1 = valid[firstblank][1][0] - valid[firstblank - 1][1][0]
        gap2 = valid[lastblank + 1][1][0] - valid[lastblank][1][0]
        if p.system == "PAL" and inrange(np.abs(gap2 - gap1), 0, p.freq * 1):
            isfirst = bool(inrange(gap1 / p.linelen, 0.45, 0.55))
        elif p.system == "NTSC" and inrange(np.abs(gap2 + gap1), p.linelen * 1.4, p.linelen * 1.6):
            isfirst = bool(inrange(gap1 / p.linelen, 0.95, 1.05))
        else:
            return None, None, None, 0, True
        return valid[firstblank - 1][1][0], isfirst, firstblank, 50, False
    return None, None, None, 0, False


def compute_line_len(p, valid):
    """Field.computeLineLen (core.py:2091-2119)."""
    longrun = [-1, -1]
    currun = None
    for i, v in enumerate([q[0] for q in valid]):
        if v != 0:
            if currun is not None and currun[1] > longrun[1]:
                longrun = currun
            currun = None
        elif currun is None:
            currun = [i, 0]
        else:
            currun[1] += 1
    if currun is not None and currun[1] > longrun[1]:
        longrun = currun
    lens = []
    for i in range(longrun[0] + 1, longrun[0] + longrun[1]):
        ll = valid[i][1][0] - valid[i - 1][1][0]
        if inrange(ll / p.linelen, 0.95, 1.05):
            lens.append(ll)
    return float(np.mean(lens)) if lens else p.linelen


# ------------------------------------------------------------- one field
@dataclass
class PrevField:
    """What getLine0 / skip_check read from the previous field."""
    valid: bool
    startloc: int
    linelocs: np.ndarray        # its final line locations
    linecount: int
    isFirstField: bool
    sync_confidence: int
    skip_score: int             # its skip_check()


def skip_check(p, demod, linelocs, lineoffset=0):
    """Field.skip_check (core.py:2122-2155) over a field's demod buffer and
    final line locations."""
    score = 0
    vsync_lines = 0
    length = p.usectoinpx(p.line_period)
    for line in range(p.outlinecount, p.outlinecount + 8):
        b = linelocs[line + lineoffset]
        sl = slice(int(b), int(b + length + 1))
        ire = p.hztoire(numba_median(demod[sl]))
        if inrange(ire, p.vsync_ire - 10, p.vsync_ire / 2):
            vsync_lines += 1
        elif inrange(ire, -5, 5):
            score += 1
        else:
            score -= 1
    if vsync_lines >= 2:
        return 100
    if vsync_lines == 1 and score > 0:
        return 50
    if score > 0:
        return 25
    return 0


def field_facts(p, demod, final, isfirst):
    """What a field leaves for its successor: (final linelocs[linecount],
    the eight skip_check medians) over its own buffer (core.py:2122-2155,
    :2196-2203; PAL's linecount 312 / 313 and lineoffset 2 / 3 from
    FieldPAL.process)."""
    if p.system == "PAL":
        lc, lo = (312, 2) if isfirst else (313, 3)
    else:
        lc, lo = (263 if isfirst else 262), 0
    length = p.usectoinpx(p.line_period)
    meds = []
    for line in range(p.outlinecount, p.outlinecount + 8):
        b = final[line + lo]
        meds.append(numba_median(demod[int(b):int(b + length + 1)]))
    return float(final[lc]), meds


def skip_score_from(p, meds):
    """skip_check's score from its eight medians (core.py:2139-2155)."""
    score = vsync_lines = 0
    for m in meds:
        ire = p.hztoire(m)
        if inrange(ire, p.vsync_ire - 10, p.vsync_ire / 2):
            vsync_lines += 1
        elif inrange(ire, -5, 5):
            score += 1
        else:
            score -= 1
    if vsync_lines >= 2:
        return 100
    if vsync_lines == 1 and score > 0:
        return 50
    if score > 0:
        return 25
    return 0


@dataclass
class FieldLines:
    """compute_linelocs results plus the state later fields read."""
    linelocs1: Optional[np.ndarray]
    linebad: Optional[np.ndarray]
    nextfieldoffset: Optional[float]
    isFirstField: Optional[bool] = None
    linecount: int = 262
    sync_confidence: int = 100
    vblank_next: Optional[int] = None
    line0loc: Optional[float] = None
    meanlinelen: Optional[float] = None
    skipdetected: bool = False
    new_ire0: Optional[float] = None
    validpulses: list = dc_field(default_factory=list)
    LT: Optional[dict] = None


def get_line0(p, valid, meanlinelen, prev: Optional[PrevField], startloc):
    """Field.getLine0 (core.py:2157-2239).  Returns (line0loc, vblank_next,
    isFirstField, sync_confidence)."""
    conf = 100
    limit = 100 if (prev is not None and prev.skip_score >= 50) else None
    l0_local, first_local, fb_local, conf_local, zeroed = process_vblank(p, valid, 0, limit)
    if zeroed:
        conf = 0
    l0_next, first_next, conf_next = None, None, None
    vblank_next = None
    if l0_local is not None:
        vblank_next, notfirst_next, _, conf_next, zeroed = process_vblank(p, valid, fb_local + 40)
        if zeroed:
            conf = 0
        if vblank_next is not None:
            first_next = not notfirst_next
            fieldlen = meanlinelen * p.field_lines[0 if first_next else 1]
            l0_next = nb_round(vblank_next - fieldlen)
            if l0_next < 0:
                conf = 10
    l0_prev, first_prev, conf_prev = None, None, None
    if prev is not None and prev.valid:
        frameoffset = startloc - prev.startloc
        l0_prev = prev.linelocs[prev.linecount] - frameoffset
        first_prev = not prev.isFirstField
        conf_prev = prev.sync_confidence
    if l0_local is not None and l0_next is not None and l0_prev is not None:
        first_all = (first_local + first_prev + first_next) >= 2
        return float(np.median([l0_local, l0_next, l0_prev])), vblank_next, bool(first_all), conf
    if l0_local is not None and conf_local > 50:
        return l0_local, vblank_next, first_local, min(conf, 90)
    if l0_prev is not None:
        return l0_prev, vblank_next, first_prev, min(conf, max(conf_prev - 10, 10))
    if l0_next is not None:
        return l0_next, vblank_next, first_next, conf_next
    return None, None, None, conf


def compute_linelocs(p, demod_05, prev: Optional[PrevField] = None, startloc=0, fields_written=1):
    """Field.compute_linelocs (core.py:2325-2481) over one field buffer."""
    starts, lens, new_ire0 = getpulses(p, demod_05, fields_written)
    if starts is None or len(starts) == 0:
        off = None if fields_written else int(p.freq * 1e6)
        return FieldLines(None, None, off, new_ire0=new_ire0)
    LT, valid = refinepulses(p, starts, lens)
    meanlinelen = compute_line_len(p, valid)
    line0loc, lastlineloc, isfirst, conf = get_line0(p, valid, meanlinelen, prev, startloc)
    linecount = 263 if isfirst else 262
    proclines = p.proclines
    res = FieldLines(None, None, None, isFirstField=isfirst, linecount=linecount, sync_confidence=conf,
                     vblank_next=lastlineloc, line0loc=line0loc, meanlinelen=meanlinelen, new_ire0=new_ire0,
                     validpulses=valid, LT=LT)
    if lastlineloc is not None:
        skipdetected = (lastlineloc - line0loc) / p.linelen < (linecount - 5)
    else:
        skipdetected = False
    res.skipdetected = skipdetected
    if line0loc is None:
        res.nextfieldoffset = p.linelen * 200
        return res
    lastline = (int(starts[-1]) - line0loc) / meanlinelen
    if lastline < proclines:
        res.nextfieldoffset = line0loc - (meanlinelen * 20)
        return res
    ldict, ldist = {}, {}
    for v in valid:
        lineloc = (v[1][0] - line0loc) / meanlinelen
        rlineloc = nb_round(lineloc)
        dist = np.abs(lineloc - rlineloc)
        if skipdetected:
            le = linecount - ((lastlineloc - v[1][0]) / meanlinelen)
            rle = nb_round(le)
            de = np.abs(le - rle)
            if v[0] == 0 and rlineloc > 23 and de < dist:
                lineloc, rlineloc, dist = le, rle, de
        if dist > p.hsync_tolerance or (rlineloc in ldict and dist > ldist[rlineloc]):
            continue
        if rlineloc > 0 and not v[2]:
            if v[0] > 0 or (v[0] == 0 and rlineloc < 10):
                continue
        ldict[rlineloc] = v[1][0]
        ldist[rlineloc] = dist
    err = np.zeros(proclines, dtype=bool)
    ll = [ldict[l] if l in ldict else -1 for l in range(proclines)]
    fill = list(ll)
    oc = p.outlinecount
    if fill[0] < 0:
        nv = next((i for i in range(0, oc + 1) if ll[i] > 0), None)
        if nv is None:
            res.nextfieldoffset = line0loc + (p.linelen * oc - 7)
            return res
        fill[0] = fill[nv] - nv * meanlinelen
        if fill[0] < p.linelen:
            res.nextfieldoffset = line0loc + (p.linelen * oc - 7)
            return res
    for l in range(1, proclines):
        if fill[l] < 0:
            err[l] = True
            pv = next((i for i in range(l, -1, -1) if ll[i] > 0), None)
            nv = next((i for i in range(l, oc + 1) if ll[i] > 0), None)
            if pv is None:
                fill[l] = ll[nv] - p.linelen * (nv - l)
            elif nv is not None:
                avg = (ll[nv] - ll[pv]) / (nv - pv)
                fill[l] = ll[pv] + avg * (l - pv)
            else:
                fill[l] = ll[pv] + p.linelen * (l - pv)
    res.linelocs1 = np.array(fill, dtype=np.float64)
    res.linebad = err
    res.nextfieldoffset = fill[oc - 7] if lastlineloc is None else lastlineloc - (p.linelen * 8)
    return res


# ---------------------------------------------------------------- walk
def field_window(p, start):
    """decodefield's read (core.py:4030-4041) and DemodCache.read's block
    range (core.py:1358-1409): (readloc, startloc, nblocks) where the field
    buffer is the stitched output of blocks [startloc // S, + nblocks)."""
    readloc = max(int(start - p.blockcut), 0)
    begin = (readloc // p.blocklen) * p.blocklen
    length = (p.readlen // p.blocklen + 2) * p.blocklen
    S = p.block_stride
    b0, b1 = begin // S, (begin + length) // S
    return readloc, b0 * S, b1 - b0 + 1


def walk_fields(p, demod_05, nfields, start, finish, demod=None):
    """LDdecode.readfield / decodefield's field-to-field walk without the
    MTF / AGC redo: field k+1 starts at field k's start + offset
    (core.py:4089-4098, :4163-4191).  demod_05 / demod: stitched buffers
    (index i = raw sample i + blockcut).  finish(k, res, startloc) returns
    the field's final line locations (hsync + burst / pilot refinement).
    Without demod, skip_check is taken as 100.  A failed field leaves no
    previous field and jumps by its own offset (core.py:1640-1644,
    :4094-4095); an rf ire0 recalibrated by getpulses' retry is kept for
    the fields after it.  Returns one dict per field."""
    out = []
    prev = None
    written = 0
    S = p.block_stride
    for _ in range(nfields):
        readloc, startloc, nb = field_window(p, start)
        if startloc + nb * S > len(demod_05):
            break
        buf = demod_05[startloc:startloc + nb * S]
        res = compute_linelocs(p, buf, prev, startloc, written)
        if res.new_ire0 is not None:
            p = LocParams(**{**p.__dict__, "ire0": res.new_ire0})
        valid = res.linelocs1 is not None
        rec = dict(start=start, startloc=startloc, nblocks=nb, res=res, valid=valid)
        if valid:
            final = finish(len(out), res, startloc)
            lc = res.linecount
            if p.system == "PAL":                   # FieldPAL.process, core.py:3230-3231
                lc = 312 if res.isFirstField else 313
            skip = 100
            if demod is not None:
                _, meds = field_facts(p, demod[startloc:startloc + nb * S], final, res.isFirstField)
                skip = skip_score_from(p, meds)
            prev = PrevField(True, startloc, final, lc, res.isFirstField, res.sync_confidence, skip)
            rec["final"] = final
            offset = res.nextfieldoffset - (readloc - startloc)
            written += 1
        else:
            prev = None
            offset = res.nextfieldoffset if res.nextfieldoffset is not None else p.linelen * 200
        out.append(rec)
        start = start + offset
    return out
