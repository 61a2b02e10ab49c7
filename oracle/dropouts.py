"""FP64 / float32 restatement of ld-decode's dropout detection: the
per-sample error mask over a field buffer, its run grouping and the
conversion to TBC line coordinates (Field.dropout_detect_demod,
build_errlist, dropout_errlist_to_tbc, dropout_detect; core.py:2808-2996).
Test infrastructure only.

Conventions kept from the reference:
  * nb_std is numba's np.std of the float32 rfhpf buffer: the mean from a
    float32 sum in order divided in float64, then the float64 sum of squared
    deviations in order, its square root.
  * The level arrays are float32 (np.full_like of the float32 buffers), so
    demod / demod_05 are compared in float32 with float32 levels; rfhpf is
    compared with the float32 of +-3 std (NumPy keeps float32 with a Python
    float); demod_raw with freq_hz / 2 in float64 (numba promotes).
  * The sync-area minimum covers [int(ll[l]), int(ll[l]) + hsync_len) on
    ordinary lines and [int(ll[l]), int(ll[l+1])) on the vertical-sync
    lines of get_vsync_lines, painted line by line in order.
"""

import numpy as np


def nb_std_f32_loop(x):
    """numba's np.std of a float32 array (numba's array_var, utils.py:1268
    nb_std): the mean is a float32 running sum divided in float64, then the
    float64 running sum of squared deviations.  Scalar loops: small inputs."""
    x = np.asarray(x, dtype=np.float32)
    c = np.float32(0.0)
    for v in x:
        c = np.float32(c + v)
    m = float(c) / len(x)
    ssd = 0.0
    for v in x.astype(np.float64):
        d = v - m
        ssd += d * d
    return (ssd / len(x)) ** 0.5


def nb_std_f32(x):
    """nb_std_f32_loop with the running sums as cumulative sums (numpy's
    cumsum adds in order, without the pairwise blocking of np.sum)."""
    x = np.asarray(x, dtype=np.float32)
    m = float(np.cumsum(x, dtype=np.float32)[-1]) / len(x)
    d = x.astype(np.float64) - m
    return float(np.cumsum(d * d)[-1] / len(x)) ** 0.5


def get_vsync_lines(system, is_first):
    """Field.get_vsync_lines (core.py:2816-2827)."""
    rv = list(range(1, 10 if is_first else 9))
    if system == "PAL":
        rv += list(range(311 if is_first else 310, 318))
    return rv


def dropout_detect_demod(p, rfhpf, demod, demod_raw, demod_05, linelocs, hsync_max, is_first, lineoffset,
                         linecount, video_rot, rfstd=None):
    """core.py:2830-2887 over one field buffer.  p: LocParams-like (system,
    ire0, hz_ire, vsync_ire, freq); hsync_max: LT['hsync'][1] of its
    compute_linelocs; rfstd: nb_std(rfhpf) (computed when None)."""
    pal = p.system == "PAL"
    f32 = np.float32
    if rfstd is None:
        rfstd = nb_std_f32(rfhpf)
    thr = f32(rfstd * 3)
    rf = np.asarray(rfhpf, dtype=np.float32)
    iserr_rf1 = (rf < -thr) | (rf > thr)
    iserr = np.zeros(len(rf), dtype=bool)
    iserr[video_rot:] = iserr_rf1[:-video_rot]
    n = len(demod)
    valid_min = np.full(n, f32(p.iretohz(-70 if pal else -50)), np.float32)
    valid_max = np.full(n, f32(p.iretohz(150 if pal else 160)), np.float32)
    valid_min05 = np.full(n, f32(p.iretohz(-30)), np.float32)
    valid_max05 = np.full(n, f32(p.iretohz(115)), np.float32)
    hsync_len = int(hsync_max)
    vsync_ire = p.vsync_ire
    vlines = get_vsync_lines(p.system, is_first)
    sync_min = f32(p.iretohz(vsync_ire - 60 if pal else vsync_ire - 35))
    sync_min_05 = f32(p.iretohz(vsync_ire - 10))
    for l in range(1, len(linelocs)):
        a = int(linelocs[l])
        if l in vlines:
            b = int(linelocs[l + 1])
        else:
            b = a + hsync_len
        valid_min[a:b] = sync_min
        valid_min05[a:b] = sync_min_05
    iserr |= np.asarray(demod_raw, dtype=np.float32).astype(np.float64) > p.freq_hz_half
    d = np.asarray(demod, dtype=np.float32)
    iserr |= (d < valid_min) | (d > valid_max)
    d5 = np.asarray(demod_05, dtype=np.float32)
    iserr |= (d5 < valid_min05) | (d5 > valid_max05)
    iserr[:int(linelocs[lineoffset + 1])] = False
    iserr[int(linelocs[lineoffset + linecount + 1]):] = False
    return iserr


def build_errlist(errmap, linelocs, lineoffset, freq):
    """Field.build_errlist (core.py:2890-2908)."""
    errlist = []
    firsterr = errmap[np.nonzero(errmap >= linelocs[lineoffset])[0][0]]
    curerr = (firsterr, firsterr)
    for e in errmap:
        if e > curerr[0] and e <= (curerr[1] + 20):
            pad = (e - curerr[0]) * 1.7
            pad = min(pad, freq * 12)
            curerr = (curerr[0], curerr[0] + pad)
        elif e > firsterr:
            errlist.append((curerr[0] - 8, curerr[1] + 4))
            curerr = (e, e)
    errlist.append(curerr)
    return errlist


def errlist_to_tbc(errlist, linelocs, lineoffset, linecount, outlinelen):
    """Field.dropout_errlist_to_tbc (core.py:2911-2975)."""
    dropouts = []
    if len(errlist) == 0:
        return dropouts
    errlistc = list(errlist)
    lo = -lineoffset
    curerr = errlistc.pop(0)
    while len(errlistc) > 0 and curerr[0] < linelocs[lineoffset]:
        curerr = errlistc.pop(0)
    for line in range(lineoffset, linecount + lineoffset):
        while curerr is not None and (linelocs[line] <= curerr[0] <= linelocs[line + 1]):
            span = linelocs[line + 1] - linelocs[line]
            start_linepos = int(((curerr[0] - linelocs[line]) / span) * outlinelen)
            end_linepos = int(np.round(((curerr[1] - linelocs[line]) / span) * outlinelen))
            first_line = line + 1 + lo
            if end_linepos > outlinelen:
                num_lines = end_linepos // outlinelen
                dropouts.append((first_line, start_linepos, outlinelen))
                for k in range(num_lines - 1):
                    dropouts.append((first_line + k + 1, 0, outlinelen))
                dropouts.append((first_line + num_lines, 0, np.remainder(end_linepos, outlinelen)))
            else:
                dropouts.append((first_line, start_linepos, end_linepos))
            curerr = errlistc.pop(0) if len(errlistc) else None
    return dropouts


def dropout_detect(iserr, linelocs, lineoffset, linecount, outlinelen, freq):
    """Field.dropout_detect (core.py:2978-2996): (lines, starts, ends)."""
    errmap = np.nonzero(iserr)[0]
    lines, starts, ends = [], [], []
    if len(errmap) > 0 and errmap[-1] > linelocs[lineoffset]:
        errlist = build_errlist(errmap, linelocs, lineoffset, freq)
        for r in errlist_to_tbc(errlist, linelocs, lineoffset, linecount, outlinelen):
            lines.append(r[0] - 1)
            starts.append(int(r[1]))
            ends.append(int(r[2]))
    return lines, starts, ends
