"""Small host-side helpers used by set-up and per-line host logic.

These operate on a few thousand values at most (filter-delay calibration, the
sync-pulse state machine); sample streams never pass through them.
"""

import numpy as np


def genwave(rate, freq, initialphase=0.0):
    """FM synthesis by phase accumulation (utils.py:1009-1023): out[i] =
    sin(phi_i), phi_{i+1} = phi_i + pi * rate[i] / freq.  The reference wraps
    phi into (-pi, pi]; sin() is unchanged by the wrap."""
    rate = np.asarray(rate, dtype=np.float64)
    phi = np.empty(len(rate))
    if len(rate):
        phi[0] = initialphase
        np.cumsum(np.pi * rate[:-1] / freq, out=phi[1:])
        phi[1:] += initialphase
    return np.sin(phi)


def calczc(data, start_offset, target, edge=0, count=16):
    """First crossing of `target` in data[s : s+count+1], s = int(start_offset),
    with linear interpolation between the bracketing samples (utils.py:795-841).
    edge 0 picks rising if data[s] < target, else falling.  None if absent."""
    s = int(start_offset)
    seg = np.asarray(data[s:s + int(count + 1)], dtype=np.float64)
    if len(seg) < 2:
        return None
    if edge == 0:
        edge = 1 if seg[0] < target else -1
    if edge == 1:
        hits = np.flatnonzero((seg[:-1] < target) & (seg[1:] >= target))
    else:
        hits = np.flatnonzero((seg[:-1] > target) & (seg[1:] <= target))
    if not len(hits):
        return None
    i = int(hits[0])
    a, b = seg[i] - target, seg[i + 1] - target
    frac = -a / (-a + b) if (b - a) != 0 else 0.0
    return s + i + frac
