"""The packed 10-bit capture format (4 samples in 5 bytes, big-endian bit
order) of lddecode/utils.py:356-410, restated in numpy.  Test
infrastructure only."""

import numpy as np


def unpack_data_4_40(indata, readlen, offset):
    """utils.py:371-390: unpack bytes (starting at a 5-byte group) and return
    readlen int16 samples (code - 512) << 6 from sample `offset` of the first
    group."""
    d = np.asarray(indata, dtype=np.uint8)
    g = len(d) // 5
    d16 = d[:g * 5].astype(np.uint16).reshape(g, 5)
    u = np.empty((g, 4), dtype=np.uint16)
    u[:, 0] = (d16[:, 0] << 2) | ((d16[:, 1] >> 6) & 0x03)
    u[:, 1] = ((d16[:, 1] & 0x3F) << 4) | ((d16[:, 2] >> 4) & 0x0F)
    u[:, 2] = ((d16[:, 2] & 0x0F) << 6) | ((d16[:, 3] >> 2) & 0x3F)
    u[:, 3] = ((d16[:, 3] & 0x03) << 8) | d16[:, 4]
    flat = u.reshape(-1)[offset:offset + readlen]
    return ((flat.astype(np.int32) - 512) << 6).astype(np.int16)


def load_packed(buf, sample, readlen):
    """load_packed_data_4_40 (utils.py:394-410) over an in-memory byte
    buffer: None when the buffer is too short."""
    start = (sample // 4) * 5
    offset = sample % 4
    needed = int(np.ceil(readlen * 5 // 4)) + 5
    indata = np.asarray(buf, dtype=np.uint8)[start:start + needed]
    if len(indata) < needed:
        return None
    return unpack_data_4_40(indata, readlen, offset)


def pack_4_40(codes):
    """Pack 10-bit codes (0..1023) four to five bytes, the layout
    unpack_data_4_40 reads (utils.py:356-369): the inverse used to build
    test captures."""
    c = np.asarray(codes, dtype=np.uint16)
    pad = (-len(c)) % 4
    c = np.concatenate([c, np.zeros(pad, np.uint16)]).reshape(-1, 4).astype(np.uint32)
    out = np.empty((len(c), 5), dtype=np.uint8)
    out[:, 0] = c[:, 0] >> 2
    out[:, 1] = ((c[:, 0] & 0x03) << 6) | (c[:, 1] >> 4)
    out[:, 2] = ((c[:, 1] & 0x0F) << 4) | (c[:, 2] >> 6)
    out[:, 3] = ((c[:, 2] & 0x3F) << 2) | (c[:, 3] >> 8)
    out[:, 4] = c[:, 3] & 0xFF
    return out.reshape(-1)


def codes_of(int16_samples):
    """10-bit codes of int16 capture samples (code - 512) << 6."""
    return ((np.asarray(int16_samples, dtype=np.int32) >> 6) + 512).astype(np.uint16)
