"""The packed 10-bit format restatement (oracle/packed.py) against hand-made
bytes and the bit layout the reference documents (utils.py:356-369)."""

import numpy as np

from oracle import packed as PK


def test_known_group():
    # samples 0x3FF, 0x000, 0x155, 0x2AA:
    # bits 1111111111 0000000000 0101010101 1010101010
    # -> 11111111 11000000 00000101 01010110 10101010
    b = np.array([0xFF, 0xC0, 0x05, 0x56, 0xAA], np.uint8)
    got = PK.unpack_data_4_40(b, 4, 0)
    codes = np.array([0x3FF, 0x000, 0x155, 0x2AA])
    np.testing.assert_array_equal(got, ((codes - 512) << 6).astype(np.int16))
    np.testing.assert_array_equal(PK.pack_4_40(codes), b)


def test_round_trip_and_offsets():
    rng = np.random.default_rng(3)
    codes = rng.integers(0, 1024, size=4001)
    buf = PK.pack_4_40(codes)
    want = ((codes.astype(np.int32) - 512) << 6).astype(np.int16)
    for sample in (0, 1, 2, 3, 5, 998):
        got = PK.load_packed(buf, sample, 1000)
        np.testing.assert_array_equal(got, want[sample:sample + 1000])
    assert PK.load_packed(buf, 3500, 1000) is None        # past the end
    np.testing.assert_array_equal(PK.codes_of(want), codes)


def test_synth_packer_round_trips_through_the_oracle():
    """The package's packer (used to build packed test captures) against the
    oracle's unpack_data_4_40."""
    from lddecode_amd import synth
    cap = synth.make_capture("NTSC", 10_000, seed=2, snr_db=40.0)
    buf = synth.pack_10bit(cap.data)
    np.testing.assert_array_equal(PK.unpack_data_4_40(buf, 10_000, 0), cap.data)
    np.testing.assert_array_equal(buf, PK.pack_4_40(PK.codes_of(cap.data)))
