"""GPU parity of dropout detection: the error sample positions of every field
of a located batch against the oracle's dropout_detect_demod over the same
device buffers and final line locations, bit for bit, and the dropout lists
(build_errlist + dropout_errlist_to_tbc) against the oracle's.
"""

import numpy as np
import pytest

from lddecode_amd import capture, dropouts, synth
from oracle import dropouts as D
from oracle import dsp as O
from oracle import linelocs as LL

pytestmark = pytest.mark.gpu

_DEC = {}


def _spans(system):
    fl = 667_334 if system == "NTSC" else 800_000
    # carrier losses of a few samples to a line and more, inside the fields
    return [(fl + 300_000, 200), (fl + 500_000, 3000), (2 * fl + 100_000, 40), (2 * fl + 400_000, 6),
            (3 * fl + 200_000, 9000)]


def _decoded(gpu_rf, system):
    if system not in _DEC:
        import torch
        rf = gpu_rf(system, audio=False)
        fl = 667_334 if system == "NTSC" else 800_000
        cap = synth.make_capture(system, 5 * fl, seed=3, snr_db=40.0, jitter_ns=50, audio=False,
                                 dropouts=_spans(system))
        dec = capture.CaptureDecoder(rf, torch.from_numpy(cap.data).cuda(), locate=True, max_fields=8)
        dec.step()
        dec.dropouts()
        per = dec.dropout_det.positions_per_field()
        lists = dec.dropout_lists()
        torch.cuda.synchronize()
        host = {k: v.cpu().numpy() for k, v in dec.buffers.items()}
        _DEC[system] = (rf, dec, per, lists, host, dec.linelocs.cpu().numpy())
    return _DEC[system]


def _oracle_field(rf, system, f, host, final):
    p = LL.LocParams.from_rf(O.OracleRF(system, decode_analog_audio=False))
    sl = slice(f["demod_offset"], f["demod_offset"] + f["demod_len"])
    ll = final[f["lineloc_offset"]:f["lineloc_offset"] + f["nlines"]]
    iserr = D.dropout_detect_demod(p, host["rfhpf"][sl], host["demod"][sl], host["demod_raw"][sl],
                                   host["demod_05"][sl], ll, f["hsync_max"], bool(f["first_field"]),
                                   f["lineoffset"], f["linecount"], rf.delays["video_rot"])
    want = D.dropout_detect(iserr, ll, f["lineoffset"], f["linecount"], rf.SysParams["outlinelen"], rf.freq)
    return np.flatnonzero(iserr), want


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_dropout_positions_and_lists_match_oracle(gpu_rf, system):
    rf, dec, per, lists, host, final = _decoded(gpu_rf, system)
    assert len(dec.fields) >= 3
    nfound = 0
    for k, f in enumerate(dec.fields):
        pos, want = _oracle_field(rf, system, f, host, final)
        np.testing.assert_array_equal(per[k], pos)
        assert tuple(lists[k]) == tuple(want)
        nfound += len(want[0])
    # every injected loss inside a located field is reported
    raw = np.concatenate([p + f["demod_offset"] + 1024 for p, f in zip(per, dec.fields)])
    for a0, ln in _spans(system):
        inside = any(f["demod_offset"] + 1024 + 3000 < a0 < f["demod_offset"] + 1024 + f["demod_len"] - 3000
                     for f in dec.fields)
        if inside:
            assert ((raw >= a0 - 100) & (raw <= a0 + ln + 300)).any(), (a0, ln)
    assert nfound > 0


def test_position_buffer_grows(gpu_rf):
    """A position buffer smaller than the batch's error count is grown and
    the search repeated: the same positions."""
    rf, dec, per, lists, host, final = _decoded(gpu_rf, "NTSC")
    n = sum(len(p) for p in per)
    assert n > 16
    det = dropouts.DropoutDetector(rf, cap=16)
    m = sum(f["nlines"] for f in dec.fields)
    det.detect(dec.buffers, dec.fields, dec.linelocs[:m])
    got = det.positions_per_field()
    assert det.cap >= n
    for a, b in zip(got, per):
        np.testing.assert_array_equal(a, b)


def test_field_dropout_detect(gpu_rf):
    """Field.dropout_detect on one field's own buffers (the reference's
    per-field call) gives the batched result for that field."""
    import torch
    from lddecode_amd import field as fieldmod
    rf, dec, per, lists, host, final = _decoded(gpu_rf, "NTSC")
    k = max(range(len(dec.fields)), key=lambda j: len(lists[j][0]))
    f = dec.fields[k]
    sl = slice(f["demod_offset"], f["demod_offset"] + f["demod_len"])
    video = {n: torch.from_numpy(np.ascontiguousarray(host[n][sl])).cuda() for n in rf.video_fields}
    fd = fieldmod.Field(rf, {"video": video, "rfhpf": torch.from_numpy(np.ascontiguousarray(host["rfhpf"][sl])).cuda()},
                        linelocs=final[f["lineloc_offset"]:f["lineloc_offset"] + f["nlines"]],
                        lineoffset=f["lineoffset"])
    fd.isFirstField, fd.linecount, fd.hsync_max = bool(f["first_field"]), f["linecount"], f["hsync_max"]
    got = fd.dropout_detect()
    assert len(got[0]) > 0
    assert tuple(got) == tuple(lists[k])
