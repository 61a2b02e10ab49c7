"""GPU parity of analogue audio stage 2 and its 44.1 kHz output.

Stage 2 (FP32 transforms) is compared with the oracle's FP64
audio_phase2 on the GPU's own stage-1 audio of every located field read,
within a stated tolerance; the output-rate pass is exact arithmetic, so on
the GPU's own stage-2 audio its int16 pairs, failure flags and next offsets
must equal the oracle's bit for bit; end to end (oracle stage 2 + oracle
output vs the GPU) at most 1 LSB may differ, on a small fraction of pairs.
"""

import numpy as np
import pytest

from lddecode_amd import audio as A
from lddecode_amd import capture, synth
from oracle import audio as OA
from oracle import dsp as O

pytestmark = pytest.mark.gpu

# FP32 transforms over a block of ~100 kHz-deviation audio
STAGE2_MAX_HZ = 10.0
STAGE2_RMS_HZ = 1.0

_DEC = {}


def _decoded(gpu_rf, system):
    if system not in _DEC:
        import torch
        rf = gpu_rf(system, audio=True)
        fl = 667_334 if system == "NTSC" else 800_000
        cap = synth.make_capture(system, 5 * fl, seed=7, snr_db=40.0, jitter_ns=50, audio=True)
        dec = capture.CaptureDecoder(rf, torch.from_numpy(cap.data).cuda(), locate=True, audio=True, max_fields=8)
        dec.step()
        offs = [A.time_offset(rf.SysParams, k, bool(f["first_field"]), 44100) for k, f in enumerate(dec.fields)]
        dec.audio_output(44100, offs)
        res = dec.audio_results()
        torch.cuda.synchronize()
        host = {k: dec.buffers[k].cpu().numpy() for k in ("audio_left", "audio_right")}
        a2 = [t.cpu().numpy() for t in dec.audio_out.a2]
        _DEC[system] = (rf, dec, offs, res, host, a2, dec.linelocs.cpu().numpy())
    return _DEC[system]


@pytest.mark.parametrize("system", ["NTSC", "PAL"])
def test_stage2_and_output_match_oracle(gpu_rf, system):
    rf, dec, offs, res, host, a2, final = _decoded(gpu_rf, system)
    orf = O.OracleRF(system, decode_analog_audio=True)
    filters = {ch: OA.audio2_filter(orf, ch) for ch in ("left", "right")}
    sp = rf.SysParams
    fdiv = int(rf.Filters["audio_fdiv"])
    assert len(dec.fields) >= 3 and len(res) == len(dec.fields)
    exact = total = 0
    for k, f in enumerate(dec.fields):
        a0, n = f["demod_offset"] // fdiv, f["demod_len"] // fdiv
        seg = {c: host[c][a0:a0 + n] for c in ("audio_left", "audio_right")}
        want = OA.audio_phase2(orf, seg, filters)
        got = {"audio_left": a2[0][f["a2_offset"]:f["a2_offset"] + n],
               "audio_right": a2[1][f["a2_offset"]:f["a2_offset"] + n]}
        for c in got:
            d = got[c] - want[c]
            assert np.abs(d).max() <= STAGE2_MAX_HZ, (k, c, np.abs(d).max())
            assert np.sqrt(np.mean(d * d)) <= STAGE2_RMS_HZ, (k, c)
        ll = final[f["lineloc_offset"]:f["lineloc_offset"] + f["nlines"]]
        pcm, nxt, failed = res[k]
        # the output pass on the GPU's own stage-2 audio: exact
        o_pcm, o_nxt, o_failed = OA.downscale_audio(got, ll, sp, rf.linelen, fdiv, f["linecount"], offs[k], 44100)
        np.testing.assert_array_equal(pcm, o_pcm)
        assert failed == o_failed and nxt == o_nxt
        assert len(pcm) == 2 * (A.ticks(sp["line_period"], f["linecount"], offs[k], 44100)[2] - 1)
        # end to end against the oracle's own stage 2
        e_pcm, _, _ = OA.downscale_audio(want, ll, sp, rf.linelen, fdiv, f["linecount"], offs[k], 44100)
        dd = np.abs(pcm.astype(np.int32) - e_pcm.astype(np.int32))
        assert dd.max() <= 1
        exact += int((dd == 0).sum())
        total += len(dd)
    assert exact >= 0.97 * total
    # the synthetic carriers carry tones: the output is not silent
    assert max(np.abs(r[0]).max() for r in res) > 100


def test_stage2_clip_removal(gpu_rf):
    """Spikes above 500 kHz in the left channel zero both channels around
    them before the filter, as findpeaks finds them."""
    import torch
    rf = gpu_rf("NTSC", audio=True)
    orf = O.OracleRF("NTSC", decode_analog_audio=True)
    sp = rf.SysParams
    n = 29 * 991
    t = np.arange(n) / orf.audio["left"]["a1_freq"]
    rng = np.random.default_rng(4)
    left = (sp["audio_lfreq"] + 60000 * np.sin(2 * np.pi * 1000 * t) + rng.normal(size=n) * 50).astype(np.float32)
    right = (sp["audio_rfreq"] + 45000 * np.sin(2 * np.pi * 700 * t)).astype(np.float32)
    for p in (5, 4000, 17000, n - 4):      # rising pairs: findpeaks' peaks
        left[p] = np.float32(sp["audio_lfreq"] + 800000)
        left[p + 1] = np.float32(sp["audio_lfreq"] + 900000)
    assert len(OA.findpeaks(left - np.float32(sp["audio_lfreq"]), 500000)) == 4
    got = rf.audio_phase2({"audio_left": torch.from_numpy(left).cuda(), "audio_right": torch.from_numpy(right).cuda()})
    want = OA.audio_phase2(orf, {"audio_left": left, "audio_right": right})
    for c in ("audio_left", "audio_right"):
        d = got[c].cpu().numpy() - want[c]
        assert got[c].dtype == torch.float64
        assert np.abs(d).max() <= STAGE2_MAX_HZ and np.sqrt(np.mean(d * d)) <= STAGE2_RMS_HZ, c


def test_field_downscale_audio(gpu_rf):
    """Field.downscale(audio=44100) on one field read's own buffers gives
    the batched output for that field."""
    import torch
    from lddecode_amd import field as fieldmod
    rf, dec, offs, res, host, a2, final = _decoded(gpu_rf, "NTSC")
    k = 1
    f = dec.fields[k]
    fdiv = int(rf.Filters["audio_fdiv"])
    a0, n = f["demod_offset"] // fdiv, f["demod_len"] // fdiv
    aud = {c: torch.from_numpy(np.ascontiguousarray(host[c][a0:a0 + n])).cuda() for c in ("audio_left", "audio_right")}
    sl = slice(f["demod_offset"], f["demod_offset"] + f["demod_len"])
    video = {"demod": dec.buffers["demod"][sl].contiguous()}
    fd = fieldmod.Field(rf, {"video": video, "audio": aud},
                        linelocs=final[f["lineloc_offset"]:f["lineloc_offset"] + f["nlines"]])
    fd.linecount, fd.isFirstField = f["linecount"], bool(f["first_field"])
    pcm = fd.downscale_audio(44100, timeoffset=offs[k])
    np.testing.assert_array_equal(pcm, res[k][0])
    assert fd.audio_next_offset == res[k][1]


def test_pipelined_audio_equals_sequential(gpu_rf):
    """PipelinedDecoder(audio_rate=44100) runs the audio stage in each
    pass's stage B: every pass of the same capture gives the sequential
    decoder's int16 pairs, next offsets and failure flags."""
    import torch
    rf = gpu_rf("NTSC", audio=True)
    fl = 667_334
    cap = synth.make_capture("NTSC", 5 * fl, seed=7, snr_db=40.0, jitter_ns=50, audio=True)
    x = torch.from_numpy(cap.data).cuda()
    seq = capture.CaptureDecoder(rf, x, locate=True, audio=True, max_fields=8, audio_rate=44100)
    seq.step()
    want = seq.audio_results()
    assert len(want) >= 3
    pipe = capture.PipelinedDecoder(rf, x, locate=True, audio=True, max_fields=8, audio_rate=44100)
    outs = [pipe.step() for _ in range(3)]
    outs = [o for o in outs if o is not None] + pipe.join()
    torch.cuda.synchronize()
    assert len(outs) == 3
    for d in outs:
        got = d.audio_results()
        assert len(got) == len(want)
        for (a, an, af), (b, bn, bf) in zip(got, want):
            np.testing.assert_array_equal(a, b)
            assert an == bn and af == bf
