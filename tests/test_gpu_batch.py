"""The batched path (many captures, NTSC and PAL mixed, ld_demod_batch in
one launch per wave, optionally replayed as a HIP graph) against
(a) the per-capture decode (CaptureDecoder over each capture alone): the
    same fields, line locations and uint16 codes, bit for bit, since every
    block, walk and refinement sees the same inputs;
(b) the FP64 oracle on sampled fields: demod channels within 25 Hz and 1e-4
    relative, TBC codes within +-1 LSB on >= 99.9% of pixels when the
    oracle resamples its own demod at the GPU's final line locations.
"""

import numpy as np
import pytest

from lddecode_amd import batch as batchmod
from lddecode_amd import capture, synth
from oracle import dsp as O
from oracle import tbc as T
from oracle.params import out_scale

pytestmark = pytest.mark.gpu

# two NTSC and two PAL captures of uneven lengths (about 3 to 8 fields each)
PLAN = [("NTSC", 2_700_000), ("PAL", 3_900_000), ("NTSC", 4_300_000), ("PAL", 2_500_000)]
_CACHE = {}


def _batch(torch):
    if "batch" not in _CACHE:
        _CACHE["batch"] = synth.make_batch(PLAN, "cuda", seed=11, ntemplates=1)
    return _CACHE["batch"]


def _decode(gpu_rf, torch, graph, wave_blocks=40, min_fields=1, groups=1):
    b = _batch(torch)
    rfs = {"NTSC": gpu_rf("NTSC", audio=True), "PAL": gpu_rf("PAL", audio=True)}
    got = {}

    def on_wave_of(index):            # decoder lane -> capture index
        def on_wave(info):
            torch.cuda.synchronize()
            for s in ("NTSC", "PAL"):
                if s not in info:
                    continue
                d = info[s]
                W, lines = rfs[s].SysParams["outlinelen"], rfs[s].SysParams["frame_lines"] // 2 + 1
                L = rfs[s].SysParams["frame_lines"] // 2 + 1 + 10 + (3 if s == "PAL" else 0)
                for f in d["fields"]:
                    b0, nb, slot0 = d["lanes"][f["lane"]]
                    start = f["demod_offset"] - slot0 * rfs[s].block_stride + b0 * rfs[s].block_stride
                    row = f["out_offset"] // (W * lines)
                    got[(index[f["lane"]], start)] = dict(
                        u16=d["out_u16"][row * W * lines:(row + 1) * W * lines].cpu().numpy(),
                        ll=d["linelocs"][f["lineloc_offset"]:f["lineloc_offset"] + L].cpu().numpy(),
                        demod=d["buffers"]["demod"][f["demod_offset"]:f["demod_offset"] + f["demod_len"]]
                        .cpu().numpy(),
                        first=f["first_field"])
        return on_wave

    decs = []
    for grp in batchmod.split_captures(b.lengths, groups):
        decs.append(batchmod.BatchDecoder(rfs, b.data, [b.systems[k] for k in grp], [b.offsets[k] for k in grp],
                                          [b.lengths[k] for k in grp], wave_blocks=wave_blocks,
                                          min_fields=min_fields, graph=graph, on_wave=on_wave_of(grp)))
    n = batchmod.run_interleaved(decs) if groups > 1 else decs[0].run()
    torch.cuda.synchronize()
    return decs[0] if groups == 1 else decs, got, n


def _per_capture(gpu_rf, torch, k):
    b = _batch(torch)
    rf = gpu_rf(b.systems[k], audio=True)
    dec = capture.CaptureDecoder(rf, b.capture(k), locate=True, max_fields=64)
    out = dec.step().clone()
    torch.cuda.synchronize()
    W, lines = dec.W, dec.linesout
    res = {}
    for j, f in enumerate(dec.fields):
        res[f["demod_offset"]] = dict(u16=out[j * W * lines:(j + 1) * W * lines].cpu().numpy(),
                                      ll=dec.linelocs[f["lineloc_offset"]:f["lineloc_offset"] + f["nlines"]]
                                      .cpu().numpy())
    return res


@pytest.mark.parametrize("graph,groups", [(False, 1), (True, 1), (True, 2)])
def test_batch_equals_per_capture_decode(gpu_rf, torch_cuda, graph, groups):
    torch = torch_cuda
    dec, got, n = _decode(gpu_rf, torch, graph, groups=groups)
    decs = dec if groups > 1 else [dec]
    waves = min(d.waves for d in decs)
    print(f"batch: {n} fields in {[d.waves for d in decs]} waves, relocated {[d.relocated for d in decs]}",
          flush=True)
    assert waves >= 2, "the test batch should take several waves"
    total = 0
    for k in range(len(PLAN)):
        want = _per_capture(gpu_rf, torch, k)
        mine = {start: v for (lane, start), v in got.items() if lane == k}
        assert sorted(mine) == sorted(want), f"capture {k}: fields {sorted(mine)} vs {sorted(want)}"
        for start, w in want.items():
            np.testing.assert_array_equal(mine[start]["ll"], w["ll"], err_msg=f"capture {k} field {start}")
            np.testing.assert_array_equal(mine[start]["u16"], w["u16"], err_msg=f"capture {k} field {start}")
        total += len(want)
    assert n == total and total >= 12


def test_batch_sampled_fields_match_oracle(gpu_rf, torch_cuda, oracle_rf):
    torch = torch_cuda
    dec, got, n = _decode(gpu_rf, torch, True)
    b = _batch(torch)
    lut = T.build_kaiser_lut()
    checked = set()
    for (lane, start), g in sorted(got.items()):
        s = b.systems[lane]
        if s in checked:
            continue
        checked.add(s)
        orf = oracle_rf(s, audio=True)
        S, N = O.block_stride(orf), orf.blocklen
        raw = b.capture(lane).cpu().numpy().astype(np.float64)
        nblk = -(-len(g["demod"]) // S)
        blocks = [orf.demodblock(data=raw[start + i * S:start + i * S + N], cut=True)["video"]["demod"]
                  for i in range(nblk)]
        ref = np.concatenate(blocks)[:len(g["demod"])]
        err = np.abs(g["demod"].astype(np.float64) - ref)
        assert err.max() <= 25.0 and np.all(err <= 1e-4 * np.abs(ref)), f"{s}: demod max err {err.max()} Hz"
        sp, dp = orf.SysParams, orf.DecoderParams
        W, lines = sp["outlinelen"], sp["frame_lines"] // 2 + 1
        lo = (2 if g["first"] else 3) if s == "PAL" else 0
        locs, wow = T.computewow_scaled(g["ll"], orf.linelen, W, lines, lo)
        ds = T.scale_field(ref.astype(np.float32), locs, wow, lut, lo, W, lines * W, coord_f32=True)
        u16 = T.hz_to_output_array(ds.astype(np.float32), dp["ire0"], dp["hz_ire"], sp["outputZero"],
                                   dp["vsync_ire"], out_scale(s, dp["vsync_ire"]))
        d = np.abs(g["u16"].astype(np.int64) - u16.astype(np.int64))
        print(f"{s} field at {start}: demod max {err.max():.2f} Hz, TBC max {d.max()} LSB, "
              f"{np.mean(d <= 1):.5f} within 1", flush=True)
        assert np.mean(d <= 1) >= 0.999
    assert checked == {"NTSC", "PAL"}
