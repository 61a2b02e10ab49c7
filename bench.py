#!/usr/bin/env python3
"""Throughput of the RF demod + TBC hot path on one MI355X.

Workload (one "step"): full NTSC field decode, batched -- a synthetic NTSC
capture of `--fields` consecutive fields (667,333 int16 samples at 40 MSPS
each, sync + burst + 75% bars, audio FM carriers, 40 dB AWGN, +-50 ns line
jitter) resident in HBM is demodulated block by block (32768-sample blocks
at stride 31712, FP32, all outputs of RFDecode.demodblock incl. audio
stage 1), every field's line locations are refined in FP64, and every field
is resampled to a 910 x 263 uint16 4fsc frame.  The "plain" figure decodes a
single field per step the same way.

Prints one JSON line.  value = usable RF samples demodulated per second
(Msamples/s, whole job, inputs already in HBM).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

import __graft_entry__

METRIC = "RF Msamples/s (×realtime @40 MSPS) demod+TBC, batched vs plain, % HBM roofline"
HBM_PEAK_GBS = 8000.0
NTSC_FIELD_SAMPLES = 667_334


def algorithmic_bytes(nblocks, S, nfields, pixels_per_field, field_samples, audio):
    """Contract bytes: input read once, every demod output written once, the
    TBC re-read of demod and the TBC write (block overlap recompute, FFT
    intermediates and tables excluded)."""
    U = nblocks * S
    demod = U * (2 + 4 * 4 + 4) + (U // 32) * 2 * 4 * (1 if audio else 0)
    tbc = nfields * field_samples * 4 + nfields * pixels_per_field * 2
    return demod, tbc, U


def run_gpu(args, torch, rf, cap, nfields, steps, warmup, single=False):
    from lddecode_amd import capture
    fields, ll1 = capture.truth_fields(rf, cap, nfields, first_field=1)
    data = cap.data
    if single:
        # one field read: only the blocks its lines need, starting at its origin
        S, N = rf.block_stride, rf.blocklen
        org = fields[0]["demod_offset"]
        nb = int(np.ceil((ll1.max() + 3000) / S)) + 1
        data = data[org: org + (nb - 1) * S + N]
        fields = [dict(fields[0], demod_offset=0)]
    x = torch.from_numpy(np.ascontiguousarray(data)).to(rf.device)
    audio = True
    dec = capture.CaptureDecoder(rf, x, fields, ll1, audio=audio)
    stream = torch.cuda.current_stream(rf.device)

    def step():
        dec.demod()
        dec.refine()
        dec.tbc()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        e = ev[i]
        e[0].record(stream)
        dec.demod()
        e[1].record(stream)
        dec.refine()
        e[2].record(stream)
        dec.tbc()
        e[3].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    demod_ms = np.mean([e[0].elapsed_time(e[1]) for e in ev])
    hsync_ms = np.mean([e[1].elapsed_time(e[2]) for e in ev])
    tbc_ms = np.mean([e[2].elapsed_time(e[3]) for e in ev])
    step_ms = np.mean([e[0].elapsed_time(e[3]) for e in ev])
    W, lines = rf.SysParams["outlinelen"], rf.SysParams["frame_lines"] // 2 + 1
    db, tb, U = algorithmic_bytes(dec.nblocks, rf.block_stride, nfields, W * lines, NTSC_FIELD_SAMPLES, audio)
    return dict(wall_s=wall, step_ms=step_ms, demod_ms=demod_ms, hsync_ms=hsync_ms, tbc_ms=tbc_ms,
                demod_bytes=db, tbc_bytes=tb, usable=U, nblocks=dec.nblocks,
                bad_lines=float(dec.linebad.float().mean().item()))


def cpu_baseline(cap, nblocks, workers):
    """FP64 numpy/scipy oracle on the host: `nblocks` blocks demodulated by a
    thread pool (one block per task, as DemodCache's workers) plus the TBC of
    one field.  Returns (Msamples/s, seconds)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import dsp as O
    from oracle import tbc as T
    orf = O.OracleRF("NTSC", decode_analog_audio=True)
    S, N = O.block_stride(orf), orf.blocklen
    data = cap.data.astype(np.float64)
    lut = T.build_kaiser_lut()

    def one(b):
        return orf.demodblock(data=data[b * S:b * S + N], cut=True)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        outs = list(ex.map(one, range(nblocks)))
    demod = np.concatenate([o["video"]["demod"] for o in outs])
    ll = 2000 + np.arange(273) * 2542.22          # one field's line map
    locs, wow = T.computewow_scaled(ll, orf.linelen, 910, 263, 0)
    T.scale_field(demod.astype(np.float32), locs, wow, lut, 0, 910, 263 * 910)
    dt = time.perf_counter() - t0
    return nblocks * S / dt / 1e6, dt


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fields", type=int, default=64, help="NTSC fields per batched step")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--cpu-blocks", type=int, default=64, help="blocks in the CPU baseline sample (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (captures shard across GPUs "
              "with no communication; run one process per GPU on separate captures).", file=sys.stderr)
        sys.exit(2)
    __graft_entry__.load_package()
    import torch
    from lddecode_amd import synth
    from lddecode_amd.rfdecode import RFDecode

    rf = RFDecode(system="NTSC", decode_analog_audio=44100, device=0)
    t = time.perf_counter()
    nf = args.fields
    cap = synth.make_capture("NTSC", (nf + 2) * NTSC_FIELD_SAMPLES + 100_000, seed=args.seed, snr_db=40.0)
    gen_s = time.perf_counter() - t
    print(f"# generated {len(cap.data)} samples in {gen_s:.1f} s", file=sys.stderr, flush=True)

    batched = run_gpu(args, torch, rf, cap, nf, args.steps, args.warmup)
    print(f"# batched: {batched['step_ms']:.3f} ms/step", file=sys.stderr, flush=True)
    plain = run_gpu(args, torch, rf, cap, 1, args.steps, args.warmup, single=True)
    print(f"# plain: {plain['step_ms']:.3f} ms/step", file=sys.stderr, flush=True)

    value = batched["usable"] / (batched["step_ms"] / 1e3) / 1e6
    plain_value = plain["usable"] / (plain["step_ms"] / 1e3) / 1e6
    demod_gbs = batched["demod_bytes"] / (batched["demod_ms"] / 1e3) / 1e9
    tbc_gbs = batched["tbc_bytes"] / (batched["tbc_ms"] / 1e3) / 1e9
    path_gbs = (batched["demod_bytes"] + batched["tbc_bytes"]) / (batched["step_ms"] / 1e3) / 1e9

    cpu = None
    if args.cpu_blocks > 0:
        workers = min(os.cpu_count() or 1, 16)
        rate, secs = cpu_baseline(cap, args.cpu_blocks, workers)
        cpu = {"value": round(rate, 3), "unit": "Msamples/s", "cores": workers, "kind": "port",
               "sample": f"FP64 numpy/scipy oracle: {args.cpu_blocks} NTSC blocks demodulated on a "
                         f"{workers}-thread pool + TBC of one field, {secs:.1f} s wall"}

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batched["step_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"C2 full NTSC field decode, batched: {nf} fields per step "
                        f"({batched['usable']} usable samples, {batched['nblocks']} blocks of 32768), "
                        "FP32 demod + FP64 hsync line positions + sinc-LUT TBC to 910x263 uint16",
            "system": "NTSC",
            "blocklen": 32768,
            "fields_per_step": nf,
        },
        "x_realtime": round(value / 40.0, 1),
        "plain": {"value": round(plain_value, 2), "unit": "Msamples/s", "ms_per_step": round(plain["step_ms"], 4),
                  "fields_per_step": 1, "blocks_per_step": plain["nblocks"],
                  "x_realtime": round(plain_value / 40.0, 1)},
        "stages_ms": {"demod": round(batched["demod_ms"], 4), "hsync": round(batched["hsync_ms"], 4),
                      "tbc": round(batched["tbc_ms"], 4)},
        "roofline": {"bound": "hbm", "achieved": round(demod_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(demod_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "ld_demod_blocks launch sequence (22.25 B per usable sample)"},
        "roofline_tbc": {"achieved": round(tbc_gbs, 1), "frac": round(tbc_gbs / HBM_PEAK_GBS, 4)},
        "roofline_path": {"achieved": round(path_gbs, 1), "frac": round(path_gbs / HBM_PEAK_GBS, 4)},
        "bad_line_fraction": round(batched["bad_lines"], 4),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
