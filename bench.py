#!/usr/bin/env python3
"""Throughput of the RF demod + TBC hot path on one MI355X.

Workload (one "step"): full NTSC field decode, batched -- a synthetic NTSC
capture (667,333 int16 samples at 40 MSPS per field, sync + burst + 75%
bars, audio FM carriers, 40 dB AWGN, +-50 ns line jitter) resident in HBM is
demodulated block by block (32768-sample blocks at stride 31712, FP32, all
outputs of RFDecode.demodblock incl. audio stage 1); `--fields` fields are
located in it (sync pulses on the GPU, the reference's field walk and line-0
logic on the host, verified against each field's final line positions),
every field's line locations are refined in FP64 (hsync edge, colour-burst
alignment, bad-line repair, subcarrier shift), and every field is resampled
to a 910 x 263 uint16 4fsc frame.  Consecutive steps overlap on two streams
with two buffer sets (stream A demodulates step k while the host walks step
k-1 and stream B aligns and resamples it); the per-stage times come from a
separate sequential pass.  The "plain" figure decodes a single field per
step (one field read of 29 blocks), sequentially.

Prints one JSON line.  value = usable RF samples demodulated per second
(Msamples/s, whole job, inputs already in HBM).
"""

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

import __graft_entry__

METRIC = "RF Msamples/s (×realtime @40 MSPS) demod+TBC, batched vs plain, % HBM roofline"
HBM_PEAK_GBS = 8000.0
DEMOD_THREADS = 1024      # threads per k_demod_block workgroup (the traffic file is keyed by grid size)
NTSC_FIELD_SAMPLES = 667_334


def algorithmic_bytes(nblocks, S, nfields, pixels_per_field, field_samples, audio):
    """Contract bytes: input read once, every demod output written once, the
    TBC re-read of demod and the TBC write (block overlap recompute, FFT
    intermediates and tables excluded)."""
    U = nblocks * S
    demod = U * (2 + 4 * 4 + 4) + (U // 32) * 2 * 4 * (1 if audio else 0)
    tbc = nfields * field_samples * 4 + nfields * pixels_per_field * 2
    return demod, tbc, U


def capture_samples(nfields):
    """Capture length holding `nfields` complete field reads after the
    partial field at its start (each read spans 29 blocks, about 1.4 fields)."""
    return (nfields + 2) * NTSC_FIELD_SAMPLES + 100_000


def run_gpu(args, torch, rf, cap, nfields, steps, warmup, single=None, packed=False):
    """Time the located decode.  single: (start, startloc, buflen) of one
    field read -- the plain figure decodes just that field per step.
    packed: the capture in the packed 10-bit format (4 samples in 5 bytes)."""
    from lddecode_amd import capture
    data = cap.data
    start = 0.0
    if single is not None:
        st, org, n = single
        data = data[org: org + n + rf.blocklen - rf.block_stride]
        start = st - org
    if packed:
        from lddecode_amd import synth
        data = synth.pack_10bit(data)
    x = torch.from_numpy(np.ascontiguousarray(data)).to(rf.device)
    audio = True
    dec = capture.CaptureDecoder(rf, x, audio=audio, locate=True, max_fields=nfields, start=start, audio_rate=44100)
    stream = torch.cuda.current_stream(rf.device)

    for _ in range(warmup):
        dec.step()
        dec.dropouts()
        dec.dropout_lists()
    torch.cuda.synchronize()
    if packed:
        # the pipelined figure only
        steps_seq = 0
    else:
        steps_seq = steps
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(9)] for _ in range(max(steps_seq, 1))]
    walk_s, lists_s = [], []
    t0 = time.perf_counter()
    for i in range(len(ev)):
        e = ev[i]
        e[0].record(stream)
        dec.demod()
        e[1].record(stream)
        dec.pulses()
        e[2].record(stream)
        t = time.perf_counter()
        dec.locate()
        walk_s.append(time.perf_counter() - t)
        dec.refine()
        e[3].record(stream)
        dec.lines()
        e[4].record(stream)
        dec.tbc()
        e[5].record(stream)
        dec.audio_stage()
        e[8].record(stream)
        dec.measure_facts()
        dec.verify()
        e[6].record(stream)
        # dropout detection (on by default in the reference's decode): the
        # GPU error search, then the host grouping of the positions
        dec.dropouts()
        e[7].record(stream)
        t = time.perf_counter()
        drops = dec.dropout_lists()
        lists_s.append(time.perf_counter() - t)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    def span(a, b):
        return float(np.mean([e[a].elapsed_time(e[b]) for e in ev]))
    out = dict(wall_s=wall, step_ms=span(0, 6), demod_ms=span(0, 1), pulses_ms=span(1, 2),
               locate_hsync_ms=span(2, 3), lines_ms=span(3, 4), tbc_ms=span(4, 5), audio_ms=span(5, 8),
               facts_ms=span(8, 6),
               walk_host_ms=1e3 * float(np.mean(walk_s)), relocated=dec.relocated, dropouts_ms=span(6, 7),
               dropout_lists_host_ms=1e3 * float(np.mean(lists_s)),
               dropout_errors=sum(len(p) for p in dec.dropout_det.positions_per_field()),
               dropout_entries=sum(len(d[0]) for d in drops))
    pipe_ms = None
    if single is None:
        # consecutive passes overlapped on two streams: stream A demodulates
        # pass k and finds its pulses while the host walks pass k-1 and
        # stream B aligns and resamples it; three buffer sets
        pipe = capture.PipelinedDecoder(rf, x, audio=audio, locate=True, max_fields=nfields, start=start,
                                        audio_rate=44100)
        for _ in range(warmup):
            pipe.step()
        pipe.join()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            pipe.step()
        pipe.join()
        e1.record(stream)
        torch.cuda.synchronize()
        pipe_ms = e0.elapsed_time(e1) / steps
        out["relocated_pipe"] = sum(d.relocated for d in pipe.sets)
        del pipe
    W, lines = rf.SysParams["outlinelen"], rf.SysParams["frame_lines"] // 2 + 1
    nf = len(dec.fields)
    db, tb, U = algorithmic_bytes(dec.nblocks, rf.block_stride, nf, W * lines, NTSC_FIELD_SAMPLES, audio)
    locs = dec.walk_state[0]
    first = next(c for c in locs if c.status == 0)
    out.update(pipe_ms=pipe_ms, demod_bytes=db, tbc_bytes=tb, usable=U, nblocks=dec.nblocks, nfields=nf,
               fields=dec.fields, walk=dec.walk_state, first_field=(first.start, first.startloc, first.buflen),
               bad_lines=float(dec.linebad_final[:nf * dec.L].float().mean().item()))
    return out


def efm_demod_ms(torch, cap, steps, warmup):
    """Per-pass demodulation time of the 64-field capture with the EFM
    output (decode_digital_audio, core.py:847-851, on by default in the
    reference's command line) next to the same pass without it."""
    from lddecode_amd.rfdecode import RFDecode
    rf = RFDecode(system="NTSC", decode_analog_audio=44100, decode_digital_audio=True, device=0)
    x = torch.from_numpy(cap.data).to(rf.device)
    nblocks = (len(cap.data) - rf.blocklen) // rf.block_stride + 1
    rv = rf.demod_blocks(x, 0, nblocks, cut=True, audio=True)
    out = dict(rv["video"], rfhpf=rv["rfhpf"], efm=rv["efm"], **rv["audio"])
    stream = torch.cuda.current_stream(rf.device)
    for _ in range(warmup):
        rf.demod_blocks(x, 0, nblocks, cut=True, audio=True, out=out)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        rf.demod_blocks(x, 0, nblocks, cut=True, audio=True, out=out)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def cpu_baseline(cap, walk, nsample, workers, min_seconds, max_seconds):
    """The same step on the host with the FP64 numpy/scipy oracle: the
    blocks of the first `nsample` located fields' reads are demodulated on a
    thread pool (one block per task, as DemodCache's workers), then each field
    (one per task) is located (compute_linelocs on its buffer), its lines are
    refined (hsync edge, colour-burst alignment, bad-line repair, subcarrier
    shift) and it is resampled to uint16 (the resampler loop is the oracle's
    compiled C restatement, as the reference runs it compiled with numba);
    its audio goes through stage 2 (numpy FFTs) and to 44.1 kHz int16 (the
    oracle's numpy form of the numba-compiled output loop).
    Passes over this fixed sample repeat until `min_seconds` have elapsed
    (stopping early past `max_seconds`).  Returns (Msamples/s, seconds,
    passes, blocks per pass)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import audio as OA
    from oracle import ctbc
    from oracle import dsp as O
    from oracle import lines as L
    from oracle import linelocs as LL
    from oracle import tbc as T
    from oracle.params import out_scale
    ctbc.lib()
    orf = O.OracleRF("NTSC", decode_analog_audio=True)
    S, N = O.block_stride(orf), orf.blocklen
    sp, dp = orf.SysParams, orf.DecoderParams
    W, lines = sp["outlinelen"], sp["frame_lines"] // 2 + 1
    vs = dp["vsync_ire"]
    osc = out_scale("NTSC", vs)
    lut = T.build_kaiser_lut()
    lp = L.LineRefineParams(orf)
    pp = LL.LocParams.from_rf(orf)
    plan = [(int(c.startloc), int(c.buflen)) for c in walk[0] if c.status == 0][:nsample]
    b0 = plan[0][0] // S
    b1 = max((org + n) // S for org, n in plan)
    data = cap.data[b0 * S: b1 * S + N - S].astype(np.float64)
    blocks = list(range(b1 - b0))

    afilt = {ch: OA.audio2_filter(orf, ch) for ch in ("left", "right")}
    fdiv = orf.Filters["audio_fdiv"]

    def demod(b):
        r = orf.demodblock(data=data[b * S:b * S + N], cut=True)
        return (r["video"]["demod"], r["video"]["demod_05"], r["video"]["demod_burst"], r["audio"]["audio_left"],
                r["audio"]["audio_right"])

    def field(args):
        off, n, demod_all, d05_all, burst_all, al, ar = args
        res = LL.compute_linelocs(pp, d05_all[off:off + n], None, off, 1)
        bad = L.compute_deriv_error(res.linelocs1, res.linebad)
        ll2, bad = T.refine_linelocs_hsync(d05_all[off:], res.linelocs1, bad, "NTSC", 40.0, orf.iretohz, vs)
        bad = L.compute_deriv_error(ll2, bad)
        ll4, _, _, _ = L.ntsc_field_lines(lp, demod_all[off:], burst_all[off:], ll2, bad, res.linecount,
                                          res.isFirstField, 0.0)
        locs, wow = T.computewow_scaled(ll4, orf.linelen, W, lines, 0)
        ds = ctbc.scale_field(demod_all[off:], locs, wow, lut, 0, W, lines * W)
        seg = {"audio_left": al[off // fdiv:(off + n) // fdiv], "audio_right": ar[off // fdiv:(off + n) // fdiv]}
        a2 = OA.audio_phase2(orf, seg, afilt)
        pcm = OA.downscale_audio_np(a2, ll4, sp, orf.linelen, fdiv, res.linecount, 0.0, 44100)
        return T.hz_to_output_array(ds.astype(np.float32), dp["ire0"], dp["hz_ire"], sp["outputZero"], vs, osc), pcm

    passes = 0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        while True:
            outs = list(ex.map(demod, blocks))
            demod_all = np.concatenate([o[0] for o in outs]).astype(np.float32)
            d05_all = np.concatenate([o[1] for o in outs]).astype(np.float32)
            burst_all = np.concatenate([o[2] for o in outs]).astype(np.float32)
            al = np.concatenate([o[3] for o in outs]).astype(np.float32)
            ar = np.concatenate([o[4] for o in outs]).astype(np.float32)
            list(ex.map(field, [(org - b0 * S, n, demod_all, d05_all, burst_all, al, ar) for org, n in plan]))
            passes += 1
            dt = time.perf_counter() - t0
            if dt >= min_seconds or dt >= max_seconds:
                break
    return passes * len(blocks) * S / dt / 1e6, dt, passes, len(blocks)


TRAFFIC_FILE = pathlib.Path(__file__).resolve().parent / "profiles" / "traffic_ntsc64.json"
COPY_ELEMS = 64 << 20          # calibration copy: 256 MiB read + 256 MiB written


def profile_step(args):
    """Profiling target: the batched step only, then one plain copy kernel of
    known size (torch.neg over COPY_ELEMS float32) that calibrates the HBM
    byte counters."""
    __graft_entry__.load_package()
    import torch
    from lddecode_amd import capture, synth
    from lddecode_amd.rfdecode import RFDecode
    rf = RFDecode(system="NTSC", decode_analog_audio=44100, device=0)
    nf = args.fields
    cap = synth.make_capture("NTSC", capture_samples(nf), seed=args.seed, snr_db=40.0)
    dec = capture.CaptureDecoder(rf, torch.from_numpy(cap.data).to(rf.device), audio=True, locate=True,
                                 max_fields=nf, audio_rate=44100)
    for _ in range(args.warmup + args.steps):
        dec.step()
    torch.cuda.synchronize()
    x = torch.ones(COPY_ELEMS, dtype=torch.float32, device=rf.device)
    y = torch.empty_like(x)
    torch.neg(x, out=y)
    torch.cuda.synchronize()
    print(f"# profile step done: {dec.nblocks} blocks", file=sys.stderr, flush=True)


def _pmc_pass(counter, args, workdir):
    """One rocprofv3 counter pass over `bench.py --profile-step` (a child
    process); returns {kernel key: [values per dispatch]}."""
    import sqlite3
    import subprocess
    out = pathlib.Path(workdir) / counter.lower()
    cmd = ["rocprofv3", "--pmc", counter, "-d", str(out), "-o", "run", "--", sys.executable, "-u",
           str(pathlib.Path(__file__).resolve()), "--profile-step", "--fields", str(args.fields),
           "--steps", "2", "--warmup", "1"]
    print("# " + " ".join(cmd), file=sys.stderr, flush=True)
    env = dict(os.environ, TMPDIR="/tmp")
    subprocess.run(["timeout", "-k", "10", "300"] + cmd, check=True, cwd="/tmp", env=env)
    dbs = sorted(out.rglob("*_results.db"))
    if not dbs:
        raise RuntimeError(f"rocprofv3 wrote no results under {out}")
    con = sqlite3.connect(str(dbs[0]))
    rows = con.execute("select dispatch_id, kernel_name, grid_size, sum(value) from counters_collection "
                       "where counter_name = ? group by dispatch_id", (counter,)).fetchall()
    got = {}
    for _, name, grid, v in rows:
        short = name.split("(")[0].replace("void ", "")
        if "neg_kernel" in name:
            short = "copy"
        got.setdefault((short, int(grid)), []).append(float(v))
    return got


def measure_traffic(args):
    """HBM traffic per launch of the step's kernels: FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) from two counter passes, each scaled by the
    ratio of known bytes to the counter's reading for the calibration copy."""
    import tempfile
    work = tempfile.mkdtemp(prefix="ld_traffic_", dir="/tmp")
    fetch = _pmc_pass("FETCH_SIZE", args, work)
    write = _pmc_pass("WRITE_SIZE", args, work)
    copy_kib = COPY_ELEMS * 4 / 1024
    fc = max(v for k, vs in fetch.items() if k[0] == "copy" for v in vs)
    wc = max(v for k, vs in write.items() if k[0] == "copy" for v in vs)
    fscale, wscale = copy_kib / fc, copy_kib / wc
    kernels = {}
    for key in sorted(set(fetch) & set(write)):
        if not key[0].startswith("ldgpu::"):
            continue
        f = float(np.mean(fetch[key])) * fscale * 1024
        w = float(np.mean(write[key])) * wscale * 1024
        kernels[f"{key[0]}@{key[1]}"] = {"kernel": key[0], "grid_threads": key[1], "dispatches": len(fetch[key]),
                                         "fetch_bytes": round(f), "write_bytes": round(w),
                                         "traffic_bytes": round(f + w)}
    res = {"fields": args.fields, "copy_bytes_each_way": COPY_ELEMS * 4,
           "copy_fetch_kib": fc, "copy_write_kib": wc, "fetch_scale": fscale, "write_scale": wscale,
           "command": f"python bench.py --measure-traffic --fields {args.fields}", "kernels": kernels}
    dest = pathlib.Path(args.traffic_out) if args.traffic_out else TRAFFIC_FILE
    dest.parent.mkdir(parents=True, exist_ok=True)
    dest.write_text(json.dumps(res, indent=1) + "\n")
    print(json.dumps(res, indent=1))


def load_traffic(kernel, grid_threads):
    """Measured HBM bytes per launch for a kernel of the batched step, or None."""
    try:
        res = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    k = res.get("kernels", {}).get(f"{kernel}@{grid_threads}")
    return None if k is None else k["traffic_bytes"]


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fields", type=int, default=64, help="NTSC fields per batched step")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--cpu-fields", type=int, default=2, help="fields in the CPU baseline sample (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="minimum CPU baseline time")
    ap.add_argument("--profile-step", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--measure-traffic", action="store_true",
                    help="run the two rocprofv3 counter passes and write " + str(TRAFFIC_FILE.name))
    ap.add_argument("--traffic-out", default=None, help="where --measure-traffic writes its JSON")
    args = ap.parse_args()
    if args.profile_step:
        return profile_step(args)
    if args.measure_traffic:
        return measure_traffic(args)
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
    cap = synth.make_capture("NTSC", capture_samples(nf), seed=args.seed, snr_db=40.0)
    gen_s = time.perf_counter() - t
    print(f"# generated {len(cap.data)} samples in {gen_s:.1f} s", file=sys.stderr, flush=True)

    batched = run_gpu(args, torch, rf, cap, nf, args.steps, args.warmup)
    print(f"# batched: {batched['pipe_ms']:.3f} ms/step overlapped, {batched['step_ms']:.3f} ms/step sequential, "
          f"{batched['nfields']} fields located, host walk {batched['walk_host_ms']:.3f} ms",
          file=sys.stderr, flush=True)
    if batched["nfields"] != nf:
        raise RuntimeError(f"located {batched['nfields']} fields, expected {nf}")
    plain = run_gpu(args, torch, rf, cap, 1, args.steps, args.warmup, single=batched["first_field"])
    packed = run_gpu(args, torch, rf, cap, nf, args.steps, args.warmup, packed=True)
    print(f"# packed 10-bit input: {packed['pipe_ms']:.3f} ms/step overlapped", file=sys.stderr, flush=True)
    print(f"# plain: {plain['step_ms']:.3f} ms/step", file=sys.stderr, flush=True)
    efm_ms = efm_demod_ms(torch, cap, args.steps, args.warmup)
    print(f"# demod with the EFM output: {efm_ms:.3f} ms/pass", file=sys.stderr, flush=True)

    value = batched["usable"] / (batched["pipe_ms"] / 1e3) / 1e6
    plain_value = plain["usable"] / (plain["step_ms"] / 1e3) / 1e6
    demod_gbs = batched["demod_bytes"] / (batched["demod_ms"] / 1e3) / 1e9
    tbc_gbs = batched["tbc_bytes"] / (batched["tbc_ms"] / 1e3) / 1e9
    path_gbs = (batched["demod_bytes"] + batched["tbc_bytes"]) / (batched["pipe_ms"] / 1e3) / 1e9

    demod_traffic = load_traffic("ldgpu::k_demod_block", batched["nblocks"] * DEMOD_THREADS)
    tbc_traffic = load_traffic("ldgpu::k_tbc_pixels", nf * (rf.SysParams["frame_lines"] // 2 + 1) * 4 * 256)
    cpu = None
    if args.cpu_fields > 0:
        workers = min(os.cpu_count() or 1, 16)
        rate, secs, passes, nblk = cpu_baseline(cap, batched["walk"], args.cpu_fields, workers,
                                                args.cpu_seconds, 3 * args.cpu_seconds)
        cpu = {"value": round(rate, 3), "unit": "Msamples/s", "cores": workers, "kind": "port",
               "sample": f"FP64 numpy/scipy oracle on a {workers}-thread pool, the same step on the first "
                         f"{args.cpu_fields} located fields of the batch ({nblk} blocks demodulated; per field "
                         f"line location, hsync refine, burst alignment, TBC to uint16, audio stage 2 and "
                         f"44.1 kHz output), {passes} passes in "
                         f"{secs:.1f} s wall"}

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batched["pipe_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"C2 full NTSC field decode, batched: {nf} fields per step, consecutive steps overlapped "
                        f"({batched['usable']} usable samples, {batched['nblocks']} blocks of 32768), "
                        "FP32 demod + line location (GPU sync pulses, host field walk) + FP64 hsync / burst "
                        "line positions + sinc-LUT TBC to 910x263 uint16 + analogue audio stage 2 and 44.1 kHz "
                        "int16 output",
            "system": "NTSC",
            "blocklen": 32768,
            "fields_per_step": nf,
        },
        "x_realtime": round(value / 40.0, 1),
        "plain": {"value": round(plain_value, 2), "unit": "Msamples/s", "ms_per_step": round(plain["step_ms"], 4),
                  "fields_per_step": 1, "blocks_per_step": plain["nblocks"],
                  "x_realtime": round(plain_value / 40.0, 1)},
        "packed_input": {"value": round(batched["usable"] / (packed["pipe_ms"] / 1e3) / 1e6, 2),
                         "unit": "Msamples/s", "ms_per_step": round(packed["pipe_ms"], 4),
                         "input": "the same capture in the packed 10-bit format (1.25 B/sample), read by "
                                  "k_demod_block directly"},
        "sequential_ms_per_step": round(batched["step_ms"], 4),
        "stages_ms": {"demod": round(batched["demod_ms"], 4), "pulses": round(batched["pulses_ms"], 4),
                      "locate_walk_and_hsync": round(batched["locate_hsync_ms"], 4),
                      "burst_lines": round(batched["lines_ms"], 4), "tbc": round(batched["tbc_ms"], 4),
                      "audio_stage2_44k1": round(batched["audio_ms"], 4),
                      "facts_verify": round(batched["facts_ms"], 4)},
        "host_walk_ms": round(batched["walk_host_ms"], 4),
        "efm": {"demod_ms_with_efm": round(efm_ms, 4), "demod_ms": round(batched["demod_ms"], 4),
                "note": "k_demod_block_efm: the same pass also writing the int16 EFM output (one more transform "
                        "per block); not in ms_per_step"},
        "dropouts": {"gpu_ms": round(batched["dropouts_ms"], 4),
                     "host_lists_ms": round(batched["dropout_lists_host_ms"], 4),
                     "error_samples": batched["dropout_errors"], "entries": batched["dropout_entries"],
                     "note": "Field.dropout_detect for every field of the step, after the decode (not in "
                             "ms_per_step): GPU time of the error search; host time of the positions' copy "
                             "and the native grouping into (line, start, end) entries"},
        "relocated_passes": batched["relocated"] + batched.get("relocated_pipe", 0),
        "roofline": {"bound": "hbm", "achieved": round(demod_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(demod_gbs / HBM_PEAK_GBS, 4), "traffic": demod_traffic,
                     "algorithmic_bytes": batched["demod_bytes"],
                     "kernel": "k_demod_block, one launch per step (22.25 B per usable sample)"},
        "roofline_tbc": {"achieved": round(tbc_gbs, 1), "frac": round(tbc_gbs / HBM_PEAK_GBS, 4),
                         "traffic": tbc_traffic, "algorithmic_bytes": batched["tbc_bytes"],
                         "kernel": "k_tbc_lines + k_tbc_pixels (traffic: k_tbc_pixels)"},
        "roofline_path": {"achieved": round(path_gbs, 1), "frac": round(path_gbs / HBM_PEAK_GBS, 4)},
        "bad_line_fraction": round(batched["bad_lines"], 4),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
