#!/usr/bin/env python3
"""Throughput of the RF demod + TBC hot path on one MI355X.

Headline workload (one "step" = one pass over the whole batch): 64 synthetic
captures of uneven length (1-30 s, log-uniform from seed 4), NTSC and PAL
mixed (each system with probability 1/2), int16 at 40 MSPS, resident in HBM
(about 2.7e10 samples, 55 GB).  Each capture tiles seeded field templates
with +-50 ns per-line jitter, its FM carrier phase-accumulated over the
whole capture, with the analogue-audio carriers and 40 dB AWGN.  The batch
is decoded in waves: every wave demodulates every capture's next window of
blocks in ONE launch with both systems' filter tables resident (replayed as
a HIP graph together with the sync-pulse search), walks every
capture's fields on the host (the reference's field walk and line-0 logic,
native code), and refines (FP64 hsync, burst / pilot), resamples to uint16
4fsc fields and checks the walk, per system; two buffer sets overlap a
wave's demodulation with the previous wave's alignment and TBC.

value = RF samples of the batch decoded per second (each capture sample once;
blocks demodulated twice where windows overlap are work, not throughput).
The block length (8192, 16384, 32768 or 65536 samples) and the transform (FP32
passes or split-BF16 MFMA radix-64 stages) are the autotuner's choice
(lddecode_amd.tune: the fastest demod launch on a small mixed batch) unless
--blocklen / --fft fix them; the same batch is also decoded at the
reference's default (32768, FP32) and with 65536-sample MFMA blocks.

Also reported: the plain single-block path (one 32768-sample block per
launch), single NTSC and PAL fields, the one-capture 64-field pipelined
decode, the packed-input and EFM variants, the demod kernel's roofline, and
the FP64 numpy/scipy oracle on the host cores over a fixed subset.
"""

import argparse
import hashlib
import json
import os
import pathlib
import platform
import sys
import time

import numpy as np

import __graft_entry__

METRIC = "RF Msamples/s (×realtime @40 MSPS) demod+TBC, batched vs plain, % HBM roofline"
HBM_PEAK_GBS = 8000.0
ROOT = pathlib.Path(__file__).resolve().parent
TRAFFIC_FILE = ROOT / "profiles" / "traffic_c4.json"
COPY_ELEMS = 64 << 20          # calibration copy: 256 MiB read + 256 MiB written
KERNEL_SOURCES = ("ld-decode_amd/csrc/demod.hip", "ld-decode_amd/csrc/fft.hpp", "ld-decode_amd/csrc/internal.hpp")


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


def kernel_source_hash():
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        h.update((ROOT / f).read_bytes())
    return h.hexdigest()[:16]


def decoders(torch, blocklen=32768, fft="fp32"):
    from lddecode_amd import tune
    from lddecode_amd.rfdecode import RFDecode
    rfs = {s: RFDecode(system=s, blocklen=blocklen, decode_analog_audio=44100, device=0) for s in ("NTSC", "PAL")}
    for rf in rfs.values():
        rf.set_fft_variant(tune.VARIANTS[fft])
    return rfs


def wave_blocks_for(args, rfs):
    """args.wave_blocks counts 32768-sample blocks (31712 usable samples);
    other block lengths get the same samples per wave."""
    return max(1, round(args.wave_blocks * 31712 / rfs["NTSC"].block_stride))


def workload(torch, rfs, seed):
    from lddecode_amd import synth
    t = time.perf_counter()
    plan = synth.batch_plan(seed=seed)
    b = synth.make_batch(plan, rfs["NTSC"].device, seed=seed)
    torch.cuda.synchronize()
    log(f"generated {len(plan)} captures, {sum(b.lengths)} samples "
        f"({sum(s == 'PAL' for s in b.systems)} PAL) in {time.perf_counter() - t:.1f} s")
    return b


# ------------------------------------------------------------ C4 batch
def batch_decoders(args, rfs, b):
    """The batch split into `args.groups` decoders of about equal samples
    (LPT), each with its own streams and buffer sets, run taking turns on
    the host; their demod launches run one after the other (each fills the
    GPU), so the events around each measure that launch."""
    from lddecode_amd.batch import BatchDecoder, split_captures
    order = {}
    decs = []
    wb = wave_blocks_for(args, rfs)
    for grp in split_captures(b.lengths, args.groups):
        decs.append(BatchDecoder(rfs, b.data, [b.systems[k] for k in grp], [b.offsets[k] for k in grp],
                                 [b.lengths[k] for k in grp], wave_blocks=wb // args.groups,
                                 graph=True, audio=True, demod_order=order))
    return decs


def run_c4(args, torch, rfs, b):
    from lddecode_amd.batch import run_interleaved
    decs = batch_decoders(args, rfs, b)

    def one_pass():
        for d in decs:
            d.reset()
        return run_interleaved(decs)

    for i in range(args.warmup):
        t = time.perf_counter()
        nf = one_pass()
        torch.cuda.synchronize()
        log(f"warm-up pass {i}: {nf} fields, {sum(d.waves for d in decs)} waves, {time.perf_counter() - t:.2f} s")
    for d in decs:
        d.demod_timing = []       # events around every wave's demod graph replay
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        nf = one_pass()
        log(f"timed pass {i}: {nf} fields")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tim = [x for d in decs for x in d.demod_timing]
    ms = [e0.elapsed_time(e1) for e0, e1, _ in tim]
    by = [n for _, _, n in tim]
    for d in decs:
        d.demod_timing = None
    bys = {}
    for d in decs:
        for s, n in d.blocks_by_system.items():
            bys[s] = bys.get(s, 0) + n
    out = dict(ms_per_step=wall / args.steps * 1e3, fields=nf, waves=sum(d.waves for d in decs),
               block_stride=decs[0].S, blocklen=decs[0].N,
               blocks=sum(d.blocks_total for d in decs), blocks_by_system=bys,
               relocated=sum(d.relocated for d in decs), redone_waves=sum(d.redone_waves for d in decs),
               samples=sum(d.samples_total for d in decs))
    out.update(demod_launches=len(ms), demod_ms_avg=float(np.mean(ms)), demod_bytes_avg=float(np.mean(by)),
               demod_ms_total=float(np.sum(ms)) / args.steps)
    # the same launches, each alone on the GPU (synchronised around it): the
    # kernel's own rate without the other decoder's concurrent stage B
    iso = iso_demod(torch, decs[0])
    out.update(iso)
    del decs
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def iso_demod(torch, dec, reps=5):
    """The first wave's demod launch of a pass (a full-size wave), run alone
    on the GPU `reps` times (nothing else queued beside it)."""
    rows = []
    for _ in range(reps):
        dec.reset()
        torch.cuda.synchronize()
        dec.demod_timing = []
        it = dec.waves_iter()
        next(it)                      # plans wave 0 and queues its stage A only
        torch.cuda.synchronize()
        rows += [(e0.elapsed_time(e1), n) for e0, e1, n in dec.demod_timing]
        it.close()
    dec.demod_timing = None
    dec.reset()
    return dict(iso_ms_avg=float(np.mean([r[0] for r in rows])), iso_bytes_avg=float(np.mean([r[1] for r in rows])),
                iso_launches=len(rows))


# ------------------------------------------------------------ plain paths
def run_block(torch, rf, x, reps):
    """C1: one block per launch (RFDecode.demodblock's shape), cut outputs."""
    st = torch.cuda.current_stream(rf.device)
    out = None
    for _ in range(5):
        rv = rf.demod_blocks(x, 0, 1, cut=True, audio=True, out=out)
        out = dict(rv["video"], rfhpf=rv["rfhpf"], **rv["audio"])
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    e0.record(st)
    for _ in range(reps):
        rf.demod_blocks(x, 0, 1, cut=True, audio=True, out=out)
    e1.record(st)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t) / reps
    gpu = e0.elapsed_time(e1) / reps / 1e3
    return dict(ms=wall * 1e3, gpu_ms=gpu * 1e3, msps=rf.block_stride / wall / 1e6)


def run_field(torch, rf, x, reps):
    """C2 / C3 plain: one field read per step, located, refined and
    resampled, sequentially."""
    from lddecode_amd import capture
    dec = capture.CaptureDecoder(rf, x, locate=True, max_fields=1, audio=True)
    for _ in range(3):
        dec.step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        dec.step()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t) / reps
    return dict(ms=wall * 1e3, blocks=dec.nblocks, msps=dec.total / wall / 1e6, fields=len(dec.fields))


def run_pipelined(torch, rf, x, nfields, steps, warmup):
    """One NTSC capture, 64 fields per pass, consecutive passes overlapped."""
    from lddecode_amd import capture
    pipe = capture.PipelinedDecoder(rf, x, audio=True, locate=True, max_fields=nfields, audio_rate=44100)
    for _ in range(warmup):
        pipe.step()
    pipe.join()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        pipe.step()
    pipe.join()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) / steps * 1e3
    d = pipe.sets[0]
    return dict(ms=ms, msps=d.total / ms / 1e3, fields=len(d.fields), blocks=d.nblocks)


def efm_demod_ms(torch, x, steps):
    from lddecode_amd.rfdecode import RFDecode
    rf = RFDecode(system="NTSC", decode_analog_audio=44100, decode_digital_audio=True, device=0)
    nb = (x.numel() - rf.blocklen) // rf.block_stride + 1
    rv = rf.demod_blocks(x, 0, nb, cut=True, audio=True)
    out = dict(rv["video"], rfhpf=rv["rfhpf"], efm=rv["efm"], **rv["audio"])
    st = torch.cuda.current_stream(rf.device)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    res = {}
    for efm in (False, True):
        o = dict(out)
        if not efm:
            o.pop("efm")
        names = None if efm else [k for k in o]
        rf.demod_blocks(x, 0, nb, cut=True, audio=True, out=o, channels=names)
        e0.record(st)
        for _ in range(steps):
            rf.demod_blocks(x, 0, nb, cut=True, audio=True, out=o, channels=names)
        e1.record(st)
        torch.cuda.synchronize()
        res[efm] = e0.elapsed_time(e1) / steps
    return res[False], res[True], nb


# ------------------------------------------------------------ CPU baseline
def cpu_baseline(b, ncapt_per_system, nread_blocks, workers, min_seconds):
    """The FP64 numpy/scipy oracle on the host over a fixed subset of the
    batch: the first `nread_blocks` blocks of the first `ncapt_per_system`
    captures of each system, demodulated block by block on a thread pool of
    `workers` (DemodCache's worker model), then per capture the field walk,
    hsync refine, burst / pilot alignment and the TBC to uint16 (the
    resampler loop in the oracle's compiled C, as the reference runs it
    compiled with numba).  Passes repeat until `min_seconds` have elapsed.
    Returns (Msamples/s, seconds, passes, samples per pass)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import ctbc
    from oracle import dsp as O
    from oracle import lines as L
    from oracle import linelocs as LL
    from oracle import tbc as T
    from oracle.params import out_scale
    ctbc.lib()
    picks = []
    for s in ("NTSC", "PAL"):
        picks += [k for k, sy in enumerate(b.systems) if sy == s][:ncapt_per_system]
    orfs = {s: O.OracleRF(s, decode_analog_audio=True) for s in ("NTSC", "PAL")}
    lut = T.build_kaiser_lut()
    S, N = 31712, 32768
    caps = []
    for k in picks:
        n = min(b.lengths[k], (nread_blocks - 1) * S + N)
        caps.append((b.systems[k], b.capture(k)[:n].cpu().numpy().astype(np.float64)))
    jobs = [(ci, i) for ci, (_, x) in enumerate(caps) for i in range((len(x) - N) // S + 1)]

    def demod(job):
        ci, i = job
        s, x = caps[ci]
        r = orfs[s].demodblock(data=x[i * S:i * S + N], cut=True)
        return r["video"]

    def fields(args):
        s, vids = args
        orf = orfs[s]
        d05 = np.concatenate([v["demod_05"] for v in vids]).astype(np.float32)
        dm = np.concatenate([v["demod"] for v in vids]).astype(np.float32)
        chan = "demod_burst" if s == "NTSC" else "demod_pilot"
        aux = np.concatenate([v[chan] for v in vids]).astype(np.float32)
        p = LL.LocParams.from_rf(orf)
        lp = L.LineRefineParams(orf)
        sp, dp = orf.SysParams, orf.DecoderParams
        W, lines = sp["outlinelen"], sp["frame_lines"] // 2 + 1
        state = {"phase": 0.0}
        outs = []

        def finish(k, res, org):
            bad = L.compute_deriv_error(res.linelocs1, res.linebad)
            ll2, bad = T.refine_linelocs_hsync(d05[org:], res.linelocs1, bad, s, 40.0, orf.iretohz, dp["vsync_ire"])
            bad = L.compute_deriv_error(ll2, bad)
            if s == "NTSC":
                ll4, _, _, ph = L.ntsc_field_lines(lp, dm[org:], aux[org:], ll2, bad, res.linecount,
                                                   res.isFirstField, state["phase"])
                state["phase"] = ph
                lo = 0
            else:
                ll4, _ = L.pal_field_lines(lp, aux[org:], ll2, bad)
                lo = 2 if res.isFirstField else 3
            locs, wow = T.computewow_scaled(ll4, orf.linelen, W, lines, lo)
            ds = ctbc.scale_field(dm[org:], locs, wow, lut, lo, W, lines * W)
            outs.append(T.hz_to_output_array(ds.astype(np.float32), dp["ire0"], dp["hz_ire"], sp["outputZero"],
                                             dp["vsync_ire"], out_scale(s, dp["vsync_ire"])))
            return ll4

        LL.walk_fields(p, d05, 64, 0, finish, demod=dm)
        return len(outs)

    passes = 0
    nfields = 0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        while True:
            vids = list(ex.map(demod, jobs))
            per = []
            j = 0
            for s, x in caps:
                nb = (len(x) - N) // S + 1
                per.append((s, vids[j:j + nb]))
                j += nb
            nfields = sum(ex.map(fields, per))
            passes += 1
            dt = time.perf_counter() - t0
            if dt >= min_seconds:
                break
    nsamp = sum(len(x) for _, x in caps)
    return nsamp * passes / dt / 1e6, dt, passes, nsamp, nfields, len(jobs)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


# ------------------------------------------------------------ profiling
def profile_step(args):
    """Profiling target: one C4 pass (graph replays), then one plain copy
    kernel of known size (torch.neg over COPY_ELEMS float32) that calibrates
    the HBM byte counters."""
    __graft_entry__.load_package()
    import torch
    from lddecode_amd.batch import run_interleaved
    rfs = decoders(torch, int(args.blocklen), args.fft)
    b = workload(torch, rfs, args.seed)
    decs = batch_decoders(args, rfs, b)
    for i in range(args.steps):
        for d in decs:
            d.reset()
        nf = run_interleaved(decs)
        torch.cuda.synchronize()
        log(f"profile pass {i}: {nf} fields")
    x = torch.ones(COPY_ELEMS, dtype=torch.float32, device=rfs["NTSC"].device)
    y = torch.empty_like(x)
    torch.neg(x, out=y)
    torch.cuda.synchronize()
    log("profile step done")


def trace_summary(db, top=40):
    """Markdown table of a rocprofv3 --kernel-trace database (its `kernels`
    view): dispatches grouped by kernel and launch shape, the `top` with the
    most total time."""
    import sqlite3
    con = sqlite3.connect(str(db))
    rows = con.execute(
        "select name, grid_x, grid_y, workgroup_x, vgpr_count, accum_vgpr_count, lds_size, count(*), "
        "avg(duration), min(duration), max(duration), sum(duration) from kernels "
        "group by name, grid_x, grid_y, workgroup_x order by sum(duration) desc limit ?", (top,)).fetchall()
    out = ["| kernel | grid (wg x y) | wg | VGPR | AGPR | LDS B | calls | avg us | min us | max us | total ms |",
           "|---|---|---|---|---|---|---|---|---|---|---|"]
    for name, gx, gy, wx, vg, ag, lds, n, avg, mn, mx, tot in rows:
        short = name.split("(")[0].replace("void ", "")[:60]
        out.append(f"| {short} | {gx // max(wx, 1)} x {gy} | {wx} | {vg} | {ag} | {lds} | {n} | {avg / 1e3:.2f} | "
                   f"{mn / 1e3:.2f} | {mx / 1e3:.2f} | {tot / 1e6:.3f} |")
    return "\n".join(out)


def pass_timeline(db, kernel, waves, steps, isolated=5, top=12):
    """Markdown summary of the timed C4 passes in a rocprofv3 --kernel-trace
    database of a default bench run: the headline demod kernel's last
    `steps * waves` launches before its `isolated` lone ones bound the window;
    reported per pass: the window, the time any kernel runs (union of
    dispatch intervals), each kernel's summed duration (kernels overlap across
    the two streams, so these add up to more than the window), and the gaps
    between consecutive demod launches with what starts inside them."""
    import sqlite3
    con = sqlite3.connect(str(db))
    ks = sorted(con.execute("select start, end, name from kernels").fetchall())
    dem = [k for k in ks if k[2].split("(")[0].replace("void ", "").endswith(kernel)]
    need = steps * waves + isolated
    if len(dem) < need:
        return f"(only {len(dem)} {kernel} launches, {need} needed for {steps} passes of {waves} waves)"
    timed = dem[-need:-isolated]
    t0, t_next = timed[0][0], dem[-isolated][0]
    win = [k for k in ks if t0 <= k[0] < t_next]
    t1 = max(k[1] for k in win)
    busy, cs, ce = 0, None, None
    for s, e, _ in win:
        if ce is None or s > ce:
            if ce is not None:
                busy += ce - cs
            cs, ce = s, e
        else:
            ce = max(ce, e)
    busy += ce - cs
    tot = {}
    for s, e, n in win:
        key = n.split("(")[0].replace("void ", "")[:60]
        v = tot.setdefault(key, [0, 0])
        v[0] += e - s
        v[1] += 1
    gaps, ing = [], {}
    for (s0, e0, _), (s1, _, _) in zip(timed, timed[1:]):
        gaps.append(max(s1 - e0, 0))
        for s, e, n in win:
            if e0 <= s < s1:
                key = n.split("(")[0].replace("void ", "")[:60]
                ing[key] = ing.get(key, 0) + min(e, s1) - s
    ps = 1e6 * steps
    out = [f"Timed passes: {steps} x {waves} waves; per pass: window {(t1 - t0) / ps:.1f} ms, some kernel running "
           f"{busy / ps:.1f} ms; {kernel} launches {sum(e - s for s, e, _ in timed) / ps:.1f} ms; gaps between "
           f"consecutive demod launches {sum(gaps) / ps:.1f} ms (median {sorted(gaps)[len(gaps) // 2] / 1e3:.0f} us).",
           "", "| kernel | ms per pass (summed durations) | calls per pass | avg us | in the demod gaps, ms per pass |",
           "|---|---|---|---|---|"]
    for key, (d, c) in sorted(tot.items(), key=lambda kv: -kv[1][0])[:top]:
        out.append(f"| {key} | {d / ps:.2f} | {c / steps:.0f} | {d / c / 1e3:.1f} | {ing.get(key, 0) / ps:.2f} |")
    return "\n".join(out)


def _pmc_run(counters, target, workdir, tag):
    """One rocprofv3 counter pass (the counters collected together) over
    `bench.py <target>` in a child process; returns {kernel name: {counter:
    [(sum, max) over the counter's instances, per dispatch]}}."""
    import sqlite3
    import subprocess
    out = pathlib.Path(workdir) / tag
    cmd = ["rocprofv3", "--pmc", *counters, "-d", str(out), "-o", "run", "--", sys.executable, "-u",
           str(pathlib.Path(__file__).resolve()), *target]
    log(" ".join(cmd))
    env = dict(os.environ, TMPDIR="/tmp")
    subprocess.run(["timeout", "-k", "10", "400"] + cmd, check=True, cwd="/tmp", env=env)
    dbs = sorted(out.rglob("*_results.db"))
    if not dbs:
        raise RuntimeError(f"rocprofv3 wrote no results under {out}")
    con = sqlite3.connect(str(dbs[0]))
    got = {}
    for counter in counters:
        rows = con.execute("select dispatch_id, kernel_name, sum(value), max(value) from counters_collection "
                           "where counter_name = ? group by dispatch_id order by dispatch_id", (counter,)).fetchall()
        for _, name, vsum, vmax in rows:
            short = name.split("(")[0].replace("void ", "")
            if "neg_kernel" in name:
                short = "copy"
            got.setdefault(short, {}).setdefault(counter, []).append((float(vsum), float(vmax)))
    return got


def _pmc_pass(counter, args, workdir):
    """One counter over `bench.py --profile-step` (the C4 pass);
    returns {kernel name: [values per dispatch]}."""
    target = ["--profile-step", "--steps", "1", "--blocklen", str(args.blocklen), "--fft", args.fft,
              "--wave-blocks", str(args.wave_blocks), "--groups", str(args.groups), "--seed", str(args.seed)]
    got = _pmc_run([counter], target, workdir, counter.lower())
    return {k: [v[0] for v in c[counter]] for k, c in got.items()}


FFT_KERNELS = ("k_fft_fp32", "k_fft_mfma", "k_fft_mfma3", "k_demod_block", "k_demod_block_mfma",
               "k_demod_block_mfma3")
MFMA_COUNTERS = ("SQ_VALU_MFMA_BUSY_CYCLES", "GRBM_GUI_ACTIVE", "SQ_INSTS_VALU_MFMA_MOPS_BF16", "SQ_INSTS_MFMA",
                 "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_BUSY_CYCLES")
FFT_TRANSFORMS, DEMOD_BLOCKS = 2048, 512
CLOCK_HZ = 2.4e9               # MI355X peak engine clock


def profile_fft(args):
    """Profiling target for the matrix-core counters: the demodulator's
    16384-point transform alone (ld_fft_forward over FFT_TRANSFORMS
    transforms), then DEMOD_BLOCKS 32768-sample NTSC blocks through
    ld_demod_blocks (every video channel, rfhpf, audio), three times each
    in each transform variant."""
    __graft_entry__.load_package()
    import torch
    from lddecode_amd import _lib, synth, tune
    from lddecode_amd.rfdecode import RFDecode
    rf = RFDecode(system="NTSC", decode_analog_audio=44100, device=0)
    M = 16384
    g = torch.Generator(device=rf.device).manual_seed(args.seed)
    x = torch.randn(FFT_TRANSFORMS * M, dtype=torch.complex64, device=rf.device, generator=g)
    y = torch.empty_like(x)
    for name, v in tune.VARIANTS.items():
        for _ in range(3):
            _lib.check(rf.lib.ld_fft_forward(rf.handle, x.data_ptr(), y.data_ptr(), FFT_TRANSFORMS, v, None))
        torch.cuda.synchronize()
        log(f"fft {name} done")
    cap = torch.from_numpy(synth.make_capture("NTSC", 2_000_000, seed=args.seed, snr_db=40.0).data).to(rf.device)
    need = (DEMOD_BLOCKS - 1) * rf.block_stride + rf.blocklen
    data = cap.repeat(-(-need // cap.numel()))[:need].contiguous()
    for name, v in tune.VARIANTS.items():
        rf.set_fft_variant(v)
        for _ in range(3):
            rf.demod_blocks(data, 0, DEMOD_BLOCKS, audio=True)
        torch.cuda.synchronize()
        log(f"demod {name} done")


def measure_mfma(args):
    """Matrix-core use of the transform variants: one counter pass
    (MFMA_COUNTERS) and one kernel-trace pass over `bench.py --profile-fft`;
    per kernel the MFMA busy share of the SIMDs (SQ_VALU_MFMA_BUSY_CYCLES
    over 1024 SIMDs x the kernel's traced duration x the 2.4 GHz engine
    clock; also over GRBM_GUI_ACTIVE, rocprofv3's MfmaUtil, whose window is
    the counter pass's longer dispatch), MFMA FLOPs (MOPS_BF16 x 512), the
    instruction mix and the average duration; stamped with the kernel
    sources' hash."""
    import csv
    import subprocess
    import tempfile
    work = tempfile.mkdtemp(prefix="ld_mfma_", dir="/tmp")
    pmc = _pmc_run(list(MFMA_COUNTERS), ["--profile-fft", "--seed", str(args.seed)], work, "pmc")
    out = pathlib.Path(work) / "trace"
    cmd = ["rocprofv3", "--kernel-trace", "--stats", "--output-format", "csv", "-d", str(out), "-o", "run", "--",
           sys.executable, "-u", str(pathlib.Path(__file__).resolve()), "--profile-fft", "--seed", str(args.seed)]
    log(" ".join(cmd))
    subprocess.run(["timeout", "-k", "10", "400"] + cmd, check=True, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"))
    stats = sorted(out.rglob("*kernel_stats.csv"))
    if not stats:
        raise RuntimeError(f"rocprofv3 wrote no kernel statistics under {out}")
    avg_ns = {}
    with open(stats[0]) as f:
        for row in csv.DictReader(f):
            avg_ns[row["Name"].split("(")[0].replace("void ", "")] = float(row["AverageNs"])
    simds = 256 * 4
    kernels = {}
    for k in FFT_KERNELS:
        c = pmc.get("ldgpu::" + k)
        if c is None:
            continue
        busy = np.mean([v[0] for v in c["SQ_VALU_MFMA_BUSY_CYCLES"]])
        gui = np.mean([v[1] for v in c["GRBM_GUI_ACTIVE"]])
        mops = np.mean([v[0] for v in c["SQ_INSTS_VALU_MFMA_MOPS_BF16"]])
        ns = avg_ns.get("ldgpu::" + k)
        rec = {"dispatches": len(c["GRBM_GUI_ACTIVE"]), "avg_ns": ns, "mfma_busy_cycles": busy,
               "mfma_busy_pct_of_window": round(100.0 * busy / (gui * simds), 3),
               "mfma_flops": mops * 512, "insts_mfma": np.mean([v[0] for v in c["SQ_INSTS_MFMA"]]),
               "insts_valu": np.mean([v[0] for v in c["SQ_INSTS_VALU"]]),
               "insts_lds": np.mean([v[0] for v in c["SQ_INSTS_LDS"]]),
               "gui_active_cycles": gui}
        if ns:
            # busy SIMD cycles over the SIMD cycles of the kernel's own
            # duration (GRBM_GUI_ACTIVE spans the counter pass's longer
            # dispatch window)
            rec["mfma_busy_pct"] = round(100.0 * busy / (ns * 1e-9 * CLOCK_HZ * simds), 3)
            rec["mfma_tflops"] = round(mops * 512 / ns / 1e3, 2)
            rec["mfma_frac_of_bf16_peak"] = round(mops * 512 / ns / 1e3 / 2500.0, 5)
        kernels[k] = rec
    res = {"source_hash": kernel_source_hash(), "seed": args.seed, "fft_transforms": FFT_TRANSFORMS,
           "demod_blocks": DEMOD_BLOCKS, "simds": simds, "counters": list(MFMA_COUNTERS),
           "command": "python bench.py --measure-mfma", "kernels": kernels}
    dest = pathlib.Path(args.traffic_out) if args.traffic_out else ROOT / "profiles" / "mfma_util.json"
    dest.parent.mkdir(parents=True, exist_ok=True)
    dest.write_text(json.dumps(res, indent=1) + "\n")
    print(json.dumps(res, indent=1))


def measure_traffic(args):
    """HBM traffic per launch of the C4 pass's kernels: FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) from two counter passes, each scaled by the
    ratio of known bytes to the counter's reading for the calibration copy;
    stamped with the demod kernel's source hash."""
    import tempfile
    work = tempfile.mkdtemp(prefix="ld_traffic_", dir="/tmp")
    fetch = _pmc_pass("FETCH_SIZE", args, work)
    write = _pmc_pass("WRITE_SIZE", args, work)
    copy_kib = COPY_ELEMS * 4 / 1024
    fc, wc = max(fetch["copy"]), max(write["copy"])
    fscale, wscale = copy_kib / fc, copy_kib / wc
    kernels = {}
    for k in sorted(set(fetch) & set(write)):
        if not k.startswith("ldgpu::"):
            continue
        f = float(np.mean(fetch[k])) * fscale * 1024
        w = float(np.mean(write[k])) * wscale * 1024
        kernels[k] = {"dispatches": len(fetch[k]), "fetch_bytes": round(f), "write_bytes": round(w),
                      "traffic_bytes": round(f + w)}
    res = {"source_hash": kernel_source_hash(), "seed": args.seed, "wave_blocks": args.wave_blocks,
           "groups": args.groups, "blocklen": int(args.blocklen), "fft": args.fft,
           "copy_bytes_each_way": COPY_ELEMS * 4, "copy_fetch_kib": fc, "copy_write_kib": wc,
           "fetch_scale": fscale, "write_scale": wscale,
           "command": f"python bench.py --measure-traffic --blocklen {args.blocklen} --fft {args.fft} "
                      f"--wave-blocks {args.wave_blocks} --groups {args.groups}", "kernels": kernels}
    dest = pathlib.Path(args.traffic_out) if args.traffic_out else TRAFFIC_FILE
    dest.parent.mkdir(parents=True, exist_ok=True)
    dest.write_text(json.dumps(res, indent=1) + "\n")
    print(json.dumps(res, indent=1))


def load_traffic(kernel, wave_blocks, seed, groups, blocklen, fft):
    """Measured HBM bytes per launch of `kernel` in the C4 pass, or None when
    the file is missing or was measured on other kernel sources / workload."""
    try:
        res = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None, "no traffic file"
    if res.get("source_hash") != kernel_source_hash():
        return None, "traffic file measured on other kernel sources (stale): not used"
    if (res.get("wave_blocks") != wave_blocks or res.get("seed") != seed or res.get("groups") != groups
            or res.get("blocklen") != blocklen or res.get("fft") != fft):
        return None, "traffic file measured on another workload or configuration: not used"
    k = res.get("kernels", {}).get(kernel)
    return (None, "kernel not in traffic file") if k is None else (k["traffic_bytes"], "measured")


# ------------------------------------------------------------ main
def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed passes over the whole batch")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--wave-blocks", type=int, default=16384,
                    help="32768-sample blocks per wave of the batched decode (scaled for other block lengths)")
    ap.add_argument("--groups", type=int, default=2, help="batch decoders taking turns on the host")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="minimum CPU baseline time (0: skip)")
    ap.add_argument("--blocklen", default="auto", choices=["auto", "8192", "16384", "32768", "65536"],
                    help="block length of the batched decode (auto: the autotuner's choice)")
    ap.add_argument("--fft", default="auto", choices=["auto", "fp32", "mfma", "mfma3"],
                    help="transform variant of the batched decode (auto: the autotuner's choice)")
    ap.add_argument("--legs", default="default,c5,c5x6",
                    help="extra batched passes beside the headline: default (32768, fp32), c5 (65536, mfma: "
                         "hi + lo terms), c5x6 (65536, mfma3: hi + mid + lo terms)")
    ap.add_argument("--profile-step", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--profile-fft", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--measure-mfma", action="store_true",
                    help="run the matrix-core counter and trace passes over the transform variants")
    ap.add_argument("--measure-traffic", action="store_true",
                    help="run the two rocprofv3 counter passes and write " + str(TRAFFIC_FILE.name))
    ap.add_argument("--traffic-out", default=None, help="where --measure-traffic / --measure-mfma write their JSON")
    ap.add_argument("--trace-summary", default=None, metavar="DB",
                    help="print the kernel table of a rocprofv3 --kernel-trace database and exit")
    ap.add_argument("--trace-waves", type=int, default=0,
                    help="with --trace-summary: also summarise the timed C4 passes (waves per pass, --steps passes)")
    ap.add_argument("--trace-kernel", default="k_demod_batch8",
                    help="with --trace-waves: the headline demod kernel whose launches bound the passes")
    args = ap.parse_args()
    if args.trace_summary:
        print(trace_summary(args.trace_summary))
        if args.trace_waves > 0:
            print()
            print(pass_timeline(args.trace_summary, args.trace_kernel, args.trace_waves, args.steps))
        return
    if (args.profile_step or args.measure_traffic) and "auto" in (args.blocklen, args.fft):
        ap.error("--profile-step / --measure-traffic take a fixed --blocklen and --fft")
    if args.profile_fft:
        return profile_fft(args)
    if args.measure_mfma:
        return measure_mfma(args)
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
    from lddecode_amd import capture, tune
    # the one-capture pipeline's stream pair first, as an application that
    # decodes captures one at a time makes it: streams made after the batch
    # decoders' can share a hardware queue and then run one after the other
    capture._pipeline_streams(torch, 0)
    tuned = None
    if "auto" in (args.blocklen, args.fft):
        lengths = tune.LENGTHS if args.blocklen == "auto" else (int(args.blocklen),)
        variants = tuple(tune.VARIANTS) if args.fft == "auto" else (args.fft,)
        tuned = tune.autotune(torch, 0, wave_samples=args.wave_blocks // args.groups * 31712, lengths=lengths,
                              variants=variants, log=log)
        log(f"autotuner: blocklen {tuned['blocklen']} {tuned['fft']} ({tuned['candidates']})")
        blocklen, fft = tuned["blocklen"], tuned["fft"]
    else:
        blocklen, fft = int(args.blocklen), args.fft
    rfs = decoders(torch)                       # 32768 / fp32: the plain paths and the default leg
    b = workload(torch, rfs, args.seed)

    head_rfs = rfs if (blocklen, fft) == (32768, "fp32") else decoders(torch, blocklen, fft)
    c4 = run_c4(args, torch, head_rfs, b)
    value = c4["samples"] / (c4["ms_per_step"] / 1e3) / 1e6
    log(f"C4 ({blocklen}, {fft}): {c4['ms_per_step']:.1f} ms per pass, {value:.0f} MS/s, {c4['fields']} fields, "
        f"{c4['waves']} waves")
    legs = {}
    for leg in [x for x in args.legs.split(",") if x]:
        cfg = {"default": (32768, "fp32"), "c5": (65536, "mfma"), "c5x6": (65536, "mfma3")}[leg]
        if cfg == (blocklen, fft):
            legs[leg] = "the headline configuration"
            continue
        lrfs = rfs if cfg == (32768, "fp32") else decoders(torch, *cfg)
        r = run_c4(args, torch, lrfs, b)
        v = r["samples"] / (r["ms_per_step"] / 1e3) / 1e6
        log(f"C4 leg {leg} {cfg}: {r['ms_per_step']:.1f} ms per pass, {v:.0f} MS/s")
        legs[leg] = {"blocklen": cfg[0], "fft": cfg[1], "value": round(v, 2), "unit": "Msamples/s",
                     "ms_per_step": round(r["ms_per_step"], 3), "fields_per_step": r["fields"],
                     "demod_launch_gbs": round(r["demod_bytes_avg"] / (r["demod_ms_avg"] / 1e3) / 1e9, 1),
                     "demod_launch_ms_avg": round(r["demod_ms_avg"], 4),
                     "isolated_demod_gbs": round(r["iso_bytes_avg"] / (r["iso_ms_avg"] / 1e3) / 1e9, 1)}
        del lrfs

    # plain paths on slices of the batch's first NTSC and PAL captures
    k_ntsc = next(k for k, s in enumerate(b.systems) if s == "NTSC")
    k_pal = next(k for k, s in enumerate(b.systems) if s == "PAL")
    c1 = {s: run_block(torch, rfs[s], b.capture(k)[:32768], 200) for s, k in (("NTSC", k_ntsc), ("PAL", k_pal))}
    reads = {s: b.capture(k)[:int(2.6 * (667_334 if s == "NTSC" else 800_000))] for s, k in
             (("NTSC", k_ntsc), ("PAL", k_pal))}
    fld = {s: run_field(torch, rfs[s], reads[s], 20) for s in ("NTSC", "PAL")}
    longest = max((k for k, s in enumerate(b.systems) if s == "NTSC"), key=lambda k: b.lengths[k])
    span = b.capture(longest)[:66 * 667_334 + 100_000].contiguous()
    pipe = run_pipelined(torch, rfs["NTSC"], span, 64, 10, 2)
    log(f"one capture, 64 NTSC fields pipelined: {pipe['ms']:.3f} ms per pass")
    efm_off, efm_on, efm_nb = efm_demod_ms(torch, span, 10)
    packed = None
    try:
        from lddecode_amd import synth
        pk = torch.from_numpy(synth.pack_10bit(span.cpu().numpy())).to(rfs["NTSC"].device)
        pk_pipe = run_pipelined(torch, rfs["NTSC"], pk, 64, 10, 2)
        packed = {"value": round(pk_pipe["msps"], 1), "unit": "Msamples/s", "ms_per_step": round(pk_pipe["ms"], 3),
                  "input": "the one-capture 64-field pipelined decode reading the packed 10-bit format "
                           "(1.25 B/sample) directly"}
    except Exception as e:  # noqa: BLE001 - reported, not fatal
        log(f"packed run failed: {e}")

    achieved = c4["demod_bytes_avg"] / (c4["demod_ms_avg"] / 1e3) / 1e9
    iso_gbs = c4["iso_bytes_avg"] / (c4["iso_ms_avg"] / 1e3) / 1e9
    kname = {8192: "ldgpu::k_demod_batch8", 16384: "ldgpu::k_demod_batch16", 32768: "ldgpu::k_demod_batch",
             65536: "ldgpu::k_demod_batch64"}[blocklen]
    kname += {"fp32": "", "mfma": "_mfma", "mfma3": "_mfma3"}[fft]
    traffic, traffic_note = load_traffic(kname, args.wave_blocks, args.seed, args.groups, blocklen, fft)

    cpu = None
    if args.cpu_seconds > 0:
        workers = os.cpu_count() or 1
        rate, secs, passes, nsamp, nf, nblk = cpu_baseline(b, 2, 60, workers, args.cpu_seconds)
        rate1, secs1, _, _, _, _ = cpu_baseline(b, 2, 60, 1, 0.0)
        import scipy
        cpu = {"value": round(rate, 3), "unit": "Msamples/s", "cores": workers, "kind": "port",
               "value_1_worker": round(rate1, 3),
               "cpu_model": cpu_model(), "numpy": np.__version__, "scipy": scipy.__version__,
               "sample": f"FP64 numpy/scipy oracle, the first 60 blocks ({nsamp // 4} samples) of the batch's first "
                         f"2 NTSC and 2 PAL captures: blocks demodulated on a {workers}-thread pool, then per "
                         f"capture the field walk, hsync refine, burst / pilot alignment and TBC to uint16 "
                         f"({nf} fields); {passes} passes in {secs:.1f} s; 1 worker: one pass in {secs1:.1f} s"}
        log(f"CPU oracle: {rate:.2f} MS/s on {workers} threads, {rate1:.2f} MS/s on 1")

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(c4["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"C4 batched: 64 captures of 1-30 s (log-uniform, seed {args.seed}), NTSC and PAL mixed "
                        f"({sum(s == 'PAL' for s in b.systems)} PAL), int16 at 40 MSPS, {c4['samples']} samples "
                        "resident in HBM; one step = one pass over the whole batch: per wave one mixed-system "
                        f"demod launch ({blocklen}-sample blocks, {fft} transform, both systems' filter tables "
                        "resident, HIP graph with the sync-pulse search), host field walk per capture, FP64 hsync "
                        "+ burst/pilot line positions, sinc-LUT TBC to uint16, per system",
            "captures": len(b.lengths),
            "samples_per_step": c4["samples"],
            "blocklen": blocklen,
            "fft": fft,
            "autotuner": tuned,
            "wave_blocks": args.wave_blocks,
            "decoders": args.groups,
        },
        "x_realtime": round(value / 40.0, 1),
        "batch": {"fields_per_step": c4["fields"], "waves": c4["waves"], "blocks_demodulated": c4["blocks"],
                  "blocks_by_system": c4["blocks_by_system"],
                  "overlap_recompute": round(c4["blocks"] * c4["block_stride"] / c4["samples"] - 1, 4),
                  "relocated_rounds": c4["relocated"], "redone_waves": c4["redone_waves"],
                  "demod_launches": c4["demod_launches"],
                  "demod_ms_per_pass": round(c4["demod_ms_total"], 2)},
        "plain": {
            "c1_block_ntsc": {"value": round(c1["NTSC"]["msps"], 2), "unit": "Msamples/s",
                              "ms_per_block": round(c1["NTSC"]["ms"], 4), "gpu_ms": round(c1["NTSC"]["gpu_ms"], 4)},
            "c1_block_pal": {"value": round(c1["PAL"]["msps"], 2), "unit": "Msamples/s",
                             "ms_per_block": round(c1["PAL"]["ms"], 4), "gpu_ms": round(c1["PAL"]["gpu_ms"], 4)},
            "c2_field_ntsc": {"value": round(fld["NTSC"]["msps"], 2), "unit": "Msamples/s",
                              "ms_per_field": round(fld["NTSC"]["ms"], 3), "blocks": fld["NTSC"]["blocks"]},
            "c3_field_pal": {"value": round(fld["PAL"]["msps"], 2), "unit": "Msamples/s",
                             "ms_per_field": round(fld["PAL"]["ms"], 3), "blocks": fld["PAL"]["blocks"]},
            "note": "one block per launch (C1, demod only: the RFDecode.demodblock call shape) and one field read "
                    "per step (C2 / C3: demod, location, hsync, burst/pilot, TBC and audio), sequential",
        },
        "batched_legs": legs,
        "one_capture_pipelined": {"value": round(pipe["msps"], 2), "unit": "Msamples/s",
                                  "ms_per_step": round(pipe["ms"], 3), "fields_per_step": pipe["fields"],
                                  "note": "64 NTSC fields of one capture per pass with audio stage 2 and 44.1 kHz "
                                          "output, consecutive passes overlapped on two streams"},
        "packed_input": packed,
        "efm": {"demod_ms": round(efm_off, 4), "demod_ms_with_efm": round(efm_on, 4), "blocks": efm_nb},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_note": traffic_note,
                     "algorithmic_bytes": round(c4["demod_bytes_avg"]),
                     "kernel_ms": round(c4["demod_ms_avg"], 4),
                     "kernel": f"{kname.split('::')[1]}: events around every wave's demod graph replay (block table upload "
                               "+ the launch) in the timed passes, beside the stage B kernels; algorithmic "
                               "22.25 B per usable NTSC sample, 26.25 PAL (input, video channels, rfhpf, audio "
                               "stage 1)",
                     "isolated": {"achieved": round(iso_gbs, 1), "frac": round(iso_gbs / HBM_PEAK_GBS, 4),
                                  "kernel_ms": round(c4["iso_ms_avg"], 4), "launches": c4["iso_launches"],
                                  "note": "the first waves' launches, each alone on the GPU"}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
