"""Autotuner for the batched demodulation: block length and transform variant.

The reference takes its block length as a constructor argument
(RFDecode(blocklen=...), core.py:284-288; 32768 by default) and has one
transform.  This build offers four block lengths (8192, 16384, 32768, 65536) and
three transform variants (FP32 Stockham passes of radix 16 / 16 / 16 / 2-4,
or radix-64 stages as split-BF16 matrix products of two or three bf16
terms, LD_FFT_MFMA_BF16X3 / LD_FFT_MFMA_BF16X6), each checked against the
FP64 restatement at the same block length.  Their
speeds differ: shorter blocks have smaller transforms and no register
spills but more overlap (1056 samples of every block are not output),
longer blocks the reverse.

``autotune`` times the demod launch of a batched wave (ld_demod_batch over a
mixed NTSC / PAL batch, as BatchDecoder launches it) for every candidate and
returns the fastest by usable RF samples per second.  The choice is kept in
a JSON file keyed by the GPU architecture, the library's build and the wave
size, so later runs on the same device reuse it:

    LDGPU_TUNE_FILE   the file (default ~/.cache/lddecode_amd/tune.json)
"""

import hashlib
import json
import os
import pathlib
import time

from . import _lib

LENGTHS = (8192, 16384, 32768, 65536)
VARIANTS = {"fp32": _lib.LD_FFT_FP32, "mfma": _lib.LD_FFT_MFMA_BF16X3, "mfma3": _lib.LD_FFT_MFMA_BF16X6}
DEFAULT = {"blocklen": 32768, "fft": "fp32"}


def tune_file():
    p = os.environ.get("LDGPU_TUNE_FILE")
    return pathlib.Path(p) if p else pathlib.Path.home() / ".cache" / "lddecode_amd" / "tune.json"


def _lib_id():
    """Hash of the loaded libldgpu.so (a rebuilt library is tuned again)."""
    lib = pathlib.Path(_lib.load()._name)
    return hashlib.sha256(lib.read_bytes()).hexdigest()[:16]


def _key(torch, device, wave_samples):
    arch = torch.cuda.get_device_properties(device).gcnArchName
    return f"{arch}|{_lib_id()}|{int(wave_samples)}"


def load(torch, device, wave_samples):
    """The stored choice for this device, library and wave size, or None."""
    f = tune_file()
    try:
        db = json.loads(f.read_text())
    except (OSError, ValueError):
        return None
    return db.get(_key(torch, device, wave_samples))


def store(torch, device, wave_samples, result):
    f = tune_file()
    try:
        db = json.loads(f.read_text())
    except (OSError, ValueError):
        db = {}
    db[_key(torch, device, wave_samples)] = result
    f.parent.mkdir(parents=True, exist_ok=True)
    tmp = f.with_suffix(".tmp")
    tmp.write_text(json.dumps(db, indent=1, sort_keys=True))
    os.replace(tmp, f)


def _tuning_batch(torch, device, wave_samples, seed):
    """Four NTSC and four PAL captures, together twice one wave's samples."""
    from . import synth
    per = int(2 * wave_samples / 8) + 700_000
    plan = [("NTSC" if k % 2 == 0 else "PAL", per) for k in range(8)]
    return synth.make_batch(plan, device, seed=seed, ntemplates=1)


def measure(torch, batch, blocklen, fft, wave_samples, reps=3, audio=True):
    """Usable RF samples per second of one full wave's demod launch (the
    block table upload and ld_demod_batch, replayed as a HIP graph), alone
    on the GPU, median of `reps` runs."""
    from .batch import BatchDecoder
    from .rfdecode import RFDecode
    rfs = {s: RFDecode(system=s, blocklen=blocklen, decode_analog_audio=44100 if audio else 0,
                       device=batch.data.device) for s in ("NTSC", "PAL")}
    for rf in rfs.values():
        rf.set_fft_variant(VARIANTS[fft])
    S = rfs["NTSC"].block_stride
    dec = BatchDecoder(rfs, batch.data, batch.systems, batch.offsets, batch.lengths,
                       wave_blocks=max(1, int(wave_samples // S)), graph=True, audio=audio)
    rates = []
    for _ in range(reps + 1):                 # the first run captures the graphs
        dec.reset()
        torch.cuda.synchronize()
        dec.demod_timing = []
        it = dec.waves_iter()
        next(it)
        torch.cuda.synchronize()
        e0, e1, _ = dec.demod_timing[0]
        samples = sum(dec.sets[0]["used"].values()) * S
        rates.append(samples / (e0.elapsed_time(e1) / 1e3))
        it.close()
    dec.demod_timing = None
    del dec, rfs
    torch.cuda.synchronize()
    rates = sorted(rates[1:])
    return rates[len(rates) // 2]


def autotune(torch, device=0, wave_samples=4096 * 31712, lengths=LENGTHS, variants=tuple(VARIANTS), seed=5,
             force=False, log=None):
    """The fastest (blocklen, fft variant) for batched waves of about
    `wave_samples` samples on `device`: {"blocklen", "fft", "msps",
    "candidates": {"<blocklen>/<fft>": Msamples/s}}.  Reuses the stored
    choice unless `force`."""
    dev = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
    if not force:
        got = load(torch, dev, wave_samples)
        if got is not None:
            return dict(got, reused=True)
    t0 = time.perf_counter()
    batch = _tuning_batch(torch, dev, wave_samples, seed)
    cands = {}
    for n in lengths:
        for v in variants:
            r = measure(torch, batch, n, v, wave_samples)
            cands[f"{n}/{v}"] = round(r / 1e6, 1)
            if log:
                log(f"tune: blocklen {n} {v}: {r / 1e6:.0f} Msamples/s")
    del batch
    torch.cuda.empty_cache()
    best = max(cands, key=cands.get)
    n, v = best.split("/")
    result = {"blocklen": int(n), "fft": v, "msps": cands[best], "candidates": cands,
              "wave_samples": int(wave_samples), "seconds": round(time.perf_counter() - t0, 1)}
    store(torch, dev, wave_samples, result)
    return dict(result, reused=False)


def tuned_decoders(torch, device=0, wave_samples=4096 * 31712, systems=("NTSC", "PAL"), audio=True, **kw):
    """RFDecode per system at the tuned block length and transform variant,
    and the tuning result."""
    from .rfdecode import RFDecode
    res = autotune(torch, device, wave_samples, **kw)
    rfs = {s: RFDecode(system=s, blocklen=res["blocklen"], decode_analog_audio=44100 if audio else 0, device=device)
           for s in systems}
    for rf in rfs.values():
        rf.set_fft_variant(VARIANTS[res["fft"]])
    return rfs, res
