"""The no-allocation demod entry point (ld_demod_blocks_noalloc) against
ld_demod_blocks: same kernels, workspace from the library's static scratch
array instead of a caller buffer, so the outputs must be bit-identical; two
calls in a row, the second smaller, and a request over the array's size."""

import ctypes

import numpy as np
import pytest

from lddecode_amd import _lib, synth

pytestmark = pytest.mark.gpu


def _outputs(rf, torch, nblocks, audio):
    S = rf.block_stride
    names = list(rf.video_fields) + ["rfhpf"]
    bufs = {n: torch.full((nblocks * S,), float("nan"), dtype=torch.float32, device=rf.device) for n in names}
    o = _lib.DemodOutputs()
    for n, t in bufs.items():
        setattr(o, n, t.data_ptr())
    o.video_stride = S
    if audio:
        na = rf.filterbank.audio["left"]["nbins"]
        fdiv = rf.blocklen // na
        alen = na - rf.blockcut // fdiv - (-(-32 // fdiv))
        for n in ("audio_left", "audio_right"):
            bufs[n] = torch.full((nblocks * alen,), float("nan"), dtype=torch.float32, device=rf.device)
            setattr(o, n, bufs[n].data_ptr())
        o.audio_stride = alen
    return bufs, o


def test_noalloc_demod_equals_demod_blocks(gpu_rf):
    import torch
    rf = gpu_rf("NTSC", audio=True)
    lib = rf.lib
    cap = synth.make_capture("NTSC", 31712 * 48 + 2000, seed=21, snr_db=40.0)
    x = torch.from_numpy(cap.data).to(rf.device)
    st = torch.cuda.current_stream(rf.device).cuda_stream
    assert lib.ld_demod_noalloc_max_bytes() == 32 << 20
    for first, nblocks in ((0, 40), (3 * 31712 + 517, 7)):
        want = rf.demod_blocks(x, first, nblocks, cut=True, audio=True)
        ref = dict(want["video"], rfhpf=want["rfhpf"], **want["audio"])
        bufs, o = _outputs(rf, torch, nblocks, True)
        _lib.check(lib.ld_demod_blocks_noalloc(rf.handle, x.data_ptr(), _lib.LD_IN_INT16, x.numel(), first,
                                               rf.block_stride, nblocks, 1, ctypes.byref(o), ctypes.c_void_p(st)))
        torch.cuda.synchronize()
        print(f"noalloc call: first {first}, {nblocks} blocks", flush=True)
        for n, t in bufs.items():
            np.testing.assert_array_equal(t.cpu().numpy(), ref[n].cpu().numpy(), err_msg=n)


def test_noalloc_rejects_more_than_the_scratch(gpu_rf):
    import torch
    rf = gpu_rf("NTSC", audio=False)
    lib = rf.lib
    per_block = lib.ld_demod_workspace_bytes(rf.handle, 1)
    nmax = (32 << 20) // per_block
    x = torch.zeros(rf.block_stride * (nmax + 2) + rf.blocklen, dtype=torch.int16, device=rf.device)
    bufs, o = _outputs(rf, torch, nmax + 1, False)
    rc = lib.ld_demod_blocks_noalloc(rf.handle, x.data_ptr(), _lib.LD_IN_INT16, x.numel(), 0, rf.block_stride,
                                     nmax + 1, 1, ctypes.byref(o), None)
    assert rc == _lib.LD_EINVAL
    assert b"scratch" in lib.ld_last_error()
