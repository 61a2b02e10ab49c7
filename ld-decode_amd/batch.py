"""Many captures of both systems decoded together: the batched hot path.

The reference decodes one capture per ``LDdecode`` instance (core.py:3498-3700),
field by field (``readfield``, core.py:4018-4199), each with its own
``RFDecode`` and ``DemodCache`` worker threads (core.py:1049-1421).  Here a
batch of captures of uneven length, NTSC and PAL mixed, advances in waves:

  stage A (stream A, optionally one replayed HIP graph per buffer set):
      the block table of the wave (pinned host -> device), ONE
      ``ld_demod_batch`` launch demodulating every lane's window of blocks
      with its system's resident filter tables, and per system the sync-pulse
      search over the demod's in-pulse bits, the pulses' levels and their
      copies to pinned host memory;
  host: every lane's field walk (``LDdecode.readfield``'s location of the
      next fields, native code) on its own pulses;
  stage B (stream B, per system): FP64 hsync refine, burst (NTSC, one chain
      per capture) / pilot (PAL) alignment, TBC to uint16 and the per-field
      facts that check the walk's provisional values.

Each capture ("lane") carries from wave to wave what an ``LDdecode`` carries
from field to field: the next read position, the previous field (its final
line position and skip-check score, ``prevfield``) and the NTSC burst
phase.  Two buffer sets let wave k+1's demodulation run while wave k is
aligned and resampled.  A lane's window overlaps the previous one by the
blocks of its next field read, which are demodulated again.
"""

import ctypes
import math
import os

import numpy as np

from . import _lib
from . import field as fieldmod
from . import locate as locatemod

SYSTEMS = ("NTSC", "PAL")

# algorithmic bytes per usable RF sample of the demodulation: the int16
# input, the video channels (4 NTSC, 5 PAL) and rfhpf as float32, the two
# audio stage-1 channels at 1/32 of the rate
DEMOD_BYTES = {"NTSC": 2 + 4 * 4 + 4 + 2 * 4 / 32, "PAL": 2 + 5 * 4 + 4 + 2 * 4 / 32}


class Lane:
    """Decode state of one capture between waves."""

    def __init__(self, idx, system, offset, length):
        self.idx = idx
        self.system = system
        self.offset = int(offset)      # first sample in the batch tensor
        self.length = int(length)      # samples
        self.start = 0.0               # capture-relative raw sample of the next field read
        self.prev = None               # PrevField the next field reads
        self.prov_delta = 0.0
        self.phase_seed = 0.0          # NTSC: phase_adjust median of the last field
        self.fields_done = 0
        self.done = False
        self.walker = None
        # the wave being walked / verified
        self.res = None
        self.valid = []
        self.row0 = 0                  # first row of the lane's fields in its system's field list
        self.prev_at_start = None


class _SystemSet:
    """Device buffers of one system in one buffer set."""

    def __init__(self, rf, cap_blocks, max_fields, audio):
        torch = rf.torch
        dev = rf.device
        S = rf.block_stride
        self.rf = rf
        self.cap_blocks = cap_blocks
        self.total = cap_blocks * S
        names = list(rf.video_fields) + ["rfhpf"]
        self.buffers = {n: torch.empty(self.total, dtype=torch.float32, device=dev) for n in names}
        # in-pulse bits start at zero (no pulses): slots a wave leaves unused
        # hold zeros or an earlier wave's bits, never arbitrary memory, so the
        # system-wide pulse search stays within its edge capacity
        self.buffers["demod_05_below"] = torch.zeros(2 * -(-self.total // 64), dtype=torch.int32, device=dev)
        self.audio = audio
        if audio:
            na = rf.filterbank.audio["left"]["nbins"]
            fdiv = rf.blocklen // na
            self.alen = na - rf.blockcut // fdiv - (-(-32 // fdiv))
            for n in ("audio_left", "audio_right"):
                self.buffers[n] = torch.empty(cap_blocks * self.alen, dtype=torch.float32, device=dev)
        self.finder = locatemod.PulseFinder(rf, self.total)
        L = locatemod.proclines(rf)
        self.max_fields = max_fields
        self.L = L
        nl = max_fields * L
        self.ll1 = torch.empty(nl, dtype=torch.float64, device=dev)
        self.bad1 = torch.zeros(nl, dtype=torch.uint8, device=dev)
        self.ll2 = torch.empty(nl, dtype=torch.float64, device=dev)
        self.bad2 = torch.empty(nl, dtype=torch.uint8, device=dev)
        self.ll = torch.empty(nl, dtype=torch.float64, device=dev)
        self.bad = torch.empty(nl, dtype=torch.uint8, device=dev)
        self.phase = torch.empty(max_fields, dtype=torch.float64, device=dev)
        self.fpid = torch.empty(max_fields, dtype=torch.int32, device=dev)
        self.ll_host = torch.empty(nl, dtype=torch.float64, pin_memory=True)
        self.bad_host = torch.empty(nl, dtype=torch.uint8, pin_memory=True)
        self.phase_host = torch.empty(max_fields, dtype=torch.float64, pin_memory=True)
        self.W = rf.SysParams["outlinelen"]
        self.linesout = rf.SysParams["frame_lines"] // 2 + 1
        self.out_u16 = torch.empty(max_fields * self.linesout * self.W, dtype=torch.uint16, device=dev)
        lib, h = rf.lib, rf.handle
        self.hs_ws = torch.empty(max(1, lib.ld_hsync_workspace_bytes(h, max_fields)), dtype=torch.uint8, device=dev)
        self.ln_ws = torch.empty(max(1, lib.ld_refine_workspace_bytes(h, max_fields, nl)), dtype=torch.uint8,
                                 device=dev)
        self.tbc_ws = torch.empty(lib.ld_tbc_workspace_bytes(h, max_fields), dtype=torch.uint8, device=dev)
        fsz = ctypes.sizeof(_lib.FieldFacts)
        self.facts_dev = torch.empty(max_fields * fsz, dtype=torch.uint8, device=dev)
        self.facts_host = torch.empty(max_fields * fsz, dtype=torch.uint8, pin_memory=True)
        self.facts_ws = torch.empty(lib.ld_field_facts_workspace_bytes(h, max_fields), dtype=torch.uint8, device=dev)
        self.outputs = _lib.DemodOutputs()
        for n, t in self.buffers.items():
            setattr(self.outputs, n, t.data_ptr())
        self.outputs.video_stride = S
        self.outputs.audio_stride = self.alen if audio else 0
        self.outputs.below_high = locatemod.pulse_threshold(rf)
        self.fields = []
        self.facts_done = None


class BatchDecoder:
    """Decode a batch of captures (``synth.Batch`` or any int16 device tensor
    with per-capture offsets, lengths and systems) in waves of about
    `wave_blocks` blocks.  rfs: {"NTSC": RFDecode, "PAL": RFDecode} (the
    systems present), equal block length, on one device.

    `run()` decodes the whole batch and returns the number of fields;
    `on_wave(info)`, when given, is called after every wave's stage B has
    been queued, with the wave's per-system fields and device buffers (valid
    until the same buffer set is reused two waves later; call
    ``torch.cuda.synchronize()`` before reading them)."""

    def __init__(self, rfs, data, systems, offsets, lengths, wave_blocks=8192, min_fields=16, graph=True,
                 audio=True, coord_f32=True, on_wave=None, streams=None, demod_order=None):
        self.rfs = dict(rfs)
        rf0 = next(iter(self.rfs.values()))
        self.torch = torch = rf0.torch
        self.device = rf0.device
        self.N, self.S = rf0.blocklen, rf0.block_stride
        for rf in self.rfs.values():
            if rf.blocklen != self.N or rf.device != self.device:
                raise ValueError("the systems' decoders differ in block length or device")
        self.data = data
        self.lanes = [Lane(i, s, o, n) for i, (s, o, n) in enumerate(zip(systems, offsets, lengths))]
        for ln in self.lanes:
            if ln.system not in self.rfs:
                raise ValueError(f"capture {ln.idx}: no decoder for {ln.system}")
            if ln.length < self.N:
                ln.done = True
        self.wave_blocks = int(wave_blocks)
        self.min_fields = int(min_fields)
        self.coord_f32 = coord_f32
        self.on_wave = on_wave
        self.audio = bool(audio)
        # host threads of the lanes' native field walks (ld_locate_fields_many)
        self.walk_threads = min(8, os.cpu_count() or 1)
        S = self.S
        # per system: blocks of one field and of one field read
        self.field_blocks, self.read_blocks, self.min_blocks, self.P = {}, {}, {}, {}
        for sname, rf in self.rfs.items():
            sp = rf.SysParams
            field_samples = sp["frame_lines"] / 2 * sp["line_period"] * rf.freq
            self.field_blocks[sname] = field_samples / S
            self.read_blocks[sname] = -(-locatemod.readlen(rf) // S) + 1
            self.min_blocks[sname] = int(math.ceil(self.min_fields * self.field_blocks[sname])) + \
                self.read_blocks[sname] + 2
            self.P[sname] = locatemod.locate_params(rf)
        self.sets = []
        self.table_cap = 0
        caps = {}
        for sname, rf in self.rfs.items():
            nl = sum(1 for ln in self.lanes if ln.system == sname)
            cap = self.wave_blocks + nl * self.min_blocks[sname] if nl else 0
            caps[sname] = cap
            self.table_cap += cap
        for _ in range(2):
            sets = {}
            for sname, rf in self.rfs.items():
                if caps[sname]:
                    mf = int(caps[sname] / self.field_blocks[sname]) + 2 * sum(
                        1 for ln in self.lanes if ln.system == sname) + 8
                    sets[sname] = _SystemSet(rf, caps[sname], mf, self.audio)
            tbl_host = (_lib.BatchBlock * self.table_cap)()
            tbl_np = np.ctypeslib.as_array(tbl_host).view(np.uint8)
            tbl_pin = torch.empty(tbl_np.size, dtype=torch.uint8, pin_memory=True)
            tbl_dev = torch.empty(tbl_np.size, dtype=torch.uint8, device=self.device)
            self.sets.append({"sys": sets, "tbl_pin": tbl_pin, "tbl_dev": tbl_dev, "graph_demod": None,
                              "graph_pulses": None, "free": None, "lanes": [], "wave": -1})
            st = self.sets[-1]
            st["pulse_ws"] = torch.empty(locatemod.pulse_search_workspace(self._searches(st)), dtype=torch.uint8,
                                         device=self.device)
        self.ws = torch.empty(max(1, rf0.lib.ld_demod_workspace_bytes(rf0.handle, self.table_cap)),
                              dtype=torch.uint8, device=self.device)
        self.handles = (ctypes.c_void_p * 2)(*[self.rfs[s].handle.value if s in self.rfs else None
                                               for s in SYSTEMS])
        self.use_graph = bool(graph)
        # streams: (A, B) shared with other decoders taking turns (their
        # demodulations then run one after the other, as they would fill the
        # GPU anyway); default: a pair of this decoder's own
        if streams is None:
            streams = (torch.cuda.Stream(device=self.device), torch.cuda.Stream(device=self.device))
        self.stream_a, self.stream_b = streams
        # demod_order: a dict shared by decoders taking turns; each demod
        # launch waits for the one queued before it by any of them and for
        # that wave's pulse search (one demodulation fills the GPU; their
        # order then follows the host's)
        self.demod_order = demod_order
        # demod_timing: a list receiving (start event, end event, algorithmic
        # bytes) of every wave's demod graph (the block table upload and the
        # ld_demod_batch launch) on stream A
        self.demod_timing = None
        self.reset()

    def reset(self):
        """Start the batch again from every capture's beginning (buffers,
        graphs and streams are kept)."""
        for ln in self.lanes:
            ln.start, ln.prev, ln.prov_delta, ln.phase_seed = 0.0, None, 0.0, 0.0
            ln.fields_done, ln.res, ln.valid = 0, None, []
            ln.done = ln.length < self.N
        for st in self.sets:
            st["free"] = None
        if self.demod_order is not None:
            self.demod_order["last"] = None
        self.waves = 0
        self.fields_total = 0
        self.relocated = 0
        self.redone_waves = 0
        self.blocks_total = 0
        self.blocks_by_system = {s: 0 for s in self.rfs}

    # ------------------------------------------------------------ planning
    def _plan(self, k):
        """Assign each active lane its window of blocks for wave k (set k % 2)
        and fill the set's block table.  Returns False when no lane is left."""
        st = self.sets[k % 2]
        lanes = [ln for ln in self.lanes if not ln.done]
        if not lanes:
            return False
        S, N = self.S, self.N
        rem = {ln.idx: max(ln.length - ln.start, 0.0) for ln in lanes}
        tot = sum(rem.values()) or 1.0
        rows = np.zeros(self.table_cap, dtype=[("first", np.int64), ("sys", np.int32), ("slot", np.int32)])
        rows["sys"] = -1
        used = {s: 0 for s in self.rfs}
        i = 0
        st["lanes"] = []
        st["geom"] = {}
        for ln in lanes:
            sset = st["sys"][ln.system]
            # the block holding the next field read's first block, found as
            # the walk finds it (decodefield's read rounded down to a
            # multiple of blocklen, core.py:4030-4041, then DemodCache's block)
            readloc = max(int(ln.start - self.rfs[ln.system].blockcut), 0)
            b0 = (readloc // N) * N // S
            avail = (ln.length - N) // S + 1 - b0
            want = max(self.min_blocks[ln.system], int(self.wave_blocks * rem[ln.idx] / tot))
            nb = min(want, avail, sset.cap_blocks - used[ln.system])
            if nb < self.read_blocks[ln.system]:
                if avail < self.read_blocks[ln.system]:
                    ln.done = True          # no complete field read left in the capture
                    continue
                nb = 0                      # no room in this wave
            if nb <= 0:
                continue
            slot0 = used[ln.system]
            st["geom"][ln.idx] = (b0, nb, slot0)
            used[ln.system] += nb
            sysid = SYSTEMS.index(ln.system)
            rows["first"][i:i + nb] = ln.offset + (b0 + np.arange(nb, dtype=np.int64)) * S
            rows["sys"][i:i + nb] = sysid
            rows["slot"][i:i + nb] = slot0 + np.arange(nb, dtype=np.int32)
            i += nb
            st["lanes"].append(ln)
        if not st["lanes"]:
            return False
        st["nslots"] = (ctypes.c_int32 * 2)(*[st["sys"][s].cap_blocks if s in st["sys"] else 0 for s in SYSTEMS])
        st["tbl_pin"].numpy()[:] = rows.view(np.uint8)
        st["used"] = dict(used)
        self.blocks_total += i
        for sname, nbk in used.items():
            self.blocks_by_system[sname] += nbk
        return True

    # ------------------------------------------------------------ stage A
    def _demod_body(self, st):
        """The wave's block table upload and its one ld_demod_batch launch."""
        torch = self.torch
        cur = torch.cuda.current_stream(self.device)
        st["tbl_dev"].copy_(st["tbl_pin"], non_blocking=True)
        outs = (_lib.DemodOutputs * 2)()
        for s, sset in st["sys"].items():
            outs[SYSTEMS.index(s)] = sset.outputs
        rf0 = next(iter(self.rfs.values()))
        _lib.check(rf0.lib.ld_demod_batch(self.handles, self.data.data_ptr(), _lib.LD_IN_INT16, self.data.numel(),
                                          st["tbl_dev"].data_ptr(), self.table_cap, st["nslots"], 1, outs,
                                          self.ws.data_ptr(), self.ws.numel(), ctypes.c_void_p(cur.cuda_stream)))

    def _pulses_body(self, st):
        """Every system's pulse search over the demod's in-pulse bits, the
        pulses' levels and their export to pinned host memory, in one
        ld_pulse_search call (four launches for all systems: these kernels
        sit between one demod launch and the next)."""
        cur = self.torch.cuda.current_stream(self.device)
        locatemod.pulse_search(self._searches(st), st["pulse_ws"], cur)

    def _searches(self, st):
        """The set's pulse searches: every system's demod_05 and in-pulse bits."""
        return [(sset.finder, sset.buffers["demod_05"], sset.total, sset.buffers["demod_05_below"])
                for sset in st["sys"].values()]

    def _replay(self, st, name, body):
        """Run `body` as a HIP graph captured once per buffer set (its
        pointers and sizes are fixed per set; the table's contents change
        per wave) and replayed on stream A."""
        torch = self.torch
        if not self.use_graph:
            body(st)
            return
        if st.get(name) is None:
            g = torch.cuda.CUDAGraph()
            self.stream_a.synchronize()
            with torch.cuda.graph(g, stream=self.stream_a):
                body(st)
            st[name] = g
        st[name].replay()

    def _stage_a(self, k):
        torch = self.torch
        st = self.sets[k % 2]
        with torch.cuda.stream(self.stream_a):
            if st["free"] is not None:
                self.stream_a.wait_event(st["free"])
            if self.demod_order is not None and self.demod_order.get("last") is not None:
                self.stream_a.wait_event(self.demod_order["last"])
            timing = self.demod_timing is not None
            if timing:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(self.stream_a)
            self._replay(st, "graph_demod", self._demod_body)
            if timing:
                e1.record(self.stream_a)
                self.demod_timing.append((e0, e1, sum(n * self.S * DEMOD_BYTES[s] for s, n in st["used"].items())))
            self._replay(st, "graph_pulses", self._pulses_body)
            ev = torch.cuda.Event()
            ev.record(self.stream_a)
            if self.demod_order is not None:
                # the next decoder's demod waits for this wave's pulse search
                # too: those short kernels would otherwise wait for CUs behind
                # it, and the host's walk of this wave behind them
                self.demod_order["last"] = ev
        st["pulses_done"] = ev
        st["wave"] = k

    # ------------------------------------------------------------ host walk
    def _walk_start(self, k, facts_for=None):
        """Start the native walks of wave k's lanes (after its pulses
        arrive) on host threads; _walk_finish collects them.  facts_for:
        {lane idx: measured facts per visited field} for a repeated walk."""
        st = self.sets[k % 2]
        st["pulses_done"].synchronize()
        S = self.S
        pulses = {}
        for s, sset in st["sys"].items():
            cnt = int(sset.finder.h_count[0])
            if cnt > sset.finder.cap:
                raise RuntimeError(f"{s}: {cnt} pulse edges exceed the search capacity {sset.finder.cap}")
            pulses[s] = (sset.finder.h_edges.numpy()[:cnt], sset.finder.h_means.numpy()[:cnt // 2])
        todo = []
        for ln in st["lanes"]:
            if facts_for is not None and ln.idx not in facts_for:
                continue
            rf = self.rfs[ln.system]
            sset = st["sys"][ln.system]
            edges, means = pulses[ln.system]
            b0, nb, slot0 = st["geom"][ln.idx]
            lo, hi = slot0 * S, (slot0 + nb) * S
            # pulse pairs inside the lane's region (edges are sorted pairs)
            a = np.searchsorted(edges[0::2], lo, side="left")
            b = np.searchsorted(edges[1::2], hi, side="left")
            e = edges[2 * a:2 * b] - lo + b0 * S        # capture coordinates
            m = means[a:b]
            if ln.walker is None:
                ln.walker = locatemod.FieldWalker(rf, nb * S, 8, finder=False)
                ln.walker.allow_recalibration = False
            w = ln.walker
            w.total = (b0 + nb) * S
            w.max_fields = int(nb / self.field_blocks[ln.system]) + 4
            w.prov_delta = ln.prov_delta
            if facts_for is None:
                ln.prev_at_start = ln.prev
                ln.start_at_wave = ln.start
            fk = None if facts_for is None else facts_for[ln.idx]
            # the walk runs in capture coordinates (stitched index of the
            # capture's block 0 = 0), as over the whole capture: its read
            # positions round to multiples of blocklen
            call = w.first_call(ln.start_at_wave, ln.prev_at_start, fk, edges_means=(e, m),
                                params=locatemod.locate_params_from(self.P[ln.system], ln.prov_delta))
            todo.append((ln, sset.buffers["demod_05"][lo:hi], fk, (e, m), b0 * S, call))
        # every lane's first native walk call at once, over host threads
        return todo, locatemod.PendingWalk([t[5] for t in todo], threads=self.walk_threads)

    def _walk_finish(self, k, pending):
        st = self.sets[k % 2]
        todo, walks = pending
        per_sys = {s: [] for s in st["sys"]}
        for (ln, view, fk, em, origin, _), (args, nf) in zip(todo, walks.wait()):
            # the records and arrays are views of the lane's own walker
            # buffers: valid until its next walk
            ln.res = ln.walker.walk_records(view, ln.start_at_wave, ln.prev_at_start, fk, args, nf,
                                            edges_means=em, origin=origin)
            ln.valid = np.flatnonzero(ln.res[0]["status"] == _lib.LD_LOC_OK)
        for ln in st["lanes"]:
            per_sys[ln.system].append(ln)
        return per_sys

    def _walk(self, k, facts_for=None):
        return self._walk_finish(k, self._walk_start(k, facts_for))

    def _fields(self, st, s, lanes):
        """The system's field plan and line arrays from its lanes' walks
        (column-wise over each lane's valid fields)."""
        sset = st["sys"][s]
        rf = self.rfs[s]
        S, L = self.S, sset.L
        pal = s == "PAL"
        parts = {k: [] for k in ("demod_offset", "demod_len", "linecount", "first_field", "chain", "lane")}
        lls, bads = [], []
        chain = 0
        row = 0
        for ln in lanes:
            locs, ll1, bad1 = ln.res[:3]
            b0, _, slot0 = st["geom"][ln.idx]
            ln.row0 = row
            v = ln.valid
            if not len(v):
                continue
            base = (slot0 - b0) * S
            sl = locs["startloc"][v]
            parts["demod_offset"].append(base + sl)
            parts["demod_len"].append(np.minimum(locs["buflen"][v], sset.total - base - sl))
            parts["linecount"].append(locs["linecount"][v])
            parts["first_field"].append(locs["is_first_field"][v])
            parts["chain"].append(np.full(len(v), chain, dtype=np.int32))
            parts["lane"].append(np.full(len(v), ln.idx, dtype=np.int64))
            row += len(v)
            lls.append(ll1[v])
            bads.append(bad1[v])
            chain += 1
        if row > sset.max_fields:
            raise RuntimeError(f"{s}: {row} fields exceed the set's {sset.max_fields}")
        cols = {k: (np.concatenate(v) if v else np.zeros(0, dtype=np.int64)) for k, v in parts.items()}
        first = cols["first_field"]
        cols["lineoffset"] = np.where(first != 0, 2, 3).astype(np.int32) if pal else np.zeros(row, np.int32)
        r = np.arange(row, dtype=np.int64)
        cols.update(nlines=np.full(row, L, np.int32), linesout=np.full(row, sset.linesout, np.int32),
                    outwidth=np.full(row, sset.W, np.int32), inlinelen=np.full(row, float(rf.linelen)),
                    lineloc_offset=r * L, out_offset=r * (sset.linesout * sset.W))
        fields = fieldmod.ColumnPlan(cols)
        sset.fields = fields
        sset.nchains = chain
        if row:
            n = row * L
            ll = np.concatenate(lls)
            bad = fieldmod.compute_deriv_error_rows(ll, np.concatenate(bads).astype(bool))
            sset.ll_host.numpy()[:n] = ll.reshape(-1)
            sset.bad_host.numpy()[:n] = bad.reshape(-1)
        return fields

    # ------------------------------------------------------------ stage B
    def _stage_b(self, st, s, lanes):
        torch = self.torch
        sset = st["sys"][s]
        rf = self.rfs[s]
        fields = sset.fields
        if not fields:
            sset.facts_done = None
            return
        n = len(fields) * sset.L
        sset.ll1[:n].copy_(sset.ll_host[:n], non_blocking=True)
        sset.bad1[:n].copy_(sset.bad_host[:n], non_blocking=True)
        fieldmod.refine_linelocs_hsync_fields(rf, sset.buffers["demod_05"], fields, sset.ll1[:n], sset.bad1[:n],
                                             out=(sset.ll2, sset.bad2), workspace=sset.hs_ws)
        seeds = [ln.phase_seed for ln in lanes if len(ln.valid)] if s == "NTSC" else None
        fieldmod.refine_lines_fields(rf, sset.buffers, fields, sset.ll2[:n], sset.bad2[:n],
                                     out=(sset.ll, sset.bad, sset.phase, sset.fpid), workspace=sset.ln_ws,
                                     chain_seeds=seeds)
        fieldmod.tbc_fields(rf, sset.buffers["demod"], fields, sset.ll, coord_f32=self.coord_f32,
                            out_u16=sset.out_u16, final=True, workspace=sset.tbc_ws)
        cur = torch.cuda.current_stream(self.device)
        arr = fieldmod._line_fields(fields)
        length = rf.SysParams["line_period"] * ((rf.freq / rf.linelen) * rf.linelen)
        _lib.check(rf.lib.ld_field_facts_measure(
            rf.handle, sset.buffers["demod"].data_ptr(), sset.total, arr, len(fields), sset.ll.data_ptr(),
            sset.ll.numel(), 1 if s == "PAL" else 0, sset.linesout, float(length), sset.facts_dev.data_ptr(),
            sset.facts_ws.data_ptr(), sset.facts_ws.numel(), ctypes.c_void_p(cur.cuda_stream)))
        nb = len(fields) * ctypes.sizeof(_lib.FieldFacts)
        sset.facts_host[:nb].copy_(sset.facts_dev[:nb], non_blocking=True)
        sset.phase_host[:len(fields)].copy_(sset.phase[:len(fields)], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(cur)
        sset.facts_done = ev

    def _queue_stage_b(self, k, per_sys):
        st = self.sets[k % 2]
        torch = self.torch
        with torch.cuda.stream(self.stream_b):
            self.stream_b.wait_event(st["pulses_done"])
            for s, lanes in per_sys.items():
                self._fields(st, s, lanes)
                self._stage_b(st, s, lanes)
            ev = torch.cuda.Event()
            ev.record(self.stream_b)
        st["free"] = ev
        st["per_sys"] = per_sys

    # ------------------------------------------------------------ verify
    def _facts(self, sset):
        """The measured facts of the set's fields (a FACTS_DTYPE copy of the
        pinned buffer, which the next use of the set overwrites)."""
        n = len(sset.fields)
        raw = sset.facts_host[:n * locatemod.FACTS_DTYPE.itemsize].numpy()
        return raw.view(locatemod.FACTS_DTYPE).copy()

    def _verify(self, k, max_rounds=4):
        """Check wave k's walks against its measured facts (re-walking and
        re-running stage B of a system whose walk changes), then carry each
        lane's state to the next wave.  Returns the lanes whose next read
        position changed from what the walk first gave."""
        st = self.sets[k % 2]
        moved = set()
        for rounds in range(max_rounds + 1):
            changed = {}
            for s, sset in st["sys"].items():
                if sset.facts_done is not None:
                    sset.facts_done.synchronize()
                measured = self._facts(sset)
                rf = self.rfs[s]
                for ln in st["per_sys"][s]:
                    locs, ll1, bad1, _, nxt0, vp = ln.res
                    facts = np.zeros(len(locs), locatemod.FACTS_DTYPE)
                    facts[ln.valid] = measured[ln.row0:ln.row0 + len(ln.valid)]
                    ln.facts = facts
                    self._tune(ln, measured)
                    P = locatemod.locate_params_from(self.P[s], ln.prov_delta)
                    # the records are checked (and updated) in place
                    first = locatemod.recheck_records(rf, locs, ll1, bad1, vp, facts, ln.prev_at_start, P)
                    if first >= 0:
                        fk = facts.copy()
                        fk[first:] = 0
                        changed.setdefault(s, {})[ln.idx] = fk
            if not changed:
                break
            if rounds == max_rounds:
                raise RuntimeError("line location did not settle")
            self.relocated += 1
            # walk the changed lanes again and redo stage B of their systems
            with self.torch.cuda.stream(self.stream_b):
                for s, fl in changed.items():
                    before = {ln.idx: ln.res[4] for ln in st["per_sys"][s] if ln.idx in fl}
                    per = self._walk(k, facts_for=fl)
                    for ln in st["per_sys"][s]:
                        if ln.idx in fl and ln.res[4] != before[ln.idx]:
                            moved.add(ln.idx)
                    self._fields(st, s, per[s])
                    self._stage_b(st, s, per[s])
                ev = self.torch.cuda.Event()
                ev.record(self.stream_b)
                st["free"] = ev
        # carry state to the next wave
        for s, sset in st["sys"].items():
            phases = sset.phase_host.numpy()
            rf = self.rfs[s]
            for ln in st["per_sys"][s]:
                locs, ll1, bad1, prev_out, nxt, vp = ln.res
                P = locatemod.locate_params_from(self.P[s], ln.prov_delta)
                if len(locs):
                    last = _lib.FieldLoc.from_buffer_copy(locs[-1:].tobytes())
                    written = prev_out.fields_written - (1 if last.status == _lib.LD_LOC_OK else 0)
                    ln.prev = locatemod._prev_from(last, ll1[-1], rf, P, written,
                                                   _lib.FieldFacts.from_buffer_copy(ln.facts[-1:].tobytes()))
                if len(ln.valid) and s == "NTSC":
                    ln.phase_seed = float(phases[ln.row0 + len(ln.valid) - 1])
                ln.fields_done += len(ln.valid)
                self.fields_total += len(ln.valid)
                if nxt <= ln.start and not len(ln.valid):
                    ln.done = True            # no progress: the capture's end
                ln.start = nxt
        return moved

    def _tune(self, ln, measured):
        """The lane's provisional line-count offset: the median over its
        valid fields of the measured final position of the linecount line
        minus its linelocs1 entry."""
        v = ln.valid
        if not len(v):
            return
        locs, ll1 = ln.res[0], ln.res[1]
        if ln.system == "PAL":
            lc = np.where(locs["is_first_field"][v] != 0, 312, 313)
        else:
            lc = locs["linecount"][v]
        f = measured[ln.row0:ln.row0 + len(v)]
        ok = (f["known"] != 0) & np.isfinite(f["ll_linecount"])
        if np.any(ok):
            d = f["ll_linecount"][ok] - ll1[v[ok], lc[ok]]
            ln.prov_delta = float(np.median(d))

    # ------------------------------------------------------------ driver
    def waves_iter(self, max_waves=None, offset=0):
        """The decode as a generator that yields before each call that waits
        for the GPU (a wave's pulses before its walk, its facts before its
        check) and while its walks run on host threads, so that several
        decoders can take turns on the host while the GPU works on the
        others' queued stages (run_interleaved).  Each wave takes three
        turns (start the walks; collect them, queue stage B and the next
        demod; check the walks); `offset` empty turns at the start stagger
        decoders taking turns, so that one's walks run during another's
        check."""
        k = 0
        if not self._plan(0):
            return
        self._stage_a(0)
        for _ in range(offset):
            yield
        while True:
            yield
            pending = self._walk_start(k)
            # the walks run on host threads while the other decoders take
            # their turns
            yield
            per_sys = self._walk_finish(k, pending)
            self._queue_stage_b(k, per_sys)
            if self.on_wave is not None:
                self.on_wave(self._info(k))
            # next wave's demod from the provisional next positions, beside
            # this wave's stage B
            saved = {ln.idx: ln.start for ln in self.lanes}
            for ln in self.sets[k % 2]["lanes"]:
                ln.start = ln.res[4]
            nxt_ok = (max_waves is None or k + 1 < max_waves) and self._plan(k + 1)
            if nxt_ok:
                self._stage_a(k + 1)
            for idx, v in saved.items():
                self.lanes[idx].start = v
            yield
            moved = self._verify(k)
            self.waves += 1
            if moved and nxt_ok:
                # a re-walked lane moves: plan and demodulate the next wave again
                self.redone_waves += 1
                nxt = self.sets[(k + 1) % 2]
                nxt["pulses_done"].synchronize()
                self.blocks_total -= sum(g[1] for g in nxt["geom"].values())
                for sname, nbk in nxt["used"].items():
                    self.blocks_by_system[sname] -= nbk
                nxt_ok = self._plan(k + 1)
                if nxt_ok:
                    self._stage_a(k + 1)
            if not nxt_ok:
                break
            k += 1

    def join(self):
        """Make torch's current stream wait for everything queued."""
        cur = self.torch.cuda.current_stream(self.device)
        cur.wait_stream(self.stream_a)
        cur.wait_stream(self.stream_b)

    def run(self, max_waves=None):
        """Decode every capture to its end.  Returns the fields decoded."""
        for _ in self.waves_iter(max_waves):
            pass
        self.join()
        return self.fields_total

    def _info(self, k):
        st = self.sets[k % 2]
        out = {"wave": k}
        for s, sset in st["sys"].items():
            out[s] = {"fields": list(sset.fields), "out_u16": sset.out_u16, "linelocs": sset.ll,
                      "buffers": sset.buffers, "lanes": {ln.idx: st["geom"][ln.idx] for ln in st["lanes"]
                                                          if ln.system == s}}
        return out

    @property
    def samples_total(self):
        return sum(ln.length for ln in self.lanes)


def run_interleaved(decoders):
    """Run several BatchDecoders (each with its own streams and buffers)
    taking turns on the host: while one waits for or walks its wave, the
    GPU runs the stages the others have queued.  Returns the fields decoded."""
    its = [d.waves_iter(offset=i % 3 if len(decoders) > 1 else 0) for i, d in enumerate(decoders)]
    while its:
        nxt = []
        for it in its:
            try:
                next(it)
                nxt.append(it)
            except StopIteration:
                pass
        its = nxt
    for d in decoders:
        d.join()
    return sum(d.fields_total for d in decoders)


def split_captures(lengths, groups):
    """Indices of the captures per group, longest first onto the group with
    the fewest samples (LPT), each group keeping the batch's order."""
    load = [0] * groups
    out = [[] for _ in range(groups)]
    for k in sorted(range(len(lengths)), key=lambda k: -lengths[k]):
        g = min(range(groups), key=lambda g: load[g])
        out[g].append(k)
        load[g] += lengths[k]
    return [sorted(o) for o in out]
