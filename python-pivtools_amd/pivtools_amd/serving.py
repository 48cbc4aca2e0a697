"""Continuous-batching PIV server on one GPU.

Requests (one image pair of one camera stream each) arrive at any time.
Each stream owns a small pool of *slots*.  A slot is a resident
``MultipassPIV`` with its own batch buffers, fixed input staging buffers, a
HIP stream and captured HIP graphs of the whole pass loop.  The scheduler is
a polling loop:

* a slot whose graph has finished hands its pairs' final fields back (copied
  out of the slot, so the slot can be reused at once) and becomes free;
* a free slot takes up to ``max_batch`` waiting pairs of its stream, copies
  them into its staging buffers on its stream and replays the graph captured
  for that batch size (capturing it the first time, after one eager run);
* new slots are created while the HBM budget allows, so a stream with a
  backlog gets more batches in flight; slots of different streams and of the
  same stream run concurrently on their own HIP streams.

Admission is by HBM budget.  A slot is charged its real device bytes: its
16-bit staging frames plus every buffer of its pass loop at ``max_batch``
pairs (``MultipassPIV.footprint_bytes``; the slot allocates them once and
every batch size uses their leading part).  A slot is created only when it
fits; when it does not and a stream waits with no slot at all, idle slots of
other streams without waiting work are released until it fits.  A stream
whose single slot exceeds the budget is refused at construction.  Batches
are never padded: each slot captures one graph per batch size it has run.

A request that arrives while others are on the GPU is admitted as soon as a
slot of its stream is free.  Batches never mix streams: a stream's image size
and window schedule fix its buffers and graphs.  Streams flagged
``ensemble`` also add every pair to that stream's ensemble accumulators
(after the pass loop, on the stream's ensemble queue).

A stream can also take its pairs as raw camera payloads (``pixel_format``
"mono12p" or "bayer_rg8", with optional A and B dark frames): the slot then
expands them on its own stream into the staging frames (``ingest.decode``),
so only the packed bytes have to reach the GPU.

Every pair is correlated independently of the others in its batch, so a
pair's result does not depend on what it was batched with.  The serving tests
check that a pair's result is bit-identical alone and in the mix.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import torch

from . import ingest
from .ensemble import EnsemblePIV
from .multipass import MultipassPIV


@dataclass
class StreamSpec:
    name: str
    shape: tuple
    window_sizes: list
    overlaps: list | None = None
    peak_finder: str = "gauss3"
    ensemble: bool = False
    ensemble_window: tuple = (32, 32)
    ensemble_overlap: int = 75
    # "u16": requests carry 16-bit frames of ``shape``; "mono12p" /
    # "bayer_rg8": requests carry the sensor payload (uint8, one frame's
    # bytes) of a ``sensor_shape`` sensor, decoded into ``shape`` frames
    pixel_format: str = "u16"
    sensor_shape: tuple | None = None
    dark_a: torch.Tensor | None = None     # sensor-type dark of the A frames (device)
    dark_b: torch.Tensor | None = None     # and of the B frames
    # further MultipassPIV options of the stream (outlier_methods,
    # mid_infill, final_infill, vector_masks / pixel_mask, num_peaks, ...)
    piv_options: dict = field(default_factory=dict)


@dataclass
class Request:
    rid: int
    stream: int
    a: torch.Tensor
    b: torch.Tensor
    t_arrival: float
    t_start: float = 0.0
    t_done: float = 0.0
    ux: torch.Tensor | None = None
    uy: torch.Tensor | None = None
    nan_mask: torch.Tensor | None = None
    ready: torch.cuda.Event | None = None     # frames written (producer's stream)
    done: torch.cuda.Event | None = None      # fields written (slot's stream)

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival

    def result(self):
        """(ux, uy, nan_mask), ordered after the batch on the caller's stream."""
        torch.cuda.current_stream(self.ux.device).wait_event(self.done)
        return self.ux, self.uy, self.nan_mask


@dataclass
class _Slot:
    mp: MultipassPIV
    stream: torch.cuda.Stream
    stage_a: torch.Tensor
    stage_b: torch.Tensor
    graphs: dict = field(default_factory=dict)
    outs: dict = field(default_factory=dict)
    warm: set = field(default_factory=set)
    batch: list = field(default_factory=list)
    done_event: torch.cuda.Event | None = None
    ens_event: torch.cuda.Event | None = None
    nbytes: int = 0


class PIVServer:
    def __init__(self, streams: list[StreamSpec], max_batch: int = 8, max_slots_per_stream: int = 3,
                 hbm_budget_bytes: int = 64 << 30, use_graphs: bool = True, device=None, autotune=False):
        """``autotune``: False keeps the default correlation-kernel variants;
        True times them for every window size of the streams first; a path
        string does the same through that JSON cache (measured once per
        device and library build)."""
        self.dev = torch.device(device) if device is not None else torch.device("cuda")
        self.tuned = None
        if autotune:
            from . import autotune as AT
            sizes = sorted({int(w[0] if isinstance(w, (tuple, list)) else w) for sp in streams for w in sp.window_sizes})
            self.tuned = AT.tune_xcorr(sizes, device=self.dev, cache=autotune if isinstance(autotune, str) else None)
        self.specs = list(streams)
        for sp in self.specs:
            if sp.pixel_format == "u16":
                continue
            if sp.pixel_format not in ingest.FORMATS or sp.sensor_shape is None:
                raise ValueError(f"stream {sp.name}: pixel_format 'u16' or one of {ingest.FORMATS} "
                                 "with its sensor_shape")
            if tuple(ingest.frame_shape(sp.pixel_format, *sp.sensor_shape)) != tuple(sp.shape):
                raise ValueError(f"stream {sp.name}: a {sp.sensor_shape} {sp.pixel_format} sensor gives "
                                 f"{ingest.frame_shape(sp.pixel_format, *sp.sensor_shape)} frames, not {sp.shape}")
        self.max_batch = int(max_batch)
        self.max_slots = int(max_slots_per_stream)
        self.budget = int(hbm_budget_bytes)
        self.use_graphs = use_graphs
        self.slots: list[list[_Slot]] = [[] for _ in self.specs]
        self.pending: list[list[Request]] = [[] for _ in self.specs]
        self.ens: list[EnsemblePIV | None] = []
        self.ens_stream: list[torch.cuda.Stream | None] = []
        for s in self.specs:
            if s.ensemble:
                self.ens.append(EnsemblePIV(s.shape, s.ensemble_window, s.ensemble_overlap, "square",
                                            device=self.dev))
                self.ens_stream.append(torch.cuda.Stream(device=self.dev))
            else:
                self.ens.append(None)
                self.ens_stream.append(None)
        self.bytes_in_use = 0
        self.stats = {"batches": 0, "pairs": 0, "padded_pairs": 0, "graphs_captured": 0, "slots_created": 0,
                      "evicted_slots": 0, "peak_bytes": 0}
        for i, sp in enumerate(self.specs):
            need = self._slot_bytes(i, self._slot_piv(i))
            if need > self.budget:
                raise ValueError(f"stream {sp.name}: one slot needs {need / 2**30:.2f} GiB of HBM "
                                 f"(max_batch {self.max_batch}), more than the {self.budget / 2**30:.2f} GiB budget")
        self._rid = 0
        self.completed: list[Request] = []

    # -- admission ---------------------------------------------------------
    def submit(self, stream: int, a: torch.Tensor, b: torch.Tensor, t_arrival: float | None = None) -> Request:
        spec = self.specs[stream]
        if spec.pixel_format != "u16":
            sh, sw = spec.sensor_shape
            nbytes = ingest.payload_bytes(spec.pixel_format, sh, sw)
            if a.dtype != torch.uint8 or b.dtype != torch.uint8 or not a.is_cuda or not b.is_cuda:
                raise TypeError(f"stream {spec.name} takes {spec.pixel_format} payloads as uint8 device tensors")
            if a.numel() != nbytes or b.numel() != nbytes or not a.is_contiguous() or not b.is_contiguous():
                raise ValueError(f"stream {spec.name} expects contiguous {nbytes}-byte payloads")
        else:
            if tuple(a.shape) != tuple(spec.shape) or a.shape != b.shape:
                raise ValueError(f"stream {spec.name} expects {spec.shape} frames, got {tuple(a.shape)}")
            if a.dtype not in (torch.int16, torch.uint16) or b.dtype != a.dtype or not a.is_cuda:
                raise TypeError("requests carry 16-bit camera counts (uint16 data) in device tensors")
        r = Request(self._rid, stream, a, b, time.perf_counter() if t_arrival is None else t_arrival)
        r.ready = torch.cuda.Event()
        r.ready.record(torch.cuda.current_stream(a.device))
        self._rid += 1
        self.pending[stream].append(r)
        return r

    def _slot_piv(self, i: int) -> MultipassPIV:
        spec = self.specs[i]
        return MultipassPIV(spec.shape, spec.window_sizes, spec.overlaps, peak_finder=spec.peak_finder,
                            device=self.dev, max_batch=self.max_batch, **spec.piv_options)

    def _slot_bytes(self, i: int, mp: MultipassPIV) -> int:
        H, W = self.specs[i].shape
        return 2 * self.max_batch * H * W * 2 + mp.footprint_bytes()

    def _evict_for(self, i: int, need: int) -> None:
        """Release idle slots of other streams that have no waiting pairs
        (their pass-loop buffers and graphs) until ``need`` more bytes fit."""
        for j in sorted(range(len(self.specs)), key=lambda j: -len(self.slots[j])):
            if j == i or self.pending[j]:
                continue
            for s in list(self.slots[j]):
                if self.bytes_in_use + need <= self.budget:
                    return
                if s.batch or (s.ens_event is not None and not s.ens_event.query()):
                    continue
                self.slots[j].remove(s)
                self.bytes_in_use -= s.nbytes
                self.stats["evicted_slots"] += 1

    def _new_slot(self, i: int) -> _Slot | None:
        spec = self.specs[i]
        if len(self.slots[i]) >= self.max_slots:
            return None
        mp = self._slot_piv(i)
        est = self._slot_bytes(i, mp)
        if self.bytes_in_use + est > self.budget and not self.slots[i]:
            self._evict_for(i, est)
        if self.bytes_in_use + est > self.budget:
            return None
        H, W = spec.shape
        st = torch.cuda.Stream(device=self.dev)
        with torch.cuda.stream(st):
            sa = torch.zeros((self.max_batch, H, W), dtype=torch.int16, device=self.dev)
            sb = torch.zeros_like(sa)
        slot = _Slot(mp, st, sa, sb)
        slot.nbytes = est
        self.bytes_in_use += est
        self.stats["peak_bytes"] = max(self.stats["peak_bytes"], self.bytes_in_use)
        self.stats["slots_created"] += 1
        self.slots[i].append(slot)
        return slot

    # -- execution -----------------------------------------------------------
    def _launch(self, i: int, slot: _Slot, batch: list[Request]) -> None:
        m = len(batch)
        n = m                          # no padding: one graph per batch size
        now = time.perf_counter()
        self.stats["batches"] += 1
        self.stats["pairs"] += m
        with torch.cuda.stream(slot.stream):
            if slot.ens_event is not None:
                slot.stream.wait_event(slot.ens_event)   # previous ensemble reads of the staging
            for r in batch:
                slot.stream.wait_event(r.ready)          # frames produced on the caller's stream
            spec = self.specs[i]
            for k in range(n):
                r = batch[k]
                if spec.pixel_format == "u16":
                    slot.stage_a[k].copy_(r.a, non_blocking=True)
                    slot.stage_b[k].copy_(r.b, non_blocking=True)
                else:
                    sh, sw = spec.sensor_shape
                    ingest.decode(spec.pixel_format, r.a, sh, sw, n_frames=1, dark=spec.dark_a,
                                  out=slot.stage_a[k:k + 1])
                    ingest.decode(spec.pixel_format, r.b, sh, sw, n_frames=1, dark=spec.dark_b,
                                  out=slot.stage_b[k:k + 1])
            for r in batch:
                r.t_start = now
            a, b = slot.stage_a[:n], slot.stage_b[:n]
            if not self.use_graphs:
                res = slot.mp.run(a, b)[-1]
                out = (res.ux, res.uy, res.nan_mask)
            elif n in slot.graphs:
                slot.graphs[n].replay()
                out = slot.outs[n]
            elif n in slot.warm:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=slot.stream):
                    res = slot.mp.run(a, b)[-1]
                slot.graphs[n] = g
                slot.outs[n] = (res.ux, res.uy, res.nan_mask)
                self.stats["graphs_captured"] += 1
                g.replay()
                out = slot.outs[n]
            else:
                res = slot.mp.run(a, b)[-1]      # eager run sizes the buffers for this batch
                slot.warm.add(n)
                out = (res.ux, res.uy, res.nan_mask)
            for k, r in enumerate(batch):
                r.ux = out[0][k].clone()
                r.uy = out[1][k].clone()
                r.nan_mask = out[2][k].clone()
            ev = torch.cuda.Event()
            ev.record(slot.stream)
            slot.done_event = ev
            for r in batch:
                r.done = ev
        slot.ens_event = None
        ens = self.ens[i]
        if ens is not None:
            es = self.ens_stream[i]
            with torch.cuda.stream(es):
                es.wait_event(ev)
                ens.accumulate(slot.stage_a[:m], slot.stage_b[:m])
                e2 = torch.cuda.Event()
                e2.record(es)
                slot.ens_event = e2
        slot.batch = batch

    def poll(self) -> list[Request]:
        """Retire finished batches and dispatch waiting pairs; returns the
        requests completed by this call."""
        done = []
        now = time.perf_counter()
        for i, slots in enumerate(self.slots):
            for s in slots:
                if s.batch and s.done_event.query():
                    for r in s.batch:
                        r.t_done = now
                    done.extend(s.batch)
                    s.batch = []
        for i in range(len(self.specs)):
            while self.pending[i]:
                slot = next((s for s in self.slots[i] if not s.batch and
                             (s.ens_event is None or s.ens_event.query())), None)
                if slot is None:
                    slot = self._new_slot(i)
                if slot is None:
                    break
                batch = self.pending[i][:self.max_batch]
                del self.pending[i][:len(batch)]
                self._launch(i, slot, batch)
        self.completed.extend(done)
        return done

    def busy(self) -> bool:
        return any(self.pending) or any(s.batch for slots in self.slots for s in slots)

    def drain(self, timeout_s: float = 120.0) -> None:
        t0 = time.perf_counter()
        while self.busy():
            self.poll()
            if time.perf_counter() - t0 > timeout_s:
                raise TimeoutError("server did not drain")

    def finalize_ensembles(self) -> dict:
        """Finish every ensemble accumulator (after draining)."""
        out = {}
        for i, e in enumerate(self.ens):
            if e is not None and e.n_images:
                torch.cuda.current_stream(self.dev).wait_stream(self.ens_stream[i])
                out[self.specs[i].name] = e.finalize()
        return out
