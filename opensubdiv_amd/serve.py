"""Serving loop: continuous batching of evaluation requests over topologies
whose tables stay resident in HBM.

A request names a resident topology, carries one new frame of control
points (L floats per vertex) and asks for either the topology's uniform
refinement (its level-l stencil table) or its limit points (P, dPdu, dPdv
at the topology's resident PatchCoords, through the patch-point stencils and
the patch table).  The server:

* admits waiting requests in arrival order while their device buffers fit
  the device budget and their host result buffers fit the pinned host
  budget (first-fit arenas; no allocation per request);
* batches the requests admitted since the last launch (up to a cap on the
  points per epoch) into one epoch
  (osdhip_batch_epoch: phase A stencils, phase B patches) on a free stream
  slot, with the frame uploads before it;
* keeps up to `slots` epochs in flight, so requests that arrive while others
  are on the GPU are admitted as soon as a slot and memory free up, and their
  work overlaps on the device;
* downloads results on copy streams of their own, in chunks, a few chunks in
  flight, always the next chunk of the request with the fewest bytes left
  (or, as an option, the earliest ideal completion: its arrival plus all its
  bytes at the link rate): the host link, not the GPU, bounds serving, and
  this way a small request is not queued behind another's gigabyte of
  results;
* completes a request when its download has finished; latency = completion
  time - arrival time.

Every request's result is the same bits alone and in any mix: each output
element depends only on its own segment's table, frame and coords.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib


class Arena:
    """First-fit allocator over [0, size) in `align`-unit blocks."""

    def __init__(self, size, align=64):
        self.size, self.align = int(size), int(align)
        self.free = [(0, self.size)]             # sorted (offset, length)

    def alloc(self, n):
        n = (int(n) + self.align - 1) // self.align * self.align
        for k, (o, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (o + n, ln - n)
                return o, n
        return None

    def release(self, o, n):
        self.free.append((o, n))
        self.free.sort()
        merged = []
        for a, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == a:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((a, ln))
        self.free = merged

    def used(self):
        return self.size - sum(ln for _, ln in self.free)


@dataclass
class Resident:
    """A topology's resident device data (from workloads.DeviceMesh-like
    compilation with one frame)."""
    index: int
    nctrl: int
    npts: int
    nrefine: int
    ncoords: int
    refine_table: object
    point_table: object
    coords: object
    arrays: object
    indices: object
    params: object
    frames: np.ndarray       # (K, nctrl, L) pinned-able host frames to cycle through


@dataclass
class Request:
    rid: int
    topo: int
    kind: str                # "refine" or "limit"
    arrival: float
    frame: int = 0
    admitted: float = 0.0
    done: float = 0.0
    dev: list = field(default_factory=list)      # device arena blocks
    host: tuple | None = None                    # host arena block
    points: int = 0
    pieces: list = field(default_factory=list)   # (device, host offset, floats) left to copy
    left: int = 0                                # floats left to copy
    inflight: int = 0                            # chunks on the copy stream
    deadline: float = 0.0                        # arrival + its bytes at the link rate


class Server:
    def __init__(self, residents, length, device_budget, host_budget, slots=4, item_slices=8,
                 max_epoch_points=None, chunk_bytes=16 << 20, copy_depth=4, copy_streams=2,
                 copy_order="srpt", link_bytes_per_s=45e9):
        import torch
        self.torch = torch
        self.lib = _lib.load_hip()
        self.res = residents
        self.L = int(length)
        self.dev = torch.empty(int(device_budget) // 4, dtype=torch.float32, device="cuda")
        self.dev_arena = Arena(self.dev.numel(), 64)
        self.host = torch.empty(int(host_budget) // 4, dtype=torch.float32, pin_memory=True)
        self.host_arena = Arena(self.host.numel(), 64)
        self.item_slices = int(item_slices)
        # an epoch takes waiting requests (in order) until their points reach
        # this cap, so one huge epoch does not hold every slot's successors
        self.max_epoch_points = max_epoch_points
        self.epochs = []                         # requests per launched epoch
        self.epoch_points = []                   # points per launched epoch
        self.slots = []
        for _ in range(int(slots)):
            self.slots.append({"stream": torch.cuda.Stream(), "event": torch.cuda.Event(),
                               "reqs": None, "staging": None, "scratch": None, "cap": 0})
        # result downloads: copy_streams streams (default priority, like every
        # stream here) taking chunks of chunk_bytes in turn, at most
        # copy_depth chunks in flight
        self.copy_streams = [torch.cuda.Stream() for _ in range(max(1, int(copy_streams)))]
        self.copy_turn = 0
        # which download's chunk goes next: "srpt" the fewest bytes left
        # (the lower mean and p99 latency on the served workload); "edf" the
        # earliest ideal completion (arrival + all its bytes at
        # link_bytes_per_s), which bounds how long a large request can be
        # overtaken by later small ones (the lower maximum)
        if copy_order not in ("srpt", "edf"):
            raise ValueError(f"copy_order {copy_order!r}")
        self.copy_order = copy_order
        self.link = float(link_bytes_per_s)
        self.chunk = max(1, int(chunk_bytes) // 4)
        self.copy_depth = int(copy_depth)
        self.copying = []                        # requests whose results are downloading
        self.inflight = []                       # (event, request) per chunk, in issue order
        # host frames in pinned memory, one block per topology
        self.frames = [torch.from_numpy(np.ascontiguousarray(r.frames)).pin_memory()
                       for r in residents]

    # ---- sizes ----------------------------------------------------------
    def dev_floats(self, r: Request):
        t = self.res[r.topo]
        if r.kind == "refine":
            return [t.nctrl * self.L, t.nrefine * self.L]
        return [(t.nctrl + t.npts) * self.L] + [t.ncoords * self.L] * 3

    def points(self, r: Request):
        t = self.res[r.topo]
        return t.nrefine if r.kind == "refine" else t.ncoords

    def host_floats(self, r: Request):
        t = self.res[r.topo]
        return (t.nrefine if r.kind == "refine" else 3 * t.ncoords) * self.L

    def admit(self, r: Request):
        blocks = []
        for n in self.dev_floats(r):
            b = self.dev_arena.alloc(n)
            if b is None:
                for x in blocks:
                    self.dev_arena.release(*x)
                return False
            blocks.append(b)
        h = self.host_arena.alloc(self.host_floats(r))
        if h is None:
            for x in blocks:
                self.dev_arena.release(*x)
            return False
        r.dev, r.host = blocks, h
        return True

    def release(self, r: Request):
        for x in r.dev:
            self.dev_arena.release(*x)
        self.host_arena.release(*r.host)
        r.dev, r.host = [], None

    # ---- one epoch ------------------------------------------------------
    def _segments(self, reqs):
        L, base = self.L, self.dev.data_ptr()
        S, P = [], []
        for r in reqs:
            t = self.res[r.topo]
            vb = base + 4 * r.dev[0][0]
            if r.kind == "refine":
                S.append(_lib.BatchStencils(t.refine_table.h, vb, L, 0, base + 4 * r.dev[1][0],
                                            L, 0, 1))
            else:
                S.append(_lib.BatchStencils(t.point_table.h, vb, L, 0, vb + 4 * t.nctrl * L,
                                            L, 0, 1))
                P.append(_lib.BatchPatches(t.ncoords, t.coords.data_ptr(), t.arrays.data_ptr(),
                                           t.indices.data_ptr(), t.params.data_ptr(), vb, L, 0,
                                           base + 4 * r.dev[1][0], base + 4 * r.dev[2][0],
                                           base + 4 * r.dev[3][0], L, 0, 1))
        SA = (_lib.BatchStencils * max(1, len(S)))(*S)
        PA = (_lib.BatchPatches * max(1, len(P)))(*P)
        return len(S), SA, len(P), PA

    def launch(self, slot, reqs):
        torch = self.torch
        nS, SA, nP, PA = self._segments(reqs)
        need = self.lib.osdhip_batch_epoch_bytes(self.L, nS, ctypes.cast(SA, ctypes.c_void_p), nP,
                                                 ctypes.cast(PA, ctypes.c_void_p),
                                                 self.item_slices)
        _lib.check(0 if need > 0 else need, "osdhip_batch_epoch_bytes")
        if need > slot["cap"]:
            # grow this slot's staging / scratch (only between its epochs)
            cap = max(need, 2 * slot["cap"], 1 << 20)
            slot["staging"] = torch.empty(cap, dtype=torch.uint8, pin_memory=True)
            slot["scratch"] = torch.empty(cap, dtype=torch.uint8, device="cuda")
            slot["cap"] = cap
        st = slot["stream"]
        with torch.cuda.stream(st):
            for r in reqs:
                t = self.res[r.topo]
                src = self.frames[r.topo][r.frame % self.frames[r.topo].shape[0]].reshape(-1)
                o, _ = r.dev[0]
                self.dev[o:o + t.nctrl * self.L].copy_(src, non_blocking=True)
            rc = self.lib.osdhip_batch_epoch(self.L, nS, ctypes.cast(SA, ctypes.c_void_p), nP,
                                             ctypes.cast(PA, ctypes.c_void_p), self.item_slices,
                                             slot["staging"].data_ptr(), slot["cap"],
                                             slot["scratch"].data_ptr(), slot["cap"],
                                             st.cuda_stream)
            _lib.check(rc, "osdhip_batch_epoch")
            slot["event"].record(st)
        slot["reqs"] = reqs
        self.epochs.append(len(reqs))
        self.epoch_points.append(sum(self.points(r) for r in reqs))

    # ---- result downloads -----------------------------------------------
    def _pieces(self, r: Request):
        """(device offset, host offset, floats) of a request's results."""
        ho, _ = r.host
        if r.kind == "refine":
            o, _ = r.dev[1]
            return [(o, ho, self.res[r.topo].nrefine * self.L)]
        n = self.res[r.topo].ncoords * self.L
        return [(r.dev[1 + k][0], ho + k * n, n) for k in range(3)]

    def _start_download(self, r: Request):
        r.pieces = self._pieces(r)
        r.left = sum(p[2] for p in r.pieces)
        r.deadline = r.arrival + 4.0 * r.left / self.link
        self.copying.append(r)

    def _issue_chunks(self):
        """Keeps copy_depth chunks in flight: each the next chunk of the
        downloading request with the fewest floats left."""
        torch = self.torch
        while len(self.inflight) < self.copy_depth:
            cand = [r for r in self.copying if r.pieces]
            if not cand:
                return
            r = min(cand, key=(lambda x: x.left) if self.copy_order == "srpt"
                    else (lambda x: x.deadline))
            o, ho, n = r.pieces[0]
            m = min(n, self.chunk)
            cs = self.copy_streams[self.copy_turn]
            self.copy_turn = (self.copy_turn + 1) % len(self.copy_streams)
            with torch.cuda.stream(cs):
                self.host[ho:ho + m].copy_(self.dev[o:o + m], non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(cs)
            if m == n:
                r.pieces.pop(0)
            else:
                r.pieces[0] = (o + m, ho + m, n - m)
            r.left -= m
            r.inflight += 1
            self.inflight.append((ev, r))

    def _finished_downloads(self):
        """Requests whose every chunk has landed."""
        out, keep = [], []
        for ev, r in self.inflight:
            if ev.query():
                r.inflight -= 1
                if r.inflight == 0 and not r.pieces:
                    self.copying.remove(r)
                    out.append(r)
            else:
                keep.append((ev, r))
        self.inflight = keep
        return out

    def result(self, r: Request):
        """Host copy of a completed request's output (before release)."""
        ho, _ = r.host
        n = self.host_floats(r)
        return self.host[ho:ho + n].numpy().copy()

    # ---- the loop -------------------------------------------------------
    def run(self, requests, keep=None, max_seconds=600.0):
        """Serves `requests` (sorted by arrival, seconds from start) in real
        time.  `keep`: set of request ids whose results are returned.
        Returns (completed requests, results)."""
        pending = list(requests)
        waiting, done, results = [], [], {}
        t0 = time.perf_counter()
        i = 0
        while len(done) < len(requests):
            now = time.perf_counter() - t0
            if now > max_seconds:
                raise TimeoutError("serving loop exceeded its time limit")
            while i < len(pending) and pending[i].arrival <= now:
                waiting.append(pending[i])
                i += 1
            # computed epochs: their slots free up, their results start downloading
            for sl in self.slots:
                if sl["reqs"] is not None and sl["event"].query():
                    for r in sl["reqs"]:
                        self._start_download(r)
                    sl["reqs"] = None
            # downloaded requests complete
            for r in self._finished_downloads():
                r.done = time.perf_counter() - t0
                if keep is not None and r.rid in keep:
                    results[r.rid] = self.result(r)
                self.release(r)
                done.append(r)
            self._issue_chunks()
            # admission + launch on a free slot
            free = [sl for sl in self.slots if sl["reqs"] is None]
            if waiting and free:
                batch, pts = [], 0
                while waiting and (not batch or self.max_epoch_points is None or
                                   pts + self.points(waiting[0]) <= self.max_epoch_points) \
                        and self.admit(waiting[0]):
                    r = waiting.pop(0)
                    r.admitted = now
                    pts += self.points(r)
                    batch.append(r)
                if batch:
                    self.launch(free[0], batch)
                elif not any(sl["reqs"] is not None for sl in self.slots) and not self.copying:
                    raise MemoryError(f"request {waiting[0].rid} exceeds the serving budgets")
            if i < len(pending) and not waiting and not self.copying and \
                    all(sl["reqs"] is None for sl in self.slots):
                # idle until the next arrival
                dt = pending[i].arrival - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(min(dt, 0.001))
        return done, results


def poisson_requests(n, rate, ntopo, seed, limit_fraction=0.5, nframes=8):
    """n requests, exponential inter-arrival times of mean 1/rate, uniform
    topology, kind limit with probability `limit_fraction`."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    reqs = []
    for k in range(n):
        reqs.append(Request(k, int(rng.integers(0, ntopo)),
                            "limit" if rng.uniform() < limit_fraction else "refine",
                            float(t[k]), int(rng.integers(0, nframes))))
    return reqs


def residents_from_hosts(hosts, stream=None, layout=2, keep_csr=False):
    """Resident topologies from compiled workloads.HostMesh objects (their
    `frames` become the frame pool requests cycle through); stencil tables
    in slice layout `layout` (0 or 2; layout 2 in its default row order,
    which suits one frame per request), with or without their CSR arrays."""
    from .osd import HipStencilTable, _device_tensor
    out = []
    for h in hosts:
        opts = dict(layout=layout, keep_csr=keep_csr)
        out.append(Resident(h.spec.index, h.nctrl, int(len(h.points["sizes"])),
                            int(h.refine.num_stencils), int(len(h.coords)),
                            HipStencilTable(h.refine, stream, **opts),
                            HipStencilTable(h.points, stream, **opts),
                            _device_tensor(h.coords), _device_tensor(h.arrays),
                            _device_tensor(h.indices), _device_tensor(h.params), h.frames))
    return out
