"""Serving loop: resident matrices, admission by HBM bytes, continuous batching.

    reg = MatrixRegistry(device)
    mid = reg.register(A)                    # CSR (Float32) or BF16CSR, stays in HBM
    srv = Server(reg, hbm_budget=2 << 30, slots=4)
    stats = srv.run(requests)                # requests: arrays of arrival time, matrix, k

Inputs and outputs are row-major blocks (n x k and m x k, Julia's
transpose of a k x n matrix): the k values one stored entry needs sit in one
cache line.  Every launch takes all requests that have arrived and fit the HBM budget
(their inputs and outputs), up to `max_group`, and runs them as one
gsa_batch_spmm_tables launch on one of `slots` streams; requests that arrive
while launches are on the device wait for the next free slot, so work from
successive groups overlaps on the GPU.  Each (row, column) output is computed
by one lane in a fixed order whatever the group, so a request's result is the
same alone and in any mix (checked by the tests).

PlainServer is the comparison point: one drop-in mul! per request, in arrival
order, on one stream.
"""
from __future__ import annotations

import ctypes
import gc
import time
from collections import deque
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L
from ._lib import ArgumentError
from .matrix import GenericSparseMatrixCSR, _PlanCache, itype_of, stream_of
from .precision import BF16CSR, mul_bf16
from .ops import mul_

DESC_DTYPE = np.dtype([("rowptr", "u8"), ("colval", "u8"), ("nzval", "u8"), ("row_scale", "u8"),
                       ("B", "u8"), ("C", "u8"), ("m", "i4"), ("n", "i4"), ("k", "i4"),
                       ("ldb", "i4"), ("ldc", "i4"), ("flags", "i4")])
assert DESC_DTYPE.itemsize == 72  # sizeof(gsa_batch_desc)


def _lpr(k: int) -> int:
    p = 1
    while p < k and p < 64:
        p <<= 1
    return p


class MatrixRegistry:
    """Matrices kept resident in HBM for serving, with their per-k row tiles."""

    def __init__(self, device="cuda", rows_per_tile: int = 512, max_k: int = 16,
                 big_rows: int = 100_000):
        self.device = torch.device(device)
        self.rows_per_tile = rows_per_tile
        self.max_k = max_k
        self.big_rows = big_rows   # matrices with at least this many rows use the dedicated kernels
        self.big = []              # per matrix: its contiguous input column 0, or None
        self._plans = {}           # matrix -> plan of a big matrix (read-only, shared by slots)
        self.mats = []
        self.inputs = []     # per matrix: n x max_k column-major Float32 input block
        self.vtype = None
        self.itype = None
        self._desc = []
        self._tiles = {}     # (mid, lpr) -> int32 [T, 2]

    def register(self, A, x_block: torch.Tensor | None = None) -> int:
        """Keep A resident; x_block is its row-major n x max_k Float32 input block
        (a request with k right-hand sides reads its first k columns)."""
        vt = L.GSA_BF16 if isinstance(A, BF16CSR) else (
            L.GSA_F32 if A.nzval.dtype == torch.float32 else None)
        if vt is None:
            raise ArgumentError("served matrices hold Float32 or BF16 values")
        it = itype_of(A.rowptr)
        if self.vtype is None:
            self.vtype, self.itype = vt, it
        elif (vt, it) != (self.vtype, self.itype):
            raise ArgumentError("one registry serves one value/index type")
        m, n = A.shape
        if x_block is None:
            x_block = torch.zeros((n, self.max_k), dtype=torch.float32, device=self.device)
        if (tuple(x_block.shape) != (n, self.max_k) or not x_block.is_contiguous()
                or x_block.dtype != torch.float32):
            raise ArgumentError("x_block must be a contiguous row-major n x max_k Float32 block")
        mid = len(self.mats)
        self.mats.append(A)
        self.inputs.append(x_block)
        if m >= self.big_rows:
            self.big.append(x_block[:, 0].contiguous())
        else:
            self.big.append(None)
        d = np.zeros(1, DESC_DTYPE)
        d["rowptr"], d["colval"], d["nzval"] = (A.rowptr.data_ptr(), A.colval.data_ptr(),
                                                A.nzval.data_ptr())
        d["row_scale"] = A.row_scale.data_ptr() if isinstance(A, BF16CSR) else 0
        d["B"] = x_block.data_ptr()
        d["m"], d["n"], d["ldb"], d["ldc"] = m, n, self.max_k, 1
        d["flags"] = L.GSA_BATCH_B_ROWMAJOR | L.GSA_BATCH_C_ROWMAJOR
        self._desc.append(d[0])
        return mid

    def plan(self, mid: int):
        """Plan handle of big matrix `mid`, or None for an empty matrix.  A plan
        is read-only once built (each call takes its partial sums from scratch
        on its own stream), so one plan serves every slot, overlapping or not."""
        A = self.mats[mid]
        if A.nnz() == 0:
            return None
        pc = self._plans.get(mid)
        if pc is None:
            pc = _PlanCache()
            self._plans[mid] = pc
        return pc.get(A.rowptr, A.shape[0], A.nnz(), A.colval, A.n)

    def desc_table(self) -> np.ndarray:
        return np.array(self._desc, dtype=DESC_DTYPE)

    def tiles(self, mid: int, k: int) -> np.ndarray:
        key = (mid, _lpr(k))
        t = self._tiles.get(key)
        if t is None:
            m = self.mats[mid].shape[0]
            rows = max(16, min(1024, self.rows_per_tile // key[1]))
            r0 = np.arange(0, m, rows, dtype=np.int64)
            t = np.stack([r0, np.minimum(r0 + rows, m)], axis=1).astype(np.int32)
            self._tiles[key] = t
        return t


@dataclass
class Requests:
    """A request stream: arrival times (s), matrix ids and k per request."""
    arrival: np.ndarray
    mid: np.ndarray
    k: np.ndarray

    def __len__(self):
        return int(self.arrival.size)

    @staticmethod
    def poisson(n: int, rate: float, n_mats: int, ks=(1, 4, 16), seed: int = 5) -> "Requests":
        rng = np.random.default_rng(seed)
        arr = np.cumsum(rng.exponential(1.0 / rate, size=n))
        return Requests(arr, rng.integers(0, n_mats, size=n), rng.choice(np.array(ks), size=n))


class _Slot:
    def __init__(self, device, max_group, max_tiles, arena_bytes):
        self.stream = torch.cuda.Stream(device=device)
        # outputs of one launch are carved from this arena: no allocation in
        # the serving loop (an allocation can synchronise the device)
        self.arena = torch.empty(max(arena_bytes // 4, 1), dtype=torch.float32, device=device)
        self.h_desc = torch.empty(max_group * DESC_DTYPE.itemsize, dtype=torch.uint8, pin_memory=True)
        self.h_tiles = torch.empty(max_tiles * 3, dtype=torch.int32, pin_memory=True)
        self.d_desc = torch.empty(max_group * DESC_DTYPE.itemsize, dtype=torch.uint8, device=device)
        self.d_tiles = torch.empty(max_tiles * 3, dtype=torch.int32, device=device)
        self.d_counter = torch.zeros(1, dtype=torch.int32, device=device)
        self.done = torch.cuda.Event(enable_timing=True)   # reused for every launch
        self.index = 0
        self.event = None
        self.req = None
        self.bytes = 0
        self.out = None


class Server:
    def __init__(self, registry: MatrixRegistry, hbm_budget: int = 2 << 30, slots: int = 4,
                 max_group: int = 4096, max_tiles: int = 1 << 21, grid: int = 0):
        """hbm_budget: bytes of result buffers in flight, split evenly over the
        slots (each launch's results must fit its slot's share)."""
        self.reg = registry
        self.budget = int(hbm_budget)
        self.max_group = max_group
        self.max_tiles = max_tiles
        self.grid = grid
        dev = registry.device
        self.slot_bytes = self.budget // slots
        self.slots = [_Slot(dev, max_group, max_tiles, self.slot_bytes) for _ in range(slots)]
        self.desc = registry.desc_table()
        m = self.desc["m"].astype(np.int64)
        n = self.desc["n"].astype(np.int64)
        self._m, self._n = m, n
        self.in_ptr = self.desc["B"].copy()
        self.is_big = np.array([b is not None for b in registry.big], dtype=bool)
        # every big matrix run once on every slot's stream, so its plan exists and
        # each stream's scratch cache holds the blocks its calls take (no
        # allocation while serving)
        for si, slot in enumerate(self.slots):
            slot.index = si
            for mid in np.flatnonzero(self.is_big).tolist():
                self._big_call(slot, int(mid), 1, slot.arena.data_ptr())
        torch.cuda.synchronize(dev)

    def request_bytes(self, mids, ks):
        """HBM bytes a request holds while in flight: its m x k Float32 result."""
        return self._m[mids] * ks * 4

    def _big_call(self, slot: _Slot, mid: int, k: int, cptr: int):
        """One big-matrix request (k right-hand sides, row-major m x k result at
        device address cptr: each row's k outputs in one segment) through the
        dedicated SpMV/SpMM kernels on the slot's stream."""
        reg = self.reg
        A = reg.mats[mid]
        m = A.shape[0]
        if m > slot.arena.numel():
            raise ArgumentError("a request's result is larger than a slot's share of the budget")
        plan = reg.plan(mid)
        st = ctypes.c_void_p(slot.stream.cuda_stream)
        if k == 1:
            bptr, ldb, opB = reg.big[mid].data_ptr(), max(A.n, 1), L.GSA_OP_N
        else:
            bptr, ldb, opB = reg.inputs[mid].data_ptr(), reg.max_k, L.GSA_OP_T
        c = ctypes.c_void_p(cptr)
        if isinstance(A, BF16CSR):
            L.call("gsa_spmm_csr_bf16", st, reg.itype, m, A.n, k,
                   ctypes.c_void_p(A.rowptr.data_ptr()), ctypes.c_void_p(A.colval.data_ptr()),
                   ctypes.c_void_p(A.nzval.data_ptr()), ctypes.c_void_p(A.row_scale.data_ptr()),
                   A.nnz(), ctypes.c_void_p(bptr), ldb, opB, c, max(k, 1), L.GSA_OP_T, 1.0, 0.0,
                   plan)
        else:
            L.call("gsa_spmm_csr_ex", st, L.GSA_F32, reg.itype, m, A.n, k,
                   ctypes.c_void_p(A.rowptr.data_ptr()), ctypes.c_void_p(A.colval.data_ptr()),
                   ctypes.c_void_p(A.nzval.data_ptr()), A.nnz(), ctypes.c_void_p(bptr), ldb, opB,
                   c, max(k, 1), L.GSA_OP_T, 1.0, 0.0, plan)

    def _launch(self, slot: _Slot, idx: np.ndarray, mids, ks, out_list):
        """One group: the small-matrix requests in one batched launch, then each
        big-matrix request through the dedicated SpMV/SpMM kernels, all on the
        slot's stream, followed by the slot's completion event."""
        reg = self.reg
        ms = self._m[mids]
        sizes = ms * ks
        offs = np.concatenate([[0], np.cumsum(sizes)])
        if offs[-1] > slot.arena.numel():
            raise ArgumentError("a request's result is larger than a slot's share of the budget")
        out = slot.arena
        isbig = self.is_big[mids]
        sm = np.flatnonzero(~isbig)
        st = ctypes.c_void_p(slot.stream.cuda_stream)
        with torch.cuda.stream(slot.stream):
            if sm.size:
                smids, sks, sms = mids[sm], ks[sm], ms[sm]
                d = self.desc[smids].copy()
                d["B"] = self.in_ptr[smids]
                d["C"] = np.uint64(out.data_ptr()) + (offs[sm] * 4).astype(np.uint64)
                d["k"] = sks
                d["ldc"] = np.maximum(sks, 1)
                # tiles of every request, vectorised: request j has T_j tiles of
                # rows_j = max(16, min(1024, rows_per_tile / lpr(k_j))) rows
                lpr = np.ones_like(sks)
                while True:
                    grow = (lpr < sks) & (lpr < 64)
                    if not grow.any():
                        break
                    lpr[grow] <<= 1
                rows = np.clip(reg.rows_per_tile // lpr, 16, 1024)
                T = (sms + rows - 1) // rows
                ntiles = int(T.sum())
                if ntiles > self.max_tiles:
                    raise ArgumentError("group exceeds the tile table; raise max_tiles")
                inst = np.repeat(np.arange(sm.size, dtype=np.int64), T)
                first = np.repeat(np.cumsum(T) - T, T)
                r0 = (np.arange(ntiles, dtype=np.int64) - first) * rows[inst]
                flat = np.empty((ntiles, 3), dtype=np.int32)
                flat[:, 0] = inst
                flat[:, 1] = r0
                flat[:, 2] = np.minimum(r0 + rows[inst], sms[inst])
                nb = d.nbytes
                slot.h_desc[:nb].numpy()[:] = d.view(np.uint8)
                slot.h_tiles[:ntiles * 3].numpy()[:] = flat.ravel()
                slot.d_desc[:nb].copy_(slot.h_desc[:nb], non_blocking=True)
                slot.d_tiles[:ntiles * 3].copy_(slot.h_tiles[:ntiles * 3], non_blocking=True)
                L.call("gsa_batch_spmm_tables", st, reg.vtype, reg.itype,
                       ctypes.c_void_p(slot.d_desc.data_ptr()),
                       ctypes.c_void_p(slot.d_tiles.data_ptr()), ntiles,
                       ctypes.c_void_p(slot.d_counter.data_ptr()), 1.0, 0.0, self.grid)
            for j in np.flatnonzero(isbig).tolist():
                self._big_call(slot, int(mids[j]), int(ks[j]), out.data_ptr() + int(offs[j]) * 4)
            slot.done.record(slot.stream)
        slot.event, slot.req, slot.out = slot.done, idx, out
        slot.used_words = int(offs[-1])
        if out_list is not None:
            for j, i in enumerate(idx.tolist()):
                out_list[i] = (out, int(offs[j]), int(ms[j]), int(ks[j]), "row")

    def run(self, req: Requests, keep_outputs: bool = False, spin_sleep: float = 0.0):
        """Serve `req` in real time; returns per-request latency (s) and summary stats.

        Latency of a request = completion of its launch on the GPU clock (an
        event recorded after the launch) minus its arrival time."""
        N = len(req)
        st = {"done_ms": np.full(N, np.nan), "out": [None] * N if keep_outputs else None,
              "launches": 0, "groups": []}
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        start.record(self.slots[0].stream)
        gc_was = gc.isenabled()
        gc.disable()  # no collector pauses inside the real-time loop
        try:
            self._loop(req, start, st, keep_outputs, spin_sleep)
        finally:
            if gc_was:
                gc.enable()
        torch.cuda.synchronize()
        lat = st["done_ms"] / 1e3 - req.arrival
        span = np.nanmax(st["done_ms"]) / 1e3
        return {"latency_s": lat, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3), "req_per_s": N / span,
                "launches": st["launches"], "mean_group": float(np.mean(st["groups"])),
                "outputs": st["out"]}

    def _loop(self, req: Requests, start, st, keep_outputs, spin_sleep):
        N = len(req)
        done_ms, out_list = st["done_ms"], st["out"]
        rbytes = self.request_bytes(req.mid, req.k)
        pend = deque()
        free = list(range(len(self.slots)))
        busy = set()
        nxt = ndone = 0
        t0 = time.perf_counter()
        while ndone < N:
            now = time.perf_counter() - t0
            while nxt < N and req.arrival[nxt] <= now:
                pend.append(nxt)
                nxt += 1
            progressed = False
            for si in list(busy):
                s = self.slots[si]
                if s.event.query():
                    done_ms[s.req] = start.elapsed_time(s.event)
                    ndone += s.req.size
                    if keep_outputs:  # the arena is reused: keep a copy of the results,
                        # taken on the slot's stream so the next launch there waits for it
                        with torch.cuda.stream(s.stream):
                            copy = s.arena[:s.used_words].clone()
                        for i in s.req.tolist():
                            out_list[i] = (copy,) + out_list[i][1:]
                    s.out = None
                    busy.discard(si)
                    free.append(si)
                    progressed = True
            while pend and free:
                take = []
                b = 0
                while pend and len(take) < self.max_group:
                    i = pend[0]
                    if b + rbytes[i] > self.slot_bytes:
                        break  # admission by HBM bytes (the slot's share of the budget)
                    take.append(pend.popleft())
                    b += int(rbytes[i])
                if not take:
                    raise ArgumentError("a request's result is larger than a slot's share "
                                        "of the HBM budget")
                si = free.pop()
                idx = np.array(take, dtype=np.int64)
                self._launch(self.slots[si], idx, req.mid[idx], req.k[idx], out_list)
                busy.add(si)
                st["launches"] += 1
                st["groups"].append(len(take))
                progressed = True
            if not progressed and spin_sleep:
                time.sleep(spin_sleep)


class PlainServer:
    """One drop-in mul! per request, in arrival order, on the current stream.
    It reads the same row-major input blocks (as Julia's transpose(Xt), which
    mul! takes without a copy) and writes column-major results."""

    def __init__(self, registry: MatrixRegistry):
        self.reg = registry

    def run(self, req: Requests, keep_outputs: bool = False):
        reg = self.reg
        N = len(req)
        outs = [None] * N if keep_outputs else None
        evs = [None] * N
        x0 = [X[:, 0].contiguous() for X in reg.inputs]  # k == 1: a contiguous vector
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream(reg.device)
        start = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        gc_was = gc.isenabled()
        gc.disable()
        t0 = time.perf_counter()
        for i in range(N):
            while time.perf_counter() - t0 < req.arrival[i]:
                pass
            mid, k = int(req.mid[i]), int(req.k[i])
            A = reg.mats[mid]
            X = reg.inputs[mid][:, :k] if k > 1 else x0[mid]
            m = A.shape[0]
            C = torch.empty((k, m), dtype=torch.float32, device=reg.device).t() if k > 1 else \
                torch.empty(m, dtype=torch.float32, device=reg.device)
            if isinstance(A, BF16CSR):
                mul_bf16(C, A, X, 1.0, 0.0)
            else:
                mul_(C, A, X, 1.0, 0.0)
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(stream)
            evs[i] = ev
            if keep_outputs:
                outs[i] = C
        if gc_was:
            gc.enable()
        torch.cuda.synchronize()
        done_ms = np.array([start.elapsed_time(e) for e in evs])
        lat = done_ms / 1e3 - req.arrival
        return {"latency_s": lat, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3),
                "req_per_s": N / (done_ms.max() / 1e3), "outputs": outs}


def output_view(entry):
    """The (m x k) result of a served request (row-major, from the batched
    launch and from the dedicated kernels alike)."""
    out, off, m, k, layout = entry
    seg = out[off:off + m * k]
    return seg.view(m, k) if layout == "row" else seg.view(k, m).t()
