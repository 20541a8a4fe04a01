"""Serving loop: resident matrices, admission by HBM bytes, continuous batching.

    reg = MatrixRegistry(device)
    mid = reg.register(A)                    # CSR (Float32) or BF16CSR, stays in HBM
    srv = Server(reg, hbm_budget=2 << 30, slots=4)
    stats = srv.run(requests)                # requests: arrays of arrival time, matrix, k

Every launch takes all requests that have arrived and fit the HBM budget
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
import time
from collections import deque
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L
from ._lib import ArgumentError
from .matrix import GenericSparseMatrixCSR, itype_of, stream_of
from .precision import BF16CSR, mul_bf16
from .ops import mul_

DESC_DTYPE = np.dtype([("rowptr", "u8"), ("colval", "u8"), ("nzval", "u8"), ("row_scale", "u8"),
                       ("B", "u8"), ("C", "u8"), ("m", "i4"), ("n", "i4"), ("k", "i4"),
                       ("ldb", "i4"), ("ldc", "i4"), ("pad", "i4")])
assert DESC_DTYPE.itemsize == 72  # sizeof(gsa_batch_desc)


def _lpr(k: int) -> int:
    p = 1
    while p < k and p < 64:
        p <<= 1
    return p


class MatrixRegistry:
    """Matrices kept resident in HBM for serving, with their per-k row tiles."""

    def __init__(self, device="cuda", rows_per_tile: int = 512, max_k: int = 16):
        self.device = torch.device(device)
        self.rows_per_tile = rows_per_tile
        self.max_k = max_k
        self.mats = []
        self.inputs = []     # per matrix: n x max_k column-major Float32 input block
        self.vtype = None
        self.itype = None
        self._desc = []
        self._tiles = {}     # (mid, lpr) -> int32 [T, 2]

    def register(self, A, x_block: torch.Tensor | None = None) -> int:
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
            x_block = torch.zeros((self.max_k, n), dtype=torch.float32, device=self.device).t()
        if (tuple(x_block.shape) != (n, self.max_k) or x_block.stride(0) != 1
                or (self.max_k > 1 and x_block.stride(1) != max(n, 1))
                or x_block.dtype != torch.float32):
            raise ArgumentError("x_block must be a dense column-major n x max_k Float32 block")
        mid = len(self.mats)
        self.mats.append(A)
        self.inputs.append(x_block)
        d = np.zeros(1, DESC_DTYPE)
        d["rowptr"], d["colval"], d["nzval"] = (A.rowptr.data_ptr(), A.colval.data_ptr(),
                                                A.nzval.data_ptr())
        d["row_scale"] = A.row_scale.data_ptr() if isinstance(A, BF16CSR) else 0
        d["B"] = x_block.data_ptr()
        d["m"], d["n"], d["ldb"], d["ldc"] = m, n, max(n, 1), max(m, 1)
        self._desc.append(d[0])
        return mid

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
    def __init__(self, device, max_group, max_tiles):
        self.stream = torch.cuda.Stream(device=device)
        self.h_desc = torch.empty(max_group * DESC_DTYPE.itemsize, dtype=torch.uint8, pin_memory=True)
        self.h_tiles = torch.empty(max_tiles * 3, dtype=torch.int32, pin_memory=True)
        self.d_desc = torch.empty(max_group * DESC_DTYPE.itemsize, dtype=torch.uint8, device=device)
        self.d_tiles = torch.empty(max_tiles * 3, dtype=torch.int32, device=device)
        self.d_counter = torch.zeros(1, dtype=torch.int32, device=device)
        self.event = None
        self.req = None
        self.bytes = 0
        self.out = None


class Server:
    def __init__(self, registry: MatrixRegistry, hbm_budget: int = 2 << 30, slots: int = 4,
                 max_group: int = 2048, max_tiles: int = 1 << 20, grid: int = 0):
        self.reg = registry
        self.budget = int(hbm_budget)
        self.max_group = max_group
        self.max_tiles = max_tiles
        self.grid = grid
        dev = registry.device
        self.slots = [_Slot(dev, max_group, max_tiles) for _ in range(slots)]
        self.desc = registry.desc_table()
        m = self.desc["m"].astype(np.int64)
        n = self.desc["n"].astype(np.int64)
        self._m, self._n = m, n
        self.in_ptr = self.desc["B"].copy()

    def request_bytes(self, mids, ks):
        return (self._m[mids] + self._n[mids]) * ks * 4

    def _launch(self, slot: _Slot, idx: np.ndarray, mids, ks, out_list):
        reg = self.reg
        ms = self._m[mids]
        sizes = ms * ks
        offs = np.concatenate([[0], np.cumsum(sizes)])
        out = torch.empty(int(offs[-1]) if offs[-1] > 0 else 1, dtype=torch.float32,
                          device=reg.device)
        d = self.desc[mids].copy()
        d["B"] = self.in_ptr[mids]
        d["C"] = np.uint64(out.data_ptr()) + (offs[:-1] * 4).astype(np.uint64)
        d["k"] = ks
        tl, inst = [], []
        for j, (mid, k) in enumerate(zip(mids.tolist(), ks.tolist())):
            t = reg.tiles(mid, k)
            tl.append(t)
            inst.append(np.full(t.shape[0], j, dtype=np.int32))
        tiles = np.concatenate(tl)
        ntiles = tiles.shape[0]
        if ntiles > self.max_tiles:
            raise ArgumentError("group exceeds the tile table; raise max_tiles")
        flat = np.empty((ntiles, 3), dtype=np.int32)
        flat[:, 0] = np.concatenate(inst)
        flat[:, 1:] = tiles
        nb = d.nbytes
        slot.h_desc[:nb].numpy()[:] = d.view(np.uint8)
        slot.h_tiles[:ntiles * 3].numpy()[:] = flat.ravel()
        with torch.cuda.stream(slot.stream):
            slot.d_desc[:nb].copy_(slot.h_desc[:nb], non_blocking=True)
            slot.d_tiles[:ntiles * 3].copy_(slot.h_tiles[:ntiles * 3], non_blocking=True)
            L.call("gsa_batch_spmm_tables", ctypes.c_void_p(slot.stream.cuda_stream), reg.vtype,
                   reg.itype, ctypes.c_void_p(slot.d_desc.data_ptr()),
                   ctypes.c_void_p(slot.d_tiles.data_ptr()), ntiles,
                   ctypes.c_void_p(slot.d_counter.data_ptr()), 1.0, 0.0, self.grid)
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(slot.stream)
        slot.event, slot.req, slot.out = ev, idx, out
        if out_list is not None:
            for j, i in enumerate(idx.tolist()):
                out_list[i] = (out, int(offs[j]), int(ms[j]), int(ks[j]))

    def run(self, req: Requests, keep_outputs: bool = False, spin_sleep: float = 0.0):
        """Serve `req` in real time; returns per-request latency (s) and summary stats."""
        N = len(req)
        done_ms = np.full(N, np.nan)
        out_list = [None] * N if keep_outputs else None
        pend = deque()
        free = list(range(len(self.slots)))
        busy = {}
        used = 0
        rbytes = self.request_bytes(req.mid, req.k)
        nxt = 0
        ndone = 0
        launches = 0
        group_sizes = []
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        start.record(self.slots[0].stream)
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
                    ms = start.elapsed_time(s.event)
                    done_ms[s.req] = ms
                    ndone += s.req.size
                    used -= s.bytes
                    if not keep_outputs:
                        s.out = None
                    del busy[si]
                    free.append(si)
                    progressed = True
            while pend and free:
                take = []
                b = 0
                while pend and len(take) < self.max_group:
                    i = pend[0]
                    if used + b + rbytes[i] > self.budget and (take or used > 0):
                        break  # admission by HBM bytes
                    take.append(pend.popleft())
                    b += int(rbytes[i])
                if not take:
                    break
                si = free.pop()
                s = self.slots[si]
                idx = np.array(take, dtype=np.int64)
                s.bytes = b
                used += b
                self._launch(s, idx, req.mid[idx], req.k[idx], out_list)
                busy[si] = True
                launches += 1
                group_sizes.append(len(take))
                progressed = True
            if not progressed:
                if spin_sleep:
                    time.sleep(spin_sleep)
        torch.cuda.synchronize()
        lat = done_ms / 1e3 - req.arrival
        span = np.nanmax(done_ms) / 1e3 - 0.0
        return {"latency_s": lat, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3), "req_per_s": N / span,
                "launches": launches, "mean_group": float(np.mean(group_sizes)),
                "outputs": out_list}


class PlainServer:
    """One drop-in mul! per request, in arrival order, on the current stream."""

    def __init__(self, registry: MatrixRegistry):
        self.reg = registry

    def run(self, req: Requests, keep_outputs: bool = False):
        reg = self.reg
        N = len(req)
        outs = [None] * N if keep_outputs else None
        evs = [None] * N
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream(reg.device)
        start = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        t0 = time.perf_counter()
        for i in range(N):
            while time.perf_counter() - t0 < req.arrival[i]:
                pass
            mid, k = int(req.mid[i]), int(req.k[i])
            A = reg.mats[mid]
            X = reg.inputs[mid][:, :k] if k > 1 else reg.inputs[mid][:, 0]
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
        torch.cuda.synchronize()
        done_ms = np.array([start.elapsed_time(e) for e in evs])
        lat = done_ms / 1e3 - req.arrival
        return {"latency_s": lat, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3),
                "req_per_s": N / (done_ms.max() / 1e3), "outputs": outs}


def output_view(entry):
    """The (m x k) column-major result of a served request."""
    out, off, m, k = entry
    return out[off:off + m * k].view(k, m).t()
