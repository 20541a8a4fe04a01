"""Autotuner for the batched launch: picks the row-tile size (and launch grid)
of gsa_batch_* by timing candidates on the device, keyed on the shape of the
batch so a tuned choice is reused for batches that look alike.

Key: (log2 of total rows, the nnz-per-row histogram over the buckets
[1,4], (4,16], (16,64], (64,256], >256 as tenths of the rows, the histogram of
k over {1, 2, 4, 8, 16, >16} as tenths of the instances, value type).
"""
from __future__ import annotations

import json
import math
import os
import threading

import torch

from .batched import BatchedSpMM

ROW_BUCKETS = (4, 16, 64, 256)
K_BUCKETS = (1, 2, 4, 8, 16)


def _hist(vals, edges):
    counts = [0] * (len(edges) + 1)
    for v in vals:
        for i, e in enumerate(edges):
            if v <= e:
                counts[i] += 1
                break
        else:
            counts[-1] += 1
    tot = max(1, sum(counts))
    return tuple(int(round(10 * c / tot)) for c in counts)


def batch_key(items) -> tuple:
    rows = sum(A.shape[0] for A, _, _ in items)
    # nnz/row histogram from the row pointers (one device reduction per instance)
    edges = torch.tensor(ROW_BUCKETS, dtype=torch.int64, device=items[0][2].device)
    counts = torch.zeros(len(ROW_BUCKETS) + 1, dtype=torch.int64, device=edges.device)
    for A, _, _ in items:
        lens = (A.rowptr[1:] - A.rowptr[:-1]).to(torch.int64)
        counts += torch.bincount(torch.bucketize(lens, edges), minlength=len(ROW_BUCKETS) + 1)
    c = counts.cpu().tolist()
    tot = max(1, sum(c))
    row_hist = tuple(int(round(10 * x / tot)) for x in c)
    ks = [1 if B.dim() == 1 else B.shape[1] for _, B, _ in items]
    vt = str(items[0][0].nzval.dtype)
    return (int(math.log2(max(rows, 1))), row_hist, _hist(ks, K_BUCKETS), vt)


class BatchAutotuner:
    def __init__(self, candidates=(128, 256, 512, 1024, 2048, 4096), reps: int = 5,
                 cache_path: str | None = None):
        self.candidates = tuple(candidates)
        self.reps = reps
        self.cache_path = cache_path
        self.cache: dict[str, dict] = {}
        self._lock = threading.Lock()
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.cache = json.load(f)

    def _time(self, batch: BatchedSpMM) -> float:
        batch.run()  # warm
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(self.reps):
            batch.run()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / self.reps

    def tune(self, items) -> dict:
        """{'rows_per_tile': R, 'ms': t, 'timings': {...}} for this batch (cached by key)."""
        key = repr(batch_key(items))
        with self._lock:
            if key in self.cache:
                return self.cache[key]
        timings = {}
        for rpt in self.candidates:
            timings[rpt] = self._time(BatchedSpMM(items, rows_per_tile=rpt))
        best = min(timings, key=timings.get)
        res = {"rows_per_tile": best, "ms": timings[best],
               "timings": {str(k): v for k, v in timings.items()}}
        with self._lock:
            self.cache[key] = res
            if self.cache_path:
                with open(self.cache_path, "w") as f:
                    json.dump(self.cache, f, indent=1)
        return res

    def build(self, items) -> BatchedSpMM:
        return BatchedSpMM(items, rows_per_tile=self.tune(items)["rows_per_tile"])
