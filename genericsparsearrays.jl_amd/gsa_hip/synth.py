"""Seeded synthetic workloads (host side, numpy).  All arrays are ONE-BASED int64.

Workloads, by the shapes the benchmarks and parity tests use:
  * sprand-like uniform matrices (10k x 10k at density 1e-3, values U[0,1));
  * 1M x 1M power-law CSR: row lengths from a Zipf law truncated to [1, 5000]
    (alpha = 1.846 gives mean 10.00; alpha = 2.0 gives mean 5.53), distinct
    uniform columns, N(0,1) values;
  * the 2-D 5-point Laplacian on a g x g grid (values 4 / -1);
  * random matrices with a fixed number of distinct columns per row;
  * uneven batches: rows log-uniform in [64, 65536], Poisson(8)+1 entries per row.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class HostCSR:
    m: int
    n: int
    rowptr: np.ndarray   # int64, one-based, m+1
    colval: np.ndarray   # int64, one-based
    nzval: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.colval.size)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.nzval, self.colval - 1, self.rowptr - 1), shape=(self.m, self.n))

    @staticmethod
    def from_scipy(A) -> "HostCSR":
        B = A.tocsr(copy=True)
        B.sum_duplicates()
        B.sort_indices()
        return HostCSR(B.shape[0], B.shape[1], B.indptr.astype(np.int64) + 1,
                       B.indices.astype(np.int64) + 1, np.ascontiguousarray(B.data))


def _distinct_sorted_columns(lengths: np.ndarray, n: int, rng: np.random.Generator) -> np.ndarray:
    """For each row, `lengths[r]` distinct uniform columns in [0, n), sorted; rows
    concatenated.  Duplicates are redrawn until none remain."""
    m = lengths.size
    lengths = np.minimum(lengths, n)
    rows = np.repeat(np.arange(m, dtype=np.int64), lengths)
    cols = rng.integers(0, n, size=rows.size, dtype=np.int64)
    while True:
        key = rows * n + cols
        order = np.argsort(key, kind="stable")
        key = key[order]
        rows, cols = rows[order], cols[order]
        dup = np.zeros(key.size, dtype=bool)
        dup[1:] = key[1:] == key[:-1]
        nd = int(dup.sum())
        if nd == 0:
            return cols
        cols[dup] = rng.integers(0, n, size=nd, dtype=np.int64)


def csr_from_lengths(lengths, n, rng, values="normal", dtype=np.float64) -> HostCSR:
    lengths = np.asarray(lengths, dtype=np.int64)
    m = lengths.size
    lengths = np.minimum(lengths, n)
    cols = _distinct_sorted_columns(lengths, n, rng)
    rowptr = np.empty(m + 1, dtype=np.int64)
    rowptr[0] = 1
    np.cumsum(lengths, out=rowptr[1:])
    rowptr[1:] += 1
    nnz = cols.size
    if values == "normal":
        vals = rng.standard_normal(nnz)
    else:
        vals = rng.random(nnz)
    return HostCSR(m, n, rowptr, cols + 1, vals.astype(dtype))


def sprand(m: int, n: int, density: float, seed: int, values="uniform",
           dtype=np.float64) -> HostCSR:
    """sprand-like: every entry present independently with probability `density`
    (row counts Binomial(n, density)), values U[0,1) (or N(0,1))."""
    rng = np.random.default_rng(seed)
    lengths = rng.binomial(n, density, size=m).astype(np.int64)
    return csr_from_lengths(lengths, n, rng, values=values, dtype=dtype)


def zipf_lengths(m: int, alpha: float, lo: int, hi: int, rng) -> np.ndarray:
    k = np.arange(lo, hi + 1, dtype=np.float64)
    pmf = k ** (-alpha)
    cdf = np.cumsum(pmf)
    cdf /= cdf[-1]
    u = rng.random(m)
    return (np.searchsorted(cdf, u, side="right") + lo).astype(np.int64)


def zipf_mean(alpha: float, lo: int = 1, hi: int = 5000) -> float:
    k = np.arange(lo, hi + 1, dtype=np.float64)
    p = k ** (-alpha)
    return float((k * p).sum() / p.sum())


def powerlaw_csr(m: int, n: int, alpha: float = 1.846, seed: int = 2, lo: int = 1,
                 hi: int = 5000, dtype=np.float64) -> HostCSR:
    """Power-law CSR: row lengths ~ Zipf(alpha) truncated to [lo, hi], distinct
    uniform columns, N(0,1) values."""
    rng = np.random.default_rng(seed)
    lengths = zipf_lengths(m, alpha, lo, hi, rng)
    return csr_from_lengths(lengths, n, rng, values="normal", dtype=dtype)


def laplacian2d(g: int, dtype=np.float64) -> HostCSR:
    """5-point Laplacian on a g x g grid, natural order, diagonal 4, neighbours -1."""
    import scipy.sparse as sp
    T = sp.diags([-np.ones(g - 1), 2 * np.ones(g), -np.ones(g - 1)], [-1, 0, 1], format="csr")
    I = sp.identity(g, format="csr")
    Lm = (sp.kron(I, T) + sp.kron(T, I)).tocsr()
    Lm.sort_indices()
    return HostCSR.from_scipy(Lm.astype(dtype))


def fixed_row_csr(m: int, n: int, per_row: int, seed: int, dtype=np.float64) -> HostCSR:
    rng = np.random.default_rng(seed)
    return csr_from_lengths(np.full(m, per_row, dtype=np.int64), n, rng, "normal", dtype)


def drop_entries(A: HostCSR, p: float, seed: int) -> HostCSR:
    """Copy of A with every stored entry dropped independently with probability p."""
    rng = np.random.default_rng(seed)
    keep = rng.random(A.nnz) >= p
    rows = np.repeat(np.arange(A.m), np.diff(A.rowptr))
    kept_rows = rows[keep]
    counts = np.bincount(kept_rows, minlength=A.m)
    rowptr = np.empty(A.m + 1, dtype=np.int64)
    rowptr[0] = 1
    np.cumsum(counts, out=rowptr[1:])
    rowptr[1:] += 1
    return HostCSR(A.m, A.n, rowptr, A.colval[keep].copy(), A.nzval[keep].copy())


def uneven_batch(count: int, seed: int, lo: int = 64, hi: int = 65536, mean_extra: float = 8.0,
                 ks=(1, 2, 4, 8, 16), dtype=np.float32):
    """`count` square CSR matrices: rows log-uniform in [lo, hi], nnz/row =
    min(n, Poisson(8)+1), N(0,1) values; each with k RHS drawn from `ks`."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(np.int64)
    mats, kk = [], []
    for s in sizes:
        lengths = rng.poisson(mean_extra, size=int(s)).astype(np.int64) + 1
        mats.append(csr_from_lengths(lengths, int(s), rng, "normal", dtype))
        kk.append(int(rng.choice(ks)))
    return mats, kk


def uneven_batch_device(count: int, seed: int, device="cuda", lo: int = 64, hi: int = 65536,
                        mean_extra: float = 8.0, ks=(1, 2, 4, 8, 16), itype=None, vdtype=None):
    """The uneven batch built directly in device memory (for batches too large to
    build on the host quickly): same size/row-length/k laws as `uneven_batch`,
    distinct uniform columns per row, N(0,1) values.  Returns a list of
    (m, rowptr, colval, nzval) device tensors and the list of k.  The instances
    share one column/value buffer (views)."""
    import torch
    itype = itype or torch.int32
    vdtype = vdtype or torch.float32
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(np.int64)
    kk = [int(rng.choice(ks)) for _ in range(count)]
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    dev = torch.device(device)
    total_rows = int(sizes.sum())
    row_n = torch.as_tensor(np.repeat(sizes, sizes), device=dev)          # n of each row's matrix
    lens = torch.poisson(torch.full((total_rows,), mean_extra, device=dev), generator=g)
    lens = torch.minimum(lens.to(torch.int64) + 1, row_n)
    rows = torch.repeat_interleave(torch.arange(total_rows, device=dev), lens)
    n_of = row_n[rows]
    cols = (torch.rand(rows.numel(), generator=g, device=dev) * n_of).to(torch.int64)
    cols = torch.minimum(cols, n_of - 1)
    while True:  # distinct columns per row: sort (row, col), redraw repeats
        key = rows * 65536 + cols
        key, order = torch.sort(key)
        cols = cols[order]
        dup = torch.zeros_like(key, dtype=torch.bool)
        dup[1:] = key[1:] == key[:-1]
        nd = int(dup.sum())
        if nd == 0:
            break
        redraw = (torch.rand(nd, generator=g, device=dev) * n_of[dup]).to(torch.int64)
        cols[dup] = torch.minimum(redraw, n_of[dup] - 1)
    vals = torch.randn(rows.numel(), generator=g, device=dev).to(vdtype)
    cols1 = (cols + 1).to(itype)
    gptr = torch.zeros(total_rows + 1, dtype=torch.int64, device=dev)
    gptr[1:] = torch.cumsum(lens, 0)
    row_start = np.concatenate([[0], np.cumsum(sizes)])
    out = []
    starts = gptr[torch.as_tensor(row_start, device=dev)].cpu().numpy()
    for i in range(count):
        r0, r1 = int(row_start[i]), int(row_start[i + 1])
        p0, p1 = int(starts[i]), int(starts[i + 1])
        rp = (gptr[r0:r1 + 1] - p0 + 1).to(itype)
        out.append((int(sizes[i]), rp, cols1[p0:p1], vals[p0:p1]))
    return out, kk
