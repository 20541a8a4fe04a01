"""Benchmark of the sparse kernel layer on one MI355X.

Headline: Float64 CSR SpMV, y = A x, on the 1,000,000 x 1,000,000 power-law
matrix (row lengths ~ Zipf(1.846) truncated to [1, 5000], mean 10 entries per
row, uniform distinct columns, N(0,1) values, Int64 one-based indices), one
step = one mul! over one copy of A resident in HBM.  Three copies of (A, x, y)
are rotated so the working set (3 x 184 MB) exceeds the 256 MB Infinity Cache.

Prints ONE JSON line (rank 0).  `value` is algorithmic GB/s of the SpMV:
  bytes = nnz*(8+8) + (m+1)*8 + n*8 + m*8   (values, column indices, row
  pointer, x read once, y written once)
divided by the mean time per mul!, measured with HIP events on the stream the
kernels run on.  The CPU baseline is the oracle (a C restatement of the
reference's CPU loop, OpenMP over rows) timed on a bounded sample on the host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "genericsparsearrays.jl_amd"))

HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join(HERE, "profiles", "spmv_f64_i64_powerlaw1M_pmc.md")


def measured_traffic():
    """HBM bytes per launch of the SpMV kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE scaled by a plain copy), or None."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def spmv_bytes(m, n, nnz, sv, si):
    return nnz * (sv + si) + (m + 1) * si + n * sv + m * sv


def spmm_bytes(m, n, nnz, k, sv, si):
    return nnz * (sv + si) + (m + 1) * si + (n + m) * k * sv


def time_steps(fn, steps, warmup, stream):
    """Mean ms per step of `fn(i)` launched eagerly, events on `stream`."""
    import torch
    for i in range(warmup):
        fn(i)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for i in range(steps):
        fn(i)
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps  # ms per step


def time_steps_graph(fn, per_graph, steps, warmup, stream):
    """Mean ms per step over exactly `steps` steps: graphs of `per_graph`
    consecutive steps replayed steps // per_graph times, then one graph of the
    remaining steps, so host launch overhead does not gap the kernels."""
    import torch
    per_graph = max(1, min(per_graph, steps))
    for i in range(per_graph):  # eager warm-up (plans, allocator)
        fn(i)
    torch.cuda.synchronize()

    def capture(count):
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for i in range(count):
                fn(i)
        return g

    g = capture(per_graph)
    reps, rem = divmod(steps, per_graph)
    g_rem = capture(rem) if rem else None
    for _ in range(max(1, warmup // per_graph)):
        g.replay()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    cur = torch.cuda.current_stream()
    start.record(cur)
    for _ in range(reps):
        g.replay()
    if g_rem is not None:
        g_rem.replay()
    stop.record(cur)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps, steps


def latency_samples(fn, reps, stream):
    """Per-call latency (ms) with events around each call."""
    import torch
    out = []
    for i in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn(i)
        b.record(stream)
        b.synchronize()
        out.append(a.elapsed_time(b))
    return out


def cpu_baseline(H, x, seconds=10.0):
    """Oracle CPU SpMV (restated reference loop) timed on the host cores."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O
    threads = O.max_threads()
    O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x, threads=threads)  # warm
    reps, t0 = 0, time.perf_counter()
    while True:
        O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x, threads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    per = el / reps
    gbs = spmv_bytes(H.m, H.n, H.nnz, 8, 8) / per / 1e9
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{reps} full SpMVs of the same 1M-row Float64/Int64 matrix "
                      f"({el:.1f} s, {per * 1e3:.2f} ms each)"}


def _sync_time(fn, reps=3):
    """Median wall time (s) of fn() with device syncs around it (ops that read a
    size back to the host synchronise anyway)."""
    import torch
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def bench_structural(H, dev):
    """Structural ops: CSR<->CSC<->COO on the 1M power-law matrix; L + L', L*L,
    R*R and kron on the Laplacian / random / density-0.01 operands."""
    import torch
    import gsa_hip as G
    from gsa_hip import synth

    def dcsr(M, itype=torch.int64):
        return G.GenericSparseMatrixCSR(M.m, M.n, torch.as_tensor(M.rowptr).to(dev, itype),
                                        torch.as_tensor(M.colval).to(dev, itype),
                                        torch.as_tensor(M.nzval).to(dev), copy=False)
    def reconstruct(M):
        """The reference's constructor on the fresh result's buffers: it copies all
        three (matrix_csr.jl:55-61, matrix_csc.jl, matrix_coo.jl), as a binding
        that returns results through it would."""
        if isinstance(M, G.GenericSparseMatrixCSR):
            return G.GenericSparseMatrixCSR(M.m, M.n, M.rowptr, M.colval, M.nzval)
        if isinstance(M, G.GenericSparseMatrixCSC):
            return G.GenericSparseMatrixCSC(M.m, M.n, M.colptr, M.rowval, M.nzval)
        return G.GenericSparseMatrixCOO(M.m, M.n, M.rowind, M.colind, M.nzval)

    def timed(fn, entry):
        """Kernel-path time of fn() and, beside it, the full drop-in time of fn()
        followed by the constructor copy of its result."""
        t = _sync_time(fn)
        entry["ms"] = round(t * 1e3, 3)
        entry["dropin_with_constructor_copy_ms"] = round(_sync_time(lambda: reconstruct(fn())) * 1e3, 3)
        return t

    out = {}
    A = dcsr(H)
    nnz, m, n = H.nnz, H.m, H.n
    Ac = G.GenericSparseMatrixCSC(A)
    Ao = G.GenericSparseMatrixCOO(A)
    Aoc = G.GenericSparseMatrixCOO(Ac)
    for name, fn, byts in (
            ("csr_to_csc", lambda: G.GenericSparseMatrixCSC(A),
             nnz * 16 + (m + 1) * 8 + nnz * 16 + (n + 1) * 8),
            ("csc_to_csr", lambda: G.GenericSparseMatrixCSR(Ac),
             nnz * 16 + (n + 1) * 8 + nnz * 16 + (m + 1) * 8),
            ("csr_to_coo", lambda: G.GenericSparseMatrixCOO(A),
             nnz * 16 + (m + 1) * 8 + nnz * 24),
            ("coo_to_csr", lambda: G.GenericSparseMatrixCSR(Ao),
             nnz * 24 + nnz * 16 + (m + 1) * 8),
            # the same from column-major storage: keys out of order, the sort runs
            ("coo_colmajor_to_csr", lambda: G.GenericSparseMatrixCSR(Aoc),
             nnz * 24 + nnz * 16 + (m + 1) * 8)):
        out[name] = {}
        t = timed(fn, out[name])
        out[name]["GBs"] = round(byts / t / 1e9, 1)
    # A + transpose(A) on the same matrix (rows of up to 5000 entries): the CSR of
    # transpose(A) is the CSC's arrays
    At = G.GenericSparseMatrixCSR(n, m, Ac.colptr, Ac.rowval, Ac.nzval, copy=False)
    Cat = A + At
    out["add_powerlaw_A_At"] = {"nnzC": Cat.nnz()}
    t = timed(lambda: A + At, out["add_powerlaw_A_At"])
    out["add_powerlaw_A_At"]["GBs"] = round((2 * nnz * 16 + 2 * (m + 1) * 8 + Cat.nnz() * 16
                                             + (m + 1) * 8) / t / 1e9, 1)
    # dropzeros on the same matrix with every tenth stored value set to zero
    # (count, allocation, fill: rows spanning many 2048-entry blocks included)
    Z = A.nzval.clone()
    Z[::10] = 0
    Az = G.GenericSparseMatrixCSR(m, n, A.rowptr, A.colval, Z, copy=False)
    Dz = G.dropzeros(Az)
    out["dropzeros_powerlaw_tenth"] = {"nnzC": Dz.nnz()}
    t = timed(lambda: G.dropzeros(Az), out["dropzeros_powerlaw_tenth"])
    out["dropzeros_powerlaw_tenth"]["GBs"] = round((nnz * 16 + (m + 1) * 8 + Dz.nnz() * 16
                                                    + (m + 1) * 8) / t / 1e9, 1)
    del Cat, At, Ac, Ao, Aoc, A, Z, Az, Dz
    # host ingest: GenericSparseMatrixCSR(::SparseMatrixCSC) of the same matrix held
    # as a scipy CSC (the arrays copied as they are, transposed on the device),
    # against the host transposition the reference performs (scipy's tocsr)
    S = H.to_scipy().tocsc()
    G.convert.csr_from_scipy(S, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    G.convert.csr_from_scipy(S, device=dev)
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    S.tocsr()
    t_host = time.perf_counter() - t0
    out["ingest_csr_from_scipy_csc"] = {"ms": round(t_gpu * 1e3, 1),
                                        "host_transpose_ms": round(t_host * 1e3, 1)}
    # host egress: SparseMatrixCSC(::GenericSparseMatrixCSR) of the same matrix on the
    # device (transposed there, then copied), against the reference's route: the CSR
    # arrays collected as they are, then transposed on the host (scipy's tocsc)
    # (the first call also allocates the pinned host buffers the copies go through;
    # later calls of like sizes reuse them from torch's caching host allocator)
    Ad = G.convert.csr_from_scipy(S.tocsr(), device=dev)
    G.convert.csr_to_csc(Ad)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    G.SparseMatrixCSC(Ad)
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    G.SparseMatrixCSC(Ad)
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    G.convert.to_scipy(Ad).tocsc()
    t_ref = time.perf_counter() - t0
    out["egress_sparsematrixcsc_of_csr"] = {"ms": round(t_gpu * 1e3, 1),
                                            "first_call_ms": round(t_first * 1e3, 1),
                                            "collect_then_host_transpose_ms": round(t_ref * 1e3, 1)}
    del S, Ad
    L = synth.laplacian2d(1000)
    Lp = synth.drop_entries(L, 0.1, seed=3)
    R = synth.fixed_row_csr(200_000, 200_000, 8, seed=3)
    K1 = synth.sprand(1000, 1000, 0.01, seed=31)
    K2 = synth.sprand(1000, 1000, 0.01, seed=32)
    dL, dLp, dR, dK1, dK2 = (dcsr(M) for M in (L, Lp, R, K1, K2))
    C = dL + dLp
    out["add_L_Lp"] = {"nnzC": C.nnz()}
    t = timed(lambda: dL + dLp, out["add_L_Lp"])
    out["add_L_Lp"]["GBs"] = round(((L.nnz + Lp.nnz) * 16 + 2 * (L.m + 1) * 8 + C.nnz() * 16
                                    + (L.m + 1) * 8) / t / 1e9, 1)
    cL, cLp = G.GenericSparseMatrixCOO(dL), G.GenericSparseMatrixCOO(dLp)  # row-major COO
    Cc = cL + cLp
    out["add_coo_L_Lp"] = {"nnzC": Cc.nnz()}
    t = timed(lambda: cL + cLp, out["add_coo_L_Lp"])
    out["add_coo_L_Lp"]["GBs"] = round(((L.nnz + Lp.nnz) * 24 + Cc.nnz() * 24) / t / 1e9, 1)
    # column-major COO (findnz order, as COO(::SparseMatrixCSC) stores it): the merge
    cL = G.GenericSparseMatrixCOO(G.GenericSparseMatrixCSC(dL))
    cLp = G.GenericSparseMatrixCOO(G.GenericSparseMatrixCSC(dLp))
    Cc = cL + cLp
    out["add_coo_colmajor_L_Lp"] = {"nnzC": Cc.nnz()}
    t = timed(lambda: cL + cLp, out["add_coo_colmajor_L_Lp"])
    out["add_coo_colmajor_L_Lp"]["GBs"] = round(((L.nnz + Lp.nnz) * 24 + Cc.nnz() * 24) / t / 1e9,
                                                1)
    del cL, cLp, Cc
    for name, X in (("spgemm_LL", dL), ("spgemm_RR", dR)):
        Cx = X * X
        out[name] = {"nnzC": Cx.nnz()}
        t = timed(lambda: X * X, out[name])
        out[name]["GBs_lower_bound"] = round((2 * X.nnz() + Cx.nnz()) * 16 / t / 1e9, 1)
        del Cx
    Kc = G.kron(dK1, dK2)
    kn = "kron_csr_1000x1000_d0.01"
    out[kn] = {"nnzC": Kc.nnz()}
    t = timed(lambda: G.kron(dK1, dK2), out[kn])
    out[kn]["GBs"] = round((Kc.nnz() * 16 + (Kc.m + 1) * 8) / t / 1e9, 1)
    del Kc
    # sparse + dense at the reference benchmark's shape (matrix_benchmarks.jl:189-229:
    # 10000^2 at density 0.01, Float64): copy(B), then the entries added
    S = synth.sprand(10_000, 10_000, 0.01, seed=33)
    dS = dcsr(S)
    Dd = torch.randn((10_000, 10_000), dtype=torch.float64, device=dev).t()  # column-major
    sd = {"nnz": S.nnz}
    for fname, M in (("csr", dS), ("csc", G.GenericSparseMatrixCSC(dS)),
                     ("coo", G.GenericSparseMatrixCOO(dS))):
        sd[fname + "_ms"] = round(_sync_time(lambda: M + Dd) * 1e3, 3)
    # bytes: the dense operand read and written once, the entries read, their
    # elements read and written again
    sd["csr_GBs"] = round((2 * Dd.numel() * 8 + S.nnz * (16 + 16) + (S.m + 1) * 8)
                          / (sd["csr_ms"] * 1e-3) / 1e9, 1)
    out["sparse_plus_dense_10000_d0.01"] = sd
    del Dd, dS
    torch.cuda.empty_cache()
    return out


def cpu_structural(H):
    """The oracle (the reference's CPU algorithms restated in C, one thread) on the
    same structural workloads; kron on a 1e7-entry sample (the 1000^2 operand
    with a 100 x 100 operand at density 0.1).  Algorithmic bytes as on the GPU."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O
    from gsa_hip import synth

    def tm(fn):
        t0 = time.perf_counter()
        r = fn()
        return r, time.perf_counter() - t0
    out = {}
    nnz, m, n = H.nnz, H.m, H.n
    c, t = tm(lambda: O.csr_to_csc(m, n, H.rowptr, H.colval, H.nzval))
    out["csr_to_csc"] = (t, nnz * 16 + (m + 1) * 8 + nnz * 16 + (n + 1) * 8)
    _, t = tm(lambda: O.csc_to_csr(m, n, c[0], c[1], c[2]))
    out["csc_to_csr"] = (t, nnz * 16 + (n + 1) * 8 + nnz * 16 + (m + 1) * 8)
    C, t = tm(lambda: O.csr_add(m, (H.rowptr, H.colval, H.nzval), c))
    out["add_powerlaw_A_At"] = (t, 2 * nnz * 16 + 2 * (m + 1) * 8 + C[2].size * 16 + (m + 1) * 8)
    del C
    Z = H.nzval.copy()
    Z[::10] = 0
    C, t = tm(lambda: O.csr_dropzeros(m, H.rowptr, H.colval, Z))
    out["dropzeros_powerlaw_tenth"] = (t, nnz * 16 + (m + 1) * 8 + C[2].size * 16 + (m + 1) * 8)
    del C, Z
    co, t = tm(lambda: O.csr_to_coo(m, H.rowptr, H.colval, H.nzval))
    out["csr_to_coo"] = (t, nnz * 16 + (m + 1) * 8 + nnz * 24)
    _, t = tm(lambda: O.coo_to_csr(m, n, co[0], co[1], co[2]))
    out["coo_to_csr"] = (t, nnz * 24 + nnz * 16 + (m + 1) * 8)
    del c, co
    L = synth.laplacian2d(1000)
    Lp = synth.drop_entries(L, 0.1, seed=3)
    R = synth.fixed_row_csr(200_000, 200_000, 8, seed=3)
    tup = lambda M: (M.rowptr, M.colval, M.nzval)
    C, t = tm(lambda: O.csr_add(L.m, tup(L), tup(Lp)))
    out["add_L_Lp"] = (t, (L.nnz + Lp.nnz) * 16 + 2 * (L.m + 1) * 8 + C[2].size * 16 + (L.m + 1) * 8)
    cL = O.csr_to_coo(L.m, *tup(L))
    cLp = O.csr_to_coo(Lp.m, *tup(Lp))
    C, t = tm(lambda: O.coo_add(L.m, cL, cLp))
    out["add_coo_L_Lp"] = (t, (L.nnz + Lp.nnz) * 24 + C[2].size * 24)
    for name, X in (("spgemm_LL", L), ("spgemm_RR", R)):
        C, t = tm(lambda: O.spgemm_csr(X.m, X.n, tup(X), tup(X)))
        out[name] = (t, (2 * X.nnz + C[2].size) * 16)
    K1 = synth.sprand(1000, 1000, 0.01, seed=31)
    Ks = synth.sprand(100, 100, 0.1, seed=32)
    C, t = tm(lambda: O.kron_csr(K1.m, K1.n, Ks.m, Ks.n, tup(K1), tup(Ks)))
    out["kron_csr_sample_1000x1000_d0.01_x_100x100_d0.1"] = (t, C[2].size * 16 + (C[0].size) * 8)
    return {k: {"ms": round(t * 1e3, 1), "GBs": round(b / t / 1e9, 2), "threads": 1}
            for k, (t, b) in out.items()}


def _reduce_workloads(H):
    """Host inputs of the reduction benchmarks: the 1M power-law matrix
    with every third diagonal entry added (so tr has terms), x, y, and the
    symmetric 2-D Laplacian (1000 x 1000 grid)."""
    import numpy as np
    import scipy.sparse as sp
    from gsa_hip import synth
    S = H.to_scipy()
    d = np.zeros(H.m)
    d[::3] = 1.0
    Hd = synth.HostCSR.from_scipy(sp.csr_matrix(S + sp.diags(d)))
    rng = np.random.default_rng(5)
    x = rng.standard_normal(H.m)
    y = rng.standard_normal(H.n)
    return Hd, x, y, synth.laplacian2d(1000)


def _reduce_bytes(M, x_len=0, y_len=0, write=False):
    """Pointer, index and value arrays read once (plus dense vectors, plus the
    written copy for A + λI)."""
    b = (M.m + 1) * 8 + M.nnz * 16 + (x_len + y_len) * 8
    return 2 * b + M.m * 16 if write else b


def bench_reductions(H, dev):
    """tr, sum, dot(x, A, y) (A and transpose(A)) and A + λI on the 1M power-law
    matrix (with a third of its diagonal stored), issymmetric on the 1M-row
    Laplacian; Float64 values, Int64 indices."""
    import torch
    import gsa_hip as G
    Hd, x, y, L = _reduce_workloads(H)

    def dcsr(M):
        return G.GenericSparseMatrixCSR(M.m, M.n, torch.as_tensor(M.rowptr).to(dev),
                                        torch.as_tensor(M.colval).to(dev),
                                        torch.as_tensor(M.nzval).to(dev), copy=False)
    A, dL = dcsr(Hd), dcsr(L)
    # the matrices' plans (row blocks and SpMV windows), built once per matrix as
    # for mul!: tr, sum and A + lambda*I reuse its row blocks, Float64 dot its windows
    A.plan()
    dL.plan()
    xd, yd = torch.as_tensor(x).to(dev), torch.as_tensor(y).to(dev)
    out = {}
    for name, fn, byts in (
            ("tr", lambda: G.tr(A), _reduce_bytes(Hd)),
            ("sum", lambda: G.sum(A), Hd.nnz * 8),
            ("dot_x_A_y", lambda: G.dot(xd, A, yd), _reduce_bytes(Hd, Hd.m, Hd.n)),
            ("dot_x_At_y", lambda: G.dot(yd, G.transpose(A), xd), _reduce_bytes(Hd, Hd.m, Hd.n)),
            ("issymmetric_laplacian", lambda: G.issymmetric(dL), _reduce_bytes(L)),
            ("A_plus_2I", lambda: A + 2.0 * G.I, _reduce_bytes(Hd, write=True))):
        t = _sync_time(fn, reps=5)
        out[name] = {"us": round(t * 1e6, 1), "GBs": round(byts / t / 1e9, 1)}
    assert G.issymmetric(dL)
    del A, dL, xd, yd
    torch.cuda.empty_cache()
    return out


def cpu_reductions(H):
    """The oracle (one thread) on the same reduction workloads."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O
    Hd, x, y, L = _reduce_workloads(H)

    def tm(fn):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0
    out = {}
    for name, fn, byts in (
            ("tr", lambda: O.trace_ptr(Hd.m, Hd.rowptr, Hd.colval, Hd.nzval), _reduce_bytes(Hd)),
            ("sum", lambda: float(Hd.nzval.sum()), Hd.nnz * 8),
            ("dot_x_A_y", lambda: O.dot3_ptr(Hd.m, Hd.rowptr, Hd.colval, Hd.nzval, x, y, True),
             _reduce_bytes(Hd, Hd.m, Hd.n)),
            ("issymmetric_laplacian", lambda: O.symmetric_ptr(L.m, L.rowptr, L.colval, L.nzval),
             _reduce_bytes(L)),
            ("A_plus_2I", lambda: O.add_diag(Hd.m, Hd.rowptr, Hd.colval, Hd.nzval, 2.0),
             _reduce_bytes(Hd, write=True))):
        t = tm(fn)
        out[name] = {"ms": round(t * 1e3, 2), "GBs": round(byts / t / 1e9, 2), "threads": 1}
    return out


def oracle_rel_err(A, B, C):
    """max_i |C - A*B| / (|A|*|B|) per entry, the reference product from the
    oracle (the restated reference loop, host cores, Float64 on the stored
    values widened).  A is a device CSR, B and C device tensors."""
    import numpy as np
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O
    rp = A.rowptr.cpu().numpy().astype(np.int64)
    cv = A.colval.cpu().numpy().astype(np.int64)
    nz = A.nzval.double().cpu().numpy()
    Bh = B.double().cpu().numpy()
    th = O.max_threads()
    ref = O.spmm_csr_N(A.m, rp, cv, nz, Bh, threads=th)
    bound = O.spmm_csr_N(A.m, rp, cv, np.abs(nz), np.abs(Bh), threads=th)
    got = C.double().cpu().numpy().reshape(ref.shape)
    return float((np.abs(got - ref) / np.maximum(bound, 1e-300)).max()) if got.size else 0.0


def bench_batched(dev, count=4096):
    """4096 uneven Float32 CSR instances (rows log-uniform in [64, 65536],
    Poisson(8)+1 per row, k in {1,2,4,8,16}) in one graph-captured persistent
    launch, against 4096 plain mul! calls on the same operands."""
    import numpy as np
    import torch
    import gsa_hip as G
    from gsa_hip import synth
    from gsa_hip.autotune import BatchAutotuner
    from gsa_hip.batched import BatchedSpMM

    t0 = time.time()
    insts, kk = synth.uneven_batch_device(count, seed=4, device=dev)
    items = []
    nnz_tot = rows_tot = bytes_tot = flops = 0
    plain_out = []
    for (m, rp, cv, nz), k in zip(insts, kk):
        A = G.GenericSparseMatrixCSR(m, m, rp, cv, nz, copy=False)
        # row-major n x k blocks (Julia: transpose(Bt)): the k values an entry
        # needs share a cache line; mul! takes them without a copy as well
        B = torch.randn((m, k), dtype=torch.float32, device=dev)
        C = torch.empty((m, k), dtype=torch.float32, device=dev)
        items.append((A, B if k > 1 else B[:, 0].contiguous(), C if k > 1 else C[:, 0].contiguous()))
        plain_out.append(torch.empty((k, m), dtype=torch.float32, device=dev).t() if k > 1
                         else torch.empty(m, dtype=torch.float32, device=dev))
        nnz_tot += nz.numel()
        rows_tot += m
        bytes_tot += nz.numel() * 8 + (m + 1) * 4 + 2 * m * k * 4
        flops += 2 * nz.numel() * k
    print(f"[bench] batch of {count}: {rows_tot} rows, {nnz_tot} nnz built in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    tuner = BatchAutotuner(candidates=(256, 512, 1024, 2048, 4096), reps=3)
    tuned = tuner.tune(items)
    batch = BatchedSpMM(items, rows_per_tile=tuned["rows_per_tile"])
    g = batch.graph()
    for _ in range(3):
        g.replay()
    torch.cuda.synchronize()
    reps = 20
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        g.replay()
    b.record()
    b.synchronize()
    ms_batch = a.elapsed_time(b) / reps
    # plain: one mul! per instance (plans built beforehand)
    for A, _, _ in items:
        A.plan()
    torch.cuda.synchronize()

    def plain():
        for (A, B, _), C in zip(items, plain_out):
            G.mul_(C, A, B, 1.0, 0.0)
    t_plain = _sync_time(plain, reps=2)
    # accuracy of the batched results against the FP64 oracle (the restated
    # reference loop on the host, in Float64 on the stored Float32 values):
    # every 64th instance and the 16 largest
    worst = 0.0
    batch.run()
    torch.cuda.synchronize()
    sizes = np.array([it[0].m for it in items])
    sample = sorted(set(range(0, count, 64)) | set(np.argsort(sizes)[-16:].tolist()))
    for i in sample:
        A, B, C = items[i]
        worst = max(worst, oracle_rel_err(A, B, C))
    return {"instances": count, "rows": rows_tot, "nnz": nnz_tot,
            "batched_graph_ms": round(ms_batch, 3),
            "batched_req_per_s": round(count / (ms_batch * 1e-3), 1),
            "plain_mul_ms": round(t_plain * 1e3, 3),
            "plain_req_per_s": round(count / t_plain, 1),
            "speedup": round(t_plain * 1e3 / ms_batch, 2),
            "batched_GBs": round(bytes_tot / (ms_batch * 1e-3) / 1e9, 1),
            "batched_GFLOPs": round(flops / (ms_batch * 1e-3) / 1e9, 1),
            "rows_per_tile": tuned["rows_per_tile"], "autotune_ms": tuned["timings"],
            "max_rel_err_vs_fp64_oracle": worst, "error_sample_instances": len(sample)}


def bench_serving(dev, rates=(2000, 10000, 50000), small=512, big=8):
    """Served stream: 512 uneven instances + 8 of the 1M-row power-law matrices,
    Float32 and BF16 storage, Poisson arrivals, k in {1, 4, 16}; against one
    plain mul! per request."""
    import numpy as np
    import torch
    import gsa_hip as G
    from gsa_hip import synth
    from gsa_hip.precision import BF16CSR
    from gsa_hip.serving import MatrixRegistry, Server, PlainServer, Requests, output_view

    insts, _ = synth.uneven_batch_device(small, seed=5, device=dev)
    mats = [G.GenericSparseMatrixCSR(m, m, rp, cv, nz, copy=False) for m, rp, cv, nz in insts]
    for b in range(big):
        Hb = synth.powerlaw_csr(1_000_000, 1_000_000, alpha=1.846, seed=100 + b, dtype=np.float32)
        mats.append(G.GenericSparseMatrixCSR(Hb.m, Hb.n, torch.as_tensor(Hb.rowptr).to(dev, torch.int32),
                                             torch.as_tensor(Hb.colval).to(dev, torch.int32),
                                             torch.as_tensor(Hb.nzval).to(dev), copy=False))
    xs = [torch.randn((A.n, 16), dtype=torch.float32, device=dev) for A in mats]
    out = {"matrices": len(mats)}
    for fmt in ("f32", "bf16"):
        reg = MatrixRegistry(dev, rows_per_tile=1024, max_k=16)
        for A, X in zip(mats, xs):
            reg.register(BF16CSR(A) if fmt == "bf16" else A, X)
        for A in mats:
            A.plan()
        res = {}
        for rate in rates:
            nreq = int(min(rate * 1.0, 20000))
            req = Requests.poisson(nreq, rate, len(mats), ks=(1, 4, 16), seed=5)
            srv = Server(reg, hbm_budget=2 << 30, slots=4)
            sv = srv.run(req, keep_outputs=True)
            pl = PlainServer(reg).run(req)
            worst = 0.0
            checked = 0
            for i in range(0, nreq, max(1, nreq // 12)):
                mid, k = int(req.mid[i]), int(req.k[i])
                # error against the FP64 oracle on the ORIGINAL Float32 values
                worst = max(worst, oracle_rel_err(mats[mid], xs[mid][:, :k],
                                                  output_view(sv["outputs"][i])))
                checked += 1
            res[str(rate)] = {"requests": nreq,
                              "served_req_per_s": round(sv["req_per_s"], 1),
                              "served_p50_ms": round(sv["p50_ms"], 3),
                              "served_p99_ms": round(sv["p99_ms"], 3),
                              "launches": sv["launches"], "mean_group": round(sv["mean_group"], 1),
                              "plain_req_per_s": round(pl["req_per_s"], 1),
                              "plain_p50_ms": round(pl["p50_ms"], 3),
                              "plain_p99_ms": round(pl["p99_ms"], 3),
                              "max_rel_err_vs_fp64_oracle": worst,
                              "error_sample_requests": checked}
            print(f"[bench] serving {fmt} {rate}/s: served {sv['req_per_s']:.0f} req/s p99 "
                  f"{sv['p99_ms']:.2f} ms; plain {pl['req_per_s']:.0f} req/s p99 "
                  f"{pl['p99_ms']:.2f} ms; err {worst:.2e}", file=sys.stderr, flush=True)
        # capacity: every request available at t=0
        nreq = 20000
        r6 = Requests.poisson(nreq, 1e9, len(mats), ks=(1, 4, 16), seed=6)
        req = Requests(np.zeros(nreq), r6.mid, r6.k)
        sv = Server(reg, hbm_budget=2 << 30, slots=4).run(req)
        pl = PlainServer(reg).run(req)
        res["capacity"] = {"requests": nreq, "served_req_per_s": round(sv["req_per_s"], 1),
                           "plain_req_per_s": round(pl["req_per_s"], 1),
                           "ratio": round(sv["req_per_s"] / pl["req_per_s"], 2)}
        out[fmt] = res
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--rows", type=int, default=1_000_000)
    ap.add_argument("--alpha", type=float, default=1.846)
    ap.add_argument("--copies", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-structural", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-reductions", action="store_true")
    ap.add_argument("--only", choices=("precisions", "views", "alpha2", "spmm", "structural",
                                       "reductions", "batched", "serving"),
                    help="after the headline SpMV, run this extra section alone (for profiling)")
    ap.add_argument("--profile-only", action="store_true",
                    help="run only the headline SpMV steps (for rocprofv3)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run plain 256 MiB device copies (the scale for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build runs every kernel on a single "
              "MI355X (multi-GPU sharding is a design note, not a measured path).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import gsa_hip as G
    from gsa_hip import synth

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    t0 = time.time()
    H = synth.powerlaw_csr(args.rows, args.rows, alpha=args.alpha, seed=2)
    x_h = np.random.default_rng(2).standard_normal(args.rows)
    print(f"[bench] matrix {H.m}x{H.n} nnz={H.nnz} built in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)

    def device_copies(dtype, itype, copies):
        mats, xs, ys = [], [], []
        for _ in range(copies):
            mats.append(G.GenericSparseMatrixCSR(
                H.m, H.n, torch.as_tensor(H.rowptr).to(dev, itype),
                torch.as_tensor(H.colval).to(dev, itype),
                torch.as_tensor(H.nzval).to(dev, dtype), copy=False))
            xs.append(torch.as_tensor(x_h).to(dev, dtype))
            ys.append(torch.empty(H.m, dtype=dtype, device=dev))
        return mats, xs, ys

    def spmv_run(dtype, itype, steps, warmup):
        mats, xs, ys = device_copies(dtype, itype, args.copies)
        for A in mats:
            A.plan()  # the per-matrix row-block analysis, reused by every mul!
        c = len(mats)

        def step(i):
            G.mul_(ys[i % c], mats[i % c], xs[i % c], 1.0, 0.0)

        ms_eager = time_steps(step, steps, warmup, stream)
        ms, nsteps = time_steps_graph(step, c * 10, steps, warmup, stream)
        sv = 8 if dtype == torch.float64 else 4
        si = 8 if itype == torch.int64 else 4
        gbs = spmv_bytes(H.m, H.n, H.nnz, sv, si) / (ms * 1e-3) / 1e9
        lat = latency_samples(step, 50, stream)
        del mats, xs, ys
        torch.cuda.empty_cache()
        return ms, gbs, lat, ms_eager, nsteps

    ms, gbs, lat, ms_eager, nsteps = spmv_run(torch.float64, torch.int64, args.steps, args.warmup)
    print(f"[bench] f64/i64 SpMV {ms * 1e3:.1f} us/step (graph), {ms_eager * 1e3:.1f} us eager, "
          f"{gbs:.0f} GB/s", file=sys.stderr, flush=True)
    if args.calibrate:
        src = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copies: 5 x {src.numel() * 8} bytes", file=sys.stderr,
              flush=True)
    if args.profile_only:
        return

    extra = {"spmv_f64_i64_call_p50_us": round(float(np.percentile(lat, 50)) * 1e3, 2),
             "spmv_f64_i64_call_p99_us": round(float(np.percentile(lat, 99)) * 1e3, 2),
             "spmv_f64_i64_eager_us": round(ms_eager * 1e3, 2),
             "timed_steps": nsteps}

    def want(section):
        if args.no_extra or getattr(args, "no_" + section, False):
            return False
        return args.only is None or args.only == section

    if want("precisions"):
        for dt, it, tag in ((torch.float64, torch.int32, "spmv_f64_i32"),
                            (torch.float32, torch.int32, "spmv_f32_i32")):
            ms2, gbs2, _, _, _ = spmv_run(dt, it, args.steps, args.warmup)
            extra[tag + "_GBs"] = round(gbs2, 1)
            extra[tag + "_us"] = round(ms2 * 1e3, 2)
            print(f"[bench] {tag} {ms2 * 1e3:.1f} us, {gbs2:.0f} GB/s", file=sys.stderr, flush=True)
    if want("views"):
        # the same matrix as CSC and COO (row-major from CSR, and column-major
        # from CSC), f64/i64, through their row views; algorithmic bytes of each format
        A = G.GenericSparseMatrixCSR(H.m, H.n, torch.as_tensor(H.rowptr).to(dev),
                                     torch.as_tensor(H.colval).to(dev),
                                     torch.as_tensor(H.nzval).to(dev), copy=False)
        xd = torch.as_tensor(x_h).to(dev)
        yd = torch.empty(H.m, dtype=torch.float64, device=dev)
        Ac = G.GenericSparseMatrixCSC(A)
        for tag, M, byts in (
                ("spmv_csc_f64_i64", Ac, spmv_bytes(H.m, H.n, H.nnz, 8, 8) - (H.m + 1) * 8 + (H.n + 1) * 8),
                ("spmv_coo_rowmajor_f64_i64", G.GenericSparseMatrixCOO(A), H.nnz * 24 + (H.n + H.m) * 8),
                ("spmv_coo_colmajor_f64_i64", G.GenericSparseMatrixCOO(Ac), H.nnz * 24 + (H.n + H.m) * 8),
                # transposed products: the reference's atomic scatters (CSR-T, COO-T)
                ("spmv_csr_T_f64_i64", G.transpose(A), spmv_bytes(H.m, H.n, H.nnz, 8, 8)),
                ("spmv_coo_T_f64_i64", G.transpose(G.GenericSparseMatrixCOO(A)),
                 H.nnz * 24 + (H.n + H.m) * 8)):
            if hasattr(M, "plan"):
                M.plan()
            elif hasattr(M, "parent") and hasattr(M.parent, "plan"):
                M.parent.plan()
            ms4 = time_steps(lambda i: G.mul_(yd, M, xd, 1.0, 0.0), 20, 3, stream)
            extra[tag + "_us"] = round(ms4 * 1e3, 1)
            extra[tag + "_GBs"] = round(byts / (ms4 * 1e-3) / 1e9, 1)
            print(f"[bench] {tag} {ms4 * 1e3:.1f} us", file=sys.stderr, flush=True)
            del M
        if not args.no_cpu:  # the oracle's scatters (restated reference loops), one thread
            sys.path.insert(0, os.path.join(HERE, "oracle"))
            import oracle as O
            cp, rv, nzc = (t.cpu().numpy() for t in (Ac.colptr, Ac.rowval, Ac.nzval))
            t0 = time.perf_counter()
            O.spmm_csr_T(H.n, H.m, cp, rv, nzc, x_h)  # CSC A x (matrix_csc_kernels.jl:1-22)
            tc = time.perf_counter() - t0
            Ao = G.GenericSparseMatrixCOO(A)
            ri, ci, nzo = (t.cpu().numpy() for t in (Ao.rowind, Ao.colind, Ao.nzval))
            t0 = time.perf_counter()
            O.spmm_coo(False, H.m, ri, ci, nzo, x_h)  # COO A x (matrix_coo_kernels.jl:1-20)
            to = time.perf_counter() - t0
            t0 = time.perf_counter()
            O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, H.nzval, x_h)  # CSR-T (matrix_csr_kernels.jl:24-45)
            tt = time.perf_counter() - t0
            t0 = time.perf_counter()
            O.spmm_coo(True, H.n, ri, ci, nzo, x_h)  # COO-T (matrix_coo_kernels.jl:22-43)
            tot = time.perf_counter() - t0
            extra["spmv_csc_coo_cpu_oracle"] = {
                "csc_ms": round(tc * 1e3, 1), "coo_rowmajor_ms": round(to * 1e3, 1),
                "csr_T_ms": round(tt * 1e3, 1), "coo_T_ms": round(tot * 1e3, 1), "threads": 1}
            print(f"[bench] CSC / COO SpMV CPU oracle {tc * 1e3:.0f} / {to * 1e3:.0f} ms, 1 thread",
                  file=sys.stderr, flush=True)
            del cp, rv, nzc, ri, ci, nzo, Ao
        del A, Ac, xd, yd
        torch.cuda.empty_cache()
    if want("alpha2"):
        # secondary point: Zipf alpha = 2.0 (mean 5.53 entries per row)
        H2 = synth.powerlaw_csr(args.rows, args.rows, alpha=2.0, seed=2)
        A2 = G.GenericSparseMatrixCSR(H2.m, H2.n, torch.as_tensor(H2.rowptr).to(dev),
                                      torch.as_tensor(H2.colval).to(dev),
                                      torch.as_tensor(H2.nzval).to(dev), copy=False)
        x2 = torch.as_tensor(x_h).to(dev)
        y2 = torch.empty(H2.m, dtype=torch.float64, device=dev)
        A2.plan()
        ms5 = time_steps(lambda i: G.mul_(y2, A2, x2, 1.0, 0.0), 50, 5, stream)
        extra["spmv_alpha2_f64_i64_nnz"] = H2.nnz
        extra["spmv_alpha2_f64_i64_us"] = round(ms5 * 1e3, 1)
        extra["spmv_alpha2_f64_i64_GBs"] = round(spmv_bytes(H2.m, H2.n, H2.nnz, 8, 8) / (ms5 * 1e-3) / 1e9, 1)
        print(f"[bench] alpha=2.0 SpMV {ms5 * 1e3:.1f} us", file=sys.stderr, flush=True)
        del A2, x2, y2, H2
        torch.cuda.empty_cache()
    if want("spmm"):
        # SpMM with a dense 1M x 32 right-hand side (column-major, as Julia's Matrix)
        k = 32
        A = G.GenericSparseMatrixCSR(H.m, H.n, torch.as_tensor(H.rowptr).to(dev),
                                     torch.as_tensor(H.colval).to(dev),
                                     torch.as_tensor(H.nzval).to(dev), copy=False)
        B = torch.randn((k, H.n), dtype=torch.float64, device=dev).t()
        C = torch.empty((k, H.m), dtype=torch.float64, device=dev).t()
        A.plan()
        ms3 = time_steps(lambda i: G.mul_(C, A, B, 1.0, 0.0), 10, 3, stream)
        b3 = spmm_bytes(H.m, H.n, H.nnz, k, 8, 8)
        extra["spmm_k32_f64_us"] = round(ms3 * 1e3, 1)
        extra["spmm_k32_f64_GBs"] = round(b3 / (ms3 * 1e-3) / 1e9, 1)
        extra["spmm_k32_f64_GFLOPs"] = round(2 * H.nnz * k / (ms3 * 1e-3) / 1e9, 1)
        print(f"[bench] SpMM k=32 {ms3 * 1e3:.0f} us", file=sys.stderr, flush=True)
        # the scatter-shaped products (CSC * B, COO * B, transpose(CSR) * B) run
        # the same gather kernel through a plan's gather view
        Ct = torch.empty((k, H.n), dtype=torch.float64, device=dev).t()
        for tag, M, Cm in (("csc", G.GenericSparseMatrixCSC(A), C),
                           ("coo", G.GenericSparseMatrixCOO(A), C),
                           ("csr_T", G.transpose(A), Ct)):
            Bm = B if tag != "csr_T" else torch.randn((k, H.m), dtype=torch.float64, device=dev).t()
            G.mul_(Cm, M, Bm, 1.0, 0.0)  # builds the gather view
            ms5 = time_steps(lambda i: G.mul_(Cm, M, Bm, 1.0, 0.0), 10, 3, stream)
            extra[f"spmm_k32_{tag}_f64_us"] = round(ms5 * 1e3, 1)
            print(f"[bench] SpMM k=32 {tag} {ms5 * 1e3:.0f} us", file=sys.stderr, flush=True)
            del M, Bm
        del A, B, C, Ct
        torch.cuda.empty_cache()
        if not args.no_cpu:  # the oracle's SpMM (restated reference loop) on the host cores
            sys.path.insert(0, os.path.join(HERE, "oracle"))
            import oracle as O
            Bh = np.random.default_rng(3).standard_normal((k, H.n)).T  # column-major, as on the GPU
            th = O.max_threads()
            t0 = time.perf_counter()
            O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, Bh, threads=th)
            tc = time.perf_counter() - t0
            extra["spmm_k32_f64_cpu_oracle"] = {"ms": round(tc * 1e3, 1),
                                                "GBs": round(b3 / tc / 1e9, 2), "threads": th}
            print(f"[bench] SpMM k=32 CPU oracle {tc * 1e3:.0f} ms on {th} threads",
                  file=sys.stderr, flush=True)
            del Bh

    if want("structural"):
        extra["structural"] = bench_structural(H, dev)
        print(f"[bench] structural {extra['structural']}", file=sys.stderr, flush=True)
        if not args.no_cpu:
            extra["structural_cpu_oracle"] = cpu_structural(H)
            print(f"[bench] structural CPU (oracle, 1 thread) {extra['structural_cpu_oracle']}",
                  file=sys.stderr, flush=True)
    if want("reductions"):
        extra["reductions"] = bench_reductions(H, dev)
        print(f"[bench] reductions {extra['reductions']}", file=sys.stderr, flush=True)
        if not args.no_cpu:
            extra["reductions_cpu_oracle"] = cpu_reductions(H)
            print(f"[bench] reductions CPU (oracle, 1 thread) {extra['reductions_cpu_oracle']}",
                  file=sys.stderr, flush=True)
    if want("batched"):
        extra["batched"] = bench_batched(dev)
        print(f"[bench] batched {extra['batched']}", file=sys.stderr, flush=True)
    if want("serving"):
        extra["serving"] = bench_serving(dev)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(H, x_h, seconds=10.0)
        print(f"[bench] CPU baseline {cpu['value']} GB/s on {cpu['cores']} threads",
              file=sys.stderr, flush=True)

    algo = spmv_bytes(H.m, H.n, H.nnz, 8, 8)
    line = {
        "metric": "SpMV/SpMM GB/s and % HBM peak; batched req/s + p99 vs plain mul!",
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded power-law CSR, N(0,1) values and x)",
        "config": {"workload": "CSR SpMV y = A*x, Float64 values, Int64 one-based indices",
                   "rows": H.m, "cols": H.n, "nnz": H.nnz,
                   "row_law": f"Zipf(alpha={args.alpha}) truncated to [1,5000]",
                   "mean_nnz_per_row": round(H.nnz / H.m, 3),
                   "copies_rotated": args.copies, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(), "algorithmic_bytes": algo,
                     "kernel": "k_csr_spmv<double,long>",
                     "timed": "HIP events around each mul!, i.e. k_csr_spmv and the short "
                              "k_spmv_fold that adds the partial sums of rows crossing windows; "
                              "the kernel's own average is in the rocprofv3 summary"},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
