"""Parity at the benchmark's full sizes: every product and structural operation
bench.py times is checked here on the same seeded inputs at the same scale.

- C2, the 1M x 1M power-law matrix (Zipf alpha 1.846 on [1, 5000], seed 2):
  Float32 SpMV (Int32 and Int64 indices), SpMM k = 32 with column-major B and
  C, CSC and both COO storage orders as SpMV, transpose(CSR) and transpose(COO)
  SpMV, and the k = 32 SpMM of CSC, COO and transpose(CSR) (gather views).
- C2 with every third diagonal entry added: tr, sum, dot(x, A, y) and A + 2I;
  issymmetric on the 1000 x 1000 grid Laplacian; dropzeros of C2 with every
  tenth stored value zero.
- C3: R*R with R 200,000 x 200,000 at 8 entries per row (seed 3).
- C4: the seed-4 batch of 4096 uneven Float32 instances, graph-replayed.
- C5: the 520-matrix registry (512 uneven + 8 1M-row power-law), Float32 and
  BF16 storage, one Poisson rate.

Tolerances, written per test:
- Float64: |y - y_ref| <= 1e-12 * sum_j |a_ij||x_j| per entry (the product's
  normaliser); rows the kernels sum on one lane in the reference's order (at
  most 32 entries for SpMV, 64 for SpMM) must be bit-identical.
- Float32 (values and x Float32): against the Float64 oracle on the same
  values, |y - y64| <= (len_i + 2) * 2^-24 * sum_j |a_ij||x_j| per row (the
  rigorous bound of a left-to-right Float32 sum); rows of at most 32 entries
  bit-identical with the Float32 oracle (the reference's loop in Float32).
- Batched and served Float32: 1e-5 normwise against the Float64 oracle (the
  stated target); BF16: 1e-5 against the Float64 product of the dequantized
  values (the storage rounding is reported by bench.py, not hidden here).
"""
import numpy as np
import pytest
import torch

import gsa_hip as G
import oracle as O
from gsa_hip import synth
from util import dev_csr, host, assert_normwise

pytestmark = pytest.mark.gpu

TH = None


def threads():
    global TH
    if TH is None:
        TH = O.max_threads()
    return TH


@pytest.fixture(scope="module")
def C2():
    H = synth.powerlaw_csr(1_000_000, 1_000_000, alpha=1.846, seed=2)
    return H


def _f32_bound(H, v32, x32):
    lens = np.diff(H.rowptr).astype(np.float64)
    a = O.spmm_csr_N(H.m, H.rowptr, H.colval, np.abs(v32.astype(np.float64)),
                     np.abs(x32.astype(np.float64)), threads=threads())
    return (lens + 2.0) * 2.0 ** -24 * a


@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_c2_spmv_float32_full_size(dev, C2, itype):
    H = C2
    v32 = H.nzval.astype(np.float32)
    x32 = np.random.default_rng(2).standard_normal(H.n).astype(np.float32)
    A = G.GenericSparseMatrixCSR(H.m, H.n, torch.as_tensor(H.rowptr).to(dev, itype),
                                 torch.as_tensor(H.colval).to(dev, itype),
                                 torch.as_tensor(v32, device=dev), copy=False)
    y = host(A @ torch.as_tensor(x32, device=dev)).astype(np.float64)
    ref64 = O.spmm_csr_N(H.m, H.rowptr, H.colval, v32.astype(np.float64),
                         x32.astype(np.float64), threads=threads())
    bound = _f32_bound(H, v32, x32)
    bad = np.flatnonzero(np.abs(y - ref64) > bound + 1e-300)
    assert bad.size == 0, bad[:5]
    ref32 = O.spmm_csr_N(H.m, H.rowptr, H.colval, v32, x32, threads=threads())
    short = np.diff(H.rowptr) <= 32
    assert np.array_equal(y[short], ref32[short].astype(np.float64))


def test_c2_spmm_k32_colmajor_full_size(dev, C2):
    H = C2
    k = 32
    rng = np.random.default_rng(3)
    B = np.asfortranarray(rng.standard_normal((H.n, k)))
    A = dev_csr(H)
    C = host(A @ G.dense(B, device=dev))
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, B, threads=threads())
    bound = O.spmm_csr_N(H.m, H.rowptr, H.colval, np.abs(H.nzval), np.abs(B), threads=threads())
    worst = assert_normwise(C, ref, bound, 1e-12)
    short = np.diff(H.rowptr) <= 64  # one lane per (row, column), left to right
    assert np.array_equal(C[short], ref[short])
    print(f"SpMM k=32 max normalised error {worst:.3e}")


def test_c2_csc_coo_and_transposed_spmv_full_size(dev, C2):
    """CSC (row bins), COO stored row-major (row view) and column-major (row
    bins), transpose(CSR) and transpose(COO) (transposed bins): each against the
    oracle's scatter in storage order, 1e-12 normwise."""
    H = C2
    rng = np.random.default_rng(4)
    x = rng.standard_normal(H.n)
    xt = rng.standard_normal(H.m)
    A = dev_csr(H)
    Ac = G.GenericSparseMatrixCSC(A)
    Ao = G.GenericSparseMatrixCOO(A)       # row-major entries
    Aoc = G.GenericSparseMatrixCOO(Ac)     # column-major entries
    refN = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x, threads=threads())
    bN = O.spmm_csr_N(H.m, H.rowptr, H.colval, np.abs(H.nzval), np.abs(x), threads=threads())
    xd = torch.as_tensor(x, device=dev)
    for M in (Ac, Ao, Aoc):
        assert_normwise(host(M @ xd), refN, bN, 1e-12)
    refT = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, H.nzval, xt)
    bT = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, np.abs(H.nzval), np.abs(xt))
    xtd = torch.as_tensor(xt, device=dev)
    for M in (A, Ao):
        assert_normwise(host(G.transpose(M) @ xtd), refT, bT, 1e-12)


def test_c2_scatter_shaped_spmm_k32_full_size(dev, C2):
    """CSC * B, COO * B and transpose(CSR) * B with k = 32 (the gather views)."""
    H = C2
    k = 32
    rng = np.random.default_rng(5)
    A = dev_csr(H)
    B = np.asfortranarray(rng.standard_normal((H.n, k)))
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, B, threads=threads())
    bound = O.spmm_csr_N(H.m, H.rowptr, H.colval, np.abs(H.nzval), np.abs(B), threads=threads())
    Bd = G.dense(B, device=dev)
    for M in (G.GenericSparseMatrixCSC(A), G.GenericSparseMatrixCOO(A)):
        assert_normwise(host(M @ Bd), ref, bound, 1e-12)
        del M
        torch.cuda.empty_cache()
    del Bd
    Bt = np.asfortranarray(rng.standard_normal((H.m, k)))
    refT = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, H.nzval, Bt)
    bT = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, np.abs(H.nzval), np.abs(Bt))
    assert_normwise(host(G.transpose(A) @ G.dense(Bt, device=dev)), refT, bT, 1e-12)


def test_c2_reductions_and_dropzeros_full_size(dev, C2):
    """The reduction and dropzeros workloads of bench.py at their sizes, through
    the matrices' cached plans as timed there: sums within 1e-12 of the sum of
    |term|, A + 2I and dropzeros index for index and bit for bit."""
    import scipy.sparse as sp
    S = C2.to_scipy()
    d = np.zeros(C2.m)
    d[::3] = 1.0
    Hd = synth.HostCSR.from_scipy(sp.csr_matrix(S + sp.diags(d)))
    A = dev_csr(Hd)
    A.plan()
    t, bound = O.trace_ptr(Hd.m, Hd.rowptr, Hd.colval, Hd.nzval)
    assert abs(G.tr(A) - t) <= 1e-12 * bound
    assert abs(G.sum(A) - Hd.nzval.sum()) <= 1e-12 * np.abs(Hd.nzval).sum()
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal(Hd.m), rng.standard_normal(Hd.n)
    ref, b2 = O.dot3_ptr(Hd.m, Hd.rowptr, Hd.colval, Hd.nzval, x, y, True)
    got = G.dot(torch.as_tensor(x).to(dev), A, torch.as_tensor(y).to(dev))
    assert abs(got - ref) <= 1e-12 * b2
    rp, cv, nz = O.add_diag(Hd.m, Hd.rowptr, Hd.colval, Hd.nzval, 2.0)
    B = A + 2.0 * G.I
    assert np.array_equal(host(B.rowptr), rp) and np.array_equal(host(B.colval), cv)
    assert np.array_equal(host(B.nzval), nz)
    del A, B
    L = synth.laplacian2d(1000)
    dL = dev_csr(L)
    dL.plan()
    assert O.symmetric_ptr(L.m, L.rowptr, L.colval, L.nzval)
    assert G.issymmetric(dL)
    Z = C2.nzval.copy()
    Z[::10] = 0
    rp, cv, nz = O.csr_dropzeros(C2.m, C2.rowptr, C2.colval, Z)
    Az = G.GenericSparseMatrixCSR(C2.m, C2.n, torch.as_tensor(C2.rowptr).to(dev),
                                  torch.as_tensor(C2.colval).to(dev),
                                  torch.as_tensor(Z).to(dev), copy=False)
    D = G.dropzeros(Az)
    assert np.array_equal(host(D.rowptr), rp) and np.array_equal(host(D.colval), cv)
    assert np.array_equal(host(D.nzval), nz)


def test_c3_spgemm_random_full_size(dev):
    """R*R for R 200,000 x 200,000 at 8 distinct uniform columns per row, N(0,1)
    values: pattern and values bit-exact with the oracle's Gustavson order."""
    R = synth.fixed_row_csr(200_000, 200_000, 8, seed=3)
    A = dev_csr(R)
    C = A * A
    rp, cv, nz = O.spgemm_csr(R.m, R.n, (R.rowptr, R.colval, R.nzval),
                              (R.rowptr, R.colval, R.nzval))
    assert C.nnz() == cv.size
    assert np.array_equal(host(C.rowptr), rp)
    assert np.array_equal(host(C.colval), cv)
    assert np.array_equal(host(C.nzval), nz)


def test_c4_batch_4096_graph_replay_full_size(dev):
    """The seed-4 batch: 4096 square Float32 instances, rows log-uniform in
    [64, 65536], Poisson(8)+1 entries per row, k in {1, 2, 4, 8, 16}; one
    graph-captured persistent launch, replayed; every instance against the
    Float64 oracle at 1e-5 normwise."""
    from gsa_hip.batched import BatchedSpMM
    insts, kk = synth.uneven_batch_device(4096, seed=4, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(44)
    items = []
    for (m, rp, cv, nz), k in zip(insts, kk):
        A = G.GenericSparseMatrixCSR(m, m, rp, cv, nz, copy=False)
        B = torch.randn((m, k), dtype=torch.float32, device=dev, generator=gen)
        C = torch.empty((m, k), dtype=torch.float32, device=dev)
        items.append((A, B if k > 1 else B[:, 0].contiguous(),
                      C if k > 1 else C[:, 0].contiguous()))
    batch = BatchedSpMM(items, rows_per_tile=1024)
    g = batch.graph()
    for _ in range(2):
        for _, _, C in items[::97]:
            C.fill_(float("nan"))
        g.replay()
    torch.cuda.synchronize()
    worst = 0.0
    for i, ((A, B, C), k) in enumerate(zip(items, kk)):
        rp = host(A.rowptr).astype(np.int64)
        cv = host(A.colval).astype(np.int64)
        v = host(A.nzval).astype(np.float64)
        Bh = host(B).astype(np.float64)
        ref = O.spmm_csr_N(A.m, rp, cv, v, Bh, threads=threads())
        bound = O.spmm_csr_N(A.m, rp, cv, np.abs(v), np.abs(Bh), threads=threads())
        got = host(C).astype(np.float64).reshape(ref.shape)
        worst = max(worst, assert_normwise(got, ref, bound, 1e-5))
    print(f"4096 instances: max normalised error {worst:.3e}")


@pytest.mark.parametrize("fmt", ["f32", "bf16"])
def test_c5_served_registry_full_size(dev, fmt):
    """520 registered matrices (512 of the uneven distribution, seed 5, and 8 of
    the 1M-row power-law law, seeds 100-107), Poisson arrivals at 10k req/s, k
    in {1, 4, 16}: every result is the same alone and in the mix, and a sample
    of requests (including every request on a 1M-row matrix in the sample range)
    matches the Float64 oracle."""
    from gsa_hip.precision import BF16CSR
    from gsa_hip.serving import MatrixRegistry, Server, Requests, output_view
    insts, _ = synth.uneven_batch_device(512, seed=5, device=dev)
    mats = [G.GenericSparseMatrixCSR(m, m, rp, cv, nz, copy=False) for m, rp, cv, nz in insts]
    for b in range(8):
        Hb = synth.powerlaw_csr(1_000_000, 1_000_000, alpha=1.846, seed=100 + b, dtype=np.float32)
        mats.append(G.GenericSparseMatrixCSR(
            Hb.m, Hb.n, torch.as_tensor(Hb.rowptr).to(dev, torch.int32),
            torch.as_tensor(Hb.colval).to(dev, torch.int32),
            torch.as_tensor(Hb.nzval).to(dev), copy=False))
    gen = torch.Generator(device=dev)
    gen.manual_seed(55)
    xs = [torch.randn((A.n, 16), dtype=torch.float32, device=dev, generator=gen) for A in mats]
    reg = MatrixRegistry(dev, rows_per_tile=1024, max_k=16)
    for A, X in zip(mats, xs):
        reg.register(BF16CSR(A) if fmt == "bf16" else A, X)
    req = Requests.poisson(1500, 10_000, len(mats), ks=(1, 4, 16), seed=5)
    mixed = Server(reg, hbm_budget=2 << 30, slots=4).run(req, keep_outputs=True)
    alone = Server(reg, hbm_budget=2 << 30, slots=1, max_group=1).run(req, keep_outputs=True)
    for i in range(len(req)):
        assert np.array_equal(host(output_view(mixed["outputs"][i])),
                              host(output_view(alone["outputs"][i]))), i
    big = [i for i in range(len(req)) if int(req.mid[i]) >= 512]
    sample = sorted(set(range(0, len(req), 25)) | set(big[:12]))
    worst = 0.0
    for i in sample:
        mid, k = int(req.mid[i]), int(req.k[i])
        A = mats[mid]
        rp = host(A.rowptr).astype(np.int64)
        cv = host(A.colval).astype(np.int64)
        v = (reg.mats[mid].dequantized_values() if fmt == "bf16"
             else host(A.nzval)).astype(np.float64)
        Bh = host(xs[mid][:, :k]).astype(np.float64)
        ref = O.spmm_csr_N(A.m, rp, cv, v, Bh, threads=threads())
        bound = O.spmm_csr_N(A.m, rp, cv, np.abs(v), np.abs(Bh), threads=threads())
        got = host(output_view(mixed["outputs"][i])).astype(np.float64).reshape(ref.shape)
        worst = max(worst, assert_normwise(got, ref, bound, 1e-5))
    print(f"served {fmt}: {len(sample)} requests checked, max normalised error {worst:.3e}")
