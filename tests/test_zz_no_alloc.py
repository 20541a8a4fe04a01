"""The allocation-free SpMV entry point (gsa_spmv_csr_noalloc) against the
existing one (gsa_spmm_csr, opA = N, k = 1, no plan), both called through the
C ABI: the same kernels in the same order, so the results must agree bit for
bit.  Two calls in a row, the second on a smaller matrix (the fixed scratch
array keeps the first call's contents), for each value and index type; the
oracle checks the first result to the usual Float64 / Float32 tolerance.
"""
import ctypes

import numpy as np
import pytest
import torch

import gsa_hip._lib as L
import oracle as O
from gsa_hip import synth
from util import host, norm_bound, assert_normwise

pytestmark = pytest.mark.gpu

_VT = {torch.float64: L.GSA_F64, torch.float32: L.GSA_F32}
_IT = {torch.int64: L.GSA_I64, torch.int32: L.GSA_I32}


def _dev(H, dtype, itype, dev):
    return (torch.as_tensor(H.rowptr).to(device=dev, dtype=itype),
            torch.as_tensor(H.colval).to(device=dev, dtype=itype),
            torch.as_tensor(H.nzval).to(device=dev, dtype=dtype))


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _both(H, x, y0, dtype, itype, dev, alpha, beta):
    rp, cv, nz = _dev(H, dtype, itype, dev)
    xd = torch.as_tensor(x, device=dev, dtype=dtype)
    y_ref = torch.as_tensor(y0, device=dev, dtype=dtype).clone()
    y_new = y_ref.clone()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib = L.load()
    ct = ctypes.c_double if dtype == torch.float64 else ctypes.c_float
    a_ref, b_ref = ct(alpha), ct(beta)  # host scalars of the value type
    L.check(lib.gsa_spmm_csr(s, _VT[dtype], _IT[itype], L.GSA_OP_N, H.m, H.n, 1, _p(rp), _p(cv),
                             _p(nz), H.nnz, _p(xd), H.n, L.GSA_OP_N, _p(y_ref), H.m,
                             ctypes.byref(a_ref), ctypes.byref(b_ref), None), "gsa_spmm_csr")
    L.check(lib.gsa_spmv_csr_noalloc(s, _VT[dtype], _IT[itype], H.m, H.n, _p(rp), _p(cv),
                                     _p(nz), H.nnz, _p(xd), _p(y_new), alpha, beta),
            "gsa_spmv_csr_noalloc")
    torch.cuda.synchronize()
    return host(y_ref), host(y_new)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_noalloc_matches_allocating_entry_point(dev, dtype, itype):
    npd = np.float64 if dtype == torch.float64 else np.float32
    rng = np.random.default_rng(11)
    # first call: 300k power-law rows (long rows cross many windows), then a
    # smaller matrix whose windows reuse the front of the same array
    for m, seed, alpha, beta in ((300_000, 21, 1.0, 0.0), (40_000, 22, -0.5, 1.5)):
        H = synth.powerlaw_csr(m, m, alpha=1.846, seed=seed, dtype=npd)
        x = rng.standard_normal(m).astype(npd)
        y0 = rng.standard_normal(m).astype(npd)
        if beta == 0.0:
            y0[::5] = np.nan  # beta = 0 clears NaN
        ref, new = _both(H, x, y0, dtype, itype, dev, alpha, beta)
        assert np.array_equal(ref, new, equal_nan=False), (m, dtype, itype)
        want = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x, C=y0.astype(np.float64),
                            alpha=alpha, beta=beta)
        bound = norm_bound(H.rowptr, H.colval, H.nzval, x) * abs(alpha) + \
            (abs(beta) * np.abs(y0) if beta != 0.0 else 0.0)
        assert_normwise(new, want, bound, 1e-12 if dtype == torch.float64 else 2e-6)


def test_noalloc_edges_and_limit(dev):
    lib = L.load()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # no stored entries: only beta applies
    rp = torch.ones(4, dtype=torch.int64, device=dev)
    e = torch.empty(0, dtype=torch.int64, device=dev)
    v = torch.empty(0, dtype=torch.float64, device=dev)
    x = torch.ones(3, dtype=torch.float64, device=dev)
    y = torch.tensor([1.0, float("nan"), 2.0], dtype=torch.float64, device=dev)
    L.check(lib.gsa_spmv_csr_noalloc(s, L.GSA_F64, L.GSA_I64, 3, 3, _p(rp), _p(e), _p(v), 0,
                                     _p(x), _p(y), 1.0, 0.0), "gsa_spmv_csr_noalloc")
    torch.cuda.synchronize()
    assert host(y).tolist() == [0.0, 0.0, 0.0]
    # above the limit: rejected before any device work (the arrays are never read)
    st = lib.gsa_spmv_csr_noalloc(s, L.GSA_F64, L.GSA_I64, 3, 3, _p(rp), _p(e), _p(v),
                                  L.GSA_NOALLOC_MAX_NNZ + 1, _p(x), _p(y), 1.0, 0.0)
    assert st == L.GSA_ERR_INVALID
    assert b"at most" in lib.gsa_last_error()
    # the allocation-free call caches no scratch on its stream
    before = lib.gsa_scratch_cached(s)
    H = synth.powerlaw_csr(50_000, 50_000, alpha=1.846, seed=23)
    ref, new = _both(H, np.ones(H.n), np.zeros(H.m), torch.float64, torch.int64, dev, 1.0, 0.0)
    assert np.array_equal(ref, new)
    rp2, cv2, nz2 = _dev(H, torch.float64, torch.int64, dev)
    x2 = torch.ones(H.n, dtype=torch.float64, device=dev)
    y2 = torch.zeros(H.m, dtype=torch.float64, device=dev)
    mid = lib.gsa_scratch_cached(s)
    L.check(lib.gsa_spmv_csr_noalloc(s, L.GSA_F64, L.GSA_I64, H.m, H.n, _p(rp2), _p(cv2),
                                     _p(nz2), H.nnz, _p(x2), _p(y2), 1.0, 0.0), "noalloc")
    torch.cuda.synchronize()
    assert lib.gsa_scratch_cached(s) == mid and mid >= before
    assert np.array_equal(host(y2), ref)
