"""Helpers shared by the GPU parity tests: host <-> device matrices, error norms."""
import numpy as np
import torch

import gsa_hip as G
import oracle as O


def dev_csr(H, dtype=torch.float64, itype=torch.int64, device="cuda"):
    return G.GenericSparseMatrixCSR(
        H.m, H.n, torch.as_tensor(H.rowptr).to(device=device, dtype=itype),
        torch.as_tensor(H.colval).to(device=device, dtype=itype),
        torch.as_tensor(H.nzval).to(device=device, dtype=dtype), copy=False)


def dev_csc_of(H, dtype=torch.float64, itype=torch.int64, device="cuda"):
    """CSC storage of the host CSR matrix H (transposed through scipy on the host)."""
    S = H.to_scipy().tocsc()
    S.sort_indices()
    return G.GenericSparseMatrixCSC(
        H.m, H.n, torch.as_tensor(S.indptr.astype(np.int64) + 1).to(device=device, dtype=itype),
        torch.as_tensor(S.indices.astype(np.int64) + 1).to(device=device, dtype=itype),
        torch.as_tensor(S.data).to(device=device, dtype=dtype), copy=False)


def host(t):
    return t.detach().cpu().numpy()


def norm_bound(rowptr, colval, nzval, x, trans=False, m=None, n=None):
    """Σ_j |a_ij| |x_j| per output row (the normaliser of the tolerance)."""
    a = np.abs(nzval).astype(np.float64)
    xa = np.abs(np.asarray(x, dtype=np.float64))
    if not trans:
        return O.spmm_csr_N(len(rowptr) - 1, rowptr, colval, a, xa)
    return O.spmm_csr_T(len(rowptr) - 1, n, rowptr, colval, a, xa)


def assert_normwise(got, ref, bound, tol):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    err = np.abs(got - ref)
    lim = tol * np.asarray(bound) + 1e-300
    bad = np.flatnonzero(err > lim)
    assert bad.size == 0, (f"{bad.size} entries over tolerance {tol}; worst rel "
                           f"{(err / np.maximum(bound, 1e-300)).max():.3e} at {bad[:5]}")
    return float((err / np.maximum(bound, 1e-300)).max()) if err.size else 0.0
