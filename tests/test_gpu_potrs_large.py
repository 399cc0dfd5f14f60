"""Cholesky::solve_mut (src/linalg/cholesky.rs:122-129) at large n through
nrocm_dpotrs: the forward sweep L y = b (solve.rs:500-519) on one 64-row block
per workgroup, then L^T x = y (solve.rs:697-755).

EXACT is bitwise against the oracle's restatement of nalgebra's loops on the
same factor, FAST within 1e-12 relative.  Factors are synthetic (any lower
triangle with a nonzero diagonal is a valid input of the solve) and their
strict upper triangle holds NaN, which the solve must never read; n = 20000
takes the path that keeps x in global memory instead of LDS."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

EXACT, FAST = 0, 1


def synthetic_factor(n, seed):
    rng = np.random.default_rng(seed)
    L = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, n)) / np.sqrt(n))
    L[np.arange(n), np.arange(n)] = rng.uniform(1.0, 2.0, n)
    for j in range(1, n):
        L[:j, j] = np.nan
    return L


def gpu_potrs(L, B, mode):
    from nalgebra_amd._lib import lib
    from nalgebra_amd.linalg import handle
    n, nrhs = B.shape
    d_l = torch.from_numpy(L.ravel(order="F")).cuda()
    d_b = torch.from_numpy(np.asfortranarray(B).ravel(order="F").copy()).cuda()
    st = lib().nrocm_dpotrs(handle(), n, nrhs, C.c_void_p(d_l.data_ptr()), n, C.c_void_p(d_b.data_ptr()), n, mode)
    assert st == 0, st
    torch.cuda.synchronize()
    return d_b.cpu().numpy().reshape(nrhs, n).T


@pytest.mark.parametrize("n,nrhs", [(1024, 1), (1500, 3), (4096, 1), (8192, 2)])
def test_potrs_large_exact(nl, orc, n, nrhs):
    L = synthetic_factor(n, 900 + n)
    B = np.random.default_rng(n).standard_normal((n, nrhs))
    print(f"potrs exact n={n} nrhs={nrhs}", flush=True)
    x = gpu_potrs(L, B, EXACT)
    xr = orc.potrs(L, B)
    assert np.array_equal(x.view(np.uint64), xr.view(np.uint64))
    # the solution solves the system (pins what the oracle computes)
    Lt = np.tril(np.nan_to_num(L, nan=0.0))
    res = Lt @ (Lt.T @ x) - B
    assert np.linalg.norm(res) / np.linalg.norm(B) <= 1e-13


@pytest.mark.parametrize("n", [1024, 3000, 8192])
def test_potrs_large_fast(nl, orc, n):
    L = synthetic_factor(n, 1700 + n)
    B = np.random.default_rng(n + 1).standard_normal((n, 2))
    print(f"potrs fast n={n}", flush=True)
    x = gpu_potrs(L, B, FAST)
    xr = orc.potrs(L, B)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 1e-12


def test_potrs_exact_beyond_lds(nl, orc):
    """n = 20000 > 19968: x lives in b itself while the transposed solve runs."""
    n = 20000
    L = synthetic_factor(n, 31)
    B = np.random.default_rng(32).standard_normal((n, 1))
    print("potrs exact n=20000", flush=True)
    x = gpu_potrs(L, B, EXACT)
    xr = orc.potrs(L, B)
    assert np.array_equal(x.view(np.uint64), xr.view(np.uint64))


def test_potrf_potrs_exact_2048(nl, orc):
    """GPU Cholesky::new (exact) then the large-n solve, against the oracle on the same input."""
    n = 2048
    rng = np.random.default_rng(77)
    G = rng.standard_normal((n, n))
    A = np.asfortranarray(G @ G.T + n * np.eye(n))
    B = rng.standard_normal((n, 2))
    ch = nl.DMatrix.from_numpy(A).cholesky(nl.EXACT)
    assert ch is not None
    x = ch.solve(nl.DMatrix.from_numpy(B)).numpy()
    Lr, info = orc.potrf(A)
    assert info == 0
    xr = orc.potrs(Lr, B)
    assert np.array_equal(x.view(np.uint64), xr.view(np.uint64))
