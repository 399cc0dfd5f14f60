"""LU::solve_mut (src/linalg/lu.rs:242-260) at large n: the swaps, the
unit-lower forward sweep (solve.rs:106-133) and the upper backward sweep
(solve.rs:137-182), one 64-row block per workgroup (n >= 1024).

EXACT is bitwise against the oracle's restatement of nalgebra's loops on the
same factors; FAST within 1e-12 relative.  Factors are synthetic (any packed
unit-L \\ U with pivots ipiv[i] >= i is a valid input of the solve), so the
largest sizes need no O(n^3) factorisation; one case factors with the GPU's
exact LU first."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

EXACT, FAST = 0, 1


def synthetic_factors(n, seed):
    """Packed L \\ U with |L| <= 1/2 below a unit diagonal and a dominant U
    diagonal (a well-conditioned solve), plus LAPACK-style pivots ipiv[i] >= i."""
    rng = np.random.default_rng(seed)
    LU = rng.uniform(-0.5, 0.5, (n, n)) / np.sqrt(n)
    LU[np.arange(n), np.arange(n)] = rng.uniform(1.0, 2.0, n) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    ipiv = np.array([rng.integers(i, n) if rng.random() < 0.7 else i for i in range(n)], dtype=np.int32)
    return np.asfortranarray(LU), ipiv


def gpu_getrs(nl, LU, ipiv, B, mode, noalloc=False):
    from nalgebra_amd._lib import lib
    from nalgebra_amd.linalg import handle
    n, nrhs = B.shape
    d_lu = torch.from_numpy(np.asfortranarray(LU).ravel(order="F").copy()).cuda()
    d_ip = torch.from_numpy(ipiv).cuda()
    d_b = torch.from_numpy(np.asfortranarray(B).ravel(order="F").copy()).cuda()
    info = torch.full((1,), 77, dtype=torch.int32, device="cuda")
    fn = lib().nrocm_dgetrs_noalloc if noalloc else lib().nrocm_dgetrs
    st = fn(handle(), n, nrhs, C.c_void_p(d_lu.data_ptr()), max(n, 1), C.c_void_p(d_ip.data_ptr()),
            C.c_void_p(d_b.data_ptr()), max(n, 1), C.c_void_p(info.data_ptr()), mode)
    assert st == 0, st
    return d_b.cpu().numpy().reshape(nrhs, n).T, int(info.item())


@pytest.mark.parametrize("n,nrhs", [(1024, 1), (1025, 3), (2000, 2), (4113, 1), (8192, 1), (16384, 1)])
def test_getrs_large_exact(nl, orc, n, nrhs):
    LU, ipiv = synthetic_factors(n, 500 + n)
    B = np.random.default_rng(n).standard_normal((n, nrhs))
    print(f"getrs n={n} nrhs={nrhs}", flush=True)
    x, info = gpu_getrs(nl, LU, ipiv, B, EXACT)
    xr, ok = orc.getrs(LU, ipiv, B)
    assert ok and info == 0
    assert np.array_equal(x.view(np.uint64), xr.view(np.uint64))


@pytest.mark.parametrize("n", [1024, 3000])
def test_getrs_large_fast(nl, orc, n):
    LU, ipiv = synthetic_factors(n, 700 + n)
    B = np.random.default_rng(n).standard_normal((n, 2))
    x, info = gpu_getrs(nl, LU, ipiv, B, FAST)
    xr, _ = orc.getrs(LU, ipiv, B)
    assert info == 0
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 1e-12


def test_getrs_large_zero_diagonal(nl, orc):
    """solve_mut returns false at the first exact zero the backward sweep
    meets, i.e. the largest i with U_ii == 0 (solve.rs:160-165)."""
    n = 1500
    LU, ipiv = synthetic_factors(n, 9)
    LU[700, 700] = 0.0
    LU[1300, 1300] = -0.0
    _, info = gpu_getrs(nl, LU, ipiv, np.ones((n, 1)), EXACT)
    _, ok = orc.getrs(LU, ipiv, np.ones((n, 1)))
    assert not ok and info == 1301


def test_getrf_getrs_exact_2048(nl, orc):
    """GPU LU::new (exact) then the large-n solve, against the oracle's
    LU::new + solve_mut of the same matrix: identical pivots and bits."""
    import gen
    n = 2048
    A = gen.uniform(2048, (n, n))
    b = gen.uniform(2049, (n, 1))
    lu = nl.DMatrix.from_numpy(A).lu(nl.EXACT)
    x = nl.DMatrix.from_numpy(b)
    assert lu.solve_mut(x)
    LUr, ipr, _ = orc.getrf(A)
    xr, _ = orc.getrs(LUr, ipr, b)
    assert np.array_equal(lu.ipiv.cpu().numpy(), ipr)
    assert np.array_equal(x.numpy().view(np.uint64), xr.view(np.uint64))
