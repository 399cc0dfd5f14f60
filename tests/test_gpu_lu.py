"""GPU parity for LU with partial pivoting + solve, and gemm.

Bar: NROCM_EXACT LU factors, pivots and solutions are bit-identical to the
oracle (nalgebra's LU::new / solve_mut order); gemm (whose reference,
matrixmultiply, has no pinned summation order) is within 1e-12 relative.
"""
import json
import os

import numpy as np
import pytest
import torch

import gen

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _known():
    with open(os.path.join(GOLD, "reference_known_answers.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("shape", [(1, 1), (2, 2), (7, 7), (64, 64), (65, 65), (200, 200), (512, 512),
                                   (300, 130), (130, 300), (1500, 1500)])
def test_lu_exact_bitwise(nl, orc, shape):
    m, n = shape
    print(f"lu exact {shape}", flush=True)
    A = gen.uniform(100 + m + n, shape)
    lu = nl.DMatrix.from_numpy(A).lu(nl.EXACT)
    LUr, ipr, infor = orc.getrf(A)
    k = min(m, n)
    assert np.array_equal(lu.ipiv[:k].cpu().numpy(), ipr), "pivot sequence differs"
    assert lu.info == infor
    assert np.array_equal(_bits(lu.lu.numpy()), _bits(LUr)), "LU bits differ"
    if m == n:
        b = gen.uniform(200 + n, (n, 2))
        x = nl.DMatrix.from_numpy(b)
        assert lu.solve_mut(x)
        xr, ok = orc.getrs(LUr, ipr, b)
        assert ok
        assert np.array_equal(_bits(x.numpy()), _bits(xr)), "solution bits differ"


def _lu_case(kind, shape):
    m, n = shape
    rng = np.random.default_rng(7 * m + n)
    if kind == "pm1":          # every first-column entry ties in magnitude
        return rng.choice([-1.0, 1.0], size=shape)
    if kind == "int2":         # entries in {-2..2}: ties in most pivot searches
        return rng.integers(-2, 3, size=shape).astype(np.float64)
    if kind == "zerocols":     # zero columns stay zero under elimination: zero pivots, info != 0
        A = gen.uniform(300 + m, shape)
        A[:, [5, n // 2, n - 3]] = 0.0
        return A
    return gen.uniform(400 + m + n, shape)


@pytest.mark.parametrize("kind,shape", [("pm1", (64, 64)), ("pm1", (300, 300)), ("pm1", (700, 700)),
                                        ("int2", (129, 129)), ("int2", (600, 600)), ("int2", (400, 250)),
                                        ("zerocols", (90, 90)), ("zerocols", (260, 260)),
                                        ("uniform", (500, 1)), ("uniform", (1, 500)),
                                        ("uniform", (2000, 70)), ("uniform", (70, 2000))])
def test_lu_exact_ties_and_shapes(nl, orc, kind, shape):
    """Pivot ties keep the first row of greatest magnitude (icamax,
    `src/base/min_max.rs:221-241`) and a zero pivot skips its step, with no
    swap and no elimination (`lu.rs:107-110`): pivots, factors and info
    identical to the oracle, bit for bit."""
    m, n = shape
    print(f"lu exact {kind} {shape}", flush=True)
    A = _lu_case(kind, shape)
    lu = nl.DMatrix.from_numpy(A).lu(nl.EXACT)
    LUr, ipr, infor = orc.getrf(A)
    k = min(m, n)
    assert np.array_equal(lu.ipiv[:k].cpu().numpy(), ipr), "pivot sequence differs"
    assert lu.info == infor
    assert np.array_equal(_bits(lu.lu.numpy()), _bits(LUr)), "LU bits differ"
    if kind == "zerocols":
        assert infor != 0
    if m == n and infor == 0:
        b = gen.uniform(500 + n, (n, 3))
        x = nl.DMatrix.from_numpy(b)
        assert lu.solve_mut(x)
        xr, ok = orc.getrs(LUr, ipr, b)
        assert ok
        assert np.array_equal(_bits(x.numpy()), _bits(xr)), "solution bits differ"


@pytest.mark.parametrize("shape,where", [((40, 40), "pivot"), ((40, 40), "below"), ((90, 90), "row"),
                                         ((200, 200), "below"), ((300, 300), "pivot")])
def test_lu_exact_nan_bits(nl, orc, shape, where):
    """NaN inputs: LU::new never fails, so its NaN factors are observable.  A
    NaN at the pivot position wins the search (min_max.rs:227-240), other
    NaNs never do: pivots, NaN positions and every other bit match the
    oracle.  The sign of a NaN produced from two NaN operands is left open
    (x86 returns the first operand, the GPU does not), so NaNs compare as NaNs."""
    m, n = shape
    A = gen.uniform(900 + m, shape)
    if where == "pivot":
        A[0, 0] = np.nan
    elif where == "below":
        A[m // 2, 3] = np.nan
    else:
        A[5, :] = np.nan
    lu = nl.DMatrix.from_numpy(A).lu(nl.EXACT)
    LUr, ipr, infor = orc.getrf(A)
    assert np.array_equal(lu.ipiv[:min(m, n)].cpu().numpy(), ipr)
    G = lu.lu.numpy()
    assert np.array_equal(np.isnan(G), np.isnan(LUr)), "NaN positions differ"
    ok = ~np.isnan(LUr)
    assert np.isnan(LUr).any() and np.array_equal(_bits(G[ok]), _bits(LUr[ok])), "LU bits differ"


def test_config1_gemm_lu_solve(nl, orc):
    """f64 512x512, entries U[-1,1]: C = A*B, LU of A, solve A x = b."""
    n = 512
    A = gen.uniform(1, (n, n)); B = gen.uniform(2, (n, n)); b = gen.uniform(3, (n, 1))
    dA, dB = nl.DMatrix.from_numpy(A), nl.DMatrix.from_numpy(B)
    Cg = (dA @ dB).numpy()
    Cr = orc.gemm(1.0, A, B)
    assert _rel(Cg, Cr) <= 1e-12
    lu = dA.lu(nl.EXACT)
    x = nl.DMatrix.from_numpy(b)
    assert lu.solve_mut(x)
    LUr, ipr, _ = orc.getrf(A)
    xr, _ = orc.getrs(LUr, ipr, b)
    assert np.array_equal(lu.ipiv.cpu().numpy(), ipr)
    assert np.array_equal(_bits(x.numpy()), _bits(xr))
    # LU of C too (C is reported in the reference workload as well)
    luc = nl.DMatrix.from_numpy(Cr).lu(nl.EXACT)
    LUc, ipc, _ = orc.getrf(Cr)
    assert np.array_equal(luc.ipiv.cpu().numpy(), ipc)
    assert np.array_equal(_bits(luc.lu.numpy()), _bits(LUc))


@pytest.mark.parametrize("name,det", [("lu_simple", 4.0), ("lu_simple_with_pivot", -4.0)])
def test_lu_known_determinants(nl, name, det):
    m = np.array(_known()[name]["m"])
    assert nl.DMatrix.from_numpy(m).lu().determinant() == det


def test_lu_matrix5_try_inverse(nl):
    k = _known()["matrix5_try_inverse"]
    inv = nl.DMatrix.from_numpy(np.array(k["a"])).lu().try_inverse().numpy()
    exp = np.array(k["expected_inverse"])
    tol = np.maximum(1e-4 * np.maximum(np.abs(inv), np.abs(exp)), np.finfo(float).eps)
    assert np.all(np.abs(inv - exp) <= tol)


def test_lu_zero_column_and_singular_solve(nl, orc):
    A = gen.uniform(7, (70, 70)); A[:, 5] = 0.0
    lu = nl.DMatrix.from_numpy(A).lu()
    LUr, ipr, infor = orc.getrf(A)
    assert lu.info == infor == 6
    assert np.array_equal(lu.ipiv.cpu().numpy(), ipr)
    assert np.array_equal(_bits(lu.lu.numpy()), _bits(LUr))
    assert lu.solve(nl.DMatrix.from_numpy(np.ones(70))) is None


def test_lu_fast_mode(nl, orc):
    n = 700
    A = gen.uniform(8, (n, n)); b = gen.uniform(9, (n, 1))
    lu = nl.DMatrix.from_numpy(A).lu(nl.FAST)
    LUr, ipr, _ = orc.getrf(A)
    agree = float(np.mean(lu.ipiv.cpu().numpy() == ipr))
    print(f"FAST pivot agreement {agree:.4f}")
    assert agree > 0.99
    x = nl.DMatrix.from_numpy(b)
    assert lu.solve_mut(x)
    xv = x.numpy()
    berr = np.max(np.abs(b - A @ xv)) / (np.max(np.sum(np.abs(A), 1)) * np.max(np.abs(xv)) + np.max(np.abs(b)))
    assert berr < 1e-14


# ------------------------------------------------------------------ gemm

def test_gemm_small_integers_bitwise(nl):
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 9, size=(67, 45)).astype(np.float64)
    B = rng.integers(-8, 9, size=(45, 131)).astype(np.float64)
    C = nl.DMatrix.from_numpy(np.full((67, 131), np.nan))
    C.gemm(1.0, nl.DMatrix.from_numpy(A), nl.DMatrix.from_numpy(B), 0.0)   # beta = 0: C not read
    assert np.array_equal(C.numpy(), A @ B)
    C0 = rng.integers(-8, 9, size=(67, 131)).astype(np.float64)
    C = nl.DMatrix.from_numpy(C0)
    C.gemm(2.0, nl.DMatrix.from_numpy(A), nl.DMatrix.from_numpy(B), -3.0)
    assert np.array_equal(C.numpy(), 2.0 * (A @ B) - 3.0 * C0)


def test_gemm_known_answers(nl):
    k = _known()
    s = k["simple_mul"]
    got = (nl.DMatrix.from_numpy(np.array(s["a"])) @ nl.DMatrix.from_numpy(np.array(s["b"]))).numpy()
    assert np.array_equal(got, np.array(s["expected"]))
    d = k["gemm_doc"]
    a, b = np.array(d["a"]), np.array(d["b"])
    c = nl.DMatrix.from_numpy(np.eye(2, 4))
    c.gemm(d["alpha"], nl.DMatrix.from_numpy(a), nl.DMatrix.from_numpy(b), d["beta"])
    assert np.allclose(c.numpy(), a @ b * 10.0 + np.eye(2, 4) * 5.0, rtol=1e-15, atol=1e-15)


def test_gemm_strides_and_f32(nl, orc):
    from nalgebra_amd import _lib
    from nalgebra_amd.linalg import handle, _ptr
    m, k, n = 96, 70, 33
    A = gen.uniform(50, (m, k)); B = gen.uniform(51, (k, n))
    # A stored row-major (rsa = k, csa = 1), B column-major, C row-major
    Ad = torch.from_numpy(np.ascontiguousarray(A)).cuda().view(-1)
    Bd = torch.from_numpy(B.ravel(order="F").copy()).cuda()
    Cd = torch.zeros(m * n, dtype=torch.float64, device="cuda")
    h = handle()
    st = _lib.lib().nrocm_dgemm(h, m, k, n, 1.0, _ptr(Ad), k, 1, _ptr(Bd), 1, k, 0.0, _ptr(Cd), n, 1)
    assert st == 0
    assert _rel(Cd.cpu().numpy().reshape(m, n), A @ B) <= 1e-14
    As, Bs = A.astype(np.float32), B.astype(np.float32)
    Cs = (nl.DMatrix.from_numpy(As) @ nl.DMatrix.from_numpy(Bs)).numpy()
    assert Cs.dtype == np.float32
    assert _rel(Cs, As.astype(np.float64) @ Bs.astype(np.float64)) <= 1e-6


def test_gemm_large_vs_oracle(nl, orc):
    m, k, n = 777, 1030, 513
    A = gen.uniform(60, (m, k)); B = gen.uniform(61, (k, n))
    Cg = (nl.DMatrix.from_numpy(A) @ nl.DMatrix.from_numpy(B)).numpy()
    assert _rel(Cg, orc.gemm(1.0, A, B)) <= 1e-12
