"""Pins the CPU oracle before anything is checked against it.

The reference (Rust) cannot be built here, so the oracle is pinned by:
  * the reference's own known-answer tests (tests/linalg/lu.rs:5-39,
    tests/linalg/inverse.rs:63-81, tests/linalg/cholesky.rs:5-12,182-218),
    carried as data in tests/golden/reference_known_answers.json;
  * LAPACK via numpy/scipy on the same inputs (dgetrf uses the same
    first-maximum pivot rule as nalgebra's icamax);
  * exact rational arithmetic (fractions.Fraction) on small cases;
  * inputs on which every operation is exact (dyadic entries), where any
    correct algorithm must return the exact factors bit for bit;
  * hand-derived rounding cases for the 8-accumulator dot and icamax.
Each check is chosen so that a dropped term, a wrong sign or index, or a
transposed operand in the oracle fails it.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg as sla

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _known():
    with open(os.path.join(GOLD, "reference_known_answers.json")) as f:
        return json.load(f)


def _rowmajor(rows):
    return np.array(rows, dtype=np.float64)


# ------------------------------------------------------------------ dot / axpy / icamax

def test_dot_accumulator_assignment(orc):
    # element 8 joins accumulator 0 before acc0+acc4 is formed: 1e16 - 1e16 = 0, + 1
    a = np.zeros(16); a[0] = 1e16; a[4] = 1.0; a[8] = -1e16
    assert orc.dot(a, np.ones(16)) == 1.0          # a sequential sum gives 0.0


def test_dot_combine_order(orc):
    # res = 0 + (acc0+acc4) = fl(1e16+1) = 1e16 ; res += (acc1+acc5) = fl(-1e16+1) = -1e16
    a = np.zeros(8); a[0] = 1e16; a[4] = 1.0; a[1] = -1e16; a[5] = 1.0
    assert orc.dot(a, np.ones(8)) == 0.0           # the exact sum (and a sequential sum) is 2.0


def test_dot_tail_added_last(orc):
    # len 9: the 9th product is added after the 4 combined pairs
    a = np.zeros(9); a[0] = 1e16; a[8] = 1.0; a[1] = -1e16
    # combined: fl(1e16 + 0) + fl(-1e16 + 0) = 0, then + 1 = 1
    assert orc.dot(a, np.ones(9)) == 1.0
    a = np.zeros(9); a[0] = 1e16; a[8] = 1.0
    assert orc.dot(a, np.ones(9)) == 1e16          # fl(1e16 + 1) = 1e16 (tie to even)


def test_dot_short_and_empty(orc):
    assert orc.dot(np.zeros(0), np.zeros(0)) == 0.0
    x = np.array([3.0, -4.0, 0.5])
    assert orc.dot(x, x) == 25.25


def test_axcpy_beta_zero_does_not_read_y(orc):
    y = np.full(3, np.nan)
    out = orc.axcpy(y, 2.0, np.array([1.0, 2.0, 3.0]), 0.5, 0.0)
    assert np.array_equal(out, [1.0, 2.0, 3.0])
    out = orc.axcpy(np.array([1.0, 1.0]), -1.0, np.array([2.0, 4.0]), 1.0, 1.0)
    assert np.array_equal(out, [-1.0, -3.0])


def test_icamax_rules(orc):
    assert orc.icamax(np.array([1.0, -3.0, 3.0, 2.0])) == 1     # ties keep the lowest index
    assert orc.icamax(np.array([0.0, -0.0, 0.0])) == 0
    assert orc.icamax(np.array([np.nan, 5.0])) == 0             # NaN first is never beaten
    assert orc.icamax(np.array([1.0, np.nan, 5.0])) == 2        # a later NaN never wins


# ------------------------------------------------------------------ gemm

def test_gemm_small_integers_exact(orc):
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 9, size=(13, 7)).astype(np.float64)
    B = rng.integers(-8, 9, size=(7, 11)).astype(np.float64)
    Cm = rng.integers(-8, 9, size=(13, 11)).astype(np.float64)
    assert np.array_equal(orc.gemm(1.0, A, B), A @ B)
    assert np.array_equal(orc.gemm(2.0, A, B, -3.0, Cm), 2.0 * (A @ B) - 3.0 * Cm)
    assert np.array_equal(orc.gemm(1.0, A, B, 0.0, np.full((13, 11), np.nan)), A @ B)


def test_gemm_simple_mul_known_answer(orc):
    k = _known()["simple_mul"]
    got = orc.gemm(1.0, _rowmajor(k["a"]), _rowmajor(k["b"]))
    assert np.array_equal(got, _rowmajor(k["expected"]))


def test_gemm_vs_numpy(orc):
    rng = np.random.default_rng(1)
    A = rng.uniform(-1, 1, (67, 45)); B = rng.uniform(-1, 1, (45, 31))
    ref = A @ B
    assert np.linalg.norm(orc.gemm(1.0, A, B) - ref) / np.linalg.norm(ref) < 1e-14
    As = A.astype(np.float32); Bs = B.astype(np.float32)
    got = orc.gemm(np.float32(1.0), As, Bs)
    assert got.dtype == np.float32
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-6


# ------------------------------------------------------------------ LU

@pytest.mark.parametrize("name,det", [("lu_simple", 4.0), ("lu_simple_with_pivot", -4.0)])
def test_lu_known_determinants(orc, name, det):
    m = _rowmajor(_known()[name]["m"])
    LU, ipiv, info = orc.getrf(m)
    assert info == 0
    assert orc.lu_determinant(LU, ipiv) == det      # assert_eq! in the reference
    L = np.tril(LU, -1) + np.eye(3); U = np.triu(LU)
    P = np.eye(3)
    for i, p in orc.perm_sequence(ipiv):
        P[[i, p]] = P[[p, i]]
    assert np.allclose(P.T @ (L @ U), m, atol=1e-7)


def test_lu_matrix5_try_inverse(orc):
    k = _known()["matrix5_try_inverse"]
    a = _rowmajor(k["a"])
    LU, ipiv, info = orc.getrf(a)
    inv, ok = orc.getrs(LU, ipiv, np.eye(5))
    assert ok and info == 0
    exp = _rowmajor(k["expected_inverse"])
    # approx::relative_eq with max_relative = 1e-4 (tests/linalg/inverse.rs:80)
    diff = np.abs(inv - exp)
    tol = np.maximum(1e-4 * np.maximum(np.abs(inv), np.abs(exp)), np.finfo(float).eps)
    assert np.all(diff <= tol)


def _lu_dyadic(n, rng, perm):
    """A = P^T L U with |L_ij| < 1 dyadic and power-of-two diagonal of U, so
    partial pivoting must find P and every rounding is exact."""
    L = np.tril(rng.choice([-0.75, -0.5, -0.25, 0.0, 0.25, 0.5, 0.75], size=(n, n)), -1) + np.eye(n)
    U = np.triu(rng.choice([-2.0, -1.0, 0.0, 1.0, 2.0], size=(n, n)), 1)
    U[np.arange(n), np.arange(n)] = rng.choice([-8.0, -4.0, 4.0, 8.0], size=n)
    A = (L @ U)[np.argsort(perm)]
    return A, L, U


def test_lu_exact_on_dyadic_inputs(orc):
    rng = np.random.default_rng(7)
    for n in (4, 9, 16):
        perm = np.arange(n)
        A, L, U = _lu_dyadic(n, rng, perm)
        LU, ipiv, info = orc.getrf(A)
        assert info == 0
        assert np.array_equal(np.tril(LU, -1) + np.eye(n), L)
        assert np.array_equal(np.triu(LU), U)
        assert np.array_equal(ipiv, np.arange(n))


def _exact_lu(A):
    """Partial-pivot LU in exact rational arithmetic, same pivot rule."""
    n = A.shape[0]
    M = [[Fraction(float(A[r, c])) for c in range(n)] for r in range(n)]
    piv = []
    for i in range(n):
        p = max(range(i, n), key=lambda r: (abs(M[r][i]), -r))
        piv.append(p)
        M[i], M[p] = M[p], M[i]
        for r in range(i + 1, n):
            M[r][i] = M[r][i] / M[i][i]
            for c in range(i + 1, n):
                M[r][c] -= M[r][i] * M[i][c]
    return np.array([[float(v) for v in row] for row in M]), np.array(piv)


def test_lu_vs_exact_rationals(orc):
    rng = np.random.default_rng(11)
    for n in (2, 3, 5, 6):
        for _ in range(5):
            A = rng.uniform(-1, 1, (n, n))
            LU, ipiv, _ = orc.getrf(A)
            ex, piv = _exact_lu(A)
            assert np.array_equal(ipiv, piv)
            assert np.max(np.abs(LU - ex)) <= 1e-13 * max(1.0, np.max(np.abs(ex)))


@pytest.mark.parametrize("n", [64, 200, 512])
def test_lu_pivots_match_lapack(orc, n):
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n))
    LU, ipiv, info = orc.getrf(A)
    lu_ref, piv_ref = sla.lu_factor(A)
    assert info == 0
    assert np.array_equal(ipiv, piv_ref)
    assert np.linalg.norm(LU - lu_ref) / np.linalg.norm(lu_ref) < 1e-12
    b = rng.uniform(-1, 1, n)
    x, ok = orc.getrs(LU, ipiv, b)
    assert ok
    xr = sla.lu_solve((lu_ref, piv_ref), b)
    assert np.linalg.norm(x[:, 0] - xr) / np.linalg.norm(xr) < 1e-11


def test_lu_rectangular_and_singular(orc):
    rng = np.random.default_rng(3)
    for shape in ((9, 5), (5, 9)):
        A = rng.uniform(-1, 1, shape)
        LU, ipiv, info = orc.getrf(A)
        lu_ref, piv_ref = sla.lu_factor(A)
        assert np.array_equal(ipiv, piv_ref)
        assert np.allclose(LU, lu_ref, rtol=0, atol=1e-13)
    # a zero column skips the step (lu.rs:107-110), records no swap, solve reports failure
    A = rng.uniform(-1, 1, (6, 6)); A[:, 2] = 0.0
    LU, ipiv, info = orc.getrf(A)
    assert info == 3 and ipiv[2] == 2
    assert np.all(LU[:, 2][3:] == 0.0)
    _, ok = orc.getrs(LU, ipiv, np.ones(6))
    assert not ok


# ------------------------------------------------------------------ Cholesky

def test_cholesky_with_substitute_known_answer(orc):
    k = _known()["cholesky_with_substitute"]
    m = _rowmajor(k["m"])
    _, info = orc.potrf(m)
    assert info == 2                                      # None: the NaN upper entry is never read
    L, info = orc.potrf(m, substitute=k["substitute"])
    assert info == 0
    assert L[1, 1] == math.sqrt(1e-8) and L[1, 0] == 1.0 and L[0, 0] == 1.0
    assert np.isnan(L[0, 1])                              # upper triangle left as it was


def test_cholesky_non_pd_returns_none(orc):
    _, info = orc.potrf(_rowmajor(_known()["cholesky_non_pd"]["m"]))
    assert info == 1
    _, info = orc.potrf(np.array([[1.0, 2.0], [2.0, 1.0]]))
    assert info == 2


def test_cholesky_closed_form_2x2(orc):
    L, info = orc.potrf(np.array([[4.0, 99.0], [2.0, 5.0]]))     # upper entry ignored
    assert info == 0
    assert np.array_equal(L, [[2.0, 99.0], [1.0, 2.0]])
    a, b, c = 2.0, 0.3, 1.7
    L, _ = orc.potrf(np.array([[a, b], [b, c]]))
    d0 = math.sqrt(a); l10 = b / d0
    assert L[0, 0] == d0 and L[1, 0] == l10 and L[1, 1] == math.sqrt(c - l10 * l10)


@pytest.mark.parametrize("n", [1, 5, 33, 64])
def test_cholesky_vs_lapack(orc, n):
    rng = np.random.default_rng(100 + n)
    G = rng.standard_normal((n, n))
    A = G @ G.T + n * np.eye(n)
    L, info = orc.potrf(A)
    assert info == 0
    Lr = np.linalg.cholesky(A)
    assert np.linalg.norm(np.tril(L) - Lr) / np.linalg.norm(Lr) < 1e-14
    B = rng.standard_normal((n, 4))
    X = orc.potrs(L, B)
    Xr = np.linalg.solve(A, B)
    assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) < 1e-13
    assert np.allclose(A @ X, B, atol=1e-7)              # tests/linalg/cholesky.rs:51-77


def test_cholesky_solve_exact_on_dyadic(orc):
    # L with power-of-two diagonal and dyadic off-diagonals: every step is exact
    L = np.array([[2.0, 0, 0, 0], [1.0, 4.0, 0, 0], [-0.5, 2.0, 1.0, 0], [3.0, -1.0, 0.5, 2.0]])
    A = L @ L.T
    Lo, info = orc.potrf(A)
    assert info == 0 and np.array_equal(np.tril(Lo), L)
    x = np.array([1.0, -2.0, 0.5, 4.0])
    X = orc.potrs(Lo, (A @ x).reshape(-1, 1))
    assert np.array_equal(X[:, 0], x)


# ------------------------------------------------------------------ triangular solves (via LU)

def test_lu_solve_multiple_rhs(orc):
    rng = np.random.default_rng(5)
    A = rng.uniform(-1, 1, (4, 4)); B = rng.uniform(-1, 1, (4, 5))
    LU, ipiv, _ = orc.getrf(A)
    X, ok = orc.getrs(LU, ipiv, B)
    assert ok and np.allclose(A @ X, B, atol=1e-7)


# ------------------------------------------------------------------ QR

@pytest.mark.parametrize("shape", [(5, 3), (3, 5), (4, 4), (40, 12)])
def test_qr_invariants_and_lapack(orc, shape):
    rng = np.random.default_rng(sum(shape))
    A = rng.uniform(-1, 1, shape)
    QR, diag = orc.geqrf(A)
    Q = orc.qr_q(QR, diag); R = orc.qr_r(QR, diag)
    assert np.allclose(Q @ R, A, atol=1e-12)                       # tests/linalg/qr.rs:15-24
    k = min(shape)
    assert np.allclose(Q.T @ Q, np.eye(k), atol=1e-12)
    assert np.all(np.diag(R) > 0)                                  # qr.rs:87
    Rr = np.linalg.qr(A, mode="r")
    Rr = Rr * np.sign(np.diag(Rr))[:, None]
    assert np.allclose(R, Rr, atol=1e-12)
    # unit-norm axes stored from the diagonal down
    for i in range(k):
        assert abs(np.linalg.norm(QR[i:, i]) - 1.0) < 1e-14


def test_qr_qtmul_matches_q(orc):
    rng = np.random.default_rng(21)
    A = rng.uniform(-1, 1, (30, 8)); B = rng.uniform(-1, 1, (30, 3))
    QR, diag = orc.geqrf(A)
    QtB = orc.qr_qtmul(QR, diag, B)
    Q = orc.qr_q(QR, diag)
    assert np.allclose(QtB[:8], Q.T @ B, atol=1e-13)


def test_qr_qmul_pins(orc):
    """Q B: on the identity it rebuilds QR::q (value for value; q() leaves the
    entries left of each reflector alone, so only zero signs may differ),
    it inverts q_tr_mul, and Q [R; 0] reproduces A."""
    rng = np.random.default_rng(22)
    for shape in ((30, 8), (6, 6), (5, 9)):
        m, n = shape
        A = rng.uniform(-1, 1, shape)
        QR, diag = orc.geqrf(A)
        k = min(shape)
        QI = orc.qr_qmul(QR, diag, np.eye(m))
        assert np.array_equal(QI[:, :k], orc.qr_q(QR, diag))
        B = rng.uniform(-1, 1, (m, 3))
        assert np.allclose(orc.qr_qmul(QR, diag, orc.qr_qtmul(QR, diag, B)), B, atol=1e-14)
        R = orc.qr_r(QR, diag)
        Rpad = np.zeros((m, n)); Rpad[:k] = R
        assert np.allclose(orc.qr_qmul(QR, diag, Rpad), A, atol=1e-14)


def test_qr_rmul_pins(orc):
    """R B in TRMM order: exact on small integers (any order is), the skipped
    zero keeps its sign, and real inputs agree with numpy's product."""
    rng = np.random.default_rng(23)
    n = 9
    QR = rng.integers(-5, 6, (12, n)).astype(np.float64)
    diag = rng.integers(-5, 6, n).astype(np.float64)
    B = rng.integers(-5, 6, (n, 4)).astype(np.float64)
    R = np.triu(QR[:n]); R[np.arange(n), np.arange(n)] = np.abs(diag)
    assert np.array_equal(orc.qr_rmul(QR, diag, B), R @ B)
    # a zero b_k contributes nothing: -0.0 stays -0.0
    got = orc.qr_rmul(np.array([[7.0, 3.0], [1.0, 7.0], [0.0, 0.0]]), np.array([2.0, -5.0]),
                      np.array([[-0.0], [0.0]]))
    assert np.signbit(got[0, 0]) and not np.signbit(got[1, 0])
    A = rng.uniform(-1, 1, (40, 12))
    QR, diag = orc.geqrf(A)
    B = rng.uniform(-1, 1, (12, 3))
    ref = orc.qr_r(QR, diag) @ B
    assert np.linalg.norm(orc.qr_rmul(QR, diag, B) - ref) / np.linalg.norm(ref) < 1e-15


def test_cholesky_determinants_closed_form(orc):
    L = np.array([[2.0, 0.0, 0.0], [0.3, 0.5, 0.0], [-1.0, 4.0, 3.0]])
    assert orc.diag_product(L) == 3.0                        # 1 * 2 * 0.5 * 3
    assert orc.diag_product(L) ** 2 == 9.0                   # Cholesky::determinant
    assert abs(orc.chol_ln_determinant(L) - math.log(9.0)) < 1e-15
    rng = np.random.default_rng(24)
    G = rng.standard_normal((20, 20)); A = G @ G.T + 20 * np.eye(20)
    Lf, info = orc.potrf(A)
    assert info == 0
    sign, logdet = np.linalg.slogdet(A)
    assert sign == 1.0 and abs(orc.chol_ln_determinant(Lf) - logdet) < 1e-12 * abs(logdet)
    assert abs(orc.diag_product(Lf) ** 2 / np.linalg.det(A) - 1.0) < 1e-12


def test_qr_square_solve(orc):
    rng = np.random.default_rng(22)
    A = rng.uniform(-1, 1, (9, 9)); b = rng.uniform(-1, 1, (9, 2))
    QR, diag = orc.geqrf(A)
    y = orc.qr_qtmul(QR, diag, b)
    x, ok = orc.qr_rsolve(QR, diag, y)
    assert ok and np.allclose(A @ x, b, atol=1e-10)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 2e-3)])
def test_qr_least_squares(orc, dtype, tol):
    rng = np.random.default_rng(23)
    m, n = 300, 64
    A = rng.standard_normal((m, n)); xt = rng.standard_normal(n)
    b = A @ xt + 1e-3 * rng.standard_normal(m)
    x, QR, diag, qtb, ok = orc.qr_ls(A.astype(dtype), b.astype(dtype))
    assert ok and x.dtype == dtype
    xr = np.linalg.lstsq(A, b, rcond=None)[0]
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < tol
    # normal equations (nalgebra-lapack/tests/linalg/qr.rs:45-61)
    r = A.T @ (A @ x.astype(np.float64) - b)
    assert np.linalg.norm(r) / (np.linalg.norm(A) ** 2 * np.linalg.norm(x)) < tol


# ------------------------------------------------------------------ adjoint

def test_adjoint_closed_form_and_fd(orc):
    rng = np.random.default_rng(4)
    n = 8
    G = rng.standard_normal((n, n)); A = G @ G.T + n * np.eye(n); b = rng.standard_normal(n)

    def loss(Am, bv):
        L, info = orc.potrf(Am)
        assert info == 0
        x = orc.potrs(L, bv)[:, 0]
        return float(np.sum(x * x)), L, x

    f0, L, x = loss(A, b)
    dA, db = orc.potrs_adjoint(L, x, 2 * x)
    lam = np.linalg.solve(A, 2 * x)
    assert np.allclose(db, lam, rtol=1e-12)
    assert np.allclose(dA, -0.5 * (np.outer(lam, x) + np.outer(x, lam)), rtol=1e-12)
    h = 1e-6
    for (i, j) in [(0, 0), (3, 1), (7, 7), (5, 2)]:
        Ap = A.copy(); Am = A.copy()
        Ap[i, j] += h; Am[i, j] -= h        # lower entry: perturbs the symmetric matrix at (i,j),(j,i)
        fd = (loss(Ap, b)[0] - loss(Am, b)[0]) / (2 * h)
        an = dA[i, j] * (1 if i == j else 2)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
    for i in (0, 4):
        bp = b.copy(); bm = b.copy(); bp[i] += h; bm[i] -= h
        fd = (loss(A, bp)[0] - loss(A, bm)[0]) / (2 * h)
        assert abs(fd - db[i]) <= 1e-6 * max(1.0, abs(db[i]))


def test_lu_adjoint_closed_form_and_fd(orc):
    """Adjoint of x = A^{-1} b through LU::new's factors: dB = A^{-T} g and
    dA = -dB x^T against the closed form (numpy, 1e-12) and central finite
    differences of loss = sum(g * x) in every entry of A and b."""
    rng = np.random.default_rng(14)
    for n, nrhs in [(1, 1), (5, 2), (9, 1), (33, 3)]:
        A = rng.standard_normal((n, n)) + 0.5 * np.eye(n)
        B = rng.standard_normal((n, nrhs)); g = rng.standard_normal((n, nrhs))

        def loss(Am, Bm):
            LU, ip, _ = orc.getrf(Am)
            X, ok = orc.getrs(LU, ip, Bm)
            assert ok
            return float(np.sum(g * X)), LU, ip, X

        f0, LU, ip, X = loss(A, B)
        dA, dB, ok = orc.getrs_adjoint(LU, ip, X, g)
        assert ok
        lam = np.linalg.solve(A.T, g)
        assert np.allclose(dB, lam, rtol=1e-12, atol=1e-13)
        assert np.allclose(dA, -lam @ X.T, rtol=1e-12, atol=1e-13)
        if n <= 9:
            h = 1e-6
            for i in range(n):
                for j in range(n):
                    Ap = A.copy(); Am = A.copy(); Ap[i, j] += h; Am[i, j] -= h
                    fd = (loss(Ap, B)[0] - loss(Am, B)[0]) / (2 * h)
                    assert abs(fd - dA[i, j]) <= 1e-6 * max(1.0, abs(dA[i, j]))
                for q in range(nrhs):
                    Bp = B.copy(); Bm = B.copy(); Bp[i, q] += h; Bm[i, q] -= h
                    fd = (loss(A, Bp)[0] - loss(A, Bm)[0]) / (2 * h)
                    assert abs(fd - dB[i, q]) <= 1e-6 * max(1.0, abs(dB[i, q]))


def test_lu_adjoint_exact_dyadic(orc):
    """Factors whose every operation is exact (small integers, unit L, U
    diagonal of +-1 and +-2, pivots ipiv[i] >= i): the adjoint must equal the
    exact rational result bit for bit, which pins the transposed solves'
    algebra and the direction of the permutation (P^T, swaps in reverse)."""
    from fractions import Fraction as Fr
    rng = np.random.default_rng(15)
    n = 6
    L = np.tril(rng.integers(-1, 2, (n, n)).astype(float), -1) + np.eye(n)
    U = np.triu(rng.integers(-2, 3, (n, n)).astype(float), 1)
    U[np.arange(n), np.arange(n)] = [1.0, -2.0, 2.0, -1.0, 2.0, 1.0]
    LU = np.tril(L, -1) + U
    ipiv = np.array([2, 1, 5, 4, 4, 5], dtype=np.int32)
    X = rng.integers(-3, 4, (n, 2)).astype(float)
    g = rng.integers(-3, 4, (n, 2)).astype(float)
    dA, dB, ok = orc.getrs_adjoint(LU, ipiv, X, g)
    assert ok
    # exact: A = P^{-1} L U with P = swaps (i, ipiv[i]) applied in order
    P = np.eye(n)
    for i, p in enumerate(ipiv):
        P[[i, p]] = P[[p, i]]
    A = np.linalg.inv(P) @ L @ U
    Af = [[Fr(int(round(v))) for v in row] for row in A]
    # solve A^T lam = g exactly by Gauss-Jordan on fractions
    for q in range(2):
        M = [[Af[j][i] for j in range(n)] + [Fr(int(g[i, q]))] for i in range(n)]
        for c in range(n):
            piv = next(r for r in range(c, n) if M[r][c] != 0)
            M[c], M[piv] = M[piv], M[c]
            M[c] = [v / M[c][c] for v in M[c]]
            for r in range(n):
                if r != c and M[r][c] != 0:
                    M[r] = [a - M[r][c] * b for a, b in zip(M[r], M[c])]
        lam = [M[i][n] for i in range(n)]
        assert [Fr(v) for v in dB[:, q]] == lam
    ref = -(dB @ X.T)
    assert np.array_equal(dA, ref)


def test_lu_adjoint_singular(orc):
    LU = np.triu(np.ones((4, 4))); LU[2, 2] = 0.0
    _, _, ok = orc.getrs_adjoint(LU, np.arange(4, dtype=np.int32), np.ones((4, 1)), np.ones((4, 1)))
    assert not ok
