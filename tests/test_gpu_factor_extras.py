"""GPU parity for the factor methods around the solvers:
Cholesky::new_with_substitute, Cholesky::inverse, Cholesky::determinant and
ln_determinant, LU::determinant (src/linalg/cholesky.rs:146-218,
src/linalg/lu.rs:301-314), and Q B / R B on nalgebra's QR factors (the roles
of nalgebra-lapack's q_mul_mut / r_mul_mut, qr/abstraction.rs:83-146).

Bar: bitwise against the oracle, except ln_determinant, whose logarithm is
the device's (within 1 ulp per term of libm's): 1e-14 relative.
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


def _known():
    with open(os.path.join(GOLD, "reference_known_answers.json")) as f:
        return json.load(f)


def _rowmajor(rows):
    return np.array(rows, dtype=np.float64)


def test_substitute_known_answer(nl, orc):
    k = _known()["cholesky_with_substitute"]
    m = _rowmajor(k["m"])
    assert nl.DMatrix.from_numpy(m).cholesky() is None
    ch = nl.Cholesky.new_with_substitute(nl.DMatrix.from_numpy(m), k["substitute"])
    assert ch is not None
    L = ch.l_dirty().numpy()
    Lr, info = orc.potrf(m, substitute=k["substitute"])
    assert info == 0
    assert np.array_equal(_bits(L), _bits(Lr))          # the NaN upper entry stays as it was


@pytest.mark.parametrize("n", [5, 40, 64, 65, 200])
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_substitute_indefinite(nl, orc, n, mode):
    """A symmetric matrix with negative and zero pivots: every such pivot takes
    the substitute (n, so that L keeps the scale of its other columns); a
    substitute <= 0 or NaN leaves Cholesky::new's None."""
    G = gen.normal(300 + n, (n, n))
    A = G @ G.T + n * np.eye(n)
    for j in range(0, n, 7):
        A[j, j] = -1.0
    A[n // 2, n // 2] = 0.0
    m = nl.EXACT if mode == "exact" else nl.FAST
    sub = float(n)
    Lr, info = orc.potrf(A, substitute=sub)
    assert info == 0 and np.all(np.isfinite(np.tril(Lr)))
    ch = nl.Cholesky.new_with_substitute(nl.DMatrix.from_numpy(A), sub, m)
    assert ch is not None
    L = ch.l_dirty().numpy()
    if mode == "exact":
        assert np.array_equal(_bits(L), _bits(Lr))
    else:
        lo = np.tril(L); lr = np.tril(Lr)
        assert np.linalg.norm(lo - lr) / np.linalg.norm(lr) <= 1e-12
    for bad in (0.0, -2.0, float("nan")):
        assert nl.Cholesky.new_with_substitute(nl.DMatrix.from_numpy(A), bad, m) is None


@pytest.mark.parametrize("n", [1, 7, 64, 130])
def test_inverse_and_determinants(nl, orc, n):
    G = gen.normal(400 + n, (n, n))
    A = G @ G.T + n * np.eye(n)
    ch = nl.DMatrix.from_numpy(A).cholesky(nl.EXACT)
    L = ch.l_dirty().numpy()
    Lr, info = orc.potrf(A)
    assert info == 0 and np.array_equal(_bits(L), _bits(Lr))
    inv = ch.inverse().numpy()
    assert np.array_equal(_bits(inv), _bits(orc.potrs(Lr, np.eye(n))))   # identity + solve_mut
    p = orc.diag_product(Lr)
    assert ch.determinant() == p * p
    ld = orc.chol_ln_determinant(Lr)
    assert abs(ch.ln_determinant() - ld) <= 1e-14 * max(1.0, abs(ld))
    assert abs(ch.ln_determinant() - np.linalg.slogdet(A)[1]) <= 1e-10 * max(1.0, abs(ld))


def test_lu_determinant_on_device(nl, orc):
    k = _known()
    for name, det in (("lu_simple", 4.0), ("lu_simple_with_pivot", -4.0)):
        lu = nl.DMatrix.from_numpy(_rowmajor(k[name]["m"])).lu(nl.EXACT)
        assert lu.determinant() == det                     # assert_eq! in the reference tests
    A = gen.uniform(77, (50, 50))
    lu = nl.DMatrix.from_numpy(A).lu(nl.EXACT)
    LUr, ipr, _ = orc.getrf(A)
    assert lu.determinant() == orc.lu_determinant(LUr, ipr)


@pytest.mark.parametrize("shape", [(30, 8), (64, 64), (300, 40), (9, 13)])
def test_qr_qmul_bitwise(nl, orc, shape):
    m, n = shape
    A = gen.uniform(500 + m, shape)
    B = gen.uniform(600 + n, (m, 3))
    qr = nl.DMatrix.from_numpy(A).qr()
    QRr, dr = orc.geqrf(A)
    Bg = nl.DMatrix.from_numpy(B)
    qr.q_mul(Bg)
    assert np.array_equal(_bits(Bg.numpy()), _bits(orc.qr_qmul(QRr, dr, B)))
    # Q^T (Q B) returns B to rounding
    qr.q_tr_mul(Bg)
    assert np.allclose(Bg.numpy(), B, atol=1e-13)


@pytest.mark.parametrize("shape", [(30, 8), (64, 64), (700, 300)])
def test_qr_rmul_bitwise(nl, orc, shape):
    m, n = shape
    A = gen.uniform(700 + m, shape)
    B = gen.uniform(800 + n, (n, 5))
    B[3, :] = 0.0
    B[5, 1] = -0.0
    qr = nl.DMatrix.from_numpy(A).qr()
    QRr, dr = orc.geqrf(A)
    Bg = nl.DMatrix.from_numpy(B)
    qr.r_mul(Bg)
    assert np.array_equal(_bits(Bg.numpy()), _bits(orc.qr_rmul(QRr, dr, B)))


def test_qr_rmul_rejects_wide(nl):
    qr = nl.DMatrix.from_numpy(gen.uniform(1, (4, 6))).qr()
    with pytest.raises(ValueError):
        qr.r_mul(nl.DMatrix.from_numpy(np.ones((6, 1))))
