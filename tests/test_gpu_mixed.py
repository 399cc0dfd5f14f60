"""Mixed-precision LU + FP64 iterative refinement (no reference counterpart;
checked against the FP64 oracle LU::new + solve_mut and LAPACK).

Bar: converged solutions have normwise backward error <= 1e-14 and a forward
error against the FP64 oracle solution within kappa * 1e-14."""
import numpy as np
import pytest

import gen

pytestmark = pytest.mark.gpu


def haar(seed, n):
    """Haar-random orthogonal matrix: QR of a Gaussian with the sign fix of
    Mezzadri (2007)."""
    q, r = np.linalg.qr(gen.normal(seed, (n, n)))
    return q * np.sign(np.diag(r))


def kappa_matrix(seed, n, kappa=1e4):
    U, V = haar(seed, n), haar(seed + 1, n)
    s = np.logspace(0, -np.log10(kappa), n)
    return (U * s) @ V.T


def berr(A, x, b):
    r = b - A @ x
    return np.max(np.abs(r)) / (np.max(np.sum(np.abs(A), axis=1)) * np.max(np.abs(x)) + np.max(np.abs(b)))


@pytest.mark.parametrize("n", [300, 1024])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_mixed_ir_converges(nl, orc, n, prec):
    A = kappa_matrix(70 + n, n)
    b = gen.normal(80 + n, (n,))
    p = {"fp64": nl.PREC_FP64, "fp32": nl.PREC_FP32, "bf16x3": nl.PREC_BF16X3}[prec]
    print(f"mixed IR n={n} {prec}", flush=True)
    x, iters, be, ok = nl.solve_mixed_ir(nl.DMatrix.from_numpy(A), nl.DMatrix.from_numpy(b.reshape(-1, 1)), p,
                                         tol=1e-14, maxit=30)
    xv = x.numpy()[:, 0]
    print(f"  iters={iters} berr={be:.3e} ok={ok}")
    assert ok and be <= 1e-14
    assert berr(A, xv, b) <= 2e-14
    LU, ip, _ = orc.getrf(A)
    xr, _ = orc.getrs(LU, ip, b)
    fe = np.max(np.abs(xv - xr[:, 0])) / np.max(np.abs(xr))
    assert fe <= 1e4 * 1e-14, fe
    if prec == "fp32":
        assert iters <= 6


@pytest.mark.parametrize("n", [300, 777, 1024])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_mixed_ir_window_converges(nl, orc, n, prec):
    """Pivots searched in the next 128 rows of each 64-column block only: same
    bars as the pivoted solver, and at most 8 refinement steps (the growth
    stays close to partial pivoting's on these matrices)."""
    A = kappa_matrix(370 + n, n)
    b = gen.normal(380 + n, (n,))
    p = {"fp64": nl.PREC_FP64, "fp32": nl.PREC_FP32, "bf16x3": nl.PREC_BF16X3}[prec] | nl.PREC_WINDOW
    print(f"window IR n={n} {prec}", flush=True)
    x, iters, be, ok = nl.solve_mixed_ir(nl.DMatrix.from_numpy(A), nl.DMatrix.from_numpy(b.reshape(-1, 1)), p,
                                         tol=1e-14, maxit=30)
    xv = x.numpy()[:, 0]
    print(f"  iters={iters} berr={be:.3e} ok={ok}")
    assert ok and be <= 1e-14
    assert berr(A, xv, b) <= 2e-14
    LU, ip, _ = orc.getrf(A)
    xr, _ = orc.getrs(LU, ip, b)
    fe = np.max(np.abs(xv - xr[:, 0])) / np.max(np.abs(xr))
    assert fe <= 1e4 * 1e-14, fe
    assert iters <= (2 if prec == "fp64" else 8)


@pytest.mark.parametrize("n", [641, 656, 777])
@pytest.mark.parametrize("window", [False, True])
def test_mixed_ir_short_last_block(nl, n, window):
    """A last 64-row block of 1, 16 or 9 rows: the backward triangular solve
    consumes that short block in every other block (regression: stale LDS
    slots past it were multiplied by zero, which gives NaN for a stale NaN)."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + n * np.eye(n)
    b = rng.standard_normal(n)
    p = nl.PREC_FP32 | (nl.PREC_WINDOW if window else 0)
    x, iters, be, ok = nl.solve_mixed_ir(nl.DMatrix.from_numpy(A), nl.DMatrix.from_numpy(b.reshape(-1, 1)), p,
                                         tol=1e-14, maxit=10)
    xv = x.numpy()[:, 0]
    assert np.all(np.isfinite(xv))
    assert ok and be <= 1e-14 and berr(A, xv, b) <= 2e-14


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 127, 128, 129, 200, 255, 256, 257, 320, 383, 511, 512, 513])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_mixed_ir_window_edge_sizes(nl, n, prec):
    """Windowed pivoting at sizes around one block (64 columns), one window (128
    rows), the first look-ahead steps and up to eight blocks with a short
    last one, on well-conditioned matrices whose largest entries sit off the
    diagonal (so the window has to pivot).  The windowed factors must do
    without the partial-pivoting fallback."""
    rng = np.random.default_rng(1000 + n)
    A = rng.standard_normal((n, n))
    A[np.arange(n), (np.arange(n) + 1) % n] += 4.0 * n ** 0.5  # dominant entries off the diagonal
    b = rng.standard_normal(n)
    p = {"fp64": nl.PREC_FP64, "fp32": nl.PREC_FP32, "bf16x3": nl.PREC_BF16X3}[prec] | nl.PREC_WINDOW
    x, iters, be, ok = nl.solve_mixed_ir(nl.DMatrix.from_numpy(A), nl.DMatrix.from_numpy(b.reshape(-1, 1)), p,
                                         tol=1e-14, maxit=20)
    xv = x.numpy()[:, 0]
    assert ok and be <= 1e-14, (iters, be)
    assert not nl.last_ir_fallback()
    assert berr(A, xv, b) <= 2e-14
    assert np.allclose(xv, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_mixed_ir_window_singular(nl):
    """An all-zero column cannot be pivoted inside any window: reported as a
    singular factorisation, not as a converged solve."""
    n = 256
    A = kappa_matrix(395, n)
    A[:, 70] = 0.0
    b = gen.normal(396, (n,))
    _, _, _, ok = nl.solve_mixed_ir(nl.DMatrix.from_numpy(A), nl.DMatrix.from_numpy(b.reshape(-1, 1)),
                                    nl.PREC_FP32 | nl.PREC_WINDOW, tol=1e-14, maxit=5)
    assert not ok


@pytest.mark.parametrize("variant", ["fp32_window", "bf16x3_window", "fp32"])
def test_gmres_ir_converges(nl, orc, variant):
    """FGMRES refinement preconditioned by the low-precision factors reaches
    the 1e-14 bar."""
    n = 1024
    A = kappa_matrix(270, n)
    b = gen.normal(280, (n,))
    base = {"fp32_window": nl.PREC_FP32 | nl.PREC_WINDOW,
            "bf16x3_window": nl.PREC_BF16X3 | nl.PREC_WINDOW, "fp32": nl.PREC_FP32}[variant]
    print(f"gmres IR n={n} {variant}", flush=True)
    x, iters, be, ok = nl.solve_mixed_ir(nl.DMatrix.from_numpy(A), nl.DMatrix.from_numpy(b.reshape(-1, 1)),
                                         base | nl.PREC_GMRES, tol=1e-14, maxit=60)
    xv = x.numpy()[:, 0]
    print(f"  solves={iters} berr={be:.3e} ok={ok}")
    assert ok and be <= 1e-14
    assert berr(A, xv, b) <= 2e-14
    LU, ip, _ = orc.getrf(A)
    xr, _ = orc.getrs(LU, ip, b)
    fe = np.max(np.abs(xv - xr[:, 0])) / np.max(np.abs(xr))
    assert fe <= 1e4 * 1e-14, fe


def test_mixed_ir_plain_bf16_reports(nl):
    """Plain BF16 operands at kappa = 1e4 sit outside classical IR's
    convergence region (kappa * u_bf16 ~ 39); the call must still return a
    finite backward error and an honest convergence flag."""
    n = 512
    A = kappa_matrix(90, n)
    b = gen.normal(91, (n,))
    x, iters, be, ok = nl.solve_mixed_ir(nl.DMatrix.from_numpy(A), nl.DMatrix.from_numpy(b.reshape(-1, 1)),
                                         nl.PREC_BF16, tol=1e-14, maxit=10)
    print(f"plain bf16: iters={iters} berr={be:.3e} converged={ok}")
    assert iters >= 1
    if ok:
        assert be <= 1e-14 and berr(A, x.numpy()[:, 0], b) <= 2e-14


def test_mixed_ir_well_conditioned_bf16(nl):
    """With kappa = 10 plain BF16 trailing updates do converge."""
    n = 512
    A = kappa_matrix(92, n, kappa=10.0)
    b = gen.normal(93, (n,))
    x, iters, be, ok = nl.solve_mixed_ir(nl.DMatrix.from_numpy(A), nl.DMatrix.from_numpy(b.reshape(-1, 1)),
                                         nl.PREC_BF16, tol=1e-14, maxit=40)
    print(f"bf16 kappa=10: iters={iters} berr={be:.3e}")
    assert ok and berr(A, x.numpy()[:, 0], b) <= 2e-14


def _window_trap(kind, n):
    """Nonsingular, well-conditioned matrices on which pivots searched only in
    the next 128 rows fail: their factors are singular or unstable, so the
    solver has to factor again with partial pivoting."""
    rng = np.random.default_rng({"anti": 11, "swap": 12, "rankdef": 13}[kind])
    if kind == "anti":
        # every column's only nonzero lies in the mirrored row, below the window
        return np.fliplr(np.eye(n)) * rng.uniform(1.0, 2.0, n)
    if kind == "swap":
        # [[0, B], [C, 0]]: the first half of the columns has nothing in the
        # first half of the rows, so the first windows are exactly singular
        h = n // 2
        A = np.zeros((n, n))
        A[:h, h:] = haar(401, h)
        A[h:, :h] = haar(402, h)
        return A
    # the window rows restricted to the block's 64 columns are rank 63 up to
    # 1e-12, while the rows below and the other columns are O(1): the window's
    # last pivot is ~1e-12 and L21 = A21 U11^{-1} ~ 1e12 (growth ~ 1e12);
    # partial pivoting takes that pivot from the rows below instead
    A = rng.standard_normal((n, n))
    X = rng.standard_normal((128, 63))
    Y = rng.standard_normal((64, 63))
    A[:128, :64] = X @ Y.T / np.sqrt(63) + 1e-12 * rng.standard_normal((128, 64))
    return A


@pytest.mark.parametrize("kind", ["anti", "swap", "rankdef"])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_mixed_ir_window_falls_back(nl, kind, prec):
    """Windowed pivoting is not partial pivoting: where its factors are
    singular or unstable the call must still converge to the 1e-14 bar, by
    factoring again with partial pivoting (lu.rs:104 searches every remaining
    row, min_max.rs:221-241)."""
    n = 512
    A = _window_trap(kind, n)
    b = np.random.default_rng(7).standard_normal(n)
    p = {"fp64": nl.PREC_FP64, "fp32": nl.PREC_FP32, "bf16x3": nl.PREC_BF16X3}[prec] | nl.PREC_WINDOW
    print(f"window trap {kind} {prec}", flush=True)
    x, iters, be, ok = nl.solve_mixed_ir(nl.DMatrix.from_numpy(A), nl.DMatrix.from_numpy(b.reshape(-1, 1)), p,
                                         tol=1e-14, maxit=30)
    fb = nl.last_ir_fallback()
    xv = x.numpy()[:, 0]
    print(f"  iters={iters} berr={be:.3e} ok={ok} fallback={fb}")
    assert ok and be <= 1e-14
    assert berr(A, xv, b) <= 2e-14
    assert np.allclose(xv, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)
    if kind in ("anti", "swap"):
        assert fb, "a singular window must trigger the partial-pivoting factorisation"


def test_mixed_ir_no_fallback_on_haar(nl):
    """The kappa = 1e4 Haar matrices keep their windowed factors."""
    n = 512
    A = kappa_matrix(470, n)
    b = gen.normal(471, (n,))
    x, iters, be, ok = nl.solve_mixed_ir(nl.DMatrix.from_numpy(A), nl.DMatrix.from_numpy(b.reshape(-1, 1)),
                                         nl.PREC_FP32 | nl.PREC_WINDOW, tol=1e-14, maxit=30)
    assert ok and not nl.last_ir_fallback()
