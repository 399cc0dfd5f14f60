"""Argument errors of the C ABI with a live handle: every entry point returns -k
for an invalid argument k before it touches device memory or launches work
(nalgebra panics in these cases: `lu.rs:93-122` asserts the shapes, the solvers
assert matching dimensions), and the handle stays usable afterwards."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

EXACT, FAST = 0, 1


@pytest.fixture(scope="module")
def env(nl):
    from nalgebra_amd._lib import lib
    from nalgebra_amd.linalg import handle
    L = lib()
    h = handle()
    d = torch.zeros(64, dtype=torch.float64, device="cuda")
    f = torch.zeros(64, dtype=torch.float32, device="cuda")
    i = torch.zeros(16, dtype=torch.int32, device="cuda")
    return L, h, C.c_void_p(d.data_ptr()), C.c_void_p(f.data_ptr()), C.c_void_p(i.data_ptr()), (d, f, i)


def test_gemm_argument_errors(env):
    L, h, d, f, _, _ = env
    big = 1 << 31
    assert L.nrocm_dgemm(h, big, 2, 2, 1.0, d, 1, 2, d, 1, 2, 0.0, d, 1, 2) == -2
    assert L.nrocm_dgemm(h, 2, big, 2, 1.0, d, 1, 2, d, 1, 2, 0.0, d, 1, 2) == -3
    assert L.nrocm_dgemm(h, 2, 2, big, 1.0, d, 1, 2, d, 1, 2, 0.0, d, 1, 2) == -4
    assert L.nrocm_dgemm(h, 2, 2, 2, 1.0, None, 1, 2, d, 1, 2, 0.0, d, 1, 2) == -6
    assert L.nrocm_dgemm(h, 2, 2, 2, 1.0, d, 1, 2, None, 1, 2, 0.0, d, 1, 2) == -9
    assert L.nrocm_dgemm(h, 2, 2, 2, 1.0, d, 1, 2, d, 1, 2, 0.0, None, 1, 2) == -13
    assert L.nrocm_sgemm(h, 2, 2, 2, 1.0, f, 1, 2, f, 1, 2, 0.0, None, 1, 2) == -13
    # an empty product is a no-op, whatever the pointers
    assert L.nrocm_dgemm(h, 0, 2, 2, 1.0, None, 1, 2, None, 1, 2, 0.0, None, 1, 2) == 0


def test_lu_cholesky_argument_errors(env):
    L, h, d, _, i, _ = env
    # nrocm_dgetrf(h, m, n, a, lda, ipiv, info, mode)
    assert L.nrocm_dgetrf(h, -1, 2, d, 2, i, i, EXACT) == -2
    assert L.nrocm_dgetrf(h, 2, -1, d, 2, i, i, EXACT) == -3
    assert L.nrocm_dgetrf(h, 2, 2, None, 2, i, i, EXACT) == -4
    assert L.nrocm_dgetrf(h, 3, 2, d, 2, i, i, EXACT) == -5
    assert L.nrocm_dgetrf(h, 2, 2, d, 2, None, i, EXACT) == -6
    assert L.nrocm_dgetrf(h, 2, 2, d, 2, i, None, EXACT) == -7
    assert L.nrocm_dgetrf(h, 2, 2, d, 2, i, i, 7) == -8
    # nrocm_dgetrs(h, n, nrhs, lu, lda, ipiv, b, ldb, info, mode)
    assert L.nrocm_dgetrs(h, -1, 1, d, 2, i, d, 2, i, EXACT) == -2
    assert L.nrocm_dgetrs(h, 2, -1, d, 2, i, d, 2, i, EXACT) == -3
    assert L.nrocm_dgetrs(h, 2, 1, None, 2, i, d, 2, i, EXACT) == -4
    assert L.nrocm_dgetrs(h, 2, 1, d, 1, i, d, 2, i, EXACT) == -5
    assert L.nrocm_dgetrs(h, 2, 1, d, 2, None, d, 2, i, EXACT) == -6
    assert L.nrocm_dgetrs(h, 2, 1, d, 2, i, None, 2, i, EXACT) == -7
    assert L.nrocm_dgetrs(h, 2, 1, d, 2, i, d, 1, i, EXACT) == -8
    assert L.nrocm_dgetrs(h, 2, 1, d, 2, i, d, 2, None, EXACT) == -9
    assert L.nrocm_dgetrs(h, 2, 1, d, 2, i, d, 2, i, -1) == -10
    # nrocm_dpotrf(h, n, a, lda, info, mode)
    assert L.nrocm_dpotrf(h, -1, d, 2, i, EXACT) == -2
    assert L.nrocm_dpotrf(h, 2, None, 2, i, EXACT) == -3
    assert L.nrocm_dpotrf(h, 2, d, 1, i, EXACT) == -4
    assert L.nrocm_dpotrf(h, 2, d, 2, None, EXACT) == -5
    assert L.nrocm_dpotrf(h, 2, d, 2, i, 2) == -6
    # nrocm_dpotrs(h, n, nrhs, l, lda, b, ldb, mode)
    assert L.nrocm_dpotrs(h, -1, 1, d, 2, d, 2, EXACT) == -2
    assert L.nrocm_dpotrs(h, 2, -1, d, 2, d, 2, EXACT) == -3
    assert L.nrocm_dpotrs(h, 2, 1, None, 2, d, 2, EXACT) == -4
    assert L.nrocm_dpotrs(h, 2, 1, d, 1, d, 2, EXACT) == -5
    assert L.nrocm_dpotrs(h, 2, 1, d, 2, None, 2, EXACT) == -6
    assert L.nrocm_dpotrs(h, 2, 1, d, 2, d, 1, EXACT) == -7
    assert L.nrocm_dpotrs(h, 2, 1, d, 2, d, 2, 5) == -8
    # nrocm_dpotrf_substitute(h, n, a, lda, substitute, info, mode)
    assert L.nrocm_dpotrf_substitute(h, -1, d, 2, 1.0, i, EXACT) == -2
    assert L.nrocm_dpotrf_substitute(h, 2, None, 2, 1.0, i, EXACT) == -3
    assert L.nrocm_dpotrf_substitute(h, 2, d, 1, 1.0, i, EXACT) == -4
    assert L.nrocm_dpotrf_substitute(h, 2, d, 2, 1.0, None, EXACT) == -6
    assert L.nrocm_dpotrf_substitute(h, 2, d, 2, 1.0, i, 9) == -7
    # nrocm_ddiag_det(h, n, a, lda, out)
    assert L.nrocm_ddiag_det(h, -1, d, 2, d) == -2
    assert L.nrocm_ddiag_det(h, 2, None, 2, d) == -3
    assert L.nrocm_ddiag_det(h, 2, d, 1, d) == -4
    assert L.nrocm_ddiag_det(h, 2, d, 2, None) == -5


def test_qr_and_mixed_argument_errors(env):
    L, h, d, f, i, _ = env
    for geqrf, p in ((L.nrocm_dgeqrf, d), (L.nrocm_sgeqrf, f)):
        assert geqrf(h, -1, 2, p, 2, p) == -2
        assert geqrf(h, 2, -1, p, 2, p) == -3
        assert geqrf(h, 2, 2, None, 2, p) == -4
        assert geqrf(h, 3, 2, p, 2, p) == -5
        assert geqrf(h, 2, 2, p, 2, None) == -6
    # nrocm_dqr_qtmul(h, m, n, qr, lda, diag, nrhs, b, ldb)
    assert L.nrocm_dqr_qtmul(h, 2, 2, d, 2, d, -1, d, 2) == -7
    assert L.nrocm_dqr_qtmul(h, 2, 2, d, 2, d, 1, None, 2) == -8
    assert L.nrocm_dqr_qtmul(h, 2, 2, d, 2, d, 1, d, 1) == -9
    # nrocm_dqr_qmul / nrocm_dqr_rmul(h, m, n, qr, lda, diag, nrhs, b, ldb)
    assert L.nrocm_dqr_qmul(h, 2, 2, d, 2, d, -1, d, 2) == -7
    assert L.nrocm_dqr_qmul(h, 2, 2, d, 2, d, 1, d, 1) == -9
    assert L.nrocm_dqr_rmul(h, 2, 3, d, 2, d, 1, d, 3) == -3     # R is square only for m >= n
    assert L.nrocm_dqr_rmul(h, 3, 2, d, 2, d, 1, d, 2) == -5
    assert L.nrocm_dqr_rmul(h, 3, 2, d, 3, d, 1, d, 1) == -9
    # nrocm_dqr_rsolve(h, n, qr, lda, diag, nrhs, b, ldb, info)
    assert L.nrocm_dqr_rsolve(h, 2, d, 1, d, 1, d, 2, i) == -4
    assert L.nrocm_dqr_rsolve(h, 2, d, 2, d, 1, d, 2, None) == -9
    # nrocm_dgesv_ir(h, n, a, lda, b, x, precision, tol, maxit, iters, berr)
    it, be = C.c_int(), C.c_double()
    args = lambda n, lda, prec, tol, maxit: L.nrocm_dgesv_ir(h, n, d, lda, d, d, prec, tol, maxit,
                                                             C.byref(it), C.byref(be))
    assert args(-1, 2, 0, 1e-14, 10) == -2
    assert args(3, 2, 0, 1e-14, 10) == -4
    assert args(2, 2, 4, 1e-14, 10) == -7
    assert args(2, 2, 1 | 64 | 8, 1e-14, 10) == -7
    assert args(2, 2, 1 | 64, float("nan"), 10) == -8
    assert args(2, 2, 1 | 64, -1.0, 10) == -8
    assert args(2, 2, 1 | 64, 1e-14, 0) == -9


def test_batch_argument_errors(env):
    L, h, d, f, i, _ = env
    ns = np.array([5, 9, 3], dtype=np.int32)
    nsp = ns.ctypes.data_as(C.POINTER(C.c_int))
    bp = C.c_void_p()
    assert L.nrocm_batch_create(h, -1, nsp, 1, C.byref(bp)) == -2
    assert L.nrocm_batch_create(h, 3, None, 1, C.byref(bp)) == -3
    bad = np.array([5, -1, 3], dtype=np.int32)
    assert L.nrocm_batch_create(h, 3, bad.ctypes.data_as(C.POINTER(C.c_int)), 1, C.byref(bp)) == -3
    assert L.nrocm_batch_create(h, 3, nsp, -1, C.byref(bp)) == -4
    assert L.nrocm_batch_create(h, 3, nsp, 1, None) == -5
    assert L.nrocm_batch_create(h, 3, nsp, 1, C.byref(bp)) == 0
    bt = bp.value
    h2 = C.c_void_p()
    assert L.nrocm_create(C.byref(h2)) == 0
    try:
        # a batch belongs to the handle that created it
        assert L.nrocm_dpotrf_potrs_batched(h2, bt, d, d, i, EXACT) == -2
        assert L.nrocm_dpotrs_backward_batched(h2, bt, d, d, d, d, d) == -2
        assert L.nrocm_sgels_batched(h2, bt, 2, d, d, d, 1, i) == -2
    finally:
        assert L.nrocm_destroy(h2) == 0
    assert L.nrocm_dpotrf_potrs_batched(h, bt, None, d, i, EXACT) == -3
    assert L.nrocm_dpotrf_potrs_batched(h, bt, d, None, i, EXACT) == -4
    assert L.nrocm_dpotrf_potrs_batched(h, bt, d, d, None, EXACT) == -5
    assert L.nrocm_dpotrf_potrs_batched(h, bt, d, d, i, 3) == -6
    for k in range(5):
        ptrs = [d] * 5
        ptrs[k] = None
        assert L.nrocm_dpotrs_backward_batched(h, bt, *ptrs) == -(3 + k)
    assert L.nrocm_sgels_batched(h, bt, 0, d, d, d, 1, i) == -3
    assert L.nrocm_sgels_batched(h, bt, 65, d, d, d, 1, i) == -3
    assert L.nrocm_sgels_batched(h, bt, 2, d, d, d, -1, i) == -7
    # a system with fewer rows than unknowns makes the batch invalid for that n
    assert L.nrocm_sgels_batched(h, bt, 4, d, d, d, 1, i) == -2
    lanes = (C.c_int * 8)()
    assert L.nrocm_batch_set_schedule(bt, 9, lanes) == -2
    assert L.nrocm_batch_set_schedule(bt, EXACT, None) == -3
    lanes[7] = 32  # padded size 64 needs a whole wavefront
    assert L.nrocm_batch_set_schedule(bt, EXACT, lanes) == -3
    lanes[7] = 0
    lanes[0] = 4   # not a group size
    assert L.nrocm_batch_set_schedule(bt, EXACT, lanes) == -3
    lanes[0] = 0
    lanes[1] = 100  # NROCM_SCHED_TILE exists for padded sizes 24..64 only
    assert L.nrocm_batch_set_schedule(bt, EXACT, lanes) == -3
    lanes[1] = 101  # NROCM_SCHED_TILE2 for padded sizes 24 and 32 only
    assert L.nrocm_batch_set_schedule(bt, EXACT, lanes) == -3
    lanes[1] = 0
    lanes[4] = 101
    assert L.nrocm_batch_set_schedule(bt, EXACT, lanes) == -3
    lanes[4] = 0
    assert L.nrocm_batch_schedule(bt, EXACT, None) == -3
    assert L.nrocm_batch_autotune(h, bt, 2, lanes) == -3
    assert L.nrocm_batch_destroy(bt) == 0
    # sizes outside what the batched kernels hold: refused before any launch
    big = np.array([65, 3000], dtype=np.int32)
    assert L.nrocm_batch_create(h, 2, big.ctypes.data_as(C.POINTER(C.c_int)), 1, C.byref(bp)) == 0
    bt = bp.value
    assert L.nrocm_dpotrf_potrs_batched(h, bt, d, d, i, EXACT) == 3       # NROCM_ERR_UNSUPPORTED
    assert L.nrocm_dpotrs_backward_batched(h, bt, d, d, d, d, d) == 3
    assert L.nrocm_sgels_batched(h, bt, 4, d, d, d, 1, i) == 3
    assert b"64" in L.nrocm_last_error(h) or b"2048" in L.nrocm_last_error(h)
    # the schedule calls hold 8 size classes (n <= 64): nothing is read or written
    guard = (C.c_int * 16)(*([-7] * 16))
    assert L.nrocm_batch_schedule(bt, EXACT, guard) == 3
    assert L.nrocm_batch_set_schedule(bt, EXACT, guard) == 3
    assert list(guard) == [-7] * 16
    assert L.nrocm_batch_destroy(bt) == 0


def test_handle_still_works_after_errors(env, nl):
    L, h, d, _, i, _ = env
    assert L.nrocm_dgetrf(h, 3, 2, d, 2, i, i, EXACT) == -5
    A = np.array([[1.0, 2.0], [3.0, 4.0]])
    lu = nl.DMatrix.from_numpy(A).lu(nl.EXACT)
    x = nl.DMatrix.from_numpy(np.array([[1.0], [2.0]]))
    assert lu.solve_mut(x)
    assert np.allclose(x.numpy()[:, 0], np.linalg.solve(A, [1.0, 2.0]), rtol=1e-14)


def test_host_layer_panics_match_nalgebra(nl):
    """The Python host layer raises where nalgebra panics, with nalgebra's
    message, before anything reaches the device."""
    rng = np.random.default_rng(3)
    sq = nl.DMatrix.from_numpy(rng.standard_normal((4, 4)) + 4 * np.eye(4))
    tall = nl.DMatrix.from_numpy(rng.standard_normal((4, 3)))
    b3 = nl.DMatrix.from_numpy(np.ones((3, 1)))
    b4 = nl.DMatrix.from_numpy(np.ones((4, 1)))
    b5 = nl.DMatrix.from_numpy(np.ones((5, 1)))
    # lu.rs:248-255
    with pytest.raises(ValueError, match="LU solve matrix dimension mismatch."):
        sq.lu().solve_mut(b3)
    with pytest.raises(ValueError, match="LU solve: unable to solve a non-square system."):
        tall.lu().solve_mut(b4)
    with pytest.raises(ValueError, match="non-square"):
        tall.lu().determinant()
    # cholesky.rs:222
    with pytest.raises(ValueError, match="The input matrix must be square."):
        tall.cholesky()
    spd = nl.DMatrix.from_numpy(np.eye(4) * 2.0).cholesky()
    with pytest.raises(ValueError, match="dimension mismatch"):
        spd.solve_mut(b3)
    # qr.rs:210-217; q_tr_mul panics in the reflection's dot product (blas.rs:32-35)
    with pytest.raises(ValueError, match="QR solve matrix dimension mismatch."):
        sq.qr().solve_mut(b5)
    with pytest.raises(ValueError, match="QR solve: unable to solve a non-square system."):
        tall.qr().solve_mut(b4)
    with pytest.raises(ValueError, match="Dot product dimensions mismatch"):
        sq.qr().q_tr_mul(b5)
    # blas.rs:788-796
    c = nl.DMatrix.zeros(4, 1)
    with pytest.raises(ValueError, match="gemm: dimensions mismatch for multiplication."):
        c.gemm(1.0, sq, b3, 0.0)
    with pytest.raises(ValueError, match="gemm: dimensions mismatch for addition."):
        c.gemm(1.0, tall, nl.DMatrix.from_numpy(np.ones((3, 2))), 0.0)
    with pytest.raises(ValueError):
        nl.solve_mixed_ir(tall, b4)
    # none of the above left the handle in an error state
    x = nl.DMatrix.from_numpy(np.ones((4, 1)))
    assert sq.lu().solve_mut(x)
    assert np.allclose(sq.numpy() @ x.numpy(), 1.0, rtol=1e-13)
