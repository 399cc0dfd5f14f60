"""GPU parity for Householder QR.

Bars: the nalgebra-format QR (QR::new, q_tr_mul, the upper solve) is
bit-identical to the oracle in f64 and f32; the batched tall least squares
(blocked WY on FP32 MFMA + one FP64 refinement step, no nalgebra bit order to
follow) is checked against an FP64 LAPACK least-squares solution.
"""
import numpy as np
import pytest
import torch

import gen

pytestmark = pytest.mark.gpu


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


@pytest.mark.parametrize("shape", [(5, 3), (3, 5), (4, 4), (40, 12), (300, 64), (64, 64), (129, 70),
                                   (1, 1), (1, 5), (5, 1), (600, 2), (2, 600)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_qr_exact_bitwise(nl, orc, shape, dtype):
    A = gen.uniform(300 + shape[0] * 7 + shape[1], shape).astype(dtype)
    qr = nl.DMatrix.from_numpy(A).qr()
    QRr, dr = orc.geqrf(A)
    k = min(shape)
    assert np.array_equal(_bits(qr.diag[:k].cpu().numpy()), _bits(dr)), "diag bits differ"
    assert np.array_equal(_bits(qr.qr.numpy()), _bits(QRr)), "QR bits differ"
    R = qr.r()
    assert np.allclose(R, orc.qr_r(QRr, dr))


@pytest.mark.parametrize("shape", [(30, 8), (64, 64), (257, 33)])
def test_qtmul_and_solve_bitwise(nl, orc, shape):
    m, n = shape
    A = gen.uniform(400 + m, shape)
    B = gen.uniform(401 + m, (m, 3))
    qr = nl.DMatrix.from_numpy(A).qr()
    QRr, dr = orc.geqrf(A)
    Bg = nl.DMatrix.from_numpy(B)
    qr.q_tr_mul(Bg)
    Br = orc.qr_qtmul(QRr, dr, B)
    assert np.array_equal(_bits(Bg.numpy()), _bits(Br))
    if m == n:
        b = gen.uniform(402, (n, 2))
        xg = nl.DMatrix.from_numpy(b)
        assert qr.solve_mut(xg)
        y = orc.qr_qtmul(QRr, dr, b)
        xr, ok = orc.qr_rsolve(QRr, dr, y)
        assert ok and np.array_equal(_bits(xg.numpy()), _bits(xr))
    else:
        x = qr.solve_least_squares(nl.DMatrix.from_numpy(B[:, :1]))
        xr, _, _, _, ok = orc.qr_ls(A, B[:, 0])
        assert ok and np.array_equal(_bits(x.numpy()[:, 0]), _bits(xr))


def test_qr_square_solve_singular(nl):
    A = gen.uniform(410, (6, 6)); A[:, 2] = 0.0
    qr = nl.DMatrix.from_numpy(A).qr()
    assert not qr.solve_mut(nl.DMatrix.from_numpy(np.ones(6)))


def _ls_batch(seed, ms, n):
    ms = np.asarray(ms, dtype=np.int32)
    As, bs, xt = [], [], []
    for j, m in enumerate(ms):
        A = gen.normal(seed + 3 * j, (m, n))
        x = gen.normal(seed + 3 * j + 1, (n,))
        b = A @ x + 1e-3 * gen.normal(seed + 3 * j + 2, (m,))
        As.append(A.astype(np.float32)); bs.append(b.astype(np.float32)); xt.append(x)
    return ms, As, bs


@pytest.mark.parametrize("n", [64, 24])
def test_batched_least_squares(nl, n):
    ms = [256, 2048, 777, 1000, 64 + n, 513, 1536, 300] * 2
    ms, As, bs = _ls_batch(500, ms, n)
    bt = nl.LeastSquaresBatch(ms, n)
    Ad = torch.from_numpy(np.concatenate([a.ravel(order="F") for a in As])).cuda()
    bd = torch.from_numpy(np.concatenate(bs)).cuda()
    A0, b0 = Ad.clone(), bd.clone()
    x0, info0 = bt.solve(Ad, bd, refine=0)
    x1, info1 = bt.solve(Ad, bd, refine=1)
    torch.cuda.synchronize()
    assert torch.equal(Ad, A0) and torch.equal(bd, b0), "inputs were modified"
    assert not info0.cpu().numpy().any() and not info1.cpu().numpy().any()
    x0 = x0.cpu().numpy(); x1 = x1.cpu().numpy()
    worst0 = worst1 = 0.0
    for j in range(len(ms)):
        A = As[j].astype(np.float64); b = bs[j].astype(np.float64)
        xr = np.linalg.lstsq(A, b, rcond=None)[0]
        e0 = np.linalg.norm(x0[j] - xr) / np.linalg.norm(xr)
        e1 = np.linalg.norm(x1[j] - xr) / np.linalg.norm(xr)
        worst0, worst1 = max(worst0, e0), max(worst1, e1)
    print(f"n={n}: FP32 QR error {worst0:.2e}, after one FP64 refinement {worst1:.2e}")
    assert worst0 < 1e-4
    assert worst1 < 1e-7
    assert worst1 < worst0


@pytest.mark.parametrize("n", [64, 24])
def test_batched_least_squares_square(nl, n):
    """m_j = n, the fewest rows allowed: the last reflector has an empty tail
    (tau = 0, the identity) and the least-squares solution is the square solve."""
    ms, As, bs = _ls_batch(700, [n] * 4 + [n + 1, n + 3], n)
    bt = nl.LeastSquaresBatch(ms, n)
    Ad = torch.from_numpy(np.concatenate([a.ravel(order="F") for a in As])).cuda()
    bd = torch.from_numpy(np.concatenate(bs)).cuda()
    x0, info0 = bt.solve(Ad, bd, refine=0)
    x1, info1 = bt.solve(Ad, bd, refine=1)
    torch.cuda.synchronize()
    assert not info0.cpu().numpy().any() and not info1.cpu().numpy().any()
    x0 = x0.cpu().numpy(); x1 = x1.cpu().numpy()
    for j in range(len(ms)):
        A = As[j].astype(np.float64); b = bs[j].astype(np.float64)
        xr = np.linalg.lstsq(A, b, rcond=None)[0]
        kappa = np.linalg.cond(A)
        e0 = np.linalg.norm(x0[j] - xr) / np.linalg.norm(xr)
        e1 = np.linalg.norm(x1[j] - xr) / np.linalg.norm(xr)
        print(f"m={ms[j]} n={n} kappa={kappa:.0f}: {e0:.2e} -> {e1:.2e}")
        assert e0 <= 1e-5 * kappa, (j, e0, kappa)       # FP32 Householder QR
        assert e1 <= 1e-6 and e1 <= max(e0, 1e-12), (j, e1)


def test_batched_least_squares_rank_deficient(nl):
    n = 16
    ms, As, bs = _ls_batch(600, [100, 120], n)
    As[1][:, 5] = 0.0
    bt = nl.LeastSquaresBatch(ms, n)
    Ad = torch.from_numpy(np.concatenate([a.ravel(order="F") for a in As])).cuda()
    bd = torch.from_numpy(np.concatenate(bs)).cuda()
    _, info = bt.solve(Ad, bd, refine=1)
    info = info.cpu().numpy()
    assert info[0] == 0 and info[1] == 6
