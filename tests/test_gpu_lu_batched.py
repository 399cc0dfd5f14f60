"""Batched uneven general systems: LU::new (src/linalg/lu.rs:93-122) then
LU::solve_mut (lu.rs:242-260) per member, n_j <= 64, one launch per size class.

Bar: NROCM_EXACT gives the oracle's pivots, factors and solutions bit for bit
(the oracle restates nalgebra's loops); NROCM_FAST is within 1e-12 relative."""
import numpy as np
import pytest
import torch

import gen

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def make_batch(ns, nrhs, seed, kind="uniform"):
    ns = np.asarray(ns, dtype=np.int32)
    n64 = ns.astype(np.int64)
    offA = np.concatenate([[0], np.cumsum(n64 * n64)[:-1]]).astype(np.int64)
    offB = np.concatenate([[0], np.cumsum(n64 * nrhs)[:-1]]).astype(np.int64)
    A = np.zeros(int((n64 * n64).sum()))
    B = np.zeros(int((n64 * nrhs).sum()))
    rng = np.random.default_rng(seed)
    for j, n in enumerate(ns):
        if n == 0:
            continue
        if kind == "uniform":
            M = gen.uniform(seed * 100003 + j, (n, n))
        elif kind == "pm1":  # pivot ties everywhere
            M = np.where(rng.random((n, n)) < 0.5, -1.0, 1.0)
        else:  # small integers with zeros: ties, exact zero pivots, skipped steps
            M = rng.integers(-2, 3, (n, n)).astype(np.float64)
            if j % 5 == 3:
                M[:, n // 2] = 0.0  # a zero column: its step is skipped, U is singular
        A[offA[j]:offA[j] + n * n] = M.ravel(order="F")
        if nrhs:
            B[offB[j]:offB[j] + n * nrhs] = gen.uniform(seed * 7919 + j, (n, nrhs)).ravel(order="F")
    return ns, A, offA, B, offB


def oracle_batch(orc, ns, A, offA, B, offB, nrhs):
    LUs, Ps, Xs, infos = [], [], [], []
    for j, n in enumerate(ns):
        M = A[offA[j]:offA[j] + n * n].reshape(n, n, order="F")
        LU, ip, _ = orc.getrf(M) if n else (M.copy(), np.zeros(0, np.int32), 0)
        d = np.diag(LU) if n else np.zeros(0)
        zero = np.nonzero(d == 0.0)[0]
        info = int(zero.max()) + 1 if len(zero) else 0
        X = B[offB[j]:offB[j] + n * nrhs].reshape(n, nrhs, order="F").copy()
        if nrhs and n and info == 0:
            X, ok = orc.getrs(LU, ip, X)
            assert ok
        LUs.append(LU); Ps.append(np.asarray(ip, np.int32)); Xs.append(X); infos.append(info)
    return LUs, Ps, Xs, infos


def run(nl, ns, A, B, nrhs, mode):
    bt = nl.LUBatch(ns, nrhs)
    Ad = torch.from_numpy(A).cuda()
    Bd = torch.from_numpy(B if B.size else np.zeros(1)).cuda()
    P, info = bt.solve(Ad, Bd, mode)
    torch.cuda.synchronize()
    return bt, Ad.cpu().numpy(), Bd.cpu().numpy(), P.cpu().numpy(), info.cpu().numpy()


NS = list(range(1, 65)) + [64] * 6 + [0, 3, 0] + [37] * 4 + list(range(64, 0, -9))


@pytest.mark.parametrize("nrhs", [4, 1, 6, 0])
@pytest.mark.parametrize("kind", ["uniform", "pm1", "int2"])
def test_batched_lu_exact_bitwise(nl, orc, nrhs, kind):
    print(f"batched LU exact {kind} nrhs={nrhs}", flush=True)
    ns, A, offA, B, offB = make_batch(NS, nrhs, 31 + nrhs, kind)
    bt, Ag, Bg, Pg, info = run(nl, ns, A, B, nrhs, nl.EXACT)
    LUs, Ps, Xs, infos = oracle_batch(orc, ns, A, offA, B, offB, nrhs)
    assert list(info) == infos
    for j, n in enumerate(ns):
        assert np.array_equal(Pg[bt.offP[j]:bt.offP[j] + n], Ps[j]), (j, n)
        assert np.array_equal(_bits(Ag[offA[j]:offA[j] + n * n]), _bits(LUs[j].ravel(order="F"))), (j, n)
        if nrhs and infos[j] == 0:
            xg = Bg[offB[j]:offB[j] + n * nrhs]
            assert np.array_equal(_bits(xg), _bits(Xs[j].ravel(order="F"))), (j, n)
    if kind == "int2":
        assert any(infos), "the integer batch holds singular members"


def test_batched_lu_singular_and_nan(nl, orc):
    """A zero column (its step is skipped, U_ii = 0, solve_mut returns false and
    the right-hand side is left unchanged), a NaN at the pivot position and a NaN
    below it (icamax: a NaN at position i wins, other NaNs never do)."""
    ns = [8, 20, 33, 64, 5]
    ns, A, offA, B, offB = make_batch(ns, 2, 41)
    M0 = A[offA[0]:offA[0] + 64].reshape(8, 8, order="F"); M0[:, 3] = 0.0
    M1 = A[offA[1]:offA[1] + 400].reshape(20, 20, order="F"); M1[0, 0] = np.nan
    M2 = A[offA[2]:offA[2] + 33 * 33].reshape(33, 33, order="F"); M2[20, 5] = np.nan
    bt, Ag, Bg, Pg, info = run(nl, ns, A, B, 2, nl.EXACT)
    LUs, Ps, Xs, infos = oracle_batch(orc, ns, A, offA, B, offB, 2)
    assert list(info) == infos and infos[0] == 4
    assert np.array_equal(Bg[offB[0]:offB[0] + 16], B[offB[0]:offB[0] + 16])
    for j, n in enumerate(ns):
        assert np.array_equal(Pg[bt.offP[j]:bt.offP[j] + n], Ps[j])
        # NaNs compare as NaNs (the sign of a NaN made from two NaNs is left open)
        G, Rf = Ag[offA[j]:offA[j] + n * n], LUs[j].ravel(order="F")
        assert np.array_equal(np.isnan(G), np.isnan(Rf))
        assert np.array_equal(_bits(G[~np.isnan(Rf)]), _bits(Rf[~np.isnan(Rf)]))
        if infos[j] == 0:
            Xg, Xr = Bg[offB[j]:offB[j] + 2 * n], Xs[j].ravel(order="F")
            assert np.array_equal(np.isnan(Xg), np.isnan(Xr))
            assert np.array_equal(_bits(Xg[~np.isnan(Xr)]), _bits(Xr[~np.isnan(Xr)]))


def test_batched_lu_fast_within_1e12(nl, orc):
    ns, A, offA, B, offB = make_batch(NS, 4, 51)
    bt, Ag, Bg, Pg, info = run(nl, ns, A, B, 4, nl.FAST)
    LUs, Ps, Xs, infos = oracle_batch(orc, ns, A, offA, B, offB, 4)
    assert not any(infos) and not info.any()
    for j, n in enumerate(ns):
        xg = Bg[offB[j]:offB[j] + 4 * n]
        xr = Xs[j].ravel(order="F")
        assert np.linalg.norm(xg - xr) <= 1e-12 * np.linalg.norm(xr), (j, n)


def test_batched_lu_many_members(nl, orc):
    """4,096 members with n uniform in [4, 64]: every member bit for bit."""
    u = gen.uniform(61, (4096,), 0.0, 1.0)
    ns = np.minimum(4 + np.floor(u * 61), 64).astype(np.int32)
    ns, A, offA, B, offB = make_batch(ns, 4, 61)
    bt, Ag, Bg, Pg, info = run(nl, ns, A, B, 4, nl.EXACT)
    LUs, Ps, Xs, infos = oracle_batch(orc, ns, A, offA, B, offB, 4)
    assert list(info) == infos
    for j, n in enumerate(ns):
        assert np.array_equal(Pg[bt.offP[j]:bt.offP[j] + n], Ps[j])
        assert np.array_equal(_bits(Bg[offB[j]:offB[j] + 4 * n]), _bits(Xs[j].ravel(order="F")))


@pytest.mark.parametrize("nrhs", [1, 4, 6, 0])
def test_batched_lu_adjoint_bitwise(nl, orc, nrhs):
    """The LU-solve adjoint (dB = A^{-T} dX, dA = -dB X^T) is bit-identical to
    the oracle's composition of nalgebra's transposed solves and
    inv_permute_rows on every member."""
    ns = [1, 2, 7, 8, 9, 16, 23, 31, 32, 33, 40, 47, 48, 55, 63, 64, 64, 5]
    ns, A, offA, B, offB = make_batch(ns, nrhs, 71 + nrhs)
    bt, Ag, Bg, Pg, info = run(nl, ns, A, B, nrhs, nl.EXACT)
    assert not info.any()
    dX = np.zeros_like(B)
    for j, n in enumerate(ns):
        if n and nrhs:
            dX[offB[j]:offB[j] + n * nrhs] = gen.normal(900 + j, (n * nrhs,))
    LUd = torch.from_numpy(Ag).cuda(); Pd = torch.from_numpy(Pg).cuda()
    Xd = torch.from_numpy(Bg).cuda(); dXd = torch.from_numpy(dX if dX.size else np.zeros(1)).cuda()
    dA, dB = bt.backward(LUd, Pd, Xd, dXd)
    dA = dA.cpu().numpy(); dB = dB.cpu().numpy()
    for j, n in enumerate(ns):
        LU = Ag[offA[j]:offA[j] + n * n].reshape(n, n, order="F")
        X = Bg[offB[j]:offB[j] + n * nrhs].reshape(n, nrhs, order="F")
        G = dX[offB[j]:offB[j] + n * nrhs].reshape(n, nrhs, order="F")
        if nrhs == 0:
            assert not dA[offA[j]:offA[j] + n * n].any()
            continue
        dAr, dBr, ok = orc.getrs_adjoint(LU, Pg[bt.offP[j]:bt.offP[j] + n], X, G)
        assert ok
        assert np.array_equal(_bits(dB[offB[j]:offB[j] + n * nrhs]), _bits(dBr.ravel(order="F"))), (j, n)
        assert np.array_equal(_bits(dA[offA[j]:offA[j] + n * n]), _bits(dAr.ravel(order="F"))), (j, n)


def test_batched_lu_adjoint_finite_differences(nl):
    """Central differences of loss = sum(g * x), x = A^{-1} b, in entries of A
    and b (h = 1e-6, relative 1e-6), through the GPU forward and adjoint."""
    rng = np.random.default_rng(81)
    n = 12
    A0 = rng.standard_normal((n, n)) + 2 * np.eye(n)
    b0 = rng.standard_normal(n); g = rng.standard_normal(n)

    def fwd(Am, bv):
        bt = nl.LUBatch([n], 1)
        Ad = torch.from_numpy(Am.ravel(order="F").copy()).cuda()
        Bd = torch.from_numpy(bv.copy()).cuda()
        P, info = bt.solve(Ad, Bd, nl.EXACT)
        return bt, Ad, P, Bd

    bt, LUd, Pd, Xd = fwd(A0, b0)
    dA, dB = bt.backward(LUd, Pd, Xd, torch.from_numpy(g).cuda())
    dA = dA.cpu().numpy().reshape(n, n, order="F"); dB = dB.cpu().numpy()

    def loss(Am, bv):
        return float(np.dot(g, fwd(Am, bv)[3].cpu().numpy()))

    h = 1e-6
    for (i, j) in [(0, 0), (3, 7), (11, 2), (5, 5), (8, 11)]:
        Ap = A0.copy(); Am = A0.copy(); Ap[i, j] += h; Am[i, j] -= h
        fd = (loss(Ap, b0) - loss(Am, b0)) / (2 * h)
        assert abs(fd - dA[i, j]) <= 1e-6 * max(1.0, abs(dA[i, j]))
    for i in (0, 6, 11):
        bp = b0.copy(); bm = b0.copy(); bp[i] += h; bm[i] -= h
        fd = (loss(A0, bp) - loss(A0, bm)) / (2 * h)
        assert abs(fd - dB[i]) <= 1e-6 * max(1.0, abs(dB[i]))
