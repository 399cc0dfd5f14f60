"""Parity of the batched workloads at their full benchmark sizes, every member checked.

* 65,536 SPD systems, n_j uniform in [4, 64], 4 right-hand sides, seed 2 (the
  batched Cholesky benchmark batch): NROCM_EXACT factors and solutions bit for
  bit against the oracle's Cholesky::new + solve_mut on all members, with the
  autotuned schedule; NROCM_FAST within 1e-12 relative per member.
* 4,096 SPD 32 x 32 systems, seed 4, loss = sum |x|^2: forward and adjoint
  (dA, db) bit for bit against the oracle on all members; central differences
  through the batch API on sampled entries.
* 2,048 tall f32 least-squares systems, m_j uniform in [256, 2048], n = 64:
  every solution (FP32 QR + one FP64 refinement step) against FP64 LAPACK
  lstsq of the same f32 data.
"""
import ctypes as C
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

import gen

pytestmark = pytest.mark.gpu


def _sizes(seed, count, lo, hi):
    u = gen.uniform(seed, (count,), 0.0, 1.0)
    return np.minimum(lo + np.floor(u * (hi - lo + 1)), hi).astype(np.int32)


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _oracle_batch(orc, ns, A, offA, B, offB, nrhs, threads=16):
    """The oracle over member slices on several threads (its C loop releases the GIL)."""
    Ar, Br = A.copy(), B.copy()
    bounds = np.linspace(0, len(ns), threads + 1).astype(int)
    infos = [None] * threads

    def run(k):
        s = slice(bounds[k], bounds[k + 1])
        infos[k] = orc.potrf_potrs_batched(ns[s], Ar, offA[s], Br, offB[s], nrhs)

    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(run, range(threads)))
    return Ar, Br, np.concatenate(infos)


@pytest.fixture(scope="module")
def cholesky_65536(orc):
    ns = _sizes(2, 65536, 4, 64)
    A, offA, B, offB = gen.spd_batch(2, ns, 4)
    Ar, Br, info_r = _oracle_batch(orc, ns, A, offA, B, offB, 4)
    assert not info_r.any()
    return ns, A, offA, B, offB, Ar, Br


def _solve(nl, ns, A, B, nrhs, mode):
    bt = nl.CholeskyBatch(ns, nrhs)
    bt.autotune(mode)
    Ad = torch.from_numpy(A).cuda(); Bd = torch.from_numpy(B).cuda()
    info = bt.solve(Ad, Bd, mode)
    torch.cuda.synchronize()
    return Ad.cpu().numpy(), Bd.cpu().numpy(), info.cpu().numpy()


def test_cholesky_65536_exact_every_member(nl, cholesky_65536):
    ns, A, offA, B, offB, Ar, Br = cholesky_65536
    print(f"65,536-member EXACT batch, {int((ns.astype(np.int64) ** 2).sum())} factor entries", flush=True)
    Ag, Bg, info = _solve(nl, ns, A, B, 4, nl.EXACT)
    assert not info.any()
    assert np.array_equal(_bits(Ag), _bits(Ar)), "factor bits differ"
    assert np.array_equal(_bits(Bg), _bits(Br)), "solution bits differ"


def test_cholesky_65536_fast_every_member(nl, cholesky_65536):
    ns, A, offA, B, offB, Ar, Br = cholesky_65536
    print("65,536-member FAST batch", flush=True)
    _, Bg, info = _solve(nl, ns, A, B, 4, nl.FAST)
    assert not info.any()
    starts = offB.astype(np.int64)
    num = np.add.reduceat((Bg - Br) ** 2, starts)
    den = np.add.reduceat(Br ** 2, starts)
    worst = float(np.sqrt(num / den).max())
    print(f"worst member relative error {worst:.3e}")
    assert worst <= 1e-12


def test_adjoint_4096x32_every_member(nl, orc):
    count, n = 4096, 32
    ns = np.full(count, n, dtype=np.int32)
    A, offA, B, offB = gen.spd_batch(4, ns, 1)
    print("4,096 x 32^2 adjoint batch", flush=True)
    bt = nl.CholeskyBatch(ns, 1)
    Ad = torch.from_numpy(A).cuda(); Bd = torch.from_numpy(B).cuda()
    info = bt.solve(Ad, Bd, nl.EXACT)
    dA, dB = bt.backward(Ad, Bd, 2.0 * Bd)
    torch.cuda.synchronize()
    assert not info.cpu().numpy().any()
    Lg = Ad.cpu().numpy(); Xg = Bd.cpu().numpy(); dAg = dA.cpu().numpy(); dBg = dB.cpu().numpy()
    Ar, Br, info_r = _oracle_batch(orc, ns, A, offA, B, offB, 1)
    assert not info_r.any()
    assert np.array_equal(_bits(Lg), _bits(Ar)) and np.array_equal(_bits(Xg), _bits(Br))
    for j in range(count):
        L = Ar[offA[j]:offA[j] + n * n].reshape(n, n, order="F")
        x = Br[offB[j]:offB[j] + n]
        dAr, dBr = orc.potrs_adjoint(L, x, 2 * x)
        assert np.array_equal(_bits(dBg[offB[j]:offB[j] + n]), _bits(dBr)), j
        assert np.array_equal(_bits(dAg[offA[j]:offA[j] + n * n]), _bits(dAr.ravel(order="F"))), j

    # central differences of loss_j = |x_j|^2 through the batch API, sampled entries
    def losses(Ah, Bh):
        Adp = torch.from_numpy(Ah).cuda(); Bdp = torch.from_numpy(Bh).cuda()
        bt.solve(Adp, Bdp, nl.EXACT)
        return (Bdp.view(count, n) ** 2).sum(dim=1).cpu().numpy()

    h = 1e-6
    members = [7, 2048, 4095]
    for (i, k) in [(0, 0), (9, 4), (31, 31), (20, 7)]:
        Ap = A.copy(); Am = A.copy()
        for j in members:
            Ap[offA[j] + i + k * n] += h; Am[offA[j] + i + k * n] -= h
        fd = (losses(Ap, B) - losses(Am, B)) / (2 * h)
        for j in members:
            an = dAg[offA[j] + i + k * n] * (1 if i == k else 2)
            assert abs(fd[j] - an) <= 1e-6 * max(1.0, abs(an)), (j, i, k, fd[j], an)
    for i in (0, 17, 31):
        Bp = B.copy(); Bm = B.copy()
        for j in members:
            Bp[offB[j] + i] += h; Bm[offB[j] + i] -= h
        fd = (losses(A, Bp) - losses(A, Bm)) / (2 * h)
        for j in members:
            an = dBg[offB[j] + i]
            assert abs(fd[j] - an) <= 1e-6 * max(1.0, abs(an)), (j, i, fd[j], an)


def test_least_squares_2048_every_system(nl):
    from nalgebra_amd import _lib
    lib = _lib.lib()
    count, n = 2048, 64
    ms = _sizes(5, count, 256, 2048)
    m64 = ms.astype(np.int64)
    tot = int(m64.sum())
    Ah = np.zeros(tot * n); xt = np.zeros(count * n); noise = np.zeros(tot)
    lib.nrocm_gen_normal(5, tot * n, Ah.ctypes.data_as(C.c_void_p))
    lib.nrocm_gen_normal(6, count * n, xt.ctypes.data_as(C.c_void_p))
    lib.nrocm_gen_normal(7, tot, noise.ctypes.data_as(C.c_void_p))
    offA = np.concatenate([[0], np.cumsum(m64 * n)[:-1]])
    offb = np.concatenate([[0], np.cumsum(m64)[:-1]])
    bh = np.zeros(tot)
    for j in range(count):
        Aj = Ah[offA[j]:offA[j] + ms[j] * n].reshape(ms[j], n, order="F")
        bh[offb[j]:offb[j] + ms[j]] = Aj @ xt[j * n:(j + 1) * n] + 1e-3 * noise[offb[j]:offb[j] + ms[j]]
    A32 = Ah.astype(np.float32); b32 = bh.astype(np.float32)
    del Ah
    print(f"2,048 least-squares systems, {tot} rows", flush=True)
    bt = nl.LeastSquaresBatch(ms, n)
    Ad = torch.from_numpy(A32).cuda(); bd = torch.from_numpy(b32).cuda()
    x, info = bt.solve(Ad, bd, refine=1)
    torch.cuda.synchronize()
    assert not info.cpu().numpy().any()
    xg = x.cpu().numpy().reshape(count, n)
    worst = 0.0
    for j in range(count):
        Aj = A32[offA[j]:offA[j] + ms[j] * n].reshape(ms[j], n, order="F").astype(np.float64)
        xr = np.linalg.lstsq(Aj, b32[offb[j]:offb[j] + ms[j]].astype(np.float64), rcond=None)[0]
        worst = max(worst, float(np.linalg.norm(xg[j] - xr) / np.linalg.norm(xr)))
    print(f"worst relative error vs FP64 lstsq over all 2,048 systems: {worst:.3e}")
    assert worst <= 1e-7
