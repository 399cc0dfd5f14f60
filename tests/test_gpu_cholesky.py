"""GPU parity for Cholesky + solve (single and batched uneven) and the adjoint.

Bar: NROCM_EXACT is bit-identical to the oracle (nalgebra's operation order);
NROCM_FAST is within 1e-12 relative (Frobenius) of it.
"""
import numpy as np
import pytest
import torch

import gen

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _rel(a, b):
    d = np.linalg.norm(np.asarray(a) - np.asarray(b))
    return d / max(np.linalg.norm(b), 1e-300)


NS_ALL = list(range(1, 65)) + [64] * 9 + [33] * 5 + [4] * 11 + list(range(64, 0, -7))


def _run_batch(nl, ns, nrhs, seed, mode, graph=True, schedule=None):
    A, offA, B, offB = gen.spd_batch(seed, ns, nrhs)
    bt = nl.CholeskyBatch(ns, nrhs)
    bt.set_graph(graph)
    if schedule:
        bt.set_schedule(schedule, mode)
    Ad = torch.from_numpy(A).cuda()
    Bd = torch.from_numpy(B).cuda()
    info = bt.solve(Ad, Bd, mode)
    torch.cuda.synchronize()
    return A, offA, B, offB, Ad.cpu().numpy(), Bd.cpu().numpy(), info.cpu().numpy()


@pytest.mark.parametrize("nrhs", [4, 1, 6, 0, 3])
@pytest.mark.parametrize("sched", ["default", "rows", "tile", "tile2"])
def test_batched_exact_is_bitwise(nl, orc, nrhs, sched):
    ns = NS_ALL
    print(f"batched exact: {len(ns)} members, nrhs={nrhs}, schedule {sched}", flush=True)
    sch = {"default": None, "rows": {nb: 64 for nb in range(40, 65, 8)},
           "tile": {nb: TILE for nb in range(24, 65, 8)} if nrhs <= 4 else None,
           "tile2": {24: TILE2, 32: TILE2} if nrhs <= 4 else None}[sched]
    A, offA, B, offB, Ag, Bg, info = _run_batch(nl, ns, nrhs, 11, nl.EXACT, schedule=sch)
    Ar, Br = A.copy(), B.copy()
    info_r = orc.potrf_potrs_batched(ns, Ar, offA, Br, offB, nrhs)
    assert np.array_equal(info, info_r) and not info.any()
    assert np.array_equal(_bits(Ag), _bits(Ar)), "factor bits differ"
    assert np.array_equal(_bits(Bg), _bits(Br)), "solution bits differ"


TILE = 100  # NROCM_SCHED_TILE: one wavefront per matrix, 8 x 8 lane grid (padded sizes 24..64)
TILE2 = 101  # NROCM_SCHED_TILE2: two matrices per wavefront, 8 x 4 lanes each (padded sizes 24, 32)


@pytest.mark.parametrize("lanes", [8, 16, 32, 64, TILE, TILE2])
def test_every_schedule_is_bitwise(nl, orc, lanes):
    """Each compiled schedule (the autotuner may pick any) reproduces the
    oracle bit for bit; size classes that do not fit `lanes` keep their
    default."""
    ns = NS_ALL
    A, offA, B, offB = gen.spd_batch(21, ns, 4)
    bt = nl.CholeskyBatch(ns, 4)
    if lanes == TILE:
        want = {nb: TILE for nb in range(24, 65, 8)}
    elif lanes == TILE2:
        want = {24: TILE2, 32: TILE2}
    else:
        want = {nb: lanes for nb in range(8, 65, 8) if nb <= lanes}
    bt.set_schedule(want, nl.EXACT)
    got = bt.schedule(nl.EXACT)
    for nb, g in got.items():
        assert g == want.get(nb, g)
    Ad = torch.from_numpy(A).cuda(); Bd = torch.from_numpy(B).cuda()
    info = bt.solve(Ad, Bd, nl.EXACT).cpu().numpy()
    Ar, Br = A.copy(), B.copy()
    orc.potrf_potrs_batched(ns, Ar, offA, Br, offB, 4)
    assert not info.any()
    assert np.array_equal(_bits(Ad.cpu().numpy()), _bits(Ar))
    assert np.array_equal(_bits(Bd.cpu().numpy()), _bits(Br))


def test_autotune_keeps_parity(nl, orc):
    ns = NS_ALL
    A, offA, B, offB = gen.spd_batch(22, ns, 4)
    bt = nl.CholeskyBatch(ns, 4)
    chosen = bt.autotune(nl.EXACT)
    print("autotuned lanes per matrix:", chosen, flush=True)
    assert set(chosen) == {8 * ((n + 7) // 8) for n in ns}
    for nb, g in chosen.items():
        assert (g in (8, 16, 32, 64) and g >= nb) or (g == TILE and nb >= 24) or (g == TILE2 and nb in (24, 32))
    assert bt.schedule(nl.EXACT) == chosen
    Ad = torch.from_numpy(A).cuda(); Bd = torch.from_numpy(B).cuda()
    info = bt.solve(Ad, Bd, nl.EXACT).cpu().numpy()
    Ar, Br = A.copy(), B.copy()
    orc.potrf_potrs_batched(ns, Ar, offA, Br, offB, 4)
    assert not info.any()
    assert np.array_equal(_bits(Ad.cpu().numpy()), _bits(Ar))
    assert np.array_equal(_bits(Bd.cpu().numpy()), _bits(Br))
    with pytest.raises(nl.NrocmError):
        bt.set_schedule({64: 32}, nl.EXACT)


@pytest.mark.parametrize("sched", ["default", "rows", "tile", "tile2"])
def test_batched_fast_within_1e12(nl, orc, sched):
    ns = NS_ALL
    sch = {"default": None, "rows": {nb: 64 for nb in range(40, 65, 8)},
           "tile": {nb: TILE for nb in range(24, 65, 8)},
           "tile2": {24: TILE2, 32: TILE2}}[sched]
    A, offA, B, offB, Ag, Bg, info = _run_batch(nl, ns, 4, 12, nl.FAST, schedule=sch)
    Ar, Br = A.copy(), B.copy()
    orc.potrf_potrs_batched(ns, Ar, offA, Br, offB, 4)
    assert not info.any()
    for j, n in enumerate(ns):
        xa = Bg[offB[j]:offB[j] + 4 * n]; xr = Br[offB[j]:offB[j] + 4 * n]
        assert _rel(xa, xr) <= 1e-12, (j, n, _rel(xa, xr))


@pytest.mark.parametrize("nrhs", [0, 3])
def test_batched_zero_sizes_and_factor_only(nl, orc, nrhs):
    """Members of size 0 (an empty DMatrix: Cholesky::new succeeds, nothing to
    store) mixed into a batch, with and without right-hand sides."""
    ns = [0, 5, 0, 64, 1, 0, 33, 8, 0]
    A, offA, B, offB, Ag, Bg, info = _run_batch(nl, ns, nrhs, 21, nl.EXACT)
    Ar, Br = A.copy(), B.copy()
    info_r = orc.potrf_potrs_batched(ns, Ar, offA, Br, offB, nrhs)
    assert np.array_equal(info, info_r) and not info.any()
    assert np.array_equal(_bits(Ag), _bits(Ar)), "factor bits differ"
    assert np.array_equal(_bits(Bg), _bits(Br)), "solution bits differ"


def test_batched_empty(nl):
    bt = nl.CholeskyBatch([], 4)
    A, B = bt.alloc()
    info = bt.solve(A, B, nl.EXACT)
    torch.cuda.synchronize()
    assert info.numel() == 0


@pytest.mark.parametrize("sched", ["rows", "tile"])
def test_batched_not_positive_definite(nl, orc, sched):
    ns = [5, 17, 40, 64, 3, 57, 48]
    A, offA, B, offB = gen.spd_batch(13, ns, 4)
    # member 2: make column 7's pivot negative after elimination; member 4: NaN
    # diagonal; member 5: NaN in column 30 below the diagonal (pivot 31 turns NaN);
    # member 6: an exactly zero pivot at column 47
    M = A[offA[2]:offA[2] + 1600].reshape(40, 40, order="F")
    M[7, 7] = -100.0
    A[offA[4]] = np.nan
    A[offA[5] + 40 + 30 * 57] = np.nan
    M6 = A[offA[6]:offA[6] + 48 * 48].reshape(48, 48, order="F")
    L6 = np.linalg.cholesky(M6)
    M6[47, 47] -= L6[47, 47] ** 2  # the last pivot becomes (about) zero or negative
    bt = nl.CholeskyBatch(ns, 4)
    bt.set_schedule({nb: (64 if sched == "rows" else TILE) for nb in range(24, 65, 8)}, nl.EXACT)
    Ad = torch.from_numpy(A).cuda(); Bd = torch.from_numpy(B).cuda()
    info = bt.solve(Ad, Bd, nl.EXACT).cpu().numpy()
    Ar, Br = A.copy(), B.copy()
    info_r = orc.potrf_potrs_batched(ns, Ar, offA, Br, offB, 4)
    assert list(info) == list(info_r)
    assert list(info_r[:5]) == [0, 0, 8, 0, 1] and info_r[5] == 41 and info_r[6] in (0, 48)
    Bg = Bd.cpu().numpy()
    # a failed member's right-hand side is left unchanged; others are solved exactly
    assert np.array_equal(Bg[offB[2]:offB[2] + 160], B[offB[2]:offB[2] + 160])
    for j in [j for j in range(len(ns)) if info_r[j] == 0]:
        sl = slice(offB[j], offB[j] + 4 * ns[j])
        assert np.array_equal(_bits(Bg[sl]), _bits(Br[sl]))


@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("nrhs", [4, 2, 0])
def test_tile2_members_pair_independently(nl, orc, mode, nrhs):
    """Two matrices share a wavefront in NROCM_SCHED_TILE2: pairs of different
    sizes, a failing member beside a healthy one, and an odd member count (the
    last wavefront holds one member).  EXACT matches the oracle bit for bit,
    FAST matches NROCM_SCHED_TILE bit for bit (the two run the same operations
    per element)."""
    ns = [20, 30, 25, 17, 32, 24, 31, 18, 29]
    A, offA, B, offB = gen.spd_batch(31, ns, nrhs)
    # member 1: NaN diagonal; member 2: negative pivot at column 3;
    # member 4: NaN below the diagonal in its last column
    A[offA[1]] = np.nan
    M2 = A[offA[2]:offA[2] + 625].reshape(25, 25, order="F")
    M2[3, 3] = -50.0
    A[offA[4] + 31 + 30 * 32] = np.nan
    m = nl.EXACT if mode == "exact" else nl.FAST
    out = {}
    for sched in (TILE, TILE2):
        bt = nl.CholeskyBatch(ns, nrhs)
        bt.set_schedule({24: sched, 32: sched}, m)
        assert bt.schedule(m)[24] == sched and bt.schedule(m)[32] == sched
        Ad = torch.from_numpy(A).cuda(); Bd = torch.from_numpy(B).cuda()
        info = bt.solve(Ad, Bd, m).cpu().numpy()
        out[sched] = (info, Ad.cpu().numpy(), Bd.cpu().numpy())
    info2, A2, B2 = out[TILE2]
    assert np.array_equal(out[TILE][0], info2)
    assert np.array_equal(_bits(out[TILE][2]), _bits(B2)), "TILE2 solutions differ from TILE"
    Ar, Br = A.copy(), B.copy()
    info_r = orc.potrf_potrs_batched(ns, Ar, offA, Br, offB, nrhs)
    assert list(info2) == list(info_r)
    assert info_r[1] == 1 and info_r[2] == 4 and info_r[4] == 32 and info_r[0] == 0
    for j in range(len(ns)):
        sl = slice(offB[j], offB[j] + nrhs * ns[j])
        if info_r[j] != 0:
            assert np.array_equal(_bits(B2[sl]), _bits(B[sl])), f"failed member {j} changed"
        elif mode == "exact":
            assert np.array_equal(_bits(B2[sl]), _bits(Br[sl])), f"member {j}"
            sa = slice(offA[j], offA[j] + ns[j] * ns[j])
            L2 = np.tril(A2[sa].reshape(ns[j], ns[j], order="F"))
            Lr = np.tril(Ar[sa].reshape(ns[j], ns[j], order="F"))
            assert np.array_equal(_bits(L2), _bits(Lr)), f"factor of member {j}"
        else:
            assert _rel(B2[sl], Br[sl]) <= 1e-12, j


def test_batched_graph_replays_match_eager(nl):
    ns = [64, 50, 31, 17, 8, 2] * 20
    A, offA, B, offB = gen.spd_batch(14, ns, 4)
    bt = nl.CholeskyBatch(ns, 4)
    Ad = torch.from_numpy(A).cuda(); Bd = torch.from_numpy(B).cuda()
    ptrs = bt.pointers(Ad, Bd)
    bt.set_graph(False)
    bt.solve(Ad, Bd, nl.EXACT, ptrs=ptrs)
    torch.cuda.synchronize()
    eager = Bd.cpu().numpy().copy()
    bt.set_graph(True)
    for rep in range(4):
        Ad.copy_(torch.from_numpy(A)); Bd.copy_(torch.from_numpy(B))
        info = bt.solve(Ad, Bd, nl.EXACT, ptrs=ptrs)
        torch.cuda.synchronize()
        assert not info.cpu().numpy().any()
        assert np.array_equal(_bits(Bd.cpu().numpy()), _bits(eager)), f"replay {rep} differs"
        _ = torch.ones(10, device="cuda").sum().item()   # unrelated eager work between replays


@pytest.mark.parametrize("n", [1, 2, 7, 64, 65, 130, 257])
def test_single_cholesky_exact(nl, orc, n):
    G = gen.normal(20 + n, (n, n))
    A = G @ G.T + n * np.eye(n)
    A[0, n - 1] = 12345.0 if n > 1 else A[0, 0]             # upper triangle must not be read
    b = gen.normal(30 + n, (n, 3))
    ch = nl.DMatrix.from_numpy(A).cholesky(nl.EXACT)
    assert ch is not None
    Lr, info = orc.potrf(A)
    assert info == 0
    Lg = ch.l_dirty().numpy()
    assert np.array_equal(_bits(Lg), _bits(Lr))
    x = ch.solve(nl.DMatrix.from_numpy(b)).numpy()
    xr = orc.potrs(Lr, b)
    assert np.array_equal(_bits(x), _bits(xr))


def test_single_cholesky_none(nl):
    A = np.array([[1.0, 2.0], [2.0, 1.0]])
    assert nl.DMatrix.from_numpy(A).cholesky() is None
    A = np.eye(100); A[70, 70] = -1.0
    assert nl.DMatrix.from_numpy(A).cholesky() is None


# ------------------------------------------------------------------ adjoint

def test_adjoint_exact_matches_oracle(nl, orc):
    ns = [32] * 64 + [5, 17, 64, 1]
    A, offA, B, offB = gen.spd_batch(40, ns, 1)
    bt = nl.CholeskyBatch(ns, 1)
    Ad = torch.from_numpy(A).cuda(); Bd = torch.from_numpy(B).cuda()
    info = bt.solve(Ad, Bd, nl.EXACT)
    dX = 2.0 * Bd                                      # loss = sum |x|^2
    dA, dB = bt.backward(Ad, Bd, dX)
    torch.cuda.synchronize()
    assert not info.cpu().numpy().any()
    Lg = Ad.cpu().numpy(); Xg = Bd.cpu().numpy(); dAg = dA.cpu().numpy(); dBg = dB.cpu().numpy()
    for j, n in enumerate(ns):
        L = Lg[offA[j]:offA[j] + n * n].reshape(n, n, order="F")
        x = Xg[offB[j]:offB[j] + n]
        dAr, dBr = orc.potrs_adjoint(L, x, 2 * x)
        assert np.array_equal(dBg[offB[j]:offB[j] + n], dBr)
        assert np.array_equal(dAg[offA[j]:offA[j] + n * n].reshape(n, n, order="F"), dAr)


def test_adjoint_several_rhs_and_zero_sizes(nl, orc):
    """nrhs = 6 spans two register chunks of 4: dB is the per-column adjoint and
    dA sums the per-column terms in column order, ((dA_0 + dA_1) + ...) + dA_5.
    Size-0 members are skipped by every load and store."""
    ns = [0, 7, 0, 33, 64, 1, 12, 0, 40]
    nrhs = 6
    A, offA, B, offB = gen.spd_batch(41, ns, nrhs)
    bt = nl.CholeskyBatch(ns, nrhs)
    Ad = torch.from_numpy(A).cuda(); Bd = torch.from_numpy(B).cuda()
    info = bt.solve(Ad, Bd, nl.EXACT)
    dXh = gen.uniform(42, (B.size,)).ravel()
    dA, dB = bt.backward(Ad, Bd, torch.from_numpy(dXh).cuda())
    torch.cuda.synchronize()
    assert not info.cpu().numpy().any()
    Lg = Ad.cpu().numpy(); Xg = Bd.cpu().numpy(); dAg = dA.cpu().numpy(); dBg = dB.cpu().numpy()
    for j, n in enumerate(ns):
        if n == 0:
            continue
        L = Lg[offA[j]:offA[j] + n * n].reshape(n, n, order="F")
        X = Xg[offB[j]:offB[j] + n * nrhs].reshape(n, nrhs, order="F")
        G = dXh[offB[j]:offB[j] + n * nrhs].reshape(n, nrhs, order="F")
        dA_ref = None
        dB_ref = np.zeros((n, nrhs), order="F")
        for q in range(nrhs):
            dAq, dBq = orc.potrs_adjoint(L, X[:, q], G[:, q])
            dB_ref[:, q] = dBq
            dA_ref = dAq if dA_ref is None else dA_ref + dAq
        got_dB = dBg[offB[j]:offB[j] + n * nrhs].reshape(n, nrhs, order="F")
        got_dA = dAg[offA[j]:offA[j] + n * n].reshape(n, n, order="F")
        assert np.array_equal(_bits(got_dB), _bits(dB_ref)), (j, n)
        assert np.array_equal(_bits(got_dA), _bits(dA_ref)), (j, n)


def test_adjoint_central_differences(nl):
    n = 32
    G = gen.normal(41, (n, n)); A = G @ G.T + n * np.eye(n); b = gen.normal(42, (n,))

    def loss(Am, bv):
        ch = nl.DMatrix.from_numpy(Am).cholesky(nl.EXACT)
        x = ch.solve(nl.DMatrix.from_numpy(bv.reshape(-1, 1))).numpy()[:, 0]
        return float(np.sum(x * x)), ch, x

    _, ch, x = loss(A, b)
    dA, db = ch.backward(nl.DMatrix.from_numpy(x.reshape(-1, 1)), nl.DMatrix.from_numpy(2 * x.reshape(-1, 1)))
    dA = dA.numpy(); db = db.numpy()[:, 0]
    h = 1e-6
    for (i, j) in [(0, 0), (5, 3), (31, 31), (20, 7), (31, 0)]:
        Ap = A.copy(); Am = A.copy(); Ap[i, j] += h; Am[i, j] -= h
        fd = (loss(Ap, b)[0] - loss(Am, b)[0]) / (2 * h)
        an = dA[i, j] * (1 if i == j else 2)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (i, j, fd, an)
    for i in (0, 13, 31):
        bp = b.copy(); bm = b.copy(); bp[i] += h; bm[i] -= h
        fd = (loss(A, bp)[0] - loss(A, bm)[0]) / (2 * h)
        assert abs(fd - db[i]) <= 1e-6 * max(1.0, abs(db[i])), (i, fd, db[i])
