"""Mixed-precision LU + FP64 iterative refinement at the sizes of the
n = 4096 / 8192 / 16384 workload: A = U diag(sigma) V^T with U, V Haar-random
orthogonal and sigma log-spaced from 1 to 1e-4 (kappa = 1e4), b ~ N(0,1).

Bars (no reference counterpart; the plain method is LU::new + solve_mut,
src/linalg/lu.rs:93-122,242-260):
  - every variant reaches normwise backward error |b - Ax|_inf /
    (|A|_inf |x|_inf + |b|_inf) <= 1e-14, recomputed here in FP64;
  - n = 4096: forward error against the oracle's FP64 LU::new + solve_mut
    solution <= kappa * 1e-14;
  - n = 8192, 16384: forward error against the GPU FP64 LU (partial pivoting)
    solution <= kappa * 1e-14.
"""
import numpy as np
import pytest
import torch

import gen

pytestmark = pytest.mark.gpu

KAPPA = 1e4
_cache = {}


def haar_gpu(seed, n):
    """Haar-random orthogonal matrix (QR of a Gaussian with the sign fix of
    Mezzadri 2007), formed on the GPU with torch: input generation only."""
    G = torch.from_numpy(gen.normal(seed, (n, n))).cuda()
    q, r = torch.linalg.qr(G)
    return q * torch.sign(torch.diagonal(r)).unsqueeze(0)


def workload(n):
    """(A column-major as an n x n torch tensor A[i, j], b) on the GPU."""
    if n not in _cache:
        _cache.clear()
        U, V = haar_gpu(3000 + n, n), haar_gpu(3001 + n, n)
        sig = torch.logspace(0, -np.log10(KAPPA), n, dtype=torch.float64, device="cuda")
        A = (U * sig.unsqueeze(0)) @ V.T.contiguous()
        b = torch.from_numpy(gen.normal(3002 + n, (n,))).cuda()
        _cache[n] = (A, b)
        torch.cuda.synchronize()
    return _cache[n]


def as_dmatrix(nl, A):
    n = A.shape[0]
    return nl.DMatrix(A.T.contiguous().reshape(-1), n, n)  # column-major storage


def backward_error(A, x, b):
    r = b - A @ x
    return (r.abs().max() / (A.abs().sum(dim=1).max() * x.abs().max() + b.abs().max())).item()


VARIANTS = {
    "fp64": "PREC_FP64",
    "fp64_window": ("PREC_FP64", "PREC_WINDOW"),
    "fp32_window": ("PREC_FP32", "PREC_WINDOW"),
    "bf16x3_window": ("PREC_BF16X3", "PREC_WINDOW"),
}


def prec_of(nl, v):
    names = VARIANTS[v]
    names = (names,) if isinstance(names, str) else names
    p = 0
    for nm in names:
        p |= getattr(nl, nm)
    return p


def solve(nl, n, v):
    A, b = workload(n)
    print(f"config-3 n={n} {v}", flush=True)
    x, iters, be, ok = nl.solve_mixed_ir(as_dmatrix(nl, A), nl.DMatrix(b.clone(), n, 1), prec_of(nl, v),
                                         tol=1e-14, maxit=30)
    xv = x.data.view(-1)
    print(f"  iters={iters} berr={be:.3e} ok={ok} fallback={nl.last_ir_fallback()}", flush=True)
    return xv, iters, be, ok


@pytest.mark.parametrize("v", list(VARIANTS))
def test_config3_n4096_vs_oracle(nl, orc, v):
    n = 4096
    x, iters, be, ok = solve(nl, n, v)
    A, b = workload(n)
    assert ok and be <= 1e-14
    assert backward_error(A, x, b) <= 2e-14
    key = ("oracle", n)
    if key not in _cache:
        print("  oracle LU::new + solve_mut (FP64, one core)", flush=True)
        Ah = A.cpu().numpy()
        LU, ip, _ = orc.getrf(Ah)
        xr, _ = orc.getrs(LU, ip, b.cpu().numpy())
        _cache[key] = xr[:, 0]
    xr = _cache[key]
    fe = np.max(np.abs(x.cpu().numpy() - xr)) / np.max(np.abs(xr))
    print(f"  forward error vs oracle {fe:.3e}")
    assert fe <= KAPPA * 1e-14


@pytest.mark.parametrize("n", [8192, 16384])
def test_config3_large(nl, n):
    xs = {}
    for v in VARIANTS:
        x, iters, be, ok = solve(nl, n, v)
        A, b = workload(n)
        assert ok and be <= 1e-14, (v, iters, be)
        eta = backward_error(A, x, b)
        print(f"  {v}: recomputed backward error {eta:.3e}", flush=True)
        assert eta <= 2e-14, (v, eta)
        xs[v] = x
    ref = xs["fp64"]
    for v, x in xs.items():
        fe = ((x - ref).abs().max() / ref.abs().max()).item()
        print(f"  {v}: forward error vs FP64 LU {fe:.3e}")
        assert fe <= KAPPA * 1e-14, (v, fe)
