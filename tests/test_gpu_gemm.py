"""GPU parity for Matrix::gemm through nrocm_dgemm / nrocm_sgemm.

The column-major path picks one of three tile shapes by problem size
(128 x 128, 64 x 64, and 32 x 32 with four k-groups); each is exercised with
interior tiles, ragged edges in m, n and k, column strides that rule out the
16-byte loads, a misaligned base pointer, row-major operands (computed as the
transposed product) and the degenerate extents k = 0 and 1.  Reference:
nalgebra's gemm has no pinned summation order (matrixmultiply blocks as it
likes), so products of small integers, exact in any order, are compared bit
for bit and real-valued products within 1e-12 (FP64) / 1e-5 (FP32) relative
of an FP64 product (the oracle's for the mid-size case).
"""
import numpy as np
import pytest
import torch

import gen

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def _dev(flat, dtype):
    return torch.from_numpy(np.ascontiguousarray(flat)).to("cuda", dtype)


def _call(nl, m, k, n, alpha, A, rsa, csa, B, rsb, csb, beta, Cd, rsc, csc):
    from nalgebra_amd import _lib
    from nalgebra_amd.linalg import handle, _ptr
    fn = _lib.lib().nrocm_dgemm if Cd.dtype == torch.float64 else _lib.lib().nrocm_sgemm
    st = fn(handle(), m, k, n, alpha, _ptr(A), rsa, csa, _ptr(B), rsb, csb, beta, _ptr(Cd), rsc, csc)
    assert st == 0
    torch.cuda.synchronize()


def _colmajor(X, ld=None, lead=0):
    """X stored column-major with leading dimension ld, after `lead` unused elements."""
    m, n = X.shape
    ld = ld or m
    buf = np.full(lead + ld * max(n, 1), np.nan)
    for j in range(n):
        buf[lead + j * ld: lead + j * ld + m] = X[:, j]
    return buf


# (m, k, n): 128-tiles >= 256 (2048^2), 64-tiles >= 256 (1024^2, 1100 x 1000),
# split-k shape (512^3, ragged small)
SHAPES = [(2048, 96, 2048), (2050, 37, 2047), (1024, 160, 1024), (1100, 75, 1000), (512, 512, 512),
          (333, 129, 77), (40, 1000, 3)]


@pytest.mark.parametrize("shape", SHAPES)
def test_gemm_integer_products_bitwise(nl, shape):
    m, k, n = shape
    print(f"gemm integers {shape}", flush=True)
    rng = np.random.default_rng(m + 3 * k + 7 * n)
    A = rng.integers(-8, 9, size=(m, k)).astype(np.float64)
    B = rng.integers(-8, 9, size=(k, n)).astype(np.float64)
    C0 = rng.integers(-8, 9, size=(m, n)).astype(np.float64)
    Cd = _dev(np.full(m * n, np.nan), torch.float64)
    _call(nl, m, k, n, 1.0, _dev(_colmajor(A), torch.float64), 1, m, _dev(_colmajor(B), torch.float64), 1, k,
          0.0, Cd, 1, m)
    got = Cd.cpu().numpy().reshape(n, m).T
    assert np.array_equal(got, A @ B)  # beta = 0: the NaN in C is never read
    Cd = _dev(_colmajor(C0), torch.float64)
    _call(nl, m, k, n, 2.0, _dev(_colmajor(A), torch.float64), 1, m, _dev(_colmajor(B), torch.float64), 1, k,
          -3.0, Cd, 1, m)
    assert np.array_equal(Cd.cpu().numpy().reshape(n, m).T, 2.0 * (A @ B) - 3.0 * C0)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_gemm_real_vs_fp64(nl, shape, dtype):
    m, k, n = shape
    print(f"gemm real {dtype} {shape}", flush=True)
    A = gen.uniform(m + k, (m, k)); B = gen.uniform(n + k + 1, (k, n)); C0 = gen.uniform(m + n + 2, (m, n))
    tdt, npdt, tol = (torch.float64, np.float64, 1e-12) if dtype == "f64" else (torch.float32, np.float32, 1e-5)
    A, B, C0 = A.astype(npdt), B.astype(npdt), C0.astype(npdt)
    ref = 0.5 * (A.astype(np.float64) @ B.astype(np.float64)) + 0.25 * C0.astype(np.float64)
    Cd = _dev(_colmajor(C0).astype(npdt), tdt)
    _call(nl, m, k, n, 0.5, _dev(_colmajor(A).astype(npdt), tdt), 1, m, _dev(_colmajor(B).astype(npdt), tdt), 1,
          k, 0.25, Cd, 1, m)
    assert _rel(Cd.cpu().numpy().reshape(n, m).T, ref) <= tol


@pytest.mark.parametrize("shape", [(2048, 64, 2048), (1024, 100, 1024), (200, 300, 100)])
def test_gemm_odd_leading_dimensions_and_misaligned_base(nl, shape):
    """Leading dimensions m + 1 and k + 3 (odd: no 16-byte loads) and operands
    that start one element into their buffers (misaligned for 16-byte loads)."""
    m, k, n = shape
    print(f"gemm strides {shape}", flush=True)
    A = gen.uniform(7, (m, k)); B = gen.uniform(8, (k, n))
    ref = A @ B
    for lda, ldb, lead in ((m + 1, k + 3, 0), (m, k, 1), (m + 2, k + 2, 1)):
        Ad = _dev(_colmajor(A, lda, lead), torch.float64)
        Bd = _dev(_colmajor(B, ldb, lead), torch.float64)
        Cd = torch.zeros(m * n + 1, dtype=torch.float64, device="cuda")
        _call(nl, m, k, n, 1.0, Ad[lead:], 1, lda, Bd[lead:], 1, ldb, 0.0, Cd[lead:], 1, m)
        assert _rel(Cd[lead:lead + m * n].cpu().numpy().reshape(n, m).T, ref) <= 1e-12, (lda, ldb, lead)


@pytest.mark.parametrize("shape", [(2048, 64, 2048), (512, 512, 512), (96, 70, 33)])
def test_gemm_row_major_operands(nl, shape):
    m, k, n = shape
    print(f"gemm row-major {shape}", flush=True)
    A = gen.uniform(9, (m, k)); B = gen.uniform(10, (k, n))
    Cd = torch.zeros(m * n, dtype=torch.float64, device="cuda")
    _call(nl, m, k, n, 1.0, _dev(A.ravel(), torch.float64), k, 1, _dev(B.ravel(), torch.float64), n, 1, 0.0, Cd,
          n, 1)
    assert _rel(Cd.cpu().numpy().reshape(m, n), A @ B) <= 1e-12


def test_gemm_degenerate_extents(nl):
    print("gemm degenerate extents", flush=True)
    rng = np.random.default_rng(5)
    for m, k, n in ((5, 0, 7), (300, 0, 300), (1, 1, 1), (1, 300, 1), (300, 1, 300), (1, 64, 500), (500, 64, 1)):
        A = rng.standard_normal((m, k)); B = rng.standard_normal((k, n)); C0 = rng.standard_normal((m, n))
        for beta in (0.0, 1.5):
            Cd = _dev(_colmajor(C0), torch.float64)
            Ad = _dev(_colmajor(A) if k else np.zeros(1), torch.float64)
            Bd = _dev(_colmajor(B) if k else np.zeros(1), torch.float64)
            _call(nl, m, k, n, 2.0, Ad, 1, max(m, 1), Bd, 1, max(k, 1), beta, Cd, 1, m)
            ref = 2.0 * (A @ B) + (beta * C0 if beta else 0.0)
            assert _rel(Cd.cpu().numpy().reshape(n, m).T, ref) <= 1e-14 or np.array_equal(
                Cd.cpu().numpy().reshape(n, m).T, ref), (m, k, n, beta)


def test_gemm_mid_size_vs_oracle(nl, orc):
    m, k, n = 640, 700, 600
    A = gen.uniform(61, (m, k)); B = gen.uniform(62, (k, n))
    Cg = (nl.DMatrix.from_numpy(A) @ nl.DMatrix.from_numpy(B)).numpy()
    assert _rel(Cg, orc.gemm(1.0, A, B)) <= 1e-12


def test_gemm_repeatable(nl):
    """The split-k shape adds its k-groups in a fixed order: two runs agree bit for bit."""
    A = nl.DMatrix.from_numpy(gen.uniform(63, (512, 512)))
    B = nl.DMatrix.from_numpy(gen.uniform(64, (512, 512)))
    x = (A @ B).numpy(); y = (A @ B).numpy()
    assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
