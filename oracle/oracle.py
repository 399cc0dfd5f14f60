"""ctypes wrapper around the C oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's ``cpu_baseline`` leg, always as the checker / the timed CPU
baseline, never as the thing measured or shipped.

All functions take numpy arrays, work on column-major (Fortran-ordered)
copies and return new arrays, mirroring the reference's by-value API:
``LU::new`` (src/linalg/lu.rs:93), ``LU::solve_mut`` (lu.rs:242),
``Cholesky::new`` (src/linalg/cholesky.rs:196), ``Cholesky::solve_mut``
(cholesky.rs:122), ``QR::new`` (src/linalg/qr.rs:55), ``QR::q_tr_mul``
(qr.rs:157) and ``Matrix::gemm`` (src/base/blas.rs:729).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_L = C.c_long


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        cmd = ["make", "-s", "-C", _HERE, "liboracle.so"]
        if force:
            cmd.insert(1, "-B")
        subprocess.run(cmd, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        for suf, T, P in (("d", C.c_double, _dp), ("s", C.c_float, _fp)):
            f = getattr(_lib, f"nro_{suf}_dot"); f.restype = T; f.argtypes = [_L, P, _L, P, _L]
            f = getattr(_lib, f"nro_{suf}_axcpy"); f.restype = None; f.argtypes = [_L, P, _L, T, P, _L, T, T]
            f = getattr(_lib, f"nro_{suf}_icamax"); f.restype = _L; f.argtypes = [_L, P, _L]
            f = getattr(_lib, f"nro_{suf}_gemm"); f.restype = None
            f.argtypes = [_L, _L, _L, T, P, _L, _L, P, _L, _L, T, P, _L, _L]
            f = getattr(_lib, f"nro_{suf}_getrf"); f.restype = C.c_int; f.argtypes = [_L, _L, P, _L, _ip]
            f = getattr(_lib, f"nro_{suf}_getrs"); f.restype = C.c_int; f.argtypes = [_L, _L, P, _L, _ip, P, _L]
            f = getattr(_lib, f"nro_{suf}_potrf"); f.restype = C.c_int; f.argtypes = [_L, P, _L, C.c_int, T]
            f = getattr(_lib, f"nro_{suf}_potrs"); f.restype = None; f.argtypes = [_L, _L, P, _L, P, _L]
            f = getattr(_lib, f"nro_{suf}_geqrf"); f.restype = None; f.argtypes = [_L, _L, P, _L, P]
            f = getattr(_lib, f"nro_{suf}_qr_qtmul"); f.restype = None; f.argtypes = [_L, _L, P, _L, P, _L, P, _L]
            f = getattr(_lib, f"nro_{suf}_qr_qmul"); f.restype = None; f.argtypes = [_L, _L, P, _L, P, _L, P, _L]
            f = getattr(_lib, f"nro_{suf}_qr_rmul"); f.restype = None; f.argtypes = [_L, P, _L, P, _L, P, _L]
            f = getattr(_lib, f"nro_{suf}_qr_rsolve"); f.restype = C.c_int; f.argtypes = [_L, P, _L, P, _L, P, _L]
            f = getattr(_lib, f"nro_{suf}_qr_ls"); f.restype = C.c_int; f.argtypes = [_L, _L, P, _L, P, P, P]
        _lib.nro_d_potrs_adjoint.restype = None
        _lib.nro_d_potrs_adjoint.argtypes = [_L, _dp, _L, _dp, _dp, _dp, _L, _dp]
        _lib.nro_d_getrs_adjoint.restype = C.c_int
        _lib.nro_d_getrs_adjoint.argtypes = [_L, _L, _dp, _L, _ip, _dp, _L, _dp, _L, _dp, _L, _dp, _L]
        _lib.nro_d_getrf_getrs_batched.restype = None
        _lib.nro_d_getrf_getrs_batched.argtypes = [_L, _ip, _i64p, _dp, _i64p, _dp, _i64p, _ip, _L, _ip]
        _lib.nro_d_potrf_potrs_batched.restype = None
        _lib.nro_d_potrf_potrs_batched.argtypes = [_L, _ip, _i64p, _dp, _i64p, _dp, _L, _ip]
        _lib.nro_s_qr_ls_batched.restype = None
        _lib.nro_s_qr_ls_batched.argtypes = [_L, _ip, _L, _i64p, _fp, _i64p, _fp, _fp, _fp, _ip]
    return _lib


def _suffix(dtype) -> str:
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return "d"
    if dtype == np.float32:
        return "s"
    raise TypeError(f"oracle supports float64/float32, got {dtype}")


def _ptr(a: np.ndarray):
    P = _dp if a.dtype == np.float64 else _fp
    return a.ctypes.data_as(P)


def _fmat(a, dtype=None) -> np.ndarray:
    a = np.asarray(a, dtype=dtype)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    return np.array(a, order="F", copy=True)


def _fn(dtype, name):
    return getattr(lib(), f"nro_{_suffix(dtype)}_{name}")


# ---------------------------------------------------------------- BLAS-1/2

def dot(a, b):
    """8-accumulator dot (src/base/blas.rs:93-167)."""
    a = np.ascontiguousarray(a); b = np.ascontiguousarray(b, dtype=a.dtype)
    return _fn(a.dtype, "dot")(len(a), _ptr(a), 1, _ptr(b), 1)


def axcpy(y, a, x, c, b):
    """y = a*x*c + b*y, unfused (src/base/blas_uninit.rs:46-75)."""
    y = np.array(y, copy=True); x = np.ascontiguousarray(x, dtype=y.dtype)
    _fn(y.dtype, "axcpy")(len(y), _ptr(y), 1, a, _ptr(x), 1, c, b)
    return y


def icamax(x):
    """First index of the strict max |x_i| (src/base/min_max.rs:221-241)."""
    x = np.ascontiguousarray(x)
    return int(_fn(x.dtype, "icamax")(len(x), _ptr(x), 1))


def gemm(alpha, A, B, beta=0.0, Cm=None):
    """C = alpha*A*B + beta*C in nalgebra's gemv loop order."""
    A = _fmat(A); B = _fmat(B, A.dtype)
    m, k = A.shape; k2, n = B.shape
    assert k == k2, "gemm: dimensions mismatch"
    out = np.zeros((m, n), dtype=A.dtype, order="F") if Cm is None else _fmat(Cm, A.dtype)
    _fn(A.dtype, "gemm")(m, k, n, alpha, _ptr(A), 1, m, _ptr(B), 1, k, beta, _ptr(out), 1, m)
    return out


# ---------------------------------------------------------------- LU

def getrf(A):
    """LU::new.  Returns (packed L\\U, ipiv (0-based pivot row per step), info)."""
    A = _fmat(A)
    m, n = A.shape
    ipiv = np.zeros(min(m, n), dtype=np.int32)
    info = _fn(A.dtype, "getrf")(m, n, _ptr(A), m, ipiv.ctypes.data_as(_ip))
    return A, ipiv, int(info)


def getrs(LU, ipiv, B):
    """LU::solve_mut.  Returns (X, ok)."""
    LU = _fmat(LU); X = _fmat(B, LU.dtype)
    n = LU.shape[0]
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    ok = _fn(LU.dtype, "getrs")(n, X.shape[1], _ptr(LU), n, ipiv.ctypes.data_as(_ip), _ptr(X), n)
    return X, bool(ok)


def perm_sequence(ipiv):
    """PermutationSequence contents: the (i, piv) pairs with i != piv, in order
    (src/linalg/permutation_sequence.rs:84-93)."""
    return [(i, int(p)) for i, p in enumerate(ipiv) if int(p) != i]


def lu_determinant(LU, ipiv):
    """LU::determinant (lu.rs:301-314): product of U's diagonal, times the
    permutation's sign (permutation_sequence.rs:158-164)."""
    res = LU.dtype.type(1)
    for i in range(LU.shape[0]):
        res = res * LU[i, i]
    sign = -1.0 if len(perm_sequence(ipiv)) % 2 else 1.0
    return res * LU.dtype.type(sign)


# ---------------------------------------------------------------- Cholesky

def potrf(A, substitute=None):
    """Cholesky::new / new_with_substitute.  Returns (L with dirty upper, info)."""
    A = _fmat(A)
    n = A.shape[0]
    has = substitute is not None
    info = _fn(A.dtype, "potrf")(n, _ptr(A), n, int(has), float(substitute) if has else 0.0)
    return A, int(info)


def potrs(L, B):
    """Cholesky::solve_mut."""
    L = _fmat(L); X = _fmat(B, L.dtype)
    n = L.shape[0]
    _fn(L.dtype, "potrs")(n, X.shape[1], _ptr(L), n, _ptr(X), n)
    return X


def potrs_adjoint(L, x, g):
    """(dLoss/dA symmetrised, dLoss/db) for x = A^{-1} b, g = dLoss/dx."""
    L = _fmat(L, np.float64); n = L.shape[0]
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    g = np.ascontiguousarray(g, dtype=np.float64).ravel()
    dA = np.zeros((n, n), order="F"); db = np.zeros(n)
    lib().nro_d_potrs_adjoint(n, _ptr(L), n, _ptr(x), _ptr(g), _ptr(dA), n, _ptr(db))
    return dA, db


def getrs_adjoint(LU, ipiv, X, G):
    """Adjoint of X = A^{-1} B from LU::new's factors: (dA, dB, ok) for the
    upstream gradient G = dLoss/dX (n x nrhs): dB = A^{-T} G, dA = -dB X^T."""
    LU = _fmat(LU, np.float64); n = LU.shape[0]
    X = _fmat(np.asarray(X, dtype=np.float64).reshape(n, -1), np.float64)
    G = _fmat(np.asarray(G, dtype=np.float64).reshape(n, -1), np.float64)
    nrhs = X.shape[1]
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    dA = np.zeros((n, n), order="F"); dB = np.zeros((n, nrhs), order="F")
    ok = lib().nro_d_getrs_adjoint(n, nrhs, _ptr(LU), max(n, 1), ipiv.ctypes.data_as(_ip), _ptr(X), max(n, 1),
                                   _ptr(G), max(n, 1), _ptr(dA), max(n, 1), _ptr(dB), max(n, 1))
    return dA, dB, bool(ok)


# ---------------------------------------------------------------- QR

def geqrf(A):
    """QR::new: (packed unit axes, signed diag)."""
    A = _fmat(A)
    m, n = A.shape
    diag = np.zeros(min(m, n), dtype=A.dtype)
    _fn(A.dtype, "geqrf")(m, n, _ptr(A), m, _ptr(diag))
    return A, diag


def qr_qtmul(QR, diag, B):
    QR = _fmat(QR); X = _fmat(B, QR.dtype)
    m, n = QR.shape
    diag = np.ascontiguousarray(diag, dtype=QR.dtype)
    _fn(QR.dtype, "qr_qtmul")(m, n, _ptr(QR), m, _ptr(diag), X.shape[1], _ptr(X), m)
    return X


def qr_rsolve(QR, diag, B):
    """Upper solve with R_ii = |diag_i| on the leading n x n block."""
    QR = _fmat(QR); X = _fmat(B, QR.dtype)
    m, n = QR.shape
    diag = np.ascontiguousarray(diag, dtype=QR.dtype)
    ok = _fn(QR.dtype, "qr_rsolve")(n, _ptr(QR), m, _ptr(diag), X.shape[1], _ptr(X), n)
    return X, bool(ok)


def qr_r(QR, diag):
    """QR::r (qr.rs:81-89): upper triangle with |diag| on the diagonal."""
    m, n = QR.shape
    k = min(m, n)
    R = np.triu(np.array(QR[:k, :], copy=True))
    R[np.arange(k), np.arange(k)] = np.abs(diag)
    return R


def qr_qmul(QR, diag, B):
    """Q B with QR::q's Q (qr.rs:108-129), the reflectors applied in reverse."""
    QR = _fmat(QR); X = _fmat(B, QR.dtype)
    m, n = QR.shape
    diag = np.ascontiguousarray(diag, dtype=QR.dtype)
    _fn(QR.dtype, "qr_qmul")(m, n, _ptr(QR), m, _ptr(diag), X.shape[1], _ptr(X), m)
    return X


def qr_rmul(QR, diag, B):
    """R B with R = QR::r()'s n x n upper triangle (m >= n), reference TRMM order."""
    QR = _fmat(QR); X = _fmat(B, QR.dtype)
    m, n = QR.shape
    assert m >= n and X.shape[0] == n
    diag = np.ascontiguousarray(diag, dtype=QR.dtype)
    _fn(QR.dtype, "qr_rmul")(n, _ptr(QR), m, _ptr(diag), X.shape[1], _ptr(X), n)
    return X


def diag_product(M):
    """1 * m_00 * m_11 * ... in index order: LU::determinant's product
    (lu.rs:301-314) and Cholesky::determinant's (cholesky.rs:155-164, which
    then squares it)."""
    M = np.asarray(M)
    res = M.dtype.type(1)
    for i in range(min(M.shape)):
        res = res * M[i, i]
    return res


def chol_ln_determinant(L):
    """Cholesky::ln_determinant (cholesky.rs:166-184): 0 + ln(L_00^2) + ... in order."""
    L = np.asarray(L, dtype=np.float64)
    res = 0.0
    for i in range(L.shape[0]):
        res = res + float(np.log(L[i, i] * L[i, i]))
    return res


def qr_q(QR, diag):
    """QR::q (qr.rs:108-129): apply the reflectors in reverse to the identity."""
    m, n = QR.shape
    k = min(m, n)
    Q = np.zeros((m, k), dtype=QR.dtype, order="F")
    Q[np.arange(k), np.arange(k)] = 1
    for i in reversed(range(k)):
        axis = np.ascontiguousarray(QR[i:, i])
        sign = QR.dtype.type(np.copysign(1.0, diag[i]))
        for c in range(i, k):
            col = np.ascontiguousarray(Q[i:, c])
            f = (dot(axis, col) - 0) * (sign * QR.dtype.type(-2))
            Q[i:, c] = axcpy(col, f, axis, QR.dtype.type(1), sign)
    return Q


def qr_ls(A, b):
    """Tall least squares: QR::new + q_tr_mul + upper solve on rows 0..n.
    Returns (x, QR, diag, Q^T b, ok)."""
    A = _fmat(A)
    m, n = A.shape
    b = np.array(np.asarray(b, dtype=A.dtype).ravel(), copy=True)
    diag = np.zeros(n, dtype=A.dtype)
    x = np.zeros(n, dtype=A.dtype)
    ok = _fn(A.dtype, "qr_ls")(m, n, _ptr(A), m, _ptr(diag), _ptr(b), _ptr(x))
    return x, A, diag, b, bool(ok)


# ---------------------------------------------------------------- batched loops

def potrf_potrs_batched(ns, A_flat, offA, B_flat, offB, nrhs):
    """Per-matrix Cholesky::new + solve_mut over packed column-major storage
    (in place on A_flat / B_flat).  Returns info per matrix."""
    ns = np.ascontiguousarray(ns, dtype=np.int32)
    offA = np.ascontiguousarray(offA, dtype=np.int64)
    offB = np.ascontiguousarray(offB, dtype=np.int64)
    info = np.zeros(len(ns), dtype=np.int32)
    lib().nro_d_potrf_potrs_batched(len(ns), ns.ctypes.data_as(_ip), offA.ctypes.data_as(_i64p),
                                    _ptr(A_flat), offB.ctypes.data_as(_i64p), _ptr(B_flat),
                                    nrhs, info.ctypes.data_as(_ip))
    return info


def getrf_getrs_batched(ns, A_flat, offA, B_flat, offB, offP, nrhs):
    """Per-member LU::new + solve_mut in place; returns (pivots packed, info)."""
    ns = np.ascontiguousarray(ns, dtype=np.int32)
    offA = np.ascontiguousarray(offA, dtype=np.int64); offB = np.ascontiguousarray(offB, dtype=np.int64)
    offP = np.ascontiguousarray(offP, dtype=np.int64)
    piv = np.zeros(max(int(ns.astype(np.int64).sum()), 1), dtype=np.int32)
    info = np.zeros(len(ns), dtype=np.int32)
    lib().nro_d_getrf_getrs_batched(len(ns), ns.ctypes.data_as(_ip), offA.ctypes.data_as(_i64p), _ptr(A_flat),
                                    offB.ctypes.data_as(_i64p), _ptr(B_flat), offP.ctypes.data_as(_i64p),
                                    piv.ctypes.data_as(_ip), nrhs, info.ctypes.data_as(_ip))
    return piv, info


def qr_ls_batched(ms, n, A_flat, offA, b_flat, offb):
    ms = np.ascontiguousarray(ms, dtype=np.int32)
    offA = np.ascontiguousarray(offA, dtype=np.int64)
    offb = np.ascontiguousarray(offb, dtype=np.int64)
    cnt = len(ms)
    diag = np.zeros(cnt * n, dtype=np.float32)
    x = np.zeros(cnt * n, dtype=np.float32)
    info = np.zeros(cnt, dtype=np.int32)
    lib().nro_s_qr_ls_batched(cnt, ms.ctypes.data_as(_ip), n, offA.ctypes.data_as(_i64p), _ptr(A_flat),
                              offb.ctypes.data_as(_i64p), _ptr(b_flat), _ptr(diag), _ptr(x),
                              info.ctypes.data_as(_ip))
    return x.reshape(cnt, n), diag.reshape(cnt, n), info
