"""Seeded synthetic inputs for the tests, drawn with the library's host-side
generators (xoshiro256** / SplitMix64, identical on every machine)."""
import ctypes as C

import numpy as np


def _lib():
    from nalgebra_amd import _lib
    return _lib.lib()


def uniform(seed, shape, lo=-1.0, hi=1.0):
    n = int(np.prod(shape))
    out = np.zeros(n)
    _lib().nrocm_gen_uniform(seed, n, lo, hi, out.ctypes.data_as(C.c_void_p))
    return out.reshape(shape, order="F")


def normal(seed, shape):
    n = int(np.prod(shape))
    out = np.zeros(n)
    _lib().nrocm_gen_normal(seed, n, out.ctypes.data_as(C.c_void_p))
    return out.reshape(shape, order="F")


def spd_batch(seed, ns, nrhs):
    """Packed A (A_j = G G^T + n_j I) and B for a batch; returns (A, offA, B, offB)."""
    ns = np.ascontiguousarray(ns, dtype=np.int32)
    n64 = ns.astype(np.int64)
    offA = np.concatenate([[0], np.cumsum(n64 * n64)[:-1]]).astype(np.int64)
    offB = np.concatenate([[0], np.cumsum(n64 * nrhs)[:-1]]).astype(np.int64)
    A = np.zeros(int((n64 * n64).sum()))
    B = np.zeros(int((n64 * nrhs).sum()))
    _lib().nrocm_gen_spd_batch(seed, len(ns), ns.ctypes.data_as(C.POINTER(C.c_int)), nrhs,
                               A.ctypes.data_as(C.c_void_p), offA.ctypes.data_as(C.c_void_p),
                               B.ctypes.data_as(C.c_void_p), offB.ctypes.data_as(C.c_void_p))
    return A, offA, B, offB
