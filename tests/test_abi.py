"""CPU-side checks of the C-ABI library: it loads, exports every function the
public header declares, and its host-side input generators are deterministic.
No GPU compute is called here."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "nalgebra_amd", "libnalgebra_rocm.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libnalgebra_rocm.so is not built (run __graft_entry__.build())")
    from nalgebra_amd import _lib
    return _lib.lib()


def test_exports_every_header_function(lib):
    from nalgebra_amd._lib import header_functions, _PROTOS
    names = header_functions()
    assert len(names) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, f"header functions not exported: {missing}"
    # every declared function also has a ctypes prototype in the host layer
    assert set(names) <= set(_PROTOS), set(names) - set(_PROTOS)


def test_header_is_plain_c():
    hdr = os.path.join(ROOT, "include", "nalgebra_rocm.h")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-x", "c", hdr], check=True)


def test_status_strings_and_version(lib):
    assert lib.nrocm_version() >= 100
    assert lib.nrocm_status_string(0) == b"ok"
    assert lib.nrocm_status_string(-3) == b"invalid argument"


def test_null_handle_is_rejected(lib):
    # argument checks run before any device call
    assert lib.nrocm_dgemm(None, 2, 2, 2, 1.0, None, 1, 2, None, 1, 2, 0.0, None, 1, 2) == -1
    assert lib.nrocm_dgetrf(None, 2, 2, None, 2, None, None, 0) == -1
    assert lib.nrocm_destroy(None) == -1


def test_generators_deterministic(lib):
    a = np.zeros(1000); b = np.zeros(1000)
    lib.nrocm_gen_uniform(1, 1000, -1.0, 1.0, a.ctypes.data_as(C.c_void_p))
    lib.nrocm_gen_uniform(1, 1000, -1.0, 1.0, b.ctypes.data_as(C.c_void_p))
    assert np.array_equal(a, b)
    assert a.min() >= -1.0 and a.max() < 1.0 and abs(a.mean()) < 0.1
    z = np.zeros(200000)
    lib.nrocm_gen_normal(7, len(z), z.ctypes.data_as(C.c_void_p))
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    lib.nrocm_gen_uniform(2, 1000, -1.0, 1.0, b.ctypes.data_as(C.c_void_p))
    assert not np.array_equal(a, b)


def test_spd_batch_generator(lib):
    ns = np.array([4, 9, 1], dtype=np.int32)
    offA = np.array([0, 16, 16 + 81], dtype=np.int64)
    offB = np.array([0, 8, 8 + 18], dtype=np.int64)
    A = np.zeros(16 + 81 + 1); B = np.zeros(8 + 18 + 2)
    lib.nrocm_gen_spd_batch(5, 3, ns.ctypes.data_as(C.POINTER(C.c_int)), 2, A.ctypes.data_as(C.c_void_p),
                            offA.ctypes.data_as(C.c_void_p), B.ctypes.data_as(C.c_void_p),
                            offB.ctypes.data_as(C.c_void_p))
    for j, n in enumerate(ns):
        M = A[offA[j]:offA[j] + n * n].reshape(n, n, order="F")
        assert np.array_equal(M, M.T)
        assert np.all(np.linalg.eigvalsh(M) >= n - 1e-9)    # G G^T + n I
    # member j is independent of the others (seed + j)
    A2 = np.zeros(81); B2 = np.zeros(18)
    one = np.array([9], dtype=np.int32); z = np.zeros(1, dtype=np.int64)
    lib.nrocm_gen_spd_batch(6, 1, one.ctypes.data_as(C.POINTER(C.c_int)), 2, A2.ctypes.data_as(C.c_void_p),
                            z.ctypes.data_as(C.c_void_p), B2.ctypes.data_as(C.c_void_p), z.ctypes.data_as(C.c_void_p))
    assert np.array_equal(A2, A[16:16 + 81]) and np.array_equal(B2, B[8:8 + 18])
