"""nrocm_dgetrs_noalloc: the allocation-free LU solve (its scratch is a static
device array of the library) against nrocm_dgetrs on the same factors, to the
same bar (bitwise in EXACT mode), on two calls in a row, the second smaller."""
import numpy as np
import pytest

from test_gpu_getrs_large import EXACT, FAST, gpu_getrs, synthetic_factors

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [EXACT, FAST])
def test_noalloc_matches_getrs(nl, mode):
    for n, nrhs, seed in [(3000, 2, 1), (700, 1, 2), (64, 3, 3), (1, 1, 4)]:
        LU, ipiv = synthetic_factors(n, seed)
        B = np.random.default_rng(seed).standard_normal((n, nrhs))
        ref, info_ref = gpu_getrs(nl, LU, ipiv, B, mode)
        x, info = gpu_getrs(nl, LU, ipiv, B, mode, noalloc=True)
        assert info == info_ref == 0
        if mode == EXACT:
            assert np.array_equal(x.view(np.uint64), ref.view(np.uint64)), (n, nrhs)
        else:
            assert np.linalg.norm(x - ref) <= 1e-12 * np.linalg.norm(ref), (n, nrhs)


def test_noalloc_max_size_and_errors(nl):
    import ctypes as C
    from nalgebra_amd._lib import lib
    from nalgebra_amd.linalg import handle
    L = lib()
    h = handle()
    # beyond the static scratch: refused before any launch, with a message
    assert L.nrocm_dgetrs_noalloc(h, 16385, 1, C.c_void_p(8), 16385, C.c_void_p(8), C.c_void_p(8), 16385,
                                  C.c_void_p(8), EXACT) == 3
    assert b"16384" in L.nrocm_last_error(h)
    assert L.nrocm_dgetrs_noalloc(h, 4, 1, None, 4, C.c_void_p(8), C.c_void_p(8), 4, C.c_void_p(8), EXACT) == -4
    # the largest size it holds, then a small one on the same scratch
    for n in (16384, 100):
        LU, ipiv = synthetic_factors(n, 11 + n)
        B = np.random.default_rng(n).standard_normal((n, 1))
        ref, _ = gpu_getrs(nl, LU, ipiv, B, EXACT)
        x, info = gpu_getrs(nl, LU, ipiv, B, EXACT, noalloc=True)
        assert info == 0 and np.array_equal(x.view(np.uint64), ref.view(np.uint64))
