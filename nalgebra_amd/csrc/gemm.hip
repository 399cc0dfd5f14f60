// nrocm_dgemm / nrocm_sgemm (drop-in for matrixmultiply::dgemm / sgemm as
// called from src/base/blas_uninit.rs:276-313) and the GEMM launchers the
// factorizations use for their trailing updates.
#include "gemm_kernels.hpp"
#include "internal_launch.hpp"

using namespace nrocm_gemm;

hipError_t nrocm_gemm_d(hipStream_t st, int m, int n, int k, double alpha, const double *A, long rsa,
                        long csa, const double *B, long rsb, long csb, double beta, double *C, long rsc,
                        long csc)
{
    return gemm<double, MODE_F64>(st, m, n, k, alpha, A, rsa, csa, B, rsb, csb, beta, C, rsc, csc);
}

hipError_t nrocm_gemm_s(hipStream_t st, int m, int n, int k, float alpha, const float *A, long rsa,
                        long csa, const float *B, long rsb, long csb, float beta, float *C, long rsc,
                        long csc)
{
    return gemm<float, MODE_F32>(st, m, n, k, alpha, A, rsa, csa, B, rsb, csb, beta, C, rsc, csc);
}

hipError_t nrocm_rank_update_d(hipStream_t st, int m, int n, int k, const double *A, long lda, const double *B,
                               long ldb, double *C, long ldc, int max_groups)
{
    return rank_update<double, MODE_F64>(st, m, n, k, A, lda, B, ldb, C, ldc, max_groups);
}

hipError_t nrocm_rank_update_s(hipStream_t st, int m, int n, int k, const float *A, long lda, const float *B,
                               long ldb, float *C, long ldc, int mode, int max_groups)
{
    switch (mode) {
    case MODE_BF16: return rank_update<float, MODE_BF16>(st, m, n, k, A, lda, B, ldb, C, ldc, max_groups);
    case MODE_BF16X3: return rank_update<float, MODE_BF16X3>(st, m, n, k, A, lda, B, ldb, C, ldc, max_groups);
    default: return rank_update<float, MODE_F32>(st, m, n, k, A, lda, B, ldb, C, ldc, max_groups);
    }
}

namespace {

template <typename T, typename F>
int gemm_entry(nrocm_handle h, size_t m, size_t k, size_t n, T alpha, const T *a, ptrdiff_t rsa,
               ptrdiff_t csa, const T *b, ptrdiff_t rsb, ptrdiff_t csb, T beta, T *c, ptrdiff_t rsc,
               ptrdiff_t csc, F launch)
{
    if (!h) return -1;
    const size_t lim = 0x7fffffff;
    if (m > lim) return -2;
    if (k > lim) return -3;
    if (n > lim) return -4;
    if (m == 0 || n == 0) return NROCM_OK;
    if (k > 0 && !a) return -6;
    if (k > 0 && !b) return -9;
    if (!c) return -13;
    const hipError_t e = launch(h->stream, (int)m, (int)n, (int)k, alpha, a, (long)rsa, (long)csa, b,
                                (long)rsb, (long)csb, beta, c, (long)rsc, (long)csc);
    if (e != hipSuccess) return nrocm_fail_hip(h, e, "gemm launch");
    return NROCM_OK;
}

}  // namespace

extern "C" {

int nrocm_dgemm(nrocm_handle h, size_t m, size_t k, size_t n, double alpha, const double *a,
                ptrdiff_t rsa, ptrdiff_t csa, const double *b, ptrdiff_t rsb, ptrdiff_t csb, double beta,
                double *c, ptrdiff_t rsc, ptrdiff_t csc)
{
    return gemm_entry(h, m, k, n, alpha, a, rsa, csa, b, rsb, csb, beta, c, rsc, csc, nrocm_gemm_d);
}

int nrocm_sgemm(nrocm_handle h, size_t m, size_t k, size_t n, float alpha, const float *a,
                ptrdiff_t rsa, ptrdiff_t csa, const float *b, ptrdiff_t rsb, ptrdiff_t csb, float beta,
                float *c, ptrdiff_t rsc, ptrdiff_t csc)
{
    return gemm_entry(h, m, k, n, alpha, a, rsa, csa, b, rsb, csb, beta, c, rsc, csc, nrocm_gemm_s);
}

}  // extern "C"
