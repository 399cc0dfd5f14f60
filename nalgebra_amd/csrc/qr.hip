// Householder QR.
//
// 1. nalgebra's representation, exact arithmetic (nrocm_dgeqrf / sgeqrf,
//    nrocm_dqr_qtmul, nrocm_dqr_rsolve): one workgroup per matrix walks the
//    columns in order exactly as QR::new (src/linalg/qr.rs:55-76) ->
//    householder::clear_column_unchecked (householder.rs:61-85) ->
//    reflection_axis_mut (householder.rs:19-53) -> reflect_with_sign
//    (src/geometry/reflection.rs:70-83).  Every dot product uses dotx's eight
//    accumulators (src/base/blas.rs:93-167); the accumulators are independent
//    chains, one thread each, and the trailing columns are spread over the
//    workgroup.
// 2. Batched tall least squares (nrocm_sgels_batched): see qr_batched below.
#include "gemm_kernels.hpp"
#include "internal_launch.hpp"

#include <algorithm>
#include <cmath>

namespace {

// Unfused y = fl(fl(fl(a*x)*c) + fl(b*y)) (blas_uninit.rs:46-52).
template <typename T>
__device__ __forceinline__ T axcpy1(T a, T x, T c, T b, T y)
{
    T p = a * x;
    p = p * c;
    T q = b * y;
    return p + q;
}

// dotx over len elements of x (stride 1) and y (stride 1) using the 8 threads
// t0..t0+7 of the calling group; the result is returned to thread t0 only.
// `part` is 8 slots of shared scratch owned by the group.
template <typename T>
__device__ __forceinline__ T group_dot8(const T *x, const T *y, int len, int sub, T *part)
{
    const int full = len & ~7;
    T acc = 0;
    for (int e = sub; e < full; e += 8) acc += x[e] * y[e];
    part[sub] = acc;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    T res = 0;
    if (sub == 0) {
        res += part[0] + part[4];
        res += part[1] + part[5];
        res += part[2] + part[6];
        res += part[3] + part[7];
        for (int e = full; e < len; ++e) res += x[e] * y[e];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    return res;
}

template <typename T>
__device__ __forceinline__ T signum_rs(T x)
{
    // Rust f64::signum: NaN -> NaN, +-0 -> +-1
    if (x != x) return x;
    return signbit(x) ? T(-1) : T(1);
}

// Reflects columns [c0, c1) of C (rows i.., ldc) through axis x (len rows)
// with the given sign; one 8-thread group per column.
template <typename T>
__device__ void reflect_cols(const T *x, int len, T *C, int ldc, int c0, int c1, T sign, T *scratch)
{
    const int grp = threadIdx.x >> 3, sub = threadIdx.x & 7, ngrp = blockDim.x >> 3;
    T *part = scratch + grp * 8;
    const T m_two = sign * T(-2);
    // uniform trip count so every group reaches the wave barriers together
    for (int cb = c0; cb < c1; cb += ngrp) {
        const int c = cb + grp;
        T *col = C + (size_t)(c < c1 ? c : c0) * ldc;
        T d = group_dot8<T>(x, col, len, sub, part);
        if (sub == 0) part[0] = d;  // broadcast within the group
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        const T factor = (part[0] - T(0)) * m_two;
        __builtin_amdgcn_wave_barrier();
        if (c < c1)
            for (int r = sub; r < len; r += 8) col[r] = axcpy1<T>(factor, x[r], T(1), sign, col[r]);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void geqrf_exact_kernel(T *A, int m, int n, int lda, T *diag)
{
    __shared__ T scratch[256];
    __shared__ T sh[4];
    const int mn = m < n ? m : n;
    for (int i = 0; i < mn; ++i) {
        T *x = A + i + (size_t)i * lda;
        const int len = m - i;
        // reflection_axis_mut: sq = 0 + dotc(x, x)
        if (threadIdx.x < 8) {
            T d = group_dot8<T>(x, x, len, threadIdx.x, scratch);
            if (threadIdx.x == 0) {
                const T sq = T(0) + d;
                const T nrm = sqrt(sq);
                const T x0 = x[0];
                const T mod = x0 >= T(0) ? x0 : -x0;
                const T sgn = x0 >= T(0) ? T(1) : T(-1);
                const T signed_norm = sgn * nrm;
                const T factor = (sq + mod * nrm) * T(2);
                x[0] = x0 + signed_norm;
                sh[0] = factor;
                sh[1] = signed_norm;
            }
        }
        __syncthreads();
        const T factor = sh[0], signed_norm = sh[1];
        const bool not_zero = factor != T(0);
        if (not_zero) {
            const T s = sqrt(factor);
            for (int r = threadIdx.x; r < len; r += blockDim.x) x[r] = x[r] / s;
            __syncthreads();
            if (threadIdx.x < 8) {
                T d = group_dot8<T>(x, x, len, threadIdx.x, scratch);
                if (threadIdx.x == 0) sh[2] = sqrt(T(0) + d);
            }
            __syncthreads();
            const T nn = sh[2];
            for (int r = threadIdx.x; r < len; r += blockDim.x) x[r] = x[r] / nn;
            __syncthreads();
            const T ret = -signed_norm;
            reflect_cols<T>(x, len, A + i, lda, i + 1, n, signum_rs(ret), scratch);
            if (threadIdx.x == 0) diag[i] = ret;
        } else if (threadIdx.x == 0) {
            diag[i] = signed_norm;
        }
        __syncthreads();
    }
}

// QR::q_tr_mul: B (m x nrhs) <- Q^T B.
__global__ __launch_bounds__(256) void qtmul_exact_kernel(const double *QR, int m, int n, int lda, const double *diag,
                                                          double *B, int ldb, int nrhs)
{
    __shared__ double scratch[256];
    const int mn = m < n ? m : n;
    for (int i = 0; i < mn; ++i) {
        reflect_cols<double>(QR + i + (size_t)i * lda, m - i, B + i, ldb, 0, nrhs, signum_rs(diag[i]), scratch);
        __syncthreads();
    }
}

// B (m x nrhs) <- Q B, Q as QR::q builds it (qr.rs:108-129): the reflectors
// in reverse order, each with the sign of its diagonal entry.  The product
// with the identity reproduces QR::q (up to the sign of zero entries left of
// the diagonal, which q() never touches).
__global__ __launch_bounds__(256) void qmul_exact_kernel(const double *QR, int m, int n, int lda, const double *diag,
                                                         double *B, int ldb, int nrhs)
{
    __shared__ double scratch[256];
    const int mn = m < n ? m : n;
    for (int i = mn - 1; i >= 0; --i) {
        reflect_cols<double>(QR + i + (size_t)i * lda, m - i, B + i, ldb, 0, nrhs, signum_rs(diag[i]), scratch);
        __syncthreads();
    }
}

// B (n x nrhs) <- R B with R the n x n upper triangle of QR::r (qr.rs:81-89:
// |diag| on the diagonal), in the reference BLAS TRMM order (side left,
// upper, no transpose): row i is ((b_i r_ii + b_{i+1} r_{i,i+1}) + ...) over
// the input values, a zero b_k contributing nothing.  One thread per row; the
// input column is read from B0 (a copy), the result written to B.
__global__ __launch_bounds__(256) void rmul_kernel(const double *QR, int n, int lda, const double *diag,
                                                   const double *B0, double *B, int ldb, int nrhs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    for (int c = blockIdx.y; c < nrhs; c += gridDim.y) {
        if (i >= n) continue;
        const double *b0 = B0 + (size_t)c * n;
        const double bi = b0[i];
        double acc = (bi != 0.0) ? bi * fabs(diag[i]) : bi;
        const double *rrow = QR + i;
        for (int k = i + 1; k < n; ++k) {
            const double bk = b0[k];
            if (bk != 0.0) acc = acc + bk * rrow[(size_t)k * lda];
        }
        B[i + (size_t)c * ldb] = acc;
    }
}

__global__ void copy_cols_kernel(const double *B, int ldb, int n, int nrhs, double *B0)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (size_t)n * nrhs) B0[e] = B[(e % n) + (e / n) * (size_t)ldb];
}

// QR's upper solve (qr.rs:224-256) on the leading n x n block.
__global__ __launch_bounds__(256) void rsolve_exact_kernel(const double *QR, int n, int lda, const double *diag,
                                                           double *B, int ldb, int nrhs, int *info)
{
    __shared__ double coeff_s;
    __shared__ int fail;
    if (threadIdx.x == 0) fail = 0;
    __syncthreads();
    for (int c = 0; c < nrhs && !fail; ++c) {
        double *b = B + (size_t)c * ldb;
        for (int i = n - 1; i >= 0; --i) {
            if (threadIdx.x == 0) {
                const double d = fabs(diag[i]);
                if (d == 0.0) {
                    fail = i + 1;
                } else {
                    coeff_s = b[i] / d;
                    b[i] = coeff_s;
                }
            }
            __syncthreads();
            if (fail) break;
            const double coeff = coeff_s;
            for (int r = threadIdx.x; r < i; r += blockDim.x)
                b[r] = axcpy1<double>(-coeff, QR[r + (size_t)i * lda], 1.0, 1.0, b[r]);
            __syncthreads();
        }
    }
    if (threadIdx.x == 0 && info) *info = fail;
}

}  // namespace

// ---------------------------------------------------------------- batched tall least squares
//
// Per system (one 256-thread workgroup), with [A | b] copied into an f32
// workspace W (m x (n+1)) so that A stays intact for the FP64 residual:
//   for each 16-column block J:
//     panel:    unblocked Householder (LAPACK larfg convention, v(0) = 1) on
//               W[16J.., 16J..16J+16) held in registers, R block and V written
//               back, T (16 x 16 upper, compact WY) from V^T V and tau
//     trailing: C = W[16J.., 16J+16..n] (b included) <- (I - V T V^T)^T C
//               computed as Wt = V^T C and C -= V (T^T Wt), both on
//               v_mfma_f32_16x16x4f32 with 64-row chunks staged through LDS
//   x0 = R^{-1} (Q^T b)[0..n)                      (f32 R, f64 arithmetic)
//   refine: r = b - A x0 (FP64), x = x0 + R^{-1} (Q^T r)[0..n)
#include "qr_batched.inc"

extern "C" {

int nrocm_dgeqrf(nrocm_handle h, int m, int n, double *a, int lda, double *diag)
{
    if (!h) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (!a && m > 0 && n > 0) return -4;
    if (lda < std::max(1, m)) return -5;
    if (!diag && std::min(m, n) > 0) return -6;
    if (std::min(m, n) == 0) return NROCM_OK;
    hipLaunchKernelGGL((geqrf_exact_kernel<double>), dim3(1), dim3(256), 0, h->stream, a, m, n, lda, diag);
    NROCM_CHECK_LAUNCH(h);
    return NROCM_OK;
}

int nrocm_sgeqrf(nrocm_handle h, int m, int n, float *a, int lda, float *diag)
{
    if (!h) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (!a && m > 0 && n > 0) return -4;
    if (lda < std::max(1, m)) return -5;
    if (!diag && std::min(m, n) > 0) return -6;
    if (std::min(m, n) == 0) return NROCM_OK;
    hipLaunchKernelGGL((geqrf_exact_kernel<float>), dim3(1), dim3(256), 0, h->stream, a, m, n, lda, diag);
    NROCM_CHECK_LAUNCH(h);
    return NROCM_OK;
}

int nrocm_dqr_qtmul(nrocm_handle h, int m, int n, const double *qr, int lda, const double *diag, int nrhs,
                    double *b, int ldb)
{
    if (!h) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (!qr && m > 0 && n > 0) return -4;
    if (lda < std::max(1, m)) return -5;
    if (!diag && std::min(m, n) > 0) return -6;
    if (nrhs < 0) return -7;
    if (!b && m > 0 && nrhs > 0) return -8;
    if (ldb < std::max(1, m)) return -9;
    if (std::min(m, n) == 0 || nrhs == 0) return NROCM_OK;
    hipLaunchKernelGGL(qtmul_exact_kernel, dim3(1), dim3(256), 0, h->stream, qr, m, n, lda, diag, b, ldb, nrhs);
    NROCM_CHECK_LAUNCH(h);
    return NROCM_OK;
}

int nrocm_dqr_qmul(nrocm_handle h, int m, int n, const double *qr, int lda, const double *diag, int nrhs, double *b,
                   int ldb)
{
    if (!h) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (!qr && m > 0 && n > 0) return -4;
    if (lda < std::max(1, m)) return -5;
    if (!diag && std::min(m, n) > 0) return -6;
    if (nrhs < 0) return -7;
    if (!b && m > 0 && nrhs > 0) return -8;
    if (ldb < std::max(1, m)) return -9;
    if (std::min(m, n) == 0 || nrhs == 0) return NROCM_OK;
    hipLaunchKernelGGL(qmul_exact_kernel, dim3(1), dim3(256), 0, h->stream, qr, m, n, lda, diag, b, ldb, nrhs);
    NROCM_CHECK_LAUNCH(h);
    return NROCM_OK;
}

int nrocm_dqr_rmul(nrocm_handle h, int m, int n, const double *qr, int lda, const double *diag, int nrhs, double *b,
                   int ldb)
{
    if (!h) return -1;
    if (m < 0) return -2;
    if (n < 0 || n > m) return -3;
    if (!qr && n > 0) return -4;
    if (lda < std::max(1, m)) return -5;
    if (!diag && n > 0) return -6;
    if (nrhs < 0) return -7;
    if (!b && n > 0 && nrhs > 0) return -8;
    if (ldb < std::max(1, n)) return -9;
    if (n == 0 || nrhs == 0) return NROCM_OK;
    void *ws = nullptr;
    const int s = nrocm_workspace(h, (size_t)n * nrhs * sizeof(double), &ws);
    if (s != NROCM_OK) return s;
    double *b0 = (double *)ws;
    const size_t tot = (size_t)n * nrhs;
    hipLaunchKernelGGL(copy_cols_kernel, dim3(nrocm_cdiv(tot, 256)), dim3(256), 0, h->stream, b, ldb, n, nrhs, b0);
    hipLaunchKernelGGL(rmul_kernel, dim3(nrocm_cdiv(n, 256), std::min(nrhs, 65535)), dim3(256), 0, h->stream, qr, n,
                       lda, diag, b0, b, ldb, nrhs);
    NROCM_CHECK_LAUNCH(h);
    return NROCM_OK;
}

int nrocm_dqr_rsolve(nrocm_handle h, int n, const double *qr, int lda, const double *diag, int nrhs, double *b,
                     int ldb, int *info)
{
    if (!h) return -1;
    if (n < 0) return -2;
    if (!qr && n > 0) return -3;
    if (lda < std::max(1, n)) return -4;
    if (!diag && n > 0) return -5;
    if (nrhs < 0) return -6;
    if (!b && n > 0 && nrhs > 0) return -7;
    if (ldb < std::max(1, n)) return -8;
    if (!info) return -9;
    hipLaunchKernelGGL(rsolve_exact_kernel, dim3(1), dim3(256), 0, h->stream, qr, n, lda, diag, b, ldb, nrhs, info);
    NROCM_CHECK_LAUNCH(h);
    return NROCM_OK;
}

}  // extern "C"
