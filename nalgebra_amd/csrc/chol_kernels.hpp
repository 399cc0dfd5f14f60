// Device code of the batched small-matrix Cholesky path (see chol_batched.hip
// for the layout and exactness argument).  Included by chol_nbXX.hip, one
// translation unit per padded size so that the heavily unrolled
// instantiations compile in parallel.
#pragma once
#include "nrocm_internal.hpp"

namespace nrocm_chol {

constexpr int CH = 4;  // right-hand sides held in registers at a time

// Occupancy target per padded size: a[NB] doubles live in registers.
template <int NB>
constexpr int waves_per_eu() { return NB <= 16 ? 8 : NB <= 32 ? 5 : NB <= 48 ? 4 : 3; }

template <int NB>
constexpr int group_for() { return NB <= 8 ? 8 : NB <= 16 ? 16 : NB <= 32 ? 32 : 64; }

template <bool EXACT>
__device__ __forceinline__ double sub_prod(double acc, double x, double y)
{
    // nalgebra axpy with factor -x: fl(fl(-x*y) + acc) == fl(acc - fl(x*y)).
    // The library is compiled with -ffp-contract=off, so this stays unfused.
    if constexpr (EXACT) {
        double p = x * y;
        return acc - p;
    } else {
        return __builtin_fma(-x, y, acc);
    }
}

template <int G>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}

template <int MPW>
struct Smem {
    double col[64];
    double diag[64];
    double xch[2][64 * CH];       // b broadcast (forward) / products (backward)
    double acc[2][MPW * 8 * CH];  // 8 accumulators per group and right-hand side
};

// Loads the lower triangle of row r (identity padding) into a[].
template <int NB>
__device__ __forceinline__ void load_lower(double (&a)[NB], const double *A, int lda, int n, int r)
{
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        double v = (r == c) ? 1.0 : 0.0;
        if (c <= r && r < n) v = A[r + (size_t)c * lda];
        a[c] = v;
    }
}

template <int NB>
__device__ __forceinline__ void store_lower(const double (&a)[NB], double *A, int lda, int n, int r)
{
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c <= r && r < n) A[r + (size_t)c * lda] = a[c];
}

// Right-looking factorisation in registers; returns 0 or 1 + first failing column.
template <int NB, int G, bool EXACT>
__device__ __forceinline__ int factor(double (&a)[NB], Smem<64 / G> &s, int n, int nmax, int g, int r, int lane)
{
    int fail = 0;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        if (k < nmax) {
            s.col[lane] = a[k];
            __syncthreads();
            const double d = s.col[g * G + k];
            if (!(d > 0.0) && fail == 0 && k < n) fail = k + 1;
            const double den = sqrt(d);
            if (r == k) {
                a[k] = den;
                s.diag[lane] = den;
            } else if (r > k) {
                a[k] = a[k] / den;
            }
            // second slot: the divided column (col is still being read above)
            s.xch[0][lane] = a[k];
            __syncthreads();
#pragma unroll
            for (int j = k + 1; j < NB; ++j) {
                const double lj = s.xch[0][g * G + j];
                a[j] = sub_prod<EXACT>(a[j], lj, a[k]);
            }
        }
    }
    return fail;
}

// Register element a[i] for a wave-uniform runtime index i.  The switch lowers
// to uniform scalar branches, so a[] stays in VGPRs (a plain a[i] with runtime
// i would demote the whole array to scratch memory).
template <int NB>
__device__ __forceinline__ double reg_at(const double (&a)[NB], int i)
{
    double v = 0.0;
    switch (i) {
#define NROCM_RC(k) \
    case k:         \
        if constexpr (k < NB) v = a[k]; \
        break;
        NROCM_RC(0) NROCM_RC(1) NROCM_RC(2) NROCM_RC(3) NROCM_RC(4) NROCM_RC(5) NROCM_RC(6)
        NROCM_RC(7) NROCM_RC(8) NROCM_RC(9) NROCM_RC(10) NROCM_RC(11) NROCM_RC(12) NROCM_RC(13)
        NROCM_RC(14) NROCM_RC(15) NROCM_RC(16) NROCM_RC(17) NROCM_RC(18) NROCM_RC(19) NROCM_RC(20)
        NROCM_RC(21) NROCM_RC(22) NROCM_RC(23) NROCM_RC(24) NROCM_RC(25) NROCM_RC(26) NROCM_RC(27)
        NROCM_RC(28) NROCM_RC(29) NROCM_RC(30) NROCM_RC(31) NROCM_RC(32) NROCM_RC(33) NROCM_RC(34)
        NROCM_RC(35) NROCM_RC(36) NROCM_RC(37) NROCM_RC(38) NROCM_RC(39) NROCM_RC(40) NROCM_RC(41)
        NROCM_RC(42) NROCM_RC(43) NROCM_RC(44) NROCM_RC(45) NROCM_RC(46) NROCM_RC(47) NROCM_RC(48)
        NROCM_RC(49) NROCM_RC(50) NROCM_RC(51) NROCM_RC(52) NROCM_RC(53) NROCM_RC(54) NROCM_RC(55)
        NROCM_RC(56) NROCM_RC(57) NROCM_RC(58) NROCM_RC(59) NROCM_RC(60) NROCM_RC(61) NROCM_RC(62)
        NROCM_RC(63)
#undef NROCM_RC
    default:
        break;
    }
    return v;
}

__device__ __forceinline__ double reg_q(const double (&b)[CH], int q)
{
    return q == 0 ? b[0] : q == 1 ? b[1] : q == 2 ? b[2] : b[3];
}

// Forward substitution L y = b (solve.rs:488-519) for CH right-hand sides:
// coeff = b_i / L_ii; b_i = coeff; b_r = fl(fl(-coeff * L_ri) + b_r) for r > i.
template <int NB, int G, bool EXACT>
__device__ __forceinline__ void forward(const double (&a)[NB], double (&b)[CH], Smem<64 / G> &s, int nmax,
                                        int g, int r, int lane)
{
    for (int i = 0; i < nmax; ++i) {
        const double li = reg_at<NB>(a, i);
        double *buf = s.xch[i & 1];
#pragma unroll
        for (int q = 0; q < CH; ++q) buf[lane * CH + q] = b[q];
        __syncthreads();
        const double di = s.diag[g * G + i];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const double coeff = buf[(g * G + i) * CH + q] / di;
            if (r == i) b[q] = coeff;
            else if (r > i) b[q] = sub_prod<EXACT>(b[q], coeff, li);
        }
    }
    __syncthreads();
}

// Backward substitution L^T x = y (solve.rs:697-755) for CH right-hand sides:
// b_i = (b_i - dotc(L[i+1..,i], b[i+1..])) / L_ii.  EXACT reproduces dotx's
// eight accumulators (acc_t sums positions t, t+8, ... of the full blocks of 8,
// then 0 + (acc0+acc4) + (acc1+acc5) + (acc2+acc6) + (acc3+acc7), then the tail).
template <int NB, int G, bool EXACT>
__device__ __forceinline__ void backward(const double (&a)[NB], double (&b)[CH], Smem<64 / G> &s, int n,
                                         int nmax, int g, int r, int lane)
{
    for (int i = nmax - 1; i >= 0; --i) {
        const double li = reg_at<NB>(a, i);
        if constexpr (EXACT) {
            double *p = s.xch[i & 1];
            double *acc = s.acc[i & 1];
#pragma unroll
            for (int q = 0; q < CH; ++q) p[lane * CH + q] = li * b[q];
            __syncthreads();
            const int len = n - i - 1;
            const int full = len & ~7;
            if (r < 8 && len > 0) {
#pragma unroll
                for (int q = 0; q < CH; ++q) {
                    double t = 0.0;
                    for (int e = r; e < full; e += 8) t += p[(g * G + i + 1 + e) * CH + q];
                    acc[(g * 8 + r) * CH + q] = t;
                }
            }
            __syncthreads();
            if (r == i && i < n) {
#pragma unroll
                for (int q = 0; q < CH; ++q) {
                    double res = 0.0;
                    if (len > 0) {
                        const double *ac = acc + g * 8 * CH + q;
                        res += ac[0 * CH] + ac[4 * CH];
                        res += ac[1 * CH] + ac[5 * CH];
                        res += ac[2 * CH] + ac[6 * CH];
                        res += ac[3 * CH] + ac[7 * CH];
                        for (int e = full; e < len; ++e) res += p[(g * G + i + 1 + e) * CH + q];
                    }
                    b[q] = (b[q] - res) / li;
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                double v = (r > i && r < n) ? li * b[q] : 0.0;
                v = group_sum<G>(v);
                if (r == i && i < n) b[q] = (b[q] - v) / li;
            }
        }
    }
    __syncthreads();
}

// One launch per size bin; matrices order[begin..end) (sorted by decreasing n).
// With order == nullptr a single matrix (A1, B1, n1, lda1, ldb1) is processed.
template <int NB, int G, bool EXACT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(waves_per_eu<NB>(), 8))) void potrf_potrs_kernel(
    const int *__restrict__ order, int begin, int end, const int *__restrict__ nsz,
    double *const *__restrict__ Aptr, double *const *__restrict__ Bptr, int nrhs,
    int *__restrict__ info, double *A1, double *B1, int n1, int lda1, int ldb1)
{
    constexpr int MPW = 64 / G;
    __shared__ Smem<MPW> s;
    const int lane = threadIdx.x;
    const int g = lane / G, r = lane % G;
    const bool has_b = order ? (Bptr != nullptr) : (B1 != nullptr);  // uniform
    int jm = 0, n = 0, lda = 0, ldb = 0;
    double *A = nullptr, *B = nullptr;
    bool valid;
    if (order) {
        const int item = begin + (int)blockIdx.x * MPW + g;
        valid = item < end;
        if (valid) {
            jm = order[item];
            n = nsz[jm];
            A = Aptr[jm];
            B = Bptr ? Bptr[jm] : nullptr;
            lda = ldb = n;
        }
    } else {
        valid = g == 0;
        if (valid) {
            n = n1;
            A = A1;
            B = B1;
            lda = lda1;
            ldb = ldb1;
        }
    }
    // Members are sorted by decreasing size, so group 0 holds the largest n.
    const int nmax = __builtin_amdgcn_readfirstlane(n);

    double a[NB];
    load_lower<NB>(a, A, lda, n, r);
    const int fail = factor<NB, G, EXACT>(a, s, n, nmax, g, r, lane);
    store_lower<NB>(a, A, lda, n, r);

    if (has_b && nrhs > 0) {
        for (int c0 = 0; c0 < nrhs; c0 += CH) {
            double b[CH];
#pragma unroll
            for (int q = 0; q < CH; ++q)
                b[q] = (valid && r < n && c0 + q < nrhs && fail == 0) ? B[r + (size_t)(c0 + q) * ldb] : 0.0;
            forward<NB, G, EXACT>(a, b, s, nmax, g, r, lane);
            backward<NB, G, EXACT>(a, b, s, n, nmax, g, r, lane);
            if (valid && fail == 0 && r < n) {
#pragma unroll
                for (int q = 0; q < CH; ++q)
                    if (c0 + q < nrhs) B[r + (size_t)(c0 + q) * ldb] = b[q];
            }
        }
    }
    if (valid && r == 0 && info) info[order ? jm : 0] = fail;
}

// Adjoint of X = A^{-1} B given L: Lambda = (L L^T)^{-1} dX (same arithmetic as
// Cholesky::solve_mut), dB = Lambda, dA = -1/2 (Lambda X^T + X Lambda^T).
template <int NB, int G, bool EXACT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(waves_per_eu<NB>(), 8))) void potrs_backward_kernel(
    const int *__restrict__ order, int begin, int end, const int *__restrict__ nsz,
    const double *const *__restrict__ Lptr, const double *const *__restrict__ Xptr,
    const double *const *__restrict__ dXptr, double *const *__restrict__ dAptr,
    double *const *__restrict__ dBptr, int nrhs)
{
    constexpr int MPW = 64 / G;
    __shared__ Smem<MPW> s;
    __shared__ double lam_s[2][64 * CH];
    const int lane = threadIdx.x;
    const int g = lane / G, r = lane % G;
    const int item = begin + (int)blockIdx.x * MPW + g;
    const bool valid = item < end;
    int jm = valid ? order[item] : 0;
    const int n = valid ? nsz[jm] : 0;
    const double *L = valid ? Lptr[jm] : nullptr;
    const double *X = valid ? Xptr[jm] : nullptr;
    const double *dX = valid ? dXptr[jm] : nullptr;
    double *dA = valid ? dAptr[jm] : nullptr;
    double *dB = valid ? dBptr[jm] : nullptr;
    const int nmax = __builtin_amdgcn_readfirstlane(n);

    double a[NB];
    load_lower<NB>(a, L, n, n, r);
    // the diagonal of L is a[r] of lane r; publish it for forward()
#pragma unroll
    for (int k = 0; k < NB; ++k)
        if (r == k) s.diag[lane] = a[k];
    __syncthreads();

    for (int c0 = 0; c0 < nrhs; c0 += CH) {
        double b[CH], x[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const bool in = r < n && c0 + q < nrhs;
            b[q] = in ? dX[r + (size_t)(c0 + q) * n] : 0.0;
            x[q] = in ? X[r + (size_t)(c0 + q) * n] : 0.0;
        }
        forward<NB, G, EXACT>(a, b, s, nmax, g, r, lane);
        backward<NB, G, EXACT>(a, b, s, n, nmax, g, r, lane);
        if (r < n) {
#pragma unroll
            for (int q = 0; q < CH; ++q)
                if (c0 + q < nrhs) dB[r + (size_t)(c0 + q) * n] = b[q];
        }
        // dA[r, c] = -1/2 (lam_r x_c + x_r lam_c), summed over right-hand sides in order
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            lam_s[0][lane * CH + q] = b[q];
            lam_s[1][lane * CH + q] = x[q];
        }
        __syncthreads();
        const int nq = nrhs - c0 < CH ? nrhs - c0 : CH;
        if (r < n) {
            for (int c = 0; c < n; ++c) {
                double v = c0 > 0 ? dA[r + (size_t)c * n] : 0.0;
                for (int q = 0; q < nq; ++q) {
                    const double lc = lam_s[0][(g * G + c) * CH + q];
                    const double xc = lam_s[1][(g * G + c) * CH + q];
                    const double p = reg_q(b, q) * xc;
                    const double t = reg_q(x, q) * lc;
                    const double d = -0.5 * (p + t);
                    v = (c0 == 0 && q == 0) ? d : v + d;
                }
                dA[r + (size_t)c * n] = v;
            }
        }
        __syncthreads();
    }
}

}  // namespace nrocm_chol

#define NROCM_CHOL_DECLARE(NB)                                                                    \
    int nrocm_chol_fwd_launch_##NB(bool exact, unsigned grid, hipStream_t st, const int *order,   \
                                   int begin, int end, const int *nsz, double *const *A,          \
                                   double *const *B, int nrhs, int *info, double *A1, double *B1, \
                                   int n1, int lda1, int ldb1);                                   \
    int nrocm_chol_bwd_launch_##NB(unsigned grid, hipStream_t st, const int *order, int begin,    \
                                   int end, const int *nsz, const double *const *L,               \
                                   const double *const *X, const double *const *dX,               \
                                   double *const *dA, double *const *dB, int nrhs);

#define NROCM_CHOL_DEFINE(NB)                                                                     \
    int nrocm_chol_fwd_launch_##NB(bool exact, unsigned grid, hipStream_t st, const int *order,   \
                                   int begin, int end, const int *nsz, double *const *A,          \
                                   double *const *B, int nrhs, int *info, double *A1, double *B1, \
                                   int n1, int lda1, int ldb1)                                    \
    {                                                                                             \
        using namespace nrocm_chol;                                                               \
        constexpr int G = group_for<NB>();                                                        \
        if (exact)                                                                                \
            hipLaunchKernelGGL((potrf_potrs_kernel<NB, G, true>), dim3(grid), dim3(64), 0, st,    \
                               order, begin, end, nsz, A, B, nrhs, info, A1, B1, n1, lda1, ldb1); \
        else                                                                                      \
            hipLaunchKernelGGL((potrf_potrs_kernel<NB, G, false>), dim3(grid), dim3(64), 0, st,   \
                               order, begin, end, nsz, A, B, nrhs, info, A1, B1, n1, lda1, ldb1); \
        return hipGetLastError() == hipSuccess ? 0 : 1;                                           \
    }                                                                                             \
    int nrocm_chol_bwd_launch_##NB(unsigned grid, hipStream_t st, const int *order, int begin,    \
                                   int end, const int *nsz, const double *const *L,               \
                                   const double *const *X, const double *const *dX,               \
                                   double *const *dA, double *const *dB, int nrhs)                \
    {                                                                                             \
        using namespace nrocm_chol;                                                               \
        constexpr int G = group_for<NB>();                                                        \
        hipLaunchKernelGGL((potrs_backward_kernel<NB, G, true>), dim3(grid), dim3(64), 0, st,     \
                           order, begin, end, nsz, L, X, dX, dA, dB, nrhs);                       \
        return hipGetLastError() == hipSuccess ? 0 : 1;                                           \
    }

NROCM_CHOL_DECLARE(8)
NROCM_CHOL_DECLARE(16)
NROCM_CHOL_DECLARE(24)
NROCM_CHOL_DECLARE(32)
NROCM_CHOL_DECLARE(40)
NROCM_CHOL_DECLARE(48)
NROCM_CHOL_DECLARE(56)
NROCM_CHOL_DECLARE(64)
