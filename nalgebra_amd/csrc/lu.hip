// Blocked LU with partial pivoting (nrocm_dgetrf), LU solve (nrocm_dgetrs),
// and the single-matrix Cholesky entry points (nrocm_dpotrf / nrocm_dpotrs).
//
// getrf schedule (right-looking, outer panel width 64):
//   for each 64-wide panel [j0, j0+nb):
//     for each base block [b0, b0+w) of the panel:
//       panel_kernel        one workgroup factors the block's columns (rows b0..m)
//       laswp               the block's swaps on the other panel columns
//       trsm + update       U12 and the rank-w update of the rest of the panel
//     swap plan + apply     the panel's swaps, composed into one row plan,
//                           on all columns outside it
//     trsm                  U12 = L11^{-1} A12 for the trailing columns
//     trailing update       A22 -= L21 * U12: EXACT = ordered unfused SIMT,
//                           FAST = FP64 MFMA, float working copies: FP32 or BF16 MFMA
#include "gemm_kernels.hpp"
#include "internal_launch.hpp"
#include "lu_kernels.hpp"

#include <algorithm>

using namespace nrocm_lu;

namespace {

template <typename T>
struct PanelCfg {
    int rpt, w;
};

// Rows per thread and width of the one-workgroup panel for `rows` rows.
template <typename T>
PanelCfg<T> panel_cfg(int rows)
{
    int rpt = 1;
    while (rpt * PANEL_THREADS < rows) rpt *= 2;
    return {rpt, panel_width(rpt, (int)sizeof(T))};
}

template <typename T, int RPT>
hipError_t launch_panel_rpt(hipStream_t st, int w_cap, T *A, int lda, int m, int col0, int w, int *ipiv,
                            int *info)
{
    constexpr int WMAX = panel_width(RPT, (int)sizeof(T));
    (void)w_cap;
    hipLaunchKernelGGL((panel_kernel<T, RPT, WMAX>), dim3(1), dim3(PANEL_THREADS), 0, st, A, lda, m, col0, w,
                       ipiv, info);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_panel(hipStream_t st, int rpt, int wcap, T *A, int lda, int m, int col0, int w, int *ipiv,
                        int *info)
{
    switch (rpt) {
    case 1: return launch_panel_rpt<T, 1>(st, wcap, A, lda, m, col0, w, ipiv, info);
    case 2: return launch_panel_rpt<T, 2>(st, wcap, A, lda, m, col0, w, ipiv, info);
    case 4: return launch_panel_rpt<T, 4>(st, wcap, A, lda, m, col0, w, ipiv, info);
    case 8: return launch_panel_rpt<T, 8>(st, wcap, A, lda, m, col0, w, ipiv, info);
    case 16: return launch_panel_rpt<T, 16>(st, wcap, A, lda, m, col0, w, ipiv, info);
    case 32: return launch_panel_rpt<T, 32>(st, wcap, A, lda, m, col0, w, ipiv, info);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t laswp(hipStream_t st, T *A, int lda, int c0, int c1, const int *ipiv, int k0, int k1)
{
    if (c1 <= c0 || k1 <= k0) return hipSuccess;
    hipLaunchKernelGGL((laswp_kernel<T>), dim3(nrocm_cdiv(c1 - c0, 256)), dim3(256), 0, st, A, lda, c0, c1, ipiv,
                       k0, k1);
    return hipGetLastError();
}

template <typename T>
hipError_t trsm(hipStream_t st, T *A, int lda, int r0, int k, int c0, int c1, bool exact)
{
    if (c1 <= c0 || k <= 1) return hipSuccess;
    const unsigned grid = std::min<unsigned>(nrocm_cdiv(c1 - c0, 4), 2048);
    hipLaunchKernelGGL((trsm_unit_lower_kernel<T>), dim3(grid), dim3(256), 0, st, A, lda, r0, k, c0, c1,
                       exact ? 1 : 0);
    return hipGetLastError();
}

template <typename T, bool EXACT, bool LOWER>
hipError_t simt_update(hipStream_t st, T *C, int ldc, int m, int nc, const T *Lp, int ldl, const T *Up, long rsu,
                       long csu, int K, int diag_off)
{
    if (m <= 0 || nc <= 0 || K <= 0) return hipSuccess;
    dim3 grid(nrocm_cdiv(m, 64), nrocm_cdiv(nc, 64));
    hipLaunchKernelGGL((update_kernel<T, EXACT, LOWER>), grid, dim3(256), 0, st, C, ldc, m, nc, Lp, ldl, Up, rsu,
                       csu, K, diag_off);
    return hipGetLastError();
}

// A22 (rows r0.., cols c0..n) -= L21 (rows r0.., cols k0..k0+K) * U12 (rows k0.., cols c0..)
template <typename T>
hipError_t trailing(hipStream_t st, NrocmTrailing trail, T *A, int lda, int r0, int m, int c0, int n, int k0,
                    int K, int max_groups = 0)
{
    const int mm = m - r0, nn = n - c0;
    if (mm <= 0 || nn <= 0 || K <= 0) return hipSuccess;
    T *C = A + r0 + (size_t)c0 * lda;
    const T *Lp = A + r0 + (size_t)k0 * lda;
    const T *Up = A + k0 + (size_t)c0 * lda;
    if (trail == TRAIL_EXACT) return simt_update<T, true, false>(st, C, lda, mm, nn, Lp, lda, Up, 1, lda, K, 0);
    if constexpr (sizeof(T) == 8) {
        return nrocm_rank_update_d(st, mm, nn, K, (const double *)Lp, lda, (const double *)Up, lda, (double *)C, lda,
                                   max_groups);
    } else {
        const int mode = trail == TRAIL_BF16X3 ? 3 : trail == TRAIL_BF16 ? 2 : 1;
        return nrocm_rank_update_s(st, mm, nn, K, (const float *)Lp, lda, (const float *)Up, lda, (float *)C, lda,
                                   mode, max_groups);
    }
}

#define LU_TRY(x)                                      \
    do {                                               \
        hipError_t e__ = (x);                          \
        if (e__ != hipSuccess) return nrocm_fail_hip(h, e__, #x); \
    } while (0)

template <typename T>
int getrf_impl(nrocm_handle h, hipStream_t st, int m, int n, T *A, int lda, int *ipiv, int *info,
               NrocmTrailing trail)
{
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, st, info, 1, 0);
    LU_TRY(hipGetLastError());
    const int mn = std::min(m, n);
    if (mn == 0) return NROCM_OK;
    if (m > 32 * PANEL_THREADS) return nrocm_fail(h, NROCM_ERR_UNSUPPORTED, "getrf: more than 16384 rows");
    const bool exact = trail == TRAIL_EXACT;
    constexpr int NB = 64;
    for (int j0 = 0; j0 < mn; j0 += NB) {
        const int nb = std::min(NB, mn - j0);
        const int j1 = j0 + nb;
        for (int b0 = j0; b0 < j1;) {
            const PanelCfg<T> cfg = panel_cfg<T>(m - b0);
            const int w = std::min(cfg.w, j1 - b0);
            LU_TRY(launch_panel<T>(st, cfg.rpt, cfg.w, A, lda, m, b0, w, ipiv, info));
            // swaps for panel columns [j0, b0); swaps + U12 + rank-w update for
            // [b0+w, j1) -- always in exact order (the panel is narrow)
            const int ncols = (b0 - j0) + (j1 - b0 - w);
            if (ncols > 0) {
                hipLaunchKernelGGL((panel_cols_kernel<T>), dim3(ncols), dim3(256), 0, st, A, lda, m, b0, w, j0,
                                   b0 + w, j1, ipiv, 1);
                LU_TRY(hipGetLastError());
            }
            b0 += w;
        }
        // the panel's swaps on every column outside it, through a composed row plan
        if (j0 > 0 || j1 < n) {
            int *plan = (int *)h->small + NROCM_SMALL_PLAN;
            hipLaunchKernelGGL(swap_plan_kernel, dim3(1), dim3(64), 0, st, ipiv, j0, j1, plan);
            LU_TRY(hipGetLastError());
            const int ncol = j0 + (n - j1);
            hipLaunchKernelGGL((apply_plan_kernel<T>), dim3(nrocm_cdiv(ncol, 2)), dim3(256), 0, st, A, lda, plan, 0,
                               j0, j1, n);
            LU_TRY(hipGetLastError());
        }
        if (j1 < n) {
            LU_TRY(trsm<T>(st, A, lda, j0, nb, j1, n, exact));
            LU_TRY(trailing<T>(st, trail, A, lda, j1, m, j1, n, j0, nb));
        }
    }
    return NROCM_OK;
}

// ---------------------------------------------------------------- getrs (one workgroup)
// LU::solve_mut (lu.rs:242-260): swaps, unit-lower forward solve
// (solve.rs:106-133), upper backward solve (solve.rs:137-182), all in
// nalgebra's per-element order.  b lives in LDS; one barrier per step.
template <typename T>
__global__ __launch_bounds__(1024) void getrs_kernel(const T *LU, int lda, int n, const int *ipiv, T *B, int ldb,
                                                     int nrhs, int *info, int exact)
{
    extern __shared__ unsigned char smem_raw[];
    T *b = reinterpret_cast<T *>(smem_raw);
    __shared__ int fail;
    const int t = threadIdx.x;
    if (t == 0) fail = 0;
    for (int c = 0; c < nrhs; ++c) {
        T *bc = B + (size_t)c * ldb;
        for (int r = t; r < n; r += blockDim.x) b[r] = bc[r];
        __syncthreads();
        if (t == 0) {
            for (int i = 0; i < n; ++i) {
                const int p = ipiv[i];
                if (p != i) {
                    const T v = b[i];
                    b[i] = b[p];
                    b[p] = v;
                }
            }
        }
        __syncthreads();
        for (int i = 0; i + 1 < n; ++i) {
            const T coeff = b[i];  // b_i / 1
            for (int r = i + 1 + t; r < n; r += blockDim.x)
                b[r] = sub_prod(b[r], coeff, LU[r + (size_t)i * lda], exact != 0);
            __syncthreads();
        }
        for (int i = n - 1; i >= 0; --i) {
            const T d = LU[i + (size_t)i * lda];
            if (d == T(0)) {
                if (t == 0 && fail == 0) fail = i + 1;
                break;
            }
            const T coeff = b[i] / d;
            __syncthreads();
            if (t == 0) b[i] = coeff;
            for (int r = t; r < i; r += blockDim.x)
                b[r] = sub_prod(b[r], coeff, LU[r + (size_t)i * lda], exact != 0);
            __syncthreads();
        }
        __syncthreads();
        for (int r = t; r < n; r += blockDim.x) bc[r] = b[r];
        __syncthreads();
        if (fail) break;
    }
    if (t == 0 && info) *info = fail;
}

// ---------------------------------------------------------------- blocked Cholesky for n > 64
// Rows below a factored 64-wide diagonal block: for each row r and panel
// column j in order, a_rj = (a_rj - sum_{k<j} fl(L_jk * a_rk)) / L_jj with the
// products subtracted one at a time in increasing k (cholesky.rs:226-262).
__global__ __launch_bounds__(256) void chol_panel_rows_kernel(double *A, int lda, int j0, int nb, int r0, int r1,
                                                              int exact)
{
    __shared__ double Ld[64][65];
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
        const int rr = e % nb, cc = e / nb;
        Ld[rr][cc] = A[(j0 + rr) + (size_t)(j0 + cc) * lda];
    }
    __syncthreads();
    const int r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= r1) return;
    double x[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) x[j] = j < nb ? A[r + (size_t)(j0 + j) * lda] : 0.0;
#pragma unroll
    for (int j = 0; j < 64; ++j) {
        if (j < nb) {
#pragma unroll
            for (int k = 0; k < j; ++k) x[j] = sub_prod(x[j], Ld[j][k], x[k], exact != 0);
            x[j] = x[j] / Ld[j][j];
        }
    }
#pragma unroll
    for (int j = 0; j < 64; ++j)
        if (j < nb) A[r + (size_t)(j0 + j) * lda] = x[j];
}

__global__ void chol_info_merge_kernel(int *info, const int *blk, int j0)
{
    if (*info == 0 && *blk != 0) *info = j0 + *blk;
}

int potrf_blocked(nrocm_handle h, int n, double *A, int lda, int *info, bool exact, double substitute)
{
    hipStream_t st = h->stream;
    int *blk = (int *)h->small;  // per-block info scratch
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, st, info, 1, 0);
    LU_TRY(hipGetLastError());
    for (int j0 = 0; j0 < n; j0 += 64) {
        const int nb = std::min(64, n - j0);
        double *D = A + j0 + (size_t)j0 * lda;
        int st2 = nrocm_potrf_small(h, nb, D, lda, nullptr, 0, 0, blk, exact, substitute);
        if (st2 != NROCM_OK) return st2;
        hipLaunchKernelGGL(chol_info_merge_kernel, dim3(1), dim3(1), 0, st, info, blk, j0);
        LU_TRY(hipGetLastError());
        const int r0 = j0 + nb;
        if (r0 < n) {
            hipLaunchKernelGGL(chol_panel_rows_kernel, dim3(nrocm_cdiv(n - r0, 256)), dim3(256), 0, st, A, lda, j0,
                               nb, r0, n, exact ? 1 : 0);
            LU_TRY(hipGetLastError());
            // A22 lower -= L21 * L21^T (U element (k, c) = L21[c, k])
            double *C = A + r0 + (size_t)r0 * lda;
            const double *L21 = A + r0 + (size_t)j0 * lda;
            if (exact)
                LU_TRY((simt_update<double, true, true>(st, C, lda, n - r0, n - r0, L21, lda, L21, lda, 1, nb, 0)));
            else
                LU_TRY((simt_update<double, false, true>(st, C, lda, n - r0, n - r0, L21, lda, L21, lda, 1, nb, 0)));
        }
    }
    return NROCM_OK;
}

// out[0] = the product of the diagonal, 1 * a_00 * a_11 * ... in index order
// (LU::determinant's product, lu.rs:301-314, and Cholesky::determinant's,
// cholesky.rs:155-164), out[1] = out[0] * out[0] (Cholesky::determinant's
// modulus squared), out[2] = 0 + ln(a_00^2) + ln(a_11^2) + ... in index order
// (Cholesky::ln_determinant, cholesky.rs:166-184).  Products and sums are
// one dependent chain: lane 0 walks it while the others prefetch the diagonal.
__global__ __launch_bounds__(64) void diag_det_kernel(const double *A, int lda, int n, double *out)
{
    __shared__ double d[64];
    const int lane = threadIdx.x;
    double prod = 1.0, lnsum = 0.0;
    for (int i0 = 0; i0 < n; i0 += 64) {
        d[lane] = (i0 + lane < n) ? A[(size_t)(i0 + lane) * ((size_t)lda + 1)] : 1.0;
        __syncthreads();
        if (lane == 0) {
            const int cnt = min(64, n - i0);
            for (int q = 0; q < cnt; ++q) {
                prod = prod * d[q];
                lnsum = lnsum + log(d[q] * d[q]);
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        out[0] = prod;
        out[1] = prod * prod;
        out[2] = lnsum;
    }
}

// Cholesky solve for n > 64 on one workgroup: forward (solve.rs:488-519) and
// backward with the 8-accumulator dotc order (solve.rs:697-755) or a tree sum.
__global__ __launch_bounds__(256) void potrs_kernel(const double *L, int lda, int n, double *B, int ldb, int nrhs,
                                                    int exact)
{
    extern __shared__ unsigned char smem_raw[];
    double *b = reinterpret_cast<double *>(smem_raw);
    __shared__ double part[256];
    const int t = threadIdx.x;
    for (int c = 0; c < nrhs; ++c) {
        double *bc = B + (size_t)c * ldb;
        for (int r = t; r < n; r += blockDim.x) b[r] = bc[r];
        __syncthreads();
        for (int i = 0; i < n; ++i) {
            const double coeff = b[i] / L[i + (size_t)i * lda];
            __syncthreads();
            if (t == 0) b[i] = coeff;
            for (int r = i + 1 + t; r < n; r += blockDim.x)
                b[r] = sub_prod(b[r], coeff, L[r + (size_t)i * lda], exact != 0);
            __syncthreads();
        }
        for (int i = n - 1; i >= 0; --i) {
            const int len = n - i - 1;
            const double *li = L + (i + 1) + (size_t)i * lda;
            const double *bi = b + i + 1;
            double dot = 0.0;
            if (exact) {
                if (t < 8) {
                    const int full = len & ~7;
                    double a = 0.0;
                    for (int e = t; e < full; e += 8) a += li[e] * bi[e];
                    part[t] = a;
                }
                __syncthreads();
                if (t == 0) {
                    const int full = len & ~7;
                    double res = 0.0;
                    res += part[0] + part[4];
                    res += part[1] + part[5];
                    res += part[2] + part[6];
                    res += part[3] + part[7];
                    for (int e = full; e < len; ++e) res += li[e] * bi[e];
                    part[8] = res;
                }
            } else {
                double a = 0.0;
                for (int e = t; e < len; e += blockDim.x) a = fma(li[e], bi[e], a);
                part[t] = a;
                __syncthreads();
                for (int s = 128; s > 0; s >>= 1) {
                    if (t < s) part[t] += part[t + s];
                    __syncthreads();
                }
                if (t == 0) part[8] = part[0];
            }
            __syncthreads();
            if (t == 0) b[i] = (b[i] - part[8]) / L[i + (size_t)i * lda];
            __syncthreads();
        }
        for (int r = t; r < n; r += blockDim.x) bc[r] = b[r];
        __syncthreads();
    }
}

// The panel step of the windowed LU for the block at k0: the window factor,
// then A21 U11^{-1} for the rows below the window.  Touches only the block's
// own columns, besides ipiv, perm, info and the row plan.
template <typename T>
hipError_t window_panel(hipStream_t st, T *A, int lda, int n, int k0, int *ipiv, int *info, int *plan, int *perm)
{
    const int nb = std::min(64, n - k0);
    const int kw = std::min(n, k0 + WINDOW_ROWS);
    hipLaunchKernelGGL((window_panel_kernel<T>), dim3(1), dim3(WINDOW_ROWS), 0, st, A, lda, k0, nb, kw, ipiv, info,
                       plan, perm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || kw >= n) return e;
    hipLaunchKernelGGL((trsm_right_upper_kernel<T>), dim3(nrocm_cdiv(n - kw, TRSM_RU_THREADS)),
                       dim3(TRSM_RU_THREADS), 0, st, A, lda, k0, nb, kw, n);
    return hipGetLastError();
}

// Blocked LU with the pivot search of each 64-column block restricted to a
// window of WINDOW_ROWS rows (window_panel_kernel), then the window's row
// plan on the other columns, A21 U11^{-1} for the rows below the window, U12
// and the trailing update.  ipiv has the LAPACK meaning, as for getrf_impl;
// perm (optional) receives the composed row permutation.
//
// Look-ahead of depth one: the panel step of block k+1 needs only the
// trailing update of its own 64 columns, so that part of block k's update
// runs first and the panel step then runs on an auxiliary stream beside the
// update of the remaining columns.  The row plan of block k+1 is applied
// after both have finished (it touches those columns too).
template <typename T>
int getrf_window_impl(nrocm_handle h, hipStream_t st, int n, T *A, int lda, int *ipiv, int *perm, int *info,
                      NrocmTrailing trail)
{
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, st, info, 1, 0);
    LU_TRY(hipGetLastError());
    if (perm) {
        hipLaunchKernelGGL(iota_kernel, dim3(nrocm_cdiv(n, 256)), dim3(256), 0, st, perm, n);
        LU_TRY(hipGetLastError());
    }
    int *plan = (int *)h->small + NROCM_SMALL_PLAN;
    constexpr int NB = 64;
    LU_TRY(window_panel<T>(st, A, lda, n, 0, ipiv, info, plan, perm));
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = std::min(NB, n - k0);
        const int k1 = k0 + nb;
        const int ncol = k0 + (n - k1);
        if (ncol > 0) {
            hipLaunchKernelGGL((apply_plan_kernel<T>), dim3(nrocm_cdiv(ncol, 2)), dim3(256), 0, st, A, lda, plan, 0,
                               k0, k1, n);
            LU_TRY(hipGetLastError());
        }
        if (k1 >= n) break;
        LU_TRY(trsm<T>(st, A, lda, k0, nb, k1, n, false));
        const int k2 = std::min(n, k1 + NB);
        LU_TRY(trailing<T>(st, trail, A, lda, k1, n, k1, k2, k0, nb));
        int s = nrocm_fork_on(h, st, 1);
        if (s != NROCM_OK) return s;
        LU_TRY(window_panel<T>(h->aux[0], A, lda, n, k1, ipiv, info, plan, perm));
        // FP32: two workgroups per CU on all but a few CUs, which stay free for
        // the panel step running beside this update (measured: FP32 gains, the
        // FP64 update with its smaller tiles does not)
        LU_TRY(trailing<T>(st, trail, A, lda, k1, n, k2, n, k0, nb,
                           sizeof(T) == 4 ? 2 * std::max(1, h->num_cus - 8) : 0));
        s = nrocm_join_on(h, st, 1);
        if (s != NROCM_OK) return s;
    }
    return NROCM_OK;
}

}  // namespace

// ---------------------------------------------------------------- internal entry points

// The blocked factorisation is thousands of small launches at large n, so the
// sequence is captured once per (shape, buffers, flavour) and replayed.
template <typename T>
int getrf_graph(nrocm_handle h, int m, int n, T *A, int lda, int *ipiv, int *info, NrocmTrailing trail)
{
    if (std::min(m, n) <= 64) return getrf_impl<T>(h, h->stream, m, n, A, lda, ipiv, info, trail);
    const long long key[8] = {1000 + (long long)sizeof(T), m, n, lda, (long long)A, (long long)ipiv,
                              (long long)info, (long long)trail};
    return nrocm_graph_run(h, key, [&](hipStream_t st) {
        return getrf_impl<T>(h, st, m, n, A, lda, ipiv, info, trail);
    });
}

int nrocm_getrf_d(nrocm_handle h, hipStream_t st, int m, int n, double *A, int lda, int *ipiv, int *info,
                  NrocmTrailing trail)
{
    (void)st;
    return getrf_graph<double>(h, m, n, A, lda, ipiv, info, trail);
}

int nrocm_getrf_s(nrocm_handle h, hipStream_t st, int m, int n, float *A, int lda, int *ipiv, int *info,
                  NrocmTrailing trail)
{
    (void)st;
    return getrf_graph<float>(h, m, n, A, lda, ipiv, info, trail);
}

template <typename T>
int getrf_window_graph(nrocm_handle h, int n, T *A, int lda, int *ipiv, int *perm, int *info, NrocmTrailing trail)
{
    if (n <= 64) return getrf_window_impl<T>(h, h->stream, n, A, lda, ipiv, perm, info, trail);
    const int s = nrocm_aux_init(h);  // streams and events exist before the capture
    if (s != NROCM_OK) return s;
    const long long key[8] = {3000 + (long long)sizeof(T) + 16 * (long long)trail, n, lda, (long long)A,
                              (long long)ipiv, (long long)perm, (long long)info, 0};
    return nrocm_graph_run(
        h, key, [&](hipStream_t st) { return getrf_window_impl<T>(h, st, n, A, lda, ipiv, perm, info, trail); });
}

int nrocm_getrf_window_d(nrocm_handle h, int n, double *A, int lda, int *ipiv, int *perm, int *info,
                         NrocmTrailing trail)
{
    return getrf_window_graph<double>(h, n, A, lda, ipiv, perm, info, trail);
}

int nrocm_getrf_window_s(nrocm_handle h, int n, float *A, int lda, int *ipiv, int *perm, int *info,
                         NrocmTrailing trail)
{
    return getrf_window_graph<float>(h, n, A, lda, ipiv, perm, info, trail);
}

// ---------------------------------------------------------------- C ABI

extern "C" {

int nrocm_dgetrf(nrocm_handle h, int m, int n, double *a, int lda, int *ipiv, int *info, int mode)
{
    if (!h) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (!a && m > 0 && n > 0) return -4;
    if (lda < std::max(1, m)) return -5;
    if (!ipiv && std::min(m, n) > 0) return -6;
    if (!info) return -7;
    if (mode != NROCM_EXACT && mode != NROCM_FAST) return -8;
    return getrf_graph<double>(h, m, n, a, lda, ipiv, info, mode == NROCM_EXACT ? TRAIL_EXACT : TRAIL_F64);
}

int nrocm_dgetrs(nrocm_handle h, int n, int nrhs, const double *lu, int lda, const int *ipiv, double *b, int ldb,
                 int *info, int mode)
{
    if (!h) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (!lu && n > 0) return -4;
    if (lda < std::max(1, n)) return -5;
    if (!ipiv && n > 0) return -6;
    if (!b && n > 0 && nrhs > 0) return -7;
    if (ldb < std::max(1, n)) return -8;
    if (!info) return -9;
    if (mode != NROCM_EXACT && mode != NROCM_FAST) return -10;
    if (n >= NROCM_GETRS_LOOKBACK_MIN) {
        // one workgroup per 64-row block, pipelined through published blocks
        void *ws = nullptr;
        const int s = nrocm_workspace(h, nrocm_getrs_scratch_bytes(n), &ws);
        if (s != NROCM_OK) return s;
        return nrocm_getrs_lookback(h, h->stream, n, nrhs, lu, lda, ipiv, b, ldb, info, mode == NROCM_EXACT, ws);
    }
    hipLaunchKernelGGL((getrs_kernel<double>), dim3(1), dim3(1024), (size_t)std::max(n, 1) * sizeof(double),
                       h->stream, lu, lda, n, ipiv, b, ldb, nrhs, info, mode == NROCM_EXACT ? 1 : 0);
    NROCM_CHECK_LAUNCH(h);
    return NROCM_OK;
}

int nrocm_dpotrf(nrocm_handle h, int n, double *a, int lda, int *info, int mode)
{
    if (!h) return -1;
    if (n < 0) return -2;
    if (!a && n > 0) return -3;
    if (lda < std::max(1, n)) return -4;
    if (!info) return -5;
    if (mode != NROCM_EXACT && mode != NROCM_FAST) return -6;
    if (n <= 64) return nrocm_potrf_small(h, n, a, lda, nullptr, 0, 0, info, mode == NROCM_EXACT);
    return potrf_blocked(h, n, a, lda, info, mode == NROCM_EXACT, 0.0);
}

int nrocm_dpotrf_substitute(nrocm_handle h, int n, double *a, int lda, double substitute, int *info, int mode)
{
    if (!h) return -1;
    if (n < 0) return -2;
    if (!a && n > 0) return -3;
    if (lda < std::max(1, n)) return -4;
    if (!info) return -6;
    if (mode != NROCM_EXACT && mode != NROCM_FAST) return -7;
    if (n <= 64) return nrocm_potrf_small(h, n, a, lda, nullptr, 0, 0, info, mode == NROCM_EXACT, substitute);
    return potrf_blocked(h, n, a, lda, info, mode == NROCM_EXACT, substitute);
}

int nrocm_ddiag_det(nrocm_handle h, int n, const double *a, int lda, double *out)
{
    if (!h) return -1;
    if (n < 0) return -2;
    if (!a && n > 0) return -3;
    if (lda < std::max(1, n)) return -4;
    if (!out) return -5;
    hipLaunchKernelGGL(diag_det_kernel, dim3(1), dim3(64), 0, h->stream, a, lda, n, out);
    NROCM_CHECK_LAUNCH(h);
    return NROCM_OK;
}

int nrocm_dpotrs(nrocm_handle h, int n, int nrhs, const double *l, int lda, double *b, int ldb, int mode)
{
    if (!h) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (!l && n > 0) return -4;
    if (lda < std::max(1, n)) return -5;
    if (!b && n > 0 && nrhs > 0) return -6;
    if (ldb < std::max(1, n)) return -7;
    if (mode != NROCM_EXACT && mode != NROCM_FAST) return -8;
    if (n == 0 || nrhs == 0) return NROCM_OK;
    if (n >= NROCM_GETRS_LOOKBACK_MIN) {
        void *ws = nullptr;
        const int s = nrocm_workspace(h, nrocm_getrs_scratch_bytes(n), &ws);
        if (s != NROCM_OK) return s;
        return nrocm_potrs_lookback(h, h->stream, n, nrhs, l, lda, b, ldb, mode == NROCM_EXACT, ws);
    }
    hipLaunchKernelGGL(potrs_kernel, dim3(1), dim3(256), (size_t)n * sizeof(double), h->stream, l, lda, n, b, ldb,
                       nrhs, mode == NROCM_EXACT ? 1 : 0);
    NROCM_CHECK_LAUNCH(h);
    return NROCM_OK;
}

}  // extern "C"
