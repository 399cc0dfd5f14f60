// Device kernels of the blocked LU factorisation and the triangular solves
// shared by LU (double) and the mixed-precision path (float working copy).
//
// Exactness argument (NROCM_EXACT).  nalgebra's LU::new (src/linalg/lu.rs:
// 93-122, 337-389) updates every trailing element (r, c) once per step k as
// A[r,c] = fl(A[r,c] - fl(u_kc * l_rk)), in increasing k, and swaps whole rows.
// A right-looking blocked schedule that (a) factors a panel column by column
// with those same operations, (b) applies the panel's row swaps to the other
// columns, (c) forms U12 by forward substitution in increasing k and (d)
// applies the rank-nb trailing update to each element in increasing k, with
// no fused multiply-add anywhere, performs for every element the identical
// sequence of IEEE operations.  Hence identical values and identical pivots.
#pragma once
#include "nrocm_internal.hpp"

namespace nrocm_lu {

// Two wavefronts per SIMD (up to 256 VGPRs per thread); the per-column
// barrier synchronises eight wavefronts.
constexpr int PANEL_THREADS = 512;

// Panel width per rows-per-thread: RPT * W values of the panel plus the W
// values of the pivot row stay in registers without spilling.
__host__ __device__ constexpr int panel_width(int rpt, int elem_bytes)
{
    return elem_bytes == 8 ? (rpt <= 1 ? 32 : rpt == 2 ? 16 : rpt == 4 ? 8 : rpt == 8 ? 4 : rpt == 16 ? 2 : 1)
                           : (rpt <= 2 ? 32 : rpt == 4 ? 16 : rpt == 8 ? 8 : rpt == 16 ? 4 : 2);
}

// pivot candidate: |value| and global row; `first-index strict max` order.
struct Cand {
    double v;
    int idx;
};

// a beats b under icamax's rule (src/base/min_max.rs:221-241): larger |x|,
// ties to the lower index.  NaN never wins except at the column head, which
// the caller encodes as +inf with the lowest index.
__device__ __forceinline__ bool beats(double av, int ai, double bv, int bi)
{
    return av > bv || (av == bv && ai < bi);
}

// Returns v unchanged.  Copy loops whose values pass through this are not
// turned into memcpy calls, which would keep register arrays in scratch.
template <typename T>
__device__ __forceinline__ T pass(T v)
{
    asm volatile("" : "+v"(v));
    return v;
}

template <typename T>
__device__ __forceinline__ T sub_prod(T acc, T x, T y, bool exact)
{
    if (exact) {
        T p = x * y;
        return acc - p;
    }
    return fma(-x, y, acc);
}

// Unblocked LU of the panel A[row0.., col0..col0+w) held in registers by one
// PANEL_THREADS-thread workgroup; thread t holds rows row0 + t + PANEL_THREADS*s, s < RPT.
// Pivot rows are recorded as global row indices in ipiv[col0 + c].
template <typename T, int RPT, int W>
__global__ __launch_bounds__(PANEL_THREADS) void panel_kernel(T *A, int lda, int m, int col0, int w,
                                                              int *ipiv, int *info)
{
    __shared__ double red_v[PANEL_THREADS / 64];
    __shared__ int red_i[PANEL_THREADS / 64];
    __shared__ T rowbuf[2][2][W];  // [parity][0: pivot row, 1: row i]
    const int t = threadIdx.x;
    const int lane = t & 63, wv = t >> 6;
    const int row0 = col0;

    T x[RPT][W];
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
        const int r = row0 + t + PANEL_THREADS * s;
#pragma unroll
        for (int c = 0; c < W; ++c) x[s][c] = (r < m && c < w) ? A[r + (size_t)(col0 + c) * lda] : T(0);
    }

    // No break/continue in this loop: it must unroll completely so that x
    // stays in registers.
#pragma unroll
    for (int c = 0; c < W; ++c) {
        if (c < w) {
            const int i = row0 + c;  // pivot row position (global)
            // ---- pivot search over rows i..m-1 of column c
            double bv = -1.0;
            int bi = 0x7fffffff;
#pragma unroll
            for (int s = 0; s < RPT; ++s) {
                const int r = row0 + t + PANEL_THREADS * s;
                if (r >= i && r < m) {
                    double v = fabs((double)x[s][c]);
                    if (v != v) v = (r == i) ? __builtin_inf() : -1.0;  // NaN rule
                    if (beats(v, r, bv, bi)) { bv = v; bi = r; }
                }
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const double ov = __shfl_xor(bv, off);
                const int oi = __shfl_xor(bi, off);
                if (beats(ov, oi, bv, bi)) { bv = ov; bi = oi; }
            }
            if (lane == 0) { red_v[wv] = bv; red_i[wv] = bi; }
            __syncthreads();
            bv = red_v[0];
            bi = red_i[0];
#pragma unroll
            for (int q = 1; q < PANEL_THREADS / 64; ++q)
                if (beats(red_v[q], red_i[q], bv, bi)) { bv = red_v[q]; bi = red_i[q]; }
            const int piv = bi;
            // ---- exchange rows i and piv through LDS
            T(*rb)[W] = rowbuf[c & 1];
#pragma unroll
            for (int s = 0; s < RPT; ++s) {
                const int r = row0 + t + PANEL_THREADS * s;
                if (r == piv) {
#pragma unroll
                    for (int q = 0; q < W; ++q) rb[0][q] = pass(x[s][q]);
                }
                if (r == i) {
#pragma unroll
                    for (int q = 0; q < W; ++q) rb[1][q] = pass(x[s][q]);
                }
            }
            __syncthreads();
            const T d = rb[0][c];  // the pivot value
            if (t == 0) ipiv[col0 + c] = (d == T(0)) ? i : piv;
            if (d == T(0)) {
                // lu.rs:107-110: all-zero column, nothing is swapped or updated
                if (t == 0 && info && *info == 0) *info = i + 1;
            } else {
                const T inv = T(1) / d;
#pragma unroll
                for (int s = 0; s < RPT; ++s) {
                    const int r = row0 + t + PANEL_THREADS * s;
                    if (r == i) {
#pragma unroll
                        for (int q = 0; q < W; ++q) x[s][q] = pass(rb[0][q]);
                    } else if (r == piv) {
#pragma unroll
                        for (int q = 0; q < W; ++q) x[s][q] = pass(rb[1][q]);
                    }
                    if (r > i && r < m) {
                        const T l = x[s][c] * inv;
                        x[s][c] = l;
#pragma unroll
                        for (int q = 0; q < W; ++q)  // constant trip count: unrolls before c does
                            if (q > c) x[s][q] = sub_prod(x[s][q], rb[0][q], l, true);
                    }
                }
            }
        }
    }
    // A copy of t the compiler cannot see through, so the load addresses and
    // row predicates are recomputed here instead of living in registers.
    int tt = t;
    asm volatile("" : "+v"(tt));
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
        const int r = row0 + tt + PANEL_THREADS * s;
        if (r < m) {
#pragma unroll
            for (int c = 0; c < W; ++c)
                if (c < w) A[r + (size_t)(col0 + c) * lda] = x[s][c];
        }
    }
}

// Row swaps ipiv[k0..k1) applied in order to columns [c0, c1) (one thread per column).
template <typename T>
__global__ void laswp_kernel(T *A, int lda, int c0, int c1, const int *ipiv, int k0, int k1)
{
    const int c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c1) return;
    T *col = A + (size_t)c * lda;
    for (int i = k0; i < k1; ++i) {
        const int p = ipiv[i];
        if (p != i) {
            const T a = col[i];
            col[i] = col[p];
            col[p] = a;
        }
    }
}

// U12 = L11^{-1} A12 for a unit-lower L11 (k x k, k <= 64) at A[r0, r0] and the
// columns [c0, c1): one wavefront per column, lane r holds row r0 + r.
// Per element: A[r,c] = fl(A[r,c] - fl(u_ic * l_ri)) for i = 0..r-1 in order.
template <typename T>
__global__ __launch_bounds__(256) void trsm_unit_lower_kernel(T *A, int lda, int r0, int k, int c0,
                                                              int c1, int exact)
{
    __shared__ T L[64][65];
    for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
        const int rr = e % k, cc = e / k;
        L[rr][cc] = A[(r0 + rr) + (size_t)(r0 + cc) * lda];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    for (int c = c0 + blockIdx.x * 4 + wv; c < c1; c += gridDim.x * 4) {
        T *col = A + (size_t)c * lda + r0;
        T v = lane < k ? col[lane] : T(0);
        for (int i = 0; i < k - 1; ++i) {
            const T u = __shfl(v, i);
            if (lane > i && lane < k) v = sub_prod(v, u, L[lane][i], exact != 0);
        }
        if (lane < k) col[lane] = v;
    }
}

// C (m x nc, ldc) -= L (m x K, ldl) * U (K x nc; element (q, c) at
// U[q*rsu + c*csu]).  SIMT register tiles: 256 threads per 64 x 64 tile,
// 4 x 4 elements per thread, each element's K terms applied in increasing k
// (EXACT: fl(C - fl(u*l)), otherwise fma).  LOWER: only elements with
// r + diag_off >= c are written (the strict upper triangle stays untouched).
template <typename T, bool EXACT, bool LOWER>
__global__ __launch_bounds__(256) void update_kernel(T *C, int ldc, int m, int nc, const T *Lp, int ldl,
                                                     const T *Up, long rsu, long csu, int K, int diag_off)
{
    constexpr int TM = 64, TN = 64, KC = 16;
    __shared__ T Ls[KC][TM + 1];
    __shared__ T Us[KC][TN + 1];
    const int tr = blockIdx.x * TM;
    const int tc = blockIdx.y * TN;
    if (LOWER && tr + TM - 1 + diag_off < tc) return;  // tile entirely above the diagonal
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    T acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int r = tr + tx + 16 * a, c = tc + ty + 16 * b;
            acc[a][b] = (r < m && c < nc) ? C[r + (size_t)c * ldc] : T(0);
        }
    for (int kk = 0; kk < K; kk += KC) {
        for (int e = threadIdx.x; e < KC * TM; e += 256) {
            const int rr = e % TM, q = e / TM;
            const int r = tr + rr, kq = kk + q;
            Ls[q][rr] = (r < m && kq < K) ? Lp[r + (size_t)kq * ldl] : T(0);
        }
        for (int e = threadIdx.x; e < KC * TN; e += 256) {
            const int q = e % KC, cc = e / KC;
            const int c = tc + cc, kq = kk + q;
            Us[q][cc] = (c < nc && kq < K) ? Up[(long)kq * rsu + (long)c * csu] : T(0);
        }
        __syncthreads();
        const int kmax = (K - kk) < KC ? (K - kk) : KC;
        for (int q = 0; q < kmax; ++q) {
            T lv[4], uv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) lv[a] = Ls[q][tx + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; ++b) uv[b] = Us[q][ty + 16 * b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = sub_prod(acc[a][b], uv[b], lv[a], EXACT);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int r = tr + tx + 16 * a, c = tc + ty + 16 * b;
            if (r < m && c < nc && (!LOWER || r + diag_off >= c)) C[r + (size_t)c * ldc] = acc[a][b];
        }
}

// For each column c of [c0, c1) (one workgroup each) after the base block
// [b0, b0+w) has been factored: apply its w row swaps to the column; when
// c >= b0 + w also form U12 (rows b0..b0+w, unit-lower solve with L11) and
// apply the rank-w update to rows b0+w..m.  Every element's terms are applied
// in increasing k (EXACT: fl(a - fl(u*l))), as nalgebra's gauss steps do.
template <typename T>
__global__ __launch_bounds__(256) void panel_cols_kernel(T *A, int lda, int m, int b0, int w, int cl, int cr0,
                                                         int cr1, const int *ipiv, int exact)
{
    __shared__ int pos[64];      // touched row positions
    __shared__ int who[64];      // original position whose value ends at pos[q]
    __shared__ int piv[32];
    __shared__ int nt;
    __shared__ T L11[32][33];
    __shared__ T U[32];
    const int t = threadIdx.x;
    const int idx = blockIdx.x;
    // columns [cl, b0) swap only; columns [cr0, cr1) swap + solve + update
    const int c = idx < b0 - cl ? cl + idx : cr0 + (idx - (b0 - cl));
    const bool full = c >= cr0;
    T *col = A + (size_t)c * lda;
    if (t < w) piv[t] = ipiv[b0 + t];
    if (full)
        for (int e = t; e < w * w; e += blockDim.x)
            L11[e % w][e / w] = A[(b0 + e % w) + (size_t)(b0 + e / w) * lda];
    __syncthreads();
    // compose the w swaps on indices only
    if (t == 0) {
        int n_ = 0;
        for (int i = 0; i < w; ++i) {
            const int p = piv[i];
            int a = -1, b = -1;
            for (int q = 0; q < n_; ++q) {
                if (pos[q] == b0 + i) a = q;
                if (pos[q] == p) b = q;
            }
            if (a < 0) { a = n_; pos[n_] = b0 + i; who[n_] = b0 + i; ++n_; }
            if (b < 0) { b = n_; pos[n_] = p; who[n_] = p; ++n_; }
            const int tw = who[a];
            who[a] = who[b];
            who[b] = tw;
        }
        nt = n_;
    }
    __syncthreads();
    T v = T(0);
    if (t < nt) v = col[who[t]];
    __syncthreads();
    if (t < nt) col[pos[t]] = v;
    __syncthreads();
    if (!full) return;
    // U12 for this column: rows b0..b0+w, unit lower L11 (one wavefront)
    if (t < 64) {
        T u = t < w ? col[b0 + t] : T(0);
        for (int i = 0; i < w - 1; ++i) {
            const T ui = __shfl(u, i);
            if (t > i && t < w) u = sub_prod(u, ui, L11[t][i], exact != 0);
        }
        if (t < w) {
            col[b0 + t] = u;
            U[t] = u;
        }
    }
    __syncthreads();
    // rank-w update of rows b0+w..m, terms in increasing k; RU rows per
    // thread have their loads in flight together
    constexpr int RU = 8;
    for (int rb = b0 + w + t; rb < m; rb += 256 * RU) {
        T a[RU];
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const int r = rb + 256 * u;
            a[u] = r < m ? col[r] : T(0);
        }
        for (int q = 0; q < w; ++q) {
            const T uq = U[q];
            const T *Lq = A + (size_t)(b0 + q) * lda;
            T l[RU];
#pragma unroll
            for (int u = 0; u < RU; ++u) {
                const int r = rb + 256 * u;
                l[u] = r < m ? Lq[r] : T(0);
            }
#pragma unroll
            for (int u = 0; u < RU; ++u) a[u] = sub_prod(a[u], uq, l[u], exact != 0);
        }
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const int r = rb + 256 * u;
            if (r < m) col[r] = a[u];
        }
    }
}

// Composition of the row swaps ipiv[k0..k1) (k1 - k0 <= 64, applied in
// order, as LU::new swaps whole rows at every step) into a row plan:
// plan[0] = number nt of touched positions; after the swaps, position
// plan[1 + s] holds the value that was at position plan[1 + PLAN_MAX + s].
constexpr int PLAN_MAX = 128;

__global__ __launch_bounds__(128) void swap_plan_kernel(const int *ipiv, int k0, int k1, int *plan)
{
    __shared__ int pos[PLAN_MAX], who[PLAN_MAX], pv[PLAN_MAX / 2];
    __shared__ int cnt, sa, sb;
    const int t = threadIdx.x;
    if (t < k1 - k0) pv[t] = ipiv[k0 + t];
    if (t == 0) cnt = 0;
    __syncthreads();
    for (int i = k0; i < k1; ++i) {
        const int p = pv[i - k0];
        if (p == i) continue;  // uniform: no swap at this step
        if (t == 0) {
            sa = -1;
            sb = -1;
        }
        __syncthreads();
        if (t < cnt) {
            const int ps = pos[t];  // positions are distinct: at most one match each
            if (ps == i) sa = t;
            if (ps == p) sb = t;
        }
        __syncthreads();
        if (t == 0) {
            int a = sa, b = sb;
            if (a < 0) {
                a = cnt++;
                pos[a] = i;
                who[a] = i;
            }
            if (b < 0) {
                b = cnt++;
                pos[b] = p;
                who[b] = p;
            }
            const int tw = who[a];
            who[a] = who[b];
            who[b] = tw;
        }
        __syncthreads();
    }
    const int nt = cnt;
    if (t == 0) plan[0] = nt;
    if (t < nt) {
        plan[1 + t] = pos[t];
        plan[1 + PLAN_MAX + t] = who[t];
    }
}

// Applies a row plan to the columns [ca0, ca1) and [cb0, cb1): two columns
// per workgroup, one thread per touched position; every value is read before
// any is written.
template <typename T>
__global__ __launch_bounds__(256) void apply_plan_kernel(T *A, int lda, const int *plan, int ca0, int ca1, int cb0,
                                                         int cb1)
{
    const int t = threadIdx.x;
    const int slot = t & (PLAN_MAX - 1);
    const int na = ca1 > ca0 ? ca1 - ca0 : 0;
    const int nbc = cb1 > cb0 ? cb1 - cb0 : 0;
    const int ci = blockIdx.x * 2 + (t >> 7);
    const bool valid = ci < na + nbc;
    const int c = ci < na ? ca0 + ci : cb0 + (ci - na);
    const int nt = plan[0];
    T v = T(0);
    if (valid && slot < nt) v = A[plan[1 + PLAN_MAX + slot] + (size_t)c * lda];
    __syncthreads();
    if (valid && slot < nt) A[plan[1 + slot] + (size_t)c * lda] = v;
}

__global__ void fill_int_kernel(int *p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace nrocm_lu
