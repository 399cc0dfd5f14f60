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

// a beats b under icamax's rule (src/base/min_max.rs:221-241): larger |x|,
// ties to the lower index.  NaN never wins except at the column head, which
// the caller encodes as +inf with the lowest index.
template <typename V>
__device__ __forceinline__ bool beats_v(V av, int ai, V bv, int bi)
{
    return av > bv || (av == bv && ai < bi);
}

// 32-bit DPP move with `old` kept where the source lane is outside the pattern.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp32(int old, int v)
{
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWMASK, 0xf, false);
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double old, double v)
{
    const long long o = __double_as_longlong(old), x = __double_as_longlong(v);
    const int lo = dpp32<CTRL, ROWMASK>((int)o, (int)x);
    const int hi = dpp32<CTRL, ROWMASK>((int)(o >> 32), (int)(x >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_f32(float old, float v)
{
    return __int_as_float(dpp32<CTRL, ROWMASK>(__float_as_int(old), __float_as_int(v)));
}

template <typename V, int CTRL, int ROWMASK>
__device__ __forceinline__ void argmax_dpp_step(V &bv, int &bi)
{
    V ov;
    if constexpr (sizeof(V) == 8) ov = dpp_f64<CTRL, ROWMASK>(V(-1), bv);
    else ov = dpp_f32<CTRL, ROWMASK>(V(-1), bv);
    const int oi = dpp32<CTRL, ROWMASK>(0x7fffffff, bi);
    if (beats_v(ov, oi, bv, bi)) {
        bv = ov;
        bi = oi;
    }
}

// Wavefront argmax under the first-index strict-max rule; lane 63 ends with
// the result.  Row scans by row_shr 1, 2, 4, 8, then row_bcast 15 into rows 1
// and 3 and row_bcast 31 into rows 2 and 3.  The rule is a total order on
// (value, row) pairs, so the combination order does not change the winner.
template <typename V>
__device__ __forceinline__ void argmax_wave(V &bv, int &bi)
{
    argmax_dpp_step<V, 0x111, 0xf>(bv, bi);
    argmax_dpp_step<V, 0x112, 0xf>(bv, bi);
    argmax_dpp_step<V, 0x114, 0xf>(bv, bi);
    argmax_dpp_step<V, 0x118, 0xf>(bv, bi);
    argmax_dpp_step<V, 0x142, 0xa>(bv, bi);
    argmax_dpp_step<V, 0x143, 0xc>(bv, bi);
}

// Returns v unchanged.  Copy loops whose values pass through this are not
// turned into memcpy calls, which would keep register arrays in scratch.
template <typename T>
__device__ __forceinline__ T pass(T v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// nalgebra's axpy with factor -x: fl(fl((-x) * y) + acc), the negation on x
// before the product as in gauss_step (lu.rs:353, :386).  In FP64 the operand
// order is pinned (-x first in the product, the product first in the sum), as
// the reference's x86 code has it, so that NaN results carry the same sign
// (with two NaN operands the first one is returned); NROCM_FAST fuses it.
template <typename T>
__device__ __forceinline__ T sub_prod(T acc, T x, T y, bool exact)
{
    if (exact) {
        if constexpr (sizeof(T) == 8) {
            double r;
            asm("v_mul_f64 %0, -%1, %2\n\tv_add_f64 %0, %0, %3" : "=&v"(r) : "v"(x), "v"(y), "v"(acc));
            return r;
        } else {
            const T p = (-x) * y;
            return p + acc;
        }
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
    __shared__ T red_v[PANEL_THREADS / 64];
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
            T bv = T(-1);
            int bi = 0x7fffffff;
#pragma unroll
            for (int s = 0; s < RPT; ++s) {
                const int r = row0 + t + PANEL_THREADS * s;
                T v = fabs(x[s][c]);
                v = (v != v) ? ((r == i) ? T(__builtin_inf()) : T(-1)) : v;  // NaN rule
                v = (r >= i && r < m) ? v : T(-1);
                // a thread visits its rows in increasing order: a tie keeps the lower row
                if (v > bv) { bv = v; bi = r; }
            }
            argmax_wave<T>(bv, bi);
            if (lane == 63) { red_v[wv] = bv; red_i[wv] = bi; }
            lds_barrier();
            bv = red_v[0];
            bi = red_i[0];
#pragma unroll
            for (int q = 1; q < PANEL_THREADS / 64; ++q)
                if (beats_v(red_v[q], red_i[q], bv, bi)) { bv = red_v[q]; bi = red_i[q]; }
            const int piv = bi;
            // ---- exchange rows i and piv through LDS (only their owners copy)
            T(*rb)[W] = rowbuf[c & 1];
            const int opiv = piv - row0 - t, oi = i - row0 - t;  // multiples of PANEL_THREADS for the owners
            if (opiv >= 0 && opiv % PANEL_THREADS == 0) {
#pragma unroll
                for (int s = 0; s < RPT; ++s)
                    if (opiv == PANEL_THREADS * s) {
#pragma unroll
                        for (int q = 0; q < W; ++q) rb[0][q] = pass(x[s][q]);
                    }
            }
            if (oi >= 0 && oi % PANEL_THREADS == 0) {
#pragma unroll
                for (int s = 0; s < RPT; ++s)
                    if (oi == PANEL_THREADS * s) {
#pragma unroll
                        for (int q = 0; q < W; ++q) rb[1][q] = pass(x[s][q]);
                    }
            }
            lds_barrier();
            const T d = rb[0][c];  // the pivot value
            if (t == 0) ipiv[col0 + c] = (d == T(0)) ? i : piv;
            if (d == T(0)) {
                // lu.rs:107-110: all-zero column, nothing is swapped or updated
                if (t == 0 && info && *info == 0) *info = i + 1;
            } else {
                const T inv = T(1) / d;
                // rows i and piv trade places (their owners only)
                if (oi >= 0 && oi % PANEL_THREADS == 0) {
#pragma unroll
                    for (int s = 0; s < RPT; ++s)
                        if (oi == PANEL_THREADS * s) {
#pragma unroll
                            for (int q = 0; q < W; ++q) x[s][q] = pass(rb[0][q]);
                        }
                }
                if (piv != i && opiv >= 0 && opiv % PANEL_THREADS == 0) {
#pragma unroll
                    for (int s = 0; s < RPT; ++s)
                        if (opiv == PANEL_THREADS * s) {
#pragma unroll
                            for (int q = 0; q < W; ++q) x[s][q] = pass(rb[1][q]);
                        }
                }
#pragma unroll
                for (int s = 0; s < RPT; ++s) {
                    const int r = row0 + t + PANEL_THREADS * s;
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
    {
        // thread t stages row t & 63, columns (t >> 6) + 4 i: 16 loads in
        // flight, clamped in-bounds addresses (entries past k are never read)
        const int rr = threadIdx.x & 63, cq = threadIdx.x >> 6;
        const T *lrow = A + (r0 + (rr < k ? rr : 0));
        T v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = lrow[(size_t)(r0 + (cq + 4 * i < k ? cq + 4 * i : 0)) * lda];
#pragma unroll
        for (int i = 0; i < 16; ++i) L[rr][cq + 4 * i] = v[i];
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

// One wavefront; lane l keeps slots l and l + 64 of the plan in registers, and
// the slots of positions i and p are found by ballot, so no barriers are needed.
__global__ __launch_bounds__(64) void swap_plan_kernel(const int *ipiv, int k0, int k1, int *plan)
{
    const int l = threadIdx.x;
    int pos0 = -1, who0 = -1, pos1 = -1, who1 = -1;
    const int myp = (l < k1 - k0) ? ipiv[k0 + l] : 0;
    int cnt = 0;
    for (int i = k0; i < k1; ++i) {
        const int p = __shfl(myp, i - k0);
        if (p == i) continue;  // uniform: no swap at this step
        const unsigned long long ma0 = __ballot(pos0 == i), ma1 = __ballot(pos1 == i);
        const unsigned long long mb0 = __ballot(pos0 == p), mb1 = __ballot(pos1 == p);
        int a = ma0 ? __builtin_ctzll(ma0) : ma1 ? 64 + __builtin_ctzll(ma1) : -1;
        int b = mb0 ? __builtin_ctzll(mb0) : mb1 ? 64 + __builtin_ctzll(mb1) : -1;
        if (a < 0) {
            a = cnt++;
            if (l == (a & 63)) {
                if (a < 64) { pos0 = i; who0 = i; } else { pos1 = i; who1 = i; }
            }
        }
        if (b < 0) {
            b = cnt++;
            if (l == (b & 63)) {
                if (b < 64) { pos0 = p; who0 = p; } else { pos1 = p; who1 = p; }
            }
        }
        // swap who[a] and who[b]
        const int wa = __shfl(a < 64 ? who0 : who1, a & 63);
        const int wb = __shfl(b < 64 ? who0 : who1, b & 63);
        if (l == (a & 63)) { if (a < 64) who0 = wb; else who1 = wb; }
        if (l == (b & 63)) { if (b < 64) who0 = wa; else who1 = wa; }
    }
    if (l == 0) plan[0] = cnt;
    if (l < cnt) {
        plan[1 + l] = pos0;
        plan[1 + PLAN_MAX + l] = who0;
    }
    if (64 + l < cnt) {
        plan[1 + 64 + l] = pos1;
        plan[1 + PLAN_MAX + 64 + l] = who1;
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

// A[r0.., k0..k0+nb) <- A[r0.., k0..k0+nb) U11^{-1} (U11 the upper triangle of
// the factored diagonal block): one thread per row, the row in registers,
// x_q = (a_q - sum_{p<q} x_p U_pq) / U_qq with four partial sums.  U is staged
// transposed in LDS so that a column is contiguous; each column is read in two
// halves of 32 whose reads are all issued before the first use (one LDS round
// trip per half instead of one per pair of values).
constexpr int TRSM_RU_THREADS = 64;

template <typename T>
__global__ __launch_bounds__(TRSM_RU_THREADS) void trsm_right_upper_kernel(T *A, int lda, int k0, int nb, int r0,
                                                                           int m)
{
    __shared__ __align__(16) T Ut[64][64];  // Ut[q][p] = U[p][q]
    __builtin_amdgcn_s_setprio(3);  // on the look-ahead critical path, beside the trailing update
    const int t = threadIdx.x;
    // the row to solve is loaded first, so that its loads overlap the staging
    const int r = r0 + blockIdx.x * TRSM_RU_THREADS + t;
    const bool live = r < m;
    T *row = A + (live ? r : r0);
    T x[64];
#pragma unroll
    for (int q = 0; q < 64; ++q) x[q] = row[(size_t)(k0 + (q < nb ? q : 0)) * lda];
    // thread t stages row t of U; 16 loads in flight (clamped in-bounds
    // addresses, zeros selected afterwards) instead of one load per wait
    const T *urow = A + (k0 + (t < nb ? t : 0));
#pragma unroll
    for (int q0 = 0; q0 < 64; q0 += 16) {
        T v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = urow[(size_t)(k0 + (q0 + j < nb ? q0 + j : 0)) * lda];
#pragma unroll
        for (int j = 0; j < 16; ++j) Ut[q0 + j][t] = (t <= q0 + j && q0 + j < nb) ? v[j] : T(0);
    }
    __syncthreads();
    // x[q] past nb holds column 0 again and is never read: the solve below
    // touches x[p] only for p < q < nb, and rows past m are not stored
#pragma unroll
    for (int q = 0; q < 64; ++q) {
        if (q < nb) {  // uniform
            T s[4] = {x[q], T(0), T(0), T(0)};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (32 * h < q) {
                    T u[32];
#pragma unroll
                    for (int j = 0; j < 32; ++j)
                        if (32 * h + j < q) u[j] = Ut[q][32 * h + j];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int j = 0; j < 32; ++j)
                        if (32 * h + j < q) s[j & 3] = fma(-x[32 * h + j], u[j], s[j & 3]);
                }
            }
            x[q] = ((s[0] + s[1]) + (s[2] + s[3])) / Ut[q][q];
        }
    }
    int nbs = nb;
    asm volatile("" : "+s"(nbs));  // recompute the column masks instead of keeping 64 alive
    if (live) {
#pragma unroll
        for (int q = 0; q < 64; ++q)
            if (q < nbs) row[(size_t)(k0 + q) * lda] = x[q];
    }
}

// ---------------------------------------------------------------- LU with windowed pivoting
// (mixed-precision solver only; pivots differ from nalgebra's by design)
//
// Partial pivoting restricted to a window: for the nb (<= 64) columns at k0,
// every pivot is the largest |a| among the rows [k0, kw) not yet chosen,
// kw - k0 <= WINDOW_ROWS.  Rows below the window are eliminated afterwards by
// A21 U11^{-1}.  With a 128-row window at least 65 candidates remain for the
// last column of a block, which keeps |L| and the growth close to full
// partial pivoting on the dense test matrices while the pivot search never
// leaves one workgroup.
constexpr int WINDOW_ROWS = 128;

// Pivot key of window position pos (< 128) for a magnitude av: its top bits
// with the low 7 replaced by 127 - pos.  The unsigned maximum of the keys is
// the largest magnitude to within 2^-16 relative (2^-13 for FP64) and, among
// magnitudes that close, the lowest position: a pivot within that factor of
// the largest is as good for this solver, and the search becomes one integer
// maximum per DPP step.
template <typename T>
__device__ __forceinline__ unsigned pivot_key(T av, int pos)
{
    const unsigned hi =
        sizeof(T) == 8 ? (unsigned)((unsigned long long)__double_as_longlong((double)av) >> 32) : __float_as_uint((float)av);
    return (hi & ~127u) | (unsigned)(127 - pos);
}

// Wavefront maximum of an unsigned value; lane 63 ends with it (same DPP
// pattern as argmax_wave).
__device__ __forceinline__ unsigned wave_max_u32(unsigned k)
{
    k = max(k, (unsigned)dpp32<0x111, 0xf>(0, (int)k));
    k = max(k, (unsigned)dpp32<0x112, 0xf>(0, (int)k));
    k = max(k, (unsigned)dpp32<0x114, 0xf>(0, (int)k));
    k = max(k, (unsigned)dpp32<0x118, 0xf>(0, (int)k));
    k = max(k, (unsigned)dpp32<0x142, 0xa>(0, (int)k));
    k = max(k, (unsigned)dpp32<0x143, 0xc>(0, (int)k));
    return k;
}

// One thread per window row, the row's nb values in registers.  Rows are not
// moved during the factorisation: each thread carries the logical position of
// its row, only the pivot row passes through LDS, and every row is written to
// its final position at the end.  Records ipiv (absolute, LAPACK order: step
// k0+c exchanged rows k0+c and ipiv[k0+c]) and the composed row plan of the
// window for apply_plan_kernel; with perm non-null it also applies the
// window's permutation to perm[k0, kw) (row i of the factors is row perm[i]
// of the input).  A zero pivot column sets info (first one) and is skipped.
template <typename T>
__global__ __launch_bounds__(WINDOW_ROWS) void window_panel_kernel(T *A, int lda, int k0, int nb, int kw, int *ipiv,
                                                                   int *info, int *plan, int *perm)
{
    __shared__ T prow[2][64];
    __shared__ unsigned wk[2][2];
    __shared__ unsigned long long wmask[2];
    // this kernel is the critical path of the look-ahead pipeline and shares
    // its CU with trailing-update wavefronts: issue first among them
    __builtin_amdgcn_s_setprio(3);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wr = kw - k0;
    const bool live = t < wr;
    T a[64];
    const T *row = A + (k0 + (live ? t : 0));
#pragma unroll
    for (int q = 0; q < 64; ++q) a[q] = (live && q < nb) ? row[(size_t)(k0 + q) * lda] : T(0);
    const int myperm = (perm && live) ? perm[k0 + t] : 0;
    int pos = t;
    int nbl = nb;
    asm volatile("" : "+s"(nbl));  // separate from the load masks, which then need not stay alive
#pragma unroll
    for (int c = 0; c < 64; ++c) {
        if (c < nbl) {  // uniform
            const int par = c & 1;
            unsigned key = (live && pos >= c) ? pivot_key(fabs(a[c]), pos) : 0u;
            key = wave_max_u32(key);
            if (lane == 63) wk[par][wave] = key;
            lds_barrier();
            const unsigned kmax = max(wk[par][0], wk[par][1]);
            int P = 127 - (int)(kmax & 127u);
            const unsigned mag = kmax & ~127u;
            // an all-zero (or non-finite) column: nothing to exchange or eliminate
            const bool zero = mag == 0u || mag >= (sizeof(T) == 8 ? 0x7ff00000u : 0x7f800000u);
            if (zero) {
                P = c;  // nothing to exchange or eliminate
                if (t == 0 && *info == 0) *info = k0 + c + 1;
            }
            if (t == 0) ipiv[k0 + c] = k0 + P;
            if (pos == P) {
#pragma unroll
                for (int q = 0; q < 64; ++q)
                    if (q >= c) prow[par][q] = a[q];
            }
            const int old = pos;
            pos = old == P ? c : old == c ? P : old;
            lds_barrier();
            if (!zero && live && pos > c) {
                const T l = a[c] * (T(1) / prow[par][c]);
                a[c] = l;
#pragma unroll
                for (int q = 0; q < 64; ++q)
                    if (q > c) a[q] = fma(-l, prow[par][q], a[q]);
            }
        }
    }
    if (perm && live) perm[k0 + pos] = myperm;
    if (live) {
        int nbs = nb;
        asm volatile("" : "+s"(nbs));
        T *dst = A + (k0 + pos);
#pragma unroll
        for (int q = 0; q < 64; ++q)
            if (q < nbs) dst[(size_t)(k0 + q) * lda] = a[q];
    }
    // row plan: the window row now at position pos came from position t
    const bool moved = live && pos != t;
    const unsigned long long m = __ballot(moved);
    if (lane == 0) wmask[wave] = m;
    lds_barrier();
    const int before = __popcll(m & ((1ull << lane) - 1)) + (wave ? __popcll(wmask[0]) : 0);
    if (moved) {
        plan[1 + before] = k0 + pos;
        plan[1 + PLAN_MAX + before] = k0 + t;
    }
    if (t == 0) plan[0] = __popcll(wmask[0]) + __popcll(wmask[1]);
}

__global__ void iota_kernel(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void fill_int_kernel(int *p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace nrocm_lu
