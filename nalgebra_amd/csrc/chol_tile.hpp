// Batched Cholesky factor + solve for padded sizes 24..64 with one wavefront
// per matrix and the lower triangle spread over an 8 x 8 grid of lanes.
//
// Lane (p, q) = (lane % 8, lane / 8) holds the elements (p + 8 ri, q + 8 ci),
// ri >= ci, of the lower triangle (upper elements of diagonal tiles are held
// but never used), so a right-looking step k updates only the trailing
// triangle: tiles with ci > k / 8, and those with ci == k / 8 on lanes q > k % 8.
// Step k of the factor (src/linalg/cholesky.rs:226-262, per element the same
// ordered unfused sequence a_rj = fl(a_rj - fl(L_jk * L_rk)), then
// L_jj = sqrt(a_jj) and L_rj = a_rj / L_jj):
//   1. the diagonal a_kk is read from lane (k%8, k%8), L_kk = sqrt(a_kk);
//   2. the lanes of column class k%8 write column k (and the right-hand-side
//      rows) to LDS; lane l divides element l (one division per lane) and
//      stores L_lk straight to global memory (one coalesced column store);
//   3. every lane reads the multipliers of its rows and columns from LDS and
//      updates its tiles.
// The forward substitution L y = b (solve.rs:488-519: coeff = b_i / L_ii,
// b_r = fl(b_r - fl(coeff * L_ri))) is the same elimination applied to the
// right-hand sides as extra rows below the matrix: lane (p, q) holds
// y_{q + 8 ci} of right-hand side p, so right-hand side p's coefficient of
// step k is divided in lane p and its rows are updated with the column
// multipliers.
// Lane l stores L_lk of every step itself, so in the backward substitution
// L^T x = y (solve.rs:697-755), one row per lane, each lane re-reads only
// values it wrote (program order makes them visible), a few steps ahead.  The
// backward keeps dotx's eight accumulators (src/base/blas.rs:93-167):
// products through LDS, one lane per accumulator and right-hand side, DPP for
// the pairs and the ordered combine.
#pragma once
#include "chol_kernels.hpp"

namespace nrocm_chol_tile {

using nrocm_chol::div_den_ok;
using nrocm_chol::div_with;
using nrocm_chol::lane_value;
using nrocm_chol::recip_refined;

constexpr int MAXRHS = 4;  // right-hand sides carried through the factor (one per lane row)

template <int R>
struct TileSmem {
    double colraw[64 + MAXRHS];  // column k before division (rows), then the right-hand sides
    double colbuf[8][R + 1];     // L_rk at [r % 8][r / 8] (rows padded: no bank conflicts at R = 8)
    double ycol[MAXRHS];         // coefficient of step k per right-hand side
    double ysave[MAXRHS][64];    // y = L^{-1} b (forward result)
    double dg[64];               // L_kk
    double inv[64];              // its refined reciprocal (EXACT) or 1 / L_kk
    double prod[MAXRHS][64 + 4]; // backward products
    double xo[MAXRHS];           // x_i of the current backward step
};

__device__ const double kZeroCellT[2] = {0.0, 1.0};  // a zero, then a one for padded diagonals

// Global-address-space views of the member pointers (which come from a pointer
// array, so the compiler would otherwise emit flat instructions, and flat
// instructions also count against the LDS counter that every step's LDS
// barrier drains: each step would wait for its own column store).
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) const char gcchar;

// Element at a 32-bit byte offset from a wavefront-uniform base: the scalar-
// base addressing form (one vector add per access instead of a 64-bit address
// per lane).  A member's matrix (n <= 64) spans far less than 4 GiB.  (Used
// for the backward's loads; the same form for the factor's column stores
// costs the padded-64 kernel 16 more spilled registers.)
__device__ __forceinline__ double gload_at(gcdouble *base, unsigned byte_off)
{
    return *(gcdouble *)((gcchar *)base + byte_off);
}

// 64-bit DPP row shift left by N within rows of 16 lanes; lanes whose source
// leaves the row get an unspecified value (their results are not used).
template <int N>
__device__ __forceinline__ double dpp_shl(double v)
{
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)x, 0x100 + N, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(x >> 32), 0x100 + N, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The update a = fl(fl((-x) * y) + a) in place (nalgebra's axpy with factor
// -x), or one fused multiply-add in FAST.  Updating the accumulator register
// itself keeps exec-masked updates free of result copies.
template <bool EXACT>
__device__ __forceinline__ void upd(double &acc, double x, double y)
{
    if constexpr (EXACT) {
        double tmp;
        asm("v_mul_f64 %1, -%2, %3\n\tv_add_f64 %0, %1, %0" : "+v"(acc), "=&v"(tmp) : "v"(x), "v"(y));
    } else {
        asm("v_fma_f64 %0, -%1, %2, %0" : "+v"(acc) : "v"(x), "v"(y));
    }
}

// Numerator range of div_with from the exponent field alone: 2^-900 <= |x| <
// 2^600, or +0 (for which the three-instruction form also gives +0).
__device__ __forceinline__ bool num_in_range(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned e = (unsigned)(b >> 52) & 0x7ffu;
    return (e - 123u) <= 1499u || b == 0ull;
}

// Exact quotient v / den for a wavefront-uniform den with reciprocal y: the
// three-instruction form when every lane's operands are in its range, the
// compiler's division otherwise (see div_with).
template <bool EXACT>
__device__ __forceinline__ double quotient(double v, double den, double y, bool den_ok)
{
    if constexpr (EXACT) {
        const bool ok = den_ok && num_in_range(v);
        if (__ballot(!ok) == 0) return div_with(v, den, y);
        return v / den;
    } else {
        return v * y;
    }
}

// Wavefronts per SIMD of each padded size (the tiles shrink with NB); at
// NB = 64 the fourth wavefront costs 3 spilled vector registers.
template <int NB>
constexpr int tile_waves() { return NB >= 56 ? 4 : NB == 48 ? 5 : NB == 40 ? 6 : 8; }

template <int NB, bool EXACT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(tile_waves<NB>())))
void potrf_potrs_tile_kernel(const int *__restrict__ order, int begin, int end, const int *__restrict__ nsz,
                             double *const *__restrict__ Aptr, double *const *__restrict__ Bptr, int nrhs,
                             int *__restrict__ info)
{
    constexpr int R = NB / 8;
    __shared__ TileSmem<R> s;
    const int lane = threadIdx.x;
    const int p = lane & 7, q = lane >> 3;
    const int item = begin + (int)blockIdx.x;
    if (item >= end) return;  // one matrix per workgroup: the whole workgroup leaves
    const int jm = order[item];
    const int n = nsz[jm];
    gdouble *const A = (gdouble *)Aptr[jm];
    const int lda = n;
    const int nr = (Bptr != nullptr) ? nrhs : 0;
    gdouble *const B = nr > 0 ? (gdouble *)Bptr[jm] : nullptr;
    gcdouble *const zc = (gcdouble *)kZeroCellT;
    const int ldb = n;

    // ---- load: tiles of the lower triangle (identity padding), right-hand-side rows
    double T[R][R];
#pragma unroll
    for (int ri = 0; ri < R; ++ri) {
#pragma unroll
        for (int ci = 0; ci <= ri; ++ci) {
            const int r = p + 8 * ri, c = q + 8 * ci;
            const bool in = r < n && c < n && r >= c;
            // padding rows and columns hold the identity
            gcdouble *src = in ? A + r + (size_t)c * lda : zc + (r == c ? 1 : 0);
            const double v = *src;
            T[ri][ci] = v;
        }
        // a tile row of loads at a time: all 36 addresses at once would need
        // more registers than the factor itself
        __builtin_amdgcn_sched_barrier(0);
    }
    double Y[R];
#pragma unroll
    for (int ci = 0; ci < R; ++ci) {
        const int c = q + 8 * ci;
        const bool in = p < nr && c < n;
        gcdouble *src = in ? B + c + (size_t)p * ldb : zc;
        Y[ci] = *src;
    }

    // ---- factor + forward substitution
    unsigned long long bad = 0;  // bit k: pivot k (< n) is not > 0
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int kq = k & 7, kc = k >> 3;
        // the lane index through an empty asm statement: comparisons with it
        // stay inside their step instead of being hoisted out of the unrolled
        // loop as 64-bit lane masks (which spill)
        int lane = threadIdx.x;
        asm volatile("" : "+v"(lane));
        const int p = lane & 7, q = lane >> 3;
        // 1. pivot
        const double d = lane_value(T[kc][kc], kq * 9);
        bad |= (unsigned long long)(k < n && !(d > 0.0)) << k;
        const double den = sqrt(d);
        const double y = EXACT ? recip_refined(den) : 1.0 / den;
        const bool den_ok = div_den_ok(den);
        if (lane == 0) {
            s.dg[k] = den;
            s.inv[k] = y;
        }
        // 2. column k (raw) to LDS
        if (q == kq) {
#pragma unroll
            for (int ri = kc; ri < R; ++ri) s.colraw[p + 8 * ri] = T[ri][kc];
            if (p < nr) s.colraw[64 + p] = Y[kc];
        }
        lds_barrier();
        double lk;
        // rows that are not divided (above the diagonal, padding) use a benign
        // numerator so that they never force the slow division
        const bool row = lane > k && lane < n;
        if (k >= MAXRHS - 1) {
            // lanes 0..k have no row to divide: lane c < nr divides right-hand
            // side c's coefficient instead, one quotient per lane for both
            const bool rhs = lane < nr;
            const double v = s.colraw[row ? lane : (rhs ? 64 + lane : 0)];
            const double qv = quotient<EXACT>((row || rhs) ? v : 1.0, den, y, den_ok);
            lk = row ? qv : (lane == k ? den : 0.0);
            if (rhs) {
                s.ycol[lane] = qv;
                s.ysave[lane][k] = qv;
            }
        } else {
            const double v = row ? s.colraw[lane] : 1.0;
            const double qv = quotient<EXACT>(v, den, y, den_ok);
            lk = row ? qv : (lane == k ? den : 0.0);
            if (lane < nr) {
                const double yk = quotient<EXACT>(s.colraw[64 + lane], den, y, den_ok);
                s.ycol[lane] = yk;
                s.ysave[lane][k] = yk;
            }
        }
        if (k < n) {
            gdouble *const colk = A + (size_t)k * lda;  // uniform base, lane offset
            if (lane >= k && lane < n) colk[lane] = lk;
        }
        if (q < R) s.colbuf[p][q] = lk;  // row `lane` = p + 8 q sits at [p][q]
        lds_barrier();
        // 3. multipliers and the trailing update
        if (kc < R) {
            double lr[R], lc[R];
#pragma unroll
            for (int t = kc; t < R; ++t) {
                lr[t] = s.colbuf[p][t];
                lc[t] = s.colbuf[q][t];
            }
            const double yv = p < nr ? s.ycol[p] : 0.0;
#pragma unroll
            for (int ci = kc; ci < R; ++ci) {
                if (ci > kc || q > kq) {
#pragma unroll
                    for (int ri = ci; ri < R; ++ri) upd<EXACT>(T[ri][ci], lc[ci], lr[ri]);
                    upd<EXACT>(Y[ci], yv, lc[ci]);
                }
                // one tile column at a time: left free, the scheduler
                // interleaves whole steps and runs out of registers
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    // nalgebra stops at the first pivot that is not > 0 (cholesky.rs:240-270)
    const int fail = bad ? (int)__builtin_ctzll(bad) + 1 : 0;
    if (lane == 0) info[jm] = fail;
    if (nr == 0 || fail != 0) return;  // failed members keep their right-hand sides

    // ---- backward substitution L^T x = y, one row per lane
    const int r = lane;
    const bool row_ok = r < n;
    double b[MAXRHS];
#pragma unroll
    for (int c = 0; c < MAXRHS; ++c) b[c] = (c < nr) ? s.ysave[c][r] : 0.0;
    // L_ri for the steps i = n-1, n-2, ...: PF of them in flight.  Every lane
    // loads from a valid address (row min(r, n-1), column max(i, 0)); values of
    // rows r <= i (the upper triangle) or r >= n only land in product slots
    // that no dot reads.
    constexpr int PF = 8;
    const int rl = row_ok ? r : n - 1;
    auto lval = [&](int i) {
        return gload_at(A, (unsigned)((i > 0 ? i : 0) * lda + rl) * 8u);
    };
    double cur[PF], nxt[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) cur[u] = lval(n - 1 - u);
    const int t = lane & 7, cq = lane >> 3;  // lane t + 8 cq builds accumulator t of right-hand side cq
    for (int i0 = n - 1; i0 >= 0; i0 -= PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) nxt[u] = lval(i0 - PF - u);
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int i = i0 - u;
            if (i < 0) break;
#pragma unroll
            for (int c = 0; c < MAXRHS; ++c)
                if (c < nr) s.prod[c][r] = cur[u] * b[c];
            lds_barrier();
            const int len = n - i - 1;
            const int full = len & ~7;
            const double *pc = &s.prod[cq < MAXRHS ? cq : 0][i + 1 + t];
            double acc = 0.0;
            if (cq < nr) {
                // acc_t: positions t, t + 8, ... of the full blocks of 8 (EXACT)
                // or of all positions (FAST), in order; a uniform trip count
                const int nm = EXACT ? (full >> 3) : ((len + 7) >> 3);
                for (int m = 0; m < nm; ++m) {
                    const double v = pc[8 * m];
                    acc += (EXACT || t + 8 * m < len) ? v : 0.0;
                }
            }
            // lanes t < 4: acc_t + acc_{t+4}; lane t == 0 combines in order
            const double pair = acc + dpp_shl<4>(acc);
            const double p1 = dpp_shl<1>(pair), p2 = dpp_shl<2>(pair), p3 = dpp_shl<3>(pair);
            if (cq < nr && t == 0) {
                double res = 0.0;
                res += pair;
                res += p1;
                res += p2;
                res += p3;
                if constexpr (EXACT)
                    for (int e = full; e < len; ++e) res += pc[e];  // the tail, in order (lane t == 0)
                const double num = s.ysave[cq][i] - res;
                const double dii = s.dg[i];
                double xq;
                if constexpr (EXACT) {
                    const bool ok = div_den_ok(dii) && num_in_range(num);
                    xq = ok ? div_with(num, dii, s.inv[i]) : num / dii;
                } else {
                    xq = num * s.inv[i];
                }
                s.xo[cq] = xq;
            }
            lds_barrier();
            if (r == i) {
#pragma unroll
                for (int c = 0; c < MAXRHS; ++c)
                    if (c < nr) b[c] = s.xo[c];
            }
        }
#pragma unroll
        for (int u = 0; u < PF; ++u) cur[u] = nxt[u];
    }
    if (row_ok) {
#pragma unroll
        for (int c = 0; c < MAXRHS; ++c)
            if (c < nr) B[r + (size_t)c * ldb] = b[c];
    }
}

template <int NB>
int tile_launch(bool exact, hipStream_t st, const int *order, int begin, int end, const int *nsz, double *const *A,
                double *const *B, int nrhs, int *info)
{
    const unsigned grid = (unsigned)(end - begin);
    if (grid == 0) return 0;
    if (exact)
        hipLaunchKernelGGL((potrf_potrs_tile_kernel<NB, true>), dim3(grid), dim3(64), 0, st, order, begin, end, nsz,
                           A, B, nrhs, info);
    else
        hipLaunchKernelGGL((potrf_potrs_tile_kernel<NB, false>), dim3(grid), dim3(64), 0, st, order, begin, end, nsz,
                           A, B, nrhs, info);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace nrocm_chol_tile

#define NROCM_CHOL_TILE_DECLARE(NB)                                                                   \
    int nrocm_chol_tile_launch_##NB(bool exact, hipStream_t st, const int *order, int begin, int end, \
                                    const int *nsz, double *const *A, double *const *B, int nrhs, int *info);
#define NROCM_CHOL_TILE_DEFINE(NB)                                                                    \
    int nrocm_chol_tile_launch_##NB(bool exact, hipStream_t st, const int *order, int begin, int end, \
                                    const int *nsz, double *const *A, double *const *B, int nrhs, int *info) \
    {                                                                                                 \
        return nrocm_chol_tile::tile_launch<NB>(exact, st, order, begin, end, nsz, A, B, nrhs, info); \
    }

NROCM_CHOL_TILE_DECLARE(24)
NROCM_CHOL_TILE_DECLARE(32)
NROCM_CHOL_TILE_DECLARE(40)
NROCM_CHOL_TILE_DECLARE(48)
NROCM_CHOL_TILE_DECLARE(56)
NROCM_CHOL_TILE_DECLARE(64)
