// Batched Cholesky factor + solve for padded sizes 24 and 32 with two
// matrices per wavefront: lanes 32 h .. 32 h + 31 serve member h, and inside a
// half lane (p, q) = (l % 8, l / 8), l = lane % 32, holds the elements
// (p + 8 ri, q + 4 ci) of the lower triangle (tiles with ci <= 2 ri + 1; the
// upper elements of the tiles crossing the diagonal are held but never used).
//
// Each element goes through exactly the operations of the one-matrix element-
// tile kernel (chol_tile.hpp), in the same order: step k of the factor
// (src/linalg/cholesky.rs:226-262) takes L_kk = sqrt(a_kk) and
// L_rk = a_rk / L_kk (one quotient per lane, row l of the member), then
// a_rc = fl(a_rc - fl(L_ck * L_rk)) on the trailing tiles; the right-hand
// sides ride along as extra rows (forward substitution, solve.rs:488-519); the
// backward substitution L^T x = y (solve.rs:697-755) keeps dotx's eight
// accumulators (src/base/blas.rs:93-167), one lane per accumulator and
// right-hand side of each member.  Results are therefore bit-identical to the
// one-matrix kernel's in NROCM_EXACT and NROCM_FAST alike.
//
// What the pairing buys: every step's wavefront-wide work (the square root,
// the reciprocal, the quotient and its range test, the column store, the LDS
// broadcasts, the ordered backward reduction) now serves two matrices, and at
// padded sizes 24 and 32 that work, not the tile updates, is most of a step.
#pragma once
#include "chol_tile.hpp"

namespace nrocm_chol_tile2 {

using nrocm_chol::div_den_ok;
using nrocm_chol::div_with;
using nrocm_chol::lane_value;
using nrocm_chol::recip_refined;
using nrocm_chol_tile::dpp_shl;
using nrocm_chol_tile::gcdouble;
using nrocm_chol_tile::gdouble;
using nrocm_chol_tile::kZeroCellT;
using nrocm_chol_tile::MAXRHS;
using nrocm_chol_tile::num_in_range;
using nrocm_chol_tile::quotient;
using nrocm_chol_tile::upd;

template <int NB>
struct Tile2Smem {
    static constexpr int RR = NB / 8;
    double colraw[2][32 + MAXRHS];     // column k before division (rows), then the right-hand sides
    double colbuf[2][8][RR + 1];       // L_rk at [r % 8][r / 8]
    double ycol[2][MAXRHS];            // coefficient of step k per right-hand side
    double ysave[2][MAXRHS][32];       // y = L^{-1} b (forward result)
    double dg[2][32];                  // L_kk
    double inv[2][32];                 // its refined reciprocal (EXACT) or 1 / L_kk
    double prod[2][MAXRHS][32 + 8];    // backward products (a block of 8 read past position n - 1 stays inside)
    double xo[2][MAXRHS];              // x_i of the current backward step
};

template <int NB, bool EXACT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NB == 24 ? 6 : 5)))
void potrf_potrs_tile2_kernel(const int *__restrict__ order, int begin, int end, const int *__restrict__ nsz,
                              double *const *__restrict__ Aptr, double *const *__restrict__ Bptr, int nrhs,
                              int *__restrict__ info)
{
    static_assert(NB == 24 || NB == 32, "two members per wavefront: padded sizes 24 and 32");
    constexpr int RR = NB / 8;  // row tiles
    constexpr int CC = NB / 4;  // column tiles
    __shared__ Tile2Smem<NB> s;
    const int item0 = begin + 2 * (int)blockIdx.x;
    if (item0 >= end) return;  // the whole workgroup leaves
    const bool valid1 = item0 + 1 < end;
    // member data, wavefront-uniform per member
    const int jm0 = order[item0];
    const int jm1 = valid1 ? order[item0 + 1] : jm0;
    const int n0 = nsz[jm0];
    const int n1 = valid1 ? nsz[jm1] : 0;
    const int nr = (Bptr != nullptr) ? nrhs : 0;

    const int lane = threadIdx.x;
    const int h = lane >> 5, l = lane & 31;
    const int p = l & 7, q = l >> 3;
    const int nh = h ? n1 : n0;
    gdouble *const A = (gdouble *)(h ? (valid1 ? Aptr[jm1] : nullptr) : Aptr[jm0]);
    gdouble *const B = (nr > 0 && (h == 0 || valid1)) ? (gdouble *)Bptr[h ? jm1 : jm0] : nullptr;
    gcdouble *const zc = (gcdouble *)kZeroCellT;

    // ---- load: tiles of the lower triangle (identity padding), right-hand-side rows
    double T[RR][CC];
#pragma unroll
    for (int ri = 0; ri < RR; ++ri) {
#pragma unroll
        for (int ci = 0; ci < CC; ++ci) {
            if (ci > 2 * ri + 1) continue;
            const int r = p + 8 * ri, c = q + 4 * ci;
            const bool in = r < nh && c < nh && r >= c;  // nh == 0 for a missing second member
            gcdouble *src = in ? A + r + (size_t)c * nh : zc + (r == c ? 1 : 0);
            T[ri][ci] = *src;
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    double Y[CC];
#pragma unroll
    for (int ci = 0; ci < CC; ++ci) {
        const int c = q + 4 * ci;
        const bool in = p < nr && c < nh;
        gcdouble *src = in ? B + c + (size_t)p * nh : zc;
        Y[ci] = *src;
    }

    // ---- factor + forward substitution
    unsigned bad0 = 0, bad1 = 0;  // bit k: pivot k (< n) of member 0 / 1 is not > 0
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int kq = k & 3, kc = k >> 2;  // column class and column tile of column k
        const int kp = k & 7, kr = k >> 3;  // row class and row tile of row k
        int ln = threadIdx.x;  // opaque each step: lane comparisons stay inside it
        asm volatile("" : "+v"(ln));
        const int h = ln >> 5, l = ln & 31;
        const int p = l & 7, q = l >> 3;
        // 1. pivots of both members, L_kk per lane
        const double d0 = lane_value(T[kr][kc], kp + 8 * kq);
        const double d1 = lane_value(T[kr][kc], 32 + kp + 8 * kq);
        bad0 |= (unsigned)(k < n0 && !(d0 > 0.0)) << k;
        bad1 |= (unsigned)(k < n1 && !(d1 > 0.0)) << k;
        const double d = h ? d1 : d0;
        const double den = sqrt(d);
        const double y = EXACT ? recip_refined(den) : 1.0 / den;
        const bool den_ok = div_den_ok(den);
        if (l == 0) {
            s.dg[h][k] = den;
            s.inv[h][k] = y;
        }
        // 2. column k (raw) to LDS
        if (q == kq) {
#pragma unroll
            for (int ri = kc / 2; ri < RR; ++ri) s.colraw[h][p + 8 * ri] = T[ri][kc];
            if (p < nr) s.colraw[h][32 + p] = Y[kc];
        }
        lds_barrier();
        double lk;
        // rows that are not divided (above the diagonal, padding) use a benign
        // numerator so that they never force the slow division
        const bool row = l > k && l < nh;
        if (k >= MAXRHS - 1) {
            // lanes 0..k of a member have no row to divide: lane c < nr divides
            // right-hand side c's coefficient instead
            const bool rhs = l < nr;
            const double v = s.colraw[h][row ? l : (rhs ? 32 + l : 0)];
            const double qv = quotient<EXACT>((row || rhs) ? v : 1.0, den, y, den_ok);
            lk = row ? qv : (l == k ? den : 0.0);
            if (rhs) {
                s.ycol[h][l] = qv;
                s.ysave[h][l][k] = qv;
            }
        } else {
            const double v = row ? s.colraw[h][l] : 1.0;
            const double qv = quotient<EXACT>(v, den, y, den_ok);
            lk = row ? qv : (l == k ? den : 0.0);
            if (l < nr) {
                const double yk = quotient<EXACT>(s.colraw[h][32 + l], den, y, den_ok);
                s.ycol[h][l] = yk;
                s.ysave[h][l][k] = yk;
            }
        }
        if (l >= k && l < nh) A[(size_t)k * nh + l] = lk;
        if (q < RR) s.colbuf[h][p][q] = lk;  // row l = p + 8 q
        lds_barrier();
        // 3. multipliers and the trailing update
        double lr[RR];
#pragma unroll
        for (int ri = kc / 2; ri < RR; ++ri) lr[ri] = s.colbuf[h][p][ri];
        const double yv = p < nr ? s.ycol[h][p] : 0.0;
#pragma unroll
        for (int ci = kc; ci < CC; ++ci) {
            if (ci > kc || q > kq) {
                // L of column q + 4 ci: row class q + 4 (ci & 1), row tile ci / 2
                const double lc = s.colbuf[h][q + 4 * (ci & 1)][ci >> 1];
#pragma unroll
                for (int ri = ci / 2; ri < RR; ++ri) upd<EXACT>(T[ri][ci], lc, lr[ri]);
                upd<EXACT>(Y[ci], yv, lc);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // nalgebra stops at the first pivot that is not > 0 (cholesky.rs:240-270)
    const int fail0 = bad0 ? (int)__builtin_ctz(bad0) + 1 : 0;
    const int fail1 = bad1 ? (int)__builtin_ctz(bad1) + 1 : 0;
    if (lane == 0) info[jm0] = fail0;
    if (lane == 32 && valid1) info[jm1] = fail1;
    const bool solve0 = nr > 0 && fail0 == 0;
    const bool solve1 = nr > 0 && valid1 && fail1 == 0;
    if (!solve0 && !solve1) return;  // failed members keep their right-hand sides

    // ---- backward substitution L^T x = y, one row per lane of each member
    const bool solve = h ? solve1 : solve0;
    const int r = l;
    double b[MAXRHS];
#pragma unroll
    for (int c = 0; c < MAXRHS; ++c) b[c] = (c < nr) ? s.ysave[h][c][r] : 0.0;
    // L_ri for the steps i = nmax-1, nmax-2, ...: PF of them in flight.  A
    // lane of a solving member loads row min(r, n-1), column min(max(i, 0),
    // n-1) of its member; a lane of a member that does not solve loads the
    // zero cell.  Values of rows r <= i or r >= n only land in product slots
    // that no dot reads.
    constexpr int PF = 8;
    const int nmax = (solve0 ? n0 : 0) > (solve1 ? n1 : 0) ? (solve0 ? n0 : 0) : (solve1 ? n1 : 0);
    const int rl = r < nh ? r : nh - 1;
    gcdouble *const Lb = solve ? (gcdouble *)A + rl : zc;
    const int ldl = solve ? nh : 0;
    auto lval = [&](int i) {
        const int ic = i < 0 ? 0 : (i < nh ? i : nh - 1);
        return Lb[(size_t)(solve ? ic : 0) * ldl];
    };
    double cur[PF], nxt[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) cur[u] = lval(nmax - 1 - u);
    const int t = l & 7, cq = l >> 3;  // lane t + 8 cq of a member builds its accumulator t of right-hand side cq
    for (int i0 = nmax - 1; i0 >= 0; i0 -= PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) nxt[u] = lval(i0 - PF - u);
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int i = i0 - u;
            if (i < 0) break;
            const bool live = solve && i < nh;  // this member has step i
#pragma unroll
            for (int c = 0; c < MAXRHS; ++c)
                if (c < nr) s.prod[h][c][r] = cur[u] * b[c];
            lds_barrier();
            const int len = live ? nh - i - 1 : 0;
            const int full = len & ~7;
            const double *pc = &s.prod[h][cq][i + 1 + t];  // i + 1 + t <= 39
            double acc = 0.0;
            if (live && cq < nr) {
                // acc_t: positions t, t + 8, ... of the full blocks of 8 (EXACT)
                // or of all positions (FAST), in order
                const int nm = EXACT ? (full >> 3) : ((len + 7) >> 3);
                for (int m = 0; m < nm; ++m) {
                    const double v = pc[8 * m];
                    acc += (EXACT || t + 8 * m < len) ? v : 0.0;
                }
            }
            // lanes t < 4: acc_t + acc_{t+4}; lane t == 0 combines in order
            const double pair = acc + dpp_shl<4>(acc);
            const double p1 = dpp_shl<1>(pair), p2 = dpp_shl<2>(pair), p3 = dpp_shl<3>(pair);
            if (live && cq < nr && t == 0) {
                double res = 0.0;
                res += pair;
                res += p1;
                res += p2;
                res += p3;
                if constexpr (EXACT)
                    for (int e = full; e < len; ++e) res += pc[e];  // the tail, in order
                const double num = s.ysave[h][cq][i] - res;
                const double dii = s.dg[h][i];
                double xq;
                if constexpr (EXACT) {
                    const bool ok = div_den_ok(dii) && num_in_range(num);
                    xq = ok ? div_with(num, dii, s.inv[h][i]) : num / dii;
                } else {
                    xq = num * s.inv[h][i];
                }
                s.xo[h][cq] = xq;
            }
            lds_barrier();
            if (live && r == i) {
#pragma unroll
                for (int c = 0; c < MAXRHS; ++c)
                    if (c < nr) b[c] = s.xo[h][c];
            }
        }
#pragma unroll
        for (int u = 0; u < PF; ++u) cur[u] = nxt[u];
    }
    if (solve && r < nh) {
#pragma unroll
        for (int c = 0; c < MAXRHS; ++c)
            if (c < nr) B[r + (size_t)c * nh] = b[c];
    }
}

template <int NB>
int tile2_launch(bool exact, hipStream_t st, const int *order, int begin, int end, const int *nsz,
                 double *const *A, double *const *B, int nrhs, int *info)
{
    const unsigned grid = (unsigned)((end - begin + 1) / 2);
    if (grid == 0) return 0;
    if (exact)
        hipLaunchKernelGGL((potrf_potrs_tile2_kernel<NB, true>), dim3(grid), dim3(64), 0, st, order, begin, end,
                           nsz, A, B, nrhs, info);
    else
        hipLaunchKernelGGL((potrf_potrs_tile2_kernel<NB, false>), dim3(grid), dim3(64), 0, st, order, begin, end,
                           nsz, A, B, nrhs, info);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace nrocm_chol_tile2

#define NROCM_CHOL_TILE2_DECLARE(NB)                                                                   \
    int nrocm_chol_tile2_launch_##NB(bool exact, hipStream_t st, const int *order, int begin, int end, \
                                     const int *nsz, double *const *A, double *const *B, int nrhs, int *info);
#define NROCM_CHOL_TILE2_DEFINE(NB)                                                                    \
    int nrocm_chol_tile2_launch_##NB(bool exact, hipStream_t st, const int *order, int begin, int end, \
                                     const int *nsz, double *const *A, double *const *B, int nrhs, int *info) \
    {                                                                                                  \
        return nrocm_chol_tile2::tile2_launch<NB>(exact, st, order, begin, end, nsz, A, B, nrhs, info); \
    }

NROCM_CHOL_TILE2_DECLARE(24)
NROCM_CHOL_TILE2_DECLARE(32)
