// Device code of the batched small-matrix Cholesky path (see chol_batched.hip
// for the layout and exactness argument).  Included by chol_nbXX.hip, one
// translation unit per padded size so that the heavily unrolled
// instantiations compile in parallel.
#pragma once
#include "nrocm_internal.hpp"

namespace nrocm_chol {

constexpr int CH = 4;  // right-hand sides held in registers at a time

template <int NB>
constexpr int group_for() { return NB <= 8 ? 8 : NB <= 16 ? 16 : NB <= 32 ? 32 : 64; }

template <bool EXACT>
__device__ __forceinline__ double sub_prod(double acc, double x, double y)
{
    // nalgebra axpy with factor -x: fl(fl((-x)*y) + acc), the negation on x
    // before the product as there (a NaN coming from y keeps its sign).
    // One asm block per update keeps the product next to its addition: as
    // separate IR operations the scheduler hoists every product of a step and
    // spills the row array to scratch.
    double r;
    if constexpr (EXACT) {
        asm("v_mul_f64 %0, -%1, %2\n\tv_add_f64 %0, %0, %3" : "=&v"(r) : "v"(x), "v"(y), "v"(acc));
    } else {
        // as an asm statement for the same reason: a plain fma lets the
        // scheduler raise register use to 256 VGPRs
        asm("v_fma_f64 %0, -%1, %2, %3" : "=v"(r) : "v"(x), "v"(y), "v"(acc));
    }
    return r;
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

// A zero that masked-off lanes load instead of branching around their loads.
__device__ const double kZeroCell[2] = {0.0, 0.0};

// Loads the lower triangle of row r into a[] (zeros elsewhere).  Every load is
// issued unconditionally from a selected address, so no per-column lane masks
// stay live while the loads are in flight.
template <int NB>
__device__ __forceinline__ void load_lower(double (&a)[NB], const double *A, int lda, int n, int r)
{
    const double *row = (r < n) ? A + r : kZeroCell;
    const size_t step = (r < n) ? (size_t)lda : 0;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const double *p = (c <= r) ? row + (size_t)c * step : kZeroCell;
        a[c] = *p;
    }
}

template <int NB>
__device__ __forceinline__ void store_lower(const double (&a)[NB], double *A, int lda, int n, int r)
{
    // A copy of r the compiler cannot see through, so that the 64-bit lane
    // masks of load_lower are not kept alive until here.
    int rr = r;
    asm volatile("" : "+v"(rr));
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c <= rr && rr < n) A[rr + (size_t)c * lda] = a[c];
}

// Value of v in lane src (src uniform across the wavefront).
__device__ __forceinline__ double lane_value(double v, int src)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)bits, src);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Right-looking factorisation in registers; returns 0 or 1 + first failing column.
// All NB steps run: rows and columns beyond n hold zeros (padding rows can
// only ever update padding columns, so their NaNs never reach stored values).
// The failing column is found afterwards: pivot k failed iff its square root
// is not > 0 (d <= 0 gives NaN or +-0, NaN gives NaN), and nalgebra stops at
// the first such k.  mydiag receives L_rr (the square root taken at step r).
// sub > 0 is Cholesky::new_with_substitute's substitute (cholesky.rs:216-218,
// 237-264): a pivot that is not > 0 (NaN included) is replaced by it before
// the square root; sub <= 0 or NaN substitutes nothing, as there.
template <int NB, int G, bool EXACT>
__device__ __forceinline__ int factor(double (&a)[NB], double *col, double *xch, int n, int g, int r, int lane,
                                      double &mydiag, double sub = 0.0)
{
    mydiag = 1.0;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        double dkk;
        if constexpr (G == 64) {
            dkk = lane_value(a[k], k);
        } else {
            col[lane] = a[k];
            lds_barrier();
            dkk = col[g * G + k];
        }
        const double den = sqrt((sub > 0.0 && !(dkk > 0.0)) ? sub : dkk);
        if (r == k) {
            a[k] = den;
            mydiag = den;
        } else if (r > k) {
            a[k] = a[k] / den;
        }
        xch[lane] = a[k];
        lds_barrier();
        // The multipliers are read in chunks of UC, one chunk ahead of its
        // updates; the scheduling barriers stop the compiler from hoisting
        // every read of the step (which would need 2*(NB-k) more VGPRs).
        // Smaller chunks at NB = 32..64 leave room for one more wavefront
        // per SIMD (NB = 40, 48) or avoid spills (NB = 56, 64).
        constexpr int UC = (NB == 64 || NB == 56 || NB == 48 || NB == 40 || NB == 32) ? 4 : 8;
        constexpr int NC = (NB + UC - 1) / UC;
        const double *xg = xch + g * G;
        const int cfirst = (k + 1) / UC;
        double cur[UC];
#pragma unroll
        for (int t = 0; t < UC; ++t) cur[t] = (cfirst * UC + t < NB) ? xg[cfirst * UC + t] : 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if (c >= cfirst) {
                double nxt[UC];
#pragma unroll
                for (int t = 0; t < UC; ++t)
                    nxt[t] = (c + 1 < NC && (c + 1) * UC + t < NB) ? xg[(c + 1) * UC + t] : 0.0;
#pragma unroll
                for (int t = 0; t < UC; ++t) {
                    const int j = c * UC + t;
                    if (j > k && j < NB) a[j] = sub_prod<EXACT>(a[j], cur[t], a[k]);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int t = 0; t < UC; ++t) cur[t] = nxt[t];
            }
        }
    }
    const bool bad = r < n && !(mydiag > 0.0);
    unsigned long long m = __ballot(bad);
    m = (m >> (g * G)) & (G == 64 ? ~0ull : ((1ull << G) - 1));
    return m ? (int)__builtin_ctzll(m) + 1 : 0;
}

__device__ __forceinline__ double reg_q(const double (&b)[CH], int q)
{
    return q == 0 ? b[0] : q == 1 ? b[1] : q == 2 ? b[2] : b[3];
}

constexpr int PF = 8;  // substitution steps whose L elements are loaded ahead

// Forward substitution L y = b (solve.rs:488-519) for CH right-hand sides:
// coeff = b_i / L_ii; b_i = coeff; b_r = fl(fl(-coeff * L_ri) + b_r) for r > i.
// L is read from memory (the factor kernel's output); the elements of the
// next PF steps are loaded while the current PF steps run.
template <int G, bool EXACT>
__device__ __forceinline__ void forward(const double *L, int lda, double (&b)[CH], Smem<64 / G> &s, int n,
                                        int nmax, int g, int r, int lane)
{
    const bool row_ok = r < n;
    double cur[PF], nxt[PF];
#pragma unroll
    for (int t = 0; t < PF; ++t) cur[t] = (row_ok && t < nmax && r >= t) ? L[r + (size_t)t * lda] : 0.0;
    for (int i0 = 0; i0 < nmax; i0 += PF) {
#pragma unroll
        for (int t = 0; t < PF; ++t) {
            const int i = i0 + PF + t;
            nxt[t] = (row_ok && i < nmax && r >= i) ? L[r + (size_t)i * lda] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < PF; ++t) {
            const int i = i0 + t;
            if (i < nmax) {
                double *buf = s.xch[i & 1];
#pragma unroll
                for (int q = 0; q < CH; ++q) buf[lane * CH + q] = b[q];
                lds_barrier();
                const double di = s.diag[g * G + i];
#pragma unroll
                for (int q = 0; q < CH; ++q) {
                    const double coeff = buf[(g * G + i) * CH + q] / di;
                    if (r == i) b[q] = coeff;
                    else if (r > i) b[q] = sub_prod<EXACT>(b[q], coeff, cur[t]);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < PF; ++t) cur[t] = nxt[t];
    }
    lds_barrier();
}

// One backward step i with li = L[r, i] (0 above the diagonal / outside).
template <int G, bool EXACT>
__device__ __forceinline__ void backward_step(int i, double li, double (&b)[CH], Smem<64 / G> &s, int n, int g,
                                              int r, int lane)
{
    const double di = s.diag[g * G + i];
    if constexpr (EXACT) {
        double *p = s.xch[i & 1];
        double *acc = s.acc[i & 1];
#pragma unroll
        for (int q = 0; q < CH; ++q) p[lane * CH + q] = li * b[q];
        lds_barrier();
        const int len = n - i - 1;
        const int full = len & ~7;
        // Every load of a chain is independent of the adds, so all are issued
        // at once; out-of-range slots add +0.0, an exact identity here (the
        // running sums start at +0 and cannot become -0).
        if (r < 8 && len > 0) {
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                double t = 0.0;
#pragma unroll
                for (int m = 0; m < G / 8; ++m) {
                    const int e = r + 8 * m;
                    t += (e < full) ? p[(g * G + i + 1 + e) * CH + q] : 0.0;
                }
                acc[(g * 8 + r) * CH + q] = t;
            }
        }
        lds_barrier();
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
#pragma unroll
                    for (int m = 0; m < 7; ++m) {
                        const int e = full + m;
                        res += (e < len) ? p[(g * G + i + 1 + e) * CH + q] : 0.0;
                    }
                }
                b[q] = (b[q] - res) / di;
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            double v = (r > i && r < n) ? li * b[q] : 0.0;
            v = group_sum<G>(v);
            if (r == i && i < n) b[q] = (b[q] - v) / di;
        }
    }
}

// Backward substitution L^T x = y (solve.rs:697-755) for CH right-hand sides:
// b_i = (b_i - dotc(L[i+1..,i], b[i+1..])) / L_ii.  EXACT reproduces dotx's
// eight accumulators (acc_t sums positions t, t+8, ... of the full blocks of 8,
// then 0 + (acc0+acc4) + (acc1+acc5) + (acc2+acc6) + (acc3+acc7), then the tail).
template <int G, bool EXACT>
__device__ __forceinline__ void backward(const double *L, int lda, double (&b)[CH], Smem<64 / G> &s, int n,
                                         int nmax, int g, int r, int lane)
{
    const bool row_ok = r < n;
    double cur[PF], nxt[PF];
    // blocks of PF steps counted down from nmax - 1
#pragma unroll
    for (int t = 0; t < PF; ++t) {
        const int i = nmax - 1 - t;
        cur[t] = (row_ok && i >= 0 && r >= i) ? L[r + (size_t)i * lda] : 0.0;
    }
    for (int i0 = nmax - 1; i0 >= 0; i0 -= PF) {
#pragma unroll
        for (int t = 0; t < PF; ++t) {
            const int i = i0 - PF - t;
            nxt[t] = (row_ok && i >= 0 && r >= i) ? L[r + (size_t)i * lda] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < PF; ++t) {
            const int i = i0 - t;
            if (i >= 0) backward_step<G, EXACT>(i, cur[t], b, s, n, g, r, lane);
        }
#pragma unroll
        for (int t = 0; t < PF; ++t) cur[t] = nxt[t];
    }
    lds_barrier();
}

// ------------------------------------------------------------ fused factor + solve

// Exact division with a reciprocal reused across numerators.  hipcc expands an
// f64 x / d as s = div_scale(d), y = rcp(s) refined by two Newton steps,
// q0 = x' * y, r = fma(-s, q0, x'), q = div_fmas(r, y, q0), div_fixup(q, d, x).
// When 2^-100 <= |d| <= 2^100 and 2^-900 <= |x| <= 2^600 both scalings are
// identities, div_fmas is a plain fma and the fixup returns q unchanged, so
// div_with(x, d, recip_refined(d)) is bit-identical to x / d.  Outside that
// range (zeros, infinities, NaNs, extreme exponents) callers use x / d.
__device__ __forceinline__ double recip_refined(double d)
{
    const double r0 = __builtin_amdgcn_rcp(d);
    const double e0 = __builtin_fma(-d, r0, 1.0);
    const double r1 = __builtin_fma(r0, e0, r0);
    const double e1 = __builtin_fma(-d, r1, 1.0);
    return __builtin_fma(r1, e1, r1);
}

__device__ __forceinline__ double div_with(double x, double d, double y)
{
    const double q0 = x * y;
    const double rem = __builtin_fma(-d, q0, x);
    return __builtin_fma(rem, y, q0);
}

__device__ __forceinline__ bool div_den_ok(double d)
{
    const double ad = __builtin_fabs(d);
    return ad >= 0x1p-100 && ad <= 0x1p100;
}

__device__ __forceinline__ bool div_num_ok(double x)
{
    const double ax = __builtin_fabs(x);
    return ax >= 0x1p-900 && ax <= 0x1p600;
}

// a[i] for a wavefront-uniform i (a scalar branch tree, no scratch).
template <int NB>
__device__ __forceinline__ double reg_at(const double (&a)[NB], int i)
{
    double v = 0.0;
    switch (i) {
#define NROCM_REG_CASE(k) \
    case k:               \
        if constexpr (k < NB) v = a[k]; \
        break;
        NROCM_REG_CASE(0) NROCM_REG_CASE(1) NROCM_REG_CASE(2) NROCM_REG_CASE(3)
        NROCM_REG_CASE(4) NROCM_REG_CASE(5) NROCM_REG_CASE(6) NROCM_REG_CASE(7)
        NROCM_REG_CASE(8) NROCM_REG_CASE(9) NROCM_REG_CASE(10) NROCM_REG_CASE(11)
        NROCM_REG_CASE(12) NROCM_REG_CASE(13) NROCM_REG_CASE(14) NROCM_REG_CASE(15)
        NROCM_REG_CASE(16) NROCM_REG_CASE(17) NROCM_REG_CASE(18) NROCM_REG_CASE(19)
        NROCM_REG_CASE(20) NROCM_REG_CASE(21) NROCM_REG_CASE(22) NROCM_REG_CASE(23)
        NROCM_REG_CASE(24) NROCM_REG_CASE(25) NROCM_REG_CASE(26) NROCM_REG_CASE(27)
        NROCM_REG_CASE(28) NROCM_REG_CASE(29) NROCM_REG_CASE(30) NROCM_REG_CASE(31)
        NROCM_REG_CASE(32) NROCM_REG_CASE(33) NROCM_REG_CASE(34) NROCM_REG_CASE(35)
        NROCM_REG_CASE(36) NROCM_REG_CASE(37) NROCM_REG_CASE(38) NROCM_REG_CASE(39)
        NROCM_REG_CASE(40) NROCM_REG_CASE(41) NROCM_REG_CASE(42) NROCM_REG_CASE(43)
        NROCM_REG_CASE(44) NROCM_REG_CASE(45) NROCM_REG_CASE(46) NROCM_REG_CASE(47)
        NROCM_REG_CASE(48) NROCM_REG_CASE(49) NROCM_REG_CASE(50) NROCM_REG_CASE(51)
        NROCM_REG_CASE(52) NROCM_REG_CASE(53) NROCM_REG_CASE(54) NROCM_REG_CASE(55)
        NROCM_REG_CASE(56) NROCM_REG_CASE(57) NROCM_REG_CASE(58) NROCM_REG_CASE(59)
        NROCM_REG_CASE(60) NROCM_REG_CASE(61) NROCM_REG_CASE(62) NROCM_REG_CASE(63)
#undef NROCM_REG_CASE
    }
    return v;
}

constexpr int PSTR = 64 + 4;

template <int MPW>
struct FusedSmem {
    double col[64];
    double xch[64];
    double dg[64];            // L_rr per lane
    double inv[64];           // recip_refined(L_rr) (EXACT) or 1 / L_rr (FAST)
    double p[CH][PSTR];       // backward products, [q][lane] (padded rows: no bank conflicts)
    double acc[MPW][8 * CH];  // dotx accumulators, [q][t]
    double bi[MPW][CH];       // b_i of the current step
    double x[MPW][CH];        // result of the current step
};

// Forward substitution L y = b (solve.rs:488-519) with L in registers:
// coeff = b_i / L_ii; b_i = coeff; b_r = fl(fl(-coeff * L_ri) + b_r) for r > i.
// Lane i computes the CH quotients and they are broadcast to its group
// (readlane for one group per wavefront, LDS otherwise).
template <int NB, int G, bool EXACT>
__device__ __forceinline__ void forward_reg(const double (&a)[NB], double (&b)[CH], FusedSmem<64 / G> &s, int n,
                                            int nmax, bool care, int qn, int g, int r, int lane, double mydiag,
                                            double myinv, bool mydok)
{
    for (int i = 0; i < nmax; ++i) {
        const double li = reg_at<NB>(a, i);
        double t[CH];
        // only lane i of each group that still solves decides (ok elsewhere)
        const bool rel = care && i < n && r == i;
        bool ok = !rel || mydok;
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            if constexpr (EXACT) {
                t[q] = div_with(b[q], mydiag, myinv);
                ok = ok && (!rel || q >= qn || div_num_ok(b[q]));
            } else {
                t[q] = b[q] * myinv;
            }
        }
        if constexpr (EXACT) {
            if (__ballot(!ok) != 0) {
#pragma unroll
                for (int q = 0; q < CH; ++q) t[q] = b[q] / mydiag;
            }
        }
        double y[CH];
        if constexpr (G == 64) {
#pragma unroll
            for (int q = 0; q < CH; ++q) y[q] = lane_value(t[q], i);
        } else {
            if (r == i) {
#pragma unroll
                for (int q = 0; q < CH; ++q) s.x[g][q] = t[q];
            }
            lds_barrier();
#pragma unroll
            for (int q = 0; q < CH; ++q) y[q] = s.x[g][q];
            lds_barrier();
        }
        if (r == i) {
#pragma unroll
            for (int q = 0; q < CH; ++q) b[q] = y[q];
        } else if (r > i) {
#pragma unroll
            for (int q = 0; q < CH; ++q) b[q] = sub_prod<EXACT>(b[q], y[q], li);
        }
    }
}

// Backward substitution L^T x = y (solve.rs:697-755) with L in registers:
// b_i = (b_i - dotc(L[i+1..,i], b[i+1..])) / L_ii, the dot in dotx's order
// (acc_t sums positions t, t+8, ... of the full blocks of 8, then
// 0 + (acc0+acc4) + (acc1+acc5) + (acc2+acc6) + (acc3+acc7), then the tail).
// Products go through LDS; 8*CH lanes per group build the accumulators and CH
// lanes finish the CH dots and quotients.
// 64-bit DPP row shift left by N (lane l reads lane l + N of its row of 16);
// lanes whose source leaves the row keep 0.
template <int N>
__device__ __forceinline__ double row_shl(double v)
{
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)x, 0x100 + N, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), 0x100 + N, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// One group per wavefront (n uniform): lanes (q, t) = (lane / 8, lane % 8),
// lane < 8*CH, sum acc_t with a uniform trip count; the pairs acc_t + acc_t+4
// and the ordered combine move between the 8 lanes of a right-hand side by
// DPP, so each step has two LDS round trips instead of three.
template <int NB, bool EXACT>
__device__ __forceinline__ void backward_reg64(const double (&a)[NB], double (&b)[CH], FusedSmem<1> &s, int n,
                                               bool care, int qn, int r, double mydiag, double myinv)
{
    const int q = r >> 3, t = r & 7;
    for (int i = n - 1; i >= 0; --i) {
        const double li = reg_at<NB>(a, i);
#pragma unroll
        for (int qq = 0; qq < CH; ++qq) s.p[qq][r] = li * b[qq];
        if (r == i) {
#pragma unroll
            for (int qq = 0; qq < CH; ++qq) s.bi[0][qq] = b[qq];
        }
        lds_barrier();
        const int len = n - i - 1;
        const int full = len > 0 ? (len & ~7) : 0;
        const int nm = full >> 3, tl = len > 0 ? len - full : 0;
        const int row = i + 1;  // first row of the dot
        double acc = 0.0;
        if (q < CH)
            for (int m = 0; m < nm; ++m) acc += s.p[q][row + t + 8 * m];
        const double pair = acc + row_shl<4>(acc);  // lanes t < 4: acc_t + acc_t+4
        const double p1 = row_shl<1>(pair), p2 = row_shl<2>(pair), p3 = row_shl<3>(pair);
        const double di = lane_value(mydiag, i);
        const double yi = lane_value(myinv, i);
        if (q < CH && t == 0) {
            double res = 0.0;
            res += pair;
            res += p1;
            res += p2;
            res += p3;
            for (int m = 0; m < tl; ++m) res += s.p[q][row + full + m];
            const double num = s.bi[0][q] - res;
            double xq;
            if constexpr (EXACT) {
                xq = div_with(num, di, yi);
                if (care && q < qn && !(div_num_ok(num) && div_den_ok(di))) xq = num / di;
            } else {
                xq = num * yi;
            }
            s.x[0][q] = xq;
        }
        lds_barrier();
        if (r == i) {
#pragma unroll
            for (int qq = 0; qq < CH; ++qq) b[qq] = s.x[0][qq];
        }
    }
}

template <int NB, int G, bool EXACT>
__device__ __forceinline__ void backward_reg(const double (&a)[NB], double (&b)[CH], FusedSmem<64 / G> &s, int n,
                                             int nmax, bool care, int qn, int g, int r, int lane)
{
    for (int i = nmax - 1; i >= 0; --i) {
        const double li = reg_at<NB>(a, i);
#pragma unroll
        for (int q = 0; q < CH; ++q) s.p[q][lane] = li * b[q];
        if (r == i) {
#pragma unroll
            for (int q = 0; q < CH; ++q) s.bi[g][q] = b[q];
        }
        lds_barrier();
        const int len = n - i - 1;
        const int full = len > 0 ? (len & ~7) : 0;
        const int pg = g * G + i + 1;  // first row of the dot in s.p
        // Out-of-range slots add +0.0, an exact identity here: the running sums
        // start at +0 and cannot become -0.
#pragma unroll
        for (int idx0 = 0; idx0 < 8 * CH; idx0 += G) {
            const int idx = idx0 + r;
            if (idx < 8 * CH) {
                const int q = idx >> 3, t = idx & 7;
                double acc = 0.0;
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const int e = t + 8 * m;
                    acc += (e < full) ? s.p[q][pg + e] : 0.0;
                }
                s.acc[g][idx] = acc;
            }
        }
        lds_barrier();
        if (r < CH) {
            const int q = r;
            double res = 0.0;
            if (len > 0) {
                const double *ac = s.acc[g] + q * 8;
                res += ac[0] + ac[4];
                res += ac[1] + ac[5];
                res += ac[2] + ac[6];
                res += ac[3] + ac[7];
#pragma unroll
                for (int m = 0; m < 7; ++m) {
                    const int e = full + m;
                    res += (e < len) ? s.p[q][pg + e] : 0.0;
                }
            }
            const double num = s.bi[g][q] - res;
            const double di = s.dg[g * G + i];
            double xq;
            if constexpr (EXACT) {
                xq = div_with(num, di, s.inv[g * G + i]);
                const bool ok = !care || i >= n || q >= qn || (div_num_ok(num) && div_den_ok(di));
                if (__ballot(!ok) != 0) {
                    if (!ok) xq = num / di;
                }
            } else {
                xq = num * s.inv[g * G + i];
            }
            s.x[g][q] = xq;
        }
        lds_barrier();
        if (r == i) {
#pragma unroll
            for (int q = 0; q < CH; ++q) b[q] = s.x[g][q];
        }
        lds_barrier();
    }
}

// Member lookup shared by the kernels: order[begin..end) (sorted by decreasing
// n), or with order == nullptr the single matrix (A1, B1, n1, lda1, ldb1).
struct Member {
    int jm, n, lda, ldb;
    double *A, *B;
    bool valid;
};

template <int G>
__device__ __forceinline__ Member member(const int *order, int begin, int end, const int *nsz, double *const *Aptr,
                                         double *const *Bptr, double *A1, double *B1, int n1, int lda1, int ldb1,
                                         int g)
{
    constexpr int MPW = 64 / G;
    Member mb{0, 0, 0, 0, nullptr, nullptr, false};
    if (order) {
        const int item = begin + (int)blockIdx.x * MPW + g;
        mb.valid = item < end;
        if (mb.valid) {
            mb.jm = order[item];
            mb.n = nsz[mb.jm];
            mb.A = Aptr[mb.jm];
            mb.B = Bptr ? Bptr[mb.jm] : nullptr;
            mb.lda = mb.ldb = mb.n;
        }
    } else {
        mb.valid = g == 0;
        if (mb.valid) {
            mb.n = n1;
            mb.A = A1;
            mb.B = B1;
            mb.lda = lda1;
            mb.ldb = ldb1;
        }
    }
    return mb;
}

// Factorisation: one launch per size bin, the row array in registers.
template <int NB, int G, bool EXACT>
__global__ __launch_bounds__(64) void potrf_kernel(
    const int *__restrict__ order, int begin, int end, const int *__restrict__ nsz,
    double *const *__restrict__ Aptr, int *__restrict__ info, double *A1, int n1, int lda1, double sub)
{
    __shared__ Smem<64 / G> s;
    const int lane = threadIdx.x;
    const int g = lane / G, r = lane % G;
    const Member mb = member<G>(order, begin, end, nsz, Aptr, nullptr, A1, nullptr, n1, lda1, 0, g);
    // Members are sorted by decreasing size, so group 0 holds the largest n.
    const int nmax = __builtin_amdgcn_readfirstlane(mb.n);
    double a[NB];
    load_lower<NB>(a, mb.A, mb.lda, mb.n, r);
    (void)nmax;
    double mydiag;
    const int fail = factor<NB, G, EXACT>(a, s.col, s.xch[0], mb.n, g, r, lane, mydiag, sub);
    store_lower<NB>(a, mb.A, mb.lda, mb.n, r);
    if (mb.valid && r == 0 && info) info[order ? mb.jm : 0] = fail;
}

// Substitutions (Cholesky::solve_mut) with L read from memory: few registers,
// so many wavefronts hide the latency of the step-by-step dependence chains.
// Members whose factorisation failed keep their right-hand sides.
template <int NB, int G, bool EXACT>
__global__ __launch_bounds__(64) void potrs_kernel(
    const int *__restrict__ order, int begin, int end, const int *__restrict__ nsz,
    double *const *__restrict__ Aptr, double *const *__restrict__ Bptr, int nrhs, const int *__restrict__ info,
    double *A1, double *B1, int n1, int lda1, int ldb1)
{
    __shared__ Smem<64 / G> s;
    const int lane = threadIdx.x;
    const int g = lane / G, r = lane % G;
    const Member mb = member<G>(order, begin, end, nsz, Aptr, Bptr, A1, B1, n1, lda1, ldb1, g);
    const int nmax = __builtin_amdgcn_readfirstlane(mb.n);
    const int fail = mb.valid ? info[order ? mb.jm : 0] : 1;
    const int n = mb.n;
    s.diag[lane] = (r < n) ? mb.A[r + (size_t)r * mb.lda] : 1.0;
    lds_barrier();
    for (int c0 = 0; c0 < nrhs; c0 += CH) {
        double b[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q)
            b[q] = (r < n && c0 + q < nrhs && fail == 0) ? mb.B[r + (size_t)(c0 + q) * mb.ldb] : 0.0;
        forward<G, EXACT>(mb.A, mb.lda, b, s, n, nmax, g, r, lane);
        backward<G, EXACT>(mb.A, mb.lda, b, s, n, nmax, g, r, lane);
        if (fail == 0 && r < n) {
#pragma unroll
            for (int q = 0; q < CH; ++q)
                if (c0 + q < nrhs) mb.B[r + (size_t)(c0 + q) * mb.ldb] = b[q];
        }
    }
}

// Factorisation and solve in one pass: the factored rows stay in registers
// for the substitutions, so L is written once and never read back.  Members
// whose factorisation failed keep their right-hand sides.
template <int NB, int G, bool EXACT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NB >= 56 ? 3 : (NB == 48 || NB == 40 || NB == 32) ? 4 : 1)))
void potrf_potrs_kernel(const int *__restrict__ order, int begin, int end,
                                                         const int *__restrict__ nsz,
                                                         double *const *__restrict__ Aptr,
                                                         double *const *__restrict__ Bptr, int nrhs,
                                                         int *__restrict__ info)
{
    __shared__ FusedSmem<64 / G> s;
    const int lane = threadIdx.x;
    const int g = lane / G, r = lane % G;
    const Member mb = member<G>(order, begin, end, nsz, Aptr, Bptr, nullptr, nullptr, 0, 0, 0, g);
    const int n = mb.n;
    const int nmax = __builtin_amdgcn_readfirstlane(n);
    double a[NB];
    load_lower<NB>(a, mb.A, mb.lda, n, r);
    double mydiag;
    const int fail = factor<NB, G, EXACT>(a, s.col, s.xch, n, g, r, lane, mydiag);
    store_lower<NB>(a, mb.A, mb.lda, n, r);
    if (mb.valid && r == 0) info[mb.jm] = fail;
    if (nrhs <= 0 || Bptr == nullptr) return;

    const double myinv = EXACT ? recip_refined(mydiag) : 1.0 / mydiag;
    const bool mydok = div_den_ok(mydiag);
    s.dg[lane] = mydiag;
    s.inv[lane] = myinv;
    lds_barrier();
    const bool care = mb.valid && fail == 0;
    for (int c0 = 0; c0 < nrhs; c0 += CH) {
        const int qn = nrhs - c0 < CH ? nrhs - c0 : CH;
        double b[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q)
            b[q] = (care && r < n && q < qn) ? mb.B[r + (size_t)(c0 + q) * mb.ldb] : 0.0;
        forward_reg<NB, G, EXACT>(a, b, s, n, nmax, care, qn, g, r, lane, mydiag, myinv, mydok);
        if constexpr (G == 64)
            backward_reg64<NB, EXACT>(a, b, s, n, care, qn, r, mydiag, myinv);
        else
            backward_reg<NB, G, EXACT>(a, b, s, n, nmax, care, qn, g, r, lane);
        if (care && r < n) {
#pragma unroll
            for (int q = 0; q < CH; ++q)
                if (q < qn) mb.B[r + (size_t)(c0 + q) * mb.ldb] = b[q];
        }
    }
}

// Adjoint of X = A^{-1} B given L: Lambda = (L L^T)^{-1} dX (same arithmetic as
// Cholesky::solve_mut), dB = Lambda, dA = -1/2 (Lambda X^T + X Lambda^T).
template <int NB, int G, bool EXACT>
__global__ __launch_bounds__(64) void potrs_backward_kernel(
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

    // the diagonal of L, published for the substitutions; padding rows get 1
    s.diag[lane] = (r < n) ? L[r + (size_t)r * n] : 1.0;
    lds_barrier();

    for (int c0 = 0; c0 < nrhs; c0 += CH) {
        double b[CH], x[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const bool in = r < n && c0 + q < nrhs;
            b[q] = in ? dX[r + (size_t)(c0 + q) * n] : 0.0;
            x[q] = in ? X[r + (size_t)(c0 + q) * n] : 0.0;
        }
        forward<G, EXACT>(L, n, b, s, n, nmax, g, r, lane);
        backward<G, EXACT>(L, n, b, s, n, nmax, g, r, lane);
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
        lds_barrier();
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
        lds_barrier();
    }
    // no right-hand sides: X = A^{-1} B is empty and so is its gradient
    if (nrhs == 0 && r < n)
        for (int c = 0; c < n; ++c) dA[r + (size_t)c * n] = 0.0;
}

// Launches the fused kernel for members [begin, end) of `order` with G lanes
// per matrix (64 / G matrices per wavefront).
template <int NB, int G>
int fused_launch_g(bool exact, hipStream_t st, const int *order, int begin, int end, const int *nsz,
                   double *const *A, double *const *B, int nrhs, int *info)
{
    constexpr int MPW = 64 / G;
    const unsigned grid = (unsigned)((end - begin + MPW - 1) / MPW);
    if (exact)
        hipLaunchKernelGGL((potrf_potrs_kernel<NB, G, true>), dim3(grid), dim3(64), 0, st, order, begin, end, nsz,
                           A, B, nrhs, info);
    else
        hipLaunchKernelGGL((potrf_potrs_kernel<NB, G, false>), dim3(grid), dim3(64), 0, st, order, begin, end, nsz,
                           A, B, nrhs, info);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Lanes-per-matrix schedules compiled for padded size NB: every power of two
// from the smallest group that holds NB rows up to a whole wavefront.
template <int NB>
int fused_dispatch(bool exact, int G, hipStream_t st, const int *order, int begin, int end, const int *nsz,
                   double *const *A, double *const *B, int nrhs, int *info)
{
    if constexpr (NB <= 8)
        if (G == 8) return fused_launch_g<NB, 8>(exact, st, order, begin, end, nsz, A, B, nrhs, info);
    if constexpr (NB <= 16)
        if (G == 16) return fused_launch_g<NB, 16>(exact, st, order, begin, end, nsz, A, B, nrhs, info);
    if constexpr (NB <= 32)
        if (G == 32) return fused_launch_g<NB, 32>(exact, st, order, begin, end, nsz, A, B, nrhs, info);
    if (G == 64) return fused_launch_g<NB, 64>(exact, st, order, begin, end, nsz, A, B, nrhs, info);
    return 2;  // no such schedule for this size
}

}  // namespace nrocm_chol

#define NROCM_CHOL_DECLARE(NB)                                                                    \
    int nrocm_chol_potrf_launch_##NB(bool exact, unsigned grid, hipStream_t st, const int *order, \
                                     int begin, int end, const int *nsz, double *const *A,        \
                                     int *info, double *A1, int n1, int lda1, double sub);        \
    int nrocm_chol_potrs_launch_##NB(bool exact, unsigned grid, hipStream_t st, const int *order, \
                                     int begin, int end, const int *nsz, double *const *A,        \
                                     double *const *B, int nrhs, const int *info, double *A1,     \
                                     double *B1, int n1, int lda1, int ldb1);                     \
    int nrocm_chol_bwd_launch_##NB(unsigned grid, hipStream_t st, const int *order, int begin,    \
                                   int end, const int *nsz, const double *const *L,               \
                                   const double *const *X, const double *const *dX,               \
                                   double *const *dA, double *const *dB, int nrhs);               \
    int nrocm_chol_fused_launch_##NB(bool exact, int G, hipStream_t st, const int *order,         \
                                     int begin, int end, const int *nsz, double *const *A,        \
                                     double *const *B, int nrhs, int *info);

#define NROCM_CHOL_DEFINE(NB)                                                                     \
    int nrocm_chol_potrf_launch_##NB(bool exact, unsigned grid, hipStream_t st, const int *order, \
                                     int begin, int end, const int *nsz, double *const *A,        \
                                     int *info, double *A1, int n1, int lda1, double sub)         \
    {                                                                                             \
        using namespace nrocm_chol;                                                               \
        constexpr int G = group_for<NB>();                                                        \
        if (exact)                                                                                \
            hipLaunchKernelGGL((potrf_kernel<NB, G, true>), dim3(grid), dim3(64), 0, st, order,   \
                               begin, end, nsz, A, info, A1, n1, lda1, sub);                      \
        else                                                                                      \
            hipLaunchKernelGGL((potrf_kernel<NB, G, false>), dim3(grid), dim3(64), 0, st, order,  \
                               begin, end, nsz, A, info, A1, n1, lda1, sub);                      \
        return hipGetLastError() == hipSuccess ? 0 : 1;                                           \
    }                                                                                             \
    int nrocm_chol_potrs_launch_##NB(bool exact, unsigned grid, hipStream_t st, const int *order, \
                                     int begin, int end, const int *nsz, double *const *A,        \
                                     double *const *B, int nrhs, const int *info, double *A1,     \
                                     double *B1, int n1, int lda1, int ldb1)                      \
    {                                                                                             \
        using namespace nrocm_chol;                                                               \
        constexpr int G = group_for<NB>();                                                        \
        if (exact)                                                                                \
            hipLaunchKernelGGL((potrs_kernel<NB, G, true>), dim3(grid), dim3(64), 0, st, order,   \
                               begin, end, nsz, A, B, nrhs, info, A1, B1, n1, lda1, ldb1);        \
        else                                                                                      \
            hipLaunchKernelGGL((potrs_kernel<NB, G, false>), dim3(grid), dim3(64), 0, st, order,  \
                               begin, end, nsz, A, B, nrhs, info, A1, B1, n1, lda1, ldb1);        \
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
    }                                                                                             \
    int nrocm_chol_fused_launch_##NB(bool exact, int G, hipStream_t st, const int *order,         \
                                     int begin, int end, const int *nsz, double *const *A,        \
                                     double *const *B, int nrhs, int *info)                       \
    {                                                                                             \
        return nrocm_chol::fused_dispatch<NB>(exact, G, st, order, begin, end, nsz, A, B, nrhs,   \
                                              info);                                              \
    }

NROCM_CHOL_DECLARE(8)
NROCM_CHOL_DECLARE(16)
NROCM_CHOL_DECLARE(24)
NROCM_CHOL_DECLARE(32)
NROCM_CHOL_DECLARE(40)
NROCM_CHOL_DECLARE(48)
NROCM_CHOL_DECLARE(56)
NROCM_CHOL_DECLARE(64)
