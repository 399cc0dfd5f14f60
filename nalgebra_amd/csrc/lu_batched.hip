// Batched (uneven sizes, n <= 64) LU with partial pivoting and solve:
// LU::new (src/linalg/lu.rs:93-122) then LU::solve_mut (lu.rs:242-260) per
// member, one launch per size class of the batch (padded sizes 8, 16, ..., 64).
//
// Layout: a group of G lanes serves one matrix (64 / G matrices per
// wavefront); lane r holds row r of its matrix in registers for the whole
// kernel.  Rows never move: each lane carries the current position of its
// row, and a swap of positions i and piv exchanges the two lanes' positions,
// which is nalgebra's full-row swap (lu.rs:114 swaps the left part too,
// gauss_step_swap :362-389 the rest).  Step i:
//   1. pivot: the first position >= i holding the strictly largest |a_i|
//      (icamax, src/base/min_max.rs:221-241; a NaN at position i wins, other
//      NaNs never do), reduced over the group;
//   2. a zero pivot skips the step (lu.rs:107-110);
//   3. the pivot row is broadcast through LDS; rows below take
//      l = fl(a_i * fl(1 / d)) and a_k = fl(a_k - fl(u_k * l)) for k > i, the
//      same unfused per-element sequence as gauss_step (NROCM_EXACT), or one
//      fused multiply-add (NROCM_FAST).
// The solve keeps the rows where they are: the right-hand side value of the
// row at position p is b_r of its original row r (PermutationSequence::
// permute_rows applies the same swaps), forward substitution with the unit
// lower factor (solve.rs:106-133) and backward substitution with U
// (solve.rs:137-182) broadcast the solved value of position i from the lane
// holding that position, and x is stored at the row's final position.
#include "internal_launch.hpp"

#include <algorithm>

namespace {

constexpr int CH = 4;  // right-hand sides held in registers at a time

template <int NB>
constexpr int group_for() { return NB <= 8 ? 8 : NB <= 16 ? 16 : NB <= 32 ? 32 : 64; }

template <bool EXACT>
__device__ __forceinline__ double sub_prod(double acc, double x, double y)
{
    // nalgebra's axpy with factor -x: fl(fl((-x) * y) + acc).  The negation
    // is applied to x before the product, as there, so that a NaN coming from
    // y keeps its sign.  One asm block keeps the product next to its addition
    // (as two IR operations the scheduler hoists every product of a step and
    // spills the row array), and it updates acc in place: with a separate
    // output the elimination of a step writes a second copy of the row and
    // moves it back at the end of the divergent region.
    if constexpr (EXACT) {
        double t;
        asm("v_mul_f64 %1, -%2, %3\n\tv_add_f64 %0, %1, %0" : "+v"(acc), "=&v"(t) : "v"(x), "v"(y));
    } else {
        asm("v_fma_f64 %0, -%1, %2, %0" : "+v"(acc) : "v"(x), "v"(y));
    }
    return acc;
}

// (value, position) under icamax's rule: larger value, then lower position.
__device__ __forceinline__ bool beats(double av, int ap, double bv, int bp)
{
    return av > bv || (av == bv && ap < bp);
}

// One DPP step of the pivot scan: (v, pos) against the pair the pattern CTRL
// brings in (rows outside ROWMASK, and lanes without a source, see the
// non-candidate pair (-1, INT_MAX), which never wins).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void pivot_dpp_step(double &v, int &pos)
{
    const long long x = __double_as_longlong(v);
    const long long o = __double_as_longlong(-1.0);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)x, CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), CTRL, ROWMASK, 0xf, false);
    const double ov = __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    const int op = __builtin_amdgcn_update_dpp(0x7fffffff, pos, CTRL, ROWMASK, 0xf, false);
    if (beats(ov, op, v, pos)) {
        v = ov;
        pos = op;
    }
}

// Group-wide pivot: every lane of the group ends with the winner's position.
// Groups of 16 or more lanes reduce by DPP scans (the rule is a total order,
// so the winner does not depend on the combination order): row_shr 1, 2, 4, 8
// leave each 16-lane row's winner in its lane 15; row_bcast 15 into rows 1
// and 3 only joins the rows of each 32-lane half, and row_bcast 31 the two
// halves.  Groups of 8 use a butterfly of lane shuffles (row_shr 4 would
// cross their boundary).
template <int G>
__device__ __forceinline__ void group_argmax(double &v, int &pos)
{
    if constexpr (G >= 16) {
        pivot_dpp_step<0x111, 0xf>(v, pos);
        pivot_dpp_step<0x112, 0xf>(v, pos);
        pivot_dpp_step<0x114, 0xf>(v, pos);
        pivot_dpp_step<0x118, 0xf>(v, pos);
        if constexpr (G >= 32) pivot_dpp_step<0x142, 0xa>(v, pos);
        if constexpr (G == 64) pivot_dpp_step<0x143, 0xc>(v, pos);
        const int lane = threadIdx.x & 63;
        if constexpr (G == 64) {
            pos = __builtin_amdgcn_readlane(pos, 63);
        } else if constexpr (G == 32) {
            const int p0 = __builtin_amdgcn_readlane(pos, 31), p1 = __builtin_amdgcn_readlane(pos, 63);
            pos = lane < 32 ? p0 : p1;
        } else {
            pos = __shfl(pos, lane | 15);
        }
    } else {
#pragma unroll
        for (int off = G / 2; off >= 1; off >>= 1) {
            const double ov = __shfl_xor(v, off, G);
            const int op = __shfl_xor(pos, off, G);
            if (beats(ov, op, v, pos)) {
                v = ov;
                pos = op;
            }
        }
    }
}

__device__ const double kZeroCellLU[1] = {0.0};

// v of lane src (wavefront-uniform) in every lane.
__device__ __forceinline__ double lane_value(double v, int src)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)bits, src);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int NB, int G, bool EXACT>
__global__ __launch_bounds__(64) void getrf_getrs_batched_kernel(
    const int *__restrict__ order, int begin, int end, const int *__restrict__ nsz, double *const *__restrict__ Aptr,
    int *const *__restrict__ Pptr, double *const *__restrict__ Bptr, int nrhs, int *__restrict__ info)
{
    constexpr int MPW = 64 / G;
    __shared__ __align__(16) double rowbuf[MPW][NB];
    __shared__ double bbuf[MPW][CH];
    __shared__ int where[MPW][NB];  // final position -> lane within the group
    const int lane = threadIdx.x;
    const int g = lane / G, r = lane % G;
    const int item = begin + (int)blockIdx.x * MPW + g;
    const bool valid = item < end;
    const int jm = valid ? order[item] : 0;
    const int n = valid ? nsz[jm] : 0;
    double *const A = valid ? Aptr[jm] : nullptr;
    int *const P = (valid && Pptr) ? Pptr[jm] : nullptr;
    const bool row_ok = r < n;

    double a[NB];
    {
        const double *row = row_ok ? A + r : kZeroCellLU;
        const size_t step = row_ok ? (size_t)n : 0;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const double *src = (c < n) ? row + (size_t)c * step : kZeroCellLU;
            a[c] = *src;
        }
    }
    int pos = r;
    int zmax = 0;  // 1 + the largest step with a zero pivot (U_ii == 0)
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const bool live = i < n;  // this group still has step i
        // 1. pivot search over the positions >= i
        double v = __builtin_fabs(a[i]);
        v = (v != v) ? ((pos == i) ? __builtin_inf() : -1.0) : v;
        const bool cand = live && row_ok && pos >= i;
        v = cand ? v : -1.0;
        int bp = cand ? pos : 0x7fffffff;
        group_argmax<G>(v, bp);
        const bool is_piv = cand && pos == bp;
        // 3. the pivot row (its entries from column i on) through LDS
        if (is_piv) {
#pragma unroll
            for (int k = i; k < NB; ++k) rowbuf[g][k] = a[k];
        }
        lds_barrier();
        const double d = live ? rowbuf[g][i] : 1.0;
        if (live && r == 0 && P) P[i] = (d == 0.0) ? i : bp;
        if (live && d == 0.0) zmax = i + 1;  // 2. zero pivot: nothing moves
        // zmax settled now: left to the compiler, its update sinks below the
        // loop and keeps every step's pivot value live to the end
        asm volatile("" : "+v"(zmax));
        if (live && d != 0.0) {
            if (pos == i) pos = bp;
            else if (is_piv) pos = i;
            if (pos > i && row_ok) {
                const double inv = 1.0 / d;
                const double l = a[i] * inv;
                a[i] = l;
#pragma unroll
                for (int k = i + 1; k < NB; ++k) a[k] = sub_prod<EXACT>(a[k], rowbuf[g][k], l);
            }
        }
        lds_barrier();
    }
    // the factors, each row at its final position
    if (row_ok) {
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < n) A[pos + (size_t)c * n] = a[c];
    }
    if (valid && r == 0) info[jm] = zmax;
    if (nrhs <= 0 || Bptr == nullptr) return;

    // ---- solve: skipped for a singular U (solve_mut returns false), whose
    // right-hand sides are left unchanged
    double *const B = valid ? Bptr[jm] : nullptr;
    const bool solve = valid && zmax == 0;
    if (row_ok) where[g][pos] = r;
    __syncthreads();
    for (int c0 = 0; c0 < nrhs; c0 += CH) {
        double b[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) b[q] = (solve && row_ok && c0 + q < nrhs) ? B[r + (size_t)(c0 + q) * n] : 0.0;
        // forward: unit lower, per element the steps i = 0, 1, ... in order.
        // A whole-wavefront group reads position i's values straight from the
        // lane holding it (a wavefront-uniform lane); smaller groups pass them
        // through LDS.
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const bool live = i + 1 < n;
            double bv[CH];
            if constexpr (G == 64) {
                const int src = __builtin_amdgcn_readfirstlane(live ? where[0][i] : 0);
#pragma unroll
                for (int q = 0; q < CH; ++q) bv[q] = lane_value(b[q], src);
            } else {
                if (live && where[g][i] == r) {
#pragma unroll
                    for (int q = 0; q < CH; ++q) bbuf[g][q] = b[q];
                }
                lds_barrier();
#pragma unroll
                for (int q = 0; q < CH; ++q) bv[q] = bbuf[g][q];
            }
            if (live && pos > i && row_ok) {
#pragma unroll
                for (int q = 0; q < CH; ++q) b[q] = sub_prod<EXACT>(b[q], bv[q], a[i]);
            }
            if constexpr (G != 64) lds_barrier();
            // one step at a time: interleaved across the unrolled steps, the
            // broadcasts and index reads run out of registers
            if constexpr (G == 64) __builtin_amdgcn_sched_barrier(0);
        }
        // backward: x_i = b_i / U_ii, then the positions above it
#pragma unroll
        for (int i = NB - 1; i >= 0; --i) {
            const bool live = i < n;
            double bv[CH];
            if constexpr (G == 64) {
                const int src = __builtin_amdgcn_readfirstlane(live ? where[0][i] : 0);
                if (live && src == r) {
#pragma unroll
                    for (int q = 0; q < CH; ++q) b[q] = b[q] / a[i];
                }
#pragma unroll
                for (int q = 0; q < CH; ++q) bv[q] = lane_value(b[q], src);
            } else {
                if (live && where[g][i] == r) {
#pragma unroll
                    for (int q = 0; q < CH; ++q) {
                        b[q] = b[q] / a[i];
                        bbuf[g][q] = b[q];
                    }
                }
                lds_barrier();
#pragma unroll
                for (int q = 0; q < CH; ++q) bv[q] = bbuf[g][q];
            }
            if (live && pos < i && row_ok) {
#pragma unroll
                for (int q = 0; q < CH; ++q) b[q] = sub_prod<EXACT>(b[q], bv[q], a[i]);
            }
            if constexpr (G != 64) lds_barrier();
            // one step at a time: interleaved across the unrolled steps, the
            // broadcasts and index reads run out of registers
            if constexpr (G == 64) __builtin_amdgcn_sched_barrier(0);
        }
        if (solve && row_ok) {
#pragma unroll
            for (int q = 0; q < CH; ++q)
                if (c0 + q < nrhs) B[pos + (size_t)(c0 + q) * n] = b[q];
        }
    }
}

// Accumulator t of dotx (src/base/blas.rs:93-167) over len products p[0..len)
// held in LDS: positions t, t + 8, ... of the full blocks of 8, in order.  The
// caller combines 0 + (acc0+acc4) + (acc1+acc5) + (acc2+acc6) + (acc3+acc7)
// and adds the tail.
__device__ __forceinline__ double dotx_acc(const double *p, int len, int t)
{
    const int full = len & ~7;
    double a = 0.0;
    for (int e = t; e < full; e += 8) a += p[e];
    return a;
}

// Adjoint of X = A^{-1} B from LU::new's packed factors and pivots, one
// member per group of G lanes (lane r holds row r of L \ U):
//   U^T y = g  (tr_solve_upper_triangular, solve.rs:255-276,411-440),
//   L^T z = y  (tr_solve_lower_triangular on the unit L, solve.rs:232-253,379-409),
//   Lambda = P^T z (inv_permute_rows, permutation_sequence.rs:108-116),
//   dB = Lambda, dA = -Lambda X^T (products summed over right-hand sides in
//   order), each dot in dotx's order.
template <int NB, int G>
__global__ __launch_bounds__(64) void getrs_backward_batched_kernel(
    const int *__restrict__ order, int begin, int end, const int *__restrict__ nsz,
    const double *const *__restrict__ LUptr, const int *const *__restrict__ Pptr,
    const double *const *__restrict__ Xptr, const double *const *__restrict__ dXptr,
    double *const *__restrict__ dAptr, double *const *__restrict__ dBptr, int nrhs)
{
    constexpr int MPW = 64 / G;
    __shared__ double prod[MPW][CH][NB];
    __shared__ double accs[MPW][CH][8];
    __shared__ double vbuf[MPW][CH][NB];  // z then Lambda, by position; X rows for dA
    const int lane = threadIdx.x;
    const int g = lane / G, r = lane % G;
    const int item = begin + (int)blockIdx.x * MPW + g;
    const bool valid = item < end;
    const int jm = valid ? order[item] : 0;
    const int n = valid ? nsz[jm] : 0;
    const double *const LU = valid ? LUptr[jm] : nullptr;
    const int *const P = valid ? Pptr[jm] : nullptr;
    const double *const X = valid ? Xptr[jm] : nullptr;
    const double *const dX = valid ? dXptr[jm] : nullptr;
    double *const dA = valid ? dAptr[jm] : nullptr;
    double *const dB = valid ? dBptr[jm] : nullptr;
    const bool row_ok = r < n;

    double a[NB];
    {
        const double *row = row_ok ? LU + r : kZeroCellLU;
        const size_t step = row_ok ? (size_t)n : 0;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const double *src = (c < n) ? row + (size_t)c * step : kZeroCellLU;
            a[c] = *src;
        }
    }
    for (int c0 = 0; c0 < nrhs; c0 += CH) {
        const int qn = nrhs - c0 < CH ? nrhs - c0 : CH;
        double y[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) y[q] = (row_ok && q < qn) ? dX[r + (size_t)(c0 + q) * n] : 0.0;
        // ---- U^T y = g, i ascending: y_i = (g_i - dotx(U[0..i, i], y[0..i])) / U_ii
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const bool live = i < n;
            if (r < NB) {
#pragma unroll
                for (int q = 0; q < CH; ++q) prod[g][q][r] = a[i] * y[q];
            }
            __syncthreads();
            if (live && r < 8) {
#pragma unroll
                for (int q = 0; q < CH; ++q) accs[g][q][r] = dotx_acc(prod[g][q], i, r);
            }
            __syncthreads();
            if (live && r == i) {
                const int full = i & ~7;
#pragma unroll
                for (int q = 0; q < CH; ++q) {
                    const double *ac = accs[g][q];
                    double res = 0.0;
                    res += ac[0] + ac[4];
                    res += ac[1] + ac[5];
                    res += ac[2] + ac[6];
                    res += ac[3] + ac[7];
                    for (int e = full; e < i; ++e) res += prod[g][q][e];
                    y[q] = (y[q] - res) / a[i];
                }
            }
            __syncthreads();
        }
        // ---- L^T z = y, i descending: z_i = (y_i - dotx(L[i+1.., i], z[i+1..])) / 1
#pragma unroll
        for (int i = NB - 1; i >= 0; --i) {
            const bool live = i < n;
            if (r < NB) {
#pragma unroll
                for (int q = 0; q < CH; ++q) prod[g][q][r] = a[i] * y[q];
            }
            __syncthreads();
            const int len = n - i - 1;
            if (live && r < 8) {
#pragma unroll
                for (int q = 0; q < CH; ++q) accs[g][q][r] = len > 0 ? dotx_acc(prod[g][q] + i + 1, len, r) : 0.0;
            }
            __syncthreads();
            if (live && r == i) {
                const int full = len > 0 ? (len & ~7) : 0;
#pragma unroll
                for (int q = 0; q < CH; ++q) {
                    const double *ac = accs[g][q];
                    double res = 0.0;
                    res += ac[0] + ac[4];
                    res += ac[1] + ac[5];
                    res += ac[2] + ac[6];
                    res += ac[3] + ac[7];
                    for (int e = full; e < len; ++e) res += prod[g][q][i + 1 + e];
                    y[q] = (y[q] - res) / 1.0;
                }
            }
            __syncthreads();
        }
        // ---- Lambda = P^T z: the swaps (i, ipiv[i]) in reverse order
        if (row_ok) {
#pragma unroll
            for (int q = 0; q < CH; ++q) vbuf[g][q][r] = y[q];
        }
        __syncthreads();
        if (valid && r == 0) {
            for (int i = n - 1; i >= 0; --i) {
                const int pv = P[i];
                if (pv != i) {
#pragma unroll
                    for (int q = 0; q < CH; ++q) {
                        const double t = vbuf[g][q][i];
                        vbuf[g][q][i] = vbuf[g][q][pv];
                        vbuf[g][q][pv] = t;
                    }
                }
            }
        }
        __syncthreads();
        double lam[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) lam[q] = row_ok ? vbuf[g][q][r] : 0.0;
        if (row_ok) {
#pragma unroll
            for (int q = 0; q < CH; ++q)
                if (q < qn) dB[r + (size_t)(c0 + q) * n] = lam[q];
        }
        __syncthreads();
        // ---- dA[r, c] = -(sum over right-hand sides of fl(Lambda_r * X_c)), X rows through LDS
        if (row_ok) {
#pragma unroll
            for (int q = 0; q < CH; ++q) vbuf[g][q][r] = (q < qn) ? X[r + (size_t)(c0 + q) * n] : 0.0;
        }
        __syncthreads();
        if (row_ok) {
            for (int c = 0; c < n; ++c) {
                double sacc = c0 > 0 ? -dA[r + (size_t)c * n] : 0.0;
#pragma unroll
                for (int q = 0; q < CH; ++q) {
                    if (q < qn) {
                        const double pr = lam[q] * vbuf[g][q][c];
                        sacc = (c0 == 0 && q == 0) ? pr : sacc + pr;
                    }
                }
                dA[r + (size_t)c * n] = -sacc;
            }
        }
        __syncthreads();
    }
    // no right-hand sides: X and its gradient are empty
    if (nrhs == 0 && row_ok)
        for (int c = 0; c < n; ++c) dA[r + (size_t)c * n] = 0.0;
}

template <int NB>
int launch_bwd_nb(hipStream_t st, const int *order, int begin, int end, const int *nsz, const double *const *LU,
                  const int *const *P, const double *const *X, const double *const *dX, double *const *dA,
                  double *const *dB, int nrhs)
{
    constexpr int G = group_for<NB>();
    constexpr int MPW = 64 / G;
    const unsigned grid = (unsigned)((end - begin + MPW - 1) / MPW);
    if (grid == 0) return 0;
    hipLaunchKernelGGL((getrs_backward_batched_kernel<NB, G>), dim3(grid), dim3(64), 0, st, order, begin, end, nsz,
                       LU, P, X, dX, dA, dB, nrhs);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

template <int NB>
int launch_nb(bool exact, hipStream_t st, const int *order, int begin, int end, const int *nsz, double *const *A,
              int *const *P, double *const *B, int nrhs, int *info)
{
    constexpr int G = group_for<NB>();
    constexpr int MPW = 64 / G;
    const unsigned grid = (unsigned)((end - begin + MPW - 1) / MPW);
    if (grid == 0) return 0;
    if (exact)
        hipLaunchKernelGGL((getrf_getrs_batched_kernel<NB, G, true>), dim3(grid), dim3(64), 0, st, order, begin, end,
                           nsz, A, P, B, nrhs, info);
    else
        hipLaunchKernelGGL((getrf_getrs_batched_kernel<NB, G, false>), dim3(grid), dim3(64), 0, st, order, begin,
                           end, nsz, A, P, B, nrhs, info);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

using Launch = int (*)(bool, hipStream_t, const int *, int, int, const int *, double *const *, int *const *,
                       double *const *, int, int *);
const Launch kLaunch[8] = {launch_nb<8>, launch_nb<16>, launch_nb<24>, launch_nb<32>,
                           launch_nb<40>, launch_nb<48>, launch_nb<56>, launch_nb<64>};
using BwdLaunch = int (*)(hipStream_t, const int *, int, int, const int *, const double *const *, const int *const *,
                          const double *const *, const double *const *, double *const *, double *const *, int);
const BwdLaunch kBwdLaunch[8] = {launch_bwd_nb<8>,  launch_bwd_nb<16>, launch_bwd_nb<24>, launch_bwd_nb<32>,
                                 launch_bwd_nb<40>, launch_bwd_nb<48>, launch_bwd_nb<56>, launch_bwd_nb<64>};

}  // namespace

extern "C" int nrocm_dgetrf_getrs_batched(nrocm_handle h, nrocm_batch bt, double *const *a, int *const *ipiv,
                                          double *const *b, int *info, int mode)
{
    if (!h) return -1;
    if (!bt || bt->handle != h) return -2;
    if (bt->count > 0 && !a) return -3;
    if (bt->count > 0 && bt->nrhs > 0 && !b) return -5;
    if (bt->count > 0 && !info) return -6;
    if (mode != NROCM_EXACT && mode != NROCM_FAST) return -7;
    if (bt->count == 0) return NROCM_OK;
    if (bt->max_n > 64) return nrocm_fail(h, NROCM_ERR_UNSUPPORTED, "batched LU supports n <= 64");
    // the bins are independent: dealt round-robin (sizes in decreasing order)
    // to the stream and one auxiliary stream, whose kernels then run side by side
    int s = nrocm_aux_init(h);
    if (s == NROCM_OK) s = nrocm_fork_on(h, h->stream, 1);
    if (s != NROCM_OK) return s;
    int k = 0;
    for (const auto &bin : bt->bins) {
        if (bin.end <= bin.begin) continue;
        hipStream_t bs = (k++ & 1) ? h->aux[0] : h->stream;
        if (kLaunch[bin.kernel](mode == NROCM_EXACT, bs, bt->order_dev, bin.begin, bin.end, bt->n_dev, a, ipiv,
                                bt->nrhs > 0 ? b : nullptr, bt->nrhs, info))
            NROCM_CHECK_LAUNCH(h);
    }
    return nrocm_join_on(h, h->stream, 1);
}

extern "C" int nrocm_dgetrs_backward_batched(nrocm_handle h, nrocm_batch bt, const double *const *lu,
                                             const int *const *ipiv, const double *const *x,
                                             const double *const *dx, double *const *da, double *const *db)
{
    if (!h) return -1;
    if (!bt || bt->handle != h) return -2;
    if (bt->count > 0 && !lu) return -3;
    if (bt->count > 0 && !ipiv) return -4;
    if (bt->count > 0 && bt->nrhs > 0 && !x) return -5;
    if (bt->count > 0 && bt->nrhs > 0 && !dx) return -6;
    if (bt->count > 0 && !da) return -7;
    if (bt->count > 0 && bt->nrhs > 0 && !db) return -8;
    if (bt->count == 0) return NROCM_OK;
    if (bt->max_n > 64) return nrocm_fail(h, NROCM_ERR_UNSUPPORTED, "batched LU adjoint supports n <= 64");
    for (const auto &bin : bt->bins) {
        if (bin.end <= bin.begin) continue;
        if (kBwdLaunch[bin.kernel](h->stream, bt->order_dev, bin.begin, bin.end, bt->n_dev, lu, ipiv, x, dx, da, db,
                                   bt->nrhs))
            NROCM_CHECK_LAUNCH(h);
    }
    return NROCM_OK;
}
