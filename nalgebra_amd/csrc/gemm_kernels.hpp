// LDS-tiled MFMA GEMM for gfx950: C = alpha * A * B + beta * C with arbitrary
// element strides (nalgebra / matrixmultiply convention: element (i, j) of X
// at X[i*rsx + j*csx]).
//
// Tile 64 x 64 per 256-thread workgroup; each wavefront owns a 32 x 32 block
// as 2 x 2 MFMA 16 x 16 accumulators.  Compute modes:
//   MODE_F64  : v_mfma_f64_16x16x4f64   (FP64 in, FP64 accumulate)
//   MODE_F32  : v_mfma_f32_16x16x4f32   (FP32 in, FP32 accumulate)
//   MODE_BF16 : v_mfma_f32_16x16x32_bf16 (operands rounded to BF16, FP32 accumulate)
//   MODE_BF16X3: each FP32 operand split into three BF16 parts (hi + mid + lo,
//               Henry, Tang & Heinecke, ARITH 2019); the six largest cross
//               products run on the BF16 MFMA, giving ~FP32 accuracy
// Operand layouts (16 x 16 outputs): A fragment lane l holds row l%16 and
// k-slice l/16; B fragment lane l holds column l%16 and k-slice l/16.  The
// four accumulator values v of lane l sit in column l%16 and row
// 4*(l/16)+v for the FP32/BF16 instructions, row l/16+4*v for the FP64 one
// (measured on gfx950 with a one-wavefront probe).
#pragma once
#include "nrocm_internal.hpp"

#include <hip/hip_bf16.h>

namespace nrocm_gemm {

enum { MODE_F64 = 0, MODE_F32 = 1, MODE_BF16 = 2, MODE_BF16X3 = 3 };

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));

template <int MODE> struct Acc;
template <> struct Acc<MODE_F64> { using T = d4; using S = double; };
template <> struct Acc<MODE_F32> { using T = f4; using S = float; };
template <> struct Acc<MODE_BF16> { using T = f4; using S = float; };
template <> struct Acc<MODE_BF16X3> { using T = f4; using S = float; };

__device__ __forceinline__ short to_bf16_bits(float x)
{
    // round to nearest even (finite inputs); NaN stays NaN
    unsigned u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (short)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ __forceinline__ float bf16_bits_to_float(short b)
{
    return __uint_as_float(((unsigned)(unsigned short)b) << 16);
}

template <typename TS, int MODE>
__global__ __launch_bounds__(256) void gemm_kernel(int m, int n, int k, TS alpha, const TS *__restrict__ A,
                                                   long rsa, long csa, const TS *__restrict__ B, long rsb,
                                                   long csb, TS beta, TS *__restrict__ C, long rsc, long csc)
{
    constexpr int BM = 64, BN = 64;
    static_assert(MODE == MODE_F64 || MODE == MODE_F32, "general GEMM: FP64 or FP32 MFMA");
    constexpr int BK = 16;
    using AccT = typename Acc<MODE>::T;
    using AccS = typename Acc<MODE>::S;
    __shared__ TS As[BK][BM + 1];
    __shared__ TS Bs[BK][BN + 1];

    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int bm = blockIdx.x * BM, bn = blockIdx.y * BN;
    const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;

    AccT acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[a][b][v] = 0;

    for (int k0 = 0; k0 < k; k0 += BK) {
#pragma unroll
        for (int s = 0; s < BM * BK / 256; ++s) {
            const int e = tid + 256 * s;
            const int i = e % BM, kk = e / BM;
            const int gi = bm + i, gk = k0 + kk;
            As[kk][i] = (gi < m && gk < k) ? A[(long)gi * rsa + (long)gk * csa] : TS(0);
        }
#pragma unroll
        for (int s = 0; s < BN * BK / 256; ++s) {
            const int e = tid + 256 * s;
            const int kk = e % BK, j = e / BK;
            const int gj = bn + j, gk = k0 + kk;
            Bs[kk][j] = (gj < n && gk < k) ? B[(long)gk * rsb + (long)gj * csb] : TS(0);
        }
        __syncthreads();
        {
#pragma unroll
            for (int k4 = 0; k4 < BK; k4 += 4) {
                TS af[2], bf[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    af[t] = As[k4 + lk][wm + t * 16 + li];
                    bf[t] = Bs[k4 + lk][wn + t * 16 + li];
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        if constexpr (MODE == MODE_F64)
                            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
                        else
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bf[b], acc[a][b], 0, 0, 0);
                    }
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gi = bm + wm + a * 16 + (MODE == MODE_F64 ? lk + 4 * v : lk * 4 + v);
                const int gj = bn + wn + b * 16 + li;
                if (gi < m && gj < n) {
                    TS *cp = C + (long)gi * rsc + (long)gj * csc;
                    const TS val = alpha * (TS)(AccS)acc[a][b][v];
                    *cp = (beta == TS(0)) ? val : val + beta * *cp;
                }
            }
}

// ---------------------------------------------------------------- column-major GEMM
// C = alpha * A * B + beta * C with A (m x k), B (k x n) and C (m x n) all
// column-major (unit row stride): the layout of every nalgebra DMatrix.
//
// A workgroup of four wavefronts owns a BM x BN tile of C.  The wavefronts
// form a WGM x WGN grid over the tile and KSPLIT groups along k; each holds a
// (BM / WGM) x (BN / WGN) block as 16 x 16 MFMA accumulators.  Per BK-deep
// slice of k:
//   * the next slice's A (BM x BK) and B (BK x BN) are fetched into registers
//     with 16-byte loads along the columns before the current slice is
//     multiplied, so their latency hides under the MFMAs;
//   * the current slice is multiplied out of one of two LDS buffers while the
//     fetched slice goes into the other: one barrier per slice;
//   * with KSPLIT > 1, k-group q takes the 4-deep steps s with s % KSPLIT == q
//     and the groups' partial sums are added in LDS, in group order, at the
//     end (small problems: more wavefronts per output tile, same result on
//     every run).
// The MFMA computes the transposed block (first operand from B, second from
// A), so each store instruction covers 16 contiguous rows of 4 columns of C.
// Tiles are numbered so that each of the eight XCDs (workgroups are handed
// to them round-robin) takes a contiguous run of tiles, ordered in groups of
// 8 tile-rows, which keeps the A and B panels that run together in one L2.
template <int BM_, int BN_, int WGM_, int KSPLIT_, int BK_>
struct GemmShape {
    static constexpr int BM = BM_, BN = BN_, WGM = WGM_, KSPLIT = KSPLIT_, BK = BK_;
    static constexpr int WGN = 4 / (WGM * KSPLIT);
    static_assert(WGM * WGN * KSPLIT == 4, "four wavefronts");
    static constexpr int WTM = BM / WGM, WTN = BN / WGN;  // wavefront block
    static constexpr int FM = WTM / 16, FN = WTN / 16;    // 16 x 16 MFMA blocks per wavefront
    static_assert(FM * 16 == WTM && FN * 16 == WTN, "16 x 16 blocks");
    static_assert((BK / 4) % KSPLIT == 0, "k steps split evenly");
};
using GemmLarge = GemmShape<128, 128, 2, 1, 16>;  // wavefront 64 x 64
using GemmMedium = GemmShape<64, 64, 2, 1, 32>;   // wavefront 32 x 32
using GemmSmall = GemmShape<32, 32, 1, 4, 64>;    // 4 k-groups of one 32 x 32 block

template <typename TS> struct Vec16;
template <> struct Vec16<double> { typedef double T __attribute__((ext_vector_type(2))); static constexpr int N = 2; };
template <> struct Vec16<float> { typedef float T __attribute__((ext_vector_type(4))); static constexpr int N = 4; };

// wavefronts per SIMD the register allocation is held to: two for the FP64
// 128 x 128 shape (its 64 accumulator doubles per lane fit beside the operand
// registers in 256 VGPRs, and the second wavefront hides the slice loads and
// the barrier: 44 -> 67 TFLOP/s at 8192^3), the compiler's choice otherwise
template <typename TS, typename SH>
constexpr int gemm_waves()
{
    return sizeof(TS) == 8 && SH::BM == 128 ? 2 : 1;
}

template <typename TS, int MODE, typename SH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(gemm_waves<TS, SH>()))) void gemm_cm_kernel(int m, int n, int k, TS alpha, const TS *__restrict__ A,
                                                      long lda, const TS *__restrict__ B, long ldb, TS beta,
                                                      TS *__restrict__ C, long ldc, int vec_ok)
{
    static_assert(MODE == MODE_F64 || MODE == MODE_F32, "column-major GEMM: FP64 or FP32 MFMA");
    constexpr int BM = SH::BM, BN = SH::BN, BK = SH::BK, KSPLIT = SH::KSPLIT;
    constexpr int WGM = SH::WGM, FM = SH::FM, FN = SH::FN;
    using AccT = typename Acc<MODE>::T;
    using V = typename Vec16<TS>::T;
    constexpr int VN = Vec16<TS>::N;
    // row strides of 16 banks (mod 32) keep the fragment reads conflict-free
    constexpr int PAD = sizeof(TS) == 8 ? 8 : 16;
    constexpr int LDA_S = BM + PAD, LDB_S = BN + PAD;
    // A slice: TPA threads per column, CPA columns per pass, NPA passes
    constexpr int TPA = BM / VN, CPA = 256 / TPA, NPA = BK / CPA;
    constexpr int TPB = BK / VN, CPB = 256 / TPB, NPB = BN / CPB;
    static_assert(TPA * VN == BM && CPA * TPA == 256 && NPA * CPA == BK, "A slice mapping");
    static_assert(TPB * VN == BK && CPB * TPB == 256 && NPB * CPB == BN, "B slice mapping");
    __shared__ __align__(16) TS As[2][BK][LDA_S];  // As[b][kk][i] = A(bm + i, k0 + kk)
    __shared__ __align__(16) TS Bs[2][BK][LDB_S];  // Bs[b][kk][j] = B(k0 + kk, bn + j)

    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;

    // tile numbering: contiguous runs per XCD, then groups of 8 tile-rows
    const int tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
    const int nt = tm * tn;
    int t = blockIdx.x;
    if ((nt & 7) == 0) t = (t & 7) * (nt >> 3) + (t >> 3);
    constexpr int GROUP = 8;
    const int gsz = GROUP * tn;
    const int g0 = (t / gsz) * GROUP;
    const int gm = min(tm - g0, GROUP);
    const int bm = (g0 + (t % gsz) % gm) * BM;
    const int bn = ((t % gsz) / gm) * BN;

    const int kq = w % KSPLIT;  // k-group of this wavefront
    const int wv = w / KSPLIT;
    const int wm = (wv % WGM) * SH::WTM, wn = (wv / WGM) * SH::WTN;

    const bool full_mn = vec_ok && bm + BM <= m && bn + BN <= n;
    const int nk = (k + BK - 1) / BK;

    // this thread's part of a slice
    const int ar = (tid % TPA) * VN, ac = tid / TPA;  // A rows ar.., column ac + CPA * p
    const int bk = (tid % TPB) * VN, bc = tid / TPB;  // B rows bk.., column bc + CPB * p
    V ra[NPA], rb[NPB];

    auto fetch = [&](int k0) {
        if (full_mn && k0 + BK <= k) {
#pragma unroll
            for (int p = 0; p < NPA; ++p)
                ra[p] = *(const V *)(A + (bm + ar) + (long)(k0 + ac + CPA * p) * lda);
#pragma unroll
            for (int p = 0; p < NPB; ++p)
                rb[p] = *(const V *)(B + (k0 + bk) + (long)(bn + bc + CPB * p) * ldb);
        } else {
#pragma unroll
            for (int p = 0; p < NPA; ++p) {
                const int gk = k0 + ac + CPA * p;
                const TS *col = A + (long)min(gk, k - 1) * lda;
#pragma unroll
                for (int e = 0; e < VN; ++e) {
                    const int gi = bm + ar + e;
                    const TS x = col[min(gi, m - 1)];
                    ra[p][e] = (gi < m && gk < k) ? x : TS(0);
                }
            }
#pragma unroll
            for (int p = 0; p < NPB; ++p) {
                const int gj = bn + bc + CPB * p;
                const TS *col = B + (long)min(gj, n - 1) * ldb;
#pragma unroll
                for (int e = 0; e < VN; ++e) {
                    const int gk = k0 + bk + e;
                    const TS x = col[min(gk, k - 1)];
                    rb[p][e] = (gj < n && gk < k) ? x : TS(0);
                }
            }
        }
    };
    auto stage = [&](int b) {
#pragma unroll
        for (int p = 0; p < NPA; ++p) *(V *)&As[b][ac + CPA * p][ar] = ra[p];
#pragma unroll
        for (int p = 0; p < NPB; ++p)
#pragma unroll
            for (int e = 0; e < VN; ++e) Bs[b][bk + e][bc + CPB * p] = rb[p][e];
    };

    AccT acc[FM][FN];
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[a][b][v] = 0;

    if (nk > 0) {
        fetch(0);
        stage(0);
    }
    __syncthreads();
    for (int s = 0; s < nk; ++s) {
        const int cur = s & 1;
        if (s + 1 < nk) fetch((s + 1) * BK);
#pragma unroll
        for (int k4 = 4 * kq; k4 < BK; k4 += 4 * KSPLIT) {
            TS af[FM], bf[FN];
#pragma unroll
            for (int a = 0; a < FM; ++a) af[a] = As[cur][k4 + lk][wm + a * 16 + li];
#pragma unroll
            for (int b = 0; b < FN; ++b) bf[b] = Bs[cur][k4 + lk][wn + b * 16 + li];
#pragma unroll
            for (int a = 0; a < FM; ++a)
#pragma unroll
                for (int b = 0; b < FN; ++b) {
                    if constexpr (MODE == MODE_F64)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[b], af[a], acc[a][b], 0, 0, 0);
                    else
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[b], af[a], acc[a][b], 0, 0, 0);
                }
        }
        if (s + 1 < nk) stage(cur ^ 1);
        __syncthreads();
    }

    if constexpr (KSPLIT > 1) {
        // k-groups 1.. park their sums in the (now idle) operand buffers;
        // group 0 adds them in group order
        constexpr int NV = FM * FN * 4;
        static_assert((KSPLIT - 1) * NV * 64 * sizeof(TS) <= sizeof(As), "reduction fits");
        TS *red = &As[0][0][0];
        if (kq > 0) {
#pragma unroll
            for (int a = 0; a < FM; ++a)
#pragma unroll
                for (int b = 0; b < FN; ++b)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        red[((kq - 1) * NV + (a * FN + b) * 4 + v) * 64 + lane] = acc[a][b][v];
                    }
        }
        __syncthreads();
        if (kq > 0) return;
#pragma unroll
        for (int q = 1; q < KSPLIT; ++q)
#pragma unroll
            for (int a = 0; a < FM; ++a)
#pragma unroll
                for (int b = 0; b < FN; ++b)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        acc[a][b][v] += red[((q - 1) * NV + (a * FN + b) * 4 + v) * 64 + lane];
                    }
    }

#pragma unroll
    for (int a = 0; a < FM; ++a) {
        const int row = bm + wm + a * 16 + li;
#pragma unroll
        for (int b = 0; b < FN; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int col = bn + wn + b * 16 + (MODE == MODE_F64 ? lk + 4 * v : 4 * lk + v);
                if (row < m && col < n) {
                    TS *cp = C + row + (long)col * ldc;
                    const TS val = alpha * (TS)acc[a][b][v];
                    *cp = (beta == TS(0)) ? val : val + beta * *cp;
                }
            }
    }
}

// ---------------------------------------------------------------- rank-K update
// C (m x n) -= A (m x K) * B (K x n), all column-major, K <= 64: the trailing
// update of a blocked LU.  Each workgroup owns a BM x BN tile of C:
//   1. its C values are loaded straight into the MFMA accumulators, issued
//      first so they arrive while the operands are staged.  The MFMA computes
//      the transposed block (first operand from B, second from -A), so the 16
//      lanes of a lane group hold 16 consecutive rows of one column of C and
//      every load and store instruction covers 4 columns x 16 contiguous rows;
//   2. -A (BM x K) and B (K x BN) are staged in LDS in one pass, every load
//      of a thread in flight together (clamped in-bounds addresses, zeros
//      selected past the edges);
//   3. all K steps run back to back with no further barrier;
//   4. the accumulators are stored.
// FP32/BF16 modes: 128 x 128 tiles, each wavefront 64 x 64 (4 x 4 MFMA 16 x 16
// blocks); FP64: 64 x 64 tiles, each wavefront 32 x 32.
template <int MODE> struct RankTile { static constexpr int BM = 128, BN = 128; };
template <> struct RankTile<MODE_F64> { static constexpr int BM = 64, BN = 64; };

// output row (first-operand index) of a 16 x 16 MFMA block held by lane
// group lk as value v; the output column is the lane index mod 16
template <int MODE>
__device__ __forceinline__ int acc_row(int lk, int v)
{
    return MODE == MODE_F64 ? lk + 4 * v : 4 * lk + v;
}

template <typename TS, int MODE, bool LOOP>
__global__ __launch_bounds__(256) void rank_update_kernel(int m, int n, int K, const TS *__restrict__ A, long lda,
                                                          const TS *__restrict__ B, long ldb, TS *__restrict__ C,
                                                          long ldc)
{
    // LOOP: tiles are taken in a grid-stride loop, so a launch may use fewer
    // workgroups than tiles (and leave CUs to concurrent work); otherwise one
    // tile per workgroup
    constexpr int BM = RankTile<MODE>::BM, BN = RankTile<MODE>::BN, KM = 64;
    constexpr int FM = BM / 32, FN = BN / 32;  // 16 x 16 blocks per wavefront and dimension
    using AccT = typename Acc<MODE>::T;
    // row strides of 16 banks (mod 32) keep the fragment reads conflict-free
    constexpr int PAD = sizeof(TS) == 8 ? 8 : 16;
    __shared__ TS As[KM][BM + PAD];  // As[k][i] = -A(bm + i, k)
    __shared__ TS Bs[KM][BN + PAD];  // Bs[k][j] = B(k, bn + j)

    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int tm = (m + BM - 1) / BM;
    const int ntiles = tm * ((n + BN - 1) / BN);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int bm = (tile % tm) * BM, bn = (tile / tm) * BN;
        const int wm = (w & 1) * (BM / 2), wn = (w >> 1) * (BN / 2);
        const int li = lane & 15, lk = lane >> 4;

        // 1. C into the accumulators: acc[a][b][v] of lane (li, lk) is
        //    C(bm + wm + 16 a + li, bn + wn + 16 b + acc_row(lk, v))
        AccT acc[FM][FN];
#pragma unroll
        for (int a = 0; a < FM; ++a) {
            const int row = min(bm + wm + a * 16 + li, m - 1);
#pragma unroll
            for (int b = 0; b < FN; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int col = min(bn + wn + b * 16 + acc_row<MODE>(lk, v), n - 1);
                    acc[a][b][v] = C[row + (long)col * ldc];
                }
        }

        // 2. operands into LDS
        {
            constexpr int RA = 256 / BM;  // columns of A per round
            const int i = tid % BM, k0 = tid / BM;
            const TS *ap = A + min(bm + i, m - 1);
            const bool rok = bm + i < m;
#pragma unroll
            for (int c0 = 0; c0 < KM; c0 += 16 * RA) {
                TS v[16];
#pragma unroll
                for (int s = 0; s < 16; ++s) {
                    const int k = k0 + c0 + RA * s;
                    v[s] = ap[(long)min(k, K - 1) * lda];
                }
#pragma unroll
                for (int s = 0; s < 16; ++s) {
                    const int k = k0 + c0 + RA * s;
                    As[k][i] = (rok && k < K) ? -v[s] : TS(0);
                }
            }
        }
        {
            const int k = tid & 63, j0 = tid >> 6;
            const TS *bp = B + min(k, K - 1);
#pragma unroll
            for (int c0 = 0; c0 < BN; c0 += 64) {
                TS v[16];
#pragma unroll
                for (int s = 0; s < 16; ++s) {
                    const int j = j0 + c0 + 4 * s;
                    v[s] = bp[(long)min(bn + j, n - 1) * ldb];
                }
#pragma unroll
                for (int s = 0; s < 16; ++s) {
                    const int j = j0 + c0 + 4 * s;
                    Bs[k][j] = (k < K && bn + j < n) ? v[s] : TS(0);
                }
            }
        }
        __syncthreads();

        // 3. all K steps
        if constexpr (MODE == MODE_F64 || MODE == MODE_F32) {
#pragma unroll 4
            for (int k4 = 0; k4 < KM; k4 += 4) {
                TS af[FM], bf[FN];
#pragma unroll
                for (int a = 0; a < FM; ++a) af[a] = As[k4 + lk][wm + a * 16 + li];
#pragma unroll
                for (int b = 0; b < FN; ++b) bf[b] = Bs[k4 + lk][wn + b * 16 + li];
#pragma unroll
                for (int a = 0; a < FM; ++a)
#pragma unroll
                    for (int b = 0; b < FN; ++b) {
                        if constexpr (MODE == MODE_F64)
                            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[b], af[a], acc[a][b], 0, 0, 0);
                        else
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[b], af[a], acc[a][b], 0, 0, 0);
                    }
            }
        } else {
#pragma unroll 1
            for (int k32 = 0; k32 < KM; k32 += 32) {
                if constexpr (MODE == MODE_BF16) {
                    s8 af[FM], bf[FN];
#pragma unroll
                    for (int a = 0; a < FM; ++a)
#pragma unroll
                        for (int q = 0; q < 8; ++q) af[a][q] = to_bf16_bits((float)As[k32 + lk * 8 + q][wm + a * 16 + li]);
#pragma unroll
                    for (int b = 0; b < FN; ++b)
#pragma unroll
                        for (int q = 0; q < 8; ++q) bf[b][q] = to_bf16_bits((float)Bs[k32 + lk * 8 + q][wn + b * 16 + li]);
#pragma unroll
                    for (int a = 0; a < FM; ++a)
#pragma unroll
                        for (int b = 0; b < FN; ++b)
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[b], af[a], acc[a][b], 0, 0, 0);
                } else {  // MODE_BF16X3: three-way split, six cross products, smallest first
                    s8 ah[FM], am[FM], al[FM], bh[FN], bmd[FN], bl[FN];
#pragma unroll
                    for (int a = 0; a < FM; ++a)
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            const float x = (float)As[k32 + lk * 8 + q][wm + a * 16 + li];
                            const short h0 = to_bf16_bits(x);
                            const float r1 = x - bf16_bits_to_float(h0);
                            const short h1 = to_bf16_bits(r1);
                            ah[a][q] = h0;
                            am[a][q] = h1;
                            al[a][q] = to_bf16_bits(r1 - bf16_bits_to_float(h1));
                        }
#pragma unroll
                    for (int b = 0; b < FN; ++b)
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            const float y = (float)Bs[k32 + lk * 8 + q][wn + b * 16 + li];
                            const short g0 = to_bf16_bits(y);
                            const float s1 = y - bf16_bits_to_float(g0);
                            const short g1 = to_bf16_bits(s1);
                            bh[b][q] = g0;
                            bmd[b][q] = g1;
                            bl[b][q] = to_bf16_bits(s1 - bf16_bits_to_float(g1));
                        }
#pragma unroll
                    for (int a = 0; a < FM; ++a)
#pragma unroll
                        for (int b = 0; b < FN; ++b) {
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bmd[b], am[a], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[b], ah[a], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[b], al[a], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bmd[b], ah[a], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[b], am[a], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[b], ah[a], acc[a][b], 0, 0, 0);
                        }
                }
            }
        }

        // 4. store
#pragma unroll
        for (int a = 0; a < FM; ++a) {
            const int row = bm + wm + a * 16 + li;
#pragma unroll
            for (int b = 0; b < FN; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int col = bn + wn + b * 16 + acc_row<MODE>(lk, v);
                    if (row < m && col < n) C[row + (long)col * ldc] = acc[a][b][v];
                }
        }
        if constexpr (!LOOP) break;
        __syncthreads();  // the next tile overwrites As and Bs
    }
}
// max_groups > 0 caps the workgroups (each then loops over tiles).
template <typename TS, int MODE>
hipError_t rank_update(hipStream_t st, int m, int n, int K, const TS *A, long lda, const TS *B, long ldb, TS *C,
                       long ldc, int max_groups)
{
    if (m <= 0 || n <= 0 || K <= 0) return hipSuccess;
    if (K > 64) return hipErrorInvalidValue;
    constexpr int BM = RankTile<MODE>::BM, BN = RankTile<MODE>::BN;
    const long tiles = (long)((m + BM - 1) / BM) * ((n + BN - 1) / BN);
    if (max_groups > 0 && tiles > max_groups)
        hipLaunchKernelGGL((rank_update_kernel<TS, MODE, true>), dim3((unsigned)max_groups), dim3(256), 0, st, m, n,
                           K, A, lda, B, ldb, C, ldc);
    else
        hipLaunchKernelGGL((rank_update_kernel<TS, MODE, false>), dim3((unsigned)tiles), dim3(256), 0, st, m, n, K,
                           A, lda, B, ldb, C, ldc);
    return hipGetLastError();
}

template <typename TS, int MODE, typename SH>
hipError_t gemm_cm_launch(hipStream_t st, int m, int n, int k, TS alpha, const TS *A, long lda, const TS *B,
                          long ldb, TS beta, TS *C, long ldc)
{
    constexpr int VN = Vec16<TS>::N;
    const int vec_ok = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && lda % VN == 0 && ldb % VN == 0;
    const long tiles = (long)((m + SH::BM - 1) / SH::BM) * ((n + SH::BN - 1) / SH::BN);
    hipLaunchKernelGGL((gemm_cm_kernel<TS, MODE, SH>), dim3((unsigned)tiles), dim3(256), 0, st, m, n, k, alpha, A,
                       lda, B, ldb, beta, C, ldc, vec_ok);
    return hipGetLastError();
}

// Column-major operands (or all three row-major, computed as the transposed
// product): the largest tile shape that still gives every CU a workgroup,
// else the split-k shape.
template <typename TS, int MODE>
hipError_t gemm_cm(hipStream_t st, int m, int n, int k, TS alpha, const TS *A, long lda, const TS *B, long ldb,
                   TS beta, TS *C, long ldc)
{
    constexpr long CUS = 256;
    auto tiles = [&](int b) { return (long)((m + b - 1) / b) * ((n + b - 1) / b); };
    if (tiles(GemmLarge::BM) >= CUS)
        return gemm_cm_launch<TS, MODE, GemmLarge>(st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    if (tiles(GemmMedium::BM) >= CUS)
        return gemm_cm_launch<TS, MODE, GemmMedium>(st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    return gemm_cm_launch<TS, MODE, GemmSmall>(st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// Launches C = alpha*A*B + beta*C on `st`.  Returns a hipError_t.
template <typename TS, int MODE>
hipError_t gemm(hipStream_t st, int m, int n, int k, TS alpha, const TS *A, long rsa, long csa, const TS *B,
                long rsb, long csb, TS beta, TS *C, long rsc, long csc)
{
    if (m <= 0 || n <= 0) return hipSuccess;
    // a dimension of extent 1 (or an operand never read) has no stride to honour
    if (m == 1) rsa = rsc = 1;
    if (n == 1) csb = csc = 1;
    if (k == 1) csa = rsb = 1;
    if (k == 0) rsa = csa = rsb = csb = 1;
    if (rsa == 1 && rsb == 1 && rsc == 1)
        return gemm_cm<TS, MODE>(st, m, n, k, alpha, A, csa, B, csb, beta, C, csc);
    if (csa == 1 && csb == 1 && csc == 1)  // row-major: C^T = B^T A^T, all column-major
        return gemm_cm<TS, MODE>(st, n, m, k, alpha, B, rsb, A, rsa, beta, C, rsc);
    dim3 grid((m + 63) / 64, (n + 63) / 64);
    hipLaunchKernelGGL((gemm_kernel<TS, MODE>), grid, dim3(256), 0, st, m, n, k, alpha, A, rsa, csa, B, rsb,
                       csb, beta, C, rsc, csc);
    return hipGetLastError();
}

}  // namespace nrocm_gemm
