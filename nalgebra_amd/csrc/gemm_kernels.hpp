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

// Launches C = alpha*A*B + beta*C on `st`.  Returns a hipError_t.
template <typename TS, int MODE>
hipError_t gemm(hipStream_t st, int m, int n, int k, TS alpha, const TS *A, long rsa, long csa, const TS *B,
                long rsb, long csb, TS beta, TS *C, long rsc, long csc)
{
    if (m <= 0 || n <= 0) return hipSuccess;
    dim3 grid((m + 63) / 64, (n + 63) / 64);
    hipLaunchKernelGGL((gemm_kernel<TS, MODE>), grid, dim3(256), 0, st, m, n, k, alpha, A, rsa, csa, B, rsb,
                       csb, beta, C, rsc, csc);
    return hipGetLastError();
}

}  // namespace nrocm_gemm
