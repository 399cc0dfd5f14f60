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
    constexpr int BK = (MODE == MODE_BF16 || MODE == MODE_BF16X3) ? 32 : 16;
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
        if constexpr (MODE == MODE_BF16X3) {
            s8 ah[2], am[2], al[2], bh[2], bm_[2], bl[2];
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    float x = (float)As[lk * 8 + q][wm + t * 16 + li];
                    short h0 = to_bf16_bits(x);
                    float r1 = x - bf16_bits_to_float(h0);
                    short h1 = to_bf16_bits(r1);
                    short h2 = to_bf16_bits(r1 - bf16_bits_to_float(h1));
                    ah[t][q] = h0; am[t][q] = h1; al[t][q] = h2;
                    float y = (float)Bs[lk * 8 + q][wn + t * 16 + li];
                    short g0 = to_bf16_bits(y);
                    float s1 = y - bf16_bits_to_float(g0);
                    short g1 = to_bf16_bits(s1);
                    short g2 = to_bf16_bits(s1 - bf16_bits_to_float(g1));
                    bh[t][q] = g0; bm_[t][q] = g1; bl[t][q] = g2;
                }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    // smallest terms first
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[a], bm_[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bm_[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[a], bh[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
                }
        } else if constexpr (MODE == MODE_BF16) {
            s8 af[2], bf[2];
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    af[t][q] = to_bf16_bits((float)As[lk * 8 + q][wm + t * 16 + li]);
                    bf[t][q] = to_bf16_bits((float)Bs[lk * 8 + q][wn + t * 16 + li]);
                }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bf[b], acc[a][b], 0, 0, 0);
        } else {
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
