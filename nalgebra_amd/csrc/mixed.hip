// Mixed-precision LU + FP64 iterative refinement (nrocm_dgesv_ir).
//
//   A32 = fl32(A)                      working copy (FP64 kept for residuals)
//   P A32 = L U                        blocked LU (partial pivoting, or with
//                                      NROCM_PREC_WINDOW pivots searched in the
//                                      next 128 rows), trailing updates on FP32 or
//                                      BF16 MFMA (FP32 accumulate); FP64 mode
//                                      factors a copy of A with FP64 MFMA
//   x = 0, r = b
//   repeat:  d = U^{-1} L^{-1} P r     (working precision, single-launch
//                                       look-back triangular solves with the
//                                       64x64 diagonal blocks pre-inverted)
//            x += d                    (FP64)
//            r = b - A x               (FP64, deterministic two-pass GEMV)
//            berr = |r|_inf / (|A|_inf |x|_inf + |b|_inf)
//   until berr <= tol or maxit.
#include "internal_launch.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int BS = 64;           // triangular-solve block
constexpr int GEMV_CS = 8;       // column splits of the residual GEMV
constexpr long SPIN_LIMIT = 1L << 26;
constexpr int GM_RESTART = 20;   // FGMRES cycle length

__global__ void cvt_d2s_kernel(const double *__restrict__ a, int lda, float *__restrict__ o, int ldo, int n)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t tot = (size_t)n * n;
    if (idx >= tot) return;
    const int r = (int)(idx % n), c = (int)(idx / n);
    o[r + (size_t)c * ldo] = (float)a[r + (size_t)c * lda];
}

__global__ void copy_d_kernel(const double *__restrict__ a, int lda, double *__restrict__ o, int ldo, int n)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t tot = (size_t)n * n;
    if (idx >= tot) return;
    const int r = (int)(idx % n), c = (int)(idx / n);
    o[r + (size_t)c * ldo] = a[r + (size_t)c * lda];
}

// Final row order of P: perm[i] = original row that ends at position i.
__global__ void perm_kernel(const int *ipiv, int n, int *perm)
{
    extern __shared__ int p[];
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = i;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < n; ++i) {
            const int q = ipiv[i];
            if (q != i) {
                const int t = p[i];
                p[i] = p[q];
                p[q] = t;
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = p[i];
}

template <typename T>
__global__ void gather_kernel(const double *__restrict__ r, const int *__restrict__ perm, T *__restrict__ o, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = (T)r[perm[i]];
}

__global__ void zero_int_kernel(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Inverses of the 64x64 diagonal blocks of the packed L\U factor: the unit
// lower blocks of L (LOWER) or the upper blocks of U.  One wavefront per block;
// lane c computes column c of the inverse by substitution against the block
// held in LDS (rows past n are padded with the identity).
template <typename T, bool LOWER>
__global__ __launch_bounds__(64) void inv_diag_kernel(const T *__restrict__ M, int lda, int n, T *__restrict__ Inv)
{
    __shared__ T D[BS][BS + 1];
    const int R = blockIdx.x;
    const int r0 = R * BS;
    const int rows = min(BS, n - r0);
    const int c = threadIdx.x;
    {
        // lane c stages row c: loads in flight 16 at a time (clamped in-bounds
        // addresses, identity padding selected afterwards)
        const T *mrow = M + (r0 + (c < rows ? c : 0));
#pragma unroll
        for (int b0 = 0; b0 < BS; b0 += 16) {
            T v[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = mrow[(size_t)(r0 + (b0 + j < rows ? b0 + j : 0)) * lda];
#pragma unroll
            for (int j = 0; j < 16; ++j)
                D[c][b0 + j] = (c < rows && b0 + j < rows) ? v[j] : (c == b0 + j ? T(1) : T(0));
        }
    }
    __syncthreads();
    T x[BS];
    if constexpr (LOWER) {
#pragma unroll
        for (int i = 0; i < BS; ++i) {
            T sacc = (i == c) ? T(1) : T(0);
#pragma unroll
            for (int k = 0; k < BS; ++k)  // constant trip count: unrolls before i does
                if (k < i) sacc = fma(-D[i][k], x[k], sacc);
            x[i] = sacc;
        }
    } else {
#pragma unroll
        for (int i = BS - 1; i >= 0; --i) {
            T sacc = (i == c) ? T(1) : T(0);
#pragma unroll
            for (int k = 0; k < BS; ++k)
                if (k > i) sacc = fma(-D[i][k], x[k], sacc);
            x[i] = sacc / D[i][i];
        }
    }
    T *out = Inv + (size_t)R * BS * BS + (size_t)c * BS;
#pragma unroll
    for (int i = 0; i < BS; ++i) out[i] = x[i];
}

// Single-launch blocked triangular solve of L y = v (unit lower, LOWER) or
// U y = v (non-unit upper) with the inverted diagonal blocks.  Work items are
// 64-row blocks, taken in solve order through an atomic ticket, so every block
// a workgroup waits for is owned by a workgroup that is already running.  The
// off-diagonal block of step j+1 is loaded while step j waits.
//
// A solved block is published in `tagged`: every 32-bit half of a value sits
// in one 64-bit word next to this launch's epoch, stored as one atomic word.
// A consumer lane polls its own words until they carry the epoch, so the data
// and its readiness arrive together (no flag, no fence).  The array is zeroed
// before the first launch of a solve and epochs are never reused on a handle,
// so no word carries the current epoch before its value is written.  Waits are
// bounded: on timeout the error word is set and the kernel still exits.
template <typename T>
__device__ __forceinline__ T untag(const unsigned long long (&w)[sizeof(T) / 4])
{
    if constexpr (sizeof(T) == 8)
        return __longlong_as_double((long long)(((w[1] & 0xffffffffull) << 32) | (w[0] & 0xffffffffull)));
    else
        return __uint_as_float((unsigned)w[0]);
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void trsv_lookback_kernel(const T *__restrict__ M, int lda, int n,
                                                            const T *__restrict__ Inv, T *v, int *ticket,
                                                            unsigned long long *tagged, unsigned epoch, int *err)
{
    constexpr int WPV = sizeof(T) / 4;  // tagged words per value
    constexpr int PER = BS / 4;  // columns per thread and block
    __shared__ T xs[BS];
    __shared__ T part[4][BS];
    __shared__ T ys[BS];
    __shared__ int my;
    const int tid = threadIdx.x;
    const int nblk = (n + BS - 1) / BS;
    if (tid == 0) my = atomicAdd(ticket, 1);
    __syncthreads();
    const int k = my;                               // k-th block in solve order
    const int R = LOWER ? k : nblk - 1 - k;         // row block
    const int r0 = R * BS;
    const int rows = min(BS, n - r0);
    const int rr = tid & 63, sl = tid >> 6;         // row within block, column slice
    T inv[PER];
    {
        const T *IR = Inv + (size_t)R * BS * BS;
#pragma unroll
        for (int q = 0; q < PER; ++q) inv[q] = IR[rr + (size_t)(sl + 4 * q) * BS];
    }
    auto load_block = [&](int j, T (&dst)[PER]) {
        const int J = LOWER ? j : nblk - 1 - j;
        const int c0 = J * BS;
        const int cols = min(BS, n - c0);
        const T *src = M + (r0 + rr) + (size_t)c0 * lda;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int c = sl + 4 * q;
            dst[q] = (rr < rows && c < cols) ? src[(size_t)c * lda] : T(0);
        }
    };
    // this block's own right-hand side: written only by this workgroup, so it
    // is loaded now rather than on the chain's critical path
    const T vown = (tid < rows) ? v[r0 + tid] : T(0);
    T acc = 0;
    T nxt[PER];
    if (k > 0) load_block(0, nxt);
    for (int j = 0; j < k; ++j) {
        const int J = LOWER ? j : nblk - 1 - j;
        const int c0 = J * BS;
        const int cols = min(BS, n - c0);
        T cur[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) cur[q] = nxt[q];
        if (j + 1 < k) load_block(j + 1, nxt);
        if (tid < BS) {
            // every slot is written: entries past a short last block are zero
            // (the products below read all BS slots, and 0 * stale LDS could be NaN)
            T xv = T(0);
            if (tid < cols) {
                unsigned long long w[WPV];
#pragma unroll
                for (int h = 0; h < WPV; ++h) {
                    unsigned long long *src = tagged + (size_t)(c0 + tid) * WPV + h;
                    long spins = 0;
                    unsigned long long word;
                    while (((word = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 32) !=
                           epoch) {
                        if (++spins > SPIN_LIMIT) {
                            atomicExch(err, 1);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    w[h] = word;
                }
                xv = untag<T>(w);
            }
            xs[tid] = xv;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q) acc = fma(-cur[q], xs[sl + 4 * q], acc);  // cur is 0 outside the block
        __syncthreads();
    }
    part[sl][rr] = acc;
    __syncthreads();
    if (tid < BS) ys[tid] = (tid < rows) ? vown + ((part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid])) : T(0);
    __syncthreads();
    T sacc = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) sacc = fma(inv[q], ys[sl + 4 * q], sacc);
    part[sl][rr] = sacc;
    __syncthreads();
    if (tid < rows) {
        const T y = (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
        v[r0 + tid] = y;
        unsigned long long bits;
        if constexpr (sizeof(T) == 8) bits = (unsigned long long)__double_as_longlong(y);
        else bits = __float_as_uint(y);
#pragma unroll
        for (int h = 0; h < WPV; ++h)
            __hip_atomic_store(tagged + (size_t)(r0 + tid) * WPV + h,
                               ((unsigned long long)epoch << 32) | ((bits >> (32 * h)) & 0xffffffffull),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// partial[cs][i] = sum over the cs-th column slice of A[i, c] * x[c]  (or |A| * 1)
template <bool ABS>
__global__ __launch_bounds__(256) void gemv_part_kernel(const double *__restrict__ A, int lda, int n,
                                                        const double *__restrict__ x, double *__restrict__ part)
{
    __shared__ double red[4][64];
    const int r = blockIdx.x * 64 + (threadIdx.x & 63);
    const int sl = threadIdx.x >> 6;
    const int cs = blockIdx.y;
    const int cper = (n + GEMV_CS - 1) / GEMV_CS;
    const int c0 = cs * cper, c1 = min(n, c0 + cper);
    double acc = 0.0;
    if (r < n) {
        for (int c = c0 + sl; c < c1; c += 4) {
            const double a = A[r + (size_t)c * lda];
            acc = ABS ? acc + fabs(a) : fma(a, x[c], acc);
        }
    }
    red[sl][threadIdx.x & 63] = acc;
    __syncthreads();
    if (sl == 0 && r < n) {
        const int t = threadIdx.x;
        part[(size_t)cs * n + r] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
    }
}

// r = b - sum_cs part[cs]  (fixed order), and max |r_i| via ordered-bits atomicMax.
__global__ void residual_kernel(const double *__restrict__ b, const double *__restrict__ part, int n,
                                double *__restrict__ r, unsigned long long *rmax)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int cs = 0; cs < GEMV_CS; ++cs) s += part[(size_t)cs * n + i];
    const double ri = b[i] - s;
    r[i] = ri;
    atomicMax(rmax, (unsigned long long)__double_as_longlong(fabs(ri)));
}

__global__ void rowsum_max_kernel(const double *__restrict__ part, int n, unsigned long long *amax)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int cs = 0; cs < GEMV_CS; ++cs) s += part[(size_t)cs * n + i];
    atomicMax(amax, (unsigned long long)__double_as_longlong(s));
}

template <typename T>
__global__ void update_x_kernel(double *__restrict__ x, const T *__restrict__ d, int n, unsigned long long *xmax)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = x[i] + (double)d[i];
    x[i] = xi;
    atomicMax(xmax, (unsigned long long)__double_as_longlong(fabs(xi)));
}

__global__ void absmax_kernel(const double *__restrict__ v, int n, unsigned long long *m)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicMax(m, (unsigned long long)__double_as_longlong(fabs(v[i])));
}

__global__ void zero_u64_kernel(unsigned long long *p, int n)
{
    const int i = threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void zero_words_kernel(unsigned long long *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void zero_d_kernel(double *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------- FGMRES helpers (FP64)

// h[j] = sum_i V[i + j ldv] w[i], j < k: one workgroup per j, fixed-order
// strided partial sums and a fixed tree, so results are deterministic.
__global__ __launch_bounds__(256) void dots_kernel(const double *__restrict__ V, int ldv, const double *__restrict__ w,
                                                   int n, double *__restrict__ hout)
{
    __shared__ double red[256];
    const int j = blockIdx.x, t = threadIdx.x;
    const double *v = V + (size_t)j * ldv;
    double sacc = 0.0;
    for (int i = t; i < n; i += 256) sacc = fma(v[i], w[i], sacc);
    red[t] = sacc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) hout[j] = red[0];
}

// w[i] -= sum_{j<k} V[i + j ldv] h[j]
__global__ void sub_vh_kernel(double *__restrict__ w, const double *__restrict__ V, int ldv, int k,
                              const double *__restrict__ hv, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sacc = 0.0;
    for (int j = 0; j < k; ++j) sacc = fma(V[i + (size_t)j * ldv], hv[j], sacc);
    w[i] -= sacc;
}

// o[i] = w[i] / sqrt(*nrm2)
__global__ void scale_by_norm_kernel(const double *__restrict__ w, const double *__restrict__ nrm2, int n,
                                     double *__restrict__ o)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = w[i] / sqrt(*nrm2);
}

// w = sum_cs part[cs] (the second pass of the deterministic GEMV)
__global__ void gemv_combine_kernel(const double *__restrict__ part, int n, double *__restrict__ w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sacc = 0.0;
    for (int cs = 0; cs < GEMV_CS; ++cs) sacc += part[(size_t)cs * n + i];
    w[i] = sacc;
}

// x[i] += sum_{j<k} Z[i + j ldz] y[j], with max |x_i|
__global__ void xupdate_kernel(double *__restrict__ x, const double *__restrict__ Z, int ldz, int k,
                               const double *__restrict__ y, int n, unsigned long long *xmax)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sacc = 0.0;
    for (int j = 0; j < k; ++j) sacc = fma(Z[i + (size_t)j * ldz], y[j], sacc);
    const double xi = x[i] + sacc;
    x[i] = xi;
    atomicMax(xmax, (unsigned long long)__double_as_longlong(fabs(xi)));
}

template <typename T>
__global__ void widen_kernel(const T *__restrict__ d, int n, double *__restrict__ z)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = (double)d[i];
}

#define IR_TRY(x)                                                       \
    do {                                                                \
        hipError_t e__ = (x);                                           \
        if (e__ != hipSuccess) return nrocm_fail_hip(h, e__, #x);       \
    } while (0)

template <typename T>
int trsv_pair(nrocm_handle h, hipStream_t st, const T *LU, int lda, int n, const T *InvL, const T *InvU, T *v,
              int *ticket, unsigned long long *tagged, int *err)
{
    const int nblk = (n + BS - 1) / BS;
    for (int pass = 0; pass < 2; ++pass) {
        if (++h->trsv_epoch == 0) ++h->trsv_epoch;  // 0 is the zeroed array's tag
        hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, st, ticket, 1);
        IR_TRY(hipGetLastError());
        if (pass == 0)
            hipLaunchKernelGGL((trsv_lookback_kernel<T, true>), dim3(nblk), dim3(256), 0, st, LU, lda, n, InvL, v,
                               ticket, tagged, h->trsv_epoch, err);
        else
            hipLaunchKernelGGL((trsv_lookback_kernel<T, false>), dim3(nblk), dim3(256), 0, st, LU, lda, n, InvU, v,
                               ticket, tagged, h->trsv_epoch, err);
        IR_TRY(hipGetLastError());
    }
    return NROCM_OK;
}

// Flexible GMRES(GM_RESTART) in FP64 on A x = b, right-preconditioned by the
// working-precision factors (M^{-1} v = U^{-1} L^{-1} P v): GMRES-based iterative
// refinement (Carson & Higham 2017).  Classical Gram-Schmidt applied twice;
// the Hessenberg matrix and its Givens rotations live on the host (one small
// copy per iteration).  After every cycle x is updated and the true residual
// decides convergence with the same normwise backward error as classic
// refinement; `iters` counts preconditioner solves.
template <typename T>
int fgmres(nrocm_handle h, int n, const double *A, int lda, const double *b, double *x, const T *W, T *d,
           const int *perm, const T *InvL, const T *InvU, double *r,
           double *part, int *sync, unsigned long long *tagged, int *err, unsigned long long *red, const int *info,
           double *Vk, double *Zk, double *hdev, double tol, int maxit, int *iters, double *berr_out)
{
    hipStream_t st = h->stream;
    constexpr int M = GM_RESTART;
    double *h1 = hdev, *h2 = hdev + (M + 2), *hn = hdev + 2 * (M + 2), *ydev = hdev + 3 * (M + 2);
    const unsigned nb = nrocm_cdiv(n, 256);
    int solves = 0;
    double berr = INFINITY;
    double nrmA = 0.0, nrmB = 0.0;
    int s = NROCM_OK;
    // H column-major (M+1) x M on the host, Givens (c, s), g
    std::vector<double> H((size_t)(M + 1) * M), cs(M), sn(M), g(M + 1), hbuf(3 * (M + 2)), y(M);
    for (;;) {
        // true residual r = b - A x and the backward error (x max recomputed)
        hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, red + 2, 2);
        hipLaunchKernelGGL(absmax_kernel, dim3(nb), dim3(256), 0, st, x, n, red + 3);
        hipLaunchKernelGGL((gemv_part_kernel<false>), dim3(nrocm_cdiv(n, 64), GEMV_CS), dim3(256), 0, st, A, lda, n, x,
                           part);
        hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(256), 0, st, b, part, n, r, red + 2);
        hipLaunchKernelGGL(dots_kernel, dim3(1), dim3(256), 0, st, r, n, r, n, hn);
        IR_TRY(hipGetLastError());
        unsigned long long hred[4];
        int hinfo[2];
        double beta2 = 0.0;
        IR_TRY(hipMemcpyAsync(hred, red, sizeof(hred), hipMemcpyDeviceToHost, st));
        IR_TRY(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, st));
        IR_TRY(hipMemcpyAsync(&beta2, hn, sizeof(double), hipMemcpyDeviceToHost, st));
        IR_TRY(hipStreamSynchronize(st));
        if (hinfo[1]) return nrocm_fail(h, NROCM_ERR_HIP, "triangular solve wait timed out");
        double v[4];
        for (int q = 0; q < 4; ++q) {
            long long bits = (long long)hred[q];
            std::memcpy(&v[q], &bits, sizeof(double));
        }
        nrmA = v[0];
        nrmB = v[1];
        berr = v[2] / (nrmA * v[3] + nrmB);
        if (!(berr > tol) || solves >= maxit || hinfo[0] != 0) break;
        const double beta = std::sqrt(beta2);
        if (!(beta > 0.0)) break;
        // V_0 = r / beta
        hipLaunchKernelGGL(scale_by_norm_kernel, dim3(nb), dim3(256), 0, st, r, hn, n, Vk);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        // stop a cycle once the estimated |r|_2 (>= |r|_inf) meets the bound
        const double target = tol * (nrmA * v[3] + nrmB);
        int k = 0;
        for (int j = 0; j < M && solves < maxit; ++j) {
            double *vj = Vk + (size_t)j * n, *zj = Zk + (size_t)j * n, *w = Vk + (size_t)(j + 1) * n;
            // z_j = M^{-1} v_j
            hipLaunchKernelGGL((gather_kernel<T>), dim3(nb), dim3(256), 0, st, vj, perm, d, n);
            IR_TRY(hipGetLastError());
            s = trsv_pair<T>(h, st, W, n, n, InvL, InvU, d, sync, tagged, err);
            if (s != NROCM_OK) return s;
            hipLaunchKernelGGL((widen_kernel<T>), dim3(nb), dim3(256), 0, st, d, n, zj);
            ++solves;
            // w = A z_j, orthogonalised twice against v_0..v_j
            hipLaunchKernelGGL((gemv_part_kernel<false>), dim3(nrocm_cdiv(n, 64), GEMV_CS), dim3(256), 0, st, A, lda,
                               n, zj, part);
            hipLaunchKernelGGL(gemv_combine_kernel, dim3(nb), dim3(256), 0, st, part, n, w);
            hipLaunchKernelGGL(dots_kernel, dim3(j + 1), dim3(256), 0, st, Vk, n, w, n, h1);
            hipLaunchKernelGGL(sub_vh_kernel, dim3(nb), dim3(256), 0, st, w, Vk, n, j + 1, h1, n);
            hipLaunchKernelGGL(dots_kernel, dim3(j + 1), dim3(256), 0, st, Vk, n, w, n, h2);
            hipLaunchKernelGGL(sub_vh_kernel, dim3(nb), dim3(256), 0, st, w, Vk, n, j + 1, h2, n);
            hipLaunchKernelGGL(dots_kernel, dim3(1), dim3(256), 0, st, w, n, w, n, hn);
            hipLaunchKernelGGL(scale_by_norm_kernel, dim3(nb), dim3(256), 0, st, w, hn, n, w);
            IR_TRY(hipGetLastError());
            IR_TRY(hipMemcpyAsync(hbuf.data(), hdev, 3 * (M + 2) * sizeof(double), hipMemcpyDeviceToHost, st));
            IR_TRY(hipStreamSynchronize(st));
            double *col = H.data() + (size_t)j * (M + 1);
            for (int i = 0; i <= j; ++i) col[i] = hbuf[i] + hbuf[(M + 2) + i];
            col[j + 1] = std::sqrt(hbuf[2 * (M + 2)]);
            for (int i = 0; i < j; ++i) {  // previous rotations
                const double t0 = cs[i] * col[i] + sn[i] * col[i + 1];
                col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1];
                col[i] = t0;
            }
            const double den = std::hypot(col[j], col[j + 1]);
            cs[j] = den > 0.0 ? col[j] / den : 1.0;
            sn[j] = den > 0.0 ? col[j + 1] / den : 0.0;
            col[j] = den;
            col[j + 1] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            k = j + 1;
            if (std::fabs(g[j + 1]) <= target || !(hbuf[2 * (M + 2)] > 0.0)) break;
        }
        if (k == 0) break;
        // y = H(0:k, 0:k)^{-1} g(0:k); x += Z y
        for (int i = k - 1; i >= 0; --i) {
            double t0 = g[i];
            for (int c = i + 1; c < k; ++c) t0 -= H[(size_t)c * (M + 1) + i] * y[c];
            y[i] = t0 / H[(size_t)i * (M + 1) + i];
        }
        IR_TRY(hipMemcpyAsync(ydev, y.data(), k * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, red + 3, 1);
        hipLaunchKernelGGL(xupdate_kernel, dim3(nb), dim3(256), 0, st, x, Zk, n, k, ydev, n, red + 3);
        IR_TRY(hipGetLastError());
        IR_TRY(hipStreamSynchronize(st));  // y lives on the host stack frame
    }
    if (iters) *iters = solves;
    if (berr_out) *berr_out = berr;
    int hinfo0 = 0;
    IR_TRY(hipMemcpy(&hinfo0, info, sizeof(int), hipMemcpyDeviceToHost));
    if (hinfo0 != 0) return nrocm_fail(h, NROCM_ERR_NOT_CONVERGED, "singular factorization");
    return berr <= tol ? NROCM_OK : NROCM_ERR_NOT_CONVERGED;
}

// Status a windowed attempt returns when its factors are unusable: a zero
// pivot inside a window, a non-finite backward error, or refinement that
// stops contracting.  The caller then factors again with partial pivoting.
constexpr int IR_RETRY_PARTIAL = -1000;

template <typename T>
int gesv_ir(nrocm_handle h, int n, const double *A, int lda, const double *b, double *x, NrocmTrailing trail,
            bool window, bool gmres, double tol, int maxit, int *iters, double *berr_out)
{
    hipStream_t st = h->stream;
    const int nblk = (n + BS - 1) / BS;
    // workspace layout
    const size_t szW = (size_t)n * n * sizeof(T);
    const size_t szVec = (size_t)n * sizeof(double);
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    const size_t oW = take(szW), oIpiv = take(n * sizeof(int)), oPerm = take(n * sizeof(int)),
                 oR = take(szVec), oD = take((size_t)n * sizeof(T)), oPart = take(GEMV_CS * szVec),
                 oSync = take((nblk + 2) * sizeof(int)), oRed = take(8 * sizeof(unsigned long long)),
                 oTag = take((size_t)n * (sizeof(T) / 4) * sizeof(unsigned long long)),
                 oInfo = take(2 * sizeof(int)), oInv = take(2 * (size_t)nblk * BS * BS * sizeof(T)),
                 oKry = take(gmres ? (size_t)(2 * GM_RESTART + 1) * szVec + 4 * (GM_RESTART + 2) * sizeof(double) : 0);
    void *ws = nullptr;
    int s = nrocm_workspace(h, off, &ws);
    if (s != NROCM_OK) return s;
    char *base = (char *)ws;
    T *W = (T *)(base + oW);
    int *ipiv = (int *)(base + oIpiv), *perm = (int *)(base + oPerm);
    double *r = (double *)(base + oR), *part = (double *)(base + oPart);
    T *d = (T *)(base + oD);
    int *sync = (int *)(base + oSync);
    unsigned long long *tagged = (unsigned long long *)(base + oTag);
    const size_t ntag = (size_t)n * (sizeof(T) / 4);
    hipLaunchKernelGGL(zero_words_kernel, dim3(nrocm_cdiv(ntag, 256)), dim3(256), 0, st, tagged, ntag);
    IR_TRY(hipGetLastError());
    unsigned long long *red = (unsigned long long *)(base + oRed);  // [0]=|A|, [1]=|b|, [2]=|r|, [3]=|x|
    int *info = (int *)(base + oInfo), *err = info + 1;
    T *InvL = (T *)(base + oInv), *InvU = InvL + (size_t)nblk * BS * BS;
    double *Vk = (double *)(base + oKry), *Zk = Vk + (size_t)(GM_RESTART + 1) * n;  // Krylov basis, preconditioned
    double *hdev = Zk + (size_t)GM_RESTART * n;  // h1 | h2 | norm | y, GM_RESTART + 2 each

    const size_t tot = (size_t)n * n;
    if constexpr (sizeof(T) == 4) {
        hipLaunchKernelGGL(cvt_d2s_kernel, dim3(nrocm_cdiv(tot, 256)), dim3(256), 0, st, A, lda, W, n, n);
    } else {
        hipLaunchKernelGGL(copy_d_kernel, dim3(nrocm_cdiv(tot, 256)), dim3(256), 0, st, A, lda, (double *)W, n, n);
    }
    IR_TRY(hipGetLastError());
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, st, info, 2);
    IR_TRY(hipGetLastError());
    int fs;
    if constexpr (sizeof(T) == 4)
        fs = window ? nrocm_getrf_window_s(h, n, (float *)W, n, ipiv, perm, info, trail)
                    : nrocm_getrf_s(h, st, n, n, (float *)W, n, ipiv, info, trail);
    else
        fs = window ? nrocm_getrf_window_d(h, n, (double *)W, n, ipiv, perm, info, trail)
                    : nrocm_getrf_d(h, st, n, n, (double *)W, n, ipiv, info, trail);
    if (fs != NROCM_OK) return fs;
    if (!window) {
        hipLaunchKernelGGL(perm_kernel, dim3(1), dim3(1024), n * sizeof(int), st, ipiv, n, perm);
        IR_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL((inv_diag_kernel<T, true>), dim3(nblk), dim3(64), 0, st, W, n, n, InvL);
    IR_TRY(hipGetLastError());
    hipLaunchKernelGGL((inv_diag_kernel<T, false>), dim3(nblk), dim3(64), 0, st, W, n, n, InvU);
    IR_TRY(hipGetLastError());
    // norms of A (row sums) and b
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, red, 8);
    hipLaunchKernelGGL((gemv_part_kernel<true>), dim3(nrocm_cdiv(n, 64), GEMV_CS), dim3(256), 0, st, A, lda, n,
                       nullptr, part);
    hipLaunchKernelGGL(rowsum_max_kernel, dim3(nrocm_cdiv(n, 256)), dim3(256), 0, st, part, n, red + 0);
    hipLaunchKernelGGL(absmax_kernel, dim3(nrocm_cdiv(n, 256)), dim3(256), 0, st, b, n, red + 1);
    hipLaunchKernelGGL(zero_d_kernel, dim3(nrocm_cdiv(n, 256)), dim3(256), 0, st, x, n);
    IR_TRY(hipGetLastError());
    IR_TRY(hipMemcpyAsync(r, b, szVec, hipMemcpyDeviceToDevice, st));

    double berr = INFINITY;
    int it = 0;
    unsigned long long hred[4];
    int hinfo[2];
    if (gmres) {
        s = fgmres(h, n, A, lda, b, x, W, d, perm, InvL, InvU, r, part, sync, tagged, err, red, info, Vk, Zk, hdev,
                   tol, maxit, iters, berr_out);
        return (window && s == NROCM_ERR_NOT_CONVERGED) ? IR_RETRY_PARTIAL : s;
    }
    double prev = INFINITY;
    for (it = 1; it <= maxit; ++it) {
        hipLaunchKernelGGL((gather_kernel<T>), dim3(nrocm_cdiv(n, 256)), dim3(256), 0, st, r, perm, d, n);
        IR_TRY(hipGetLastError());
        s = trsv_pair<T>(h, st, W, n, n, InvL, InvU, d, sync, tagged, err);
        if (s != NROCM_OK) return s;
        hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, red + 2, 2);
        hipLaunchKernelGGL((update_x_kernel<T>), dim3(nrocm_cdiv(n, 256)), dim3(256), 0, st, x, d, n, red + 3);
        hipLaunchKernelGGL((gemv_part_kernel<false>), dim3(nrocm_cdiv(n, 64), GEMV_CS), dim3(256), 0, st, A, lda, n, x,
                           part);
        hipLaunchKernelGGL(residual_kernel, dim3(nrocm_cdiv(n, 256)), dim3(256), 0, st, b, part, n, r, red + 2);
        IR_TRY(hipGetLastError());
        IR_TRY(hipMemcpyAsync(hred, red, sizeof(hred), hipMemcpyDeviceToHost, st));
        IR_TRY(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, st));
        IR_TRY(hipStreamSynchronize(st));
        if (hinfo[1]) return nrocm_fail(h, NROCM_ERR_HIP, "triangular solve wait timed out");
        double v[4];
        for (int q = 0; q < 4; ++q) {
            long long bits = (long long)hred[q];
            std::memcpy(&v[q], &bits, sizeof(double));
        }
        berr = v[2] / (v[0] * v[3] + v[1]);
        // windowed factors that hit a zero pivot, give a non-finite solution or
        // stop contracting the error (a step that does not halve it, from the
        // third step on) are replaced by partial-pivoting factors
        if (window && (hinfo[0] != 0 || !std::isfinite(berr) || (it >= 3 && berr > tol && berr > 0.5 * prev))) {
            if (iters) *iters = it;
            return IR_RETRY_PARTIAL;
        }
        if (!(berr > tol)) break;  // also stops on NaN
        prev = berr;
    }
    if (it > maxit) it = maxit;
    if (iters) *iters = it;
    if (window && berr > tol) return IR_RETRY_PARTIAL;
    if (berr_out) *berr_out = berr;
    if (hinfo[0] != 0) return nrocm_fail(h, NROCM_ERR_NOT_CONVERGED, "singular factorization");
    return berr <= tol ? NROCM_OK : NROCM_ERR_NOT_CONVERGED;
}

}  // namespace

extern "C" {

int nrocm_dgesv_ir(nrocm_handle h, int n, const double *a, int lda, const double *b, double *x, int precision,
                   double tol, int maxit, int *iters_host, double *berr_host)
{
    if (!h) return -1;
    if (n < 0) return -2;
    if (!a && n > 0) return -3;
    if (lda < std::max(1, n)) return -4;
    if (!b && n > 0) return -5;
    if (!x && n > 0) return -6;
    const bool gmres = (precision & NROCM_PREC_GMRES) != 0;
    const bool window = (precision & NROCM_PREC_WINDOW) != 0;
    precision &= ~(NROCM_PREC_GMRES | NROCM_PREC_WINDOW);
    if (precision < NROCM_PREC_FP64 || precision > NROCM_PREC_BF16X3) return -7;
    if (!(tol >= 0.0)) return -8;
    if (maxit < 1) return -9;
    if (n == 0) {
        if (iters_host) *iters_host = 0;
        if (berr_host) *berr_host = 0.0;
        return NROCM_OK;
    }
    const NrocmTrailing tr = precision == NROCM_PREC_FP64     ? TRAIL_F64
                             : precision == NROCM_PREC_BF16   ? TRAIL_BF16
                             : precision == NROCM_PREC_BF16X3 ? TRAIL_BF16X3
                                                              : TRAIL_F32;
    auto attempt = [&](bool win, int *it_out, double *be_out) {
        return precision == NROCM_PREC_FP64
                   ? gesv_ir<double>(h, n, a, lda, b, x, tr, win, gmres, tol, maxit, it_out, be_out)
                   : gesv_ir<float>(h, n, a, lda, b, x, tr, win, gmres, tol, maxit, it_out, be_out);
    };
    h->ir_fallback = 0;
    int it1 = 0;
    double be1 = INFINITY;
    int s = attempt(window, &it1, &be1);
    if (s == IR_RETRY_PARTIAL) {
        // the windowed pivots were not good enough for this matrix: factor
        // again with partial pivoting over all remaining rows
        h->ir_fallback = 1;
        int it2 = 0;
        s = attempt(false, &it2, &be1);
        it1 += it2;
    }
    if (iters_host) *iters_host = it1;
    if (berr_host) *berr_host = be1;
    return s;
}

int nrocm_ir_fallback(nrocm_handle h, int *fell_back_host)
{
    if (!h) return -1;
    if (!fell_back_host) return -2;
    *fell_back_host = h->ir_fallback;
    return NROCM_OK;
}

}  // extern "C"
