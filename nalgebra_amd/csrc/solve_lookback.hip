// LU::solve_mut (src/linalg/lu.rs:242-260) for any n: the row swaps, the
// unit-lower forward substitution (solve.rs:106-133) and the upper backward
// substitution (solve.rs:137-182) in nalgebra's per-element operation order,
// spread over one 64-row block per workgroup.
//
// Both substitutions are column sweeps: element r receives
// b_r = fl(b_r - fl(x_i * M_ri)) for i = 0, 1, ..., r-1 (forward) or
// i = n-1, ..., r+1 (backward), one step per solved x_i.  A row block can
// apply the steps of a finished column block as soon as that block is
// published, in the sweep's order, so the blocks form a pipeline: blocks take
// tickets in sweep order, block k waits for blocks 0..k-1 one at a time, applies
// their 64 steps, solves its own 64 x 64 diagonal block and publishes it.
//
// A published value sits in two 64-bit words, each holding a 32-bit half of
// the double next to the launch's epoch, written as one atomic word; a reader
// polls the words until they carry the epoch, so data and readiness arrive
// together.  The tag array is cleared at the start of every call and every
// launch of a call has its own epoch.  Waits are bounded: a wait that runs out
// sets info to -1 and the kernel still exits.
#include "nrocm_internal.hpp"

#include <algorithm>
#include <mutex>

namespace {

constexpr int LB = 64;                    // rows per block (one wavefront)
constexpr int PCH = 16;                   // matrix elements loaded ahead per lane
constexpr long LB_SPIN_LIMIT = 1L << 26;  // ~2 s of polling
constexpr int PERM_LDS_MAX = 16384;       // composed permutation kept in LDS up to this n

template <bool EXACT>
__device__ __forceinline__ double sub_prod(double acc, double x, double y)
{
    // nalgebra's axpy with factor -x: fl(fl((-x) * y) + acc) (the negation on
    // x before the product, so a NaN coming from y keeps its sign); FAST fuses it
    if constexpr (EXACT) {
        double r;  // operand order pinned: NaN results carry the reference's sign
        asm("v_mul_f64 %0, -%1, %2\n\tv_add_f64 %0, %0, %3" : "=&v"(r) : "v"(x), "v"(y), "v"(acc));
        return r;
    } else {
        return fma(-x, y, acc);
    }
}

__device__ __forceinline__ void publish(unsigned long long *tags, int row, unsigned epoch, double v)
{
    const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
    __hip_atomic_store(tags + 2 * (size_t)row, ((unsigned long long)epoch << 32) | (bits & 0xffffffffull),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(tags + 2 * (size_t)row + 1, ((unsigned long long)epoch << 32) | (bits >> 32),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double await(const unsigned long long *tags, int row, unsigned epoch, int *info)
{
    unsigned long long w[2];
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
        const unsigned long long *src = tags + 2 * (size_t)row + hh;
        long spins = 0;
        unsigned long long word;
        while (((word = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 32) != epoch) {
            if (++spins > LB_SPIN_LIMIT) {
                atomicExch(info, -1);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        w[hh] = word;
    }
    return __longlong_as_double((long long)(((w[1] & 0xffffffffull) << 32) | (w[0] & 0xffffffffull)));
}

__device__ __forceinline__ double lane_value(double v, int src)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)bits, src);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Applies the steps of one finished column block [c0, c0 + cols) to this
// lane's element: acc = fl(acc - fl(x_i * M[row, c0 + i])) for i ascending
// (FORWARD) or descending, x from LDS, M streamed PCH columns ahead.
template <bool EXACT, bool FORWARD>
__device__ __forceinline__ double apply_block(double acc, const double *__restrict__ Mrow, int lda, int c0, int cols,
                                              const double *xs, bool row_ok)
{
    double cur[PCH], nxt[PCH];
    auto col_of = [&](int s) { return FORWARD ? s : cols - 1 - s; };
#pragma unroll
    for (int t = 0; t < PCH; ++t)
        cur[t] = (row_ok && t < cols) ? Mrow[(size_t)(c0 + col_of(t)) * lda] : 0.0;
    for (int s0 = 0; s0 < cols; s0 += PCH) {
#pragma unroll
        for (int t = 0; t < PCH; ++t) {
            const int s = s0 + PCH + t;
            nxt[t] = (row_ok && s < cols) ? Mrow[(size_t)(c0 + col_of(s)) * lda] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < PCH; ++t) {
            const int s = s0 + t;
            if (s < cols) acc = sub_prod<EXACT>(acc, xs[col_of(s)], cur[t]);
        }
#pragma unroll
        for (int t = 0; t < PCH; ++t) cur[t] = nxt[t];
    }
    return acc;
}

// Forward substitution L y = P b (unit diagonal): block k (ticket order) holds
// rows [64k, 64k + 64); its right-hand side is b at the rows perm names.
template <bool EXACT>
__global__ __launch_bounds__(LB) void lu_forward_lookback_kernel(const double *__restrict__ LU, int lda, int n,
                                                                 const int *__restrict__ perm,
                                                                 const double *__restrict__ b, int *ticket,
                                                                 unsigned long long *tags, unsigned epoch, int *info)
{
    __shared__ double xs[LB];
    __shared__ int my;
    const int lane = threadIdx.x;
    if (lane == 0) my = atomicAdd(ticket, 1);
    __syncthreads();
    const int R = my;
    const int r0 = R * LB;
    const int rows = min(LB, n - r0);
    const bool row_ok = lane < rows;
    const int r = r0 + lane;
    const double *Mrow = LU + (row_ok ? r : r0);
    double acc = row_ok ? b[perm[r]] : 0.0;
    for (int J = 0; J < R; ++J) {
        const int c0 = J * LB;  // earlier blocks are full
        xs[lane] = await(tags, c0 + lane, epoch, info);
        __syncthreads();
        acc = apply_block<EXACT, true>(acc, Mrow, lda, c0, LB, xs, row_ok);
        __syncthreads();
    }
    // own diagonal block: x_i = b_i (division by the unit diagonal is exact)
    for (int i = 0; i + 1 < rows; ++i) {
        const double xi = lane_value(acc, i);
        const double l = (lane > i && row_ok) ? Mrow[(size_t)(r0 + i) * lda] : 0.0;
        if (lane > i) acc = sub_prod<EXACT>(acc, xi, l);
    }
    if (row_ok) publish(tags, r, epoch, acc);
}

// Backward substitution U x = y: block k (ticket order) holds rows
// [64(N-1-k), ...); y comes from the forward launch (tags, epoch_in), x is
// published with epoch_out and written to b.
template <bool EXACT>
__global__ __launch_bounds__(LB) void lu_backward_lookback_kernel(const double *__restrict__ LU, int lda, int n,
                                                                  double *__restrict__ b, int *ticket,
                                                                  unsigned long long *tags, unsigned epoch_in,
                                                                  unsigned epoch_out, int *info)
{
    __shared__ double xs[LB];
    __shared__ int my;
    const int lane = threadIdx.x;
    const int nblk = (n + LB - 1) / LB;
    if (lane == 0) my = atomicAdd(ticket, 1);
    __syncthreads();
    const int k = my;
    const int R = nblk - 1 - k;
    const int r0 = R * LB;
    const int rows = min(LB, n - r0);
    const bool row_ok = lane < rows;
    const int r = r0 + lane;
    const double *Mrow = LU + (row_ok ? r : r0);
    double acc = row_ok ? await(tags, r, epoch_in, info) : 0.0;
    for (int j = 0; j < k; ++j) {
        const int J = nblk - 1 - j;
        const int c0 = J * LB;
        const int cols = min(LB, n - c0);
        xs[lane] = lane < cols ? await(tags, c0 + lane, epoch_out, info) : 0.0;
        __syncthreads();
        acc = apply_block<EXACT, false>(acc, Mrow, lda, c0, cols, xs, row_ok);
        __syncthreads();
    }
    // own diagonal block: x_i = b_i / U_ii, then the rows above it
    for (int i = rows - 1; i >= 0; --i) {
        if (lane == i) acc = acc / Mrow[(size_t)(r0 + i) * lda];
        const double xi = lane_value(acc, i);
        const double u = (lane < i) ? Mrow[(size_t)(r0 + i) * lda] : 0.0;
        if (lane < i) acc = sub_prod<EXACT>(acc, xi, u);
    }
    if (row_ok) {
        publish(tags, r, epoch_out, acc);
        b[r] = acc;
    }
}

// ---------------------------------------------------------------- Cholesky::solve_mut
// Cholesky::solve_mut (src/linalg/cholesky.rs:122-129): L y = b by the column
// sweep of solve_lower_triangular_unchecked_mut (solve.rs:500-519,
// x_i = b_i / L_ii), then L^T x = y by ad_solve_lower_triangular_unchecked_mut
// (solve.rs:697-755), whose row i is x_i = (y_i - dot(L[i+1.., i], x[i+1..])) / L_ii.

// Forward sweep, block k (ticket order) = rows [64k, 64k + 64); y is
// published with `epoch` and written over b.
template <bool EXACT>
__global__ __launch_bounds__(LB) void chol_forward_lookback_kernel(const double *__restrict__ L, int lda, int n,
                                                                   double *__restrict__ b, int *ticket,
                                                                   unsigned long long *tags, unsigned epoch,
                                                                   int *info)
{
    __shared__ double xs[LB];
    __shared__ int my;
    const int lane = threadIdx.x;
    if (lane == 0) my = atomicAdd(ticket, 1);
    __syncthreads();
    const int R = my;
    const int r0 = R * LB;
    const int rows = min(LB, n - r0);
    const bool row_ok = lane < rows;
    const int r = r0 + lane;
    const double *Mrow = L + (row_ok ? r : r0);
    double acc = row_ok ? b[r] : 0.0;
    for (int J = 0; J < R; ++J) {
        const int c0 = J * LB;
        xs[lane] = await(tags, c0 + lane, epoch, info);
        __syncthreads();
        acc = apply_block<EXACT, true>(acc, Mrow, lda, c0, LB, xs, row_ok);
        __syncthreads();
    }
    for (int i = 0; i < rows; ++i) {
        if (lane == i) acc = acc / Mrow[(size_t)(r0 + i) * lda];
        const double xi = lane_value(acc, i);
        const double l = (lane > i && row_ok) ? Mrow[(size_t)(r0 + i) * lda] : 0.0;
        if (lane > i) acc = sub_prod<EXACT>(acc, xi, l);
    }
    if (row_ok) {
        publish(tags, r, epoch, acc);
        b[r] = acc;
    }
}

// NROCM_FAST transposed solve as a backward column sweep over U = L^T (any
// summation order is allowed): block k (ticket order) = rows
// [64(N-1-k), ...); lane r reads column r of L, U[r][c] = L[c + r lda].
__global__ __launch_bounds__(LB) void chol_backward_lookback_fast_kernel(const double *__restrict__ L, int lda, int n,
                                                                         double *__restrict__ b, int *ticket,
                                                                         unsigned long long *tags, unsigned epoch_in,
                                                                         unsigned epoch_out, int *info)
{
    __shared__ double xs[LB];
    __shared__ int my;
    const int lane = threadIdx.x;
    const int nblk = (n + LB - 1) / LB;
    if (lane == 0) my = atomicAdd(ticket, 1);
    __syncthreads();
    const int k = my;
    const int R = nblk - 1 - k;
    const int r0 = R * LB;
    const int rows = min(LB, n - r0);
    const bool row_ok = lane < rows;
    const int r = r0 + lane;
    const double *Mcol = L + (size_t)(row_ok ? r : r0) * lda;
    double acc = row_ok ? await(tags, r, epoch_in, info) : 0.0;
    for (int j = 0; j < k; ++j) {
        const int c0 = (nblk - 1 - j) * LB;
        const int cols = min(LB, n - c0);
        xs[lane] = lane < cols ? await(tags, c0 + lane, epoch_out, info) : 0.0;
        __syncthreads();
        double cur[PCH], nxt[PCH];
#pragma unroll
        for (int t = 0; t < PCH; ++t) cur[t] = (row_ok && t < cols) ? Mcol[c0 + t] : 0.0;
        for (int s0 = 0; s0 < cols; s0 += PCH) {
#pragma unroll
            for (int t = 0; t < PCH; ++t) {
                const int s = s0 + PCH + t;
                nxt[t] = (row_ok && s < cols) ? Mcol[c0 + s] : 0.0;
            }
#pragma unroll
            for (int t = 0; t < PCH; ++t)
                if (s0 + t < cols) acc = fma(-xs[s0 + t], cur[t], acc);
#pragma unroll
            for (int t = 0; t < PCH; ++t) cur[t] = nxt[t];
        }
        __syncthreads();
    }
    for (int i = rows - 1; i >= 0; --i) {
        if (lane == i) acc = acc / Mcol[r0 + i];
        const double xi = lane_value(acc, i);
        const double u = (lane < i) ? Mcol[r0 + i] : 0.0;
        if (lane < i) acc = fma(-xi, u, acc);
    }
    if (row_ok) {
        publish(tags, r, epoch_out, acc);
        b[r] = acc;
    }
}

// NROCM_EXACT transposed solve in nalgebra's dot order (the 8-accumulator dot
// of src/base/blas.rs: accumulator q sums the products of elements q, q + 8,
// ... of the full chunks in order, the sum is
// ((((0 + (a0 + a4)) + (a1 + a5)) + (a2 + a6)) + (a3 + a7)) and the remaining
// elements are added one by one).  Row i's first product needs x_{i+1}, so
// the rows form one dependent chain: one wavefront, lanes 0..7 the
// accumulators, x kept in LDS (XS_LDS) or, beyond the LDS, in b itself (each
// new x_i is then released to the L2 and the vector L1 dropped before the
// next row reads it: other lanes read values lane 0 wrote in this launch).
template <bool XS_LDS>
__global__ __launch_bounds__(64) void chol_backward_dot_kernel(const double *__restrict__ L, int lda, int n,
                                                               double *b)
{
    extern __shared__ double xs_dyn[];
    const int lane = threadIdx.x;
    double *xb = XS_LDS ? xs_dyn : b;
    if constexpr (XS_LDS) {
        for (int r = lane; r < n; r += 64) xs_dyn[r] = b[r];
        __syncthreads();
    }
    auto xload = [&](int idx) -> double { return xb[idx]; };
    for (int i = n - 1; i >= 0; --i) {
        const int len = n - i - 1;
        const int full = len & ~7;
        const double *li = L + (i + 1) + (size_t)i * lda;
        double a = 0.0;
        if (lane < 8) {
#pragma unroll 4
            for (int e = lane; e < full; e += 8) a += li[e] * xload(i + 1 + e);
        }
        double res = 0.0;
        res += lane_value(a, 0) + lane_value(a, 4);
        res += lane_value(a, 1) + lane_value(a, 5);
        res += lane_value(a, 2) + lane_value(a, 6);
        res += lane_value(a, 3) + lane_value(a, 7);
        for (int e = full; e < len; ++e) res += li[e] * xload(i + 1 + e);
        const double xi = (xload(i) - res) / L[i + (size_t)i * lda];
        __syncthreads();  // every lane has read x_i's old value
        if (lane == 0) xb[i] = xi;
        if constexpr (!XS_LDS) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if constexpr (!XS_LDS) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    if constexpr (XS_LDS) {
        for (int r = lane; r < n; r += 64) b[r] = xs_dyn[r];
    }
}

// info = 1 + the largest i with U_ii == 0 (the backward sweep meets it first
// and solve_mut returns false), or 0.
__global__ void zero_diag_kernel(const double *__restrict__ LU, int lda, int n, int *info)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && LU[i + (size_t)i * lda] == 0.0) atomicMax(info, i + 1);
}

__global__ void fill_words_kernel(unsigned long long *p, size_t count, int *ints, int nints)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
    if (i < (size_t)nints) ints[i] = 0;
}

__global__ void reset_ticket_kernel(int *ticket) { *ticket = 0; }

// perm[i] = the row of b that the swaps (i, ipiv[i]), applied in order, bring
// to position i (PermutationSequence::permute_rows, permutation_sequence.rs:97-104).
// The swaps are sequential; one thread applies them to an index array in LDS
// (n <= PERM_LDS_MAX) or in perm itself.
// A pivot outside [0, n) (nalgebra's PermutationSequence cannot hold one) is
// not applied and sets info to -2.
__global__ __launch_bounds__(1024) void compose_perm_kernel(const int *__restrict__ ipiv, int n, int *perm,
                                                            int *info)
{
    extern __shared__ int sp[];
    const bool lds = n <= PERM_LDS_MAX;
    int *p = lds ? sp : perm;
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = i;
    __syncthreads();
    if (threadIdx.x == 0) {
        int nxt = n > 0 ? ipiv[0] : 0;
        for (int i = 0; i < n; ++i) {
            const int q = nxt;
            if (i + 1 < n) nxt = ipiv[i + 1];
            if (q < 0 || q >= n) {
                atomicExch(info, -2);
            } else if (q != i) {
                const int t = p[i];
                p[i] = p[q];
                p[q] = t;
            }
        }
    }
    __syncthreads();
    if (lds)
        for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = p[i];
}

}  // namespace

// Scratch of one solve: [ticket, spare] | perm (n ints) | tags (2n words).
size_t nrocm_getrs_scratch_bytes(int n)
{
    return 256 + (((size_t)n * sizeof(int) + 255) & ~(size_t)255) + 2 * (size_t)n * sizeof(unsigned long long);
}

// Enqueues the exact-order LU solve on st with `scratch` (at least
// nrocm_getrs_scratch_bytes(n), device memory): allocates nothing.
int nrocm_getrs_lookback(nrocm_handle h, hipStream_t st, int n, int nrhs, const double *lu, int lda, const int *ipiv,
                         double *b, int ldb, int *info, bool exact, void *scratch)
{
    char *base = (char *)scratch;
    int *ticket = (int *)base;
    int *perm = (int *)(base + 256);
    unsigned long long *tags =
        (unsigned long long *)(base + 256 + (((size_t)n * sizeof(int) + 255) & ~(size_t)255));
    const size_t ntag = 2 * (size_t)n;
    hipLaunchKernelGGL(fill_words_kernel, dim3(nrocm_cdiv(std::max<size_t>(ntag, 1), 256)), dim3(256), 0, st, tags,
                       ntag, info, 1);
    NROCM_CHECK_LAUNCH(h);
    if (n == 0) return NROCM_OK;
    hipLaunchKernelGGL(zero_diag_kernel, dim3(nrocm_cdiv(n, 256)), dim3(256), 0, st, lu, lda, n, info);
    hipLaunchKernelGGL(compose_perm_kernel, dim3(1), dim3(1024), n <= PERM_LDS_MAX ? n * sizeof(int) : 0, st, ipiv,
                       n, perm, info);
    NROCM_CHECK_LAUNCH(h);
    const unsigned nblk = nrocm_cdiv(n, LB);
    unsigned epoch = 0;
    for (int c = 0; c < nrhs; ++c) {
        double *bc = b + (size_t)c * ldb;
        const unsigned e1 = ++epoch, e2 = ++epoch;
        hipLaunchKernelGGL(reset_ticket_kernel, dim3(1), dim3(1), 0, st, ticket);
        if (exact)
            hipLaunchKernelGGL(lu_forward_lookback_kernel<true>, dim3(nblk), dim3(LB), 0, st, lu, lda, n, perm, bc,
                               ticket, tags, e1, info);
        else
            hipLaunchKernelGGL(lu_forward_lookback_kernel<false>, dim3(nblk), dim3(LB), 0, st, lu, lda, n, perm, bc,
                               ticket, tags, e1, info);
        hipLaunchKernelGGL(reset_ticket_kernel, dim3(1), dim3(1), 0, st, ticket);
        if (exact)
            hipLaunchKernelGGL(lu_backward_lookback_kernel<true>, dim3(nblk), dim3(LB), 0, st, lu, lda, n, bc,
                               ticket, tags, e1, e2, info);
        else
            hipLaunchKernelGGL(lu_backward_lookback_kernel<false>, dim3(nblk), dim3(LB), 0, st, lu, lda, n, bc,
                               ticket, tags, e1, e2, info);
        NROCM_CHECK_LAUNCH(h);
    }
    return NROCM_OK;
}

// Largest n whose x the exact transposed solve keeps in LDS (beyond it, in b).
constexpr int CHOL_DOT_LDS_MAX = 19968;

// Enqueues Cholesky::solve_mut on st with `scratch` (at least
// nrocm_getrs_scratch_bytes(n), device memory): the forward sweep over one
// 64-row block per workgroup, then the transposed solve (EXACT: one
// wavefront in dot order; FAST: the backward look-back sweep over L^T).
int nrocm_potrs_lookback(nrocm_handle h, hipStream_t st, int n, int nrhs, const double *L, int lda, double *b,
                         int ldb, bool exact, void *scratch)
{
    char *base = (char *)scratch;
    int *ticket = (int *)base;
    int *timeout = ticket + 1;  // set by a look-back wait that ran out (never expected)
    unsigned long long *tags =
        (unsigned long long *)(base + 256 + (((size_t)n * sizeof(int) + 255) & ~(size_t)255));
    const size_t ntag = 2 * (size_t)n;
    hipLaunchKernelGGL(fill_words_kernel, dim3(nrocm_cdiv(std::max<size_t>(ntag, 1), 256)), dim3(256), 0, st, tags,
                       ntag, timeout, 1);
    NROCM_CHECK_LAUNCH(h);
    const unsigned nblk = nrocm_cdiv(n, LB);
    const bool lds = n <= CHOL_DOT_LDS_MAX;
    if (exact && lds) {
        static std::once_flag once;
        std::call_once(once, [] {
            (void)hipFuncSetAttribute((const void *)chol_backward_dot_kernel<true>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      CHOL_DOT_LDS_MAX * (int)sizeof(double));
        });
    }
    unsigned epoch = 0;
    for (int c = 0; c < nrhs; ++c) {
        double *bc = b + (size_t)c * ldb;
        const unsigned e1 = ++epoch, e2 = ++epoch;
        hipLaunchKernelGGL(reset_ticket_kernel, dim3(1), dim3(1), 0, st, ticket);
        if (exact)
            hipLaunchKernelGGL(chol_forward_lookback_kernel<true>, dim3(nblk), dim3(LB), 0, st, L, lda, n, bc, ticket,
                               tags, e1, timeout);
        else
            hipLaunchKernelGGL(chol_forward_lookback_kernel<false>, dim3(nblk), dim3(LB), 0, st, L, lda, n, bc,
                               ticket, tags, e1, timeout);
        if (exact) {
            if (lds)
                hipLaunchKernelGGL(chol_backward_dot_kernel<true>, dim3(1), dim3(64), (size_t)n * sizeof(double), st,
                                   L, lda, n, bc);
            else
                hipLaunchKernelGGL(chol_backward_dot_kernel<false>, dim3(1), dim3(64), 0, st, L, lda, n, bc);
        } else {
            hipLaunchKernelGGL(reset_ticket_kernel, dim3(1), dim3(1), 0, st, ticket);
            hipLaunchKernelGGL(chol_backward_lookback_fast_kernel, dim3(nblk), dim3(LB), 0, st, L, lda, n, bc, ticket,
                               tags, e1, e2, timeout);
        }
        NROCM_CHECK_LAUNCH(h);
    }
    return NROCM_OK;
}

// ---------------------------------------------------------------- the allocation-free solve
// nrocm_dgetrs_noalloc keeps the solve's scratch (ticket, permutation, tags)
// in this file-scope device array instead of the handle's workspace: each GPU
// the library runs on holds one copy, NROCM_NOALLOC_SCRATCH_BYTES bytes, until
// the process ends.  Host code never passes the array's name to a kernel (on
// the host it is a stand-in, not a device address): the device address is
// looked up once per device and kept in scratch_dev.
namespace {

constexpr size_t NROCM_NOALLOC_SCRATCH_BYTES =
    256 + (((size_t)NROCM_NOALLOC_MAX_N * sizeof(int) + 255) & ~(size_t)255) +
    2 * (size_t)NROCM_NOALLOC_MAX_N * sizeof(unsigned long long);

}  // namespace

static __device__ __align__(256) unsigned char scratch[NROCM_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void *scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the handle's device (looked up on the first
// call for that device, with the device current).
int scratch_ptr(nrocm_handle h, void **out)
{
    if (h->device < 0 || h->device >= kMaxDevices) return nrocm_fail(h, NROCM_ERR_UNSUPPORTED, "device index >= 64");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[h->device]) {
        int cur = 0;
        NROCM_CHECK_HIP(h, hipGetDevice(&cur));
        if (cur != h->device) NROCM_CHECK_HIP(h, hipSetDevice(h->device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != h->device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return nrocm_fail_hip(h, e, "hipGetSymbolAddress(scratch)");
        scratch_dev[h->device] = p;
    }
    *out = scratch_dev[h->device];
    return NROCM_OK;
}

}  // namespace

extern "C" int nrocm_dgetrs_noalloc(nrocm_handle h, int n, int nrhs, const double *lu, int lda, const int *ipiv,
                                    double *b, int ldb, int *info, int mode)
{
    if (!h) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (!lu && n > 0) return -4;
    if (lda < std::max(1, n)) return -5;
    if (!ipiv && n > 0) return -6;
    if (!b && n > 0 && nrhs > 0) return -7;
    if (ldb < std::max(1, n)) return -8;
    if (!info) return -9;
    if (mode != NROCM_EXACT && mode != NROCM_FAST) return -10;
    if (n > NROCM_NOALLOC_MAX_N)
        return nrocm_fail(h, NROCM_ERR_UNSUPPORTED,
                          "getrs_noalloc: n exceeds NROCM_NOALLOC_MAX_N (16384), the rows its static scratch holds");
    void *ptr = nullptr;
    const int s = scratch_ptr(h, &ptr);
    if (s != NROCM_OK) return s;
    return nrocm_getrs_lookback(h, h->stream, n, nrhs, lu, lda, ipiv, b, ldb, info, mode == NROCM_EXACT, ptr);
}
