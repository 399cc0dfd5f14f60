// Batched (uneven sizes, n <= 64) Cholesky factorisation + solve, the adjoint
// of that solve, and the single-matrix small-n path built on the same code.
//
// Layout: one wavefront serves 64/G matrices; within a group of G lanes, lane
// r holds row r of the lower triangle of its matrix in registers (a[0..NB)),
// padded to NB with the identity.  Right-looking elimination: at step k the
// finished column k is broadcast through LDS and every lane applies
// a[j] -= L_jk * L_rk to its row.  Per element this is the same ordered
// sequence of unfused operations as nalgebra's left-looking column loop
// (src/linalg/cholesky.rs:226-262), so NROCM_EXACT reproduces its bits; the
// back-substitution reproduces the 8-accumulator dotc order
// (src/base/blas.rs:93-167) with the products exchanged through LDS.
// NROCM_FAST uses fused multiply-adds and a butterfly reduction instead.
//
// Identity padding does not change the leading n x n results: padded rows
// and columns only ever meet zeros, and the dot lengths use the true n.
#include "chol_kernels.hpp"

#include <algorithm>
#include <numeric>

namespace {

// ------------------------------------------------------------------ dispatch tables

using FacLaunch = int (*)(bool, unsigned, hipStream_t, const int *, int, int, const int *,
                         double *const *, int *, double *, int, int);
using SolLaunch = int (*)(bool, unsigned, hipStream_t, const int *, int, int, const int *,
                         double *const *, double *const *, int, const int *, double *, double *,
                         int, int, int);
using BwdLaunch = int (*)(unsigned, hipStream_t, const int *, int, int, const int *,
                         const double *const *, const double *const *, const double *const *,
                         double *const *, double *const *, int);
using FusedLaunch = int (*)(bool, unsigned, hipStream_t, const int *, int, int, const int *,
                           double *const *, double *const *, int, int *);

const FacLaunch kFac[8] = {nrocm_chol_potrf_launch_8,  nrocm_chol_potrf_launch_16,
                           nrocm_chol_potrf_launch_24, nrocm_chol_potrf_launch_32,
                           nrocm_chol_potrf_launch_40, nrocm_chol_potrf_launch_48,
                           nrocm_chol_potrf_launch_56, nrocm_chol_potrf_launch_64};
const SolLaunch kSol[8] = {nrocm_chol_potrs_launch_8,  nrocm_chol_potrs_launch_16,
                           nrocm_chol_potrs_launch_24, nrocm_chol_potrs_launch_32,
                           nrocm_chol_potrs_launch_40, nrocm_chol_potrs_launch_48,
                           nrocm_chol_potrs_launch_56, nrocm_chol_potrs_launch_64};
const BwdLaunch kBwd[8] = {nrocm_chol_bwd_launch_8,  nrocm_chol_bwd_launch_16,
                           nrocm_chol_bwd_launch_24, nrocm_chol_bwd_launch_32,
                           nrocm_chol_bwd_launch_40, nrocm_chol_bwd_launch_48,
                           nrocm_chol_bwd_launch_56, nrocm_chol_bwd_launch_64};

const FusedLaunch kFused[8] = {nrocm_chol_fused_launch_8,  nrocm_chol_fused_launch_16,
                               nrocm_chol_fused_launch_24, nrocm_chol_fused_launch_32,
                               nrocm_chol_fused_launch_40, nrocm_chol_fused_launch_48,
                               nrocm_chol_fused_launch_56, nrocm_chol_fused_launch_64};

int mpw_of_bin(int bin)
{
    const int nb = 8 * (bin + 1);
    const int G = nb <= 8 ? 8 : nb <= 16 ? 16 : nb <= 32 ? 32 : 64;
    return 64 / G;
}

int launch_batched(nrocm_handle h, nrocm_batch bt, hipStream_t st, double *const *A,
                   double *const *B, int *info, bool exact)
{
    for (const auto &bin : bt->bins) {
        const int items = bin.end - bin.begin;
        if (items <= 0) continue;
        const int mpw = mpw_of_bin(bin.kernel);
        const unsigned grid = nrocm_cdiv(items, mpw);
        if (kFused[bin.kernel](exact, grid, st, bt->order_dev, bin.begin, bin.end, bt->n_dev, A,
                               bt->nrhs > 0 ? B : nullptr, bt->nrhs, info))
            NROCM_CHECK_LAUNCH(h);
    }
    return NROCM_OK;
}

}  // namespace

// ------------------------------------------------------------------ internal entry for single matrices

int nrocm_potrf_small(nrocm_handle h, int n, double *a, int lda, double *b, int ldb, int nrhs,
                      int *info, bool exact)
{
    const int bin = std::max(0, (n + 7) / 8 - 1);
    if (kFac[bin](exact, 1, h->stream, nullptr, 0, 1, nullptr, nullptr, info, a, n, lda))
        NROCM_CHECK_LAUNCH(h);
    if (b && nrhs > 0 &&
        kSol[bin](exact, 1, h->stream, nullptr, 0, 1, nullptr, nullptr, nullptr, nrhs, info, a, b, n, lda, ldb))
        NROCM_CHECK_LAUNCH(h);
    return NROCM_OK;
}

// ------------------------------------------------------------------ C ABI

extern "C" {

int nrocm_batch_create(nrocm_handle h, int count, const int *n_host, int nrhs, nrocm_batch *out)
{
    if (!h) return -1;
    if (count < 0) return -2;
    if (count > 0 && !n_host) return -3;
    if (nrhs < 0) return -4;
    if (!out) return -5;
    nrocm_batch bt = new nrocm_batch_s();
    bt->handle = h;
    bt->count = count;
    bt->nrhs = nrhs;
    bt->n_host.assign(n_host, n_host + count);
    for (int j = 0; j < count; ++j) {
        if (n_host[j] < 0) {
            delete bt;
            return -3;
        }
        bt->max_n = std::max(bt->max_n, n_host[j]);
    }
    bt->order_host.resize(count);
    std::iota(bt->order_host.begin(), bt->order_host.end(), 0);
    std::stable_sort(bt->order_host.begin(), bt->order_host.end(),
                     [&](int x, int y) { return n_host[x] > n_host[y]; });
    // Bins of width 8; each bin runs the kernel instantiated for its padded size.
    int pos = 0;
    while (pos < count) {
        const int n = n_host[bt->order_host[pos]];
        const int bin = std::max(0, (n + 7) / 8 - 1);
        int endp = pos;
        while (endp < count && std::max(0, (n_host[bt->order_host[endp]] + 7) / 8 - 1) == bin)
            ++endp;
        bt->bins.push_back({8 * (bin + 1), pos, endp, bin});
        pos = endp;
    }
    if (count > 0) {
        if (hipMalloc(&bt->order_dev, sizeof(int) * count) != hipSuccess ||
            hipMalloc(&bt->n_dev, sizeof(int) * count) != hipSuccess) {
            nrocm_batch_destroy(bt);
            return NROCM_ERR_ALLOC;
        }
        if (hipMemcpy(bt->order_dev, bt->order_host.data(), sizeof(int) * count,
                      hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(bt->n_dev, n_host, sizeof(int) * count, hipMemcpyHostToDevice) != hipSuccess) {
            nrocm_batch_destroy(bt);
            return NROCM_ERR_HIP;
        }
    }
    *out = bt;
    return NROCM_OK;
}

int nrocm_batch_destroy(nrocm_batch bt)
{
    if (!bt) return -1;
    for (auto &g : bt->graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    if (bt->graph_stream) (void)hipStreamDestroy(bt->graph_stream);
    if (bt->order_dev) (void)hipFree(bt->order_dev);
    if (bt->n_dev) (void)hipFree(bt->n_dev);
    delete bt;
    return NROCM_OK;
}

int nrocm_batch_set_graph(nrocm_batch bt, int enable)
{
    if (!bt) return -1;
    bt->use_graph = enable != 0;
    return NROCM_OK;
}

int nrocm_dpotrf_potrs_batched(nrocm_handle h, nrocm_batch bt, double *const *A, double *const *B,
                               int *info, int mode)
{
    if (!h) return -1;
    if (!bt || bt->handle != h) return -2;
    if (bt->count > 0 && !A) return -3;
    if (bt->count > 0 && bt->nrhs > 0 && !B) return -4;
    if (bt->count > 0 && !info) return -5;
    if (mode != NROCM_EXACT && mode != NROCM_FAST) return -6;
    if (bt->count == 0) return NROCM_OK;
    if (bt->max_n > 64)
        return nrocm_fail(h, NROCM_ERR_UNSUPPORTED, "batched Cholesky supports n <= 64");
    const bool exact = mode == NROCM_EXACT;
    if (!bt->use_graph) return launch_batched(h, bt, h->stream, A, B, info, exact);

    const void *key[3] = {A, B, info};
    nrocm_batch_s::GraphEntry *hit = nullptr;
    for (auto &g : bt->graphs)
        if (g.key[0] == key[0] && g.key[1] == key[1] && g.key[2] == key[2] && g.mode == mode) hit = &g;
    if (!hit) {
        constexpr size_t kMaxGraphs = 64;
        if (bt->graphs.size() >= kMaxGraphs) {
            NROCM_CHECK_HIP(h, hipStreamSynchronize(h->stream));
            NROCM_CHECK_HIP(h, hipGraphExecDestroy(bt->graphs.front().exec));
            NROCM_CHECK_HIP(h, hipGraphDestroy(bt->graphs.front().graph));
            bt->graphs.erase(bt->graphs.begin());
        }
        if (!bt->graph_stream)
            NROCM_CHECK_HIP(h, hipStreamCreateWithFlags(&bt->graph_stream, hipStreamNonBlocking));
        NROCM_CHECK_HIP(h, hipStreamBeginCapture(bt->graph_stream, hipStreamCaptureModeThreadLocal));
        int st = launch_batched(h, bt, bt->graph_stream, A, B, info, exact);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(bt->graph_stream, &graph);
        if (st != NROCM_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        if (e != hipSuccess) return nrocm_fail_hip(h, e, "hipStreamEndCapture");
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return nrocm_fail_hip(h, e, "hipGraphInstantiate");
        }
        bt->graphs.push_back({{key[0], key[1], key[2]}, mode, graph, exec});
        hit = &bt->graphs.back();
    }
    NROCM_CHECK_HIP(h, hipGraphLaunch(hit->exec, h->stream));
    return NROCM_OK;
}

int nrocm_dpotrs_backward_batched(nrocm_handle h, nrocm_batch bt, const double *const *L,
                                  const double *const *X, const double *const *dX,
                                  double *const *dA, double *const *dB)
{
    if (!h) return -1;
    if (!bt || bt->handle != h) return -2;
    if (bt->count > 0 && (!L)) return -3;
    if (bt->count > 0 && (!X)) return -4;
    if (bt->count > 0 && (!dX)) return -5;
    if (bt->count > 0 && (!dA)) return -6;
    if (bt->count > 0 && (!dB)) return -7;
    if (bt->count == 0) return NROCM_OK;
    if (bt->max_n > 64)
        return nrocm_fail(h, NROCM_ERR_UNSUPPORTED, "batched Cholesky adjoint supports n <= 64");
    for (const auto &bin : bt->bins) {
        const int items = bin.end - bin.begin;
        if (items <= 0) continue;
        const int mpw = mpw_of_bin(bin.kernel);
        if (kBwd[bin.kernel](nrocm_cdiv(items, mpw), h->stream, bt->order_dev, bin.begin, bin.end,
                             bt->n_dev, L, X, dX, dA, dB, bt->nrhs))
            NROCM_CHECK_LAUNCH(h);
    }
    return NROCM_OK;
}

}  // extern "C"
