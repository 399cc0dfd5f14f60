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
#include "chol_tile.hpp"
#include "chol_tile2.hpp"

#include <algorithm>
#include <numeric>

namespace {

// ------------------------------------------------------------------ dispatch tables

using FacLaunch = int (*)(bool, unsigned, hipStream_t, const int *, int, int, const int *,
                         double *const *, int *, double *, int, int, double);
using SolLaunch = int (*)(bool, unsigned, hipStream_t, const int *, int, int, const int *,
                         double *const *, double *const *, int, const int *, double *, double *,
                         int, int, int);
using BwdLaunch = int (*)(unsigned, hipStream_t, const int *, int, int, const int *,
                         const double *const *, const double *const *, const double *const *,
                         double *const *, double *const *, int);
using FusedLaunch = int (*)(bool, int, hipStream_t, const int *, int, int, const int *, double *const *,
                           double *const *, int, int *);

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

using TileLaunch = int (*)(bool, hipStream_t, const int *, int, int, const int *, double *const *, double *const *,
                          int, int *);
// the element-tile schedule of padded sizes 24..64 (bins 2..7)
const TileLaunch kTile[6] = {nrocm_chol_tile_launch_24, nrocm_chol_tile_launch_32, nrocm_chol_tile_launch_40,
                             nrocm_chol_tile_launch_48, nrocm_chol_tile_launch_56,
                             nrocm_chol_tile_launch_64};

bool tile_ok(int bin, int nrhs) { return bin >= 2 && bin < 8 && nrhs <= nrocm_chol_tile::MAXRHS; }

// the element-tile schedule with two matrices per wavefront, padded sizes 24 and 32 (bins 2, 3)
const TileLaunch kTile2[2] = {nrocm_chol_tile2_launch_24, nrocm_chol_tile2_launch_32};

bool tile2_ok(int bin, int nrhs) { return (bin == 2 || bin == 3) && nrhs <= nrocm_chol_tile::MAXRHS; }

// Whether schedule G can run bin `bin` (padded size 8 (bin + 1)) with nrhs right-hand sides.
bool schedule_ok(int G, int bin, int nrhs)
{
    if (G == NROCM_SCHED_TILE) return tile_ok(bin, nrhs);
    if (G == NROCM_SCHED_TILE2) return tile2_ok(bin, nrhs);
    return (G == 8 || G == 16 || G == 32 || G == 64) && G >= 8 * (bin + 1);
}

// One fused launch of a bin's members [begin, end) of `order` with schedule G.
int launch_bin(bool exact, int G, int bin, hipStream_t st, const int *order, int begin, int end, const int *nsz,
               double *const *A, double *const *B, int nrhs, int *info)
{
    if (G == NROCM_SCHED_TILE) return kTile[bin - 2](exact, st, order, begin, end, nsz, A, B, nrhs, info);
    if (G == NROCM_SCHED_TILE2) return kTile2[bin - 2](exact, st, order, begin, end, nsz, A, B, nrhs, info);
    return kFused[bin](exact, G, st, order, begin, end, nsz, A, B, nrhs, info);
}

// Default lanes per matrix: the smallest power-of-two group holding the rows.
int default_lanes(int bin)
{
    const int nb = 8 * (bin + 1);
    return nb <= 8 ? 8 : nb <= 16 ? 16 : nb <= 32 ? 32 : 64;
}

int mpw_of_bin(int bin) { return 64 / default_lanes(bin); }

// The bins are independent.  With `split`, the bins (sizes in decreasing
// order) are dealt round-robin to the stream and kBranches - 1 auxiliary
// streams, so that in a captured graph they form parallel branches whose
// kernels run side by side (measured on the 65,536-member step: one branch
// 1.134 ms, two 0.911, three 0.950, four 0.937).
constexpr int kBranches = 2;
static_assert(kBranches >= 1 && kBranches <= nrocm_handle_s::kAux + 1, "branches");

int launch_batched(nrocm_handle h, nrocm_batch bt, hipStream_t st, double *const *A,
                   double *const *B, int *info, bool exact, bool split = false)
{
    split = split && kBranches > 1;
    if (split) {
        const int s = nrocm_fork_on(h, st, kBranches - 1);
        if (s != NROCM_OK) return s;
    }
    int k = 0;
    for (const auto &bin : bt->bins) {
        const int items = bin.end - bin.begin;
        if (items <= 0) continue;
        const int G = bt->lanes[exact ? 0 : 1][&bin - bt->bins.data()];
        const int br = split ? k++ % kBranches : 0;
        hipStream_t bs = br ? h->aux[br - 1] : st;
        if (launch_bin(exact, G, bin.kernel, bs, bt->order_dev, bin.begin, bin.end, bt->n_dev, A,
                       bt->nrhs > 0 ? B : nullptr, bt->nrhs, info))
            NROCM_CHECK_LAUNCH(h);
    }
    if (split) {
        const int s = nrocm_join_on(h, st, kBranches - 1);
        if (s != NROCM_OK) return s;
    }
    return NROCM_OK;
}

// ------------------------------------------------------------------ autotuning

// Deterministic strictly diagonally dominant SPD test members for timing:
// A_rc = A_cr = h(r, c) / n with |h| <= 1/2, A_rr = n + 1 + h(r, r) / n, and
// B entries in [-1, 1).  One workgroup per member.
__device__ __forceinline__ double tune_hash(unsigned a, unsigned b, unsigned c)
{
    unsigned x = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u ^ (c + 0x165667B1u) * 0xC2B2AE3Du;
    x ^= x >> 15;
    x *= 0x2C1B3C6Du;
    x ^= x >> 12;
    x *= 0x297A2D39u;
    x ^= x >> 15;
    return (double)(x >> 8) * 0x1p-24 - 0.5;  // [-0.5, 0.5)
}

__global__ void tune_fill_kernel(const int *nsz, double *const *Aptr, double *const *Bptr, int nrhs)
{
    const int j = blockIdx.x;
    const int n = nsz[j];
    double *A = Aptr[j];
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int r = e % n, c = e / n;
        const int lo = r < c ? r : c, hi = r < c ? c : r;
        const double off = tune_hash((unsigned)j, (unsigned)lo, (unsigned)hi) / n;
        A[e] = (r == c) ? (double)(n + 1) + off : off;
    }
    if (Bptr)
        for (int e = threadIdx.x; e < n * nrhs; e += blockDim.x)
            Bptr[j][e] = 2.0 * tune_hash((unsigned)j, 0x10000u + (unsigned)e, 7u);
}

// Times every compiled lanes-per-matrix schedule of each bin on up to
// kTuneSample members of that bin's sizes (generated, so the caller's data
// is untouched) and keeps the fastest.
int autotune(nrocm_handle h, nrocm_batch bt, bool exact)
{
    constexpr int kTuneSample = 4096, kReps = 3;
    const int mi = exact ? 0 : 1;
    hipStream_t st = h->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    NROCM_CHECK_HIP(h, hipEventCreate(&e0));
    NROCM_CHECK_HIP(h, hipEventCreate(&e1));
    int status = NROCM_OK;
    for (size_t bi = 0; bi < bt->bins.size() && status == NROCM_OK; ++bi) {
        const auto &bin = bt->bins[bi];
        const int cand[6] = {8, 16, 32, 64, NROCM_SCHED_TILE, NROCM_SCHED_TILE2};
        auto usable = [&](int G) { return schedule_ok(G, bin.kernel, bt->nrhs); };
        int ncand = 0;
        for (int G : cand) ncand += usable(G) ? 1 : 0;
        if (ncand < 2) continue;
        const int k = std::min(bin.end - bin.begin, kTuneSample);
        std::vector<int> ns(k);
        std::vector<size_t> offA(k), offB(k);
        size_t totA = 0, totB = 0;
        for (int q = 0; q < k; ++q) {
            ns[q] = bt->n_host[bt->order_host[bin.begin + q]];
            offA[q] = totA;
            offB[q] = totB;
            totA += (size_t)ns[q] * ns[q];
            totB += (size_t)ns[q] * bt->nrhs;
        }
        double *dA = nullptr, *dB = nullptr;
        int *dn = nullptr, *dord = nullptr, *dinfo = nullptr;
        double **pA = nullptr, **pB = nullptr;
        std::vector<double *> hpA(k), hpB(k);
        std::vector<int> ord(k);
        auto cleanup = [&]() {
            (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dn); (void)hipFree(dord);
            (void)hipFree(dinfo); (void)hipFree(pA); (void)hipFree(pB);
        };
        if (hipMalloc(&dA, totA * sizeof(double)) != hipSuccess ||
            (totB && hipMalloc(&dB, totB * sizeof(double)) != hipSuccess) ||
            hipMalloc(&dn, k * sizeof(int)) != hipSuccess || hipMalloc(&dord, k * sizeof(int)) != hipSuccess ||
            hipMalloc(&dinfo, k * sizeof(int)) != hipSuccess || hipMalloc(&pA, k * sizeof(double *)) != hipSuccess ||
            hipMalloc(&pB, k * sizeof(double *)) != hipSuccess) {
            cleanup();
            status = nrocm_fail(h, NROCM_ERR_ALLOC, "autotune scratch allocation failed");
            break;
        }
        for (int q = 0; q < k; ++q) {
            hpA[q] = dA + offA[q];
            hpB[q] = totB ? dB + offB[q] : nullptr;
            ord[q] = q;
        }
        double *const *Bp = totB ? pB : nullptr;
        hipError_t e = hipMemcpy(dn, ns.data(), k * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dord, ord.data(), k * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(pA, hpA.data(), k * sizeof(double *), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(pB, hpB.data(), k * sizeof(double *), hipMemcpyHostToDevice);
        float best = 0.0f, t_tile2 = 0.0f;
        int bestG = bt->lanes[mi][bi];
        for (int G : cand) {
            if (!usable(G) || e != hipSuccess) continue;
            float tmin = 1e30f;
            for (int rep = 0; rep <= kReps && e == hipSuccess; ++rep) {  // rep 0 warms up
                hipLaunchKernelGGL(tune_fill_kernel, dim3(k), dim3(256), 0, st, dn, pA, Bp, bt->nrhs);
                e = hipGetLastError();
                if (e == hipSuccess) e = hipEventRecord(e0, st);
                if (e == hipSuccess && launch_bin(exact, G, bin.kernel, st, dord, 0, k, dn, pA, Bp, bt->nrhs, dinfo))
                    e = hipErrorLaunchFailure;
                if (e == hipSuccess) e = hipEventRecord(e1, st);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e == hipSuccess && rep > 0) tmin = std::min(tmin, ms);
            }
            if (e == hipSuccess && (best == 0.0f || tmin < best)) {
                best = tmin;
                bestG = G;
            }
            if (e == hipSuccess && G == NROCM_SCHED_TILE2) t_tile2 = tmin;
        }
        // Two members per wavefront launches half the wavefronts; inside the
        // branch-parallel step that leaves room to the other branch, which a
        // bin timed alone does not show (65,536-member step: 0.88-0.90 ms with
        // it at padded 24 and 32, 0.915-0.92 ms with the one-member tile
        // kernel that the sample timed 1 % faster).  Kept within 2 % of the best.
        if (e == hipSuccess && t_tile2 > 0.0f && t_tile2 <= best * 1.02f) bestG = NROCM_SCHED_TILE2;
        cleanup();
        if (e != hipSuccess) {
            status = nrocm_fail_hip(h, e, "autotune timing");
            break;
        }
        bt->lanes[mi][bi] = bestG;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return status;
}

void drop_graphs(nrocm_batch bt)
{
    for (auto &g : bt->graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    bt->graphs.clear();
}

void report_lanes(nrocm_batch bt, int mi, int *lanes)
{
    if (!lanes) return;
    for (int c = 0; c < 8; ++c) lanes[c] = 0;
    for (size_t bi = 0; bi < bt->bins.size(); ++bi)
        if (bt->bins[bi].kernel < 8) lanes[bt->bins[bi].kernel] = bt->lanes[mi][bi];
}

}  // namespace

// ------------------------------------------------------------------ internal entry for single matrices

int nrocm_potrf_small(nrocm_handle h, int n, double *a, int lda, double *b, int ldb, int nrhs,
                      int *info, bool exact, double substitute)
{
    const int bin = std::max(0, (n + 7) / 8 - 1);
    if (kFac[bin](exact, 1, h->stream, nullptr, 0, 1, nullptr, nullptr, info, a, n, lda, substitute))
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
        bt->lanes[0].push_back(default_lanes(bin));
        bt->lanes[1].push_back(default_lanes(bin));
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
    drop_graphs(bt);
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

int nrocm_batch_autotune(nrocm_handle h, nrocm_batch bt, int mode, int *lanes)
{
    if (!h) return -1;
    if (!bt || bt->handle != h) return -2;
    if (mode != NROCM_EXACT && mode != NROCM_FAST) return -3;
    if (bt->max_n > 64)
        return nrocm_fail(h, NROCM_ERR_UNSUPPORTED, "batched Cholesky supports n <= 64");
    NROCM_CHECK_HIP(h, hipStreamSynchronize(h->stream));
    const int st = autotune(h, bt, mode == NROCM_EXACT);
    drop_graphs(bt);  // captured launches used the previous schedule
    report_lanes(bt, mode == NROCM_EXACT ? 0 : 1, lanes);
    return st;
}

int nrocm_batch_set_schedule(nrocm_batch bt, int mode, const int *lanes)
{
    if (!bt) return -1;
    if (mode != NROCM_EXACT && mode != NROCM_FAST) return -2;
    if (!lanes) return -3;
    if (bt->max_n > 64)
        return nrocm_fail(bt->handle, NROCM_ERR_UNSUPPORTED, "batched Cholesky supports n <= 64");
    for (int c = 0; c < 8; ++c) {
        const int G = lanes[c];
        if (G != 0 && !schedule_ok(G, c, bt->nrhs)) return -3;
    }
    const int mi = mode == NROCM_EXACT ? 0 : 1;
    for (size_t bi = 0; bi < bt->bins.size(); ++bi)
        if (lanes[bt->bins[bi].kernel]) bt->lanes[mi][bi] = lanes[bt->bins[bi].kernel];
    drop_graphs(bt);
    return NROCM_OK;
}

int nrocm_batch_schedule(nrocm_batch bt, int mode, int *lanes)
{
    if (!bt) return -1;
    if (mode != NROCM_EXACT && mode != NROCM_FAST) return -2;
    if (!lanes) return -3;
    if (bt->max_n > 64)
        return nrocm_fail(bt->handle, NROCM_ERR_UNSUPPORTED, "batched Cholesky supports n <= 64");
    report_lanes(bt, mode == NROCM_EXACT ? 0 : 1, lanes);
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
        const int sa = nrocm_aux_init(h);  // streams and events exist before the capture
        if (sa != NROCM_OK) return sa;
        NROCM_CHECK_HIP(h, hipStreamBeginCapture(bt->graph_stream, hipStreamCaptureModeThreadLocal));
        int st = launch_batched(h, bt, bt->graph_stream, A, B, info, exact, true);
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
