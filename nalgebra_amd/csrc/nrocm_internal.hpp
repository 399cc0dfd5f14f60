// Internal declarations shared by the HIP translation units of libnalgebra_rocm.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/nalgebra_rocm.h"

// Workgroup barrier ordering LDS accesses only.  __syncthreads() also waits
// for every outstanding global load and store of the calling wavefront, which
// stalls kernels that keep global stores in flight across their barriers.
// Use this only where threads exchange data through LDS, not global memory.
// (This header is also read by the host-only runtime, built without hipcc.)
#ifdef __HIPCC__
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
#endif

constexpr int NROCM_SMALL_PLAN = 256;  // int offset of the LU row plan (257 ints) in small

struct nrocm_handle_s {
    hipStream_t stream = nullptr;   // caller's stream (nullptr = legacy default stream)
    int device = 0;
    int num_cus = 0;
    std::string last_error;

    // Growable scratch on the device, reused across calls (ordered on `stream`).
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // Small device-side scalars (counters, flags, norms).
    void *small = nullptr;          // 4 KiB: int[0] Cholesky block info, ints from NROCM_SMALL_PLAN: LU row plan
    // Streams for independent launches that may run side by side (created on
    // first use, default priority), with the events that fork them from and
    // join them back into `stream`.
    static constexpr int kAux = 4;
    hipStream_t aux[kAux] = {};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[kAux] = {};

    // hipGraph cache for long launch sequences (keyed by every argument that
    // shapes the sequence), captured on a private stream, replayed on `stream`.
    struct Graph {
        long long key[8];
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    hipStream_t capture_stream = nullptr;
    // Epoch of the last look-back triangular solve launched on this handle
    // (tags its published values; never reused while the handle lives).
    unsigned trsv_epoch = 0;
    // 1 when the last nrocm_dgesv_ir call replaced its windowed factors by
    // partial-pivoting factors.
    int ir_fallback = 0;
};

// Runs `body(stream)` through the handle's graph cache: on a miss the launches
// are captured into a new graph; either way the graph is launched on h->stream.
template <typename F>
int nrocm_graph_run(nrocm_handle h, const long long (&key)[8], F body);

struct nrocm_batch_s {
    nrocm_handle handle = nullptr;
    int count = 0;
    int nrhs = 0;
    int max_n = 0;
    std::vector<int> n_host;        // sizes in caller order
    std::vector<int> order_host;    // member indices, sorted by decreasing size
    // Bins: [begin, end) ranges of `order` sharing one kernel instantiation.
    struct Bin { int nb; int begin; int end; int kernel; };
    std::vector<Bin> bins;
    // Lanes per matrix of each bin's fused kernel, per mode (EXACT, FAST):
    // the default heuristic until nrocm_batch_autotune measures them.
    std::vector<int> lanes[2];
    int *order_dev = nullptr;       // device copy of order_host
    int *n_dev = nullptr;           // device copy of n_host (caller order)
    // hipGraph replays of captured launch sequences, keyed by their arguments
    // (a bounded cache, oldest entry evicted first).
    struct GraphEntry {
        const void *key[3];
        int mode;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    bool use_graph = true;
    std::vector<GraphEntry> graphs;
    hipStream_t graph_stream = nullptr;
};

// Records a message on the handle and returns NROCM_ERR_HIP.
int nrocm_fail_hip(nrocm_handle h, hipError_t e, const char *what);
int nrocm_fail(nrocm_handle h, int status, const char *msg);
// Ensures at least `bytes` of workspace; synchronises the stream when it grows.
int nrocm_workspace(nrocm_handle h, size_t bytes, void **ptr);
// Forks the handle's auxiliary streams from its stream (nrocm_fork) and joins
// them back (nrocm_join): work launched on aux[i] in between runs concurrently
// with the other auxiliary streams and after everything queued before the fork.
int nrocm_fork(nrocm_handle h);
int nrocm_join(nrocm_handle h);
// The same for the first `count` auxiliary streams and any origin stream st
// (also a stream under graph capture: the auxiliary work joins the capture;
// call nrocm_aux_init before the capture begins so that it creates nothing).
int nrocm_aux_init(nrocm_handle h);
int nrocm_fork_on(nrocm_handle h, hipStream_t st, int count);
int nrocm_join_on(nrocm_handle h, hipStream_t st, int count);

#define NROCM_CHECK_HIP(h, call)                                    \
    do {                                                            \
        hipError_t e__ = (call);                                    \
        if (e__ != hipSuccess) return nrocm_fail_hip((h), e__, #call); \
    } while (0)

#define NROCM_CHECK_LAUNCH(h)                                       \
    do {                                                            \
        hipError_t e__ = hipGetLastError();                         \
        if (e__ != hipSuccess) return nrocm_fail_hip((h), e__, "kernel launch"); \
    } while (0)

static inline unsigned nrocm_cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

template <typename F>
int nrocm_graph_run(nrocm_handle h, const long long (&key)[8], F body)
{
    nrocm_handle_s::Graph *hit = nullptr;
    for (auto &g : h->graphs) {
        bool same = true;
        for (int i = 0; i < 8; ++i) same = same && g.key[i] == key[i];
        if (same) hit = &g;
    }
    if (!hit) {
        if (h->graphs.size() >= 32) {
            NROCM_CHECK_HIP(h, hipStreamSynchronize(h->stream));
            NROCM_CHECK_HIP(h, hipGraphExecDestroy(h->graphs.front().exec));
            NROCM_CHECK_HIP(h, hipGraphDestroy(h->graphs.front().graph));
            h->graphs.erase(h->graphs.begin());
        }
        if (!h->capture_stream)
            NROCM_CHECK_HIP(h, hipStreamCreateWithFlags(&h->capture_stream, hipStreamNonBlocking));
        NROCM_CHECK_HIP(h, hipStreamBeginCapture(h->capture_stream, hipStreamCaptureModeThreadLocal));
        const int st = body(h->capture_stream);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(h->capture_stream, &graph);
        if (st != NROCM_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        if (e != hipSuccess) return nrocm_fail_hip(h, e, "hipStreamEndCapture");
        hipGraphExec_t exec = nullptr;
        const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e2 != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return nrocm_fail_hip(h, e2, "hipGraphInstantiate");
        }
        nrocm_handle_s::Graph g;
        for (int i = 0; i < 8; ++i) g.key[i] = key[i];
        g.graph = graph;
        g.exec = exec;
        h->graphs.push_back(g);
        hit = &h->graphs.back();
    }
    NROCM_CHECK_HIP(h, hipGraphLaunch(hit->exec, h->stream));
    return NROCM_OK;
}
