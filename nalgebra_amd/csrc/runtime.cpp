// Handle, error reporting, workspace management and the seeded input
// generators of libnalgebra_rocm.
#include "nrocm_internal.hpp"

#include <cmath>
#include <cstring>

int nrocm_fail_hip(nrocm_handle h, hipError_t e, const char *what)
{
    if (h) {
        h->last_error = std::string(what) + ": " + hipGetErrorString(e);
    }
    return NROCM_ERR_HIP;
}

int nrocm_fail(nrocm_handle h, int status, const char *msg)
{
    if (h) h->last_error = msg;
    return status;
}

int nrocm_workspace(nrocm_handle h, size_t bytes, void **ptr)
{
    if (bytes > h->ws_bytes) {
        if (h->ws) {
            // Work already queued may still use the old buffer.
            NROCM_CHECK_HIP(h, hipStreamSynchronize(h->stream));
            NROCM_CHECK_HIP(h, hipFree(h->ws));
            h->ws = nullptr;
            h->ws_bytes = 0;
        }
        size_t want = bytes + bytes / 4;
        if (hipMalloc(&h->ws, want) != hipSuccess) {
            h->ws = nullptr;
            return nrocm_fail(h, NROCM_ERR_ALLOC, "workspace allocation failed");
        }
        h->ws_bytes = want;
    }
    *ptr = h->ws;
    return NROCM_OK;
}

int nrocm_aux_init(nrocm_handle h)
{
    if (!h->fork_ev) {
        NROCM_CHECK_HIP(h, hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming));
        for (int i = 0; i < nrocm_handle_s::kAux; ++i) {
            NROCM_CHECK_HIP(h, hipStreamCreateWithFlags(&h->aux[i], hipStreamNonBlocking));
            NROCM_CHECK_HIP(h, hipEventCreateWithFlags(&h->join_ev[i], hipEventDisableTiming));
        }
    }
    return NROCM_OK;
}

int nrocm_fork_on(nrocm_handle h, hipStream_t st, int count)
{
    const int s = nrocm_aux_init(h);
    if (s != NROCM_OK) return s;
    NROCM_CHECK_HIP(h, hipEventRecord(h->fork_ev, st));
    for (int i = 0; i < count; ++i) NROCM_CHECK_HIP(h, hipStreamWaitEvent(h->aux[i], h->fork_ev, 0));
    return NROCM_OK;
}

int nrocm_join_on(nrocm_handle h, hipStream_t st, int count)
{
    for (int i = 0; i < count; ++i) {
        NROCM_CHECK_HIP(h, hipEventRecord(h->join_ev[i], h->aux[i]));
        NROCM_CHECK_HIP(h, hipStreamWaitEvent(st, h->join_ev[i], 0));
    }
    return NROCM_OK;
}

int nrocm_fork(nrocm_handle h) { return nrocm_fork_on(h, h->stream, nrocm_handle_s::kAux); }

int nrocm_join(nrocm_handle h) { return nrocm_join_on(h, h->stream, nrocm_handle_s::kAux); }

extern "C" {

int nrocm_version(void) { return 100; }

int nrocm_create(nrocm_handle *handle)
{
    if (!handle) return -1;
    nrocm_handle h = new nrocm_handle_s();
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return NROCM_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) {
        delete h;
        return NROCM_ERR_HIP;
    }
    h->num_cus = prop.multiProcessorCount;
    if (hipMalloc(&h->small, 4096) != hipSuccess) {
        delete h;
        return NROCM_ERR_ALLOC;
    }
    *handle = h;
    return NROCM_OK;
}

int nrocm_destroy(nrocm_handle h)
{
    if (!h) return -1;
    (void)hipStreamSynchronize(h->stream);
    for (auto &g : h->graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    if (h->capture_stream) (void)hipStreamDestroy(h->capture_stream);
    for (int i = 0; i < nrocm_handle_s::kAux; ++i) {
        if (h->aux[i]) (void)hipStreamDestroy(h->aux[i]);
        if (h->join_ev[i]) (void)hipEventDestroy(h->join_ev[i]);
    }
    if (h->fork_ev) (void)hipEventDestroy(h->fork_ev);
    if (h->ws) (void)hipFree(h->ws);
    if (h->small) (void)hipFree(h->small);
    delete h;
    return NROCM_OK;
}

int nrocm_set_stream(nrocm_handle h, void *stream)
{
    if (!h) return -1;
    h->stream = (hipStream_t)stream;
    return NROCM_OK;
}

int nrocm_get_stream(nrocm_handle h, void **stream)
{
    if (!h) return -1;
    if (!stream) return -2;
    *stream = (void *)h->stream;
    return NROCM_OK;
}

const char *nrocm_status_string(int status)
{
    switch (status) {
    case NROCM_OK: return "ok";
    case NROCM_ERR_HIP: return "HIP runtime error";
    case NROCM_ERR_ALLOC: return "device allocation failed";
    case NROCM_ERR_UNSUPPORTED: return "unsupported request";
    case NROCM_ERR_NOT_CONVERGED: return "iterative refinement did not converge";
    default: return status < 0 ? "invalid argument" : "unknown status";
    }
}

const char *nrocm_last_error(nrocm_handle h)
{
    return h ? h->last_error.c_str() : "null handle";
}

// ---------------------------------------------------------------- generators
// xoshiro256** (Blackman & Vigna, 2018) seeded from SplitMix64 (Steele et
// al., 2014), both public-domain published algorithms.

namespace {

struct Rng {
    uint64_t s[4];
    double spare = 0.0;
    bool has_spare = false;

    static uint64_t splitmix(uint64_t &x)
    {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

    explicit Rng(uint64_t seed)
    {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) s[i] = splitmix(x);
    }
    uint64_t next()
    {
        uint64_t result = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    // u53 * 2^-53 in [0, 1)
    double uniform01() { return (double)(next() >> 11) * 0x1.0p-53; }
    // Box-Muller on two 53-bit uniforms; both outputs are used.
    double normal()
    {
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        double u1 = ((double)(next() >> 11) + 1.0) * 0x1.0p-53;   // (0, 1]
        double u2 = (double)(next() >> 11) * 0x1.0p-53;
        double r = std::sqrt(-2.0 * std::log(u1));
        double t = 6.283185307179586 * u2;
        spare = r * std::sin(t);
        has_spare = true;
        return r * std::cos(t);
    }
};

}  // namespace

void nrocm_gen_uniform(uint64_t seed, size_t count, double lo, double hi, double *out)
{
    Rng r(seed);
    for (size_t i = 0; i < count; ++i) out[i] = lo + (hi - lo) * r.uniform01();
}

void nrocm_gen_normal(uint64_t seed, size_t count, double *out)
{
    Rng r(seed);
    for (size_t i = 0; i < count; ++i) out[i] = r.normal();
}

void nrocm_gen_spd_batch(uint64_t seed, int count, const int *n_host, int nrhs, double *a,
                         const int64_t *offa, double *b, const int64_t *offb)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int j = 0; j < count; ++j) {
        const int n = n_host[j];
        Rng r(seed + (uint64_t)j);
        std::vector<double> g((size_t)n * n);
        for (size_t i = 0; i < g.size(); ++i) g[i] = r.normal();
        double *bj = b ? b + offb[j] : nullptr;
        for (int i = 0; i < n * nrhs; ++i) {
            double v = r.normal();
            if (bj) bj[i] = v;
        }
        double *aj = a + offa[j];
        for (int c = 0; c < n; ++c)
            for (int rr = c; rr < n; ++rr) {
                double s = 0.0;
                for (int k = 0; k < n; ++k) s += g[rr + (size_t)k * n] * g[c + (size_t)k * n];
                if (rr == c) s += (double)n;
                aj[rr + (size_t)c * n] = s;
                aj[c + (size_t)rr * n] = s;
            }
    }
}

}  // extern "C"
