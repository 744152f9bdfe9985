/*
 * hist.hip -- CausalBP::run(tol, minIters, maxIters, histLength), the
 * 4-argument run (csrc/ext/causal_bp.cu:1006-1143, HostBeliefHistory
 * 28-78), with the belief history kept on the device.
 *
 * The reference copies every sweep's beliefs to the host, counts the
 * variables that moved by more than tol and pushes each finite belief into
 * a per-variable ring of histLength values; when it stops on the fraction
 * test or on its limits it returns each variable's mean over the ring.
 * Here one kernel per sweep does the count and the push on the device
 * (the ring is [histLength][V], so a sweep's pushes are coalesced), and
 * only the count crosses to the host each sweep.  The loop's tests and
 * their order are the reference's; the mean adds the stored values oldest
 * first and divides by their number, as HostBeliefHistory::mean does, so
 * the low-pass beliefs are its bits.
 *
 * Two entry points share this: flbp_plan_run_hist allocates the history
 * per call; flbp_plan_run_hist_noalloc takes it from a fixed array in
 * device memory (`scratch` below) and allocates nothing.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <limits>
#include <mutex>

#include "fastlbp.h"
#include "flbp_internal.h"
#include "plan_internal.h"

namespace {

constexpr unsigned kHistBlock = 256;

/* moved = |prev - cur| > tol (false for NaN, as on the host); finite
 * beliefs enter the ring; prev <- cur */
__global__ __launch_bounds__(kHistBlock) void hist_step(const double *__restrict__ cur, double *__restrict__ prev,
                                                        double *__restrict__ ring, uint32_t *__restrict__ next,
                                                        uint32_t *__restrict__ count, uint64_t V, uint32_t L,
                                                        double tol, unsigned long long *__restrict__ moved) {
    const uint64_t v = (uint64_t)blockIdx.x * kHistBlock + threadIdx.x;
    bool m = false;
    if (v < V) {
        const double x = cur[v];
        m = fabs(prev[v] - x) > tol;
        prev[v] = x;
        if (__builtin_isfinite(x)) {
            const uint32_t s = next[v];
            ring[(uint64_t)s * V + v] = x;
            next[v] = s + 1 == L ? 0u : s + 1;
            if (count[v] < L) count[v] = count[v] + 1;
        }
    }
    const unsigned long long n = __popcll(__ballot(m));
    if ((threadIdx.x & 63u) == 0 && n) atomicAdd(moved, n);
}

__global__ __launch_bounds__(kHistBlock) void hist_mean(const double *__restrict__ ring,
                                                        const uint32_t *__restrict__ next,
                                                        const uint32_t *__restrict__ count, uint64_t V, uint32_t L,
                                                        double *__restrict__ out) {
    const uint64_t v = (uint64_t)blockIdx.x * kHistBlock + threadIdx.x;
    if (v >= V) return;
    const uint32_t c = count[v];
    uint32_t s = c == L ? next[v] : 0u;
    double sum = 0.0;
    for (uint32_t k = 0; k < c; k++) {
        sum += ring[(uint64_t)s * V + v];
        s = s + 1 == L ? 0u : s + 1;
    }
    if (c > 0) sum /= (double)c;
    out[v] = sum;
}

/* u32 array <- 0 (the history's counters; a kernel rather than a memset) */
__global__ __launch_bounds__(kHistBlock) void hist_zero_u32(uint32_t *__restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * kHistBlock + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void hist_zero_moved(unsigned long long *moved) { *moved = 0ull; }

/* out[perm[dv]] = src[dv]: device order -> variable order */
__global__ __launch_bounds__(kHistBlock) void hist_unpermute(const double *__restrict__ src,
                                                             const uint32_t *__restrict__ perm,
                                                             double *__restrict__ out, uint64_t V) {
    const uint64_t dv = (uint64_t)blockIdx.x * kHistBlock + threadIdx.x;
    if (dv < V) out[perm[dv]] = src[dv];
}

/* The history's device memory: ring [hist_len][V] doubles, prev [V], low
 * [V], next [V] and count [V] (u32), the moved counter; each range 256-byte
 * aligned. */
struct HistDev {
    double *ring = nullptr, *prev = nullptr, *low = nullptr;
    uint32_t *next = nullptr, *count = nullptr;
    unsigned long long *moved = nullptr;
};

struct HistLayout {
    uint64_t moved, prev, low, next, count, ring, total;
};

inline uint64_t up256(uint64_t x) { return (x + 255u) & ~uint64_t(255); }

/* byte offsets of the ranges for V variables and hist_len L; the caller has
 * checked that 8 * max(V, 1) * L does not overflow */
HistLayout hist_layout(uint64_t V, uint64_t L) {
    const uint64_t nb = V ? V : 1;
    HistLayout h;
    uint64_t off = 0;
    h.moved = off;
    off += 256;
    h.prev = off;
    off += up256(8 * nb);
    h.low = off;
    off += up256(8 * nb);
    h.next = off;
    off += up256(4 * nb);
    h.count = off;
    off += up256(4 * nb);
    h.ring = off;
    off += up256(8 * nb * L);
    h.total = off;
    return h;
}

HistDev hist_at(unsigned char *base, const HistLayout &h) {
    HistDev B;
    B.moved = reinterpret_cast<unsigned long long *>(base + h.moved);
    B.prev = reinterpret_cast<double *>(base + h.prev);
    B.low = reinterpret_cast<double *>(base + h.low);
    B.next = reinterpret_cast<uint32_t *>(base + h.next);
    B.count = reinterpret_cast<uint32_t *>(base + h.count);
    B.ring = reinterpret_cast<double *>(base + h.ring);
    return B;
}

/* device memory of one flbp_plan_run_hist call, freed on every return */
struct HistAlloc {
    unsigned char *base = nullptr;
    ~HistAlloc() {
        if (base) (void)hipFree(base);
    }
};

/* The no-allocation entry point's history: one array per device, resolved
 * to its device address once per device (hipGetSymbolAddress, with that
 * device current) and kept for the life of the process.  The array's name
 * in host code is not a device address: only the resolved pointer is
 * passed to kernels and copies. */
constexpr uint64_t kHistScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kHistScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

/* the current device's scratch address (hipSetDevice done by the caller) */
int scratch_on(int device, unsigned char **out) {
    if (device < 0 || device >= kMaxDevices)
        return flbp::fail(FLBP_EINVAL, "flbp_plan_run_hist_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return flbp::fail(FLBP_EHIP, std::string("flbp_plan_run_hist_noalloc: hipGetSymbolAddress: ") +
                                             hipGetErrorString(e));
        }
        g_scratch_ptr[device] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[device];
    return FLBP_OK;
}

}  // namespace

using flbp::Plan;
using flbp::grid_for;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return flbp::fail(FLBP_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

/* argument checks shared by both entry points */
int hist_args(Plan *P, const flbp_hist_result *res, uint64_t hist_len, const char *who) {
    if (!P || !res) return flbp::fail(FLBP_EINVAL, std::string(who) + ": null argument");
    if (hist_len == 0)
        return flbp::fail(FLBP_EINVAL, std::string(who) + ": belief history length must be positive");
    const uint64_t V = P->V;
    if (hist_len > 0xffffffffull || (V && hist_len > std::numeric_limits<uint64_t>::max() / 16 / V))
        return flbp::fail(FLBP_ERANGE, std::string(who) + ": belief history too large");
    return FLBP_OK;
}

/* the loop of causal_bp.cu:1006-1143 on the history B */
int run_hist_on(Plan *P, double tol, uint64_t min_iters, uint64_t max_iters, uint32_t L, double maxtime,
                double damping, const HistDev &B, flbp_hist_result *res, double *low_pass) {
    flbp_plan *plan = reinterpret_cast<flbp_plan *>(P);
    const uint64_t V = P->V;
    const uint64_t nb = V ? V : 1;
    const auto tic = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(hist_zero_u32, dim3(grid_for(nb, kHistBlock)), dim3(kHistBlock), 0, P->stream, B.next, nb);
    hipLaunchKernelGGL(hist_zero_u32, dim3(grid_for(nb, kHistBlock)), dim3(kHistBlock), 0, P->stream, B.count, nb);
    HIP_TRY(hipGetLastError());
    /* the beliefs before the first sweep, as the reference reads them */
    if (V) HIP_TRY(hipMemcpyAsync(B.prev, P->belief, V * sizeof(double), hipMemcpyDeviceToDevice, P->stream));
    uint64_t num_iters = 0;
    double max_diff = INFINITY, yet = 1.0, node_frac = 0.0;
    int reason = 2;
    for (;; num_iters++) {
        if (num_iters >= min_iters) node_frac = (double)(num_iters - min_iters) / (double)(max_iters - min_iters);
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - tic).count();
        if (max_diff <= tol) {
            reason = 0;
            break;
        } else if (num_iters > min_iters && yet < node_frac) {
            reason = 1;
            break;
        } else if (num_iters > max_iters || elapsed > maxtime) {
            reason = 2;
            break;
        }
        flbp_run_result r;
        if (int rc = flbp_plan_run_fixed(plan, 1, damping, &r)) return rc;
        max_diff = r.maxdiff;
        unsigned long long moved = 0;
        hipLaunchKernelGGL(hist_zero_moved, dim3(1), dim3(1), 0, P->stream, B.moved);
        if (V)
            hipLaunchKernelGGL(hist_step, dim3(grid_for(V, kHistBlock)), dim3(kHistBlock), 0, P->stream, P->belief,
                               B.prev, B.ring, B.next, B.count, V, L, tol, B.moved);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&moved, B.moved, sizeof(moved), hipMemcpyDeviceToHost, P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
        yet = (double)moved / (double)V;
    }
    res->yet_to_converge = yet;
    res->maxdiff = max_diff;
    res->iters = num_iters;
    res->reason = reason;
    if (low_pass && V) {
        const double *src = P->belief;
        if (reason != 0) {
            hipLaunchKernelGGL(hist_mean, dim3(grid_for(V, kHistBlock)), dim3(kHistBlock), 0, P->stream, B.ring,
                               B.next, B.count, V, L, B.low);
            HIP_TRY(hipGetLastError());
            src = B.low;
        }
        /* into variable order on the device (prev is free now), then one copy */
        hipLaunchKernelGGL(hist_unpermute, dim3(grid_for(V, kHistBlock)), dim3(kHistBlock), 0, P->stream, src,
                           P->var_perm, B.prev, V);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(low_pass, B.prev, V * sizeof(double), hipMemcpyDeviceToHost, P->stream));
    }
    HIP_TRY(hipStreamSynchronize(P->stream));
    return FLBP_OK;
}

}  // namespace

extern "C" int flbp_plan_run_hist(flbp_plan *plan, double tol, uint64_t min_iters, uint64_t max_iters,
                                  uint64_t hist_len, double maxtime, double damping, flbp_hist_result *res,
                                  double *low_pass) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (int rc = hist_args(P, res, hist_len, "flbp_plan_run_hist")) return rc;
    HIP_TRY(hipSetDevice(P->device));
    const HistLayout h = hist_layout(P->V, hist_len);
    HistAlloc A;
    if (hipMalloc((void **)&A.base, h.total) != hipSuccess) {
        A.base = nullptr;
        (void)hipGetLastError();
        return flbp::fail(FLBP_ENOMEM, "flbp_plan_run_hist: belief history allocation failed");
    }
    return run_hist_on(P, tol, min_iters, max_iters, (uint32_t)hist_len, maxtime, damping, hist_at(A.base, h), res,
                       low_pass);
}

extern "C" int flbp_plan_run_hist_noalloc(flbp_plan *plan, double tol, uint64_t min_iters, uint64_t max_iters,
                                          uint64_t hist_len, double maxtime, double damping, flbp_hist_result *res,
                                          double *low_pass) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (int rc = hist_args(P, res, hist_len, "flbp_plan_run_hist_noalloc")) return rc;
    const HistLayout h = hist_layout(P->V, hist_len);
    if (h.total > kHistScratchBytes)
        return flbp::fail(FLBP_ERANGE, "flbp_plan_run_hist_noalloc: the belief history needs " +
                                           std::to_string(h.total) + " bytes, above the fixed " +
                                           std::to_string(kHistScratchBytes) + " (use flbp_plan_run_hist)");
    HIP_TRY(hipSetDevice(P->device));
    unsigned char *base = nullptr;
    if (int rc = scratch_on(P->device, &base)) return rc;
    /* sweeps by direct launches: no graph is instantiated here */
    const bool graphs = P->use_graphs;
    P->use_graphs = false;
    const int rc = run_hist_on(P, tol, min_iters, max_iters, (uint32_t)hist_len, maxtime, damping,
                               hist_at(base, h), res, low_pass);
    P->use_graphs = graphs;
    return rc;
}
