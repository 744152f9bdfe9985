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
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <limits>
#include <vector>

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

/* device buffers of one call, freed on every return */
struct HistBufs {
    double *ring = nullptr, *prev = nullptr, *low = nullptr;
    uint32_t *next = nullptr, *count = nullptr;
    unsigned long long *moved = nullptr;
    ~HistBufs() {
        for (void *p : {(void *)ring, (void *)prev, (void *)low, (void *)next, (void *)count, (void *)moved})
            if (p) (void)hipFree(p);
    }
};

}  // namespace

using flbp::Plan;
using flbp::grid_for;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return flbp::fail(FLBP_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" int flbp_plan_run_hist(flbp_plan *plan, double tol, uint64_t min_iters, uint64_t max_iters,
                                  uint64_t hist_len, double maxtime, double damping, flbp_hist_result *res,
                                  double *low_pass) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P || !res) return flbp::fail(FLBP_EINVAL, "flbp_plan_run_hist: null argument");
    if (hist_len == 0) return flbp::fail(FLBP_EINVAL, "flbp_plan_run_hist: belief history length must be positive");
    const uint64_t V = P->V;
    if (hist_len > 0xffffffffull || (V && hist_len > std::numeric_limits<uint64_t>::max() / 8 / V))
        return flbp::fail(FLBP_ERANGE, "flbp_plan_run_hist: belief history too large");
    HIP_TRY(hipSetDevice(P->device));
    const auto tic = std::chrono::steady_clock::now();
    const uint32_t L = (uint32_t)hist_len;
    HistBufs B;
    const uint64_t nb = V ? V : 1;
    if (hipMalloc((void **)&B.ring, nb * hist_len * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&B.prev, nb * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&B.low, nb * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&B.next, nb * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void **)&B.count, nb * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void **)&B.moved, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipGetLastError();
        return flbp::fail(FLBP_ENOMEM, "flbp_plan_run_hist: belief history allocation failed");
    }
    HIP_TRY(hipMemsetAsync(B.next, 0, nb * sizeof(uint32_t), P->stream));
    HIP_TRY(hipMemsetAsync(B.count, 0, nb * sizeof(uint32_t), P->stream));
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
        HIP_TRY(hipMemsetAsync(B.moved, 0, sizeof(unsigned long long), P->stream));
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
        std::vector<double> dev(V);
        HIP_TRY(hipMemcpyAsync(dev.data(), src, V * sizeof(double), hipMemcpyDeviceToHost, P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
        for (uint64_t dv = 0; dv < V; dv++) low_pass[P->h_perm[dv]] = dev[dv];
    }
    HIP_TRY(hipStreamSynchronize(P->stream));
    return FLBP_OK;
}
