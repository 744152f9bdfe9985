/*
 * plan.hip -- the plain PARALL solve on one GPU: device layout, the two
 * sweep kernels, and the chunked run loop with a device-side stop test.
 *
 * Device layout (all arrays owned by the plan, sized once at creation):
 *   fac_off  u32[F+1]   factor-major edge ranges (h_row_ptr_vf)
 *   fac_meta u32[F]     type | head-clamped bit | head slot << 8
 *   fac_p    f64[F]     prob(); fac_q f64[F] prob_default()
 *   fac_mask f64x2[F]   head_mask, read only for head-clamped factors
 *   fac_vpos u32[E]     var-major position of each factor-major slot
 *   var_off  u32[V+1]   var-major edge ranges (h_row_ptr_fv)
 *   var_edge u32[E]     factor-major slot of each var-major edge
 *   m_vf     f64x2[E]   v->f messages, (state0, state1) pairs: var-major
 *                       (variants 0, 1), factor-major (2, 3) or in variant
 *                       4's storage order
 *   m_fv     f64x2[E]   f->v messages, factor-major (in variant 3 a copy kept
 *                       for damping; fm_stale marks it out of date); in
 *                       variant 4 each staged factor tile's slots in its
 *                       work order (h_store: slot -> storage position)
 *   fac_var, var_src u32[E]  variant 4: device variable at each storage
 *                       position; storage position of each var-major edge
 *   m_fv_vm  f64x2[E]   f->v messages, var-major (variant 3)
 *   belief   f64[V]     P(x=1) of the last sweep (the reference's _oldBeliefsV)
 *   ctl      u64[chunk][64 slots][8]  per-sweep partial max |b - b_old|
 *            (IEEE bits), one 64-byte line per slot; ctl_out u64[chunk]
 *
 * Replaces the reference's per-iteration sequence in csrc/ext/causal_bp.cu:
 * calcNewMessageFusedAll (409-444), calcBeliefsV (483-515), the dist/max
 * reduce with its blocking D->H read (879-886) and the two per-iteration
 * device_vector allocations (876-877).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fastlbp.h"
#include "flbp_internal.h"
#include "plan_internal.h"
#include "staged_kernels.h"
#include "sweep_kernels.h"
#include "state_kernels.h"

using namespace flbp::dev;

namespace {

/* zero the per-sweep slots of a chunk (a kernel, so the sequence can be
 * captured in a graph) */
__global__ void clear_u64(unsigned long long *p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

/* one wave: the max over the slots of each of the chunk's n sweeps */
__global__ void ctl_compact(const unsigned long long *ctl, unsigned long long *out, int n) {
    for (int k = 0; k < n; k++) {
        const unsigned long long m = ctl_read(ctl, k);
        if (threadIdx.x == 0) out[k] = m;
    }
}

__global__ void fill_pairs(double2 *p, uint64_t n, double c) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_double2(c, c);
}

/* factor-major f->v copy from the var-major array */
__global__ void gather_pairs(const double2 *__restrict__ src, const uint32_t *__restrict__ pos,
                             double2 *__restrict__ dst, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[pos[i]];
}

/* dst[pos[i]] = src[i] */
__global__ void scatter_pairs(const double2 *__restrict__ src, const uint32_t *__restrict__ pos,
                              double2 *__restrict__ dst, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[pos[i]] = src[i];
}

/* out[i] = src[idx[i]] */
__global__ void gather_f64(const double *__restrict__ src, const uint32_t *__restrict__ idx, double *__restrict__ out,
                           uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = src[idx[i]];
}

/* beliefs straight after init with no sweep: prod = (c, c), zero counts 0 */
__global__ void fill_f64(double *p, uint64_t n, double c) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = c;
}

}  // namespace

namespace flbp {

int hip_fail(hipError_t e, const char *what) {
    return fail(FLBP_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                   \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

/* variant 4: m_vf <- the current v->f messages in the storage order of the
 * f->v messages (derived from the records and the f->v messages, or the
 * init value) */
int Plan::materialize_vf() {
    if (!E) return FLBP_OK;
    if (vf_src == kVfState)
        hipLaunchKernelGGL(vf_from_state, dim3(grid_for(E, 256)), dim3(256), 0, stream, vprod, vzc, fac_var, m_fv,
                           m_vf, E);
    else if (vf_src == kVfConst)
        hipLaunchKernelGGL(fill_pairs, dim3(grid_for(E, 256)), dim3(256), 0, stream, m_vf, E, init_value);
    vf_src = kVfArray;
    HIP_TRY(hipGetLastError());
    /* callers copy m_vf or overwrite m_fv with blocking copies next, which
     * do not order against the plan's non-blocking stream */
    HIP_TRY(hipStreamSynchronize(stream));
    return FLBP_OK;
}

int Plan::fm_current() {
    if (fm_stale && E)
        hipLaunchKernelGGL(gather_pairs, dim3(grid_for(E, 256)), dim3(256), 0, stream, m_fv_vm, fac_vpos, m_fv, E);
    fm_stale = false;
    HIP_TRY(hipGetLastError());
    return FLBP_OK;
}

int Plan::vm_current() {
    if (vm_stale && E)
        hipLaunchKernelGGL(scatter_pairs, dim3(grid_for(E, 256)), dim3(256), 0, stream, m_fv, fac_vpos, m_fv_vm, E);
    vm_stale = false;
    HIP_TRY(hipGetLastError());
    return FLBP_OK;
}

int Plan::launch_sweep(int k, double tol, double damping) {
    if (seq) {   /* a sequential schedule updates the factor-major f->v array only */
        vm_stale = true;
        return launch_seq_sweep(k, tol);
    }
    if (variant == 3) {
        /* undamped sweeps of the scatter layout skip the factor-major f->v
         * copy; a damped sweep first rebuilds it from the var-major array */
        if (damping != 0.0 && fm_stale && E) {
            hipLaunchKernelGGL(gather_pairs, dim3(grid_for(E, 256)), dim3(256), 0, stream, m_fv_vm, fac_vpos,
                               m_fv, E);
            fm_stale = false;
        }
        if (damping == 0.0) fm_stale = true;
    }
    return launch_sweep_on(k, tol, damping, m_vf, m_fv, m_fv, m_vf, belief);
}

/* One sweep reading v->f messages from vf_in and the previous f->v
 * messages from fv_old, writing f->v messages to fv_out and v->f messages
 * to vf_out (the plain solve passes the same arrays in and out). */
int Plan::launch_sweep_on(int k, double tol, double damping, const double2 *vf_in,
                          const double2 *fv_old, double2 *fv_out, double2 *vf_out, double *bel) {
    const double keep = 1.0 - damping;
    const bool narrow = max_fac_deg <= (uint64_t)kFacRegDegSmall;
    static_assert(kStageItems == (uint32_t)kBlock && kStageSlots == (uint32_t)kStage, "staged block shape");
    if (variant == 4) {
        /* the state layout works in place on the plan's own arrays: the
         * factor-major f->v messages and the per-variable products */
        if (vf_in != m_vf || fv_old != m_fv || fv_out != m_fv || vf_out != m_vf)
            return fail(FLBP_EINVAL, "plan: the state layout sweeps the plan's own message arrays only");
        if (F) {
            const dim3 grid((unsigned)n_fac_blocks), block(kBlock);
#define FLBP_FACTOR_STATE(DAMP, N, MODE)                                                                         \
    hipLaunchKernelGGL((factor_pass_state<DAMP, N, MODE>), grid, block, 0, stream, fac_blk, fac_off, fac_meta,     \
                       fac_p, fac_q, fac_mask, fac_var, fac_work, vprod, vzc, m_vf, m_fv, m_vf, init_value, damping, \
                       keep, ctl, k, tol)
#define FLBP_FACTOR_STATE_M(DAMP, N)                                       \
    do {                                                                   \
        if (vf_src == kVfState) FLBP_FACTOR_STATE(DAMP, N, kVfState);      \
        else if (vf_src == kVfConst) FLBP_FACTOR_STATE(DAMP, N, kVfConst); \
        else FLBP_FACTOR_STATE(DAMP, N, kVfArray);                         \
    } while (0)
            if (damping != 0.0) {
                if (narrow) FLBP_FACTOR_STATE_M(true, kFacRegDegSmall);
                else FLBP_FACTOR_STATE_M(true, kFacRegDeg);
            } else {
                if (narrow) FLBP_FACTOR_STATE_M(false, kFacRegDegSmall);
                else FLBP_FACTOR_STATE_M(false, kFacRegDeg);
            }
#undef FLBP_FACTOR_STATE_M
#undef FLBP_FACTOR_STATE
        }
        if (profiling) HIP_TRY(hipEventRecord(ev[2 * k + 1], stream));
        if (V)
            hipLaunchKernelGGL(var_pass_state, dim3((unsigned)n_var_blocks4), dim3(kBlock), 0, stream, var_blk4,
                               var_off, var_src, var_work4, m_fv, vprod, vzc, bel, ctl, k, tol);
        if (profiling) HIP_TRY(hipEventRecord(ev[2 * k + 2], stream));
        HIP_TRY(hipGetLastError());
        /* from the second sweep on, the v->f inputs come from the products
         * the var pass just wrote (a sweep the stop test skips changes nothing) */
        vf_src = kVfState;
        return FLBP_OK;
    }
    if (F) {
        if (variant != 1) {
            const dim3 grid((unsigned)n_fac_blocks), block(kBlock);
#define FLBP_FACTOR_STAGED(DAMP, N, LAYOUT)                                                                    \
    hipLaunchKernelGGL((factor_pass_staged<DAMP, N, LAYOUT>), grid, block, 0, stream, fac_blk, fac_off, fac_meta, \
                       fac_p, fac_q, fac_mask, fac_vpos, fac_work, vf_in, fv_old, fv_out, m_fv_vm, damping, keep, \
                       ctl, k, tol)
#define FLBP_FACTOR_STAGED_N(DAMP, N)                                                   \
    do {                                                                                \
        if (variant == 0) FLBP_FACTOR_STAGED(DAMP, N, kLayoutPull);                     \
        else if (variant == 2) FLBP_FACTOR_STAGED(DAMP, N, kLayoutPush);                \
        else FLBP_FACTOR_STAGED(DAMP, N, kLayoutScatter);                               \
    } while (0)
            if (damping != 0.0) {
                if (narrow) FLBP_FACTOR_STAGED_N(true, kFacRegDegSmall);
                else FLBP_FACTOR_STAGED_N(true, kFacRegDeg);
            } else {
                if (narrow) FLBP_FACTOR_STAGED_N(false, kFacRegDegSmall);
                else FLBP_FACTOR_STAGED_N(false, kFacRegDeg);
            }
#undef FLBP_FACTOR_STAGED_N
#undef FLBP_FACTOR_STAGED
        } else {
            const dim3 grid(grid_for(F, kBlock)), block(kBlock);
#define FLBP_FACTOR_PASS(DAMP, N)                                                                     \
    hipLaunchKernelGGL((factor_pass<DAMP, N>), grid, block, 0, stream, fac_off, fac_meta, fac_p, fac_q, \
                       fac_mask, fac_vpos, vf_in, fv_old, fv_out, (uint32_t)F, damping, keep, ctl, k, tol)
            if (damping != 0.0) {
                if (narrow) FLBP_FACTOR_PASS(true, kFacRegDegSmall);
                else FLBP_FACTOR_PASS(true, kFacRegDeg);
            } else {
                if (narrow) FLBP_FACTOR_PASS(false, kFacRegDegSmall);
                else FLBP_FACTOR_PASS(false, kFacRegDeg);
            }
#undef FLBP_FACTOR_PASS
        }
    }
    if (profiling) HIP_TRY(hipEventRecord(ev[2 * k + 1], stream));
    if (V) {
        const dim3 vgrid((unsigned)n_var_blocks), vblock(kBlock);
        if (variant == 0)
            hipLaunchKernelGGL(var_pass_staged<kLayoutPull>, vgrid, vblock, 0, stream, var_blk, var_off, var_edge,
                               var_work, fv_out, vf_out, bel, ctl, k, tol);
        else if (variant == 2)
            hipLaunchKernelGGL(var_pass_staged<kLayoutPush>, vgrid, vblock, 0, stream, var_blk, var_off, var_edge,
                               var_work, fv_out, vf_out, bel, ctl, k, tol);
        else if (variant == 3)
            hipLaunchKernelGGL(var_pass_staged<kLayoutScatter>, vgrid, vblock, 0, stream, var_blk, var_off,
                               var_edge, var_work, m_fv_vm, vf_out, bel, ctl, k, tol);
        else
            hipLaunchKernelGGL(var_pass, dim3(grid_for(V, kBlock)), dim3(kBlock), 0, stream, var_off, var_edge,
                               fv_out, vf_out, bel, (uint32_t)V, ctl, k, tol);
    }
    if (profiling) HIP_TRY(hipEventRecord(ev[2 * k + 2], stream));
    HIP_TRY(hipGetLastError());
    return FLBP_OK;
}

/* Size of the next chunk of a run: sweeps that follow the stop test only
 * exit at once, but each still costs a launch, so once the max belief change
 * falls geometrically the chunk is cut to the sweeps that the decay rate
 * says are left (plus two), else it doubles up to kMaxChunk.  The stop test
 * is evaluated per sweep on the device, so the chunking changes no result. */
static int next_chunk(int chunk, double d_prev, double d_last, double tol) {
    const int grow = std::min(chunk * 2, flbp::kMaxChunk);
    if (d_last > tol && d_last < d_prev && d_last > 0 && std::isfinite(d_prev) && tol > 0) {
        const double est = std::log(tol / d_last) / std::log(d_last / d_prev);
        if (est >= 0 && est < flbp::kMaxChunk) return std::max(4, std::min((int)std::ceil(est) + 2, grow));
    }
    return grow;
}

/* Enqueue n sweeps (n <= kMaxChunk) and return how many ran before the stop
 * test fired, with the max diff of the last one that ran. */
void Plan::drop_graphs() {
    for (auto &g : graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    graphs.clear();
}

/* the captured chunk of n sweeps: clear, the sweeps, compact (a stale
 * factor-major copy is rebuilt before the replay, outside the graph) */
int Plan::chunk_graph(int n, double tol, double damping, hipGraphExec_t &exec) {
    for (auto &g : graphs)
        if (g.n == n && g.variant == variant && g.vf_src == vf_src && g.damping == damping && g.tol == tol) {
            exec = g.exec;
            if (variant == 4) vf_src = kVfState;   /* as the captured launches leave it */
            return FLBP_OK;
        }
    if (graphs.size() >= 2 * (size_t)kMaxChunk) drop_graphs();
    const int vf_src0 = vf_src;
    HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(clear_u64, dim3(16), dim3(256), 0, stream, ctl, kMaxChunk * kCtlPerSweep);
    int rc = FLBP_OK;
    for (int k = 0; k < n && !rc; k++)
        rc = seq ? launch_seq_sweep(k, tol) : launch_sweep_on(k, tol, damping, m_vf, m_fv, m_fv, m_vf, belief);
    hipLaunchKernelGGL(ctl_compact, dim3(1), dim3(64), 0, stream, ctl, ctl_out, n);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(ei, "hipGraphInstantiate");
    }
    graphs.push_back({n, variant, vf_src0, damping, tol, g, x});
    exec = x;
    return FLBP_OK;
}

int Plan::run_chunk(int n, double tol, double damping, int &ran, double &last_diff) {
    if (!profiling && use_graphs) {
        if (seq) {
            vm_stale = true;
        } else if (variant == 3) {   /* the scatter layout's damping copy, as launch_sweep keeps it */
            if (damping != 0.0 && fm_stale && E) {
                hipLaunchKernelGGL(gather_pairs, dim3(grid_for(E, 256)), dim3(256), 0, stream, m_fv_vm, fac_vpos,
                                   m_fv, E);
                fm_stale = false;
            }
            if (damping == 0.0) fm_stale = true;
        }
        hipGraphExec_t exec = nullptr;
        if (int rc = chunk_graph(n, tol, damping, exec)) return rc;
        HIP_TRY(hipGraphLaunch(exec, stream));
    } else {
        hipLaunchKernelGGL(clear_u64, dim3(16), dim3(256), 0, stream, ctl, kMaxChunk * kCtlPerSweep);
        if (profiling) HIP_TRY(hipEventRecord(ev[0], stream));
        for (int k = 0; k < n; k++) {
            int rc = launch_sweep(k, tol, damping);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(ctl_compact, dim3(1), dim3(64), 0, stream, ctl, ctl_out, n);
    }
    HIP_TRY(hipMemcpyAsync(host_ctl, ctl_out, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    ran = n;
    for (int k = 0; k < n; k++) {
        double dk;
        std::memcpy(&dk, &host_ctl[k], sizeof dk);
        prev_diff = last_diff;
        last_diff = dk;
        if (!(dk > tol) && tol != -INFINITY) {   /* -inf: fixed sweeps, none stops */
            ran = k + 1;
            break;
        }
    }
    if (profiling) {
        for (int k = 0; k < ran; k++) {
            float a = 0, b = 0;
            HIP_TRY(hipEventElapsedTime(&a, ev[2 * k], ev[2 * k + 1]));
            HIP_TRY(hipEventElapsedTime(&b, ev[2 * k + 1], ev[2 * k + 2]));
            prof_fac_ms += a;
            prof_var_ms += b;
        }
        prof_fac_n += F ? ran : 0;
        prof_var_n += V ? ran : 0;
    }
    return FLBP_OK;
}

Plan::~Plan() {
    if (device >= 0) (void)hipSetDevice(device);
    drop_graphs();
    free_seq_items();
    for (void *p : allocs) (void)hipFree(p);
    if (stage_idx) (void)hipFree(stage_idx);
    if (stage_val) (void)hipFree(stage_val);
    if (host_ctl) (void)hipHostFree(host_ctl);
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

}  // namespace flbp

using flbp::Plan;
using flbp::hip_fail;
using flbp::grid_for;

extern "C" {

int flbp_plan_create(const flbp_csr *c, int device, flbp_plan **out) {
    if (!c || !out) return flbp::fail(FLBP_EINVAL, "flbp_plan_create: null argument");
    const uint64_t V = c->n_vars, F = c->n_factors, E = c->n_edges;
    if (V >= (1ull << 32) - 1 || F >= (1ull << 32) - 1 || E >= (1ull << 32) - 1)
        return flbp::fail(FLBP_ERANGE, "flbp_plan_create: V, F and E must each be below 2^32-1");
    int rc = flbp::validate_csr(*c);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return flbp::fail(FLBP_ENODEV, "flbp_plan_create: no HIP device");
    if (device < 0 || device >= ndev) return flbp::fail(FLBP_EINVAL, "flbp_plan_create: bad device");
    HIP_TRY(hipSetDevice(device));
    auto *P = new (std::nothrow) Plan();
    if (!P) return flbp::fail(FLBP_ENOMEM, "flbp_plan_create: out of memory");
    P->device = device;
    P->V = V;
    P->F = F;
    P->E = E;
    rc = P->upload(*c);
    if (rc) {
        delete P;
        return rc;
    }
    *out = reinterpret_cast<flbp_plan *>(P);
    return FLBP_OK;
}

void flbp_plan_destroy(flbp_plan *plan) { delete reinterpret_cast<Plan *>(plan); }

int flbp_plan_init(flbp_plan *plan, int logdomain) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P) return flbp::fail(FLBP_EINVAL, "flbp_plan_init: null plan");
    HIP_TRY(hipSetDevice(P->device));
    const double c = logdomain ? 0.0 : 1.0;
    if (P->E) {
        /* the state layout keeps only the f->v messages: its first factor
         * pass takes the init value as its v->f input (vf_src), and the
         * other arrays are written whenever the plan leaves the layout
         * (flbp_plan_set_variant moves the messages) */
        if (P->variant != 4)
            hipLaunchKernelGGL(fill_pairs, dim3(grid_for(P->E, 256)), dim3(256), 0, P->stream, P->m_vf, P->E, c);
        hipLaunchKernelGGL(fill_pairs, dim3(grid_for(P->E, 256)), dim3(256), 0, P->stream, P->m_fv, P->E, c);
        if (P->variant != 4)
            hipLaunchKernelGGL(fill_pairs, dim3(grid_for(P->E, 256)), dim3(256), 0, P->stream, P->m_fv_vm, P->E, c);
    }
    if (int rc = flbp::seq_reset_state(P, c)) return rc;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(P->stream));
    P->fm_stale = false;
    P->vm_stale = false;
    P->init_value = c;
    P->vf_src = kVfConst;
    P->swept_since_init = false;
    return FLBP_OK;
}

int flbp_plan_run(flbp_plan *plan, const flbp_run_opts *o, flbp_run_result *res) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P || !o) return flbp::fail(FLBP_EINVAL, "flbp_plan_run: null argument");
    if (!(o->damping >= 0.0 && o->damping < 1.0))
        return flbp::fail(FLBP_EINVAL, "flbp_plan_run: damping must be in [0, 1)");
    if (P->seq && o->damping != 0.0)
        return flbp::fail(FLBP_EINVAL, "flbp_plan_run: damping is defined for PARALL updates only");
    HIP_TRY(hipSetDevice(P->device));
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&] {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    P->reset_profile();
    double maxdiff = INFINITY;
    uint64_t sweeps = 0;
    int chunk = o->chunk ? (int)std::min<uint32_t>(o->chunk, flbp::kMaxChunk) : 4;
    const bool adaptive = o->chunk == 0;
    while (P->iters < o->maxiter && maxdiff > o->tol && elapsed() < o->maxtime) {
        const int n = (int)std::min<uint64_t>((uint64_t)chunk, o->maxiter - P->iters);
        int ran = 0;
        double last = maxdiff;
        P->prev_diff = INFINITY;
        int rc = P->run_chunk(n, o->tol, o->damping, ran, last);
        if (rc) return rc;
        P->iters += ran;
        sweeps += ran;
        maxdiff = last;
        P->swept_since_init = true;
        if (adaptive) chunk = flbp::next_chunk(chunk, P->prev_diff, maxdiff, o->tol);
    }
    P->last_maxdiff = maxdiff;
    if (res) {
        res->iters = P->iters;
        res->sweeps = sweeps;
        res->maxdiff = maxdiff;
        res->seconds = elapsed();
    }
    return FLBP_OK;
}

int flbp_plan_run_fixed(flbp_plan *plan, uint64_t n_sweeps, double damping, flbp_run_result *res) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P) return flbp::fail(FLBP_EINVAL, "flbp_plan_run_fixed: null plan");
    if (!(damping >= 0.0 && damping < 1.0))
        return flbp::fail(FLBP_EINVAL, "flbp_plan_run_fixed: damping must be in [0, 1)");
    if (P->seq && damping != 0.0)
        return flbp::fail(FLBP_EINVAL, "flbp_plan_run_fixed: damping is defined for PARALL updates only");
    HIP_TRY(hipSetDevice(P->device));
    const auto t0 = std::chrono::steady_clock::now();
    P->reset_profile();
    uint64_t done = 0;
    double last = INFINITY;
    while (done < n_sweeps) {
        const int n = (int)std::min<uint64_t>(flbp::kMaxChunk, n_sweeps - done);
        int ran = 0;
        /* tol = -inf: the stop test never fires */
        int rc = P->run_chunk(n, -INFINITY, damping, ran, last);
        if (rc) return rc;
        done += ran;
        P->swept_since_init = true;
    }
    if (n_sweeps == 0 && !P->swept_since_init && P->V) {
        /* run(tol, 0, hist) still computes beliefs from the init products:
         * (c, c) with no zeros, i.e. c / (c + c) */
        const double c = P->init_value;
        const double b = c / (c + c);
        hipLaunchKernelGGL(fill_f64, dim3(grid_for(P->V, 256)), dim3(256), 0, P->stream, P->belief, P->V, b);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(P->stream));
    }
    P->iters += n_sweeps;
    P->last_maxdiff = last;
    if (res) {
        res->iters = P->iters;
        res->sweeps = n_sweeps;
        res->maxdiff = last;
        res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return FLBP_OK;
}

int flbp_plan_beliefs(flbp_plan *plan, double *out, uint64_t n) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P || (n && !out)) return flbp::fail(FLBP_EINVAL, "flbp_plan_beliefs: null argument");
    if (n > P->V) return flbp::fail(FLBP_EINVAL, "flbp_plan_beliefs: n > n_vars");
    HIP_TRY(hipSetDevice(P->device));
    if (n) {
        std::vector<double> dev(P->V);
        HIP_TRY(hipMemcpyAsync(dev.data(), P->belief, P->V * sizeof(double), hipMemcpyDeviceToHost, P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
        for (uint64_t dv = 0; dv < P->V; dv++)
            if (P->h_perm[dv] < n) out[P->h_perm[dv]] = dev[dv];
    }
    HIP_TRY(hipStreamSynchronize(P->stream));
    return FLBP_OK;
}

int flbp_plan_beliefs_at(flbp_plan *plan, uint64_t n, const uint64_t *vars, double *out) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P || (n && (!vars || !out))) return flbp::fail(FLBP_EINVAL, "flbp_plan_beliefs_at: null argument");
    if (!n) return FLBP_OK;
    std::vector<uint32_t> idx(n);
    for (uint64_t i = 0; i < n; i++) {
        if (vars[i] >= P->V) return flbp::fail(FLBP_EINVAL, "flbp_plan_beliefs_at: variable out of range");
        idx[i] = P->h_inv_perm[vars[i]];
    }
    HIP_TRY(hipSetDevice(P->device));
    if (int rc = P->stage(n)) return rc;
    HIP_TRY(hipMemcpyAsync(P->stage_idx, idx.data(), n * 4, hipMemcpyHostToDevice, P->stream));
    hipLaunchKernelGGL(gather_f64, dim3(grid_for(n, 256)), dim3(256), 0, P->stream, P->belief, P->stage_idx,
                       P->stage_val, n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, P->stage_val, n * 8, hipMemcpyDeviceToHost, P->stream));
    HIP_TRY(hipStreamSynchronize(P->stream));
    return FLBP_OK;
}

int flbp_plan_reset(flbp_plan *plan) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P) return flbp::fail(FLBP_EINVAL, "flbp_plan_reset: null plan");
    HIP_TRY(hipSetDevice(P->device));
    if (P->V) hipLaunchKernelGGL(fill_f64, dim3(grid_for(P->V, 256)), dim3(256), 0, P->stream, P->belief, P->V, 0.0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(P->stream));
    P->iters = 0;
    P->last_maxdiff = INFINITY;
    P->swept_since_init = false;
    return FLBP_OK;
}

int flbp_plan_get_messages(flbp_plan *plan, double *fv_varmajor, double *vf_facmajor) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P) return flbp::fail(FLBP_EINVAL, "flbp_plan_get_messages: null plan");
    HIP_TRY(hipSetDevice(P->device));
    std::vector<double> tmp(2 * P->E);
    if (fv_varmajor && P->E) {
        /* factor-major, in the state layout's storage order (variant 4), or
         * var-major in device variable order (variant 3, unless a schedule's
         * sweeps left only the factor-major array current) */
        const bool vm = P->variant == 3 && !P->vm_stale, st = P->variant == 4;
        HIP_TRY(hipMemcpy(tmp.data(), vm ? P->m_fv_vm : P->m_fv, P->E * sizeof(double2), hipMemcpyDeviceToHost));
        for (uint64_t k = 0; k < P->E; k++) {
            const uint64_t e = P->h_var_edge[k];
            const uint64_t s = vm ? P->h_fac_vpos[e] : st ? P->h_store[e] : e;
            fv_varmajor[2 * k] = tmp[2 * s];
            fv_varmajor[2 * k + 1] = tmp[2 * s + 1];
        }
    }
    if (vf_facmajor && P->E) {   /* stored var-major, factor-major in the push and scatter layouts, or in the
                                    state layout's storage order */
        if (P->variant == 4)      /* derived from the products: materialized first */
            if (int rc = P->materialize_vf()) return rc;
        HIP_TRY(hipMemcpy(tmp.data(), P->m_vf, P->E * sizeof(double2), hipMemcpyDeviceToHost));
        for (uint64_t r = 0; r < P->E; r++) {   /* reference slot r, device slot s */
            const uint64_t k = P->vf_pos(P->h_slot_dev[r]);
            vf_facmajor[2 * r] = tmp[2 * k];
            vf_facmajor[2 * r + 1] = tmp[2 * k + 1];
        }
    }
    return FLBP_OK;
}

int flbp_plan_set_messages(flbp_plan *plan, const double *fv_varmajor, const double *vf_facmajor) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P) return flbp::fail(FLBP_EINVAL, "flbp_plan_set_messages: null plan");
    if (P->seq && fv_varmajor)
        return flbp::fail(FLBP_EINVAL, "flbp_plan_set_messages: a sequential schedule keeps its products in step "
                                       "with the f->v messages; remove the schedule before setting them");
    HIP_TRY(hipSetDevice(P->device));
    /* the state layout derives v->f from the f->v messages: keep the current
     * v->f messages explicitly before the f->v messages change under them */
    if (P->variant == 4 && fv_varmajor && !vf_facmajor)
        if (int rc = P->materialize_vf()) return rc;
    if (fv_varmajor && P->E) {
        std::vector<double> tmp(2 * P->E);
        for (uint64_t k = 0; k < P->E; k++) {
            const uint64_t s = P->variant == 4 ? P->h_store[P->h_var_edge[k]] : P->h_var_edge[k];
            tmp[2 * s] = fv_varmajor[2 * k];
            tmp[2 * s + 1] = fv_varmajor[2 * k + 1];
        }
        HIP_TRY(hipMemcpy(P->m_fv, tmp.data(), P->E * sizeof(double2), hipMemcpyHostToDevice));
        for (uint64_t k = 0; k < P->E; k++) {   /* the var-major copy, device variable order */
            const uint64_t d = P->h_fac_vpos[P->h_var_edge[k]];
            tmp[2 * d] = fv_varmajor[2 * k];
            tmp[2 * d + 1] = fv_varmajor[2 * k + 1];
        }
        HIP_TRY(hipMemcpy(P->m_fv_vm, tmp.data(), P->E * sizeof(double2), hipMemcpyHostToDevice));
        P->fm_stale = false;
        P->vm_stale = false;
    }
    if (vf_facmajor && P->E) {
        std::vector<double> tmp(2 * P->E);
        for (uint64_t r = 0; r < P->E; r++) {   /* reference slot r, device slot s */
            const uint64_t k = P->vf_pos(P->h_slot_dev[r]);
            tmp[2 * k] = vf_facmajor[2 * r];
            tmp[2 * k + 1] = vf_facmajor[2 * r + 1];
        }
        HIP_TRY(hipMemcpy(P->m_vf, tmp.data(), P->E * sizeof(double2), hipMemcpyHostToDevice));
        if (P->variant == 4) P->vf_src = kVfArray;
    }
    return FLBP_OK;
}

int flbp_plan_set_p(flbp_plan *plan, const double *p, uint64_t n) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P || (n && !p)) return flbp::fail(FLBP_EINVAL, "flbp_plan_set_p: null argument");
    if (n != P->F) return flbp::fail(FLBP_EINVAL, "flbp_plan_set_p: n must equal n_factors");
    HIP_TRY(hipSetDevice(P->device));
    if (n) {
        std::vector<double> pd(n);   /* device factor order */
        for (uint64_t d = 0; d < n; d++) pd[d] = p[P->h_fperm[d]];
        HIP_TRY(hipMemcpy(P->fac_p, pd.data(), n * sizeof(double), hipMemcpyHostToDevice));
    }
    if (int rc = flbp::seq_refresh_factors(P)) return rc;
    HIP_TRY(hipStreamSynchronize(P->stream));
    return FLBP_OK;
}

}  // extern "C"

/* head clamps of device factors dev[0..n) (flbp_plan_set_head_clamps and
 * flbp_plan_clamp_vars); ref names each factor by its reference index in
 * error messages */
static int set_head_clamps_dev(Plan *P, uint64_t n, const uint64_t *dev, const uint64_t *ref, const uint8_t *values,
                               const char *who) {
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t I = dev[i];
        if ((P->h_meta[I] & 3u) == 0)
            return flbp::fail(FLBP_EINVAL, std::string(who) + ": a prior factor is clamped through its p "
                                                              "(rebuild the graph with flbp_clamp)");
        if (P->h_var_body[P->h_fac_head[I]])
            return flbp::fail(FLBP_EINVAL, std::string(who) + ": the head of factor " + std::to_string(ref[i]) +
                                               " appears in a factor body; its clamp changes the structure "
                                               "(rebuild the graph with flbp_clamp)");
        if (values[i] > 1) return flbp::fail(FLBP_EINVAL, std::string(who) + ": values are 0 or 1");
    }
    HIP_TRY(hipSetDevice(P->device));
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t I = dev[i];
        const unsigned x = values[i];
        double m0 = x == 0 ? 1.0 : 0.0, m1 = x == 1 ? 1.0 : 0.0;
        if (P->h_meta[I] & 4u) {   /* a second clamp of the head multiplies */
            m0 *= P->h_mask[I].x;
            m1 *= P->h_mask[I].y;
        }
        P->h_meta[I] |= 4u;
        P->h_mask[I] = make_double2(m0, m1);
        P->clamp_dirty.push_back(I);
        HIP_TRY(hipMemcpyAsync(P->fac_meta + I, &P->h_meta[I], 4, hipMemcpyHostToDevice, P->stream));
        HIP_TRY(hipMemcpyAsync(P->fac_mask + I, &P->h_mask[I], 16, hipMemcpyHostToDevice, P->stream));
    }
    if (int rc = flbp::seq_refresh_factors(P)) return rc;
    HIP_TRY(hipStreamSynchronize(P->stream));
    return FLBP_OK;
}

extern "C" {

int flbp_plan_set_head_clamps(flbp_plan *plan, uint64_t n, const uint64_t *factors, const uint8_t *values) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P || (n && (!factors || !values))) return flbp::fail(FLBP_EINVAL, "flbp_plan_set_head_clamps: null argument");
    std::vector<uint64_t> dev(n);
    for (uint64_t i = 0; i < n; i++) {
        if (factors[i] >= P->F) return flbp::fail(FLBP_EINVAL, "flbp_plan_set_head_clamps: factor out of range");
        dev[i] = P->h_finv[factors[i]];
    }
    return set_head_clamps_dev(P, n, dev.data(), factors, values, "flbp_plan_set_head_clamps");
}

int flbp_plan_clamp_vars(flbp_plan *plan, uint64_t n, const uint64_t *vars, const uint8_t *values) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P || (n && (!vars || !values))) return flbp::fail(FLBP_EINVAL, "flbp_plan_clamp_vars: null argument");
    std::vector<uint64_t> fac;
    std::vector<uint8_t> val;
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t v = vars[i];
        if (v >= P->V) return flbp::fail(FLBP_EINVAL, "flbp_plan_clamp_vars: variable out of range");
        if (values[i] > 1) return flbp::fail(FLBP_EINVAL, "flbp_plan_clamp_vars: values are 0 or 1");
        if (P->h_var_body[v])
            return flbp::fail(FLBP_EINVAL, "flbp_plan_clamp_vars: variable " + std::to_string(v) +
                                               " appears in a factor body; its clamp changes the structure "
                                               "(rebuild the graph with flbp_clamp)");
        if (P->h_head_off[v] == P->h_head_off[v + 1])
            return flbp::fail(FLBP_EINVAL, "flbp_plan_clamp_vars: variable " + std::to_string(v) + " heads no factor");
        for (uint32_t k = P->h_head_off[v]; k < P->h_head_off[v + 1]; k++) {
            fac.push_back(P->h_head_fac[k]);   /* device factor */
            val.push_back(values[i]);
        }
    }
    std::vector<uint64_t> ref(fac.size());
    for (size_t i = 0; i < fac.size(); i++) ref[i] = P->h_fperm[fac[i]];
    return set_head_clamps_dev(P, fac.size(), fac.data(), ref.data(), val.data(), "flbp_plan_clamp_vars");
}

int flbp_plan_reset_head_clamps(flbp_plan *plan) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P) return flbp::fail(FLBP_EINVAL, "flbp_plan_reset_head_clamps: null plan");
    HIP_TRY(hipSetDevice(P->device));
    for (uint64_t I : P->clamp_dirty) {
        P->h_meta[I] = P->h_meta0[I];
        P->h_mask[I] = P->h_mask0[I];
        HIP_TRY(hipMemcpyAsync(P->fac_meta + I, &P->h_meta[I], 4, hipMemcpyHostToDevice, P->stream));
        HIP_TRY(hipMemcpyAsync(P->fac_mask + I, &P->h_mask[I], 16, hipMemcpyHostToDevice, P->stream));
    }
    if (int rc = flbp::seq_refresh_factors(P)) return rc;
    HIP_TRY(hipStreamSynchronize(P->stream));
    P->clamp_dirty.clear();
    return FLBP_OK;
}

int flbp_plan_use_graphs(flbp_plan *plan, int enable) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P) return flbp::fail(FLBP_EINVAL, "flbp_plan_use_graphs: null plan");
    HIP_TRY(hipSetDevice(P->device));
    P->use_graphs = enable != 0;
    if (!P->use_graphs) P->drop_graphs();
    return FLBP_OK;
}

int flbp_plan_profile(flbp_plan *plan, int enable) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P) return flbp::fail(FLBP_EINVAL, "flbp_plan_profile: null plan");
    HIP_TRY(hipSetDevice(P->device));
    if (enable && !P->ev[0]) {
        for (auto &e : P->ev) HIP_TRY(hipEventCreate(&e));
    }
    P->profiling = enable != 0;
    return FLBP_OK;
}

int flbp_plan_last_timing(flbp_plan *plan, double *fac_ms, double *var_ms, uint64_t *fac_n,
                          uint64_t *var_n) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P) return flbp::fail(FLBP_EINVAL, "flbp_plan_last_timing: null plan");
    if (fac_ms) *fac_ms = P->prof_fac_ms;
    if (var_ms) *var_ms = P->prof_var_ms;
    if (fac_n) *fac_n = P->prof_fac_n;
    if (var_n) *var_n = P->prof_var_n;
    return FLBP_OK;
}

int flbp_plan_get_variant(const flbp_plan *plan) {
    const Plan *P = reinterpret_cast<const Plan *>(plan);
    return P ? P->variant : flbp::fail(FLBP_EINVAL, "flbp_plan_get_variant: null plan");
}

int flbp_plan_set_variant(flbp_plan *plan, int variant) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P || variant < 0 || variant > 4) return flbp::fail(FLBP_EINVAL, "flbp_plan_set_variant: bad argument");
    if (P->seq && variant != 3)
        return flbp::fail(FLBP_EINVAL, "flbp_plan_set_variant: a plan with a sequential schedule keeps variant 3");
    auto layout = [](int v) { return v == 1 ? 0 : v; };
    if (layout(P->variant) != layout(variant) && P->E) {
        /* move the stored messages to the new layout through their
         * reference-order form */
        std::vector<double> fv(2 * P->E), vf(2 * P->E);
        int rc = flbp_plan_get_messages(plan, fv.data(), vf.data());
        if (rc) return rc;
        P->variant = variant;
        return flbp_plan_set_messages(plan, fv.data(), vf.data());
    }
    P->variant = variant;
    return FLBP_OK;
}

uint64_t flbp_plan_device_bytes(const flbp_plan *plan) {
    const Plan *P = reinterpret_cast<const Plan *>(plan);
    return P ? P->device_bytes : 0;
}

}  // extern "C"
