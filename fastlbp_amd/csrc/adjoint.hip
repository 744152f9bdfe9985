/*
 * adjoint.hip -- checkpointed reverse-mode gradient of the damped PARALL
 * solve with respect to the factors' probabilities p.
 *
 * Forward: init, then n sweeps on the plan's buffers; before sweep c*k+1
 * the message state (m_vf, m_fv) is copied to checkpoint c.
 * Backward: for segments c = last..0, recompute sweeps c*k+1..c*k+k from
 * checkpoint c into per-sweep state slots (out of place), then walk the
 * segment backwards with var_vjp_staged and fac_vjp_staged
 * (adjoint_kernels.h) over the adjoint's own block ranges (at most
 * kAdjStage slots each) and work orders.  Device memory (m_vf and the
 * adjoint of m_fv var-major, m_fv and the adjoint of m_vf factor-major, so
 * every kernel gathers and writes contiguously): ceil(n/k) checkpoints and
 * k state slots of 32 B per edge each, three 16 B-per-edge adjoint arrays,
 * and a 64 B-per-edge scratch row when some variable has more than
 * kAdjStage edges.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "adjoint_kernels.h"
#include "fastlbp.h"
#include "flbp_internal.h"
#include "plan_internal.h"

namespace {

__global__ void fill2(double2 *p, uint64_t n, double c) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_double2(c, c);
}

__global__ void fill1(double *p, uint64_t n, double c) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = c;
}

/* dst[idx[i]] = val[i] */
__global__ void scatter_f64(double *__restrict__ dst, const uint32_t *__restrict__ idx, const double *__restrict__ val,
                            uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[idx[i]] = val[i];
}

/* one block per rule r: out[r] = sum over the rule's factors f (ascending,
 * rule_fac[rule_off[r] .. rule_off[r+1])) of pbar[f] * p[f] * (1 - p[f]),
 * summed per lane in a fixed stride and then in a fixed tree: the same
 * bits on every run */
__global__ __launch_bounds__(256) void rule_reduce(const uint64_t *__restrict__ rule_off,
                                                  const uint32_t *__restrict__ rule_fac,
                                                  const double *__restrict__ pbar, const double *__restrict__ p,
                                                  double *__restrict__ out) {
    __shared__ double part[256];
    const uint32_t r = blockIdx.x;
    double acc = 0.0;
    for (uint64_t j = rule_off[r] + threadIdx.x; j < rule_off[r + 1]; j += 256) {
        const uint32_t f = rule_fac[j];
        acc += pbar[f] * p[f] * (1.0 - p[f]);
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[r] = part[0];
}

}  // namespace

namespace flbp {

struct Adjoint {
    Plan *P = nullptr;
    int device = -1;          /* kept here: the plan may be destroyed first */
    uint64_t T = 0;
    uint32_t k = 1;
    uint32_t nck = 0;
    double damp = 0.0;
    std::vector<double2 *> ck_vf, ck_fv, st_vf, st_fv;
    double2 *avf = nullptr, *avf2 = nullptr, *afv = nullptr, *gscratch = nullptr;
    uint32_t *fac_blk = nullptr, *var_blk = nullptr;    /* block ranges of the reverse kernels */
    uint16_t *fac_work = nullptr, *var_work = nullptr;  /* per block: factors by type and degree, variables by degree */
    uint64_t n_fac_blocks = 0, n_var_blocks = 0;
    uint64_t *rule_off = nullptr;   /* per rule: its factors' range in rule_fac (flbp_adjoint_set_rules) */
    uint32_t *rule_fac = nullptr;
    double *rule_out = nullptr;
    uint64_t n_rules = 0;
    int run_backward();
    double *pbar = nullptr, *bbar = nullptr, *scratch_belief = nullptr;
    std::vector<void *> allocs;
    uint64_t bytes = 0;
    bool have_forward = false;
    hipEvent_t ev[4] = {};
    double fwd_ms = 0, bwd_ms = 0;

    template <class T_>
    int alloc(T_ *&p, uint64_t n) {
        void *q = nullptr;
        const size_t b = std::max<uint64_t>(n, 1) * sizeof(T_);
        hipError_t e = hipMalloc(&q, b);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(FLBP_ENOMEM, "adjoint: hipMalloc of " + std::to_string(b) + " bytes failed");
        }
        allocs.push_back(q);
        bytes += b;
        p = static_cast<T_ *>(q);
        return FLBP_OK;
    }
    /* block ranges and work orders of the staged reverse kernels, from the
     * plan's device-order CSR */
    int build_blocks() {
        const uint64_t F = P->F, V = P->V;
        std::vector<uint32_t> fo(F + 1), meta(F), vo(V + 1);
        if (hipMemcpy(fo.data(), P->fac_off, (F + 1) * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            (F && hipMemcpy(meta.data(), P->fac_meta, F * 4, hipMemcpyDeviceToHost) != hipSuccess) ||
            hipMemcpy(vo.data(), P->var_off, (V + 1) * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return fail(FLBP_EHIP, "adjoint: reading the plan's offsets failed");
        std::vector<uint32_t> fb, vb;
        cut_blocks(fo.data(), F, kBlock, kAdjStage, fb);
        cut_blocks(vo.data(), V, kBlock, kAdjStage, vb);
        std::vector<uint16_t> fw(F), vw(V);
        std::vector<std::pair<uint64_t, uint16_t>> key;
        for (size_t b = 0; b + 1 < fb.size(); b++) {
            key.clear();
            for (uint32_t f = fb[b]; f < fb[b + 1]; f++) {
                const uint64_t d = fo[f + 1] - fo[f], prior = (meta[f] & 3u) == 0;
                key.push_back({(prior << 40) | (d << 20) | (f - fb[b]), (uint16_t)(f - fb[b])});
            }
            std::sort(key.begin(), key.end());
            for (size_t i = 0; i < key.size(); i++) fw[fb[b] + i] = key[i].second;
        }
        for (size_t b = 0; b + 1 < vb.size(); b++) {
            key.clear();
            for (uint32_t v = vb[b]; v < vb[b + 1]; v++)
                key.push_back({((uint64_t)(vo[v + 1] - vo[v]) << 20) | (v - vb[b]), (uint16_t)(v - vb[b])});
            std::sort(key.begin(), key.end());
            for (size_t i = 0; i < key.size(); i++) vw[vb[b] + i] = key[i].second;
        }
        n_fac_blocks = fb.size() - 1;
        n_var_blocks = vb.size() - 1;
        int rc = alloc(fac_blk, fb.size());
        if (!rc) rc = alloc(var_blk, vb.size());
        if (!rc) rc = alloc(fac_work, F);
        if (!rc) rc = alloc(var_work, V);
        if (rc) return rc;
        if (hipMemcpy(fac_blk, fb.data(), fb.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(var_blk, vb.data(), vb.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            (F && hipMemcpy(fac_work, fw.data(), F * 2, hipMemcpyHostToDevice) != hipSuccess) ||
            (V && hipMemcpy(var_work, vw.data(), V * 2, hipMemcpyHostToDevice) != hipSuccess))
            return fail(FLBP_EHIP, "adjoint: uploading block ranges failed");
        return FLBP_OK;
    }
    ~Adjoint() {
        if (device >= 0) (void)hipSetDevice(device);
        for (void *q : allocs) (void)hipFree(q);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

#define TRY(expr)            \
    do {                     \
        int rc_ = (expr);    \
        if (rc_) return rc_; \
    } while (0)
#define HIP_TRY(expr)                                     \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

}  // namespace flbp

using flbp::Adjoint;
using flbp::Plan;
using flbp::grid_for;
using flbp::hip_fail;

extern "C" {

int flbp_adjoint_create(flbp_plan *plan, uint64_t n_sweeps, uint32_t checkpoint_every, double damping,
                        flbp_adjoint **out) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P || !out || n_sweeps == 0 || checkpoint_every == 0)
        return flbp::fail(FLBP_EINVAL, "flbp_adjoint_create: bad arguments");
    if (!(damping >= 0.0 && damping < 1.0))
        return flbp::fail(FLBP_EINVAL, "flbp_adjoint_create: damping must be in [0, 1)");
    if (P->seq)
        return flbp::fail(FLBP_EINVAL, "flbp_adjoint_create: the adjoint differentiates PARALL sweeps, and the plan "
                                       "has a sequential schedule");
    if (P->variant >= 2) {   /* the adjoint walks the pull layout: move the plan's messages there */
        const int rc = flbp_plan_set_variant(plan, 0);
        if (rc) return rc;
    }
    HIP_TRY(hipSetDevice(P->device));
    auto *A = new (std::nothrow) Adjoint();
    if (!A) return flbp::fail(FLBP_ENOMEM, "flbp_adjoint_create: out of memory");
    A->P = P;
    A->device = P->device;
    A->T = n_sweeps;
    A->k = (uint32_t)std::min<uint64_t>(checkpoint_every, n_sweeps);
    A->nck = (uint32_t)((n_sweeps + A->k - 1) / A->k);
    A->damp = damping;
    const uint64_t E = P->E;
    int rc = FLBP_OK;
    A->ck_vf.resize(A->nck);
    A->ck_fv.resize(A->nck);
    A->st_vf.resize(A->k);
    A->st_fv.resize(A->k);
    for (uint32_t c = 0; c < A->nck && !rc; c++) {
        rc = A->alloc(A->ck_vf[c], E);
        if (!rc) rc = A->alloc(A->ck_fv[c], E);
    }
    for (uint32_t i = 0; i < A->k && !rc; i++) {
        rc = A->alloc(A->st_vf[i], E);
        if (!rc) rc = A->alloc(A->st_fv[i], E);
    }
    if (!rc) rc = A->alloc(A->avf, E);
    if (!rc) rc = A->alloc(A->avf2, E);
    if (!rc) rc = A->alloc(A->afv, E);
    /* scratch rows (4 pairs per slot) for variables wider than a staged
     * block and for factors above the exact-degree code */
    if (!rc)
        rc = A->alloc(A->gscratch,
                      P->max_var_deg > (uint64_t)kAdjStage || P->max_fac_deg > (uint64_t)kAdjFacDeg ? 4 * E : 1);
    if (!rc) rc = A->build_blocks();
    if (!rc) rc = A->alloc(A->pbar, P->F);
    if (!rc) rc = A->alloc(A->bbar, P->V);
    if (!rc) rc = A->alloc(A->scratch_belief, P->V);
    for (int i = 0; i < 4 && !rc; i++)
        if (hipEventCreate(&A->ev[i]) != hipSuccess) rc = flbp::fail(FLBP_EHIP, "hipEventCreate failed");
    if (rc) {
        delete A;
        return rc;
    }
    *out = reinterpret_cast<flbp_adjoint *>(A);
    return FLBP_OK;
}

void flbp_adjoint_destroy(flbp_adjoint *adj) { delete reinterpret_cast<Adjoint *>(adj); }

int flbp_adjoint_forward(flbp_adjoint *adj, double *beliefs) {
    Adjoint *A = reinterpret_cast<Adjoint *>(adj);
    if (!A) return flbp::fail(FLBP_EINVAL, "flbp_adjoint_forward: null adjoint");
    if (A->P->variant >= 2)
        return flbp::fail(FLBP_EINVAL, "flbp_adjoint_forward: the plan was switched away from the pull layout");
    Plan *P = A->P;
    HIP_TRY(hipSetDevice(P->device));
    const uint64_t E = P->E;
    hipStream_t st = P->stream;
    HIP_TRY(hipEventRecord(A->ev[0], st));
    if (E) {
        hipLaunchKernelGGL(fill2, dim3(grid_for(E, 256)), dim3(256), 0, st, P->m_vf, E, 1.0);
        hipLaunchKernelGGL(fill2, dim3(grid_for(E, 256)), dim3(256), 0, st, P->m_fv, E, 1.0);
    }
    for (uint64_t t = 1; t <= A->T; t++) {
        if ((t - 1) % A->k == 0 && E) {
            const uint64_t c = (t - 1) / A->k;
            HIP_TRY(hipMemcpyAsync(A->ck_vf[c], P->m_vf, E * sizeof(double2), hipMemcpyDeviceToDevice, st));
            HIP_TRY(hipMemcpyAsync(A->ck_fv[c], P->m_fv, E * sizeof(double2), hipMemcpyDeviceToDevice, st));
        }
        TRY(P->launch_sweep_on(0, -INFINITY, A->damp, P->m_vf, P->m_fv, P->m_fv, P->m_vf, P->belief));
    }
    HIP_TRY(hipEventRecord(A->ev[1], st));
    std::vector<double> dev(beliefs ? P->V : 0);
    if (beliefs && P->V)
        HIP_TRY(hipMemcpyAsync(dev.data(), P->belief, P->V * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (beliefs)   /* device variable order -> variable order */
        for (uint64_t dv = 0; dv < P->V; dv++) beliefs[P->h_perm[dv]] = dev[dv];
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, A->ev[0], A->ev[1]));
    A->fwd_ms = ms;
    A->have_forward = true;
    P->swept_since_init = true;
    return FLBP_OK;
}

}  // extern "C"

/* the reverse walk with bbar already on the device; dL/dp is added to pbar */
int flbp::Adjoint::run_backward() {
    const uint64_t E = P->E, V = P->V, F = P->F;
    hipStream_t st = P->stream;
    HIP_TRY(hipEventRecord(ev[2], st));
    if (E) {
        hipLaunchKernelGGL(fill2, dim3(grid_for(E, 256)), dim3(256), 0, st, avf, E, 0.0);
        hipLaunchKernelGGL(fill2, dim3(grid_for(E, 256)), dim3(256), 0, st, afv, E, 0.0);
    }
    const double keep = 1.0 - damp;
    for (int64_t c = (int64_t)nck - 1; c >= 0; c--) {
        const uint64_t t0 = (uint64_t)c * k, t1 = std::min<uint64_t>(t0 + k, T);
        auto vf_state = [&](uint64_t t) -> double2 * {   /* m_vf after t sweeps, t0 <= t <= t1 */
            return t == t0 ? ck_vf[c] : st_vf[t - t0 - 1];
        };
        auto fv_state = [&](uint64_t t) -> double2 * {
            return t == t0 ? ck_fv[c] : st_fv[t - t0 - 1];
        };
        for (uint64_t t = t0 + 1; t <= t1; t++)
            TRY(P->launch_sweep_on(0, -INFINITY, damp, vf_state(t - 1), fv_state(t - 1), fv_state(t), vf_state(t),
                                   scratch_belief));
        for (uint64_t t = t1; t > t0; t--) {
            if (V)
                hipLaunchKernelGGL(var_vjp_staged, dim3((unsigned)n_var_blocks), dim3(kBlock), 0, st, var_blk,
                                   P->var_off, P->var_edge, var_work, fv_state(t), avf, t == T ? bbar : nullptr, afv,
                                   gscratch, damp);
            if (F)
                hipLaunchKernelGGL(fac_vjp_staged, dim3((unsigned)n_fac_blocks), dim3(kBlock), 0, st, fac_blk,
                                   P->fac_off, P->fac_meta, P->fac_p, P->fac_q, P->fac_mask, P->fac_vpos, fac_work,
                                   vf_state(t - 1), afv, avf2, pbar, gscratch, keep);
            std::swap(avf, avf2);
        }
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ev[3], st));
    return FLBP_OK;
}

extern "C" {

static int adjoint_check(Adjoint *A, const char *what) {
    if (!A) return flbp::fail(FLBP_EINVAL, std::string(what) + ": null adjoint");
    if (!A->have_forward) return flbp::fail(FLBP_EINVAL, std::string(what) + ": run the forward first");
    if (A->P->variant >= 2)
        return flbp::fail(FLBP_EINVAL, std::string(what) + ": the plan was switched away from the pull layout");
    return FLBP_OK;
}

static int adjoint_finish(Adjoint *A) {
    HIP_TRY(hipStreamSynchronize(A->P->stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, A->ev[2], A->ev[3]));
    A->bwd_ms = ms;
    return FLBP_OK;
}

int flbp_adjoint_backward(flbp_adjoint *adj, const double *dl_db, double *dl_dp) {
    Adjoint *A = reinterpret_cast<Adjoint *>(adj);
    if (!dl_db || !dl_dp) return flbp::fail(FLBP_EINVAL, "flbp_adjoint_backward: null argument");
    TRY(adjoint_check(A, "flbp_adjoint_backward"));
    Plan *P = A->P;
    HIP_TRY(hipSetDevice(P->device));
    const uint64_t V = P->V, F = P->F;
    std::vector<double> dev_db(V);
    for (uint64_t dv = 0; dv < V; dv++) dev_db[dv] = dl_db[P->h_perm[dv]];
    if (V) HIP_TRY(hipMemcpyAsync(A->bbar, dev_db.data(), V * sizeof(double), hipMemcpyHostToDevice, P->stream));
    if (F) hipLaunchKernelGGL(fill1, dim3(grid_for(F, 256)), dim3(256), 0, P->stream, A->pbar, F, 0.0);
    TRY(A->run_backward());
    if (F) {   /* device factor order -> factor order */
        std::vector<double> dev_dp(F);
        HIP_TRY(hipMemcpyAsync(dev_dp.data(), A->pbar, F * sizeof(double), hipMemcpyDeviceToHost, P->stream));
        HIP_TRY(hipStreamSynchronize(P->stream));
        for (uint64_t I = 0; I < F; I++) dl_dp[I] = dev_dp[P->h_finv[I]];
    }
    return adjoint_finish(A);
}

int flbp_adjoint_backward_sparse(flbp_adjoint *adj, uint64_t n, const uint64_t *vars, const double *dl_db) {
    Adjoint *A = reinterpret_cast<Adjoint *>(adj);
    if (n && (!vars || !dl_db)) return flbp::fail(FLBP_EINVAL, "flbp_adjoint_backward_sparse: null argument");
    TRY(adjoint_check(A, "flbp_adjoint_backward_sparse"));
    Plan *P = A->P;
    HIP_TRY(hipSetDevice(P->device));
    /* device positions, repeated variables summed in the given order */
    std::vector<std::pair<uint32_t, uint64_t>> key(n);
    for (uint64_t i = 0; i < n; i++) {
        if (vars[i] >= P->V) return flbp::fail(FLBP_EINVAL, "flbp_adjoint_backward_sparse: variable out of range");
        key[i] = {P->h_inv_perm[vars[i]], i};
    }
    std::sort(key.begin(), key.end());
    std::vector<uint32_t> idx;
    std::vector<double> val;
    for (uint64_t i = 0; i < n; i++) {
        if (!idx.empty() && idx.back() == key[i].first) val.back() += dl_db[key[i].second];
        else {
            idx.push_back(key[i].first);
            val.push_back(dl_db[key[i].second]);
        }
    }
    if (P->V) hipLaunchKernelGGL(fill1, dim3(grid_for(P->V, 256)), dim3(256), 0, P->stream, A->bbar, P->V, 0.0);
    if (!idx.empty()) {
        TRY(P->stage(idx.size()));
        HIP_TRY(hipMemcpyAsync(P->stage_idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, P->stream));
        HIP_TRY(hipMemcpyAsync(P->stage_val, val.data(), val.size() * 8, hipMemcpyHostToDevice, P->stream));
        hipLaunchKernelGGL(scatter_f64, dim3(grid_for(idx.size(), 256)), dim3(256), 0, P->stream, A->bbar,
                           P->stage_idx, P->stage_val, (uint64_t)idx.size());
    }
    TRY(A->run_backward());
    return adjoint_finish(A);   /* also keeps the host vectors alive until the copies are done */
}

int flbp_adjoint_clear_grad(flbp_adjoint *adj) {
    Adjoint *A = reinterpret_cast<Adjoint *>(adj);
    if (!A) return flbp::fail(FLBP_EINVAL, "flbp_adjoint_clear_grad: null adjoint");
    HIP_TRY(hipSetDevice(A->device));
    if (A->P->F)
        hipLaunchKernelGGL(fill1, dim3(grid_for(A->P->F, 256)), dim3(256), 0, A->P->stream, A->pbar, A->P->F, 0.0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(A->P->stream));
    return FLBP_OK;
}

int flbp_adjoint_get_dl_dp(flbp_adjoint *adj, double *dl_dp) {
    Adjoint *A = reinterpret_cast<Adjoint *>(adj);
    if (!A || (A->P->F && !dl_dp)) return flbp::fail(FLBP_EINVAL, "flbp_adjoint_get_dl_dp: null argument");
    HIP_TRY(hipSetDevice(A->device));
    const uint64_t F = A->P->F;
    if (F) {   /* device factor order -> factor order */
        std::vector<double> dev_dp(F);
        HIP_TRY(hipMemcpyAsync(dev_dp.data(), A->pbar, F * sizeof(double), hipMemcpyDeviceToHost, A->P->stream));
        HIP_TRY(hipStreamSynchronize(A->P->stream));
        for (uint64_t I = 0; I < F; I++) dl_dp[I] = dev_dp[A->P->h_finv[I]];
    }
    return FLBP_OK;
}

int flbp_adjoint_set_rules(flbp_adjoint *adj, const int32_t *rule, uint64_t n_rules) {
    Adjoint *A = reinterpret_cast<Adjoint *>(adj);
    const uint64_t F = A ? A->P->F : 0;
    if (!A || (F && !rule) || n_rules == 0 || n_rules >= (1ull << 31))
        return flbp::fail(FLBP_EINVAL, "flbp_adjoint_set_rules: bad argument");
    std::vector<uint64_t> off(n_rules + 1, 0);
    for (uint64_t I = 0; I < F; I++) {
        if (rule[I] < 0) continue;
        if ((uint64_t)rule[I] >= n_rules) return flbp::fail(FLBP_EINVAL, "flbp_adjoint_set_rules: rule id out of range");
        off[rule[I] + 1]++;
    }
    for (uint64_t r = 0; r < n_rules; r++) off[r + 1] += off[r];
    /* each rule's factors as device indices, listed in factor order (the
     * order the reduction adds them in) */
    std::vector<uint32_t> fac(off[n_rules]);
    std::vector<uint64_t> pos(off.begin(), off.end() - 1);
    for (uint64_t I = 0; I < F; I++)
        if (rule[I] >= 0) fac[pos[rule[I]]++] = A->P->h_finv[I];
    HIP_TRY(hipSetDevice(A->device));
    for (void *q : {(void *)A->rule_off, (void *)A->rule_fac, (void *)A->rule_out}) {
        if (!q) continue;
        auto it = std::find(A->allocs.begin(), A->allocs.end(), q);
        if (it != A->allocs.end()) A->allocs.erase(it);
        (void)hipFree(q);
    }
    A->rule_off = nullptr;
    A->rule_fac = nullptr;
    A->rule_out = nullptr;
    TRY(A->alloc(A->rule_off, n_rules + 1));
    TRY(A->alloc(A->rule_fac, fac.size()));
    TRY(A->alloc(A->rule_out, n_rules));
    HIP_TRY(hipMemcpy(A->rule_off, off.data(), off.size() * 8, hipMemcpyHostToDevice));
    if (!fac.empty()) HIP_TRY(hipMemcpy(A->rule_fac, fac.data(), fac.size() * 4, hipMemcpyHostToDevice));
    A->n_rules = n_rules;
    return FLBP_OK;
}

int flbp_adjoint_rule_grad(flbp_adjoint *adj, double *dl_dtheta) {
    Adjoint *A = reinterpret_cast<Adjoint *>(adj);
    if (!A || !dl_dtheta) return flbp::fail(FLBP_EINVAL, "flbp_adjoint_rule_grad: null argument");
    if (!A->n_rules) return flbp::fail(FLBP_EINVAL, "flbp_adjoint_rule_grad: call flbp_adjoint_set_rules first");
    HIP_TRY(hipSetDevice(A->device));
    hipLaunchKernelGGL(rule_reduce, dim3((unsigned)A->n_rules), dim3(256), 0, A->P->stream, A->rule_off, A->rule_fac,
                       A->pbar, A->P->fac_p, A->rule_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(dl_dtheta, A->rule_out, A->n_rules * 8, hipMemcpyDeviceToHost, A->P->stream));
    HIP_TRY(hipStreamSynchronize(A->P->stream));
    return FLBP_OK;
}

int flbp_adjoint_last_timing(flbp_adjoint *adj, double *forward_ms, double *backward_ms) {
    Adjoint *A = reinterpret_cast<Adjoint *>(adj);
    if (!A) return flbp::fail(FLBP_EINVAL, "flbp_adjoint_last_timing: null adjoint");
    if (forward_ms) *forward_ms = A->fwd_ms;
    if (backward_ms) *backward_ms = A->bwd_ms;
    return FLBP_OK;
}

uint64_t flbp_adjoint_device_bytes(const flbp_adjoint *adj) {
    const Adjoint *A = reinterpret_cast<const Adjoint *>(adj);
    return A ? A->bytes : 0;
}

int flbp_tied_p(uint64_t n_factors, const double *base_p, const int32_t *rule, uint64_t n_rules,
                const double *theta, double *p_out) {
    if (n_factors && (!base_p || !rule || !p_out)) return flbp::fail(FLBP_EINVAL, "flbp_tied_p: null argument");
    for (uint64_t I = 0; I < n_factors; I++) {
        const int32_t r = rule[I];
        if (r < 0) {
            p_out[I] = base_p[I];
        } else {
            if ((uint64_t)r >= n_rules || !theta) return flbp::fail(FLBP_EINVAL, "flbp_tied_p: rule id out of range");
            p_out[I] = 1.0 / (1.0 + std::exp(-theta[r]));
        }
    }
    return FLBP_OK;
}

int flbp_rule_grad(uint64_t n_factors, const double *dl_dp, const double *p, const int32_t *rule,
                   uint64_t n_rules, double *dl_dtheta) {
    if ((n_factors && (!dl_dp || !p || !rule)) || (n_rules && !dl_dtheta))
        return flbp::fail(FLBP_EINVAL, "flbp_rule_grad: null argument");
    for (uint64_t r = 0; r < n_rules; r++) dl_dtheta[r] = 0.0;
    for (uint64_t I = 0; I < n_factors; I++) {
        const int32_t r = rule[I];
        if (r < 0) continue;
        if ((uint64_t)r >= n_rules) return flbp::fail(FLBP_EINVAL, "flbp_rule_grad: rule id out of range");
        dl_dtheta[r] += dl_dp[I] * p[I] * (1.0 - p[I]);
    }
    return FLBP_OK;
}

}  // extern "C"
