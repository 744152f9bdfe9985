/*
 * seq.hip -- a plan's sweeps under a sequential schedule (updates = SEQFIX /
 * TOPO): upload of the schedule's items grouped by dependency level, the
 * per-sweep launches, and the C ABI that attaches a schedule to a plan.
 *
 * Device state added to the plan (variant-3 message orders, which are the
 * reference's: v->f factor-major m_vf, f->v var-major m_fv_vm):
 *   seq_prod f64x2[V]  products (prod_fv_0, prod_fv_1), maintained in place
 *   seq_nz   u64x2[V]  zero counts (num_zeros_fv_0/1)
 *   seq_vf   u32x4[n_vf]  v->f items (slot, var-major position, variable)
 *   seq_fv   u32x4[n_fv]  f->v items (var-major position, slot, factor, variable)
 * Items are grouped by level; within a level v->f items are ordered by
 * slot and f->v items by factor, so neighbouring lanes touch neighbouring
 * messages.
 *
 * Replaces the reference's per-wave launches (csrc/ext/causal_bp.cu:446-477,
 * 868-872; the CUDA-graph replay of 82-112, 1028-1035) -- two kernels per
 * wave -- by one launch per level plus the belief kernel per sweep.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "fastlbp.h"
#include "flbp_internal.h"
#include "plan_internal.h"
#include "seq_internal.h"
#include "seq_kernels.h"

namespace {

__global__ void fill_state(double2 *prod, ulonglong2 *nz, uint64_t n, double c) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        prod[i] = make_double2(c, c);
        nz[i] = make_ulonglong2(0ull, 0ull);
    }
}

}  // namespace

namespace flbp {

#define HIP_TRY(expr)                                     \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

void Plan::free_seq_items() {
    if (seq_vf) (void)hipFree(seq_vf);
    if (seq_fv) (void)hipFree(seq_fv);
    seq_vf = seq_fv = nullptr;
    seq_lvl_vf.clear();
    seq_lvl_fv.clear();
}

/* One sweep: every level, then beliefs and the max-diff of sweep k. */
int Plan::launch_seq_sweep(int k, double tol) {
    const size_t L = seq_lvl_fv.empty() ? 0 : seq_lvl_fv.size() - 1;
    for (size_t l = 0; l < L; l++) {
        const uint32_t v0 = seq_lvl_vf[l], nv = seq_lvl_vf[l + 1] - v0;
        const uint32_t f0 = seq_lvl_fv[l], nf = seq_lvl_fv[l + 1] - f0;
        if (!nv && !nf) continue;
        const unsigned vb = grid_for(nv, dev::kBlock), fb = grid_for(nf, dev::kBlock);
        hipLaunchKernelGGL(seq_level, dim3(vb + fb), dim3(dev::kBlock), 0, stream, seq_vf + v0, nv, seq_fv + f0, nf,
                           vb, fac_off, fac_meta, fac_p, fac_q, fac_mask, m_vf, m_fv_vm, seq_prod, seq_nz, ctl, k,
                           tol);
    }
    if (profiling) HIP_TRY(hipEventRecord(ev[2 * k + 1], stream));
    if (V)
        hipLaunchKernelGGL(seq_beliefs, dim3(grid_for(V, dev::kBlock)), dim3(dev::kBlock), 0, stream, seq_prod,
                           seq_nz, belief, (uint32_t)V, ctl, k, tol);
    if (profiling) HIP_TRY(hipEventRecord(ev[2 * k + 2], stream));
    HIP_TRY(hipGetLastError());
    return FLBP_OK;
}

}  // namespace flbp

using flbp::Plan;
using flbp::grid_for;
using flbp::hip_fail;

extern "C" {

int flbp_plan_set_schedule(flbp_plan *plan, const flbp_schedule *sched) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P) return flbp::fail(FLBP_EINVAL, "flbp_plan_set_schedule: null plan");
    HIP_TRY(hipSetDevice(P->device));
    HIP_TRY(hipStreamSynchronize(P->stream));
    P->drop_graphs();
    if (!sched) {
        P->free_seq_items();
        P->seq = false;
        P->seq_updates = -1;
        P->seq_waves = 0;
        return FLBP_OK;
    }
    const flbp::Schedule &S = *flbp::schedule_of(sched);
    if (S.V != P->V || S.F != P->F || S.E != P->E)
        return flbp::fail(FLBP_EINVAL, "flbp_plan_set_schedule: schedule built for another graph");
    if (P->variant != 3) {
        if (int rc = flbp_plan_set_variant(plan, 3)) return rc;
    }
    /* items grouped by level (counting sort, stable in wave order), then
     * ordered by slot / factor inside each level */
    const uint32_t nl = S.n_levels;
    std::vector<uint32_t> lv(nl + 2, 0), lf(nl + 2, 0);
    for (uint32_t l : S.vf_level) lv[l + 1]++;
    for (uint32_t l : S.fv_level) lf[l + 1]++;
    for (uint32_t l = 0; l <= nl; l++) {
        lv[l + 1] += lv[l];
        lf[l + 1] += lf[l];
    }
    std::vector<uint4> vf(S.vf_slot.size()), fv(S.fv_edge.size());
    {
        std::vector<uint32_t> at(lv.begin(), lv.end());
        for (size_t x = 0; x < S.vf_slot.size(); x++)
            vf[at[S.vf_level[x]]++] = make_uint4(S.vf_slot[x], S.vf_pos[x], S.vf_var[x], 0u);
        std::vector<uint32_t> bt(lf.begin(), lf.end());
        for (size_t x = 0; x < S.fv_edge.size(); x++)
            fv[bt[S.fv_level[x]]++] = make_uint4(S.fv_edge[x], S.fv_slot[x], S.fv_fac[x], S.fv_var[x]);
    }
    for (uint32_t l = 1; l <= nl; l++) {
        std::sort(vf.begin() + lv[l], vf.begin() + lv[l + 1], [](const uint4 &a, const uint4 &b) { return a.x < b.x; });
        std::sort(fv.begin() + lf[l], fv.begin() + lf[l + 1], [](const uint4 &a, const uint4 &b) {
            return a.z != b.z ? a.z < b.z : a.x < b.x;
        });
    }
    P->free_seq_items();
    const size_t nvf = std::max<size_t>(vf.size(), 1), nfv = std::max<size_t>(fv.size(), 1);
    if (hipMalloc(&P->seq_vf, nvf * sizeof(uint4)) != hipSuccess ||
        hipMalloc(&P->seq_fv, nfv * sizeof(uint4)) != hipSuccess) {
        (void)hipGetLastError();
        P->free_seq_items();
        return flbp::fail(FLBP_ENOMEM, "flbp_plan_set_schedule: item upload allocation failed");
    }
    if (!vf.empty()) HIP_TRY(hipMemcpy(P->seq_vf, vf.data(), vf.size() * sizeof(uint4), hipMemcpyHostToDevice));
    if (!fv.empty()) HIP_TRY(hipMemcpy(P->seq_fv, fv.data(), fv.size() * sizeof(uint4), hipMemcpyHostToDevice));
    /* level l (1-based) -> index l-1 */
    P->seq_lvl_vf.assign(lv.begin() + 1, lv.end());
    P->seq_lvl_fv.assign(lf.begin() + 1, lf.end());
    if (!P->seq_prod) {
        if (int rc = P->alloc(P->seq_prod, P->V)) return rc;
        if (int rc = P->alloc(P->seq_nz, P->V)) return rc;
    }
    if (P->V)
        hipLaunchKernelGGL(fill_state, dim3(grid_for(P->V, 256)), dim3(256), 0, P->stream, P->seq_prod, P->seq_nz, P->V,
                           P->init_value);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(P->stream));
    P->seq = true;
    P->seq_updates = S.updates;
    P->seq_waves = S.n_waves;
    return FLBP_OK;
}

int flbp_plan_get_updates(const flbp_plan *plan) {
    const Plan *P = reinterpret_cast<const Plan *>(plan);
    if (!P) return flbp::fail(FLBP_EINVAL, "flbp_plan_get_updates: null plan");
    return P->seq ? P->seq_updates : FLBP_UPDATES_PARALL;
}

}  // extern "C"

namespace flbp {

/* products back to the values init() leaves: (c, c), no zeros */
int seq_reset_state(Plan *P, double c) {
    if (!P->seq_prod || !P->V) return FLBP_OK;
    hipLaunchKernelGGL(fill_state, dim3(grid_for(P->V, 256)), dim3(256), 0, P->stream, P->seq_prod, P->seq_nz, P->V, c);
    HIP_TRY(hipGetLastError());
    return FLBP_OK;
}

}  // namespace flbp
