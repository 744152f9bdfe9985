/*
 * seq.hip -- a plan's sweeps under a sequential schedule (updates = SEQFIX /
 * TOPO): upload of the schedule's items grouped by dependency level, the
 * per-sweep launches, and the C ABI that attaches a schedule to a plan.
 *
 * Device state added to the plan (variant-3 message orders, which are the
 * reference's: v->f factor-major m_vf, f->v var-major m_fv_vm):
 *   seq_state  {f64 p0, p1; u64 z0, z1}[V]  products and zero counts
 *              (prod_fv_0/1, num_zeros_fv_0/1), maintained in place
 *   seq_vf     u32x4[n_vf]  v->f items (slot, var-major position, variable, 0)
 *   seq_fv     u32x4[n_fv]  f->v items (var-major position, variable, factor,
 *              first slot of the factor)
 *   seq_fv_aux u32x4[n_fv]  (degree | factor type << 24, target slot, head slot, 0)
 *   seq_fac    {f64 p, q, m0, m1}[F]  factor records, rebuilt on the device
 *              from fac_p / fac_q / fac_meta / fac_mask whenever they change
 *   seq_lvl_*_dev u32[levels+1]  first item of each level
 * Items are grouped by level; within a level v->f items are ordered by
 * slot and f->v items by factor, so neighbouring lanes touch neighbouring
 * messages.  The trailing levels of at most kSeqTailItems items each run
 * in one single-workgroup launch (seq_tail).
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

__global__ void fill_state(flbp::dev::VarState *st, uint64_t n, double c) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st[i] = flbp::dev::VarState{c, c, 0ull, 0ull};
}

/* trailing levels up to this many items each go to the single-workgroup
 * tail kernel */
constexpr uint32_t kSeqTailItems = flbp::dev::kSeqTailPer * flbp::dev::kSeqTailBlock;

}  // namespace

namespace flbp {

#define HIP_TRY(expr)                                     \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

void Plan::free_seq_items() {
    for (void *p : {(void *)seq_vf, (void *)seq_fv, (void *)seq_fv_aux, (void *)seq_lvl_vf_dev,
                    (void *)seq_lvl_fv_dev})
        if (p) (void)hipFree(p);
    seq_vf = seq_fv = seq_fv_aux = nullptr;
    seq_lvl_vf_dev = seq_lvl_fv_dev = nullptr;
    seq_lvl_vf.clear();
    seq_lvl_fv.clear();
    seq_tail_from = 0;
}

/* One sweep: every level, then beliefs and the max-diff of sweep k. */
int Plan::launch_seq_sweep(int k, double tol) {
    const size_t L = seq_lvl_fv.empty() ? 0 : seq_lvl_fv.size() - 1;
    SeqArgs A{seq_vf, seq_fv, seq_fv_aux, seq_fac, m_vf, m_fv_vm, seq_state};
    for (size_t l = 0; l < seq_tail_from; l++) {
        const uint32_t v0 = seq_lvl_vf[l], nv = seq_lvl_vf[l + 1] - v0;
        const uint32_t f0 = seq_lvl_fv[l], nf = seq_lvl_fv[l + 1] - f0;
        if (!nv && !nf) continue;
        const unsigned vb = grid_for(nv, dev::kBlock), fb = grid_for(nf, dev::kBlock);
        SeqArgs a = A;
        a.vf_items = seq_vf + v0;
        a.fv_items = seq_fv + f0;
        a.fv_aux = seq_fv_aux + f0;
        hipLaunchKernelGGL(seq_level, dim3(vb + fb), dim3(dev::kBlock), 0, stream, a, nv, nf, vb, ctl, k, tol);
    }
    for (size_t l0 = seq_tail_from; l0 < L; l0 += dev::kSeqTailLevels) {
        const size_t l1 = std::min(L, l0 + dev::kSeqTailLevels);
        hipLaunchKernelGGL(seq_tail, dim3(1), dim3(dev::kSeqTailBlock), 0, stream, A, seq_lvl_vf_dev,
                           seq_lvl_fv_dev, (uint32_t)l0, (uint32_t)l1, ctl, k, tol);
    }
    if (profiling) HIP_TRY(hipEventRecord(ev[2 * k + 1], stream));
    if (V)
        hipLaunchKernelGGL(seq_beliefs, dim3(grid_for(V, dev::kBlock)), dim3(dev::kBlock), 0, stream, seq_state,
                           belief, (uint32_t)V, ctl, k, tol);
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
    /* f->v items as (key = factor, edge; payload = position in wave order) */
    std::vector<uint4> vf(S.vf_slot.size());
    std::vector<std::pair<uint64_t, uint32_t>> fk(S.fv_edge.size());
    {
        std::vector<uint32_t> at(lv.begin(), lv.end());
        for (size_t x = 0; x < S.vf_slot.size(); x++)
            vf[at[S.vf_level[x]]++] = make_uint4(S.vf_slot[x], S.vf_pos[x], S.vf_var[x], 0u);
        std::vector<uint32_t> bt(lf.begin(), lf.end());
        for (size_t x = 0; x < S.fv_edge.size(); x++)
            fk[bt[S.fv_level[x]]++] = {((uint64_t)S.fv_fac[x] << 32) | S.fv_edge[x], (uint32_t)x};
    }
    for (uint32_t l = 1; l <= nl; l++) {
        std::sort(vf.begin() + lv[l], vf.begin() + lv[l + 1], [](const uint4 &a, const uint4 &b) { return a.x < b.x; });
        std::sort(fk.begin() + lf[l], fk.begin() + lf[l + 1]);
    }
    std::vector<uint4> fv(fk.size()), faux(fk.size());
    for (size_t y = 0; y < fk.size(); y++) {
        const uint32_t x = fk[y].second, I = S.fv_fac[x];
        const uint32_t a = (uint32_t)P->h_fac_off[I], d = (uint32_t)(P->h_fac_off[I + 1] - a);
        const uint32_t meta = P->h_meta[I];   /* type and head slot never change */
        fv[y] = make_uint4(S.fv_edge[x], S.fv_var[x], I, a);
        faux[y] = make_uint4(d | ((meta & 3u) << 24), S.fv_slot[x] - a, meta >> 8, 0u);
    }
    P->free_seq_items();
    const size_t nvf = std::max<size_t>(vf.size(), 1), nfv = std::max<size_t>(fv.size(), 1);
    if (hipMalloc(&P->seq_vf, nvf * sizeof(uint4)) != hipSuccess ||
        hipMalloc(&P->seq_fv, nfv * sizeof(uint4)) != hipSuccess ||
        hipMalloc(&P->seq_fv_aux, nfv * sizeof(uint4)) != hipSuccess ||
        hipMalloc(&P->seq_lvl_vf_dev, (nl + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&P->seq_lvl_fv_dev, (nl + 1) * sizeof(uint32_t)) != hipSuccess) {
        (void)hipGetLastError();
        P->free_seq_items();
        return flbp::fail(FLBP_ENOMEM, "flbp_plan_set_schedule: item upload allocation failed");
    }
    if (!vf.empty()) HIP_TRY(hipMemcpy(P->seq_vf, vf.data(), vf.size() * sizeof(uint4), hipMemcpyHostToDevice));
    if (!fv.empty()) {
        HIP_TRY(hipMemcpy(P->seq_fv, fv.data(), fv.size() * sizeof(uint4), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P->seq_fv_aux, faux.data(), faux.size() * sizeof(uint4), hipMemcpyHostToDevice));
    }
    /* level l (1-based) -> index l-1 */
    P->seq_lvl_vf.assign(lv.begin() + 1, lv.end());
    P->seq_lvl_fv.assign(lf.begin() + 1, lf.end());
    HIP_TRY(hipMemcpy(P->seq_lvl_vf_dev, P->seq_lvl_vf.data(), (nl + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(P->seq_lvl_fv_dev, P->seq_lvl_fv.data(), (nl + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    /* the tail: the longest run of trailing levels with few items each */
    size_t tail = nl;
    while (tail > 0 && (P->seq_lvl_vf[tail] - P->seq_lvl_vf[tail - 1]) + (P->seq_lvl_fv[tail] - P->seq_lvl_fv[tail - 1]) <=
                           kSeqTailItems)
        tail--;
    P->seq_tail_from = tail;
    if (!P->seq_state) {
        if (int rc = P->alloc(P->seq_state, P->V)) return rc;
        if (int rc = P->alloc(P->seq_fac, P->F)) return rc;
    }
    if (P->V)
        hipLaunchKernelGGL(fill_state, dim3(grid_for(P->V, 256)), dim3(256), 0, P->stream, P->seq_state, P->V,
                           P->init_value);
    P->seq = true;
    if (int rc = flbp::seq_refresh_factors(P)) return rc;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(P->stream));
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

int seq_refresh_factors(Plan *P) {
    if (!P->seq || !P->F) return FLBP_OK;
    hipLaunchKernelGGL(seq_fac_build, dim3(grid_for(P->F, dev::kBlock)), dim3(dev::kBlock), 0, P->stream,
                       P->fac_meta, P->fac_p, P->fac_q, P->fac_mask, P->seq_fac, (uint32_t)P->F);
    HIP_TRY(hipGetLastError());
    return FLBP_OK;
}

/* products back to the values init() leaves: (c, c), no zeros */
int seq_reset_state(Plan *P, double c) {
    if (!P->seq_state || !P->V) return FLBP_OK;
    hipLaunchKernelGGL(fill_state, dim3(grid_for(P->V, 256)), dim3(256), 0, P->stream, P->seq_state, P->V, c);
    HIP_TRY(hipGetLastError());
    return FLBP_OK;
}

}  // namespace flbp
