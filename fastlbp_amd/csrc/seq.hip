/*
 * seq.hip -- a plan's sweeps under a sequential schedule (updates = SEQFIX /
 * TOPO): upload of the schedule's items grouped by dependency level, the
 * per-sweep launches, and the C ABI that attaches a schedule to a plan.
 *
 * Messages: v->f factor-major in m_vf (the reference's order) and f->v
 * factor-major in m_fv -- every message a group item touches lies in its
 * factor's slot range -- with the var-major copy m_fv_vm rebuilt when the
 * schedule is removed or the messages are read out (vm_stale).
 * Device state added to the plan:
 *   seq_state  {f64 p0, p1; u64 z0, z1}[V]  products and zero counts
 *              (prod_fv_0/1, num_zeros_fv_0/1), maintained in place
 *   seq_vf     u32x4[n_vf]  v->f items (slot, variable, 0, 0)
 *   seq_fv     u32x4[n_fv]  f->v items (var-major position, variable, factor,
 *              first slot of the factor)
 *   seq_fv_aux u32x4[n_fv]  (degree | factor type << 24, target slot, head slot, 0)
 *   seq_grp    u32x4[lanes]  group items, one lane per slot: (factor, first
 *              slot, group_word(degree, type, head slot, updated mask,
 *              recomputed mask, slot), variable); a group never straddles a
 *              64-lane wave and each level's lanes are padded to whole waves
 *              (degree 0)
 *   seq_fac    {f64 p, q, m0, m1}[F]  factor records, rebuilt on the device
 *              from fac_p / fac_q / fac_meta / fac_mask whenever they change
 *   seq_lvl_*_dev u32[levels+1]  first item of each level
 * Items are grouped by level; within a level v->f items are ordered by
 * slot and f->v items by factor, so neighbouring lanes touch neighbouring
 * messages.  The trailing levels of at most seq_tail_max items each run
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
static_assert(kSeqTailItems == flbp::kSeqTailItemsMax, "tail kernel capacity");

}  // namespace

namespace flbp {

#define HIP_TRY(expr)                                     \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

void Plan::free_seq_items() {
    for (void *p : {(void *)seq_vf, (void *)seq_fv, (void *)seq_fv_aux, (void *)seq_grp, (void *)seq_lvl_vf_dev,
                    (void *)seq_lvl_fv_dev, (void *)seq_lvl_grp_dev})
        if (p) (void)hipFree(p);
    seq_vf = seq_fv = seq_fv_aux = seq_grp = nullptr;
    seq_lvl_vf_dev = seq_lvl_fv_dev = seq_lvl_grp_dev = nullptr;
    seq_lvl_vf.clear();
    seq_lvl_fv.clear();
    seq_lvl_grp.clear();
    seq_tail_from = 0;
}

/* The tail: the longest run of trailing levels with at most seq_tail_max
 * items (group lanes count one each) every one. */
void Plan::seq_place_tail() {
    const size_t nl = seq_lvl_fv.empty() ? 0 : seq_lvl_fv.size() - 1;
    auto level_items = [&](size_t l) {   /* items of 0-based level l */
        return (seq_lvl_vf[l + 1] - seq_lvl_vf[l]) + (seq_lvl_fv[l + 1] - seq_lvl_fv[l]) +
               (seq_lvl_grp[l + 1] - seq_lvl_grp[l]);
    };
    size_t tail = nl;
    while (tail > 0 && level_items(tail - 1) <= seq_tail_max) tail--;
    seq_tail_from = tail;
}

/* One sweep: every level, then beliefs and the max-diff of sweep k. */
int Plan::launch_seq_sweep(int k, double tol) {
    const size_t L = seq_lvl_fv.empty() ? 0 : seq_lvl_fv.size() - 1;
    SeqArgs A{seq_vf, seq_fv, seq_fv_aux, seq_grp, seq_fac, m_vf, m_fv, seq_state};
    for (size_t l = 0; l < seq_tail_from; l++) {
        const uint32_t v0 = seq_lvl_vf[l], nv = seq_lvl_vf[l + 1] - v0;
        const uint32_t f0 = seq_lvl_fv[l], nf = seq_lvl_fv[l + 1] - f0;
        const uint32_t g0 = seq_lvl_grp[l], ng = seq_lvl_grp[l + 1] - g0;
        if (!nv && !nf && !ng) continue;
        const unsigned bs = dev::kSeqLevelBlock;
        const unsigned vb = grid_for(nv, bs), fb = grid_for(nf, bs), gb = grid_for(ng, bs);
        SeqArgs a = A;
        a.vf_items = seq_vf + v0;
        a.fv_items = seq_fv + f0;
        a.fv_aux = seq_fv_aux + f0;
        a.grp_items = seq_grp + g0;
        hipLaunchKernelGGL(seq_level, dim3(vb + fb + gb), dim3(bs), 0, stream, a, nv, nf, ng, vb, fb, ctl,
                           k, tol);
    }
    for (size_t l0 = seq_tail_from; l0 < L; l0 += dev::kSeqTailLevels) {
        const size_t l1 = std::min(L, l0 + dev::kSeqTailLevels);
        hipLaunchKernelGGL(seq_tail, dim3(1), dim3(dev::kSeqTailBlock), 0, stream, A, seq_lvl_vf_dev,
                           seq_lvl_fv_dev, seq_lvl_grp_dev, (uint32_t)l0, (uint32_t)l1, ctl, k, tol);
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
        if (int rc = P->vm_current()) return rc;   /* PARALL's var pass reads the var-major array */
        HIP_TRY(hipStreamSynchronize(P->stream));
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
    /* the schedule's sweeps read and write the factor-major f->v array */
    if (int rc = P->fm_current()) return rc;
    /* the schedule's execution items, grouped by kind and level (counting
     * sort), then ordered by slot / factor inside each level; an item names
     * its variable by device index, the index of the products and beliefs */
    const uint32_t nl = S.n_levels;
    const size_t nx = S.ex_kind.size();
    std::vector<uint32_t> off[3];
    for (auto &o : off) o.assign(nl + 2, 0);
    for (size_t x = 0; x < nx; x++) off[S.ex_kind[x]][S.ex_level[x] + 1]++;
    for (auto &o : off)
        for (uint32_t l = 0; l <= nl; l++) o[l + 1] += o[l];
    std::vector<uint32_t> slot_fac(P->E);
    for (uint64_t I = 0; I < P->F; I++)
        for (uint64_t sl = P->h_fac_off[I]; sl < P->h_fac_off[I + 1]; sl++) slot_fac[sl] = (uint32_t)I;
    /* per kind: (sort key, execution item) in level order */
    std::vector<std::pair<uint64_t, uint32_t>> keyed[3];
    for (int kd = 0; kd < 3; kd++) keyed[kd].resize(off[kd][nl + 1]);
    {
        std::vector<uint32_t> at[3];
        for (int kd = 0; kd < 3; kd++) at[kd].assign(off[kd].begin(), off[kd].end());
        for (size_t x = 0; x < nx; x++) {
            const int kd = S.ex_kind[x];
            /* f->v items by factor; groups by degree, then factor, so the
             * lanes of a wave mostly run loops of one length */
            uint64_t key = S.ex_a[x];
            if (kd == flbp::kExecVF) key = P->h_slot_dev[S.ex_a[x]];
            if (kd == flbp::kExecFV) key |= (uint64_t)slot_fac[P->h_var_edge[S.ex_a[x]]] << 32;
            if (kd == flbp::kExecGroup) {
                const uint32_t I = P->h_finv[S.ex_a[x]];
                key = I | (P->h_fac_off[I + 1] - P->h_fac_off[I]) << 32;
            }
            keyed[kd][at[kd][S.ex_level[x]]++] = {key, (uint32_t)x};
        }
    }
    for (int kd = 0; kd < 3; kd++)
        for (uint32_t l = 1; l <= nl; l++) std::sort(keyed[kd].begin() + off[kd][l], keyed[kd].begin() + off[kd][l + 1]);
    std::vector<uint4> vf(keyed[0].size()), fv(keyed[1].size()), faux(keyed[1].size());
    for (size_t y = 0; y < keyed[0].size(); y++) {
        const uint32_t sl = S.ex_a[keyed[0][y].second];   /* reference slot */
        vf[y] = make_uint4(P->h_slot_dev[sl], P->h_inv_perm[S.slot_var[sl]], 0u, 0u);
    }
    for (size_t y = 0; y < keyed[1].size(); y++) {
        const uint32_t e = S.ex_a[keyed[1][y].second];
        const uint32_t own = (uint32_t)P->h_var_edge[e], I = slot_fac[own];
        const uint32_t a = (uint32_t)P->h_fac_off[I], d = (uint32_t)(P->h_fac_off[I + 1] - a);
        const uint32_t meta = P->h_meta[I];   /* type and head slot never change */
        fv[y] = make_uint4(e, P->h_inv_perm[S.slot_var[P->h_slot_ref[own]]], I, a);
        faux[y] = make_uint4(d | ((meta & 3u) << 24), own - a, meta >> 8, 0u);
    }
    /* group items: one lane (one record) per slot, a group's lanes inside
     * one 64-lane wave, each level padded to whole waves; glane[l] = first
     * lane of level l */
    std::vector<uint4> grp;
    std::vector<uint32_t> glane(nl + 2, 0);
    static_assert(flbp::kSeqGroupDeg <= 6, "group lane record: degree, head and slot in 3 bits, masks in 6");
    for (uint32_t l = 1; l <= nl; l++) {
        const size_t base = grp.size();
        auto pad_to_wave = [&]() {
            while ((grp.size() - base) % 64) grp.push_back(make_uint4(0u, 0u, 0u, 0u));
        };
        for (uint32_t y = off[2][l]; y < off[2][l + 1]; y++) {
            const size_t x = keyed[2][y].second;
            const uint32_t I = P->h_finv[S.ex_a[x]];   /* device factor */
            const uint32_t a = (uint32_t)P->h_fac_off[I], d = (uint32_t)(P->h_fac_off[I + 1] - a);
            const uint32_t meta = P->h_meta[I];
            if ((grp.size() - base) % 64 + d > 64) pad_to_wave();
            for (uint32_t k = 0; k < d; k++)
                grp.push_back(make_uint4(I, a, flbp::dev::group_word(d, meta & 3u, meta >> 8, S.ex_b[x], S.ex_c[x], k),
                                         P->h_inv_perm[S.slot_var[P->h_slot_ref[a + k]]]));
        }
        pad_to_wave();
        if (grp.size() >= 0xffffffffull)
            return flbp::fail(FLBP_EINVAL, "flbp_plan_set_schedule: too many group lanes for 32-bit offsets");
        glane[l + 1] = (uint32_t)grp.size();
    }
    P->free_seq_items();
    auto dev_alloc = [&](auto *&ptr, size_t n) {
        return hipMalloc((void **)&ptr, std::max<size_t>(n, 1) * sizeof(*ptr)) == hipSuccess;
    };
    if (!dev_alloc(P->seq_vf, vf.size()) || !dev_alloc(P->seq_fv, fv.size()) || !dev_alloc(P->seq_fv_aux, faux.size()) ||
        !dev_alloc(P->seq_grp, grp.size()) || !dev_alloc(P->seq_lvl_vf_dev, nl + 1) ||
        !dev_alloc(P->seq_lvl_fv_dev, nl + 1) || !dev_alloc(P->seq_lvl_grp_dev, nl + 1)) {
        (void)hipGetLastError();
        P->free_seq_items();
        return flbp::fail(FLBP_ENOMEM, "flbp_plan_set_schedule: item upload allocation failed");
    }
    if (!vf.empty()) HIP_TRY(hipMemcpy(P->seq_vf, vf.data(), vf.size() * sizeof(uint4), hipMemcpyHostToDevice));
    if (!fv.empty()) {
        HIP_TRY(hipMemcpy(P->seq_fv, fv.data(), fv.size() * sizeof(uint4), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P->seq_fv_aux, faux.data(), faux.size() * sizeof(uint4), hipMemcpyHostToDevice));
    }
    if (!grp.empty()) HIP_TRY(hipMemcpy(P->seq_grp, grp.data(), grp.size() * sizeof(uint4), hipMemcpyHostToDevice));
    /* level l (1-based) -> index l-1 */
    P->seq_lvl_vf.assign(off[0].begin() + 1, off[0].end());
    P->seq_lvl_fv.assign(off[1].begin() + 1, off[1].end());
    P->seq_lvl_grp.assign(glane.begin() + 1, glane.end());
    HIP_TRY(hipMemcpy(P->seq_lvl_vf_dev, P->seq_lvl_vf.data(), (nl + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(P->seq_lvl_fv_dev, P->seq_lvl_fv.data(), (nl + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(P->seq_lvl_grp_dev, P->seq_lvl_grp.data(), (nl + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    P->seq_place_tail();
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

int flbp_plan_set_seq_tail(flbp_plan *plan, uint32_t max_items) {
    Plan *P = reinterpret_cast<Plan *>(plan);
    if (!P) return flbp::fail(FLBP_EINVAL, "flbp_plan_set_seq_tail: null plan");
    if (max_items > kSeqTailItems)
        return flbp::fail(FLBP_EINVAL, "flbp_plan_set_seq_tail: at most " + std::to_string(kSeqTailItems) +
                                           " items per tail level");
    HIP_TRY(hipSetDevice(P->device));
    HIP_TRY(hipStreamSynchronize(P->stream));
    P->drop_graphs();
    P->seq_tail_max = max_items;
    P->seq_place_tail();
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
