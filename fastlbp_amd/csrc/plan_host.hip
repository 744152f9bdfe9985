/*
 * plan_host.hip -- CSR validation and the upload of a graph into a plan's
 * device layout (the H->D half of CausalBP::construct,
 * csrc/ext/causal_bp.cu:368-369, 815-816).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "fastlbp.h"
#include "flbp_internal.h"
#include "plan_internal.h"
#include "sweep_kernels.h"

namespace flbp {

int validate_csr(const flbp_csr &c) {
    const uint64_t V = c.n_vars, F = c.n_factors, E = c.n_edges;
    auto bad = [](const std::string &m) { return fail(FLBP_EINVAL, "invalid CSR: " + m); };
    if (!c.fac_off || !c.var_off || (E && (!c.edge_var || !c.var_edge)) ||
        (F && (!c.fac_type || !c.head_slot || !c.head_clamped || !c.p || !c.q || !c.mask0 || !c.mask1)))
        return bad("null array");
    if (c.fac_off[0] != 0 || c.fac_off[F] != E) return bad("fac_off must run from 0 to n_edges");
    if (c.var_off[0] != 0 || c.var_off[V] != E) return bad("var_off must run from 0 to n_edges");
    for (uint64_t I = 0; I < F; I++) {
        if (c.fac_off[I + 1] <= c.fac_off[I]) return bad("every factor needs at least one variable");
        if (c.fac_type[I] < 0 || c.fac_type[I] > 2) return bad("factor type must be 0, 1 or 2");
        if (c.head_slot[I] < c.fac_off[I] || c.head_slot[I] >= c.fac_off[I + 1])
            return bad("head slot outside its factor");
        if (c.fac_off[I + 1] - c.fac_off[I] >= (1u << 24)) return fail(FLBP_ERANGE, "factor degree >= 2^24");
    }
    for (uint64_t i = 0; i < V; i++)
        if (c.var_off[i + 1] < c.var_off[i]) return bad("var_off not monotone");
    /* var_edge must be a permutation of the slots that agrees with
     * edge_var: checked by threads over variable ranges of equal edge
     * counts, marking slots in a shared bitmap (the reads are random, so
     * threads scale them) */
    std::vector<std::atomic<uint64_t>> seen((E + 63) / 64);
    for (auto &w : seen) w.store(0, std::memory_order_relaxed);
    std::atomic<int> err{0};   /* 1 not a permutation, 2 disagreement */
    auto check = [&](uint64_t v0, uint64_t v1) {
        for (uint64_t i = v0; i < v1 && !err.load(std::memory_order_relaxed); i++) {
            for (uint64_t k = c.var_off[i]; k < c.var_off[i + 1]; k++) {
                const uint64_t s = c.var_edge[k];
                if (s >= E) {
                    err.store(1);
                    return;
                }
                const uint64_t bit = 1ull << (s & 63);
                if (seen[s >> 6].fetch_or(bit, std::memory_order_relaxed) & bit) {
                    err.store(1);
                    return;
                }
                if (c.edge_var[s] != i) {
                    err.store(2);
                    return;
                }
            }
        }
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nt = E < (1u << 20) ? 1u : hw;
    if (nt == 1) {
        check(0, V);
    } else {
        std::vector<std::thread> pool;
        uint64_t v0 = 0;
        for (unsigned t = 0; t < nt; t++) {
            /* first variable whose edges start at or after (t+1)/nt of E */
            const uint64_t target = E / nt * (t + 1);
            const uint64_t v1 = t + 1 == nt ? V
                                            : (uint64_t)(std::lower_bound(c.var_off, c.var_off + V + 1, target) - c.var_off);
            const uint64_t a = v0, b = std::max(v0, std::min(v1, V));
            pool.emplace_back(check, a, b);
            v0 = b;
        }
        for (auto &th : pool) th.join();
    }
    if (err.load() == 1) return bad("var_edge is not a permutation of the slots");
    if (err.load() == 2) return bad("var_edge and edge_var disagree");
    return FLBP_OK;
}

void var_device_order(const flbp_csr &c, int order, std::vector<uint32_t> &perm,
                      std::vector<uint32_t> &dev_off, std::vector<uint32_t> &dev_edge,
                      std::vector<uint32_t> &vpos) {
    const uint64_t V = c.n_vars;
    perm.resize(V);
    if (order == kOrderHotFirst) {
        /* variables of degree >= kHotDegree first, then the rest, label order
         * within each class */
        uint64_t k = 0;
        for (uint64_t i = 0; i < V; i++)
            if (c.var_off[i + 1] - c.var_off[i] >= kHotDegree) perm[k++] = (uint32_t)i;
        for (uint64_t i = 0; i < V; i++)
            if (c.var_off[i + 1] - c.var_off[i] < kHotDegree) perm[k++] = (uint32_t)i;
    } else if (order == kOrderDegree) {
        uint64_t maxd = 0;
        for (uint64_t i = 0; i < V; i++) maxd = std::max<uint64_t>(maxd, c.var_off[i + 1] - c.var_off[i]);
        /* counting sort by degree, heaviest first, stable in variable index */
        std::vector<uint64_t> start(maxd + 2, 0);
        for (uint64_t i = 0; i < V; i++) start[maxd - (c.var_off[i + 1] - c.var_off[i]) + 1]++;
        for (uint64_t b = 0; b <= maxd; b++) start[b + 1] += start[b];
        for (uint64_t i = 0; i < V; i++) perm[start[maxd - (c.var_off[i + 1] - c.var_off[i])]++] = (uint32_t)i;
    } else {
        for (uint64_t i = 0; i < V; i++) perm[i] = (uint32_t)i;
    }
    dev_off.assign(V + 1, 0);
    dev_edge.resize(c.n_edges);
    vpos.resize(c.n_edges);
    for (uint64_t dv = 0; dv < V; dv++) {
        const uint64_t v = perm[dv], a = c.var_off[v], b = c.var_off[v + 1];
        dev_off[dv + 1] = dev_off[dv] + (uint32_t)(b - a);
        for (uint64_t k = a; k < b; k++) {
            const uint32_t pos = dev_off[dv] + (uint32_t)(k - a);
            dev_edge[pos] = (uint32_t)c.var_edge[k];
            vpos[c.var_edge[k]] = pos;
        }
    }
}

void factor_device_order(const flbp_csr &c, const std::vector<uint32_t> &var_inv, DeviceCsr &out,
                         std::vector<uint32_t> &fperm, std::vector<uint32_t> &finv, std::vector<uint32_t> &slot_dev,
                         std::vector<uint32_t> &slot_ref) {
    const uint64_t V = c.n_vars, F = c.n_factors, E = c.n_edges;
    /* counting sort of the factors by their head's device index (stable) */
    std::vector<uint64_t> start(V + 1, 0);
    for (uint64_t I = 0; I < F; I++) start[var_inv[c.edge_var[c.head_slot[I]]] + 1]++;
    for (uint64_t v = 0; v < V; v++) start[v + 1] += start[v];
    fperm.resize(F);
    finv.resize(F);
    for (uint64_t I = 0; I < F; I++) {
        const uint64_t d = start[var_inv[c.edge_var[c.head_slot[I]]]]++;
        fperm[d] = (uint32_t)I;
        finv[I] = (uint32_t)d;
    }
    out.fac_off.assign(F + 1, 0);
    out.edge_var.resize(E);
    out.head_slot.resize(F);
    out.fac_type.resize(F);
    out.head_clamped.resize(F);
    out.p.resize(F);
    out.q.resize(F);
    out.mask0.resize(F);
    out.mask1.resize(F);
    slot_dev.resize(E);
    slot_ref.resize(E);
    for (uint64_t d = 0; d < F; d++) {
        const uint64_t I = fperm[d], a = c.fac_off[I], n = c.fac_off[I + 1] - a, b = out.fac_off[d];
        out.fac_off[d + 1] = b + n;
        for (uint64_t j = 0; j < n; j++) {
            slot_dev[a + j] = (uint32_t)(b + j);
            slot_ref[b + j] = (uint32_t)(a + j);
            out.edge_var[b + j] = c.edge_var[a + j];
        }
        out.head_slot[d] = b + (c.head_slot[I] - a);
        out.fac_type[d] = c.fac_type[I];
        out.head_clamped[d] = c.head_clamped[I];
        out.p[d] = c.p[I];
        out.q[d] = c.q[I];
        out.mask0[d] = c.mask0[I];
        out.mask1[d] = c.mask1[I];
    }
    out.var_edge.resize(E);
    for (uint64_t k = 0; k < E; k++) out.var_edge[k] = slot_dev[c.var_edge[k]];
    flbp_csr &w = out.view;
    w = c;   /* sizes, labels and var_off are the reference's */
    w.fac_off = out.fac_off.data();
    w.edge_var = out.edge_var.data();
    w.var_edge = out.var_edge.data();
    w.fac_type = out.fac_type.data();
    w.head_slot = out.head_slot.data();
    w.head_clamped = out.head_clamped.data();
    w.p = out.p.data();
    w.q = out.q.data();
    w.mask0 = out.mask0.data();
    w.mask1 = out.mask1.data();
}

void cut_blocks(const uint32_t *off, uint64_t n_items, uint32_t max_items, uint32_t max_slots,
                std::vector<uint32_t> &blk) {
    blk.clear();
    blk.push_back(0);
    uint64_t i = 0;
    while (i < n_items) {
        uint64_t j = i, slots = 0;
        while (j < n_items && j - i < max_items) {
            const uint64_t d = off[j + 1] - off[j];
            if (j > i && slots + d > max_slots) break;
            slots += d;
            j++;
            if (slots > max_slots) break;   /* a wide item stays alone */
        }
        blk.push_back((uint32_t)j);
        i = j;
    }
}

void stage_work_orders(const uint32_t *fac_off, const uint32_t *fac_meta, const std::vector<uint32_t> &fb,
                       const uint32_t *var_off, const std::vector<uint32_t> &vb, std::vector<uint32_t> &fac_work,
                       std::vector<uint16_t> &var_work, const uint32_t *slot_var, uint32_t reg_deg) {
    const uint64_t E = fac_off[fb.back()];
    fac_work.assign(E, 0);
    std::vector<std::pair<uint64_t, uint32_t>> keyed;
    for (size_t b = 0; b + 1 < fb.size(); b++) {
        const uint32_t f0 = fb[b], f1 = fb[b + 1], base = fac_off[f0];
        const uint32_t n = fac_off[f1] - base;
        if (n > kStageSlots) continue;   /* a wide factor: no work order */
        keyed.clear();
        for (uint32_t f = f0; f < f1; f++) {
            const uint32_t d = fac_off[f + 1] - fac_off[f];
            const uint32_t type = fac_meta[f] & 3u, h = fac_meta[f] >> 8;
            for (uint32_t t = 0; t < d; t++) {
                const uint32_t kind = type == 0 ? 0u : (t == h ? 1u : 2u);
                const uint32_t sl = fac_off[f] - base + t;
                const bool by_var = slot_var && kind == 2 && d <= reg_deg;
                const uint64_t cls = by_var ? 0u : d, sec = by_var ? slot_var[fac_off[f] + t] : sl;
                keyed.push_back({((uint64_t)kind << 62) | (cls << 42) | (sec << 10) | sl, sl | ((f - f0) << 10)});
            }
        }
        std::sort(keyed.begin(), keyed.end());
        for (uint32_t i = 0; i < n; i++) fac_work[base + i] = keyed[i].second;
    }
    var_work_order(var_off, vb, var_work);
}

void var_work_order(const uint32_t *var_off, const std::vector<uint32_t> &vb, std::vector<uint16_t> &var_work) {
    const uint64_t V = vb.back();
    var_work.assign(V, 0);
    std::vector<std::pair<uint64_t, uint16_t>> vk;
    for (size_t b = 0; b + 1 < vb.size(); b++) {
        const uint32_t v0 = vb[b], v1 = vb[b + 1];
        vk.clear();
        for (uint32_t v = v0; v < v1; v++)
            vk.push_back({((uint64_t)(var_off[v + 1] - var_off[v]) << 20) | (v - v0), (uint16_t)(v - v0)});
        std::sort(vk.begin(), vk.end());
        for (uint32_t i = 0; i < v1 - v0; i++) var_work[v0 + i] = vk[i].second;
    }
}

int Plan::stage(uint64_t n) {
    if (n <= stage_cap) return FLBP_OK;
    if (stage_idx) (void)hipFree(stage_idx);
    if (stage_val) (void)hipFree(stage_val);
    stage_idx = nullptr;
    stage_val = nullptr;
    stage_cap = 0;
    const uint64_t cap = std::max<uint64_t>(n, 4096);
    if (hipMalloc(&stage_idx, cap * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&stage_val, cap * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(FLBP_ENOMEM, "plan: staging buffer of " + std::to_string(cap) + " entries failed");
    }
    stage_cap = cap;
    return FLBP_OK;
}

template <class T>
int Plan::alloc(T *&p, uint64_t n) {
    const size_t bytes = std::max<uint64_t>(n, 1) * sizeof(T);
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(FLBP_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                     hipGetErrorString(e));
    }
    allocs.push_back(q);
    device_bytes += bytes;
    p = static_cast<T *>(q);
    return FLBP_OK;
}

/* element types allocated from other translation units */
template int Plan::alloc<double2>(double2 *&, uint64_t);
template int Plan::alloc<dev::VarState>(dev::VarState *&, uint64_t);
template int Plan::alloc<dev::FacRec>(dev::FacRec *&, uint64_t);

#define TRY(expr)          \
    do {                   \
        int rc_ = (expr);  \
        if (rc_) return rc_; \
    } while (0)
#define HIP_TRY(expr)                                     \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

int Plan::upload(const flbp_csr &c_ref) {
    /* factors in device order (factor_device_order); everything below works
     * on that CSR */
    DeviceCsr dcsr;
    {
        std::vector<uint32_t> perm0, off0, edge0, vpos0;
        var_device_order(c_ref, kOrderHotFirst, perm0, off0, edge0, vpos0);
        std::vector<uint32_t> inv0(V);
        for (uint64_t dv = 0; dv < V; dv++) inv0[perm0[dv]] = (uint32_t)dv;
        factor_device_order(c_ref, inv0, dcsr, h_fperm, h_finv, h_slot_dev, h_slot_ref);
    }
    const flbp_csr &c = dcsr.view;
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    TRY(alloc(fac_off, F + 1));
    TRY(alloc(fac_meta, F));
    TRY(alloc(fac_p, F));
    TRY(alloc(fac_q, F));
    TRY(alloc(fac_mask, F));
    TRY(alloc(var_off, V + 1));
    TRY(alloc(var_edge, E));
    TRY(alloc(fac_vpos, E));
    TRY(alloc(m_vf, E));
    TRY(alloc(m_fv, E));
    TRY(alloc(m_fv_vm, E));
    TRY(alloc(belief, V));
    TRY(alloc(vprod, V));
    TRY(alloc(vzc, V));
    TRY(alloc(fac_var, E));
    TRY(alloc(var_src, E));
    TRY(alloc(ctl, (uint64_t)kMaxChunk * dev::kCtlPerSweep));
    TRY(alloc(ctl_out, kMaxChunk));
    HIP_TRY(hipHostMalloc((void **)&host_ctl, sizeof(unsigned long long) * kMaxChunk, hipHostMallocDefault));

    std::vector<uint32_t> u32(std::max(std::max(F, V), E) + 1);
    for (uint64_t I = 0; I <= F; I++) u32[I] = (uint32_t)c.fac_off[I];
    HIP_TRY(hipMemcpy(fac_off, u32.data(), (F + 1) * 4, hipMemcpyHostToDevice));
    std::vector<uint32_t> meta(F);
    for (uint64_t I = 0; I < F; I++)
        meta[I] = dev::pack_meta((uint32_t)c.fac_type[I], c.head_clamped[I] != 0,
                                 (uint32_t)(c.head_slot[I] - c.fac_off[I]));
    if (F) HIP_TRY(hipMemcpy(fac_meta, meta.data(), F * 4, hipMemcpyHostToDevice));
    /* variables in device order; each variable's edges keep their nbV order,
     * so its arithmetic is unchanged */
    std::vector<uint32_t> dev_off, dev_edge;
    var_device_order(c, kOrderHotFirst, h_perm, dev_off, dev_edge, h_fac_vpos);
    HIP_TRY(hipMemcpy(var_off, dev_off.data(), (V + 1) * 4, hipMemcpyHostToDevice));
    {   /* staged-kernel block ranges */
        std::vector<uint32_t> fo(F + 1), vb, fb;
        for (uint64_t I = 0; I <= F; I++) fo[I] = (uint32_t)c.fac_off[I];
        cut_blocks(fo.data(), F, kStageItems, kStageSlots, fb);
        cut_blocks(dev_off.data(), V, kStageItems, kStageSlots, vb);
        n_fac_blocks = fb.size() - 1;
        n_var_blocks = vb.size() - 1;
        TRY(alloc(fac_blk, fb.size()));
        TRY(alloc(var_blk, vb.size()));
        HIP_TRY(hipMemcpy(fac_blk, fb.data(), fb.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(var_blk, vb.data(), vb.size() * 4, hipMemcpyHostToDevice));
        std::vector<uint32_t> fw;
        std::vector<uint16_t> vw;
        std::vector<uint32_t> inv(V), sv(E);
        for (uint64_t dv = 0; dv < V; dv++) inv[h_perm[dv]] = (uint32_t)dv;
        uint64_t maxd = 0;
        for (uint64_t I = 0; I < F; I++) maxd = std::max<uint64_t>(maxd, c.fac_off[I + 1] - c.fac_off[I]);
        for (uint64_t s = 0; s < E; s++) sv[s] = inv[c.edge_var[s]];
        /* the register degree bound the sweep launches pick (plan.hip) */
        const uint32_t reg_deg = maxd <= (uint64_t)kFacRegDegSmall ? kFacRegDegSmall : kFacRegDeg;
        stage_work_orders(fo.data(), meta.data(), fb, dev_off.data(), vb, fw, vw, sv.data(), reg_deg);
        TRY(alloc(fac_work, E));
        TRY(alloc(var_work, V));
        {   /* the state layout's variable tiles */
            std::vector<uint32_t> vb4;
            std::vector<uint16_t> vw4;
            cut_blocks(dev_off.data(), V, kStageItems, kStateVarSlots, vb4);
            var_work_order(dev_off.data(), vb4, vw4);
            n_var_blocks4 = vb4.size() - 1;
            TRY(alloc(var_blk4, vb4.size()));
            TRY(alloc(var_work4, V));
            HIP_TRY(hipMemcpy(var_blk4, vb4.data(), vb4.size() * 4, hipMemcpyHostToDevice));
            if (V) HIP_TRY(hipMemcpy(var_work4, vw4.data(), V * 2, hipMemcpyHostToDevice));
        }
        if (E) HIP_TRY(hipMemcpy(fac_work, fw.data(), E * 4, hipMemcpyHostToDevice));
        if (V) HIP_TRY(hipMemcpy(var_work, vw.data(), V * 2, hipMemcpyHostToDevice));
        /* the state layout's storage order: a staged tile's slots in its
         * work order, a wide factor's in slot order */
        h_store.resize(E);
        for (size_t b = 0; b + 1 < fb.size(); b++) {
            const uint32_t base = fo[fb[b]], n = fo[fb[b + 1]] - base;
            for (uint32_t i = 0; i < n; i++)
                h_store[n > kStageSlots ? base + i : base + (fw[base + i] & 1023u)] = base + i;
        }
    }
    if (E) {
        HIP_TRY(hipMemcpy(var_edge, dev_edge.data(), E * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(fac_vpos, h_fac_vpos.data(), E * 4, hipMemcpyHostToDevice));
        std::vector<uint32_t> inv(V), fv(E);
        for (uint64_t dv = 0; dv < V; dv++) inv[h_perm[dv]] = (uint32_t)dv;
        for (uint64_t s = 0; s < E; s++) fv[h_store[s]] = inv[c.edge_var[s]];
        HIP_TRY(hipMemcpy(fac_var, fv.data(), E * 4, hipMemcpyHostToDevice));
        for (uint64_t k = 0; k < E; k++) fv[k] = h_store[dev_edge[k]];
        HIP_TRY(hipMemcpy(var_src, fv.data(), E * 4, hipMemcpyHostToDevice));
    }
    if (F) {
        HIP_TRY(hipMemcpy(fac_p, c.p, F * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(fac_q, c.q, F * 8, hipMemcpyHostToDevice));
        std::vector<double2> mk(F);
        for (uint64_t I = 0; I < F; I++) mk[I] = make_double2(c.mask0[I], c.mask1[I]);
        HIP_TRY(hipMemcpy(fac_mask, mk.data(), F * 16, hipMemcpyHostToDevice));
        h_mask = mk;
        h_mask0.swap(mk);
    }
    /* old beliefs are value-initialised to 0.0 (causal_bp.cu:371) */
    HIP_TRY(hipMemset(belief, 0, std::max<uint64_t>(V, 1) * 8));
    /* messages are zero until init(), as the reference's freshly resized
     * device_vectors are */
    HIP_TRY(hipMemset(m_vf, 0, std::max<uint64_t>(E, 1) * 16));
    HIP_TRY(hipMemset(m_fv, 0, std::max<uint64_t>(E, 1) * 16));
    HIP_TRY(hipMemset(m_fv_vm, 0, std::max<uint64_t>(E, 1) * 16));
    h_var_edge.assign(c.var_edge, c.var_edge + E);
    h_fac_off.assign(c.fac_off, c.fac_off + F + 1);
    for (uint64_t I = 0; I < F; I++) max_fac_deg = std::max<uint64_t>(max_fac_deg, c.fac_off[I + 1] - c.fac_off[I]);
    for (uint64_t i = 0; i < V; i++) max_var_deg = std::max<uint64_t>(max_var_deg, c.var_off[i + 1] - c.var_off[i]);
    h_meta0 = meta;
    h_meta.swap(meta);
    /* head of each factor, body uses of each variable, factors per head */
    h_fac_head.assign(F, 0);
    h_var_body.assign(V, 0);
    h_head_off.assign(V + 1, 0);
    for (uint64_t I = 0; I < F; I++) {
        h_fac_head[I] = (uint32_t)c.edge_var[c.head_slot[I]];
        h_head_off[h_fac_head[I] + 1]++;
        for (uint64_t s = c.fac_off[I]; s < c.fac_off[I + 1]; s++)
            if (s != c.head_slot[I]) h_var_body[c.edge_var[s]] = 1;
    }
    for (uint64_t i = 0; i < V; i++) h_head_off[i + 1] += h_head_off[i];
    h_head_fac.assign(F, 0);
    {
        std::vector<uint32_t> fill(h_head_off.begin(), h_head_off.end() - 1);
        for (uint64_t I = 0; I < F; I++) h_head_fac[fill[h_fac_head[I]]++] = (uint32_t)I;
    }
    h_inv_perm.assign(V, 0);
    for (uint64_t dv = 0; dv < V; dv++) h_inv_perm[h_perm[dv]] = (uint32_t)dv;
    TRY(alloc(var_perm, V));
    if (V) HIP_TRY(hipMemcpy(var_perm, h_perm.data(), V * 4, hipMemcpyHostToDevice));
    /* per solve on one MI355X (generated graphs, E = 4.5 V, damping 0.3):
     * 100k variables (40 MB) scatter 1.20 ms, state 1.31 ms; 300k (120 MB)
     * state 2.03, scatter 2.23; 1M state 5.82 ms, pull 7.40; 3M state 20.7,
     * pull 23.8, scatter 29.4 */
    variant = device_bytes <= kStateMinBytes ? 3 : 4;
    return FLBP_OK;
}

}  // namespace flbp
