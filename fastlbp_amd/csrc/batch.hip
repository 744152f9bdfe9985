/*
 * batch.hip -- batched PARALL solve of many graphs and clamped variants in
 * one device state (no counterpart in the reference, whose CausalBP solves
 * one graph per object; its only capture precedent is MessageSweepGraph,
 * csrc/ext/causal_bp.cu:82-112, and a variant there is clamp -> rebuild ->
 * run, examples/fl_wrapper.cu:26-89).
 *
 * The batch holds groups (see batch_kernels.h): a group is one structure
 * and K variants of it, K-interleaved.  Independent graphs are groups of one
 * variant.  Each sweep is the per-variant stop kernel, the active-tile
 * expansion, a factor pass and a variable pass; the stop test and the list
 * of running groups live on the device, so no sweep waits for the host.
 * Sweeps are captured into hipGraphs of `chunk` sweeps (default 8) and
 * replayed back to back, at most two replays ahead of the host, which
 * reads one word per replay (how many groups still run) and stops
 * enqueuing once none does; chunk 0 replays the whole solve, up to the
 * iteration cap, as one launch (its sweeps after the last variant stopped
 * find no active tile and exit at once, but each still dispatches its
 * grid).  The tile slot bounds are chosen by flbp_batch_autotune.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "batch_kernels.h"
#include "fastlbp.h"
#include "flbp_internal.h"
#include "plan_internal.h"

namespace flbp {

struct GroupHost {
    GroupDev d;
    std::vector<uint32_t> fperm;   /* device factor -> factor (factor_device_order) */
    uint64_t n_edges = 0;
    uint64_t perm_base = 0;   /* into h_perm: device variable -> variable */
    uint64_t mask_rows = 0;
};

struct Batch {
    int device = -1;
    hipStream_t stream = nullptr;
    uint64_t G = 0, NV = 0;                  /* groups, variants */
    uint64_t Ftot = 0, Vtot = 0, Etot = 0;   /* structure totals */
    uint64_t Mtot = 0, Stot = 0;             /* K-interleaved message / variable-state totals */
    uint64_t max_fac_deg = 0;
    std::vector<GroupHost> hg;
    std::vector<uint32_t> h_fac_off, h_fac_meta, h_var_off;   /* host copies for tiling */
    /* per group, by its structure edges: device variable of each slot, slot
     * of each var-major edge, and (once tiles are cut) the storage position
     * of each slot -- a factor tile keeps its messages in its work order */
    std::vector<uint32_t> h_slot_var, h_var_slot, h_store;
    bool msgs_live = false;   /* messages set by init or a run: a new storage order moves them */
    std::vector<uint32_t> h_perm;
    std::vector<uint32_t> var_group;         /* variant -> group */
    GroupDev *groups = nullptr;
    uint32_t *fac_off = nullptr, *fac_meta = nullptr, *fac_var = nullptr, *fac_work = nullptr, *fac_crow = nullptr;
    double *fac_p = nullptr, *fac_q = nullptr;
    double2 *masks = nullptr;
    uint32_t *var_off = nullptr, *var_edge = nullptr;
    uint16_t *var_work = nullptr;
    double2 *m_fv = nullptr, *scratch = nullptr;
    double *vprod = nullptr;
    uint8_t *vzc = nullptr;
    double *belief = nullptr;
    uint8_t *vrun = nullptr, *vfresh = nullptr;
    uint32_t *budget = nullptr;
    unsigned long long *iters_dev = nullptr, *diff = nullptr;
    double *last_dev = nullptr;
    uint32_t *act_n = nullptr, *act_group = nullptr, *act_fpre = nullptr, *act_vpre = nullptr;
    Tile *ftiles = nullptr, *vtiles = nullptr;
    Tile *act_ftiles = nullptr, *act_vtiles = nullptr;   /* the running groups' tiles */
    uint32_t *act_count = nullptr;                        /* [2] */
    uint64_t cap_ftiles = 0, cap_vtiles = 0, cap_scratch = 0;
    uint32_t n_ftiles = 0, n_vtiles = 0;
    uint32_t cut_f = 0, cut_v = 0;
    double init_c = 1.0;
    /* pinned host staging */
    uint8_t *h_run = nullptr;
    uint32_t *h_budget = nullptr;
    unsigned long long *h_iters = nullptr;
    double *h_last = nullptr;
    uint32_t *h_actn = nullptr;   /* act_n after each chunk launch */
    std::vector<uint64_t> iters;
    std::vector<double> last_diff;
    std::vector<void *> allocs;
    uint64_t bytes = 0;
    flbp_batch_opts opts{};
    /* the captured solve */
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    uint32_t exec_n = 0, exec_fgrid = 0, exec_vgrid = 0;
    double exec_damp = -1, exec_tol = 0;

    template <class T>
    int alloc(T *&p, uint64_t n) {
        void *q = nullptr;
        const size_t b = std::max<uint64_t>(n, 1) * sizeof(T);
        hipError_t e = hipMalloc(&q, b);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(FLBP_ENOMEM, "batch: hipMalloc of " + std::to_string(b) + " bytes failed");
        }
        allocs.push_back(q);
        bytes += b;
        p = static_cast<T *>(q);
        return FLBP_OK;
    }
    void release(void *q) {
        auto it = std::find(allocs.begin(), allocs.end(), q);
        if (it != allocs.end()) allocs.erase(it);
        (void)hipFree(q);
    }
    void drop_exec() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~Batch() {
        if (device >= 0) (void)hipSetDevice(device);
        drop_exec();
        for (void *q : allocs) (void)hipFree(q);
        for (void *q : {(void *)h_run, (void *)h_budget, (void *)h_iters, (void *)h_last, (void *)h_actn})
            if (q) (void)hipHostFree(q);
        if (stream) (void)hipStreamDestroy(stream);
    }
    BatchArgs args() const {
        BatchArgs a;
        a.groups = groups;
        a.act_ftiles = act_ftiles;
        a.act_vtiles = act_vtiles;
        a.act_count = act_count;
        a.fac_off = fac_off;
        a.fac_meta = fac_meta;
        a.fac_var = fac_var;
        a.fac_work = fac_work;
        a.fac_crow = fac_crow;
        a.fac_p = fac_p;
        a.fac_q = fac_q;
        a.masks = masks;
        a.var_off = var_off;
        a.var_edge = var_edge;
        a.var_work = var_work;
        a.m_fv = m_fv;
        a.vprod = vprod;
        a.vzc = vzc;
        a.belief = belief;
        a.scratch = scratch;
        a.vrun = vrun;
        a.vfresh = vfresh;
        a.diff = diff;
        a.c = init_c;
        return a;
    }
    StopArgs stop_args(double tol) const {
        return StopArgs{groups, (uint32_t)G, (uint32_t)NV, vrun, vfresh, budget, iters_dev, last_dev, diff,
                        act_n, act_group, act_fpre, act_vpre, tol};
    }
    int build_tiles();
    int move_messages(const std::vector<uint32_t> &store);
    /* after batch_stop: the flat lists of the running groups' tiles */
    void expand() {
        const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n_ftiles + kBlock - 1) / kBlock, 1024));
        hipLaunchKernelGGL(batch_expand, dim3(grid), dim3(kBlock), 0, stream, groups, ftiles, vtiles, act_n, act_group,
                           act_fpre, act_vpre, act_ftiles, act_vtiles, act_count);
    }
    int record_sweeps(uint32_t n, double tol, double damping);
    int launch_solve(uint32_t n, double tol, double damping);
    int start_run(uint64_t maxiter, double tol);
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

/* Tiles of every group for the current slot bounds: a K-variant group's
 * tiles hold at most bound / K slots (and 256 / K variables per variable
 * tile), twice that for K >= kDirectK, so each tile fills a block's LDS
 * slots (batch_tile_cap) with all K variants;
 * per tile the plain solve's work orders.  A factor wider than its bound
 * gets a tile of its own and a scratch range for its v->f inputs.  A
 * factor tile stores its messages in its work order (a wide factor's in
 * slot order), which fixes fac_var and var_edge; messages already set by
 * init or a run move to the new order. */
int Batch::build_tiles() {
    const uint32_t sf = opts.tile_factors, sv = opts.tile_vars;
    if (sf == cut_f && sv == cut_v && ftiles) return FLBP_OK;
    drop_exec();
    std::vector<std::vector<Tile>> ft(G), vt(G);
    std::vector<uint32_t> store(Etot);
    /* the body outputs that share one register path go in variable order,
     * as in a plan (the register degree the sweep launches pick) */
    const uint32_t reg_deg = max_fac_deg <= (uint64_t)kFacRegDegSmall ? kFacRegDegSmall : kFacRegDeg;
    std::vector<uint64_t> scr(G, 0);
    std::vector<uint32_t> fw(Etot);
    std::vector<uint16_t> vw(Vtot);
    auto cut = [&](uint64_t g) {
        const GroupDev &d = hg[g].d;
        const uint32_t *fo = h_fac_off.data() + d.fac_off_base;
        const uint32_t *vo = h_var_off.data() + d.var_off_base;
        /* a tile of a group of K >= kDirectK variants holds twice the slots
         * (batch_tile_cap) and up to two variables' records per lane */
        const uint32_t wide = d.K >= kDirectK ? 2u : 1u;
        const uint32_t fs = std::max<uint32_t>(1, (wide * sf) >> d.lk), vs = std::max<uint32_t>(1, (wide * sv) >> d.lk);
        const uint32_t vi = std::max<uint32_t>(1, (wide * kStageItems) >> d.lk);
        std::vector<uint32_t> fb, vb, gfw;
        std::vector<uint16_t> gvw;
        cut_blocks(fo, d.n_fac, kStageItems, fs, fb);
        cut_blocks(vo, d.n_var, vi, vs, vb);
        stage_work_orders(fo, h_fac_meta.data() + d.fac_base, fb, vo, vb, gfw, gvw, h_slot_var.data() + d.edge_base,
                          reg_deg);
        std::copy(gfw.begin(), gfw.end(), fw.begin() + d.edge_base);
        for (size_t b = 0; b + 1 < fb.size(); b++) {
            const uint32_t base = fo[fb[b]], n = fo[fb[b + 1]] - base;
            const bool wide_tile = ((uint64_t)n << d.lk) > (uint64_t)batch_tile_cap(d.K);
            for (uint32_t i = 0; i < n; i++)
                store[d.edge_base + (wide_tile ? base + i : base + (gfw[base + i] & 1023u))] = base + i;
        }
        std::copy(gvw.begin(), gvw.end(), vw.begin() + d.var_base);
        for (size_t b = 0; b + 1 < fb.size(); b++) {
            Tile t{(uint32_t)g, fb[b], fb[b + 1] - fb[b], 0u};
            const uint32_t n = fo[fb[b + 1]] - fo[fb[b]];
            if (((uint64_t)n << d.lk) > (uint64_t)batch_tile_cap(d.K)) {   /* a wide factor: scratch range */
                t.aux = (uint32_t)scr[g];
                scr[g] += n;
            }
            ft[g].push_back(t);
        }
        for (size_t b = 0; b + 1 < vb.size(); b++) vt[g].push_back({(uint32_t)g, vb[b], vb[b + 1] - vb[b], 0u});
    };
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++)
        pool.emplace_back([&, t] {
            for (uint64_t g = t; g < G; g += nth) cut(g);
        });
    for (auto &th : pool) th.join();
    std::vector<Tile> fall, vall;
    uint64_t scr_tot = 0;
    for (uint64_t g = 0; g < G; g++) {
        GroupDev &d = hg[g].d;
        d.ftile_first = (uint32_t)fall.size();
        d.ftile_count = (uint32_t)ft[g].size();
        d.vtile_first = (uint32_t)vall.size();
        d.vtile_count = (uint32_t)vt[g].size();
        for (Tile &t : ft[g])
            if (((uint64_t)(h_fac_off[d.fac_off_base + t.first + t.count] - h_fac_off[d.fac_off_base + t.first])
                 << d.lk) > (uint64_t)batch_tile_cap(d.K))
                t.aux += (uint32_t)scr_tot;
        scr_tot += scr[g];
        fall.insert(fall.end(), ft[g].begin(), ft[g].end());
        vall.insert(vall.end(), vt[g].begin(), vt[g].end());
        if (fall.size() >= (1ull << 31) || vall.size() >= (1ull << 31) || scr_tot >= (1ull << 32))
            return fail(FLBP_ERANGE, "batch: too many tiles");
    }
    /* messages set by a run move to the new order first: a failure leaves
     * the batch as it was */
    if (msgs_live && Etot && store != h_store) TRY(move_messages(store));
    n_ftiles = (uint32_t)fall.size();
    n_vtiles = (uint32_t)vall.size();
    if (n_ftiles > cap_ftiles) {
        if (ftiles) release(ftiles);
        if (act_ftiles) release(act_ftiles);
        ftiles = act_ftiles = nullptr;
        TRY(alloc(ftiles, n_ftiles));
        TRY(alloc(act_ftiles, n_ftiles));
        cap_ftiles = n_ftiles;
    }
    if (n_vtiles > cap_vtiles) {
        if (vtiles) release(vtiles);
        if (act_vtiles) release(act_vtiles);
        vtiles = act_vtiles = nullptr;
        TRY(alloc(vtiles, n_vtiles));
        TRY(alloc(act_vtiles, n_vtiles));
        cap_vtiles = n_vtiles;
    }
    if (scr_tot > cap_scratch || !scratch) {
        if (scratch) release(scratch);
        scratch = nullptr;
        TRY(alloc(scratch, scr_tot));
        cap_scratch = std::max<uint64_t>(scr_tot, 1);
    }
    if (n_ftiles) HIP_TRY(hipMemcpy(ftiles, fall.data(), fall.size() * sizeof(Tile), hipMemcpyHostToDevice));
    if (n_vtiles) HIP_TRY(hipMemcpy(vtiles, vall.data(), vall.size() * sizeof(Tile), hipMemcpyHostToDevice));
    std::vector<GroupDev> gd(G);
    for (uint64_t g = 0; g < G; g++) gd[g] = hg[g].d;
    HIP_TRY(hipMemcpy(groups, gd.data(), G * sizeof(GroupDev), hipMemcpyHostToDevice));
    if (Etot) HIP_TRY(hipMemcpy(fac_work, fw.data(), Etot * 4, hipMemcpyHostToDevice));
    if (Vtot) HIP_TRY(hipMemcpy(var_work, vw.data(), Vtot * 2, hipMemcpyHostToDevice));
    h_store.swap(store);
    if (Etot) {
        std::vector<uint32_t> fv(Etot), ve(Etot);
        for (uint64_t g = 0; g < G; g++) {
            const uint64_t eb = hg[g].d.edge_base, Eg = hg[g].n_edges;
            for (uint64_t e = 0; e < Eg; e++) {
                fv[eb + h_store[eb + e]] = h_slot_var[eb + e];
                ve[eb + e] = h_store[eb + h_var_slot[eb + e]];
            }
        }
        HIP_TRY(hipMemcpy(fac_var, fv.data(), Etot * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(var_edge, ve.data(), Etot * 4, hipMemcpyHostToDevice));
    }
    cut_f = sf;
    cut_v = sv;
    return FLBP_OK;
}

/* messages from the storage order h_store to `store`, every group and
 * variant (through a temporary copy of the message array) */
int Batch::move_messages(const std::vector<uint32_t> &store) {
    uint32_t *d_from = nullptr, *d_to = nullptr;
    double2 *tmp = nullptr;
    auto cleanup = [&]() {
        for (void *q : {(void *)d_from, (void *)d_to, (void *)tmp})
            if (q) (void)hipFree(q);
    };
    if (hipMalloc(&d_from, Etot * 4) != hipSuccess || hipMalloc(&d_to, Etot * 4) != hipSuccess ||
        hipMalloc(&tmp, Mtot * 16) != hipSuccess) {
        (void)hipGetLastError();
        cleanup();
        return fail(FLBP_ENOMEM, "batch: no room to move the messages to the new tiles' order");
    }
    hipError_t e = hipMemcpy(d_from, h_store.data(), Etot * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_to, store.data(), Etot * 4, hipMemcpyHostToDevice);
    for (uint64_t g = 0; g < G && e == hipSuccess; g++) {
        const GroupDev &d = hg[g].d;
        const uint64_t n = hg[g].n_edges << d.lk;
        if (!n) continue;
        hipLaunchKernelGGL(bmove_msgs, dim3(grid_for(n, 256)), dim3(256), 0, stream, m_fv + d.msg_base,
                           tmp + d.msg_base, d_from + d.edge_base, d_to + d.edge_base, n, d.lk);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(m_fv, tmp, Mtot * 16, hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    cleanup();
    return e == hipSuccess ? FLBP_OK : hip_fail(e, "batch: moving messages to the new tiles' order");
}

/* n sweeps: factor pass, variable pass and stop kernel each */
int Batch::record_sweeps(uint32_t n, double tol, double damping) {
    const double keep = 1.0 - damping;
    const BatchArgs a = args();
    const StopArgs st = stop_args(tol);
    const uint32_t fg = std::max<uint32_t>(1, n_ftiles), vg = std::max<uint32_t>(1, n_vtiles);
    const bool narrow = max_fac_deg <= (uint64_t)kFacRegDegSmall;
    for (uint32_t k = 0; k < n; k++) {
        if (damping != 0.0) {
            if (narrow)
                hipLaunchKernelGGL((batch_factor_pass<true, kFacRegDegSmall>), dim3(fg), dim3(kBlock), 0, stream, a,
                                   damping, keep);
            else
                hipLaunchKernelGGL((batch_factor_pass<true, kFacRegDeg>), dim3(fg), dim3(kBlock), 0, stream, a,
                                   damping, keep);
        } else {
            if (narrow)
                hipLaunchKernelGGL((batch_factor_pass<false, kFacRegDegSmall>), dim3(fg), dim3(kBlock), 0, stream, a,
                                   damping, keep);
            else
                hipLaunchKernelGGL((batch_factor_pass<false, kFacRegDeg>), dim3(fg), dim3(kBlock), 0, stream, a,
                                   damping, keep);
        }
        hipLaunchKernelGGL(batch_var_pass, dim3(vg), dim3(kBlock), 0, stream, a);
        hipLaunchKernelGGL(batch_stop, dim3(1), dim3(kStopBlock), 0, stream, st, 1);
        expand();
    }
    HIP_TRY(hipGetLastError());
    return FLBP_OK;
}

/* enqueue n sweeps as one replay of the captured solve (captured on first
 * use for this length, damping, tol and grid) or as direct launches */
int Batch::launch_solve(uint32_t n, double tol, double damping) {
    if (!opts.use_graph) return record_sweeps(n, tol, damping);
    const uint32_t fg = std::max<uint32_t>(1, n_ftiles), vg = std::max<uint32_t>(1, n_vtiles);
    if (!exec || exec_n != n || exec_damp != damping || exec_tol != tol || exec_fgrid != fg || exec_vgrid != vg) {
        drop_exec();
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        const int rc = record_sweeps(n, tol, damping);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        graph = g;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            drop_exec();
            return hip_fail(ei, "hipGraphInstantiate");
        }
        exec_n = n;
        exec_damp = damping;
        exec_tol = tol;
        exec_fgrid = fg;
        exec_vgrid = vg;
    }
    HIP_TRY(hipGraphLaunch(exec, stream));
    return FLBP_OK;
}

/* per-variant budgets and run flags for a run up to maxiter, then the
 * first list of running groups */
int Batch::start_run(uint64_t maxiter, double tol) {
    for (uint64_t v = 0; v < NV; v++) {
        const uint64_t left = maxiter > iters[v] ? maxiter - iters[v] : 0;
        h_budget[v] = (uint32_t)std::min<uint64_t>(left, 0xffffffffu);
        h_run[v] = left > 0;
        h_last[v] = INFINITY;
    }
    HIP_TRY(hipMemcpyAsync(budget, h_budget, NV * 4, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(vrun, h_run, NV, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(last_dev, h_last, NV * 8, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(batch_stop, dim3(1), dim3(kStopBlock), 0, stream, stop_args(tol), 0);
    expand();
    HIP_TRY(hipGetLastError());
    return FLBP_OK;
}

}  // namespace flbp

using flbp::Batch;
using flbp::grid_for;
using flbp::hip_fail;

namespace {

int check_opts(const flbp_batch_opts &o, const char *who) {
    if (o.layout != 0) return flbp::fail(FLBP_EINVAL, std::string(who) + ": layout must be 0 (the state layout)");
    if (o.tile_factors < 64 || o.tile_factors > flbp::kStageSlots || o.tile_vars < 64 ||
        o.tile_vars > flbp::kStageSlots)
        return flbp::fail(FLBP_EINVAL, std::string(who) + ": tile bounds must be 64.." +
                                           std::to_string(flbp::kStageSlots) + " slots");
    return FLBP_OK;
}

/* host view of one group's variant clamps, checked against its structure */
struct ClampSet {
    std::vector<std::vector<std::pair<uint64_t, uint8_t>>> per_variant;   /* (variable, value), ascending variable */
};

}  // namespace

extern "C" {

void flbp_batch_default_opts(flbp_batch_opts *o) {
    if (!o) return;
    o->tile_factors = flbp::kStageSlots;
    o->tile_vars = flbp::kStageSlots;
    o->grid_blocks = 0;
    o->chunk = 8;
    o->use_graph = 1;
    o->layout = 0;
}

int flbp_batch_create_groups(uint64_t n_groups, const flbp_batch_group *gs, int device, const flbp_batch_opts *opts,
                             flbp_batch **out) {
    if (!gs || !out || n_groups == 0) return flbp::fail(FLBP_EINVAL, "flbp_batch_create_groups: bad arguments");
    if (n_groups >= (1ull << 31)) return flbp::fail(FLBP_ERANGE, "flbp_batch_create_groups: too many groups");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return flbp::fail(FLBP_ENODEV, "flbp_batch_create_groups: no HIP device");
    if (device < 0 || device >= ndev) return flbp::fail(FLBP_EINVAL, "flbp_batch_create_groups: bad device");
    flbp_batch_opts o;
    if (opts) o = *opts;
    else flbp_batch_default_opts(&o);
    if (int rc = check_opts(o, "flbp_batch_create_groups")) return rc;
    /* validation: structures, variant counts, clamps */
    uint64_t NV = 0, F = 0, V = 0, E = 0, M = 0, S = 0;
    for (uint64_t g = 0; g < n_groups; g++) {
        const flbp_batch_group &bg = gs[g];
        if (!bg.graph) return flbp::fail(FLBP_EINVAL, "flbp_batch_create_groups: null graph");
        const flbp_csr &c = *bg.graph;
        if (c.n_vars >= (1ull << 32) - 1 || c.n_factors >= (1ull << 32) - 1 || c.n_edges >= (1ull << 32) - 1)
            return flbp::fail(FLBP_ERANGE, "flbp_batch_create_groups: a graph has V, F or E >= 2^32-1");
        const uint32_t K = bg.n_variants;
        if (K == 0 || K > 64 || (K & (K - 1)))
            return flbp::fail(FLBP_EINVAL, "flbp_batch_create_groups: variants per group must be 1, 2, 4, ..., 64");
        if (bg.clamp_off && (!bg.clamp_var || !bg.clamp_val) && bg.clamp_off[K] > bg.clamp_off[0])
            return flbp::fail(FLBP_EINVAL, "flbp_batch_create_groups: clamp arrays missing");
        if (int rc = flbp::validate_csr(c)) return rc;
        NV += K;
        F += c.n_factors;
        V += c.n_vars;
        E += c.n_edges;
        M += c.n_edges * K;
        S += c.n_vars * K;
    }
    if (NV >= (1ull << 31)) return flbp::fail(FLBP_ERANGE, "flbp_batch_create_groups: too many variants");
    HIP_TRY(hipSetDevice(device));
    auto *B = new (std::nothrow) Batch();
    if (!B) return flbp::fail(FLBP_ENOMEM, "flbp_batch_create_groups: out of memory");
    std::unique_ptr<Batch> hold(B);
    B->device = device;
    B->G = n_groups;
    B->NV = NV;
    B->Ftot = F;
    B->Vtot = V;
    B->Etot = E;
    B->Mtot = M;
    B->Stot = S;
    B->opts = o;
    B->hg.resize(n_groups);
    {
        uint64_t fb = 0, vb = 0, eb = 0, mb = 0, sb = 0, nb = 0;
        for (uint64_t g = 0; g < n_groups; g++) {
            const flbp_csr &c = *gs[g].graph;
            GroupDev &d = B->hg[g].d;
            d = GroupDev{};
            d.fac_base = fb;
            d.var_base = vb;
            d.edge_base = eb;
            d.fac_off_base = fb + g;
            d.var_off_base = vb + g;
            d.msg_base = mb;
            d.vst_base = sb;
            d.n_fac = (uint32_t)c.n_factors;
            d.n_var = (uint32_t)c.n_vars;
            d.K = gs[g].n_variants;
            d.lk = (uint32_t)__builtin_ctz(d.K);
            d.vbase = (uint32_t)nb;
            B->hg[g].n_edges = c.n_edges;
            B->hg[g].perm_base = vb;
            for (uint64_t k = 0; k < d.K; k++) B->var_group.push_back((uint32_t)g);
            for (uint64_t I = 0; I < c.n_factors; I++)
                B->max_fac_deg = std::max<uint64_t>(B->max_fac_deg, c.fac_off[I + 1] - c.fac_off[I]);
            fb += c.n_factors;
            vb += c.n_vars;
            eb += c.n_edges;
            mb += c.n_edges * d.K;
            sb += c.n_vars * d.K;
            nb += d.K;
        }
    }
    HIP_TRY(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    TRY(B->alloc(B->groups, n_groups));
    TRY(B->alloc(B->fac_off, F + n_groups));
    TRY(B->alloc(B->fac_meta, F));
    TRY(B->alloc(B->fac_var, E));
    TRY(B->alloc(B->fac_work, E));
    TRY(B->alloc(B->fac_crow, F));
    TRY(B->alloc(B->fac_p, F));
    TRY(B->alloc(B->fac_q, F));
    TRY(B->alloc(B->var_off, V + n_groups));
    TRY(B->alloc(B->var_edge, E));
    TRY(B->alloc(B->var_work, V));
    TRY(B->alloc(B->m_fv, M));
    TRY(B->alloc(B->vprod, S));
    TRY(B->alloc(B->vzc, S));
    TRY(B->alloc(B->belief, S));
    TRY(B->alloc(B->vrun, NV));
    TRY(B->alloc(B->vfresh, NV));
    TRY(B->alloc(B->budget, NV));
    TRY(B->alloc(B->iters_dev, NV));
    TRY(B->alloc(B->last_dev, NV));
    TRY(B->alloc(B->diff, NV * kBatchSlots * kCtlPad));
    TRY(B->alloc(B->act_n, 1));
    TRY(B->alloc(B->act_group, n_groups));
    TRY(B->alloc(B->act_fpre, n_groups + 1));
    TRY(B->alloc(B->act_vpre, n_groups + 1));
    TRY(B->alloc(B->act_count, 2));
    HIP_TRY(hipHostMalloc((void **)&B->h_run, NV, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&B->h_budget, NV * 4, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&B->h_iters, NV * 8, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&B->h_last, NV * 8, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&B->h_actn, 4 * 64, hipHostMallocDefault));
    B->h_fac_off.resize(F + n_groups);
    B->h_fac_meta.resize(F);
    B->h_var_off.resize(V + n_groups);
    B->h_slot_var.resize(E);
    B->h_var_slot.resize(E);
    /* structure upload, group by group; head masks per clamped factor */
    std::vector<double2> mask_all;
    std::vector<uint32_t> crow(F, 0);
    for (uint64_t g = 0; g < n_groups; g++) {
        const flbp_batch_group &bg = gs[g];
        /* the group's factors in device order, as a plan keeps them
         * (factor_device_order); clamps name variables, which keep theirs */
        flbp::DeviceCsr dcsr;
        {
            const flbp_csr &c0 = *bg.graph;
            std::vector<uint32_t> perm0, off0, edge0, vpos0, finv, slot_dev, slot_ref;
            flbp::var_device_order(c0, flbp::kOrderHotFirst, perm0, off0, edge0, vpos0);
            std::vector<uint32_t> inv0(c0.n_vars);
            for (uint64_t dv = 0; dv < c0.n_vars; dv++) inv0[perm0[dv]] = (uint32_t)dv;
            flbp::factor_device_order(c0, inv0, dcsr, B->hg[g].fperm, finv, slot_dev, slot_ref);
        }
        const flbp_csr &c = dcsr.view;
        GroupDev &d = B->hg[g].d;
        const uint64_t Fg = c.n_factors, Vg = c.n_vars, Eg = c.n_edges;
        const uint32_t K = d.K;
        /* variables' head factors and body uses, for the clamps */
        std::vector<uint32_t> head_var(Fg);
        std::vector<uint8_t> body_use(Vg, 0);
        for (uint64_t I = 0; I < Fg; I++) {
            head_var[I] = (uint32_t)c.edge_var[c.head_slot[I]];
            for (uint64_t s = c.fac_off[I]; s < c.fac_off[I + 1]; s++)
                if (s != c.head_slot[I]) body_use[c.edge_var[s]] = 1;
        }
        std::vector<std::vector<uint32_t>> heads_of;   /* built on demand */
        std::vector<std::vector<std::pair<uint32_t, uint8_t>>> clamps(K);   /* (factor, value) per variant */
        if (bg.clamp_off) {
            heads_of.assign(Vg, {});
            for (uint64_t I = 0; I < Fg; I++)
                if (c.fac_type[I] != 0) heads_of[head_var[I]].push_back((uint32_t)I);
            for (uint32_t k = 0; k < K; k++) {
                std::vector<std::pair<uint64_t, uint8_t>> cv;
                for (uint64_t x = bg.clamp_off[k]; x < bg.clamp_off[k + 1]; x++) {
                    const uint64_t v = bg.clamp_var[x];
                    if (v >= Vg) return flbp::fail(FLBP_EINVAL, "flbp_batch_create_groups: clamped variable out of range");
                    if (bg.clamp_val[x] > 1) return flbp::fail(FLBP_EINVAL, "flbp_batch_create_groups: clamp values are 0 or 1");
                    if (body_use[v])
                        return flbp::fail(FLBP_EINVAL, "flbp_batch_create_groups: variable " + std::to_string(v) +
                                                           " appears in a factor body; its clamp changes the structure "
                                                           "(rebuild the graph with flbp_clamp and give it its own group)");
                    if (heads_of[v].empty())
                        return flbp::fail(FLBP_EINVAL, "flbp_batch_create_groups: variable " + std::to_string(v) +
                                                           " heads no AND/OR factor");
                    cv.push_back({v, bg.clamp_val[x]});
                }
                /* ascending variable (= label) order, as fl_wrapper applies clamps */
                std::stable_sort(cv.begin(), cv.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
                for (const auto &pv : cv)
                    for (uint32_t I : heads_of[pv.first]) clamps[k].push_back({I, pv.second});
            }
        }
        /* per factor: base clamp, and each variant's mask (gen_clamped's
         * head rule: one-hot, a repeated clamp multiplies) */
        std::vector<uint8_t> any(Fg, 0);
        for (uint64_t I = 0; I < Fg; I++) any[I] = c.head_clamped[I] != 0;
        for (uint32_t k = 0; k < K; k++)
            for (const auto &pr : clamps[k]) any[pr.first] = 1;
        std::vector<int64_t> row(Fg, -1);
        for (uint64_t I = 0; I < Fg; I++) {
            if (!any[I]) continue;
            row[I] = (int64_t)mask_all.size();
            crow[d.fac_base + I] = (uint32_t)mask_all.size();
            for (uint32_t k = 0; k < K; k++)
                mask_all.push_back(c.head_clamped[I] ? make_double2(c.mask0[I], c.mask1[I]) : make_double2(1.0, 1.0));
        }
        std::vector<uint8_t> clamped(Fg, 0);   /* per factor: clamped before the clamp at hand */
        for (uint32_t k = 0; k < K; k++) {
            for (const auto &pr : clamps[k]) clamped[pr.first] = c.head_clamped[pr.first] != 0;
            for (const auto &pr : clamps[k]) {
                const uint32_t I = pr.first;
                const uint8_t x = pr.second;
                double2 &m = mask_all[(uint64_t)row[I] + k];
                double m0 = x == 0 ? 1.0 : 0.0, m1 = x == 1 ? 1.0 : 0.0;
                if (clamped[I]) {   /* a repeated clamp multiplies (causal_factor.cpp:78-81) */
                    m0 *= m.x;
                    m1 *= m.y;
                }
                m = make_double2(m0, m1);
                clamped[I] = 1;
            }
        }
        if (mask_all.size() >= (1ull << 32)) return flbp::fail(FLBP_ERANGE, "flbp_batch_create_groups: too many clamps");
        std::vector<uint32_t> u32(std::max(std::max(Fg, Vg), Eg) + 1);
        for (uint64_t I = 0; I <= Fg; I++) u32[I] = (uint32_t)c.fac_off[I];
        std::copy(u32.begin(), u32.begin() + Fg + 1, B->h_fac_off.begin() + d.fac_off_base);
        HIP_TRY(hipMemcpy(B->fac_off + d.fac_off_base, u32.data(), (Fg + 1) * 4, hipMemcpyHostToDevice));
        for (uint64_t I = 0; I < Fg; I++)
            u32[I] = flbp::dev::pack_meta((uint32_t)c.fac_type[I], any[I] != 0, (uint32_t)(c.head_slot[I] - c.fac_off[I]));
        std::copy(u32.begin(), u32.begin() + Fg, B->h_fac_meta.begin() + d.fac_base);
        if (Fg) {
            HIP_TRY(hipMemcpy(B->fac_meta + d.fac_base, u32.data(), Fg * 4, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(B->fac_p + d.fac_base, c.p, Fg * 8, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(B->fac_q + d.fac_base, c.q, Fg * 8, hipMemcpyHostToDevice));
        }
        std::vector<uint32_t> perm, dev_off, dev_edge, vpos;
        flbp::var_device_order(c, flbp::kOrderHotFirst, perm, dev_off, dev_edge, vpos);
        B->h_perm.insert(B->h_perm.end(), perm.begin(), perm.end());
        std::copy(dev_off.begin(), dev_off.end(), B->h_var_off.begin() + d.var_off_base);
        HIP_TRY(hipMemcpy(B->var_off + d.var_off_base, dev_off.data(), (Vg + 1) * 4, hipMemcpyHostToDevice));
        if (Eg) {   /* fac_var and var_edge go up with the tiles (build_tiles) */
            std::copy(dev_edge.begin(), dev_edge.end(), B->h_var_slot.begin() + d.edge_base);
            std::vector<uint32_t> inv(Vg);
            for (uint64_t dv = 0; dv < Vg; dv++) inv[perm[dv]] = (uint32_t)dv;
            for (uint64_t s = 0; s < Eg; s++) B->h_slot_var[d.edge_base + s] = inv[c.edge_var[s]];
        }
    }
    TRY(B->alloc(B->masks, mask_all.size()));
    if (!mask_all.empty())
        HIP_TRY(hipMemcpy(B->masks, mask_all.data(), mask_all.size() * sizeof(double2), hipMemcpyHostToDevice));
    if (F) HIP_TRY(hipMemcpy(B->fac_crow, crow.data(), F * 4, hipMemcpyHostToDevice));
    /* messages zero until init, old beliefs 0.0, every variant fresh */
    HIP_TRY(hipMemset(B->m_fv, 0, std::max<uint64_t>(M, 1) * 16));
    HIP_TRY(hipMemset(B->belief, 0, std::max<uint64_t>(S, 1) * 8));
    HIP_TRY(hipMemset(B->vprod, 0, std::max<uint64_t>(S, 1) * 8));
    HIP_TRY(hipMemset(B->vzc, 0, std::max<uint64_t>(S, 1)));
    HIP_TRY(hipMemset(B->vfresh, 1, std::max<uint64_t>(NV, 1)));
    HIP_TRY(hipMemset(B->iters_dev, 0, std::max<uint64_t>(NV, 1) * 8));
    HIP_TRY(hipMemset(B->diff, 0, std::max<uint64_t>(NV, 1) * kBatchSlots * kCtlPad * 8));
    B->iters.assign(NV, 0);
    B->last_diff.assign(NV, INFINITY);
    TRY(B->build_tiles());
    *out = reinterpret_cast<flbp_batch *>(hold.release());
    return FLBP_OK;
}

int flbp_batch_create(uint64_t n_graphs, const flbp_csr *const *csrs, int device, const flbp_batch_opts *opts,
                      flbp_batch **out) {
    if (!csrs || !out || n_graphs == 0) return flbp::fail(FLBP_EINVAL, "flbp_batch_create: bad arguments");
    std::vector<flbp_batch_group> gs(n_graphs);
    for (uint64_t g = 0; g < n_graphs; g++) {
        if (!csrs[g]) return flbp::fail(FLBP_EINVAL, "flbp_batch_create: null graph");
        gs[g] = flbp_batch_group{csrs[g], 1u, nullptr, nullptr, nullptr};
    }
    return flbp_batch_create_groups(n_graphs, gs.data(), device, opts, out);
}

void flbp_batch_destroy(flbp_batch *b) { delete reinterpret_cast<Batch *>(b); }

int flbp_batch_init(flbp_batch *b, int logdomain) {
    Batch *B = reinterpret_cast<Batch *>(b);
    if (!B) return flbp::fail(FLBP_EINVAL, "flbp_batch_init: null batch");
    HIP_TRY(hipSetDevice(B->device));
    const double c = logdomain ? 0.0 : 1.0;
    if (B->Mtot) {
        const unsigned grid = (unsigned)std::min<uint64_t>(grid_for(B->Mtot, 256), 65536);
        hipLaunchKernelGGL(bfill_pairs, dim3(grid), dim3(256), 0, B->stream, B->m_fv, B->Mtot, c);
    }
    hipLaunchKernelGGL(bfill_u8, dim3((unsigned)std::min<uint64_t>(grid_for(B->NV, 256), 65536)), dim3(256), 0,
                       B->stream, B->vfresh, B->NV, (uint8_t)1);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(B->stream));
    B->init_c = c;
    B->msgs_live = false;   /* every message holds c: no storage order to keep */
    B->drop_exec();   /* the captured kernels carry the init value */
    return FLBP_OK;
}

int flbp_batch_run(flbp_batch *b, const flbp_run_opts *o, uint64_t *iters_out, double *maxdiff_out,
                   flbp_run_result *total) {
    Batch *B = reinterpret_cast<Batch *>(b);
    if (!B || !o) return flbp::fail(FLBP_EINVAL, "flbp_batch_run: null argument");
    if (!(o->damping >= 0.0 && o->damping < 1.0))
        return flbp::fail(FLBP_EINVAL, "flbp_batch_run: damping must be in [0, 1)");
    HIP_TRY(hipSetDevice(B->device));
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    const uint64_t NV = B->NV;
    uint64_t most = 0;
    for (uint64_t v = 0; v < NV; v++) most = std::max<uint64_t>(most, o->maxiter > B->iters[v] ? o->maxiter - B->iters[v] : 0);
    most = std::min<uint64_t>(most, 0xffffffffu);
    if (most) {
        B->msgs_live = true;
        TRY(B->start_run(o->maxiter, o->tol));
        /* chunk 0: the whole solve as one replay; else replays of `chunk`
         * sweeps, enqueued back to back (at most two ahead of the host,
         * which stops enqueuing once no group runs or maxtime passed) */
        const uint32_t chunk = o->chunk ? o->chunk : B->opts.chunk;
        if (!chunk || chunk >= most) {
            TRY(B->launch_solve((uint32_t)most, o->tol, o->damping));
        } else {
            const uint64_t n_launch = (most + chunk - 1) / chunk;
            std::vector<hipEvent_t> ev;
            int rc = FLBP_OK;
            for (uint64_t i = 0; i < n_launch && !rc; i++) {
                if (i >= 2) {   /* launch i-2 has finished: stop once nothing runs */
                    if (hipEventSynchronize(ev[i - 2]) != hipSuccess) {
                        rc = hip_fail(hipGetLastError(), "hipEventSynchronize");
                        break;
                    }
                    if (B->h_actn[(i - 2) % 64] == 0 || elapsed() >= o->maxtime) break;
                }
                const uint32_t n = (uint32_t)std::min<uint64_t>(chunk, most - i * chunk);
                rc = B->launch_solve(n, o->tol, o->damping);
                if (rc) break;
                hipEvent_t e = nullptr;
                if (hipMemcpyAsync(&B->h_actn[i % 64], B->act_n, 4, hipMemcpyDeviceToHost, B->stream) != hipSuccess ||
                    hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess ||
                    hipEventRecord(e, B->stream) != hipSuccess) {
                    rc = hip_fail(hipGetLastError(), "batch: chunk bookkeeping");
                    if (e) (void)hipEventDestroy(e);
                    break;
                }
                ev.push_back(e);
            }
            (void)hipStreamSynchronize(B->stream);
            for (hipEvent_t e : ev) (void)hipEventDestroy(e);
            if (rc) return rc;
        }
        HIP_TRY(hipMemcpyAsync(B->h_iters, B->iters_dev, NV * 8, hipMemcpyDeviceToHost, B->stream));
        HIP_TRY(hipMemcpyAsync(B->h_last, B->last_dev, NV * 8, hipMemcpyDeviceToHost, B->stream));
        HIP_TRY(hipStreamSynchronize(B->stream));
    }
    uint64_t sweeps = 0;
    double md_max = 0;
    for (uint64_t v = 0; v < NV; v++) {
        const uint64_t before = B->iters[v];
        if (most) {
            B->iters[v] = B->h_iters[v];
            B->last_diff[v] = B->h_last[v];
        } else {
            B->last_diff[v] = INFINITY;
        }
        sweeps = std::max<uint64_t>(sweeps, B->iters[v] - before);
        if (iters_out) iters_out[v] = B->iters[v];
        if (maxdiff_out) maxdiff_out[v] = B->last_diff[v];
        md_max = std::max(md_max, B->last_diff[v]);
    }
    if (total) {
        total->iters = sweeps;
        total->sweeps = sweeps;
        total->maxdiff = md_max;
        total->seconds = elapsed();
    }
    return FLBP_OK;
}

static int variant_slot(const Batch *B, uint64_t v, const flbp::GroupHost *&gh, uint32_t &k) {
    if (v >= B->NV) return flbp::fail(FLBP_EINVAL, "batch: variant index out of range");
    gh = &B->hg[B->var_group[v]];
    k = (uint32_t)(v - gh->d.vbase);
    return FLBP_OK;
}

int flbp_batch_beliefs(flbp_batch *b, uint64_t v, double *out, uint64_t n) {
    Batch *B = reinterpret_cast<Batch *>(b);
    if (!B || (n && !out)) return flbp::fail(FLBP_EINVAL, "flbp_batch_beliefs: bad argument");
    const flbp::GroupHost *gh = nullptr;
    uint32_t k = 0;
    TRY(variant_slot(B, v, gh, k));
    const GroupDev &d = gh->d;
    if (n > d.n_var) return flbp::fail(FLBP_EINVAL, "flbp_batch_beliefs: n > n_vars of the graph");
    HIP_TRY(hipSetDevice(B->device));
    std::vector<double> dev((uint64_t)d.n_var << d.lk);
    if (d.n_var)
        HIP_TRY(hipMemcpyAsync(dev.data(), B->belief + d.vst_base, dev.size() * 8, hipMemcpyDeviceToHost, B->stream));
    HIP_TRY(hipStreamSynchronize(B->stream));
    for (uint64_t dv = 0; dv < d.n_var; dv++) {
        const uint32_t x = B->h_perm[gh->perm_base + dv];
        if (x < n) out[x] = dev[(dv << d.lk) + k];
    }
    return FLBP_OK;
}

int flbp_batch_beliefs_at(flbp_batch *b, uint64_t v, uint64_t n, const uint64_t *vars, double *out) {
    Batch *B = reinterpret_cast<Batch *>(b);
    if (!B || (n && (!vars || !out))) return flbp::fail(FLBP_EINVAL, "flbp_batch_beliefs_at: bad argument");
    const flbp::GroupHost *gh = nullptr;
    uint32_t k = 0;
    TRY(variant_slot(B, v, gh, k));
    if (!n) return FLBP_OK;
    const GroupDev &d = gh->d;
    std::vector<uint32_t> inv;   /* variable -> device variable, for the asked ones */
    std::vector<uint64_t> idx(n);
    {
        std::vector<std::pair<uint64_t, uint64_t>> want(n);
        for (uint64_t i = 0; i < n; i++) {
            if (vars[i] >= d.n_var) return flbp::fail(FLBP_EINVAL, "flbp_batch_beliefs_at: variable out of range");
            want[i] = {vars[i], i};
        }
        std::sort(want.begin(), want.end());
        std::vector<uint32_t> pos(d.n_var);
        for (uint64_t dv = 0; dv < d.n_var; dv++) pos[B->h_perm[gh->perm_base + dv]] = (uint32_t)dv;
        for (uint64_t i = 0; i < n; i++) idx[want[i].second] = d.vst_base + ((uint64_t)pos[want[i].first] << d.lk) + k;
    }
    HIP_TRY(hipSetDevice(B->device));
    uint64_t *didx = nullptr;
    double *dout = nullptr;
    if (hipMalloc(&didx, n * 8) != hipSuccess || hipMalloc(&dout, n * 8) != hipSuccess) {
        (void)hipGetLastError();
        if (didx) (void)hipFree(didx);
        return flbp::fail(FLBP_ENOMEM, "flbp_batch_beliefs_at: staging allocation failed");
    }
    int rc = FLBP_OK;
    if (hipMemcpyAsync(didx, idx.data(), n * 8, hipMemcpyHostToDevice, B->stream) != hipSuccess) rc = FLBP_EHIP;
    if (!rc) {
        hipLaunchKernelGGL(bgather_f64, dim3(grid_for(n, 256)), dim3(256), 0, B->stream, B->belief, didx, dout, n);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, dout, n * 8, hipMemcpyDeviceToHost, B->stream) != hipSuccess ||
            hipStreamSynchronize(B->stream) != hipSuccess)
            rc = FLBP_EHIP;
    }
    (void)hipFree(didx);
    (void)hipFree(dout);
    return rc ? flbp::fail(FLBP_EHIP, "flbp_batch_beliefs_at: device gather failed") : FLBP_OK;
}

int flbp_batch_set_p(flbp_batch *b, uint64_t group, const double *p, uint64_t n) {
    Batch *B = reinterpret_cast<Batch *>(b);
    if (!B || (n && !p) || group >= (B ? B->G : 0)) return flbp::fail(FLBP_EINVAL, "flbp_batch_set_p: bad argument");
    const GroupDev &d = B->hg[group].d;
    if (n != d.n_fac) return flbp::fail(FLBP_EINVAL, "flbp_batch_set_p: n must equal the group's n_factors");
    HIP_TRY(hipSetDevice(B->device));
    if (n) {
        const std::vector<uint32_t> &fperm = B->hg[group].fperm;
        std::vector<double> pd(n);   /* device factor order */
        for (uint64_t f = 0; f < n; f++) pd[f] = p[fperm[f]];
        HIP_TRY(hipMemcpy(B->fac_p + d.fac_base, pd.data(), n * 8, hipMemcpyHostToDevice));
    }
    return FLBP_OK;
}

int flbp_batch_set_opts(flbp_batch *b, const flbp_batch_opts *o) {
    Batch *B = reinterpret_cast<Batch *>(b);
    if (!B || !o) return flbp::fail(FLBP_EINVAL, "flbp_batch_set_opts: null argument");
    if (int rc = check_opts(*o, "flbp_batch_set_opts")) return rc;
    HIP_TRY(hipSetDevice(B->device));
    B->opts = *o;
    B->drop_exec();
    return B->build_tiles();
}

int flbp_batch_get_opts(flbp_batch *b, flbp_batch_opts *o) {
    Batch *B = reinterpret_cast<Batch *>(b);
    if (!B || !o) return flbp::fail(FLBP_EINVAL, "flbp_batch_get_opts: null argument");
    *o = B->opts;
    return FLBP_OK;
}

int flbp_batch_reset_counters(flbp_batch *b) {
    Batch *B = reinterpret_cast<Batch *>(b);
    if (!B) return flbp::fail(FLBP_EINVAL, "flbp_batch_reset_counters: null batch");
    std::fill(B->iters.begin(), B->iters.end(), 0);
    std::fill(B->last_diff.begin(), B->last_diff.end(), INFINITY);
    HIP_TRY(hipSetDevice(B->device));
    HIP_TRY(hipMemsetAsync(B->belief, 0, std::max<uint64_t>(B->Stot, 1) * 8, B->stream));
    HIP_TRY(hipMemsetAsync(B->iters_dev, 0, std::max<uint64_t>(B->NV, 1) * 8, B->stream));
    HIP_TRY(hipStreamSynchronize(B->stream));
    return FLBP_OK;
}

int flbp_batch_autotune(flbp_batch *b, double damping, uint32_t trial_sweeps, flbp_batch_opts *best) {
    Batch *B = reinterpret_cast<Batch *>(b);
    if (!B || trial_sweeps == 0) return flbp::fail(FLBP_EINVAL, "flbp_batch_autotune: bad argument");
    if (!(damping >= 0.0 && damping < 1.0)) return flbp::fail(FLBP_EINVAL, "flbp_batch_autotune: bad damping");
    HIP_TRY(hipSetDevice(B->device));
    const flbp_batch_opts base = B->opts;
    const uint32_t tiles[] = {flbp::kStageSlots / 4, flbp::kStageSlots / 2, flbp::kStageSlots};
    double best_t = INFINITY;
    flbp_batch_opts best_o = base;
    for (uint32_t tf : tiles)
        for (uint32_t tv : tiles) {
            flbp_batch_opts o = base;
            o.tile_factors = tf;
            o.tile_vars = tv;
            TRY(flbp_batch_set_opts(b, &o));
            TRY(flbp_batch_reset_counters(b));
            TRY(flbp_batch_init(b, 0));
            flbp_run_opts ro{-INFINITY, trial_sweeps, INFINITY, damping, 0};
            TRY(flbp_batch_run(b, &ro, nullptr, nullptr, nullptr));   /* warm: captures the solve */
            TRY(flbp_batch_reset_counters(b));
            TRY(flbp_batch_init(b, 0));
            const auto t0 = std::chrono::steady_clock::now();
            flbp_run_result r;
            TRY(flbp_batch_run(b, &ro, nullptr, nullptr, &r));
            const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (t < best_t) {
                best_t = t;
                best_o = o;
            }
        }
    TRY(flbp_batch_set_opts(b, &best_o));
    TRY(flbp_batch_reset_counters(b));
    TRY(flbp_batch_init(b, 0));
    if (best) *best = best_o;
    return FLBP_OK;
}

uint64_t flbp_batch_device_bytes(const flbp_batch *b) {
    const Batch *B = reinterpret_cast<const Batch *>(b);
    return B ? B->bytes : 0;
}

uint64_t flbp_batch_n_variants(const flbp_batch *b) {
    const Batch *B = reinterpret_cast<const Batch *>(b);
    return B ? B->NV : 0;
}

}  // extern "C"
