/*
 * batch.hip -- batched PARALL solve of many graphs (independent graphs or
 * clamped variants of one graph) in one device state.
 *
 * Layout: every per-factor / per-variable / per-edge array of all graphs
 * concatenated (graph bases in BatchGraph, 64-bit; indices inside a graph
 * 32-bit).  A chunk of sweeps -- one clear kernel and, per sweep, one
 * persistent factor kernel and one persistent variable kernel over the
 * tile list -- is captured once into a hipGraph and replayed; per-graph
 * stop flags live in device memory, so one replay advances every unfinished
 * graph by up to `chunk` sweeps with no host round trip.  Tiles are the
 * plain solve's staged blocks cut per graph (their edge-slot bound is the
 * tile option), with per-tile work orders; messages use the scatter layout.
 * The tile bound and the grid size are chosen by flbp_batch_autotune.
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

/* one batch_permute launch over the factor tiles; returns on a launch error */
#define TRY_PERMUTE(TO_FM, SRC, DST)                                                                       \
    do {                                                                                                    \
        const uint32_t g_ = std::max<uint32_t>(1, std::min(opts.grid_blocks, n_ftiles));                    \
        hipLaunchKernelGGL(batch_permute<TO_FM>, dim3(g_), dim3(kBlock), 0, stream, graphs, ftiles, n_ftiles, \
                           fac_off, fac_vpos, SRC, DST);                                                    \
        const hipError_t e_ = hipGetLastError();                                                            \
        if (e_ != hipSuccess) return hip_fail(e_, "batch_permute");                                         \
    } while (0)

struct Batch {
    int device = -1;
    hipStream_t stream = nullptr;
    uint64_t G = 0, Ftot = 0, Vtot = 0, Etot = 0;
    uint64_t max_fac_deg = 0;
    std::vector<BatchGraph> h_graphs;
    BatchGraph *graphs = nullptr;
    uint32_t *fac_off = nullptr, *fac_meta = nullptr, *var_off = nullptr, *var_edge = nullptr;
    uint32_t *fac_vpos = nullptr;
    double *fac_p = nullptr, *fac_q = nullptr, *belief = nullptr;
    double2 *fac_mask = nullptr, *m_vf = nullptr, *m_fv = nullptr;
    double2 *m_fv_vm = nullptr;             /* f->v var-major; m_fv is the factor-major copy for damping */
    bool fm_stale = false;                  /* scatter layout: m_fv out of date after undamped sweeps */
    int layout = 0;                         /* resolved: kLayoutPull or kLayoutScatter */
    uint32_t *fac_work = nullptr;           /* per tile: its slots in work order (graph-local tile slots) */
    uint16_t *var_work = nullptr;           /* per tile: its variables by degree */
    std::vector<uint32_t> h_fac_off, h_fac_meta, h_var_off;   /* host copies, batch-concatenated */
    Tile *ftiles = nullptr, *vtiles = nullptr;           /* every graph's tiles */
    uint64_t cap_ftiles = 0, cap_vtiles = 0;
    uint32_t n_ftiles = 0, n_vtiles = 0;
    std::vector<Tile> h_ftiles, h_vtiles;                 /* host copies, grouped by graph */
    std::vector<uint32_t> h_ftile_off, h_vtile_off;       /* per graph: its tiles' range (G+1) */
    Tile *act_ftiles = nullptr, *act_vtiles = nullptr;    /* tiles of the running graphs */
    uint32_t *act_count = nullptr;                        /* device: [factor tiles, variable tiles] */
    std::vector<uint8_t> act_set;                         /* running set the active lists hold */
    uint32_t cut_f = 0, cut_v = 0;          /* slot bounds of the current tiles */
    unsigned long long *diff = nullptr;     /* [kMaxChunk][G][kBatchSlots][kCtlPad] partial maxima */
    unsigned long long *diff_out = nullptr; /* [kMaxChunk][G] compacted */
    uint8_t *start_active = nullptr;        /* device [G] */
    uint32_t *budget = nullptr;             /* device [G] */
    unsigned long long *h_diff = nullptr;   /* pinned */
    uint8_t *h_active = nullptr;            /* pinned */
    uint32_t *h_budget = nullptr;           /* pinned */
    std::vector<uint64_t> iters;
    std::vector<double> last_diff;
    std::vector<uint32_t> h_perm;           /* per graph at var_base: device variable -> variable */
    std::vector<void *> allocs;
    uint64_t bytes = 0;
    flbp_batch_opts opts{};
    /* captured chunk */
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int exec_chunk = 0;
    double exec_damp = -1, exec_tol = 0;
    uint32_t exec_fgrid = 0, exec_vgrid = 0;

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
        if (h_diff) (void)hipHostFree(h_diff);
        if (h_active) (void)hipHostFree(h_active);
        if (h_budget) (void)hipHostFree(h_budget);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int build_tiles();
    /* the sweep kernels' tile lists: the tiles of the graphs with run[g]
     * set, uploaded when that set changes */
    int set_active(const uint8_t *run) {
        if (act_set.size() == G && std::equal(act_set.begin(), act_set.end(), run)) return FLBP_OK;
        act_set.assign(run, run + G);
        std::vector<Tile> fa, va;
        for (uint64_t g = 0; g < G; g++) {
            if (!run[g]) continue;
            fa.insert(fa.end(), h_ftiles.begin() + h_ftile_off[g], h_ftiles.begin() + h_ftile_off[g + 1]);
            va.insert(va.end(), h_vtiles.begin() + h_vtile_off[g], h_vtiles.begin() + h_vtile_off[g + 1]);
        }
        const uint32_t cnt[2] = {(uint32_t)fa.size(), (uint32_t)va.size()};
        if (cnt[0] && hipMemcpyAsync(act_ftiles, fa.data(), fa.size() * sizeof(Tile), hipMemcpyHostToDevice, stream) !=
                          hipSuccess)
            return fail(FLBP_EHIP, "batch: uploading the active tiles failed");
        if (cnt[1] && hipMemcpyAsync(act_vtiles, va.data(), va.size() * sizeof(Tile), hipMemcpyHostToDevice, stream) !=
                          hipSuccess)
            return fail(FLBP_EHIP, "batch: uploading the active tiles failed");
        if (hipMemcpyAsync(act_count, cnt, sizeof cnt, hipMemcpyHostToDevice, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess)   /* the host vectors die here */
            return fail(FLBP_EHIP, "batch: uploading the active tiles failed");
        return FLBP_OK;
    }
    int resolve_layout() const {
        if (opts.layout == FLBP_LAYOUT_PULL) return kLayoutPull;
        if (opts.layout == FLBP_LAYOUT_SCATTER) return kLayoutScatter;
        return bytes <= kScatterMaxBytes ? kLayoutScatter : kLayoutPull;
    }
    /* undamped sweeps skip the factor-major f->v copy; a damped run first
     * rebuilds it from the var-major array */
    int prepare_damping(double damping) {
        if (layout != kLayoutScatter) return FLBP_OK;
        if (damping != 0.0 && fm_stale && n_ftiles) TRY_PERMUTE(true, m_fv_vm, m_fv);
        fm_stale = damping == 0.0 ? true : (fm_stale && !n_ftiles);
        return FLBP_OK;
    }
    /* move the stored messages to layout `to` (kLayoutPull / kLayoutScatter) */
    int set_layout(int to) {
        if (to == layout) return FLBP_OK;
        if (n_ftiles) {
            if (to == kLayoutScatter) {
                /* v->f var-major -> factor-major through the spare array,
                 * f->v factor-major -> var-major; the factor-major f->v
                 * copy stays valid */
                TRY_PERMUTE(true, m_vf, m_fv_vm);
                if (hipMemcpyAsync(m_vf, m_fv_vm, Etot * sizeof(double2), hipMemcpyDeviceToDevice, stream) != hipSuccess)
                    return fail(FLBP_EHIP, "batch: layout copy failed");
                TRY_PERMUTE(false, m_fv, m_fv_vm);
                fm_stale = false;
            } else {
                if (fm_stale) TRY_PERMUTE(true, m_fv_vm, m_fv);
                TRY_PERMUTE(false, m_vf, m_fv_vm);
                if (hipMemcpyAsync(m_vf, m_fv_vm, Etot * sizeof(double2), hipMemcpyDeviceToDevice, stream) != hipSuccess)
                    return fail(FLBP_EHIP, "batch: layout copy failed");
                fm_stale = false;
            }
            if (hipStreamSynchronize(stream) != hipSuccess) return fail(FLBP_EHIP, "batch: layout move failed");
        }
        layout = to;
        drop_exec();
        return FLBP_OK;
    }
    int record_chunk(int n, double tol, double damping);
    int enqueue_chunk(int n, double tol, double damping);
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

int Batch::build_tiles() {
    const uint32_t sf = std::min<uint32_t>(std::max<uint32_t>(opts.tile_factors, 64), kStageSlots);
    const uint32_t sv = std::min<uint32_t>(std::max<uint32_t>(opts.tile_vars, 64), kStageSlots);
    drop_exec();
    if (sf == cut_f && sv == cut_v && ftiles) return FLBP_OK;
    /* per graph: staged blocks and their work orders, graphs spread over
     * host threads */
    std::vector<std::vector<Tile>> ft(G), vt(G);
    std::vector<uint32_t> fw(Etot);
    std::vector<uint16_t> vw(Vtot);
    auto cut = [&](uint64_t g) {
        const BatchGraph &bg = h_graphs[g];
        const uint32_t *fo = h_fac_off.data() + bg.fac_off_base;
        const uint32_t *vo = h_var_off.data() + bg.var_off_base;
        std::vector<uint32_t> fb, vb, gfw;
        std::vector<uint16_t> gvw;
        cut_blocks(fo, bg.n_fac, kStageItems, sf, fb);
        cut_blocks(vo, bg.n_var, kStageItems, sv, vb);
        stage_work_orders(fo, h_fac_meta.data() + bg.fac_base, fb, vo, vb, gfw, gvw);
        std::copy(gfw.begin(), gfw.end(), fw.begin() + bg.edge_base);
        std::copy(gvw.begin(), gvw.end(), vw.begin() + bg.var_base);
        for (size_t b = 0; b + 1 < fb.size(); b++) ft[g].push_back({(uint32_t)g, fb[b], fb[b + 1] - fb[b]});
        for (size_t b = 0; b + 1 < vb.size(); b++) vt[g].push_back({(uint32_t)g, vb[b], vb[b + 1] - vb[b]});
    };
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++)
        pool.emplace_back([&, t] {
            for (uint64_t g = t; g < G; g += nth) cut(g);
        });
    for (auto &th : pool) th.join();
    std::vector<Tile> fall, vall;
    for (uint64_t g = 0; g < G; g++) {
        fall.insert(fall.end(), ft[g].begin(), ft[g].end());
        vall.insert(vall.end(), vt[g].begin(), vt[g].end());
    }
    if (fall.size() >= (1ull << 31) || vall.size() >= (1ull << 31)) return fail(FLBP_ERANGE, "batch: too many tiles");
    n_ftiles = (uint32_t)fall.size();
    n_vtiles = (uint32_t)vall.size();
    if (n_ftiles > cap_ftiles) {
        if (ftiles) release(ftiles);
        ftiles = nullptr;
        TRY(alloc(ftiles, n_ftiles));
        cap_ftiles = n_ftiles;
    }
    if (n_vtiles > cap_vtiles) {
        if (vtiles) release(vtiles);
        vtiles = nullptr;
        TRY(alloc(vtiles, n_vtiles));
        cap_vtiles = n_vtiles;
    }
    if (n_ftiles) HIP_TRY(hipMemcpy(ftiles, fall.data(), fall.size() * sizeof(Tile), hipMemcpyHostToDevice));
    if (n_vtiles) HIP_TRY(hipMemcpy(vtiles, vall.data(), vall.size() * sizeof(Tile), hipMemcpyHostToDevice));
    h_ftile_off.assign(G + 1, 0);
    h_vtile_off.assign(G + 1, 0);
    for (uint64_t g = 0; g < G; g++) {
        h_ftile_off[g + 1] = h_ftile_off[g] + (uint32_t)ft[g].size();
        h_vtile_off[g + 1] = h_vtile_off[g] + (uint32_t)vt[g].size();
    }
    h_ftiles.swap(fall);
    h_vtiles.swap(vall);
    if (act_ftiles) release(act_ftiles);
    if (act_vtiles) release(act_vtiles);
    act_ftiles = act_vtiles = nullptr;
    TRY(alloc(act_ftiles, n_ftiles));
    TRY(alloc(act_vtiles, n_vtiles));
    if (!act_count) TRY(alloc(act_count, 2));
    act_set.clear();
    std::vector<uint8_t> all(G, 1);
    TRY(set_active(all.data()));
    if (Etot) HIP_TRY(hipMemcpy(fac_work, fw.data(), Etot * 4, hipMemcpyHostToDevice));
    if (Vtot) HIP_TRY(hipMemcpy(var_work, vw.data(), Vtot * 2, hipMemcpyHostToDevice));
    cut_f = sf;
    cut_v = sv;
    return FLBP_OK;
}

int Batch::record_chunk(int n, double tol, double damping) {
    const double keep = 1.0 - damping;
    BatchCtl c{diff, start_active, budget, (uint32_t)G, tol};
    const uint64_t nslots = (uint64_t)n * G * kBatchSlots * kCtlPad;
    hipLaunchKernelGGL(batch_clear, dim3((unsigned)std::max<uint64_t>(1, std::min<uint64_t>(1024, (nslots + 255) / 256))),
                       dim3(256), 0, stream, diff, nslots);
    const uint32_t fg = std::max<uint32_t>(1, std::min(opts.grid_blocks, n_ftiles));
    const uint32_t vg = std::max<uint32_t>(1, std::min(opts.grid_blocks, n_vtiles));
    for (int k = 0; k < n; k++) {
        if (n_ftiles) {
            const bool narrow = max_fac_deg <= (uint64_t)kFacRegDegSmall;
#define FLBP_BATCH_FACTOR_PASS(DAMP, N, L)                                                                    \
    hipLaunchKernelGGL((batch_factor_pass<DAMP, N, L>), dim3(fg), dim3(kBlock), 0, stream, graphs, act_ftiles, \
                       act_count, fac_off, fac_meta, fac_p, fac_q, fac_mask, fac_vpos, fac_work, m_vf, m_fv,    \
                       m_fv_vm, damping, keep, c, k)
#define FLBP_BATCH_FACTOR_PASS_L(DAMP, N)                                                   \
    do {                                                                                    \
        if (layout == kLayoutScatter) FLBP_BATCH_FACTOR_PASS(DAMP, N, kLayoutScatter);      \
        else FLBP_BATCH_FACTOR_PASS(DAMP, N, kLayoutPull);                                  \
    } while (0)
            if (damping != 0.0) {
                if (narrow) FLBP_BATCH_FACTOR_PASS_L(true, kFacRegDegSmall);
                else FLBP_BATCH_FACTOR_PASS_L(true, kFacRegDeg);
            } else {
                if (narrow) FLBP_BATCH_FACTOR_PASS_L(false, kFacRegDegSmall);
                else FLBP_BATCH_FACTOR_PASS_L(false, kFacRegDeg);
            }
#undef FLBP_BATCH_FACTOR_PASS_L
#undef FLBP_BATCH_FACTOR_PASS
        }
        if (n_vtiles) {
            if (layout == kLayoutScatter)
                hipLaunchKernelGGL(batch_var_pass<kLayoutScatter>, dim3(vg), dim3(kBlock), 0, stream, graphs,
                                   act_vtiles, act_count + 1, var_off, var_edge, var_work, m_fv_vm, m_vf, belief, c, k);
            else
                hipLaunchKernelGGL(batch_var_pass<kLayoutPull>, dim3(vg), dim3(kBlock), 0, stream, graphs,
                                   act_vtiles, act_count + 1, var_off, var_edge, var_work, m_fv, m_vf, belief, c, k);
        }
    }
    const uint64_t waves = (uint64_t)n * G;
    hipLaunchKernelGGL(batch_compact, dim3((unsigned)std::max<uint64_t>(1, std::min<uint64_t>(1024, (waves * 64 + 255) / 256))),
                       dim3(256), 0, stream, c, diff_out, n);
    return FLBP_OK;
}

int Batch::enqueue_chunk(int n, double tol, double damping) {
    if (!opts.use_graph) {
        TRY(record_chunk(n, tol, damping));
        HIP_TRY(hipGetLastError());
        return FLBP_OK;
    }
    if (!exec || exec_chunk != n || exec_damp != damping || exec_tol != tol) {
        drop_exec();
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = record_chunk(n, tol, damping);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc) return rc;
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        graph = g;
        HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        exec_chunk = n;
        exec_damp = damping;
        exec_tol = tol;
    }
    HIP_TRY(hipGraphLaunch(exec, stream));
    return FLBP_OK;
}

}  // namespace flbp

using flbp::Batch;
using flbp::grid_for;
using flbp::hip_fail;

namespace {

__global__ void bfill2(double2 *p, uint64_t n, double c) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = make_double2(c, c);
}

int default_grid(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 1024;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_factor_pass<true, kFacRegDegSmall, kLayoutScatter>,
                                                     kBlock, 0) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 4;
    return prop.multiProcessorCount * per_cu;
}

}  // namespace

extern "C" {

void flbp_batch_default_opts(flbp_batch_opts *o) {
    if (!o) return;
    o->tile_factors = flbp::kStageSlots;
    o->tile_vars = flbp::kStageSlots;
    o->grid_blocks = 0;
    o->chunk = 32;
    o->use_graph = 1;
    o->layout = FLBP_LAYOUT_AUTO;
}

int flbp_batch_create(uint64_t n_graphs, const flbp_csr *const *csrs, int device, const flbp_batch_opts *opts,
                      flbp_batch **out) {
    if (!csrs || !out || n_graphs == 0) return flbp::fail(FLBP_EINVAL, "flbp_batch_create: bad arguments");
    if (n_graphs >= (1ull << 31)) return flbp::fail(FLBP_ERANGE, "flbp_batch_create: too many graphs");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return flbp::fail(FLBP_ENODEV, "flbp_batch_create: no HIP device");
    if (device < 0 || device >= ndev) return flbp::fail(FLBP_EINVAL, "flbp_batch_create: bad device");
    for (uint64_t g = 0; g < n_graphs; g++) {
        if (!csrs[g]) return flbp::fail(FLBP_EINVAL, "flbp_batch_create: null graph");
        const flbp_csr &c = *csrs[g];
        if (c.n_vars >= (1ull << 32) - 1 || c.n_factors >= (1ull << 32) - 1 || c.n_edges >= (1ull << 32) - 1)
            return flbp::fail(FLBP_ERANGE, "flbp_batch_create: a graph has V, F or E >= 2^32-1");
        int rc = flbp::validate_csr(c);
        if (rc) return rc;
    }
    HIP_TRY(hipSetDevice(device));
    auto *B = new (std::nothrow) Batch();
    if (!B) return flbp::fail(FLBP_ENOMEM, "flbp_batch_create: out of memory");
    std::unique_ptr<Batch> hold(B);
    B->device = device;
    B->G = n_graphs;
    if (opts) B->opts = *opts;
    else flbp_batch_default_opts(&B->opts);
    if (B->opts.layout > FLBP_LAYOUT_SCATTER) return flbp::fail(FLBP_EINVAL, "flbp_batch_create: bad layout");
    if (B->opts.tile_factors < 64 || B->opts.tile_factors > flbp::kStageSlots || B->opts.tile_vars < 64 ||
        B->opts.tile_vars > flbp::kStageSlots)
        return flbp::fail(FLBP_EINVAL, "flbp_batch_create: tile bounds must be 64.." +
                                           std::to_string(flbp::kStageSlots) + " slots");
    if (B->opts.chunk == 0 || B->opts.chunk > (uint32_t)flbp::kMaxChunk) B->opts.chunk = flbp::kMaxChunk;
    if (B->opts.grid_blocks == 0) B->opts.grid_blocks = (uint32_t)default_grid(device);
    B->h_graphs.resize(n_graphs);
    uint64_t F = 0, V = 0, E = 0;
    for (uint64_t g = 0; g < n_graphs; g++) {
        const flbp_csr &c = *csrs[g];
        B->h_graphs[g] = {F, V, E, F + g, V + g, (uint32_t)c.n_factors, (uint32_t)c.n_vars};
        for (uint64_t I = 0; I < c.n_factors; I++)
            B->max_fac_deg = std::max<uint64_t>(B->max_fac_deg, c.fac_off[I + 1] - c.fac_off[I]);
        F += c.n_factors;
        V += c.n_vars;
        E += c.n_edges;
    }
    B->Ftot = F;
    B->Vtot = V;
    B->Etot = E;
    HIP_TRY(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    TRY(B->alloc(B->graphs, n_graphs));
    TRY(B->alloc(B->fac_off, F + n_graphs));
    TRY(B->alloc(B->fac_meta, F));
    TRY(B->alloc(B->fac_p, F));
    TRY(B->alloc(B->fac_q, F));
    TRY(B->alloc(B->fac_mask, F));
    TRY(B->alloc(B->var_off, V + n_graphs));
    TRY(B->alloc(B->var_edge, E));
    TRY(B->alloc(B->fac_vpos, E));
    TRY(B->alloc(B->m_vf, E));
    TRY(B->alloc(B->m_fv, E));
    TRY(B->alloc(B->m_fv_vm, E));
    TRY(B->alloc(B->fac_work, E));
    TRY(B->alloc(B->var_work, V));
    TRY(B->alloc(B->belief, V));
    B->h_fac_off.resize(F + n_graphs);
    B->h_fac_meta.resize(F);
    B->h_var_off.resize(V + n_graphs);
    TRY(B->alloc(B->diff, (uint64_t)flbp::kMaxChunk * n_graphs * kBatchSlots * kCtlPad));
    TRY(B->alloc(B->diff_out, (uint64_t)flbp::kMaxChunk * n_graphs));
    TRY(B->alloc(B->start_active, n_graphs));
    TRY(B->alloc(B->budget, n_graphs));
    HIP_TRY(hipHostMalloc((void **)&B->h_diff, sizeof(unsigned long long) * flbp::kMaxChunk * n_graphs, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&B->h_active, n_graphs, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&B->h_budget, sizeof(uint32_t) * n_graphs, hipHostMallocDefault));
    HIP_TRY(hipMemcpy(B->graphs, B->h_graphs.data(), n_graphs * sizeof(BatchGraph), hipMemcpyHostToDevice));
    /* upload graph by graph through one staging vector */
    std::vector<uint32_t> u32;
    std::vector<double2> mk;
    for (uint64_t g = 0; g < n_graphs; g++) {
        const flbp_csr &c = *csrs[g];
        const BatchGraph &bg = B->h_graphs[g];
        u32.resize(std::max(std::max(c.n_factors, c.n_vars), c.n_edges) + 1);
        for (uint64_t I = 0; I <= c.n_factors; I++) u32[I] = (uint32_t)c.fac_off[I];
        std::copy(u32.begin(), u32.begin() + c.n_factors + 1, B->h_fac_off.begin() + bg.fac_off_base);
        HIP_TRY(hipMemcpy(B->fac_off + bg.fac_off_base, u32.data(), (c.n_factors + 1) * 4, hipMemcpyHostToDevice));
        for (uint64_t I = 0; I < c.n_factors; I++)
            u32[I] = flbp::dev::pack_meta((uint32_t)c.fac_type[I], c.head_clamped[I] != 0,
                                          (uint32_t)(c.head_slot[I] - c.fac_off[I]));
        std::copy(u32.begin(), u32.begin() + c.n_factors, B->h_fac_meta.begin() + bg.fac_base);
        if (c.n_factors) {
            HIP_TRY(hipMemcpy(B->fac_meta + bg.fac_base, u32.data(), c.n_factors * 4, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(B->fac_p + bg.fac_base, c.p, c.n_factors * 8, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(B->fac_q + bg.fac_base, c.q, c.n_factors * 8, hipMemcpyHostToDevice));
            mk.resize(c.n_factors);
            for (uint64_t I = 0; I < c.n_factors; I++) mk[I] = make_double2(c.mask0[I], c.mask1[I]);
            HIP_TRY(hipMemcpy(B->fac_mask + bg.fac_base, mk.data(), c.n_factors * 16, hipMemcpyHostToDevice));
        }
        std::vector<uint32_t> perm, dev_off, dev_edge, vpos;
        flbp::var_device_order(c, flbp::kOrderLabel, perm, dev_off, dev_edge, vpos);
        B->h_perm.insert(B->h_perm.end(), perm.begin(), perm.end());
        std::copy(dev_off.begin(), dev_off.end(), B->h_var_off.begin() + bg.var_off_base);
        HIP_TRY(hipMemcpy(B->var_off + bg.var_off_base, dev_off.data(), (c.n_vars + 1) * 4, hipMemcpyHostToDevice));
        if (c.n_edges) {
            HIP_TRY(hipMemcpy(B->var_edge + bg.edge_base, dev_edge.data(), c.n_edges * 4, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(B->fac_vpos + bg.edge_base, vpos.data(), c.n_edges * 4, hipMemcpyHostToDevice));
        }
    }
    HIP_TRY(hipMemset(B->belief, 0, std::max<uint64_t>(V, 1) * 8));
    HIP_TRY(hipMemset(B->m_vf, 0, std::max<uint64_t>(E, 1) * 16));
    HIP_TRY(hipMemset(B->m_fv, 0, std::max<uint64_t>(E, 1) * 16));
    HIP_TRY(hipMemset(B->m_fv_vm, 0, std::max<uint64_t>(E, 1) * 16));
    B->iters.assign(n_graphs, 0);
    B->last_diff.assign(n_graphs, INFINITY);
    TRY(B->build_tiles());
    B->layout = B->resolve_layout();
    *out = reinterpret_cast<flbp_batch *>(hold.release());
    return FLBP_OK;
}

void flbp_batch_destroy(flbp_batch *b) { delete reinterpret_cast<Batch *>(b); }

int flbp_batch_init(flbp_batch *b, int logdomain) {
    Batch *B = reinterpret_cast<Batch *>(b);
    if (!B) return flbp::fail(FLBP_EINVAL, "flbp_batch_init: null batch");
    HIP_TRY(hipSetDevice(B->device));
    const double c = logdomain ? 0.0 : 1.0;
    if (B->Etot) {
        const unsigned grid = (unsigned)std::min<uint64_t>(grid_for(B->Etot, 256), 65536);
        hipLaunchKernelGGL(bfill2, dim3(grid), dim3(256), 0, B->stream, B->m_vf, B->Etot, c);
        hipLaunchKernelGGL(bfill2, dim3(grid), dim3(256), 0, B->stream, B->m_fv, B->Etot, c);
        hipLaunchKernelGGL(bfill2, dim3(grid), dim3(256), 0, B->stream, B->m_fv_vm, B->Etot, c);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(B->stream));
    B->fm_stale = false;
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
    const uint64_t G = B->G;
    std::vector<uint8_t> active(G);
    std::vector<double> md(G, INFINITY);
    uint64_t n_active = 0, sweeps = 0;
    for (uint64_t g = 0; g < G; g++) {
        active[g] = B->iters[g] < o->maxiter;
        n_active += active[g];
    }
    const int chunk = (int)std::min<uint32_t>(o->chunk ? o->chunk : B->opts.chunk, flbp::kMaxChunk);
    TRY(B->prepare_damping(o->damping));
    while (n_active && elapsed() < o->maxtime) {
        for (uint64_t g = 0; g < G; g++) {
            B->h_active[g] = active[g];
            const uint64_t left = o->maxiter > B->iters[g] ? o->maxiter - B->iters[g] : 0;
            B->h_budget[g] = (uint32_t)std::min<uint64_t>(left, 0xffffffffu);
        }
        TRY(B->set_active(B->h_active));
        HIP_TRY(hipMemcpyAsync(B->start_active, B->h_active, G, hipMemcpyHostToDevice, B->stream));
        HIP_TRY(hipMemcpyAsync(B->budget, B->h_budget, G * 4, hipMemcpyHostToDevice, B->stream));
        TRY(B->enqueue_chunk(chunk, o->tol, o->damping));
        HIP_TRY(hipMemcpyAsync(B->h_diff, B->diff_out, sizeof(unsigned long long) * chunk * G, hipMemcpyDeviceToHost,
                               B->stream));
        HIP_TRY(hipStreamSynchronize(B->stream));
        n_active = 0;
        uint64_t most = 0;
        for (uint64_t g = 0; g < G; g++) {
            if (!active[g]) continue;
            const uint64_t cap = std::min<uint64_t>((uint64_t)chunk, B->h_budget[g]);
            uint64_t ran = cap;
            bool stop = false;
            for (uint64_t k = 0; k < cap; k++) {
                double dk;
                std::memcpy(&dk, &B->h_diff[k * G + g], sizeof dk);
                md[g] = dk;
                if (!(dk > o->tol)) {
                    ran = k + 1;
                    stop = true;
                    break;
                }
            }
            B->iters[g] += ran;
            most = std::max(most, ran);
            if (stop || B->iters[g] >= o->maxiter) active[g] = 0;
            n_active += active[g];
        }
        sweeps += most;
    }
    for (uint64_t g = 0; g < G; g++) {
        B->last_diff[g] = md[g];
        if (iters_out) iters_out[g] = B->iters[g];
        if (maxdiff_out) maxdiff_out[g] = md[g];
    }
    if (total) {
        total->iters = sweeps;
        total->sweeps = sweeps;
        total->maxdiff = 0;
        for (uint64_t g = 0; g < G; g++) total->maxdiff = std::max(total->maxdiff, md[g]);
        total->seconds = elapsed();
    }
    return FLBP_OK;
}

int flbp_batch_beliefs(flbp_batch *b, uint64_t g, double *out, uint64_t n) {
    Batch *B = reinterpret_cast<Batch *>(b);
    if (!B || (n && !out) || g >= (B ? B->G : 0)) return flbp::fail(FLBP_EINVAL, "flbp_batch_beliefs: bad argument");
    if (n > B->h_graphs[g].n_var) return flbp::fail(FLBP_EINVAL, "flbp_batch_beliefs: n > n_vars of the graph");
    HIP_TRY(hipSetDevice(B->device));
    const BatchGraph &bg = B->h_graphs[g];
    std::vector<double> dev(bg.n_var);
    if (bg.n_var)
        HIP_TRY(hipMemcpyAsync(dev.data(), B->belief + bg.var_base, bg.n_var * 8, hipMemcpyDeviceToHost, B->stream));
    HIP_TRY(hipStreamSynchronize(B->stream));
    for (uint64_t dv = 0; dv < bg.n_var; dv++) {
        const uint32_t v = B->h_perm[bg.var_base + dv];
        if (v < n) out[v] = dev[dv];
    }
    return FLBP_OK;
}

int flbp_batch_set_opts(flbp_batch *b, const flbp_batch_opts *o) {
    Batch *B = reinterpret_cast<Batch *>(b);
    if (!B || !o) return flbp::fail(FLBP_EINVAL, "flbp_batch_set_opts: null argument");
    HIP_TRY(hipSetDevice(B->device));
    if (o->layout > FLBP_LAYOUT_SCATTER) return flbp::fail(FLBP_EINVAL, "flbp_batch_set_opts: bad layout");
    if (o->tile_factors < 64 || o->tile_factors > flbp::kStageSlots || o->tile_vars < 64 ||
        o->tile_vars > flbp::kStageSlots)
        return flbp::fail(FLBP_EINVAL, "flbp_batch_set_opts: tile bounds must be 64.." +
                                           std::to_string(flbp::kStageSlots) + " slots");
    B->opts = *o;
    if (B->opts.chunk == 0 || B->opts.chunk > (uint32_t)flbp::kMaxChunk) B->opts.chunk = flbp::kMaxChunk;
    if (B->opts.grid_blocks == 0) B->opts.grid_blocks = (uint32_t)default_grid(B->device);
    TRY(B->build_tiles());
    return B->set_layout(B->resolve_layout());
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
    HIP_TRY(hipSetDevice(B->device));
    HIP_TRY(hipMemsetAsync(B->belief, 0, std::max<uint64_t>(B->Vtot, 1) * 8, B->stream));
    HIP_TRY(hipStreamSynchronize(B->stream));
    return FLBP_OK;
}

int flbp_batch_autotune(flbp_batch *b, double damping, uint32_t trial_sweeps, flbp_batch_opts *best) {
    Batch *B = reinterpret_cast<Batch *>(b);
    if (!B || trial_sweeps == 0) return flbp::fail(FLBP_EINVAL, "flbp_batch_autotune: bad argument");
    HIP_TRY(hipSetDevice(B->device));
    const flbp_batch_opts base = B->opts;
    const uint32_t g0 = (uint32_t)default_grid(B->device);
    const uint32_t tiles[] = {flbp::kStageSlots / 2, flbp::kStageSlots};
    const uint32_t grids[] = {g0, 2 * g0};
    const uint32_t layouts[] = {FLBP_LAYOUT_PULL, FLBP_LAYOUT_SCATTER};
    const int n = (int)std::min<uint32_t>(trial_sweeps, flbp::kMaxChunk);
    double best_t = INFINITY;
    flbp_batch_opts best_o = base;
    for (uint32_t lay : layouts)
    for (uint32_t tf : tiles) {
        for (uint32_t gb : grids) {
            flbp_batch_opts o = base;
            o.tile_factors = tf;
            o.tile_vars = tf;
            o.grid_blocks = gb;
            o.layout = lay;
            TRY(flbp_batch_set_opts(b, &o));
            TRY(flbp_batch_init(b, 0));
            TRY(B->prepare_damping(damping));
            for (uint64_t g = 0; g < B->G; g++) {
                B->h_active[g] = 1;
                B->h_budget[g] = (uint32_t)n;
            }
            TRY(B->set_active(B->h_active));
            HIP_TRY(hipMemcpyAsync(B->start_active, B->h_active, B->G, hipMemcpyHostToDevice, B->stream));
            HIP_TRY(hipMemcpyAsync(B->budget, B->h_budget, B->G * 4, hipMemcpyHostToDevice, B->stream));
            TRY(B->enqueue_chunk(n, -INFINITY, damping));   /* warm: captures the graph */
            HIP_TRY(hipStreamSynchronize(B->stream));
            const auto t0 = std::chrono::steady_clock::now();
            TRY(B->enqueue_chunk(n, -INFINITY, damping));
            HIP_TRY(hipStreamSynchronize(B->stream));
            const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (t < best_t) {
                best_t = t;
                best_o = o;
            }
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

}  // extern "C"
