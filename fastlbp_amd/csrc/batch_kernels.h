/*
 * batch_kernels.h -- PARALL sweeps over many graphs at once.
 *
 * A batch holds groups.  A group is one graph structure (a CSR in the
 * reference's orders) solved for K >= 1 variants: independent graphs are
 * groups of one variant; clamped feedback variants of one graph whose
 * clamps keep its structure (head clamps of variables without body uses,
 * causal_factor.cpp:75-83) share one group.  Structure arrays (offsets,
 * factor meta, p, q, slot maps, work orders) are stored once per group;
 * per-variant state is interleaved: f->v message of slot s in variant k at
 * msg_base + s*K + k, product pair / zero-count code / belief of device
 * variable v in variant k at vst_base + v*K + k.  A tile of a K-variant
 * group covers at most kStage/K slots for all K variants, so one index load
 * serves K lanes and one gathered line carries K messages (K = 8: a whole
 * 128-byte line of f->v messages per gathered slot).
 *
 * Messages use the state layout of state_kernels.h: the factor pass derives
 * each v->f input from the variable's product pair and zero-count code and
 * the slot's old f->v message, the variable pass gathers f->v messages and
 * writes per-variable products.  Arithmetic and operation order are the
 * plain solve's, so every variant's results are bit-identical to a plain
 * solve of its graph.
 *
 * Stop test, per variant and entirely on the device: vrun[v] says whether
 * variant v runs the next sweep.  After each sweep, batch_stop reads each
 * running variant's max belief change, advances its counter, clears the
 * change slots and sets vrun for the next sweep (max change > tol and
 * iteration budget left), then compacts the list of groups with a running
 * variant and the prefix sums of their tile counts.  The sweep kernels walk
 * only those groups' tiles, and lanes of a variant that stopped store
 * nothing, so a variant stops after exactly the sweeps its plain solve
 * runs.  A whole solve -- every sweep up to the iteration cap -- is one
 * captured hipGraph; the sweeps after the last variant stopped find no
 * active group and exit at once.
 */
#ifndef FLBP_BATCH_KERNELS_H
#define FLBP_BATCH_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "plan_internal.h"
#include "staged_kernels.h"
#include "state_kernels.h"
#include "sweep_kernels.h"

namespace {

using namespace flbp::dev;

struct GroupDev {
    uint64_t fac_base, var_base, edge_base;   /* structure element offsets */
    uint64_t fac_off_base, var_off_base;      /* into fac_off (F+1 entries per group) / var_off (V+1) */
    uint64_t msg_base;                        /* K-interleaved f->v messages */
    uint64_t vst_base;                        /* K-interleaved products, codes, beliefs */
    uint32_t n_fac, n_var;
    uint32_t K, lk;                           /* variants, log2 K */
    uint32_t vbase;                           /* batch index of variant 0 */
    uint32_t ftile_first, ftile_count, vtile_first, vtile_count;
    uint32_t pad;
};

struct Tile {
    uint32_t group;
    uint32_t first;    /* first local factor (factor tiles) or variable (var tiles) */
    uint32_t count;
    uint32_t aux;      /* wide factor tiles: offset of the factor's scratch range */
};

/* per variant, kBatchSlots partial maxima of the sweep's belief change one
 * 64-byte line apart (tile i of a group folds into slot i mod kBatchSlots) */
constexpr int kBatchSlots = 8;

/* Tiles of groups with K >= kDirectK variants hold up to kBatchStage
 * message slots (all variants), tiles of smaller groups kStage.  Factor
 * tiles store their messages in work order, so every lane stores its
 * outputs straight from the compute phase to contiguous positions.  Variable
 * tiles of K >= kDirectK store their records straight from the compute
 * phase too (lanes next to each other handle one variable of neighbouring
 * variants: each run of K records is one contiguous store); those of
 * smaller groups stage them in the upper half of the LDS buffer, as the
 * plain kernels do. */
constexpr uint32_t kDirectK = 4;
constexpr int kBatchStage = 2 * kStage;
constexpr int kBatchPerLane = kBatchStage / kBlock;

__host__ __device__ inline uint32_t batch_tile_cap(uint32_t K) { return K >= kDirectK ? kBatchStage : kStage; }

struct BatchFactorSmem {
    double2 buf[kBatchStage];
    uint32_t start[kBlock + 1];   /* factor offsets relative to the tile's first slot */
    uint32_t meta[kBlock];
    double p[kBlock], q[kBlock];
};

struct BatchVarSmem {
    double2 buf[kBatchStage];
    uint32_t start[kBlock + 1];
};

struct BatchArgs {
    const GroupDev *__restrict__ groups;
    const Tile *__restrict__ act_ftiles;      /* the running groups' factor tiles, in group order */
    const Tile *__restrict__ act_vtiles;      /* ... and variable tiles */
    const uint32_t *__restrict__ act_count;   /* [2]: how many of each */
    const uint32_t *__restrict__ fac_off, *__restrict__ fac_meta, *__restrict__ fac_var, *__restrict__ fac_work;
    const uint32_t *__restrict__ fac_crow;    /* clamped factor: first of its K masks */
    const double *__restrict__ fac_p, *__restrict__ fac_q;
    const double2 *__restrict__ masks;        /* per clamped factor of a group, K head masks */
    const uint32_t *__restrict__ var_off, *__restrict__ var_edge;
    const uint16_t *__restrict__ var_work;
    double2 *m_fv;
    double *__restrict__ vprod;               /* product records (pack_prod) */
    uint8_t *__restrict__ vzc;
    double *__restrict__ belief;
    double2 *__restrict__ scratch;            /* wide factors' v->f inputs */
    const uint8_t *__restrict__ vrun;         /* [n variants] */
    const uint8_t *__restrict__ vfresh;       /* [n variants] 1: no sweep since init */
    unsigned long long *__restrict__ diff;    /* [n variants][kBatchSlots][kCtlPad] */
    double c;                                 /* init value: the v->f input of a fresh variant */
};

/* Messages of one factor from a global array with stride K. */
struct StridedIn {
    const double2 *__restrict__ m;
    uint32_t lk;
    __device__ __forceinline__ double2 get(int j) const { return m[(uint64_t)j << lk]; }
    __device__ __forceinline__ double2 pick(int j) const { return m[(uint64_t)j << lk]; }
};

/* One wide factor tile (n slots, n * K > kStage): each variant in turn,
 * from global memory. */
template <bool kDamp>
__device__ __forceinline__ void batch_wide_factor(const BatchArgs &A, const GroupDev &gr, const Tile &tile,
                                                           double damp, double keep) {
    const uint32_t *fac_off = A.fac_off + gr.fac_off_base;
    const uint64_t fb = gr.fac_base, eb = gr.edge_base;
    const uint32_t f0 = tile.first;
    const uint32_t base = fac_off[f0];
    const uint32_t n = fac_off[f0 + 1] - base;
    const uint32_t lk = gr.lk, K = gr.K;
    const int tid = threadIdx.x;
    double2 *const fv = A.m_fv + gr.msg_base + ((uint64_t)base << lk);
    /* one wide factor, each variant in turn, from global memory */
    const uint32_t meta = A.fac_meta[fb + f0];
    const uint32_t type = meta & 3u;
    const bool clamped = (meta & 4u) != 0;
    const int h = (int)(meta >> 8);
    const double p = A.fac_p[fb + f0], q = type ? A.fac_q[fb + f0] : 0.0;
    double2 *const scr = A.scratch + tile.aux;
    for (uint32_t k = 0; k < K; k++) {
        if (!A.vrun[gr.vbase + k]) continue;   /* uniform over the block */
        double mask0 = 1.0, mask1 = 1.0;
        if (clamped) {
            const double2 mk = A.masks[(uint64_t)A.fac_crow[fb + f0] + k];
            mask0 = mk.x;
            mask1 = mk.y;
        }
        const bool fresh = A.vfresh[gr.vbase + k] != 0;
        if (!fresh) {
            for (uint32_t t = tid; t < n; t += kBlock) {
                const uint64_t v = gr.vst_base + ((uint64_t)A.fac_var[eb + base + t] << lk) + k;
                scr[t] = vf_of(A.vprod[v], A.vzc[v], fv[((uint64_t)t << lk) + k]);
            }
        }
        __syncthreads();
        const LdsIn in{scr};
        const ConstIn cin{A.c};
        const double2 mh = fresh ? cin.pick(h) : in.pick(h), mo = fresh ? cin.pick(h == 0 ? 1 : 0) : in.pick(h == 0 ? 1 : 0);
        for (uint32_t t = tid; t < n; t += kBlock) {
            double2 r = fresh ? fv_msg<0>(cin, type, clamped, (int)n, (int)t, h, mh, mo, p, q, mask0, mask1)
                              : fv_msg<0>(in, type, clamped, (int)n, (int)t, h, mh, mo, p, q, mask0, mask1);
            double2 *dst = fv + ((uint64_t)t << lk) + k;
            if (kDamp) {
                const double2 old = *dst;
                r.x = keep * r.x + damp * old.x;
                r.y = keep * r.y + damp * old.y;
            }
            *dst = r;
        }
        __syncthreads();   /* the scratch range is reused by the next variant */
    }
}

/* One factor tile of group gr: factors [f0, f1), all K variants. */
template <bool kDamp, int N>
__device__ __forceinline__ void batch_factor_tile(const BatchArgs &A, const GroupDev &gr, const Tile &tile,
                                                  double damp, double keep, BatchFactorSmem &sm) {
    static_assert(N <= 8, "SkipIn packs the operands of factors of degree 8 or less");
    const uint32_t *fac_off = A.fac_off + gr.fac_off_base;
    const uint64_t fb = gr.fac_base, eb = gr.edge_base;
    const uint32_t f0 = tile.first, f1 = tile.first + tile.count;
    const uint32_t base = fac_off[f0];
    const uint32_t n = fac_off[f1] - base;
    const uint32_t lk = gr.lk, K = gr.K, km = K - 1;
    const uint32_t nk = n << lk;
    const int tid = threadIdx.x;
    double2 *const fv = A.m_fv + gr.msg_base + ((uint64_t)base << lk);
    if (nk > batch_tile_cap(K)) {
        batch_wide_factor<kDamp>(A, gr, tile, damp, keep);
        return;
    }
    /* the tile's messages are stored in its work order: position p of the
     * tile holds work entry p (fac_var and fac_work are indexed by
     * position), variant k of it at p * K + k.  Lane order is storage
     * order, so every load and store of the tile is contiguous. */
    double2 *const s_in = sm.buf;   /* v->f inputs by slot: slot s, variant k at s * K + k */
    /* phase 1 */
    const uint32_t nf = f1 - f0;
    uint32_t w[kBatchPerLane], ix[kBatchPerLane];
#pragma unroll
    for (int j = 0; j < kBatchPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        w[j] = i < nk ? A.fac_work[eb + base + (i >> lk)] : 0u;
        ix[j] = i < nk ? A.fac_var[eb + base + (i >> lk)] : 0u;
    }
    if ((uint32_t)tid < nf) {
        sm.start[tid] = fac_off[f0 + tid] - base;
        sm.meta[tid] = A.fac_meta[fb + f0 + tid];
        sm.p[tid] = A.fac_p[fb + f0 + tid];
        sm.q[tid] = A.fac_q[fb + f0 + tid];
    }
    if (tid == 0) sm.start[nf] = n;
    /* whether the lane's variant starts from the init constant (all of a
     * lane's positions have variant tid & (K - 1)), read once */
    const bool fresh = A.vfresh[gr.vbase + (tid & km)] != 0;
#pragma unroll
    for (int j = 0; j < kBatchPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        if (i < nk) {
            const uint32_t k = i & km;
            const uint64_t v = gr.vst_base + ((uint64_t)ix[j] << lk) + k;
            const double2 old = fv[i];
            const double pr = A.vprod[v];
            const uint32_t zc = A.vzc[v];
            s_in[((w[j] & 1023u) << lk) + k] = fresh ? make_double2(A.c, A.c) : vf_of(pr, zc, old);
        }
    }
    /* the run flag of the lane's variant (every position of a lane,
     * tid + j * kBlock, has variant tid & (K - 1)), read once here rather
     * than before each store of phase 2 */
    const bool run_k = A.vrun[gr.vbase + (tid & km)] != 0;
    __syncthreads();
    /* phase 2: lane position i computes work entry i >> lk of variant
     * i & (K - 1) and stores it at its own position (unrolled, so each
     * entry stays in its register: the 768-graph batch 248.6 -> 245.5 ms) */
#pragma unroll
    for (int j = 0; j < kBatchPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        if (i >= nk) break;
        uint32_t wj = 0;
#pragma unroll
        for (int x = 0; x < kBatchPerLane; x++) wj |= (j == x) ? w[x] : 0u;
        const uint32_t k = i & km;
        const uint32_t sl = wj & 1023u, lf = wj >> 10;
        const uint32_t s = sm.start[lf];
        const int d = (int)(sm.start[lf + 1] - s);
        const int t = (int)(sl - s);
        const uint32_t meta = sm.meta[lf];
        const uint32_t type = meta & 3u;
        const bool clamped = (meta & 4u) != 0;
        const int h = (int)(meta >> 8);
        const double p = sm.p[lf];
        const double q = type ? sm.q[lf] : 0.0;
        double mask0 = 1.0, mask1 = 1.0;
        if (clamped) {
            const double2 mk = A.masks[(uint64_t)A.fac_crow[fb + f0 + lf] + k];
            mask0 = mk.x;
            mask1 = mk.y;
        }
        const double2 *fin = s_in + (s << lk) + k;
        double2 r;
        if (d <= N) {
            const bool head_target = (t == h);
            const int cnt = type == 0 ? 0 : d - (head_target ? 1 : 2);
            const SkipIn ops =
                SkipIn::make(fin, head_target ? t : (t < h ? t : h), head_target ? N + 1 : (t < h ? h : t), lk);
            r = fv_list<N - 1>(ops, cnt, type, clamped, d, head_target, fin[(uint32_t)h << lk], p, q, mask0, mask1);
        } else {
            const StridedIn in{fin, lk};
            const double2 mh = in.pick(h), mo = in.pick(h == 0 ? 1 : 0);
            r = fv_msg<0>(in, type, clamped, d, t, h, mh, mo, p, q, mask0, mask1);
        }
        /* the old message of this position was read by this lane in phase
         * 1 and is read back here; no other lane writes it */
        if (run_k) {
            if (kDamp) {
                const double2 old2 = fv[i];
                r.x = keep * r.x + damp * old2.x;
                r.y = keep * r.y + damp * old2.y;
            }
            fv[i] = r;
        }
    }
}

/* max over lanes with the same residue mod K (K a power of two <= 64):
 * afterwards lane l holds the max over lanes l' of the wave with
 * l' = l mod K */
__device__ __forceinline__ unsigned long long wave_max_mod(unsigned long long v, uint32_t K) {
#pragma unroll
    for (uint32_t off = 32; off > 0; off >>= 1) {
        if (off < K) break;
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ void fold_diff(const BatchArgs &A, uint32_t variant, uint32_t ti, unsigned long long m) {
    if (!m) return;
    unsigned long long *slot = A.diff + ((uint64_t)variant * kBatchSlots + (ti % kBatchSlots)) * kCtlPad;
    if (m > __atomic_load_n(slot, __ATOMIC_RELAXED)) atomicMax(slot, m);
}

/* One variable tile of group gr: variables [v0, v1), all K variants. */
__device__ __forceinline__ void batch_var_tile(const BatchArgs &A, const GroupDev &gr, const Tile &tile,
                                               uint32_t ti, BatchVarSmem &sm) {
    const uint32_t *var_off = A.var_off + gr.var_off_base;
    const uint64_t eb = gr.edge_base;
    const uint32_t v0 = tile.first, v1 = tile.first + tile.count;
    const uint32_t base = var_off[v0];
    const uint32_t n = var_off[v1] - base;
    const uint32_t lk = gr.lk, K = gr.K, km = K - 1;
    const uint32_t nk = n << lk;
    const int tid = threadIdx.x;
    const double2 *fv = A.m_fv + gr.msg_base;
    const uint64_t vst = gr.vst_base + ((uint64_t)v0 << lk);
    const bool direct = K >= kDirectK;
    const uint32_t cap = batch_tile_cap(K);
    double2 *const s_m = sm.buf;
    if (nk > cap) {
        /* one wide variable: chunks of cap / K edges of all K variants
         * staged by the block, lane k multiplying variant k's messages in
         * edge order */
        const uint32_t C = cap >> lk;
        double prod_0 = 1.0, prod_1 = 1.0;
        uint32_t z0 = 0, z1 = 0;
        for (uint32_t c = 0; c < n; c += C) {
            const uint32_t m = n - c < C ? n - c : C;
            for (uint32_t i = tid; i < (m << lk); i += kBlock)
                s_m[i] = fv[((uint64_t)A.var_edge[eb + base + c + (i >> lk)] << lk) + (i & km)];
            __syncthreads();
            if ((uint32_t)tid < K) {
                for (uint32_t e = 0; e < m; e++) {
                    const double2 x = s_m[(e << lk) + tid];
                    if (x.x == 0) z0++; else prod_0 *= x.x;
                    if (x.y == 0) z1++; else prod_1 *= x.y;
                    scale_pair(prod_0, prod_1);
                }
            }
            __syncthreads();
        }
        if ((uint32_t)tid < K && A.vrun[gr.vbase + tid]) {
            const double b0 = prod_0 * (double)(z0 == 0);
            const double b1 = prod_1 * (double)(z1 == 0);
            const double b = b1 / (b0 + b1);
            const double diff = fabs(A.belief[vst + tid] - b);
            A.belief[vst + tid] = b;
            double rest;
            A.vzc[vst + tid] = pack_prod(prod_0, prod_1, z0, z1, rest);
            A.vprod[vst + tid] = rest;
            fold_diff(A, gr.vbase + tid, ti, (unsigned long long)__double_as_longlong(diff));
        }
        return;
    }
    /* staged results of a tile of K < kDirectK (nk <= kStage), in the upper
     * half of the buffer: 2 * kBlock doubles and kBlock codes fit its 8 KiB */
    double *const s_prod = reinterpret_cast<double *>(sm.buf + kStage);
    double *const s_bel = s_prod + kBlock;
    uint32_t *const s_z = reinterpret_cast<uint32_t *>(s_bel + kBlock);
    /* phase 1 */
    const uint32_t nv = v1 - v0, nvk = nv << lk;   /* nvk <= 2 * kBlock (direct), kBlock (staged) */
    uint32_t ve[kBatchPerLane];
#pragma unroll
    for (int j = 0; j < kBatchPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        ve[j] = i < nk ? A.var_edge[eb + base + (i >> lk)] : 0u;
    }
    uint32_t lv[2] = {0u, 0u};
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const uint32_t t = tid + r * kBlock;
        if (t < nvk) lv[r] = A.var_work[gr.var_base + v0 + (t >> lk)];
    }
    if (!direct && (uint32_t)tid < nvk) s_bel[tid] = A.belief[vst + tid];
    if ((uint32_t)tid < nv) sm.start[tid] = var_off[v0 + tid] - base;
    if (tid == 0) sm.start[nv] = n;
#pragma unroll
    for (int j = 0; j < kBatchPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        if (i < nk) s_m[i] = fv[((uint64_t)ve[j] << lk) + (i & km)];
    }
    /* the run flag of the lane's variant (both of a lane's entries, and on
     * lanes 0..K-1 the fold below, belong to variant tid & (K - 1)), read
     * once here */
    const bool run_k = A.vrun[gr.vbase + (tid & km)] != 0;
    __syncthreads();
    /* phase 2: lane t (and t + kBlock in a direct tile) is work entry
     * t >> lk (a variable) of variant t & (K - 1); both entries of a lane
     * belong to the same variant, since K divides kBlock */
    unsigned long long diff_bits = 0;
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const uint32_t t = tid + r * kBlock;
        if (t >= nvk) break;
        const uint32_t k = t & km;
        const uint32_t s = sm.start[lv[r]];
        const int d = (int)(sm.start[lv[r] + 1] - s);
        double prod_0 = 1.0, prod_1 = 1.0;
        uint32_t z0 = 0, z1 = 0;
        for (int j = 0; j < d; j++) {
            const double2 x = s_m[((s + j) << lk) + k];
            if (x.x == 0) z0++; else prod_0 *= x.x;
            if (x.y == 0) z1++; else prod_1 *= x.y;
            scale_pair(prod_0, prod_1);
        }
        const uint32_t o = (lv[r] << lk) + k;
        const double b0 = prod_0 * (double)(z0 == 0);
        const double b1 = prod_1 * (double)(z1 == 0);
        const double b = b1 / (b0 + b1);
        if (direct) {
            /* runs of K neighbouring lanes store K contiguous records */
            if (run_k) {
                const double diff = fabs(A.belief[vst + o] - b);
                A.belief[vst + o] = b;
                double rest;
                A.vzc[vst + o] = pack_prod(prod_0, prod_1, z0, z1, rest);
                A.vprod[vst + o] = rest;
                const unsigned long long bits = (unsigned long long)__double_as_longlong(diff);
                diff_bits = bits > diff_bits ? bits : diff_bits;
            }
        } else {
            const double diff = fabs(s_bel[o] - b);
            s_bel[o] = b;
            diff_bits = (unsigned long long)__double_as_longlong(diff);
            double rest;
            s_z[o] = pack_prod(prod_0, prod_1, z0, z1, rest);
            s_prod[o] = rest;
        }
        if (!direct) break;
    }
    /* the wave's max change per variant (K <= 64 divides the wave) */
    diff_bits = wave_max_mod(diff_bits, K);
    const uint32_t lane = tid & 63;
    if (!direct) {
        __syncthreads();
        /* phase 3: the variants that run store products, codes and beliefs */
        if ((uint32_t)tid < nvk && run_k) {
            A.vprod[vst + tid] = s_prod[tid];
            A.vzc[vst + tid] = (uint8_t)s_z[tid];
            A.belief[vst + tid] = s_bel[tid];
        }
    }
    /* per-variant fold: lanes 0..K-1 of each wave hold their variant's max */
    if (lane < K && run_k) fold_diff(A, gr.vbase + lane, ti, diff_bits);
}

/* The sweep kernels launch one block per tile of the batch; block b takes
 * active tile b, and blocks past the active count exit at once.  (A
 * persistent grid walking the tiles in a loop was latency-bound: each tile
 * starts with a chain of dependent loads -- tile, group, offsets, slot map,
 * messages -- that a few blocks per CU cannot overlap; one block per tile
 * lets the dispatcher keep every CU full.) */
template <bool kDamp, int N>
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void batch_factor_pass(BatchArgs A, double damp,
                                                                                    double keep) {
    __shared__ BatchFactorSmem sm;
    const uint32_t ti = blockIdx.x;
    if (ti >= A.act_count[0]) return;
    const Tile tile = A.act_ftiles[ti];
    const GroupDev gr = A.groups[tile.group];
    batch_factor_tile<kDamp, N>(A, gr, tile, damp, keep, sm);
}

__global__ __launch_bounds__(kBlock) __attribute__((unused)) void batch_var_pass(BatchArgs A) {
    __shared__ BatchVarSmem sm;
    /* the active tiles in runs of kBatchVarRun per XCD, as the plain state
     * var pass gathers: the 768-graph batch 267.5 -> 247.1 ms (runs of 64:
     * 253.2 ms; runs of 16 in the factor pass: 269.0 ms) */
    const uint32_t na = A.act_count[1];
    if (blockIdx.x >= na) return;
    const uint32_t ti = xcd_runs<kBatchVarRun>(blockIdx.x, na);
    const Tile tile = A.act_vtiles[ti];
    const GroupDev gr = A.groups[tile.group];
    batch_var_tile(A, gr, tile, ti, sm);
}

/* the active group of active tile ti: pre[j] <= ti < pre[j + 1] */
__device__ __forceinline__ uint32_t tile_group(const uint32_t *__restrict__ pre, uint32_t n_act, uint32_t ti) {
    uint32_t lo = 0, hi = n_act;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= ti) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* The running groups' tiles copied into the flat lists the sweep kernels
 * walk, one thread per active tile (grid-stride). */
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void batch_expand(
    const GroupDev *__restrict__ groups, const Tile *__restrict__ ftiles, const Tile *__restrict__ vtiles,
    const uint32_t *__restrict__ act_n, const uint32_t *__restrict__ act_group, const uint32_t *__restrict__ act_fpre,
    const uint32_t *__restrict__ act_vpre, Tile *__restrict__ act_ftiles, Tile *__restrict__ act_vtiles,
    uint32_t *__restrict__ act_count) {
    const uint32_t n_act = *act_n;
    const uint32_t nf = act_fpre[n_act], nv = act_vpre[n_act];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        act_count[0] = nf;
        act_count[1] = nv;
    }
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < nf; t += stride) {
        const uint32_t j = tile_group(act_fpre, n_act, t);
        act_ftiles[t] = ftiles[groups[act_group[j]].ftile_first + (t - act_fpre[j])];
    }
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < nv; t += stride) {
        const uint32_t j = tile_group(act_vpre, n_act, t);
        act_vtiles[t] = vtiles[groups[act_group[j]].vtile_first + (t - act_vpre[j])];
    }
}

constexpr int kStopBlock = 1024;

/* After a sweep (count = 1) or before the first one (count = 0): per
 * running variant, its max change, counter and next vrun; then the list of
 * groups with a running variant and their tile prefix sums. */
struct StopArgs {
    const GroupDev *__restrict__ groups;
    uint32_t G, nvar;
    uint8_t *__restrict__ vrun;
    uint8_t *__restrict__ vfresh;
    uint32_t *__restrict__ budget;
    unsigned long long *__restrict__ iters;
    double *__restrict__ last;
    unsigned long long *__restrict__ diff;
    uint32_t *__restrict__ act_n, *__restrict__ act_group, *__restrict__ act_fpre, *__restrict__ act_vpre;
    double tol;
};

__global__ __launch_bounds__(kStopBlock) __attribute__((unused)) void batch_stop(StopArgs S, int count) {
    __shared__ uint32_t s_cnt[kStopBlock / 64][3];
    __shared__ uint32_t s_base[3];
    const uint32_t tid = threadIdx.x;
    if (count) {
        for (uint32_t v = tid; v < S.nvar; v += kStopBlock) {
            if (!S.vrun[v]) continue;
            unsigned long long m = 0;
            unsigned long long *sl = S.diff + (uint64_t)v * kBatchSlots * kCtlPad;
#pragma unroll
            for (int j = 0; j < kBatchSlots; j++) {
                const unsigned long long x = sl[j * kCtlPad];
                m = x > m ? x : m;
                sl[j * kCtlPad] = 0ull;
            }
            const double d = __longlong_as_double((long long)m);
            S.iters[v] += 1;
            S.last[v] = d;
            S.vfresh[v] = 0;
            const uint32_t left = S.budget[v] - 1;
            S.budget[v] = left;
            /* tol = -inf: a fixed number of sweeps, NaN changes included */
            S.vrun[v] = (left > 0 && (d > S.tol || S.tol == -__builtin_inf())) ? 1 : 0;
        }
        __syncthreads();
    }
    /* compaction of the groups with a running variant, in group order */
    if (tid == 0) {
        s_base[0] = 0;
        s_base[1] = 0;
        s_base[2] = 0;
    }
    __syncthreads();
    for (uint32_t g0 = 0; g0 < S.G; g0 += kStopBlock) {
        const uint32_t g = g0 + tid;
        uint32_t on = 0, nf = 0, nv = 0;
        if (g < S.G) {
            const GroupDev gr = S.groups[g];
            for (uint32_t k = 0; k < gr.K && !on; k++) on = S.vrun[gr.vbase + k];
            if (on) {
                nf = gr.ftile_count;
                nv = gr.vtile_count;
            }
        }
        /* block scan of (on, nf, nv): wave scans, then across waves */
        uint32_t x[3] = {on, nf, nv};
        const uint32_t lane = tid & 63, wave = tid >> 6;
#pragma unroll
        for (int c = 0; c < 3; c++) {
#pragma unroll
            for (uint32_t off = 1; off < 64; off <<= 1) {
                const uint32_t y = __shfl_up(x[c], off, 64);
                if (lane >= off) x[c] += y;
            }
            if (lane == 63) s_cnt[wave][c] = x[c];
        }
        __syncthreads();
        uint32_t pre[3] = {0, 0, 0};
#pragma unroll
        for (int c = 0; c < 3; c++)
            for (uint32_t w = 0; w < wave; w++) pre[c] += s_cnt[w][c];
        if (on) {
            const uint32_t slot = s_base[0] + pre[0] + x[0] - 1;
            S.act_group[slot] = g;
            S.act_fpre[slot] = s_base[1] + pre[1] + x[1] - nf;
            S.act_vpre[slot] = s_base[2] + pre[2] + x[2] - nv;
        }
        __syncthreads();
        if (tid == kStopBlock - 1) {
#pragma unroll
            for (int c = 0; c < 3; c++) s_base[c] += pre[c] + x[c];
        }
        __syncthreads();
    }
    if (tid == 0) {
        *S.act_n = s_base[0];
        S.act_fpre[s_base[0]] = s_base[1];
        S.act_vpre[s_base[0]] = s_base[2];
    }
}

__global__ __attribute__((unused)) void bfill_pairs(double2 *p, uint64_t n, double c) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = make_double2(c, c);
}

__global__ __attribute__((unused)) void bfill_u8(uint8_t *p, uint64_t n, uint8_t c) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = c;
}

/* one group's messages to a new storage order: variant k of slot e moves
 * from position from[e] to position to[e] (n = slots << lk) */
__global__ __attribute__((unused)) void bmove_msgs(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                   const uint32_t *__restrict__ from, const uint32_t *__restrict__ to,
                                                   uint64_t n, uint32_t lk) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint64_t e = i >> lk, k = i & ((1u << lk) - 1u);
        dst[((uint64_t)to[e] << lk) + k] = src[((uint64_t)from[e] << lk) + k];
    }
}

/* out[i] = src[idx[i]] */
__global__ __attribute__((unused)) void bgather_f64(const double *__restrict__ src, const uint64_t *__restrict__ idx,
                                                    double *__restrict__ out, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = src[idx[i]];
}

}  // namespace

#endif
