/*
 * state_kernels.h -- the sweep kernels of the state layout (kernel
 * variant 4).
 *
 * The reference stores both message directions per edge and recomputes all
 * of them every iteration: calcMessageFVFusedAll writes f->v, calcProdAll
 * forms each variable's zero-counted product of its f->v messages, and
 * calcMessageVFAll divides one f->v message back out of that product to get
 * each v->f message (csrc/kernel/causal_bp_parall.cu:20-92, 94-398).  The
 * v->f message of edge (I, v) is therefore a function of two values only:
 * the variable's product record and the f->v message of the same edge.
 *
 * This layout keeps exactly those two: f->v messages by factor tile (the
 * factor pass reads and rewrites them in place, contiguously) and, per
 * variable, its product pair in 8 bytes (the rescaled pair's non-unit
 * value) and a one-byte code of its zero counts and unit (pack_prod).  The
 * variable pass forms each variable's product, belief and |b - b_old| and
 * stores the record and code -- it writes no per-edge data.  The factor
 * pass derives every v->f input of its tile from the record and code of
 * the slot's variable and the slot's own
 * previous f->v message (vf_msg, the reference's division-based
 * leave-one-out), then computes and stores the new f->v messages.  The
 * factor pass's random reads go to V-sized arrays (the codes of a million
 * variables fit one XCD's L2), which caches hold far better than a
 * per-edge array.  The one per-edge random access of a sweep is the
 * variable pass's gather of its f->v messages from their storage positions
 * (var_src).  Within a factor tile the messages are stored in the tile's
 * work order, heads before bodies: a variable's head messages then sit in
 * the short head run of its factors' tile, so the gathers of a variable
 * tile touch fewer lines than with each factor's slots side by side
 * (353 against 402 MB per pass on the 1M-variable graph by a line count).
 * (Scattering a var-major copy from the factor pass instead, so that the
 * variable pass reads in order, measured 2.2x slower per factor pass on a
 * 4.5M-edge graph in HBM: 177 against 80 us.)
 *
 * The arithmetic is the reference's, so results are bit-identical to the
 * other layouts: each v->f value is vf_msg of the same product and message
 * the variable pass of the previous sweep would have used.  The first sweep
 * after init() takes the init value (c, c) as its v->f input, as the
 * reference's first calcMessageFVFusedAll does; after a caller sets v->f
 * messages explicitly, the next factor pass reads them from m_vf, stored in
 * the same order as the f->v messages.
 */
#ifndef FLBP_STATE_KERNELS_H
#define FLBP_STATE_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "plan_internal.h"
#include "staged_kernels.h"

namespace {

using namespace flbp::dev;

struct ConstIn {
    double c;
    __device__ __forceinline__ double2 get(int) const { return make_double2(c, c); }
    __device__ __forceinline__ double2 pick(int) const { return make_double2(c, c); }
};

/* the v->f message of an edge from its variable's product record (rest,
 * code: pack_prod) and the edge's f->v message m */
__device__ __forceinline__ double2 vf_of(double rest, uint32_t code, double2 m) {
    const double2 prod = unpack_prod(rest, code);
    return vf_msg(prod.x, prod.y, code & 3u, (code >> 2) & 3u, m);
}

/* LDS of one factor tile of the state layout: the v->f inputs of its slots
 * and its factors' offsets and parameters (the outputs go from registers
 * to HBM). */
struct StateFactorSmem {
    double2 in[kStage];
    uint32_t start[kBlock + 1];   /* factor offsets relative to the tile's first slot */
    uint32_t meta[kBlock];
    double p[kBlock], q[kBlock];
};

/* The factor pass over factors [f0, f1) in the state layout.
 *
 * Storage order: the f->v message of a staged tile's slot sits at the
 * tile's base plus the slot's position in the tile's work order (so a
 * tile's head outputs, which the work order puts before its body outputs,
 * form one run, next to the runs of the neighbouring tiles); a wide
 * factor's slots keep their slot order.  idx is the storage position ->
 * device variable map, m_vf the explicit v->f inputs in storage order
 * (kVfArray).  m_fv is read and rewritten in place: every lane reads the
 * old messages of the storage positions it will write, and each slot
 * belongs to one factor of one tile. */
template <bool kDamp, int N, int kMode>
__device__ __forceinline__ void factor_tile_state(
    const uint32_t *__restrict__ fac_off, const uint32_t *__restrict__ fac_meta,
    const double *__restrict__ fac_p, const double *__restrict__ fac_q, const double2 *__restrict__ fac_mask,
    const uint32_t *__restrict__ idx, const uint32_t *__restrict__ fac_work, const double *__restrict__ vprod,
    const uint8_t *__restrict__ vzc, const double2 *m_vf, double2 *m_fv, double2 *scratch, double c, double damp,
    double keep, uint32_t f0, uint32_t f1, StateFactorSmem &sm) {
    const uint32_t base = fac_off[f0];
    const uint32_t n = fac_off[f1] - base;
    const int tid = threadIdx.x;
    if (n > (uint32_t)kStage) {
        /* one wide factor, from global memory.  In the state mode its v->f
         * inputs go to scratch first (the factor's own slot range of it):
         * its outputs overwrite the f->v messages they are derived from. */
        const uint32_t meta = fac_meta[f0];
        const double p = fac_p[f0], q = (meta & 3u) ? fac_q[f0] : 0.0;
        double mask0 = 1.0, mask1 = 1.0;
        if (meta & 4u) {
            const double2 mk = fac_mask[f0];
            mask0 = mk.x;
            mask1 = mk.y;
        }
        if constexpr (kMode == kVfArray) {
            wide_factor<kDamp, kLayoutPull>(LdsIn{m_vf + base}, base, n, meta, p, q, mask0, mask1, nullptr, m_fv,
                                            m_fv, nullptr, damp, keep);
        } else if constexpr (kMode == kVfConst) {
            wide_factor<kDamp, kLayoutPull>(ConstIn{c}, base, n, meta, p, q, mask0, mask1, nullptr, m_fv, m_fv,
                                            nullptr, damp, keep);
        } else {
            for (uint32_t t = tid; t < n; t += kBlock) {
                const uint32_t v = idx[base + t];
                scratch[base + t] = vf_of(vprod[v], vzc[v], m_fv[base + t]);
            }
            __syncthreads();
            wide_factor<kDamp, kLayoutPull>(LdsIn{scratch + base}, base, n, meta, p, q, mask0, mask1, nullptr, m_fv,
                                            m_fv, nullptr, damp, keep);
        }
        return;
    }
    /* phase 1: work entries, variables and old messages of the lane's
     * storage positions into registers, the factors' offsets and parameters
     * into LDS, and each slot's v->f input (derived from the products and
     * the old f->v message) into LDS at its slot */
    const uint32_t nf = f1 - f0;
    /* the tile's arrays from wave-uniform base pointers: each lane then
     * addresses them by its own small offset, in one 32-bit operation */
    const uint32_t *t_work = fac_work + base, *t_idx = idx + base;
    double2 *t_fv = m_fv + base;
    uint32_t w[kPerLane], ix[kPerLane];
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        w[j] = i < n ? t_work[i] : 0u;
        ix[j] = (kMode == kVfState && i < n) ? t_idx[i] : 0u;
    }
    if ((uint32_t)tid < nf) {
        sm.start[tid] = (fac_off + f0)[tid] - base;
        sm.meta[tid] = (fac_meta + f0)[tid];
        sm.p[tid] = (fac_p + f0)[tid];
        sm.q[tid] = (fac_q + f0)[tid];
    }
    if (tid == 0) sm.start[nf] = n;
    double2 old[kPerLane], vin[kPerLane];
    double pr[kPerLane];
    uint32_t zc[kPerLane];
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        old[j] = make_double2(0.0, 0.0);
        if (i < n) {
            if (kDamp || kMode == kVfState) old[j] = t_fv[i];
            if constexpr (kMode == kVfState) {
                pr[j] = vprod[ix[j]];
                zc[j] = vzc[ix[j]];
            } else if constexpr (kMode == kVfArray) {
                vin[j] = (m_vf + base)[i];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        if (i < n) {
            double2 in;
            if constexpr (kMode == kVfState) in = vf_of(pr[j], zc[j], old[j]);
            else if constexpr (kMode == kVfConst) in = make_double2(c, c);
            else in = vin[j];
            sm.in[w[j] & 1023u] = in;
        }
    }
    __syncthreads();
    /* phase 2: each lane's outputs in work order, stored straight to its
     * storage positions (coalesced: lane order is storage order); unrolled,
     * so each output's work entry and old message stay in their registers
     * (57.3 against 58.0 us per launch with the loop rolled) */
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        if (i >= n) break;
        uint32_t wj = 0;
        double2 oj = make_double2(0.0, 0.0);
#pragma unroll
        for (int x = 0; x < kPerLane; x++) {
            wj |= (j == x) ? w[x] : 0u;
            if (kDamp) {
                oj.x = (j == x) ? old[x].x : oj.x;
                oj.y = (j == x) ? old[x].y : oj.y;
            }
        }
        double2 r = factor_output<N>(sm, wj & 1023u, wj >> 10, f0, fac_mask);
        if (kDamp) {
            r.x = keep * r.x + damp * oj.x;
            r.y = keep * r.y + damp * oj.y;
        }
        t_fv[i] = r;
    }
}

/* LDS of one variable tile of the state layout (kStateVarSlots slots:
 * the pass stages no outputs, so its tiles hold more messages than the
 * other layouts' 512) */
constexpr int kVarStage = (int)flbp::kStateVarSlots;
constexpr int kVarPerLane = (kVarStage + kBlock - 1) / kBlock;
struct StateVarSmem {
    double2 m[kVarStage];
    double rest[kBlock];
    uint32_t z[kBlock];
    uint32_t start[kBlock + 1];
    double bel[kBlock];
    unsigned long long red[kBlock / 64];
};

/* The variable pass over variables [v0, v1) in the state layout: the f->v
 * messages gathered from their factor slots, then each variable's product
 * pair, zero-count code and belief stored.  Returns the lane's max
 * |b - b_old| as IEEE bits. */
__device__ __forceinline__ unsigned long long var_tile_state(
    const uint32_t *__restrict__ var_off, const uint32_t *__restrict__ var_edge,
    const uint16_t *__restrict__ var_work, const double2 *__restrict__ m_fv, double *__restrict__ vprod,
    uint8_t *__restrict__ vzc, double *__restrict__ belief, uint32_t v0, uint32_t v1, StateVarSmem &sm) {
    const uint32_t base = var_off[v0];
    const uint32_t n = var_off[v1] - base;
    const int tid = threadIdx.x;
    unsigned long long diff_bits = 0;
    if (n > (uint32_t)kVarStage) {
        /* one wide variable: the product in edge order on lane 0 over
         * kVarStage-slot chunks gathered by the whole block */
        double prod_0 = 1.0, prod_1 = 1.0;
        uint32_t z0 = 0, z1 = 0;
        for (uint32_t c = 0; c < n; c += kVarStage) {
            const uint32_t m = n - c < (uint32_t)kVarStage ? n - c : (uint32_t)kVarStage;
            for (uint32_t e = tid; e < m; e += kBlock) sm.m[e] = m_fv[var_edge[base + c + e]];
            __syncthreads();
            if (tid == 0) {
                for (uint32_t e = 0; e < m; e++) {
                    const double2 x = sm.m[e];
                    if (x.x == 0) z0++; else prod_0 *= x.x;
                    if (x.y == 0) z1++; else prod_1 *= x.y;
                    scale_pair(prod_0, prod_1);
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            const double b0 = prod_0 * (double)(z0 == 0);
            const double b1 = prod_1 * (double)(z1 == 0);
            const double b = b1 / (b0 + b1);
            const double diff = fabs(belief[v0] - b);
            belief[v0] = b;
            double rest;
            vzc[v0] = pack_prod(prod_0, prod_1, z0, z1, rest);
            vprod[v0] = rest;
            diff_bits = (unsigned long long)__double_as_longlong(diff);
        }
        return diff_bits;
    }
    /* phase 1: offsets, old beliefs, work entries and the tile's messages,
     * all loads independent of each other */
    const uint32_t nv = v1 - v0;
    /* the tile's arrays from wave-uniform base pointers (one 32-bit lane
     * offset per access) */
    double *t_bel = belief + v0, *t_prod = vprod + v0;
    uint8_t *t_zc = vzc + v0;
    uint32_t lv = 0;
    uint32_t ve[kVarPerLane];
#pragma unroll
    for (int j = 0; j < kVarPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        ve[j] = i < n ? (var_edge + base)[i] : 0u;
    }
    if ((uint32_t)tid < nv) {
        lv = (var_work + v0)[tid];
        sm.start[tid] = (var_off + v0)[tid] - base;
        sm.bel[tid] = t_bel[tid];
    }
    if (tid == 0) sm.start[nv] = n;
#pragma unroll
    for (int j = 0; j < kVarPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        if (i < n) sm.m[i] = m_fv[ve[j]];
    }
    __syncthreads();
    /* phase 2: products, one lane per variable in the tile's work order */
    if ((uint32_t)tid < nv) {
        const uint32_t s = sm.start[lv];
        const int d = (int)(sm.start[lv + 1] - s);
        double prod_0 = 1.0, prod_1 = 1.0;
        uint32_t z0 = 0, z1 = 0;
        for (int j = 0; j < d; j++) {
            const double2 x = sm.m[s + j];
            if (x.x == 0) z0++; else prod_0 *= x.x;
            if (x.y == 0) z1++; else prod_1 *= x.y;
            scale_pair(prod_0, prod_1);
        }
        const double b0 = prod_0 * (double)(z0 == 0);
        const double b1 = prod_1 * (double)(z1 == 0);
        const double b = b1 / (b0 + b1);
        const double diff = fabs(sm.bel[lv] - b);
        sm.bel[lv] = b;
        diff_bits = (unsigned long long)__double_as_longlong(diff);
        double rest;
        sm.z[lv] = pack_prod(prod_0, prod_1, z0, z1, rest);
        sm.rest[lv] = rest;
    }
    __syncthreads();
    /* phase 3: products, codes and beliefs, coalesced */
    if ((uint32_t)tid < nv) {
        t_prod[tid] = sm.rest[tid];
        t_zc[tid] = (uint8_t)sm.z[tid];
        t_bel[tid] = sm.bel[tid];
    }
    return diff_bits;
}

template <bool kDamp, int N, int kMode>
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void factor_pass_state(
    const uint32_t *__restrict__ blk, const uint32_t *__restrict__ fac_off,
    const uint32_t *__restrict__ fac_meta, const double *__restrict__ fac_p,
    const double *__restrict__ fac_q, const double2 *__restrict__ fac_mask,
    const uint32_t *__restrict__ idx, const uint32_t *__restrict__ fac_work, const double *__restrict__ vprod,
    const uint8_t *__restrict__ vzc, const double2 *m_vf, double2 *m_fv, double2 *scratch, double c, double damp,
    double keep, const unsigned long long *__restrict__ ctl, int k, double tol) {
    if (chunk_stopped(ctl, k, tol)) return;
    __shared__ StateFactorSmem sm;
    /* tiles of neighbouring head variables read neighbouring product
     * records: runs of kStateFacRun tiles per XCD share them in one L2 */
    const uint32_t b = xcd_runs<kStateFacRun>(blockIdx.x, gridDim.x);
    factor_tile_state<kDamp, N, kMode>(fac_off, fac_meta, fac_p, fac_q, fac_mask, idx, fac_work, vprod, vzc, m_vf,
                                       m_fv, scratch, c, damp, keep, blk[b], blk[b + 1], sm);
}

__global__ __launch_bounds__(kBlock) __attribute__((unused)) void var_pass_state(
    const uint32_t *__restrict__ blk, const uint32_t *__restrict__ var_off,
    const uint32_t *__restrict__ var_edge, const uint16_t *__restrict__ var_work,
    const double2 *__restrict__ m_fv, double *__restrict__ vprod, uint8_t *__restrict__ vzc,
    double *__restrict__ belief, unsigned long long *__restrict__ ctl, int k, double tol) {
    if (chunk_stopped(ctl, k, tol)) return;
    __shared__ StateVarSmem sm;
    /* gathers from the factor tiles' lines: runs of kStateVarRun tiles per XCD */
    const uint32_t b = xcd_runs<kStateVarRun>(blockIdx.x, gridDim.x);
    const unsigned long long bits =
        var_tile_state(var_off, var_edge, var_work, m_fv, vprod, vzc, belief, blk[b], blk[b + 1], sm);
    block_fold_diff(ctl, k, bits, sm.red);
}

/* the v->f messages of the state layout, materialized in its storage
 * order: vf[s] = vf_msg(products of the variable at storage position s,
 * f->v[s]) */
__global__ __attribute__((unused)) void vf_from_state(const double *__restrict__ vprod, const uint8_t *__restrict__ vzc,
                                                      const uint32_t *__restrict__ fac_var,
                                                      const double2 *__restrict__ m_fv, double2 *__restrict__ vf,
                                                      uint64_t n) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) {
        const uint32_t v = fac_var[s];
        vf[s] = vf_of(vprod[v], vzc[v], m_fv[s]);
    }
}

}  // namespace

#endif
