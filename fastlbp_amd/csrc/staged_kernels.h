/*
 * staged_kernels.h -- the two sweep kernels with LDS staging.
 *
 * Each block owns a contiguous range of factors (factor pass) or variables
 * (var pass) whose edge slots form one contiguous range of at most kStage
 * slots; the host cuts the ranges (at most kBlock items each) so that every
 * block stages about the same number of edges.  A block
 *   1. gathers the messages of its whole slot range into LDS, every lane
 *      taking slots tid, tid+kBlock, ...: the index loads are coalesced and
 *      all the random gathers of the block are in flight at once, whatever
 *      the degrees of the block's factors or variables;
 *   2. computes, from LDS and in the reference's operation order, every
 *      output of the range.  Lanes take the outputs in a host-made work
 *      order that groups outputs of the same kind (prior, head, body) and
 *      factor degree, or variables of the same degree, so the lanes of a
 *      wave run the same path the same number of steps;
 *   3. writes the slot range back to HBM with coalesced stores (or, in the
 *      push layout, scatters the v->f messages to their factor slots).
 * A factor or variable with more than kStage slots gets a block of its own
 * and is processed in kStage-slot chunks (variables) or straight from
 * global memory with the block's lanes spread over its outputs (factors).
 */
#ifndef FLBP_STAGED_KERNELS_H
#define FLBP_STAGED_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sweep_kernels.h"

namespace {

using namespace flbp::dev;

constexpr int kStage = 512;     /* slots staged per block (8 KiB per message array) */
constexpr int kPerLane = kStage / kBlock;   /* staged slots per lane */

struct LdsIn {
    const double2 *m;
    __device__ __forceinline__ double2 get(int j) const { return m[j]; }
    __device__ __forceinline__ double2 pick(int j) const { return m[j]; }
};

/* The operands of one output of a factor whose slots' messages start at
 * in[0] (stride 1 << lk): operand k is the k-th slot other than the
 * skipped slots a < b (the target and, for a body target, the head; b past
 * the factor for a head target), slot k + [k >= a] + [k + 1 >= b].  The
 * skip counts of operands 0..7 are packed once per output as 2-bit fields
 * of sk (field k: [k >= a] from 0x5555 << 2a, plus [k >= b - 1]; a field
 * holds at most 2, so the two terms add without carries), so each operand
 * costs one field extract.  An operand is read only when the loop that
 * uses it runs, so a wave whose outputs have few operands loads and
 * indexes only those. */
struct SkipIn {
    const double2 *in;
    uint32_t sk;
    uint32_t lk;
    /* a < b, a <= 7, b <= 9 */
    __device__ __forceinline__ static SkipIn make(const double2 *in, int a, int b, uint32_t lk) {
        return SkipIn{in, (0x5555u << (2 * a)) + (0x5555u << (2 * (b - 1))), lk};
    }
    /* k <= 6 (operand lists of factors of degree 8 or less) */
    __device__ __forceinline__ double2 get(int k) const {
        return in[(uint32_t)(k + ((sk >> (2 * k)) & 3u)) << lk];
    }
};

/* One f->v output from a compact operand list: ops.get(0..cnt) are the
 * messages of the factor's slots other than the target t (and, for a body
 * target, other than the head h), in slot order -- exactly the slots the
 * per-output loops of fv_msg visit, in the same order -- so the loops run
 * cnt steps with no per-slot tests and the arithmetic is fv_msg's.  mh is
 * the head slot's message; for a degree-2 head output ops.get(0) is the
 * other slot.  M bounds cnt. */
template <int M, class Ops>
__device__ __forceinline__ double2 fv_list(const Ops &ops, int cnt, uint32_t type, bool clamped, int d,
                                           bool head_target, double2 mh, double p, double q, double mask0,
                                           double mask1) {
    double2 r;
    if (type == 0) {
        r.x = 1 - p;
        r.y = p;
        return r;
    }
    const bool is_and = (type == 1);
    if (!head_target) {
        /* body_msg */
        const double hx = is_and ? mh.y : mh.x, hy = is_and ? mh.x : mh.y;
        const double mx = is_and ? mask1 : mask0, my = is_and ? mask0 : mask1;
        double P = (p - q) * (hx * mx - hy * my);
        double Q = q * hx * mx + (1 - q) * hy * my;
        scale_pair(P, Q);
        if (Q != 0) {
#pragma unroll
            for (int k = 0; k < M; k++) {
                if (k < cnt) {
                    const double2 m = ops.get(k);
                    Q *= m.x + m.y;
                    P *= is_and ? m.y : m.x;
                    scale_pair(P, Q);
                }
            }
        }
        const double QP = Q + P;
        const double marg = Q + QP;
        if (marg != 0) {
            const double a = Q / marg, b = QP / marg;
            r.x = is_and ? a : b;
            r.y = is_and ? b : a;
        } else {
            r.x = 0.5;
            r.y = 0.5;
        }
        return r;
    }
    /* head_uw */
    double u, w;
    if (d == 1) {
        u = p;
        w = 1 - p;
    } else if (d == 2) {
        const double2 m = ops.get(0);
        const double x = is_and ? m.y : m.x, y = is_and ? m.x : m.y;
        u = p * x + q * y;
        w = (1 - p) * x + (1 - q) * y;
    } else {
        double px = 1.0, p10 = 1.0, e1 = 0;
#pragma unroll
        for (int k = 0; k < M; k++) {
            if (k < cnt) {
                const double2 m = ops.get(k);
                const double a0 = m.x + m.y;
                const double x = is_and ? m.y : m.x;
                const double delta = is_and ? m.x : m.y;
                px *= x;
                p10 *= a0;
                scale_pair(px, p10);
                if (x != 0 && a0 == x && delta != 0) e1 += delta / x;
            }
        }
        const double R = e1 * p10 + (p10 - px);
        u = p * px + q * R;
        w = (1 - p) * px + (1 - q) * R;
    }
    double marg0 = is_and ? w : u, marg1 = is_and ? u : w;
    if (clamped) {
        marg0 *= mask0;
        marg1 *= mask1;
    }
    const double marg = marg0 + marg1;
    if (marg != 0) {
        r.x = marg0 / marg;
        r.y = marg1 / marg;
    } else if (clamped && mask0 == 0) {
        r.x = 0.0;
        r.y = 1.0;
    } else if (clamped && mask1 == 0) {
        r.x = 1.0;
        r.y = 0.0;
    } else {
        r.x = 0.5;
        r.y = 0.5;
    }
    return r;
}

/* block-wide max of per-lane diff bits, folded into the sweep's slots */
__device__ __forceinline__ void block_fold_diff(unsigned long long *ctl, int k, unsigned long long bits,
                                                unsigned long long *red) {
    bits = wave_max_u64(bits);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = bits;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = red[0];
#pragma unroll
        for (int w = 1; w < kBlock / 64; w++) m = red[w] > m ? red[w] : m;
        ctl_fold(ctl, k, m);
    }
}

/* Message layouts of the staged kernels.  Pull: v->f var-major, f->v
 * factor-major; both passes gather their inputs and store contiguously.
 * Push: both factor-major; the factor pass is contiguous, the variable pass
 * gathers and scatters.  Scatter: v->f factor-major, f->v var-major (plus a
 * factor-major copy of f->v, kept for damping); both passes read
 * contiguously and scatter their outputs, the cheapest random access. */
enum : int { kLayoutPull = 0, kLayoutPush = 2, kLayoutScatter = 3 };

/* f->v outputs of one factor: the factor-major copy (not needed by the
 * scatter layout unless damped) and, in the scatter layout, the var-major
 * array through fac_vpos */
template <bool kDamp, int kLayout>
__device__ __forceinline__ void put_fv(double2 *m_fv, double2 *m_fv_vm, const uint32_t *fac_vpos, uint32_t slot,
                                       double2 r) {
    if (kLayout != kLayoutScatter || kDamp) m_fv[slot] = r;
    if (kLayout == kLayoutScatter) m_fv_vm[fac_vpos[slot]] = r;
}

template <bool kDamp, int kLayout, class In>
__device__ __forceinline__ void wide_factor(const In &in, uint32_t base, uint32_t n, uint32_t meta, double p,
                                            double q, double mask0, double mask1, const uint32_t *fac_vpos,
                                            const double2 *m_fv_old, double2 *m_fv, double2 *m_fv_vm,
                                            double damp, double keep) {
    const uint32_t type = meta & 3u;
    const bool clamped = (meta & 4u) != 0;
    const int h = (int)(meta >> 8);
    const double2 mh = in.pick(h), mo = in.pick(h == 0 ? 1 : 0);
    for (uint32_t t = threadIdx.x; t < n; t += kBlock) {
        double2 r = fv_msg<0>(in, type, clamped, (int)n, (int)t, h, mh, mo, p, q, mask0, mask1);
        if (kDamp) {
            const double2 old = m_fv_old[base + t];
            r.x = keep * r.x + damp * old.x;
            r.y = keep * r.y + damp * old.y;
        }
        put_fv<kDamp, kLayout>(m_fv, m_fv_vm, fac_vpos, base + t, r);
    }
}

/* LDS of one factor tile */
struct FactorSmem {
    double2 in[kStage];
    double2 out[kStage];
    uint32_t start[kBlock + 1];   /* factor offsets relative to the tile's first slot */
    uint32_t meta[kBlock];
    double p[kBlock], q[kBlock];
};

/* LDS of one variable tile */
struct VarSmem {
    double2 m[kStage];
    double prod[2][kBlock];
    uint32_t z[kBlock];            /* z0 | z1 << 16 (a staged variable has <= kStage edges) */
    uint8_t owner[kStage];         /* local variable of each staged slot */
    uint32_t start[kBlock + 1];
    double bel[kBlock];
    unsigned long long red[kBlock / 64];
};

/* One output of a staged factor tile: the f->v message of local slot sl of
 * local factor lf, from the v->f inputs of the tile's slots in sm.in (slot
 * order) and its factors' offsets and parameters in sm.start / meta / p / q. */
template <int N, class Smem>
__device__ __forceinline__ double2 factor_output(const Smem &sm, uint32_t sl, uint32_t lf, uint32_t f0,
                                                 const double2 *__restrict__ fac_mask) {
    static_assert(N <= 8, "SkipIn packs the operands of factors of degree 8 or less");
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
        const double2 mk = fac_mask[f0 + lf];
        mask0 = mk.x;
        mask1 = mk.y;
    }
    if (d <= N) {
        /* the other slots' messages, in slot order, from LDS */
        const bool head_target = (t == h);
        const int cnt = type == 0 ? 0 : d - (head_target ? 1 : 2);
        const SkipIn ops =
            SkipIn::make(sm.in + s, head_target ? t : (t < h ? t : h), head_target ? N + 1 : (t < h ? h : t), 0u);
        return fv_list<N - 1>(ops, cnt, type, clamped, d, head_target, sm.in[s + h], p, q, mask0, mask1);
    }
    LdsIn in{sm.in + s};
    const double2 mh = in.pick(h), mo = in.pick(h == 0 ? 1 : 0);
    return fv_msg<0>(in, type, clamped, d, t, h, mh, mo, p, q, mask0, mask1);
}

/* Phase 2 of a factor tile: one lane per output slot, in the tile's work
 * order (w = local slot | local factor << 10), from LDS only: sm.in holds
 * the tile's v->f inputs, sm.out receives the f->v outputs (and holds the
 * old f->v messages when damped). */
template <bool kDamp, int N>
__device__ __forceinline__ void factor_tile_compute(FactorSmem &sm, const uint32_t (&w)[kPerLane], uint32_t n,
                                                    uint32_t f0, const double2 *__restrict__ fac_mask, double damp,
                                                    double keep) {
    const int tid = threadIdx.x;
#pragma unroll 1
    for (int j = 0; j < kPerLane; j++) {
        if (tid + j * kBlock >= n) break;
        uint32_t wj = 0;
#pragma unroll
        for (int x = 0; x < kPerLane; x++) wj |= (j == x) ? w[x] : 0u;
        const uint32_t sl = wj & 1023u, lf = wj >> 10;
        double2 r = factor_output<N>(sm, sl, lf, f0, fac_mask);
        if (kDamp) {
            const double2 old = sm.out[sl];
            r.x = keep * r.x + damp * old.x;
            r.y = keep * r.y + damp * old.y;
        }
        sm.out[sl] = r;
    }
}

/* The factor pass over factors [f0, f1) of one graph (every pointer is the
 * graph's own: offsets and slots are graph-local).  The whole block calls
 * it; it ends with the block's stores issued and LDS free for reuse after a
 * barrier. */
template <bool kDamp, int N, int kLayout>
__device__ __forceinline__ void factor_tile(
    const uint32_t *__restrict__ fac_off, const uint32_t *__restrict__ fac_meta,
    const double *__restrict__ fac_p, const double *__restrict__ fac_q, const double2 *__restrict__ fac_mask,
    const uint32_t *__restrict__ fac_vpos, const uint32_t *__restrict__ fac_work,
    const double2 *__restrict__ m_vf, const double2 *m_fv_old, double2 *m_fv, double2 *__restrict__ m_fv_vm,
    double damp, double keep, uint32_t f0, uint32_t f1, FactorSmem &sm) {
    const uint32_t base = fac_off[f0];
    const uint32_t n = fac_off[f1] - base;
    const int tid = threadIdx.x;
    if (n > (uint32_t)kStage) {
        /* one wide factor: lanes spread over its outputs, inputs from HBM */
        const uint32_t meta = fac_meta[f0];
        const double p = fac_p[f0], q = (meta & 3u) ? fac_q[f0] : 0.0;
        double mask0 = 1.0, mask1 = 1.0;
        if (meta & 4u) {
            const double2 mk = fac_mask[f0];
            mask0 = mk.x;
            mask1 = mk.y;
        }
        if (kLayout == kLayoutPull)
            wide_factor<kDamp, kLayout>(GlobalIn{m_vf, fac_vpos + base}, base, n, meta, p, q, mask0, mask1,
                                        fac_vpos, m_fv_old, m_fv, m_fv_vm, damp, keep);
        else   /* LdsIn is a plain pointer read: it works on global memory too */
            wide_factor<kDamp, kLayout>(LdsIn{m_vf + base}, base, n, meta, p, q, mask0, mask1, fac_vpos,
                                        m_fv_old, m_fv, m_fv_vm, damp, keep);
        return;
    }
    /* phase 1: the tile's work entries into registers, its factors'
     * offsets and parameters into LDS, and its messages into LDS, all loads
     * independent of each other */
    const uint32_t nf = f1 - f0;
    uint32_t w[kPerLane], vp[kPerLane];   /* work entries; var-major positions (gather or scatter index) */
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        w[j] = i < n ? fac_work[base + i] : 0u;
        vp[j] = (kLayout != kLayoutPush && i < n) ? fac_vpos[base + i] : 0u;
    }
    if ((uint32_t)tid < nf) {
        sm.start[tid] = fac_off[f0 + tid] - base;
        sm.meta[tid] = fac_meta[f0 + tid];
        sm.p[tid] = fac_p[f0 + tid];
        sm.q[tid] = fac_q[f0 + tid];
    }
    if (tid == 0) sm.start[nf] = n;
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        if (i < n) {
            sm.in[i] = kLayout == kLayoutPull ? m_vf[vp[j]] : m_vf[base + i];
            if (kDamp) sm.out[i] = m_fv_old[base + i];
        }
    }
    __syncthreads();
    factor_tile_compute<kDamp, N>(sm, w, n, f0, fac_mask, damp, keep);
    __syncthreads();
    /* phase 3: stores (the scatter positions came with phase 1) */
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        if (i < n) {
            const double2 r = sm.out[i];
            if (kLayout != kLayoutScatter || kDamp) m_fv[base + i] = r;
            if (kLayout == kLayoutScatter) m_fv_vm[vp[j]] = r;
        }
    }
}

/* The variable pass over variables [v0, v1) of one graph (graph-local
 * pointers).  f->v messages are read var-major (scatter layout) or
 * gathered from their factor slots; v->f messages are stored var-major
 * (pull layout) or scattered to their factor slots.  Returns the lane's
 * max |b - b_old| as IEEE bits (0 for lanes without a variable). */
template <int kLayout>
__device__ __forceinline__ unsigned long long var_tile(
    const uint32_t *__restrict__ var_off, const uint32_t *__restrict__ var_edge,
    const uint16_t *__restrict__ var_work, const double2 *__restrict__ m_fv, double2 *__restrict__ m_vf,
    double *__restrict__ belief, uint32_t v0, uint32_t v1, VarSmem &sm) {
    const uint32_t base = var_off[v0];
    const uint32_t n = var_off[v1] - base;
    const int tid = threadIdx.x;
    unsigned long long diff_bits = 0;
    if (n > (uint32_t)kStage) {
        /* one wide variable: the product runs in edge order on lane 0 over
         * kStage-slot chunks gathered by the whole block, then every lane
         * writes its share of the leave-one-out messages */
        double prod_0 = 1.0, prod_1 = 1.0;
        uint32_t z0 = 0, z1 = 0;
        for (uint32_t c = 0; c < n; c += kStage) {
            const uint32_t m = n - c < (uint32_t)kStage ? n - c : (uint32_t)kStage;
            for (uint32_t e = tid; e < m; e += kBlock)
                sm.m[e] = kLayout == kLayoutScatter ? m_fv[base + c + e] : m_fv[var_edge[base + c + e]];
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
            sm.prod[0][0] = prod_0;
            sm.prod[1][0] = prod_1;
            sm.z[0] = z0;
            sm.z[1] = z1;
            const double b0 = prod_0 * (double)(z0 == 0);
            const double b1 = prod_1 * (double)(z1 == 0);
            const double b = b1 / (b0 + b1);
            const double diff = fabs(belief[v0] - b);
            belief[v0] = b;
            diff_bits = (unsigned long long)__double_as_longlong(diff);
        }
        __syncthreads();
        prod_0 = sm.prod[0][0];
        prod_1 = sm.prod[1][0];
        z0 = sm.z[0];
        z1 = sm.z[1];
        for (uint32_t e = tid; e < n; e += kBlock) {
            const uint32_t sl = var_edge[base + e];
            const double2 x = kLayout == kLayoutScatter ? m_fv[base + e] : m_fv[sl];
            m_vf[kLayout == kLayoutPull ? base + e : sl] = vf_msg(prod_0, prod_1, z0, z1, x);
        }
        return diff_bits;
    }
    /* phase 1: offsets, old beliefs and the work entry of each variable,
     * and the tile's messages, all loads independent of each other */
    const uint32_t nv = v1 - v0;
    uint32_t lv = 0;
    uint32_t ve[kPerLane];   /* factor-major slots: gather or scatter index */
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        ve[j] = i < n ? var_edge[base + i] : 0u;
    }
    if ((uint32_t)tid < nv) {
        lv = var_work[v0 + tid];
        sm.start[tid] = var_off[v0 + tid] - base;
        sm.bel[tid] = belief[v0 + tid];
    }
    if (tid == 0) sm.start[nv] = n;
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        if (i < n) sm.m[i] = kLayout == kLayoutScatter ? m_fv[base + i] : m_fv[ve[j]];
    }
    __syncthreads();
    /* phase 2: products, one lane per variable in the tile's work order
     * (by degree), from LDS only */
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
            sm.owner[s + j] = (uint8_t)lv;
        }
        const double b0 = prod_0 * (double)(z0 == 0);
        const double b1 = prod_1 * (double)(z1 == 0);
        const double b = b1 / (b0 + b1);
        const double diff = fabs(sm.bel[lv] - b);
        sm.bel[lv] = b;
        diff_bits = (unsigned long long)__double_as_longlong(diff);
        sm.prod[0][lv] = prod_0;
        sm.prod[1][lv] = prod_1;
        sm.z[lv] = z0 | (z1 << 16);
    }
    __syncthreads();
    /* phase 3: leave-one-out messages, one lane per slot, stored coalesced
     * (var-major) or scattered to the factor slots */
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        if (i < n) {
            const uint32_t lv2 = sm.owner[i];
            const uint32_t z = sm.z[lv2];
            const double2 r = vf_msg(sm.prod[0][lv2], sm.prod[1][lv2], z & 0xffffu, z >> 16, sm.m[i]);
            m_vf[kLayout == kLayoutPull ? base + i : ve[j]] = r;
        }
    }
    if ((uint32_t)tid < nv) belief[v0 + tid] = sm.bel[tid];
    return diff_bits;
}

/* Workgroups go to the 8 XCDs round-robin (workgroup b to XCD b mod 8).
 * Tile xcd_block(b, n) gives each XCD one contiguous run of tiles, so
 * neighbouring tiles, whose scattered stores or gathered loads touch
 * neighbouring lines, share one L2: partial-line writes merge there and a
 * gathered line is fetched once.  A bijection of [0, n): XCD x runs n / 8
 * tiles, plus one if x < n mod 8. */
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
    const uint32_t x = b & 7u, i = b >> 3, q = n >> 3, r = n & 7u;
    return x * q + (x < r ? x : r) + i;
}

/* Tile of workgroup b when each XCD takes runs of kRun consecutive tiles
 * in turn: within each group of 8 * kRun workgroups, XCD x = b mod 8 runs
 * tiles [x * kRun, (x + 1) * kRun) of the group; the last, partial group
 * keeps the plain order.  A bijection of [0, n). */
template <uint32_t kRun>
__device__ __forceinline__ uint32_t xcd_runs(uint32_t b, uint32_t n) {
    constexpr uint32_t g = 8u * kRun;
    if (b >= n - n % g) return b;
    return b - b % g + (b & 7u) * kRun + (b >> 3) % kRun;
}

/* tiles per XCD run in the pull layout's var pass */
constexpr uint32_t kPullVarRun = 64;
/* tiles per XCD run in the state layout's factor pass: on the 1M-variable
 * graph 62.2 us per launch against 64.4 in the plain order (runs of 64 and
 * 256 tiles: 64.1 and 63.5 us, one run per XCD 64.8 us) */
constexpr uint32_t kStateFacRun = 16;
/* tiles per XCD run in the state layout's var pass: on the 1M-variable
 * graph with 512-slot tiles 38.4 us per launch against 40.6 with runs of 64
 * (16: 47.9, 32: 44.6, 128: 38.7, 512: 39.2, 1,024: 40.3, one run per XCD:
 * 41.1); with 768-slot tiles runs of 512 36.0 us (128: 36.9, 256: 36.4).
 * The batched var pass takes runs of 256. */
constexpr uint32_t kStateVarRun = 512;
constexpr uint32_t kBatchVarRun = 256;

/* the factor pass of a plain solve: block b owns factors [blk[b], blk[b+1]) */
template <bool kDamp, int N, int kLayout>
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void factor_pass_staged(
    const uint32_t *__restrict__ blk, const uint32_t *__restrict__ fac_off,
    const uint32_t *__restrict__ fac_meta, const double *__restrict__ fac_p,
    const double *__restrict__ fac_q, const double2 *__restrict__ fac_mask,
    const uint32_t *__restrict__ fac_vpos, const uint32_t *__restrict__ fac_work,
    const double2 *__restrict__ m_vf, const double2 *m_fv_old, double2 *m_fv, double2 *__restrict__ m_fv_vm,
    double damp, double keep, const unsigned long long *__restrict__ ctl, int k, double tol) {
    if (chunk_stopped(ctl, k, tol)) return;
    __shared__ FactorSmem sm;
    /* the pull layout's factor pass gathers its v->f messages from the
     * var-major lines of nearby variables: XCD-local runs of tiles measured
     * 3% faster per launch at 2M and 5M variables (264.6 -> 256.2 us at 5M);
     * the scatter layout reads its inputs in order and keeps the plain order */
    const uint32_t b = kLayout == kLayoutPull ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    factor_tile<kDamp, N, kLayout>(fac_off, fac_meta, fac_p, fac_q, fac_mask, fac_vpos, fac_work, m_vf, m_fv_old,
                                   m_fv, m_fv_vm, damp, keep, blk[b], blk[b + 1], sm);
}

/* the variable pass of a plain solve: block b owns variables [blk[b], blk[b+1]) */
template <int kLayout>
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void var_pass_staged(
    const uint32_t *__restrict__ blk, const uint32_t *__restrict__ var_off,
    const uint32_t *__restrict__ var_edge, const uint16_t *__restrict__ var_work,
    const double2 *__restrict__ m_fv, double2 *__restrict__ m_vf, double *__restrict__ belief,
    unsigned long long *__restrict__ ctl, int k, double tol) {
    if (chunk_stopped(ctl, k, tol)) return;
    __shared__ VarSmem sm;
    /* the scatter layout's var pass writes its v->f messages into the
     * factor-major lines of nearby factors: XCD-local runs of tiles measured
     * ~1 us faster per launch at 1M variables.  The pull layout's var pass
     * gathers from factor-major lines: one run per XCD measured 11% slower
     * than the plain order (227.7 -> 252.6 us at 5M), runs of 64 tiles 2%
     * faster (228.0 -> 223.0 us at 5M, 94.5 -> 92.2 us at 2M; runs of 2 to
     * 1,024 tiles measured, 64 and 128 best) */
    const uint32_t b = kLayout == kLayoutScatter ? xcd_block(blockIdx.x, gridDim.x)
                       : kLayout == kLayoutPull  ? xcd_runs<kPullVarRun>(blockIdx.x, gridDim.x)
                                                 : blockIdx.x;
    const unsigned long long bits =
        var_tile<kLayout>(var_off, var_edge, var_work, m_fv, m_vf, belief, blk[b], blk[b + 1], sm);
    block_fold_diff(ctl, k, bits, sm.red);
}

}  // namespace

#endif
