/*
 * seq_kernels.h -- device code of the sequential schedules (SEQFIX / TOPO).
 *
 * One launch runs every item of one dependency level (seq_host.cpp):
 * blocks [0, vf_blocks) take its v->f items, then its f->v items, then its
 * group lanes.  Both message arrays are factor-major here: slot s of factor
 * I holds the v->f message m_vf[s] and the f->v message m_fv[s] of edge
 * (I, variable of s).
 *
 *   v->f item (slot s, variable j):
 *     calcMessageVFKernel, csrc/kernel/causal_bp_seq.cu:113-147 -- the
 *     leave-one-out message from j's product and zero counts with the
 *     f->v message of the same edge divided out, normalized with no (0, 0)
 *     guard.
 *   f->v item (var-major position, variable i, factor I, the factor's
 *     first slot a; degree d, type, target slot t and head slot h in a
 *     second word; p, q and the head mask in the factor's 32-byte record):
 *     calcMessageFVFusedKernel, causal_bp_seq.cu:149-459 -- the PARALL
 *     f->v arithmetic (fv_msg) with the sequential scale(), then
 *     update_message (34-83): the old message m_fv[a + t] divided out of
 *     i's product, the new one multiplied in (zeros counted, counts as u64
 *     that wrap like the reference's size_t), rescaled, and stored.
 *
 * A level is short on work and long on latency, so each lane's chain of
 * dependent memory accesses is kept at two: the item, then everything it
 * names (the item carries the factor's slot range, so the v->f messages do
 * not wait for the factor's offsets).  The stop test's read is issued
 * beside the item load.  Every kernel boundary empties the L2, so each
 * random access costs a line from the fabric: a variable's product and
 * zero counts share one 32-byte record, and so do a factor's p, q and head
 * mask, whose mask is (1, 1) for an unclamped factor -- multiplying by 1
 * is exact and the zero-mass fallbacks test the mask for 0, so applying
 * the mask always gives the unclamped codes' results bit for bit.
 *
 *   group item (a factor of at most kSeqGroupDeg slots in one wave, with
 *     the slots it updates and the slots whose v->f message it recomputes
 *     first): one lane per slot, the factor's lanes side by side in one
 *     hardware wave.  A recomputing lane derives its slot's v->f message
 *     from the variable's state as loaded and stores it to m_vf; the other
 *     lanes load theirs from m_vf; the lanes then exchange the messages
 *     (ds_bpermute) and each updating lane runs its f->v update.  Every
 *     state read of the item is issued before the exchange and every state
 *     write after it, so the item's reads see the states from before its
 *     own updates, as the wave's v->f kernel comes before its f->v kernel.
 *
 * Items of one level touch disjoint outputs and read nothing another item
 * of the level writes, so lanes need no ordering among themselves.
 *
 * seq_tail: the last levels, each small, in one single-workgroup launch
 * with a workgroup barrier between levels (a barrier makes the
 * workgroup's global stores visible to the whole workgroup), instead of
 * one launch each.
 *
 * seq_beliefs: getBeliefsVKernel x2 + norm_functor + dist_functor + max
 * (causal_bp.cu:483-515, 878-889) from the maintained products, folded into
 * the sweep's max-diff slots like the PARALL variable pass.
 */
#ifndef FLBP_SEQ_KERNELS_H
#define FLBP_SEQ_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "staged_kernels.h"
#include "sweep_kernels.h"

#include "flbp_internal.h"
#include "plan_internal.h"

namespace flbp {
namespace dev {

constexpr int kSeqRegDeg = 8;       /* factors up to this degree keep messages in registers */
constexpr int kSeqTailBlock = 1024; /* threads of the tail kernel (four waves per SIMD: 128 VGPRs each) */
constexpr int kSeqTailLevels = 256; /* level offsets the tail kernel keeps in LDS */
constexpr int kSeqLevelBlock = 64;  /* threads per block of a level launch (one wave) */
constexpr int kSeqTailPer = 1;      /* items per thread and level in the tail kernel */

/* a group lane's third word: degree (0 = padding lane) bits 0-2, factor
 * type 3-4, head slot 5-7, updated-slot mask 8-13, recomputed-slot mask
 * 14-19, the lane's slot 20-22 */
__host__ __device__ inline uint32_t group_word(uint32_t d, uint32_t type, uint32_t h, uint32_t tmask, uint32_t imask,
                                               uint32_t k) {
    return d | (type << 3) | (h << 5) | (tmask << 8) | (imask << 14) | (k << 20);
}

}  // namespace dev
}  // namespace flbp

namespace {

using namespace flbp::dev;

/* calcMessageVFKernel's message (no (0, 0) guard) */
__device__ __forceinline__ double2 vf_msg_seq(const VarState &st, double2 m) {
    double prod_0 = st.p0, prod_1 = st.p1;
    const unsigned long long z0 = st.z0, z1 = st.z1;
    if ((z0 >= 2) || (z0 == 1 && m.x != 0)) {
        prod_0 = 0.0;
    } else if (m.x != 0) {
        prod_0 = prod_0 / m.x;
    }
    if ((z1 >= 2) || (z1 == 1 && m.y != 0)) {
        prod_1 = 0.0;
    } else if (m.y != 0) {
        prod_1 = prod_1 / m.y;
    }
    const double s = prod_0 + prod_1;
    double2 r;
    r.x = prod_0 / s;
    r.y = prod_1 / s;
    return r;
}

struct SeqArgs {
    const uint4 *__restrict__ vf_items;   /* slot, variable */
    const uint4 *__restrict__ fv_items;   /* var-major position, variable, factor, first slot */
    const uint4 *__restrict__ fv_aux;     /* degree | type << 24, target slot, head slot */
    const uint4 *__restrict__ grp_items;  /* one record per group lane (see seq_group_lane) */
    const FacRec *__restrict__ fac;       /* per-factor p, q, head mask */
    double2 *__restrict__ m_vf;           /* factor-major */
    double2 *__restrict__ m_fv;           /* factor-major */
    VarState *__restrict__ state;
};

__device__ __forceinline__ void seq_vf_item(const SeqArgs &A, const uint4 it) {
    const double2 m = A.m_fv[it.x];
    const VarState st = A.state[it.y];
    A.m_vf[it.x] = vf_msg_seq(st, m);
}

__device__ __forceinline__ void seq_fv_item(const SeqArgs &A, const uint4 it, const uint4 aux) {
    const uint32_t I = it.z, v = it.y, a = it.w;
    const int d = (int)(aux.x & 0xffffffu), t = (int)aux.y, h = (int)aux.z;
    const uint32_t type = aux.x >> 24;
    /* everything the item names, all independent of each other */
    RegIn<kSeqRegDeg> in;
#pragma unroll
    for (int j = 0; j < kSeqRegDeg; j++) in.m[j] = j < d ? A.m_vf[a + j] : make_double2(0.0, 0.0);
    const FacRec f = A.fac[I];
    const double2 old = A.m_fv[a + t];
    VarState st = A.state[v];
    double2 r;
    if (type == 0) {
        r.x = 1 - f.p;
        r.y = f.p;
    } else if (d <= kSeqRegDeg) {
        r = fv_msg<kSeqRegDeg, true>(in, type, true, d, t, h, in.pick(h), in.pick(h == 0 ? 1 : 0), f.p, f.q,
                                     f.m0, f.m1);
    } else {
        const ContigIn gin{A.m_vf + a};
        r = fv_msg<0, true>(gin, type, true, d, t, h, gin.pick(h), gin.pick(h == 0 ? 1 : 0), f.p, f.q, f.m0,
                            f.m1);
    }
    /* update_message */
    if (old.x == 0) st.z0--; else st.p0 /= old.x;
    if (r.x == 0) st.z0++; else st.p0 *= r.x;
    if (old.y == 0) st.z1--; else st.p1 /= old.y;
    if (r.y == 0) st.z1++; else st.p1 *= r.y;
    scale_seq(st.p0, st.p1);
    A.state[v] = st;
    A.m_fv[a + t] = r;
}

constexpr int kG = (int)flbp::kSeqGroupDeg;

/* The messages of the group's slots, slot j held by lane base + j of the
 * wave: every lane of the wave takes part (ds_bpermute reads other lanes'
 * registers), indices past the group read some other lane's value, which
 * the f->v code never uses (it reads slots below the degree only). */
__device__ __forceinline__ RegIn<kG> group_messages(double2 m, int base) {
    RegIn<kG> in;
#pragma unroll
    for (int j = 0; j < kG; j++) {
        in.m[j].x = __shfl(m.x, base + j, 64);
        in.m[j].y = __shfl(m.y, base + j, 64);
    }
    return in;
}

/* One lane of a group item, slot k of factor I: w = (I, first slot a,
 * group_word(degree, type, head slot, updated mask, recomputed mask, k),
 * variable of slot k); degree 0 marks a padding lane.  The lanes of one
 * group sit side by side in one wave and all lanes of the wave run this
 * together.  Every message the group touches lies in the factor's slot
 * range (both message arrays are factor-major), so the variable's state is
 * the lane's only scattered access. */
__device__ __forceinline__ void seq_group_lane(const SeqArgs &A, const uint4 w) {
    const uint32_t I = w.x, a = w.y, g = w.z;
    const int d = (int)(g & 7u), h = (int)((g >> 5) & 7u), k = (int)((g >> 20) & 7u);
    const uint32_t type = (g >> 3) & 3u;
    const bool tk = (g >> (8 + k)) & 1u, ik = (g >> (14 + k)) & 1u;   /* masks are 0 on padding */
    /* everything the lane names, independent of each other */
    double2 m = make_double2(0.0, 0.0), old = make_double2(0.0, 0.0);
    VarState st{0.0, 0.0, 0ull, 0ull};
    FacRec f{0.0, 0.0, 0.0, 0.0};
    /* a lone target's own v->f message is not an input of its update (the
     * schedule's levels do not order it against writers of that slot) */
    const uint32_t tmask = (g >> 8) & 63u;
    if (d > 0 && !ik && !(tk && (tmask & (tmask - 1u)) == 0u)) m = A.m_vf[a + k];
    if (tk || ik) {
        st = A.state[w.w];
        old = A.m_fv[a + k];
    }
    if (tk) f = A.fac[I];
    /* slot k's v->f message of this wave, from the state before the
     * factor's own updates */
    if (ik) {
        m = vf_msg_seq(st, old);
        A.m_vf[a + k] = m;
    }
    const RegIn<kG> in = group_messages(m, (int)(threadIdx.x & 63u) - k);
    if (!tk) return;
    double2 r;
    if (type == 0) {
        r.x = 1 - f.p;
        r.y = f.p;
    } else {
        r = fv_msg<kG, true>(in, type, true, d, k, h, in.pick(h), in.pick(h == 0 ? 1 : 0), f.p, f.q, f.m0, f.m1);
    }
    if (old.x == 0) st.z0--; else st.p0 /= old.x;
    if (r.x == 0) st.z0++; else st.p0 *= r.x;
    if (old.y == 0) st.z1--; else st.p1 /= old.y;
    if (r.y == 0) st.z1++; else st.p1 *= r.y;
    scale_seq(st.p0, st.p1);
    A.state[w.w] = st;
    A.m_fv[a + k] = r;
}

__global__ __launch_bounds__(kSeqLevelBlock) __attribute__((unused)) void seq_level(
    SeqArgs A, uint32_t n_vf, uint32_t n_fv, uint32_t n_grp, uint32_t vf_blocks, uint32_t fv_blocks,
    const unsigned long long *__restrict__ ctl, int k, double tol) {
    if (blockIdx.x < vf_blocks) {
        const uint32_t x = blockIdx.x * kSeqLevelBlock + threadIdx.x;
        const uint4 it = x < n_vf ? A.vf_items[x] : make_uint4(0u, 0u, 0u, 0u);
        if (chunk_stopped(ctl, k, tol) || x >= n_vf) return;
        seq_vf_item(A, it);
        return;
    }
    if (blockIdx.x < vf_blocks + fv_blocks) {
        const uint32_t x = (blockIdx.x - vf_blocks) * kSeqLevelBlock + threadIdx.x;
        const uint4 it = x < n_fv ? A.fv_items[x] : make_uint4(0u, 0u, 0u, 0u);
        const uint4 aux = x < n_fv ? A.fv_aux[x] : make_uint4(0u, 0u, 0u, 0u);
        if (chunk_stopped(ctl, k, tol) || x >= n_fv) return;
        seq_fv_item(A, it, aux);
        return;
    }
    /* group lanes: n_grp is a multiple of 64, so a wave is wholly inside
     * or wholly past the level and every lane of a running wave reaches
     * the exchange */
    const uint32_t x = (blockIdx.x - vf_blocks - fv_blocks) * kSeqLevelBlock + threadIdx.x;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    const uint4 w = x < n_grp ? A.grp_items[x] : z;
    if (chunk_stopped(ctl, k, tol) || x >= n_grp) return;
    seq_group_lane(A, w);
}

/* one item of a tail level: x indexes the level's group lanes (a multiple
 * of 64, so each wave holds group lanes only or none), then its v->f items,
 * then its f->v items; kind 0 = none, 1 = v->f, 2 = f->v, 3 = group lane */
struct TailItem {
    uint4 w[2];
    uint32_t kind;
};

__device__ __forceinline__ TailItem tail_none() {
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    return TailItem{{z, z}, 0u};
}

__device__ __forceinline__ TailItem tail_fetch(const SeqArgs &A, const uint32_t *s_vf, const uint32_t *s_fv,
                                               const uint32_t *s_grp, uint32_t l, uint32_t x) {
    TailItem r = tail_none();
    const uint32_t nv = s_vf[l + 1] - s_vf[l], nf = s_fv[l + 1] - s_fv[l], ng = s_grp[l + 1] - s_grp[l];
    if (x < ng) {
        r.w[0] = A.grp_items[s_grp[l] + x];
        r.kind = 3u;
    } else if (x - ng < nv) {
        r.w[0] = A.vf_items[s_vf[l] + x - ng];
        r.kind = 1u;
    } else if (x - ng - nv < nf) {
        r.w[0] = A.fv_items[s_fv[l] + x - ng - nv];
        r.w[1] = A.fv_aux[s_fv[l] + x - ng - nv];
        r.kind = 2u;
    }
    return r;
}

__device__ __forceinline__ void tail_run(const SeqArgs &A, const TailItem &t) {
    if (t.kind == 3u) seq_group_lane(A, t.w[0]);
    else if (t.kind == 1u) seq_vf_item(A, t.w[0]);
    else if (t.kind == 2u) seq_fv_item(A, t.w[0], t.w[1]);
}

/* levels [l0, l1) by one workgroup, at most kSeqTailPer * kSeqTailBlock
 * items each; lvl_vf / lvl_fv / lvl_grp hold each level's first item of
 * each kind, l1 - l0 <= kSeqTailLevels.  Items are read-only, so a level's
 * items are loaded while the level before it computes. */
__global__ __launch_bounds__(kSeqTailBlock) __attribute__((unused)) void seq_tail(
    SeqArgs A, const uint32_t *__restrict__ lvl_vf, const uint32_t *__restrict__ lvl_fv,
    const uint32_t *__restrict__ lvl_grp, uint32_t l0, uint32_t l1, const unsigned long long *__restrict__ ctl,
    int k, double tol) {
    __shared__ uint32_t s_vf[kSeqTailLevels + 1], s_fv[kSeqTailLevels + 1], s_grp[kSeqTailLevels + 1];
    const uint32_t n = l1 - l0;
    for (uint32_t i = threadIdx.x; i <= n; i += kSeqTailBlock) {
        s_vf[i] = lvl_vf[l0 + i];
        s_fv[i] = lvl_fv[l0 + i];
        s_grp[i] = lvl_grp[l0 + i];
    }
    if (chunk_stopped(ctl, k, tol)) return;   /* the same answer in every wave */
    __syncthreads();
    TailItem cur[kSeqTailPer];
#pragma unroll
    for (int j = 0; j < kSeqTailPer; j++)
        cur[j] = tail_fetch(A, s_vf, s_fv, s_grp, 0, threadIdx.x + j * kSeqTailBlock);
    for (uint32_t l = 0; l < n; l++) {
        TailItem next[kSeqTailPer];
#pragma unroll
        for (int j = 0; j < kSeqTailPer; j++)
            next[j] = l + 1 < n ? tail_fetch(A, s_vf, s_fv, s_grp, l + 1, threadIdx.x + j * kSeqTailBlock)
                                : tail_none();
#pragma unroll
        for (int j = 0; j < kSeqTailPer; j++) tail_run(A, cur[j]);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kSeqTailPer; j++) cur[j] = next[j];
    }
}

/* the factor records from the plan's arrays */
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void seq_fac_build(
    const uint32_t *__restrict__ fac_meta, const double *__restrict__ fac_p, const double *__restrict__ fac_q,
    const double2 *__restrict__ fac_mask, FacRec *__restrict__ out, uint32_t F) {
    const uint32_t I = blockIdx.x * kBlock + threadIdx.x;
    if (I >= F) return;
    const uint32_t meta = fac_meta[I];
    FacRec r{fac_p[I], (meta & 3u) ? fac_q[I] : 0.0, 1.0, 1.0};
    if (meta & 4u) {
        const double2 mk = fac_mask[I];
        r.m0 = mk.x;
        r.m1 = mk.y;
    }
    out[I] = r;
}

__global__ __launch_bounds__(kBlock) __attribute__((unused)) void seq_beliefs(
    const VarState *__restrict__ state, double *__restrict__ belief, uint32_t V, unsigned long long *__restrict__ ctl,
    int k, double tol) {
    if (chunk_stopped(ctl, k, tol)) return;
    __shared__ unsigned long long red[kBlock / 64];
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    unsigned long long bits = 0;
    if (v < V) {
        const VarState st = state[v];
        const double b0 = st.p0 * (double)(st.z0 == 0);
        const double b1 = st.p1 * (double)(st.z1 == 0);
        const double b = b1 / (b0 + b1);
        const double diff = fabs(belief[v] - b);
        belief[v] = b;
        bits = (unsigned long long)__double_as_longlong(diff);
    }
    block_fold_diff(ctl, k, bits, red);
}

}  // namespace

#endif
