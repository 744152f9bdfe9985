/*
 * seq_kernels.h -- device code of the sequential schedules (SEQFIX / TOPO).
 *
 * One launch runs every item of one dependency level (seq_host.cpp):
 * blocks [0, vf_blocks) take its v->f items, the rest its f->v items.
 *
 *   v->f item (slot s, var-major position k of the same edge, variable j):
 *     calcMessageVFKernel, csrc/kernel/causal_bp_seq.cu:113-147 -- the
 *     leave-one-out message from j's product and zero counts with the
 *     f->v message k divided out, normalized with no (0, 0) guard.
 *   f->v item (var-major position k, variable i, factor I, the factor's
 *     first slot a; degree d, type, target slot t and head slot h in a
 *     second word; p, q and the head mask in the factor's 32-byte record):
 *     calcMessageFVFusedKernel, causal_bp_seq.cu:149-459 -- the PARALL
 *     f->v arithmetic (fv_msg) with the sequential scale(), then
 *     update_message (34-83): the old message k divided out of i's
 *     product, the new one multiplied in (zeros counted, counts as u64 that
 *     wrap like the reference's size_t), rescaled, and stored.
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

#include "plan_internal.h"

namespace flbp {
namespace dev {

constexpr int kSeqRegDeg = 8;       /* factors up to this degree keep messages in registers */
constexpr int kSeqTailBlock = 1024; /* threads of the tail kernel */
constexpr int kSeqTailLevels = 256; /* level offsets the tail kernel keeps in LDS */
constexpr int kSeqTailPer = 2;      /* items per thread and level in the tail kernel */

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
    const uint4 *__restrict__ vf_items;   /* slot, var-major position, variable */
    const uint4 *__restrict__ fv_items;   /* var-major position, variable, factor, first slot */
    const uint4 *__restrict__ fv_aux;     /* degree | type << 24, target slot, head slot */
    const FacRec *__restrict__ fac;       /* per-factor p, q, head mask */
    double2 *__restrict__ m_vf;           /* factor-major */
    double2 *__restrict__ m_fv_vm;        /* var-major */
    VarState *__restrict__ state;
};

__device__ __forceinline__ void seq_vf_item(const SeqArgs &A, const uint4 it) {
    const double2 m = A.m_fv_vm[it.y];
    const VarState st = A.state[it.z];
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
    const double2 old = A.m_fv_vm[it.x];
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
    A.m_fv_vm[it.x] = r;
}

__global__ __launch_bounds__(kBlock) __attribute__((unused)) void seq_level(
    SeqArgs A, uint32_t n_vf, uint32_t n_fv, uint32_t vf_blocks, const unsigned long long *__restrict__ ctl, int k,
    double tol) {
    if (blockIdx.x < vf_blocks) {
        const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
        const uint4 it = x < n_vf ? A.vf_items[x] : make_uint4(0u, 0u, 0u, 0u);
        if (chunk_stopped(ctl, k, tol) || x >= n_vf) return;
        seq_vf_item(A, it);
        return;
    }
    const uint32_t x = (blockIdx.x - vf_blocks) * kBlock + threadIdx.x;
    const uint4 it = x < n_fv ? A.fv_items[x] : make_uint4(0u, 0u, 0u, 0u);
    const uint4 aux = x < n_fv ? A.fv_aux[x] : make_uint4(0u, 0u, 0u, 0u);
    if (chunk_stopped(ctl, k, tol) || x >= n_fv) return;
    seq_fv_item(A, it, aux);
}

/* one item of a tail level: x indexes the level's v->f items, then its
 * f->v items; kind 0 = none, 1 = v->f, 2 = f->v */
struct TailItem {
    uint4 it;
    uint4 aux;
    uint32_t kind;
};

__device__ __forceinline__ TailItem tail_fetch(const SeqArgs &A, const uint32_t *s_vf, const uint32_t *s_fv,
                                               uint32_t l, uint32_t x) {
    TailItem r{make_uint4(0u, 0u, 0u, 0u), make_uint4(0u, 0u, 0u, 0u), 0u};
    const uint32_t nv = s_vf[l + 1] - s_vf[l], nf = s_fv[l + 1] - s_fv[l];
    if (x < nv) {
        r.it = A.vf_items[s_vf[l] + x];
        r.kind = 1u;
    } else if (x - nv < nf) {
        r.it = A.fv_items[s_fv[l] + x - nv];
        r.aux = A.fv_aux[s_fv[l] + x - nv];
        r.kind = 2u;
    }
    return r;
}

__device__ __forceinline__ void tail_run(const SeqArgs &A, const TailItem &t) {
    if (t.kind == 1u) seq_vf_item(A, t.it);
    else if (t.kind == 2u) seq_fv_item(A, t.it, t.aux);
}

/* levels [l0, l1) by one workgroup, at most kSeqTailPer * kSeqTailBlock
 * items each; lvl_vf / lvl_fv hold each level's first item,
 * l1 - l0 <= kSeqTailLevels.  Items are read-only, so a level's items are
 * loaded while the level before it computes. */
__global__ __launch_bounds__(kSeqTailBlock) __attribute__((unused)) void seq_tail(
    SeqArgs A, const uint32_t *__restrict__ lvl_vf, const uint32_t *__restrict__ lvl_fv, uint32_t l0, uint32_t l1,
    const unsigned long long *__restrict__ ctl, int k, double tol) {
    __shared__ uint32_t s_vf[kSeqTailLevels + 1], s_fv[kSeqTailLevels + 1];
    const uint32_t n = l1 - l0;
    for (uint32_t i = threadIdx.x; i <= n; i += kSeqTailBlock) {
        s_vf[i] = lvl_vf[l0 + i];
        s_fv[i] = lvl_fv[l0 + i];
    }
    if (chunk_stopped(ctl, k, tol)) return;   /* the same answer in every wave */
    __syncthreads();
    TailItem cur[kSeqTailPer];
#pragma unroll
    for (int j = 0; j < kSeqTailPer; j++) cur[j] = tail_fetch(A, s_vf, s_fv, 0, threadIdx.x + j * kSeqTailBlock);
    for (uint32_t l = 0; l < n; l++) {
        TailItem next[kSeqTailPer];
#pragma unroll
        for (int j = 0; j < kSeqTailPer; j++)
            next[j] = l + 1 < n ? tail_fetch(A, s_vf, s_fv, l + 1, threadIdx.x + j * kSeqTailBlock)
                                : TailItem{make_uint4(0u, 0u, 0u, 0u), make_uint4(0u, 0u, 0u, 0u), 0u};
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
