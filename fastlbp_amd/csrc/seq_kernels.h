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
 *   f->v item (var-major position k, its slot s, factor I, variable i):
 *     calcMessageFVFusedKernel, causal_bp_seq.cu:149-459 -- the PARALL
 *     f->v arithmetic (fv_msg) with the sequential scale(), then
 *     update_message (34-83): the old message k divided out of i's
 *     product, the new one multiplied in (zeros counted, counts as u64 that
 *     wrap like the reference's size_t), rescaled, and stored.
 *
 * Items of one level touch disjoint outputs and read nothing another item
 * of the level writes, so lanes need no ordering among themselves.
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

namespace {

using namespace flbp::dev;

constexpr int kSeqRegDeg = 8;   /* factors up to this degree keep messages in registers */

/* calcMessageVFKernel's message (no (0, 0) guard) */
__device__ __forceinline__ double2 vf_msg_seq(double prod_0, double prod_1, unsigned long long z0,
                                              unsigned long long z1, double2 m) {
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

__global__ __launch_bounds__(kBlock) __attribute__((unused)) void seq_level(
    const uint4 *__restrict__ vf_items, uint32_t n_vf, const uint4 *__restrict__ fv_items, uint32_t n_fv,
    uint32_t vf_blocks, const uint32_t *__restrict__ fac_off, const uint32_t *__restrict__ fac_meta,
    const double *__restrict__ fac_p, const double *__restrict__ fac_q, const double2 *__restrict__ fac_mask,
    double2 *__restrict__ m_vf, double2 *__restrict__ m_fv_vm, double2 *__restrict__ prod,
    ulonglong2 *__restrict__ nz, const unsigned long long *__restrict__ ctl, int k, double tol) {
    if (chunk_stopped(ctl, k, tol)) return;
    if (blockIdx.x < vf_blocks) {
        const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
        if (x >= n_vf) return;
        const uint4 it = vf_items[x];   /* slot, var-major position, variable */
        const double2 m = m_fv_vm[it.y];
        const double2 pr = prod[it.z];
        const ulonglong2 z = nz[it.z];
        m_vf[it.x] = vf_msg_seq(pr.x, pr.y, z.x, z.y, m);
        return;
    }
    const uint32_t x = (blockIdx.x - vf_blocks) * kBlock + threadIdx.x;
    if (x >= n_fv) return;
    const uint4 it = fv_items[x];   /* var-major position, slot, factor, variable */
    const uint32_t I = it.z, v = it.w;
    const uint32_t a = fac_off[I];
    const int d = (int)(fac_off[I + 1] - a);
    const uint32_t meta = fac_meta[I];
    const uint32_t type = meta & 3u;
    const bool clamped = (meta & 4u) != 0;
    const int h = (int)(meta >> 8);
    const int t = (int)(it.y - a);
    const double p = fac_p[I];
    const double q = type ? fac_q[I] : 0.0;
    double mask0 = 1.0, mask1 = 1.0;
    if (clamped) {
        const double2 mk = fac_mask[I];
        mask0 = mk.x;
        mask1 = mk.y;
    }
    const double2 old = m_fv_vm[it.x];
    double2 pr = prod[v];
    ulonglong2 z = nz[v];
    double2 r;
    if (type == 0) {
        r.x = 1 - p;
        r.y = p;
    } else if (d <= kSeqRegDeg) {
        RegIn<kSeqRegDeg> in;
#pragma unroll
        for (int j = 0; j < kSeqRegDeg; j++) in.m[j] = j < d ? m_vf[a + j] : make_double2(0.0, 0.0);
        r = fv_msg<kSeqRegDeg, true>(in, type, clamped, d, t, h, in.pick(h), in.pick(h == 0 ? 1 : 0), p, q,
                                     mask0, mask1);
    } else {
        const ContigIn in{m_vf + a};
        r = fv_msg<0, true>(in, type, clamped, d, t, h, in.pick(h), in.pick(h == 0 ? 1 : 0), p, q, mask0,
                            mask1);
    }
    /* update_message */
    if (old.x == 0) z.x--; else pr.x /= old.x;
    if (r.x == 0) z.x++; else pr.x *= r.x;
    if (old.y == 0) z.y--; else pr.y /= old.y;
    if (r.y == 0) z.y++; else pr.y *= r.y;
    scale_seq(pr.x, pr.y);
    prod[v] = pr;
    nz[v] = z;
    m_fv_vm[it.x] = r;
}

__global__ __launch_bounds__(kBlock) __attribute__((unused)) void seq_beliefs(
    const double2 *__restrict__ prod, const ulonglong2 *__restrict__ nz, double *__restrict__ belief,
    uint32_t V, unsigned long long *__restrict__ ctl, int k, double tol) {
    if (chunk_stopped(ctl, k, tol)) return;
    __shared__ unsigned long long red[kBlock / 64];
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    unsigned long long bits = 0;
    if (v < V) {
        const double2 pr = prod[v];
        const ulonglong2 z = nz[v];
        const double b0 = pr.x * (double)(z.x == 0);
        const double b1 = pr.y * (double)(z.y == 0);
        const double b = b1 / (b0 + b1);
        const double diff = fabs(belief[v] - b);
        belief[v] = b;
        bits = (unsigned long long)__double_as_longlong(diff);
    }
    block_fold_diff(ctl, k, bits, red);
}

}  // namespace

#endif
