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

constexpr int kStage = 1024;    /* slots staged per block (16 KiB per message array) */
constexpr int kPerLane = kStage / kBlock;   /* staged slots per lane */

struct LdsIn {
    const double2 *m;
    __device__ __forceinline__ double2 get(int j) const { return m[j]; }
    __device__ __forceinline__ double2 pick(int j) const { return m[j]; }
};

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

template <bool kDamp, int N, int kLayout>
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void factor_pass_staged(
    const uint32_t *__restrict__ blk, const uint32_t *__restrict__ fac_off,
    const uint32_t *__restrict__ fac_meta, const double *__restrict__ fac_p,
    const double *__restrict__ fac_q, const double2 *__restrict__ fac_mask,
    const uint32_t *__restrict__ fac_vpos, const uint32_t *__restrict__ fac_work,
    const double2 *__restrict__ m_vf, const double2 *m_fv_old, double2 *m_fv, double2 *__restrict__ m_fv_vm,
    double damp, double keep, const unsigned long long *__restrict__ ctl, int k, double tol) {
    if (chunk_stopped(ctl, k, tol)) return;
    __shared__ double2 s_in[kStage];
    __shared__ double2 s_out[kStage];
    const uint32_t f0 = blk[blockIdx.x], f1 = blk[blockIdx.x + 1];
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
    /* phase 1: the block's work entries into registers, its factors'
     * offsets and parameters into LDS, and its messages into LDS, all loads
     * independent of each other */
    __shared__ uint32_t s_start[kBlock + 1];   /* factor offsets relative to base */
    __shared__ uint32_t s_meta[kBlock];
    __shared__ double s_p[kBlock], s_q[kBlock];
    const uint32_t nf = f1 - f0;
    uint32_t w[kPerLane];
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const uint32_t i = tid + j * kBlock;
        w[j] = i < n ? fac_work[base + i] : 0u;
    }
    if ((uint32_t)tid < nf) {
        s_start[tid] = fac_off[f0 + tid] - base;
        s_meta[tid] = fac_meta[f0 + tid];
        s_p[tid] = fac_p[f0 + tid];
        s_q[tid] = fac_q[f0 + tid];
    }
    if (tid == 0) s_start[nf] = n;
    for (uint32_t e = tid; e < n; e += kBlock) {
        s_in[e] = kLayout == kLayoutPull ? m_vf[fac_vpos[base + e]] : m_vf[base + e];
        if (kDamp) s_out[e] = m_fv_old[base + e];
    }
    __syncthreads();
    /* phase 2: one lane per output slot, in the block's work order
     * (w = local slot | local factor << 10), from LDS only */
#pragma unroll 1
    for (int j = 0; j < kPerLane; j++) {
        if (tid + j * kBlock >= n) break;
        uint32_t wj = 0;
#pragma unroll
        for (int x = 0; x < kPerLane; x++) wj |= (j == x) ? w[x] : 0u;
        const uint32_t sl = wj & 1023u, lf = wj >> 10;
        const uint32_t s = s_start[lf];
        const int d = (int)(s_start[lf + 1] - s);
        const int t = (int)(sl - s);
        const uint32_t meta = s_meta[lf];
        const uint32_t type = meta & 3u;
        const bool clamped = (meta & 4u) != 0;
        const int h = (int)(meta >> 8);
        const double p = s_p[lf];
        const double q = type ? s_q[lf] : 0.0;
        double mask0 = 1.0, mask1 = 1.0;
        if (clamped) {
            const double2 mk = fac_mask[f0 + lf];
            mask0 = mk.x;
            mask1 = mk.y;
        }
        double2 r;
        if (d <= N) {
            RegIn<N> in;
#pragma unroll
            for (int jj = 0; jj < N; jj++) in.m[jj] = jj < d ? s_in[s + jj] : make_double2(0.0, 0.0);
            const double2 mh = in.pick(h), mo = in.pick(h == 0 ? 1 : 0);
            r = fv_msg<N>(in, type, clamped, d, t, h, mh, mo, p, q, mask0, mask1);
        } else {
            LdsIn in{s_in + s};
            const double2 mh = in.pick(h), mo = in.pick(h == 0 ? 1 : 0);
            r = fv_msg<0>(in, type, clamped, d, t, h, mh, mo, p, q, mask0, mask1);
        }
        if (kDamp) {
            const double2 old = s_out[sl];
            r.x = keep * r.x + damp * old.x;
            r.y = keep * r.y + damp * old.y;
        }
        s_out[sl] = r;
    }
    __syncthreads();
    for (uint32_t e = tid; e < n; e += kBlock) put_fv<kDamp, kLayout>(m_fv, m_fv_vm, fac_vpos, base + e, s_out[e]);
}

/* f->v messages read var-major (scatter layout) or gathered from their
 * factor slots; v->f messages stored var-major (pull layout) or scattered
 * to their factor slots */
template <int kLayout>
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void var_pass_staged(
    const uint32_t *__restrict__ blk, const uint32_t *__restrict__ var_off,
    const uint32_t *__restrict__ var_edge, const uint16_t *__restrict__ var_work,
    const double2 *__restrict__ m_fv, double2 *__restrict__ m_vf, double *__restrict__ belief,
    unsigned long long *__restrict__ ctl, int k, double tol) {
    if (chunk_stopped(ctl, k, tol)) return;
    __shared__ double2 s_m[kStage];
    __shared__ unsigned long long red[kBlock / 64];
    __shared__ double s_prod[2][kBlock];
    __shared__ uint32_t s_z[kBlock];          /* z0 | z1 << 16 (a staged variable has <= kStage edges) */
    __shared__ uint8_t s_owner[kStage];       /* local variable of each staged slot */
    const uint32_t v0 = blk[blockIdx.x], v1 = blk[blockIdx.x + 1];
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
                s_m[e] = kLayout == kLayoutScatter ? m_fv[base + c + e] : m_fv[var_edge[base + c + e]];
            __syncthreads();
            if (tid == 0) {
                for (uint32_t e = 0; e < m; e++) {
                    const double2 x = s_m[e];
                    if (x.x == 0) z0++; else prod_0 *= x.x;
                    if (x.y == 0) z1++; else prod_1 *= x.y;
                    scale_pair(prod_0, prod_1);
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            s_prod[0][0] = prod_0;
            s_prod[1][0] = prod_1;
            s_z[0] = z0;
            s_z[1] = z1;
            const double b0 = prod_0 * (double)(z0 == 0);
            const double b1 = prod_1 * (double)(z1 == 0);
            const double b = b1 / (b0 + b1);
            const double diff = fabs(belief[v0] - b);
            belief[v0] = b;
            diff_bits = (unsigned long long)__double_as_longlong(diff);
        }
        __syncthreads();
        prod_0 = s_prod[0][0];
        prod_1 = s_prod[1][0];
        z0 = s_z[0];
        z1 = s_z[1];
        for (uint32_t e = tid; e < n; e += kBlock) {
            const uint32_t sl = var_edge[base + e];
            const double2 x = kLayout == kLayoutScatter ? m_fv[base + e] : m_fv[sl];
            m_vf[kLayout == kLayoutPull ? base + e : sl] = vf_msg(prod_0, prod_1, z0, z1, x);
        }
        block_fold_diff(ctl, k, diff_bits, red);
        return;
    }
    /* phase 1: offsets, old beliefs and the work entry of each variable,
     * and the block's messages, all loads independent of each other */
    __shared__ uint32_t s_start[kBlock + 1];
    __shared__ double s_bel[kBlock];
    const uint32_t nv = v1 - v0;
    uint32_t lv = 0;
    if ((uint32_t)tid < nv) {
        lv = var_work[v0 + tid];
        s_start[tid] = var_off[v0 + tid] - base;
        s_bel[tid] = belief[v0 + tid];
    }
    if (tid == 0) s_start[nv] = n;
    for (uint32_t e = tid; e < n; e += kBlock)
        s_m[e] = kLayout == kLayoutScatter ? m_fv[base + e] : m_fv[var_edge[base + e]];
    __syncthreads();
    /* phase 2: products, one lane per variable in the block's work order
     * (by degree), from LDS only */
    if ((uint32_t)tid < nv) {
        const uint32_t s = s_start[lv];
        const int d = (int)(s_start[lv + 1] - s);
        double prod_0 = 1.0, prod_1 = 1.0;
        uint32_t z0 = 0, z1 = 0;
        for (int j = 0; j < d; j++) {
            const double2 x = s_m[s + j];
            if (x.x == 0) z0++; else prod_0 *= x.x;
            if (x.y == 0) z1++; else prod_1 *= x.y;
            scale_pair(prod_0, prod_1);
            s_owner[s + j] = (uint8_t)lv;
        }
        const double b0 = prod_0 * (double)(z0 == 0);
        const double b1 = prod_1 * (double)(z1 == 0);
        const double b = b1 / (b0 + b1);
        const double diff = fabs(s_bel[lv] - b);
        s_bel[lv] = b;
        diff_bits = (unsigned long long)__double_as_longlong(diff);
        s_prod[0][lv] = prod_0;
        s_prod[1][lv] = prod_1;
        s_z[lv] = z0 | (z1 << 16);
    }
    __syncthreads();
    /* leave-one-out messages: one lane per slot, stored coalesced (var-major)
     * or scattered to the factor slots (push layout) */
    for (uint32_t e = tid; e < n; e += kBlock) {
        const uint32_t lv = s_owner[e];
        const uint32_t z = s_z[lv];
        const double2 r = vf_msg(s_prod[0][lv], s_prod[1][lv], z & 0xffffu, z >> 16, s_m[e]);
        m_vf[kLayout == kLayoutPull ? base + e : var_edge[base + e]] = r;
    }
    if ((uint32_t)tid < nv) belief[v0 + tid] = s_bel[tid];
    block_fold_diff(ctl, k, diff_bits, red);
}

}  // namespace

#endif
