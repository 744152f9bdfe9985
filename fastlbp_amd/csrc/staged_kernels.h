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
 *   2. runs, one lane per factor/variable, the reference-order arithmetic
 *      of sweep_kernels.h on its slots in LDS, writing the outputs back into
 *      LDS;
 *   3. writes the slot range back to HBM with coalesced stores.
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

/* kPull: v->f messages stored var-major, gathered through fac_vpos; else
 * stored factor-major (read contiguously) */
template <bool kDamp, int N, bool kPull>
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void factor_pass_staged(
    const uint32_t *__restrict__ blk, const uint32_t *__restrict__ fac_off,
    const uint32_t *__restrict__ fac_meta, const double *__restrict__ fac_p,
    const double *__restrict__ fac_q, const double2 *__restrict__ fac_mask,
    const uint32_t *__restrict__ fac_vpos, const double2 *__restrict__ m_vf, const double2 *m_fv_old,
    double2 *m_fv, double damp, double keep, const unsigned long long *__restrict__ ctl, int k,
    double tol) {
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
        const uint32_t type = meta & 3u;
        const bool clamped = (meta & 4u) != 0;
        const int h = (int)(meta >> 8);
        const double p = fac_p[f0], q = type ? fac_q[f0] : 0.0;
        double mask0 = 1.0, mask1 = 1.0;
        if (clamped) {
            const double2 mk = fac_mask[f0];
            mask0 = mk.x;
            mask1 = mk.y;
        }
        if (!kPull) {
            LdsIn in{m_vf + base};   /* plain pointer read; works on global memory too */
            const double2 mh = in.pick(h), mo = in.pick(h == 0 ? 1 : 0);
            for (uint32_t t = tid; t < n; t += kBlock) {
                double2 r = fv_msg<0>(in, type, clamped, (int)n, (int)t, h, mh, mo, p, q, mask0, mask1);
                if (kDamp) {
                    const double2 old = m_fv_old[base + t];
                    r.x = keep * r.x + damp * old.x;
                    r.y = keep * r.y + damp * old.y;
                }
                m_fv[base + t] = r;
            }
            return;
        }
        GlobalIn in{m_vf, fac_vpos + base};
        const double2 mh = in.pick(h), mo = in.pick(h == 0 ? 1 : 0);
        for (uint32_t t = tid; t < n; t += kBlock) {
            double2 r = fv_msg<0>(in, type, clamped, (int)n, (int)t, h, mh, mo, p, q, mask0, mask1);
            if (kDamp) {
                const double2 old = m_fv_old[base + t];
                r.x = keep * r.x + damp * old.x;
                r.y = keep * r.y + damp * old.y;
            }
            m_fv[base + t] = r;
        }
        return;
    }
    for (uint32_t e = tid; e < n; e += kBlock) {
        s_in[e] = kPull ? m_vf[fac_vpos[base + e]] : m_vf[base + e];
        if (kDamp) s_out[e] = m_fv_old[base + e];
    }
    __syncthreads();
    const uint32_t f = f0 + tid;
    if (f < f1) {
        const uint32_t s = fac_off[f] - base;
        const int d = (int)(fac_off[f + 1] - base - s);
        const uint32_t meta = fac_meta[f];
        const uint32_t type = meta & 3u;
        const bool clamped = (meta & 4u) != 0;
        const int h = (int)(meta >> 8);
        const double p = fac_p[f];
        const double q = type ? fac_q[f] : 0.0;
        double mask0 = 1.0, mask1 = 1.0;
        if (clamped) {
            const double2 mk = fac_mask[f];
            mask0 = mk.x;
            mask1 = mk.y;
        }
        auto store = [&](int t, double2 r) __attribute__((always_inline)) {
            if (kDamp) {
                const double2 old = s_out[s + t];
                r.x = keep * r.x + damp * old.x;
                r.y = keep * r.y + damp * old.y;
            }
            s_out[s + t] = r;
        };
        if (d <= N) {
            RegIn<N> in;
#pragma unroll
            for (int j = 0; j < N; j++) in.m[j] = j < d ? s_in[s + j] : make_double2(0.0, 0.0);
            const double2 mh = in.pick(h), mo = in.pick(h == 0 ? 1 : 0);
            for (int t = 0; t < d; t++)
                store(t, fv_msg<N>(in, type, clamped, d, t, h, mh, mo, p, q, mask0, mask1));
        } else {
            LdsIn in{s_in + s};
            const double2 mh = in.pick(h), mo = in.pick(h == 0 ? 1 : 0);
            for (int t = 0; t < d; t++)
                store(t, fv_msg<0>(in, type, clamped, d, t, h, mh, mo, p, q, mask0, mask1));
        }
    }
    __syncthreads();
    for (uint32_t e = tid; e < n; e += kBlock) m_fv[base + e] = s_out[e];
}

/* kPull: write v->f messages var-major (contiguous); else scatter them to
 * their factor-major slots */
template <bool kPull>
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void var_pass_staged(
    const uint32_t *__restrict__ blk, const uint32_t *__restrict__ var_off,
    const uint32_t *__restrict__ var_edge, const double2 *__restrict__ m_fv, double2 *__restrict__ m_vf,
    double *__restrict__ belief, unsigned long long *__restrict__ ctl, int k, double tol) {
    if (chunk_stopped(ctl, k, tol)) return;
    __shared__ double2 s_m[kStage];
    __shared__ unsigned long long red[kBlock / 64];
    __shared__ double s_prod[2];
    __shared__ uint32_t s_z[2];
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
            for (uint32_t e = tid; e < m; e += kBlock) s_m[e] = m_fv[var_edge[base + c + e]];
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
            s_prod[0] = prod_0;
            s_prod[1] = prod_1;
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
        prod_0 = s_prod[0];
        prod_1 = s_prod[1];
        z0 = s_z[0];
        z1 = s_z[1];
        for (uint32_t e = tid; e < n; e += kBlock) {
            const uint32_t sl = var_edge[base + e];
            m_vf[kPull ? base + e : sl] = vf_msg(prod_0, prod_1, z0, z1, m_fv[sl]);
        }
        block_fold_diff(ctl, k, diff_bits, red);
        return;
    }
    for (uint32_t e = tid; e < n; e += kBlock) s_m[e] = m_fv[var_edge[base + e]];
    __syncthreads();
    const uint32_t v = v0 + tid;
    if (v < v1) {
        const uint32_t s = var_off[v] - base;
        const int d = (int)(var_off[v + 1] - base - s);
        double prod_0 = 1.0, prod_1 = 1.0;
        uint32_t z0 = 0, z1 = 0;
        for (int j = 0; j < d; j++) {
            const double2 x = s_m[s + j];
            if (x.x == 0) z0++; else prod_0 *= x.x;
            if (x.y == 0) z1++; else prod_1 *= x.y;
            scale_pair(prod_0, prod_1);
        }
        const double b0 = prod_0 * (double)(z0 == 0);
        const double b1 = prod_1 * (double)(z1 == 0);
        const double b = b1 / (b0 + b1);
        const double diff = fabs(belief[v] - b);
        belief[v] = b;
        diff_bits = (unsigned long long)__double_as_longlong(diff);
        for (int j = 0; j < d; j++) s_m[s + j] = vf_msg(prod_0, prod_1, z0, z1, s_m[s + j]);
    }
    __syncthreads();
    for (uint32_t e = tid; e < n; e += kBlock) m_vf[kPull ? base + e : var_edge[base + e]] = s_m[e];
    block_fold_diff(ctl, k, diff_bits, red);
}

}  // namespace

#endif
