/*
 * sweep_kernels.h -- device code of one PARALL sweep (gfx950).
 *
 * One sweep of the reference (csrc/ext/causal_bp.cu:409-444 + 483-515 +
 * 879-886) is three per-edge/per-variable kernels, two belief kernels, a
 * thrust transform and a thrust reduce.  Here it is two kernels, a factor
 * pass and a variable pass.  This file holds what every form of them
 * shares -- the per-message arithmetic (fv_msg, vf_msg), the max-diff
 * slots and the device-side stop test -- and the one-lane-per-item kernels
 * of kernel variant 1:
 *
 *   factor_pass  -- one lane per factor: gathers the factor's v->f messages
 *                   (stored var-major) into registers through the factor's
 *                   slot -> var-major position map and writes every f->v
 *                   message of the factor contiguously (factor-major),
 *                   damping fused into the store.
 *   var_pass     -- one lane per variable: gathers its f->v messages through
 *                   the var-major -> factor-major slot map, forms the product
 *                   with zero counts, the belief and |b - b_old|, then writes
 *                   the leave-one-out v->f messages at its own var-major
 *                   positions; the max |b - b_old| is reduced in the wave
 *                   (DPP shuffles), across the block in LDS and across the
 *                   grid into line-padded partial slots.
 *
 * The LDS-staged kernels of the other variants are in staged_kernels.h.
 *
 * Every message is computed with the reference's own operation order (the
 * per-output loops of calcMessageFVFusedAllKernel, csrc/kernel/
 * causal_bp_parall.cu:94-398, and calcProdAllKernel/calcMessageVFAllKernel,
 * 20-92), so results are bit-identical to the reference's FP64 arithmetic.
 * The one rewrite is scale() (causal_bp_parall.cu:14-18), which divides both
 * values by their max: the max divided by itself is exactly 1 for any finite
 * nonzero value, so only the other value is divided (half the FP64 divides);
 * zero, infinite and NaN maxima keep the reference's NaN results.
 */
#ifndef FLBP_SWEEP_KERNELS_H
#define FLBP_SWEEP_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flbp {
namespace dev {

constexpr int kFacRegDeg = 8;     /* factors up to this degree run from registers */
constexpr int kFacRegDegSmall = 5; /* register slots when no factor is wider (more waves fit) */
constexpr int kVarRegDeg = 8;     /* variables up to this degree keep messages in registers */
constexpr int kBlock = 256;

/* factor meta word: [1:0] type, [2] head clamped, [31:8] head slot in factor */
__host__ __device__ inline uint32_t pack_meta(uint32_t type, bool clamped, uint32_t hslot) {
    return (type & 3u) | (clamped ? 4u : 0u) | (hslot << 8);
}

__device__ __forceinline__ void scale_pair(double &x, double &y) {
    const double m = fmax(x, y);
    const bool x_is_max = (x == m);
    const double r = (x_is_max ? y : x) / m;
    const double one = (m != 0.0 && __builtin_isfinite(m)) ? 1.0 : __builtin_nan("");
    x = x_is_max ? one : r;
    y = x_is_max ? r : one;
}

/* scale() of the sequential kernels, csrc/kernel/causal_bp_seq.cu:13-23:
 * the smaller value over the larger, the larger set to 1; equal (or NaN)
 * values both become 1.  Written with selects (the one quotient taken is
 * the reference's, smaller / larger); branches that assign x or y made the
 * compiler pick the target through a stack address. */
__device__ __forceinline__ void scale_seq(double &x, double &y) {
    const bool lt = x < y, gt = y < x;
    const double r = (lt ? x : y) / (lt ? y : x);
    x = lt ? r : 1.0;
    y = gt ? r : 1.0;
}

template <bool kSeq>
__device__ __forceinline__ void scale_any(double &x, double &y) {
    if constexpr (kSeq) scale_seq(x, y);
    else scale_pair(x, y);
}

/* Messages of one factor, either held in registers or read from global
 * memory.  get(j) is only called with indices that are compile-time
 * constants after unrolling; pick(j) fetches a run-time slot with an AND/OR
 * select over bit patterns, which the compiler cannot turn back into an
 * indexed load (that would push the array into scratch memory). */
template <int N>
struct RegIn {
    double2 m[N];
    __device__ __forceinline__ double2 get(int j) const { return m[j]; }
    __device__ __forceinline__ double2 pick(int j) const {
        unsigned long long x = 0, y = 0;
#pragma unroll
        for (int k = 0; k < N; k++) {
            const unsigned long long sel = 0ull - (unsigned long long)(j == k);
            x |= sel & (unsigned long long)__double_as_longlong(m[k].x);
            y |= sel & (unsigned long long)__double_as_longlong(m[k].y);
        }
        return make_double2(__longlong_as_double((long long)x), __longlong_as_double((long long)y));
    }
};
struct GlobalIn {
    const double2 *__restrict__ m;     /* message array */
    const uint32_t *__restrict__ pos;  /* position of slot j in it */
    __device__ __forceinline__ double2 get(int j) const { return m[pos[j]]; }
    __device__ __forceinline__ double2 pick(int j) const { return m[pos[j]]; }
};
struct ContigIn {
    const double2 *__restrict__ m;     /* the factor's messages, slot j at m[j] */
    __device__ __forceinline__ double2 get(int j) const { return m[j]; }
    __device__ __forceinline__ double2 pick(int j) const { return m[j]; }
};

/* Runs f(j) for j = 0..d-1 in order; unrolled when the degree bound N is a
 * compile-time register count, a plain loop when N == 0. */
template <int N, class F>
__device__ __forceinline__ void for_slots(int d, F &&f) {
    if constexpr (N > 0) {
#pragma unroll
        for (int j = 0; j < N; j++)
            if (j < d) f(j);
    } else {
        for (int j = 0; j < d; j++) f(j);
    }
}

/* The head-target message (codes 1-4) before masking: returns
 * u = p*X + q*R and w = (1-p)*X + (1-q)*R where X is the product of the
 * state the CPD singles out (state 1 for AND, state 0 for OR); AND gives
 * (marg0, marg1) = (w, u), OR gives (u, w).  causal_bp_parall.cu:113-143,
 * 152-182. */
template <int N, bool kSeq = false, class In>
__device__ __forceinline__ void head_uw(const In &in, int d, int t, double2 m_other, bool is_and,
                                        double p, double q, double &u, double &w) {
    if (d == 1) {
        u = p;
        w = 1 - p;
        return;
    }
    if (d == 2) {
        const double2 m = m_other;   /* slot (t == 0 ? 1 : 0) */
        const double x = is_and ? m.y : m.x, y = is_and ? m.x : m.y;
        u = p * x + q * y;
        w = (1 - p) * x + (1 - q) * y;
        return;
    }
    double px = 1.0, p10 = 1.0, e1 = 0;
    for_slots<N>(d, [&](int j) __attribute__((always_inline)) {
        if (j != t) {
            const double2 m = in.get(j);
            const double a0 = m.x + m.y;
            const double x = is_and ? m.y : m.x;
            const double delta = is_and ? m.x : m.y;
            px *= x;
            p10 *= a0;
            scale_any<kSeq>(px, p10);
            if (x != 0 && a0 == x && delta != 0) e1 += delta / x;
        }
    });
    const double R = e1 * p10 + (p10 - px);
    u = p * px + q * R;
    w = (1 - p) * px + (1 - q) * R;
}

/* The body-target message (codes 5-8): causal_bp_parall.cu:289-393.
 * Unclamped factors use masks (1, 1), which multiply exactly. */
template <int N, bool kSeq = false, class In>
__device__ __forceinline__ double2 body_msg(const In &in, int d, int t, int h, double2 mh,
                                            bool is_and, double p, double q, double mask0,
                                            double mask1) {
    const double hx = is_and ? mh.y : mh.x, hy = is_and ? mh.x : mh.y;
    const double mx = is_and ? mask1 : mask0, my = is_and ? mask0 : mask1;
    double P = (p - q) * (hx * mx - hy * my);
    double Q = q * hx * mx + (1 - q) * hy * my;
    scale_any<kSeq>(P, Q);
    if (Q != 0) {
        for_slots<N>(d, [&](int j) __attribute__((always_inline)) {
            if (j != t && j != h) {
                const double2 m = in.get(j);
                Q *= m.x + m.y;
                P *= is_and ? m.y : m.x;
                scale_any<kSeq>(P, Q);
            }
        });
    }
    const double QP = Q + P;
    const double marg = Q + QP;
    double2 r;
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

/* The f->v message to slot t of a factor (type, degree d, head slot h).
 * mh = message of the head slot, m_other = message of slot (h == 0 ? 1 : 0)
 * (the only other slot of a degree-2 factor). */
template <int N, bool kSeq = false, class In>
__device__ __forceinline__ double2 fv_msg(const In &in, uint32_t type, bool clamped, int d, int t,
                                          int h, double2 mh, double2 m_other, double p, double q,
                                          double mask0, double mask1) {
    double2 r;
    if (type == 0) {
        r.x = 1 - p;
        r.y = p;
        return r;
    }
    const bool is_and = (type == 1);
    if (t != h) return body_msg<N, kSeq>(in, d, t, h, mh, is_and, p, q, mask0, mask1);
    double u, w;
    head_uw<N, kSeq>(in, d, t, m_other, is_and, p, q, u, w);
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

/* leave-one-out v->f message, calcMessageVFAllKernel causal_bp_parall.cu:68-90 */
__device__ __forceinline__ double2 vf_msg(double prod_0, double prod_1, uint32_t z0, uint32_t z1,
                                          double2 m) {
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
    if (prod_0 == 0 && prod_1 == 0) prod_0 = prod_1 = 0.5;
    const double s = prod_0 + prod_1;
    double2 r;
    r.x = prod_0 / s;
    r.y = prod_1 / s;
    return r;
}

/* one step of the wave max: each lane takes the max of its value and the
 * value DPP control kCtrl routes to it (lanes without a source, or outside
 * the row and bank masks, keep their own) */
template <int kCtrl, int kRowMask>
__device__ __forceinline__ unsigned long long dpp_max_step(unsigned long long v) {
    const int lo = (int)(uint32_t)v, hi = (int)(uint32_t)(v >> 32);
    const uint32_t olo = (uint32_t)__builtin_amdgcn_update_dpp(lo, lo, kCtrl, kRowMask, 0xf, false);
    const uint32_t ohi = (uint32_t)__builtin_amdgcn_update_dpp(hi, hi, kCtrl, kRowMask, 0xf, false);
    const unsigned long long o = ((unsigned long long)ohi << 32) | olo;
    return o > v ? o : v;
}

/* max over the wave's 64 lanes, returned to every lane (the whole wave
 * active): row_shr 1, 2, 4, 8 leave each 16-lane row's max in its lane 15;
 * row_bcast 15 (rows 1 and 3) and row_bcast 31 (rows 2 and 3) carry those
 * into lane 63, which is then read.  Register moves only, no LDS round
 * trips as in a shuffle reduction. */
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    v = dpp_max_step<0x111, 0xf>(v);   /* row_shr:1 */
    v = dpp_max_step<0x112, 0xf>(v);   /* row_shr:2 */
    v = dpp_max_step<0x114, 0xf>(v);   /* row_shr:4 */
    v = dpp_max_step<0x118, 0xf>(v);   /* row_shr:8 */
    v = dpp_max_step<0x142, 0xa>(v);   /* row_bcast:15 */
    v = dpp_max_step<0x143, 0xc>(v);   /* row_bcast:31 */
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}

/* Per-sweep max |b - b_old| is kept in kCtlSlots partial maxima, one cache
 * line apart: block b folds its maximum into slot b mod kCtlSlots, so the
 * grid's atomics spread over many lines instead of queueing on one address.
 * A reader takes the max of the slots with one wave-wide load and a DPP
 * reduction; the stop test needs only two wave votes. */
constexpr int kCtlSlots = 64;   /* == wavefront size */
constexpr int kCtlPad = 8;      /* u64 per slot: one 64-byte line */
constexpr int kCtlPerSweep = kCtlSlots * kCtlPad;

__device__ __forceinline__ void ctl_fold(unsigned long long *ctl, int k, unsigned long long m) {
    if (!m) return;
    unsigned long long *slot = ctl + (uint64_t)k * kCtlPerSweep + (blockIdx.x & (kCtlSlots - 1)) * kCtlPad;
    /* no read of the slot first: the atomic's result is unused, so the
     * wave issues it and ends, where a load to skip it would hold the
     * block (and its LDS) for a memory round trip -- var pass 35.1 -> 33.6
     * us on the 1M-variable graph, about 100 atomics per slot per launch */
    atomicMax(slot, m);
}

/* max over the slots of sweep k; every lane of the wave gets it */
__device__ __forceinline__ unsigned long long ctl_read(const unsigned long long *ctl, int k) {
    const int lane = threadIdx.x & 63;
    const unsigned long long v =
        __atomic_load_n(ctl + (uint64_t)k * kCtlPerSweep + lane * kCtlPad, __ATOMIC_RELAXED);
    return wave_max_u64(v);
}

/* stop test: sweep k of a chunk runs only if sweep k-1 left max diff > tol
 * (a sweep that never ran leaves its slots 0 and stops every later sweep).
 * tol = -inf means a fixed number of sweeps (the 3-argument run): no sweep
 * stops, whatever max diff the one before left -- NaN included, which no
 * comparison with tol could let through.
 *
 * The slots hold non-negative doubles or NaN, whose max as bit patterns
 * (ctl_read) is NaN when any slot is NaN and otherwise their max as
 * doubles; so !(max > tol) holds exactly when some slot is NaN or no slot
 * exceeds tol: two votes over the wave's 64 slots, in place of the
 * six-step 64-bit shuffle reduction.  Called with the whole wave active. */
__device__ __forceinline__ bool chunk_stopped(const unsigned long long *ctl, int k, double tol) {
    if (k <= 0 || tol == -__builtin_inf()) return false;
    const int lane = threadIdx.x & 63;
    const double v = __longlong_as_double(
        (long long)__atomic_load_n(ctl + (uint64_t)(k - 1) * kCtlPerSweep + lane * kCtlPad, __ATOMIC_RELAXED));
    return __ballot(v > tol) == 0 || __ballot(v != v) != 0;
}


}  // namespace dev
}  // namespace flbp

namespace {

using namespace flbp::dev;

/* The update of one factor: gathers its v->f messages vf[pos[j]], j < d
 * (registers up to N slots, global memory above), and writes every f->v
 * message fv[t] (damped with fv_old[t] when kDamp) for its slots t. */
template <int N, bool kDamp>
__device__ __forceinline__ void factor_update(const double2 *__restrict__ vf,
                                              const uint32_t *__restrict__ pos, const double2 *fv_old,
                                              double2 *fv, int d, uint32_t meta, double p, double q,
                                              double mask0, double mask1, double damp, double keep) {
    const uint32_t type = meta & 3u;
    const bool clamped = (meta & 4u) != 0;
    const int h = (int)(meta >> 8);
    auto store = [&](int t, double2 r) __attribute__((always_inline)) {
        if (kDamp) {
            const double2 old = fv_old[t];
            r.x = keep * r.x + damp * old.x;
            r.y = keep * r.y + damp * old.y;
        }
        fv[t] = r;
    };
    if (d <= N) {
        RegIn<N> in;
        uint32_t ps[N];
#pragma unroll
        for (int j = 0; j < N; j++) ps[j] = j < d ? pos[j] : 0u;
#pragma unroll
        for (int j = 0; j < N; j++) in.m[j] = j < d ? vf[ps[j]] : make_double2(0.0, 0.0);
        const double2 mh = in.pick(h), mo = in.pick(h == 0 ? 1 : 0);
        for (int t = 0; t < d; t++)
            store(t, fv_msg<N>(in, type, clamped, d, t, h, mh, mo, p, q, mask0, mask1));
    } else {
        GlobalIn in{vf, pos};
        const double2 mh = in.pick(h), mo = in.pick(h == 0 ? 1 : 0);
        for (int t = 0; t < d; t++)
            store(t, fv_msg<0>(in, type, clamped, d, t, h, mh, mo, p, q, mask0, mask1));
    }
}

/* one lane per factor; damping fused into the store.  m_fv_old may alias
 * m_fv (in place: each lane reads its own slots before writing them). */
template <bool kDamp, int N>
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void factor_pass(
    const uint32_t *__restrict__ fac_off, const uint32_t *__restrict__ fac_meta,
    const double *__restrict__ fac_p, const double *__restrict__ fac_q,
    const double2 *__restrict__ fac_mask, const uint32_t *__restrict__ fac_vpos,
    const double2 *__restrict__ m_vf, const double2 *m_fv_old, double2 *m_fv, uint32_t F,
    double damp, double keep, const unsigned long long *__restrict__ ctl, int k, double tol) {
    if (chunk_stopped(ctl, k, tol)) return;
    const uint32_t f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= F) return;
    const uint32_t s = fac_off[f];
    const uint32_t meta = fac_meta[f];
    const double p = fac_p[f];
    const double q = (meta & 3u) ? fac_q[f] : 0.0;
    double mask0 = 1.0, mask1 = 1.0;
    if (meta & 4u) {
        const double2 mk = fac_mask[f];
        mask0 = mk.x;
        mask1 = mk.y;
    }
    factor_update<N, kDamp>(m_vf, fac_vpos + s, m_fv_old + s, m_fv + s, (int)(fac_off[f + 1] - s), meta, p, q,
                            mask0, mask1, damp, keep);
}

/* The update of one variable: gathers its f->v messages m_fv[ve[j]] for
 * j < d (ve = its slice of the var-major -> factor-major slot map), forms
 * the zero-counted product in edge order (calcProdAllKernel), the belief
 * and |b - b_old| (getBeliefsVKernel, norm_functor, dist_functor), and
 * writes the leave-one-out v->f messages at its var-major positions vf[j]
 * (calcMessageVFAllKernel).  Returns |b - b_old| as IEEE bits. */
__device__ __forceinline__ unsigned long long var_update(const uint32_t *__restrict__ ve,
                                                         const double2 *__restrict__ m_fv,
                                                         double2 *__restrict__ vf, double *bel, int d) {
    double prod_0 = 1.0, prod_1 = 1.0;
    uint32_t z0 = 0, z1 = 0;
    uint32_t slot[kVarRegDeg];
    double2 msg[kVarRegDeg];
    auto accumulate = [&](double2 m) __attribute__((always_inline)) {
        if (m.x == 0) z0++; else prod_0 *= m.x;
        if (m.y == 0) z1++; else prod_1 *= m.y;
        scale_pair(prod_0, prod_1);
    };
    /* gathers in unrolled chunks so up to kVarRegDeg loads are in flight;
     * a variable of degree <= kVarRegDeg keeps its messages in registers */
    for (int base = 0; base < d; base += kVarRegDeg) {
#pragma unroll
        for (int j = 0; j < kVarRegDeg; j++)
            if (base + j < d) slot[j] = ve[base + j];
#pragma unroll
        for (int j = 0; j < kVarRegDeg; j++)
            if (base + j < d) msg[j] = m_fv[slot[j]];
#pragma unroll
        for (int j = 0; j < kVarRegDeg; j++)
            if (base + j < d) accumulate(msg[j]);
    }
    const double b0 = prod_0 * (double)(z0 == 0);
    const double b1 = prod_1 * (double)(z1 == 0);
    const double b = b1 / (b0 + b1);
    const double diff = fabs(*bel - b);
    *bel = b;
    if (d <= kVarRegDeg) {
#pragma unroll
        for (int j = 0; j < kVarRegDeg; j++)
            if (j < d) vf[j] = vf_msg(prod_0, prod_1, z0, z1, msg[j]);
    } else {
        for (int base = 0; base < d; base += kVarRegDeg) {
#pragma unroll
            for (int j = 0; j < kVarRegDeg; j++)
                if (base + j < d) slot[j] = ve[base + j];
#pragma unroll
            for (int j = 0; j < kVarRegDeg; j++)
                if (base + j < d) msg[j] = m_fv[slot[j]];
#pragma unroll
            for (int j = 0; j < kVarRegDeg; j++)
                if (base + j < d) vf[base + j] = vf_msg(prod_0, prod_1, z0, z1, msg[j]);
        }
    }
    return (unsigned long long)__double_as_longlong(diff);
}

/* one lane per variable */
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void var_pass(
    const uint32_t *__restrict__ var_off, const uint32_t *__restrict__ var_edge,
    const double2 *__restrict__ m_fv, double2 *__restrict__ m_vf, double *__restrict__ belief,
    uint32_t V, unsigned long long *__restrict__ ctl, int k, double tol) {
    if (chunk_stopped(ctl, k, tol)) return;
    __shared__ unsigned long long red[kBlock / 64];
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    unsigned long long diff_bits = 0;
    if (v < V) {
        const uint32_t s = var_off[v];
        diff_bits = var_update(var_edge + s, m_fv, m_vf + s, belief + v, (int)(var_off[v + 1] - s));
    }
    /* max over non-negative doubles == max over their bit patterns; a NaN
     * (positive after fabs) wins, which stops the loop as NaN > tol is false */
    diff_bits = wave_max_u64(diff_bits);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = diff_bits;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = red[0];
#pragma unroll
        for (int w = 1; w < kBlock / 64; w++) m = red[w] > m ? red[w] : m;
        ctl_fold(ctl, k, m);
    }
}

}  // namespace

#endif
