/*
 * adjoint_kernels.h -- reverse-mode kernels of one damped PARALL sweep.
 *
 * Forward sweep t (state after t sweeps: m_vf^t factor-major, m_fv^t):
 *   u_e     = FV_I(m_vf^{t-1} of I's other slots, p_I)     (normalized pair)
 *   m_fv^t  = (1-d) u + d m_fv^{t-1}
 *   m_vf^t  = normalize(prod over the var's other f->v messages)
 *   b^t     = P1 / (P0 + P1),  P_s = product of all the var's f->v messages
 * The reverse sweep differentiates the exact sum-product map (the forward's
 * per-step rescaling cancels in every normalized quantity and its e1 term
 * is a rounding correction of the same map, so neither appears here):
 *   var_vjp: adjoint of m_fv^t from the adjoints of m_vf^t and b^t, plus
 *            d times the adjoint of m_fv^t carried from sweep t+1;
 *   fac_vjp: adjoint of m_vf^{t-1} and of p_I from (1-d) times the adjoint
 *            of m_fv^t.
 * Leave-one-out products use prefix/suffix products and a running
 * "derivative of a product" recurrence, so nothing divides by a message
 * (exact zeros from clamps are handled like any value).
 */
#ifndef FLBP_ADJOINT_KERNELS_H
#define FLBP_ADJOINT_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sweep_kernels.h"

namespace {

using namespace flbp::dev;

constexpr int kAdjFacDeg = 8;
constexpr int kAdjVarDeg = 8;

/* Reverse of the variable pass of one sweep, one lane per variable.
 * m_fv: f->v messages of the sweep (factor-major); avf: adjoint of the
 * sweep's v->f outputs (factor-major); bbar: adjoint of its beliefs (NULL
 * when the loss does not read them); afv: in = adjoint of m_fv carried from
 * the next sweep before damping (the next sweep's total), out = total
 * adjoint of this sweep's m_fv. */
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void var_vjp(
    const uint32_t *__restrict__ var_off, const uint32_t *__restrict__ var_edge,
    const double2 *__restrict__ m_fv, const double2 *__restrict__ avf,
    const double *__restrict__ bbar, double2 *__restrict__ afv, double2 *__restrict__ gscratch,
    double damp, uint32_t V) {
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= V) return;
    const uint32_t s = var_off[v];
    const int d = (int)(var_off[v + 1] - s);
    const double bb = bbar ? bbar[v] : 0.0;
    if (d <= kAdjVarDeg) {
        uint32_t slot[kAdjVarDeg];
        double m0[kAdjVarDeg], m1[kAdjVarDeg];
#pragma unroll
        for (int j = 0; j < kAdjVarDeg; j++) {
            if (j < d) {
                slot[j] = var_edge[s + j];
                const double2 m = m_fv[slot[j]];
                m0[j] = m.x;
                m1[j] = m.y;
            } else {
                slot[j] = 0;
                m0[j] = 1.0;
                m1[j] = 1.0;
            }
        }
        /* leave-one-out products by prefix/suffix */
        double L0[kAdjVarDeg], L1[kAdjVarDeg];
        double a0 = 1.0, a1 = 1.0;
#pragma unroll
        for (int j = 0; j < kAdjVarDeg; j++) {
            L0[j] = a0;
            L1[j] = a1;
            a0 *= m0[j];
            a1 *= m1[j];
        }
        const double P0 = a0, P1 = a1;
        a0 = 1.0;
        a1 = 1.0;
#pragma unroll
        for (int j = kAdjVarDeg - 1; j >= 0; j--) {
            L0[j] *= a0;
            L1[j] *= a1;
            a0 *= m0[j];
            a1 *= m1[j];
        }
        /* adjoint of each leave-one-out product through the normalization */
        double g0[kAdjVarDeg], g1[kAdjVarDeg];
#pragma unroll
        for (int j = 0; j < kAdjVarDeg; j++) {
            g0[j] = 0.0;
            g1[j] = 0.0;
            const double S = L0[j] + L1[j];
            if (j < d && S != 0.0) {
                const double2 ob = avf[slot[j]];
                const double dot = ob.x * (L0[j] / S) + ob.y * (L1[j] / S);
                g0[j] = (ob.x - dot) / S;
                g1[j] = (ob.y - dot) / S;
            }
        }
        double G0 = 0.0, G1 = 0.0;
        if (bb != 0.0 && P0 + P1 != 0.0) {
            const double den = (P0 + P1) * (P0 + P1);
            G1 = bb * P0 / den;
            G0 = -bb * P1 / den;
        }
        /* d/dm_k of sum_e g_e prod_{j != e} m_j, for each k: a running
         * product/derivative recurrence over j != k */
#pragma unroll
        for (int k = 0; k < kAdjVarDeg; k++) {
            if (k >= d) break;
            double w0 = 0.0, w1 = 0.0, q0 = 1.0, q1 = 1.0;
#pragma unroll
            for (int j = 0; j < kAdjVarDeg; j++) {
                if (j == k) continue;
                w0 = w0 * m0[j] + g0[j] * q0;
                w1 = w1 * m1[j] + g1[j] * q1;
                q0 *= m0[j];
                q1 *= m1[j];
            }
            const double2 c = afv[slot[k]];
            afv[slot[k]] = make_double2(damp * c.x + w0 + G0 * L0[k], damp * c.y + w1 + G1 * L1[k]);
        }
        return;
    }
    /* high-degree variables: the same quantities with O(d^2) loops over
     * global memory (the lane owns these slots and its scratch row) */
    double P0 = 1.0, P1 = 1.0;
    for (int j = 0; j < d; j++) {
        const double2 m = m_fv[var_edge[s + j]];
        P0 *= m.x;
        P1 *= m.y;
    }
    double G0 = 0.0, G1 = 0.0;
    if (bb != 0.0 && P0 + P1 != 0.0) {
        const double den = (P0 + P1) * (P0 + P1);
        G1 = bb * P0 / den;
        G0 = -bb * P1 / den;
    }
    auto loo = [&](int e, double &l0, double &l1) __attribute__((always_inline)) {
        l0 = 1.0;
        l1 = 1.0;
        for (int j = 0; j < d; j++) {
            if (j == e) continue;
            const double2 m = m_fv[var_edge[s + j]];
            l0 *= m.x;
            l1 *= m.y;
        }
    };
    /* g_e of every edge into the var-major scratch row, then the
     * recurrence per edge: O(d^2) per variable */
    for (int e = 0; e < d; e++) {
        double l0, l1;
        loo(e, l0, l1);
        const double S = l0 + l1;
        double2 g = make_double2(0.0, 0.0);
        if (S != 0.0) {
            const double2 ob = avf[var_edge[s + e]];
            const double dot = ob.x * (l0 / S) + ob.y * (l1 / S);
            g = make_double2((ob.x - dot) / S, (ob.y - dot) / S);
        }
        gscratch[s + e] = g;
    }
    for (int k = 0; k < d; k++) {
        double w0 = 0.0, w1 = 0.0, q0 = 1.0, q1 = 1.0;
        for (int j = 0; j < d; j++) {
            if (j == k) continue;
            const double2 m = m_fv[var_edge[s + j]];
            const double2 g = gscratch[s + j];
            w0 = w0 * m.x + g.x * q0;
            w1 = w1 * m.y + g.y * q1;
            q0 *= m.x;
            q1 *= m.y;
        }
        double l0, l1;
        loo(k, l0, l1);
        const uint32_t sl = var_edge[s + k];
        const double2 c = afv[sl];
        afv[sl] = make_double2(damp * c.x + w0 + G0 * l0, damp * c.y + w1 + G1 * l1);
    }
}

/* The VJP of one factor output (slot t) with adjoint ub, accumulating into
 * the input adjoints acc0/acc1 and the p-adjoint pb.  x = the state the CPD
 * singles out (state 1 for AND, 0 for OR), y = the other state. */
template <int N>
__device__ __forceinline__ void fac_output_vjp(const double (&m0)[N], const double (&m1)[N], int d,
                                               int t, int h, uint32_t type, bool clamped, double p,
                                               double q, double mask0, double mask1, double2 ub,
                                               double (&acc0)[N], double (&acc1)[N], double &pb) {
    if (type == 0) {   /* (1-p, p), already normalized */
        const double u0 = 1 - p, u1 = p, S = u0 + u1;
        const double dot = ub.x * (u0 / S) + ub.y * (u1 / S);
        pb += ((ub.y - dot) - (ub.x - dot)) / S;
        return;
    }
    const bool is_and = (type == 1);
    const bool head_target = (t == h);
    const double mk0 = clamped ? mask0 : 1.0, mk1 = clamped ? mask1 : 1.0;
    double X[N], A[N];
#pragma unroll
    for (int j = 0; j < N; j++) {
        const bool excl = (j >= d) || (j == t) || (!head_target && j == h);
        const double x = is_and ? m1[j] : m0[j], y = is_and ? m0[j] : m1[j];
        X[j] = excl ? 1.0 : x;
        A[j] = excl ? 1.0 : x + y;
    }
    double looX[N], looA[N];
    double px = 1.0, pa = 1.0;
#pragma unroll
    for (int j = 0; j < N; j++) {
        looX[j] = px;
        looA[j] = pa;
        px *= X[j];
        pa *= A[j];
    }
    const double pi = px, Z = pa;
    px = 1.0;
    pa = 1.0;
#pragma unroll
    for (int j = N - 1; j >= 0; j--) {
        looX[j] *= px;
        looA[j] *= pa;
        px *= X[j];
        pa *= A[j];
    }
    double Xb, Zb;            /* adjoints of pi and Z */
    double hxb = 0.0, hyb = 0.0;
    if (head_target) {
        const double Mu = (p - q) * pi + q * Z, Mw = (q - p) * pi + (1 - q) * Z;
        const double M0 = (is_and ? Mw : Mu) * mk0, M1 = (is_and ? Mu : Mw) * mk1;
        const double S = M0 + M1;
        if (S == 0.0) return;
        const double dot = ub.x * (M0 / S) + ub.y * (M1 / S);
        const double Mb0 = (ub.x - dot) / S * mk0, Mb1 = (ub.y - dot) / S * mk1;
        const double Ub = is_and ? Mb1 : Mb0, Wb = is_and ? Mb0 : Mb1;
        Xb = (p - q) * (Ub - Wb);
        Zb = q * Ub + (1 - q) * Wb;
        pb += (Ub - Wb) * pi;
    } else {
        double m0h = 0.0, m1h = 0.0;
#pragma unroll
        for (int j = 0; j < N; j++) {
            m0h = (j == h) ? m0[j] : m0h;
            m1h = (j == h) ? m1[j] : m1h;
        }
        const double hx = is_and ? m1h * mk1 : m0h * mk0;
        const double hy = is_and ? m0h * mk0 : m1h * mk1;
        const double B = (p - q) * (hx - hy), Aq = q * hx + (1 - q) * hy;
        const double Ma = Aq * Z, Mb = Ma + B * pi;
        const double M0 = is_and ? Ma : Mb, M1 = is_and ? Mb : Ma;
        const double S = M0 + M1;
        if (S == 0.0) return;
        const double dot = ub.x * (M0 / S) + ub.y * (M1 / S);
        const double Mb0 = (ub.x - dot) / S, Mb1 = (ub.y - dot) / S;
        const double Mab = is_and ? Mb0 : Mb1, Mbb = is_and ? Mb1 : Mb0;
        Zb = (Mab + Mbb) * Aq;              /* adjoint of Z' */
        const double Ab = (Mab + Mbb) * Z;  /* adjoint of A */
        Xb = Mbb * B;                       /* adjoint of pi' */
        const double Bbar = Mbb * pi;       /* adjoint of B */
        pb += Bbar * (hx - hy);
        hxb = Bbar * (p - q) + Ab * q;
        hyb = -Bbar * (p - q) + Ab * (1 - q);
    }
#pragma unroll
    for (int j = 0; j < N; j++) {
        const bool excl = (j >= d) || (j == t) || (!head_target && j == h);
        double xb = excl ? 0.0 : Xb * looX[j] + Zb * looA[j];
        double yb = excl ? 0.0 : Zb * looA[j];
        if (!head_target && j == h) {
            /* head message of a body target */
            xb = hxb * (is_and ? mk1 : mk0);
            yb = hyb * (is_and ? mk0 : mk1);
        }
        acc1[j] += is_and ? xb : yb;
        acc0[j] += is_and ? yb : xb;
    }
}

/* Reverse of the factor pass of one sweep, one lane per factor.
 * m_vf: the sweep's v->f inputs; afv: total adjoint of the sweep's f->v
 * outputs; avf_out: adjoint of the v->f inputs (written); pbar += dL/dp. */
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void fac_vjp(
    const uint32_t *__restrict__ fac_off, const uint32_t *__restrict__ fac_meta,
    const double *__restrict__ fac_p, const double *__restrict__ fac_q,
    const double2 *__restrict__ fac_mask, const double2 *__restrict__ m_vf,
    const double2 *__restrict__ afv, double2 *__restrict__ avf_out, double *__restrict__ pbar,
    double keep, uint32_t F) {
    const uint32_t f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= F) return;
    const uint32_t s = fac_off[f];
    const int d = (int)(fac_off[f + 1] - s);
    const uint32_t meta = fac_meta[f];
    const uint32_t type = meta & 3u;
    const bool clamped = (meta & 4u) != 0;
    const int h = (int)(meta >> 8);
    const double p = fac_p[f];
    const double q = fac_q[f];
    double mask0 = 1.0, mask1 = 1.0;
    if (clamped) {
        const double2 mk = fac_mask[f];
        mask0 = mk.x;
        mask1 = mk.y;
    }
    double pb = 0.0;
    if (d <= kAdjFacDeg) {
        double m0[kAdjFacDeg], m1[kAdjFacDeg], acc0[kAdjFacDeg], acc1[kAdjFacDeg];
#pragma unroll
        for (int j = 0; j < kAdjFacDeg; j++) {
            const double2 m = j < d ? m_vf[s + j] : make_double2(1.0, 1.0);
            m0[j] = m.x;
            m1[j] = m.y;
            acc0[j] = 0.0;
            acc1[j] = 0.0;
        }
        for (int t = 0; t < d; t++) {
            const double2 a = afv[s + t];
            fac_output_vjp<kAdjFacDeg>(m0, m1, d, t, h, type, clamped, p, q, mask0, mask1,
                                       make_double2(keep * a.x, keep * a.y), acc0, acc1, pb);
        }
#pragma unroll
        for (int j = 0; j < kAdjFacDeg; j++)
            if (j < d) avf_out[s + j] = make_double2(acc0[j], acc1[j]);
    } else {
        /* factors above the register degree: not produced by the layered
         * generator; flagged by the host (adjoint creation rejects them) */
        for (int j = 0; j < d; j++) avf_out[s + j] = make_double2(0.0, 0.0);
    }
    pbar[f] += pb;
}

}  // namespace

#endif
