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

constexpr int kAdjFacDeg = 6;   /* factors up to this degree use the exact-degree reverse code */
constexpr int kAdjVarDeg = 4;   /* var_vjp_one keeps variables up to this degree in registers */

/* Reverse of the variable pass for one variable v, run by one lane.
 * m_fv: f->v messages of the sweep (factor-major); avf: adjoint of the
 * sweep's v->f outputs (factor-major); bbar: adjoint of its beliefs (NULL
 * when the loss does not read them); afv (var-major, written at the
 * variable's own positions): in = total adjoint of the next sweep's m_fv
 * (carried through damping), out = total adjoint of this sweep's m_fv.
 *
 * Per state s, with L_e = prod_{j != e} m_j and g_e the adjoint of L_e, the
 * adjoint of m_k is sum_{e != k} g_e prod_{j != e,k} m_j + G * L_k.  It is
 * formed in O(d) from prefix products P, suffix products Q and the two
 * running sums W_k = sum_{e<k} g_e prod_{j<k, j!=e} m_j (forward) and
 * U_k = sum_{e>k} g_e prod_{j>k, j!=e} m_j (backward): adjoint = W_k Q_k +
 * P_k U_k + G P_k Q_k.  Variables above kAdjVarDeg keep P, Q, g, U in a
 * lane-owned scratch row (4 pairs per edge). */
__device__ __forceinline__ void gadj(double L0, double L1, double2 ob, double &g0, double &g1) {
    /* adjoint of the unnormalized pair L from the adjoint ob of L / sum(L) */
    const double S = L0 + L1;
    g0 = g1 = 0.0;
    if (S != 0.0) {
        const double dot = ob.x * (L0 / S) + ob.y * (L1 / S);
        g0 = (ob.x - dot) / S;
        g1 = (ob.y - dot) / S;
    }
}

__device__ __forceinline__ void belief_adj(double bb, double P0, double P1, double &G0, double &G1) {
    /* adjoint of the full products from the adjoint bb of b = P1 / (P0 + P1) */
    G0 = G1 = 0.0;
    if (bb != 0.0 && P0 + P1 != 0.0) {
        const double den = (P0 + P1) * (P0 + P1);
        G1 = bb * P0 / den;
        G0 = -bb * P1 / den;
    }
}

__device__ __attribute__((noinline)) void var_vjp_one(
    const uint32_t *__restrict__ var_off, const uint32_t *__restrict__ var_edge,
    const double2 *__restrict__ m_fv, const double2 *__restrict__ avf,
    const double *__restrict__ bbar, double2 *__restrict__ afv, double2 *__restrict__ scratch,
    double damp, uint32_t v) {
    const uint32_t s = var_off[v];
    const int d = (int)(var_off[v + 1] - s);
    const double bb = bbar ? bbar[v] : 0.0;
    if (d <= kAdjVarDeg) {
        constexpr int N = kAdjVarDeg;
        uint32_t slot[N];
        double m0[N], m1[N], P0[N], P1[N], Q0[N], Q1[N], g0[N], g1[N];
#pragma unroll
        for (int j = 0; j < N; j++) slot[j] = j < d ? var_edge[s + j] : 0u;
#pragma unroll
        for (int j = 0; j < N; j++) {
            const double2 m = j < d ? m_fv[slot[j]] : make_double2(1.0, 1.0);
            m0[j] = m.x;
            m1[j] = m.y;
        }
        double a0 = 1.0, a1 = 1.0;
#pragma unroll
        for (int j = 0; j < N; j++) {
            P0[j] = a0;
            P1[j] = a1;
            a0 *= m0[j];
            a1 *= m1[j];
        }
        double G0, G1;
        belief_adj(bb, a0, a1, G0, G1);
        a0 = a1 = 1.0;
#pragma unroll
        for (int j = N - 1; j >= 0; j--) {
            Q0[j] = a0;
            Q1[j] = a1;
            a0 *= m0[j];
            a1 *= m1[j];
        }
#pragma unroll
        for (int j = 0; j < N; j++) {
            if (j < d) gadj(P0[j] * Q0[j], P1[j] * Q1[j], avf[slot[j]], g0[j], g1[j]);
            else g0[j] = g1[j] = 0.0;
        }
        /* backward running sums U_k, kept in the g arrays' partners */
        double U0[N], U1[N];
        double u0 = 0.0, u1 = 0.0;
#pragma unroll
        for (int j = N - 1; j >= 0; j--) {
            U0[j] = u0;
            U1[j] = u1;
            u0 = u0 * m0[j] + g0[j] * Q0[j];
            u1 = u1 * m1[j] + g1[j] * Q1[j];
        }
        double w0 = 0.0, w1 = 0.0;
#pragma unroll
        for (int k = 0; k < N; k++) {
            if (k < d) {
                const double2 c = afv[s + k];
                const double r0 = w0 * Q0[k] + P0[k] * U0[k] + G0 * (P0[k] * Q0[k]);
                const double r1 = w1 * Q1[k] + P1[k] * U1[k] + G1 * (P1[k] * Q1[k]);
                afv[s + k] = make_double2(damp * c.x + r0, damp * c.y + r1);
            }
            w0 = w0 * m0[k] + g0[k] * P0[k];
            w1 = w1 * m1[k] + g1[k] * P1[k];
        }
        return;
    }
    /* high degree: three passes over the edges with P, Q, g, U in the
     * lane's scratch row scratch[4*(s+j) .. 4*(s+j)+3] */
    double2 *row = scratch + 4ull * s;
    double a0 = 1.0, a1 = 1.0;
    for (int j = 0; j < d; j++) {
        const double2 m = m_fv[var_edge[s + j]];
        row[4 * j] = make_double2(a0, a1);          /* P_j */
        a0 *= m.x;
        a1 *= m.y;
    }
    double G0, G1;
    belief_adj(bb, a0, a1, G0, G1);
    a0 = a1 = 1.0;
    double u0 = 0.0, u1 = 0.0;
    for (int j = d - 1; j >= 0; j--) {
        const uint32_t sl = var_edge[s + j];
        const double2 m = m_fv[sl];
        const double2 P = row[4 * j];
        double g0, g1;
        gadj(P.x * a0, P.y * a1, avf[sl], g0, g1);
        row[4 * j + 1] = make_double2(a0, a1);      /* Q_j */
        row[4 * j + 2] = make_double2(g0, g1);      /* g_j */
        row[4 * j + 3] = make_double2(u0, u1);      /* U_j */
        u0 = u0 * m.x + g0 * a0;
        u1 = u1 * m.y + g1 * a1;
        a0 *= m.x;
        a1 *= m.y;
    }
    double w0 = 0.0, w1 = 0.0;
    for (int k = 0; k < d; k++) {
        const uint32_t sl = var_edge[s + k];
        const double2 m = m_fv[sl];
        const double2 P = row[4 * k], Q = row[4 * k + 1], g = row[4 * k + 2], U = row[4 * k + 3];
        const double2 c = afv[s + k];
        const double r0 = w0 * Q.x + P.x * U.x + G0 * (P.x * Q.x);
        const double r1 = w1 * Q.y + P.y * U.y + G1 * (P.y * Q.y);
        afv[s + k] = make_double2(damp * c.x + r0, damp * c.y + r1);
        w0 = w0 * m.x + g.x * P.x;
        w1 = w1 * m.y + g.y * P.y;
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

/* The VJP of every output of one factor of any degree d in O(d): m(j) is
 * the v->f input of slot j, a(j) the adjoint of output j (before the
 * damping weight keep), out(j, r) receives the adjoint r of input j; row
 * holds 4 pairs of scratch per slot.  With x the state the CPD singles out
 * and A = x + y, over the body slots B (every slot but the head h):
 *   head output:   pi = prod_B x, Z = prod_B A
 *   body output t: pi_t = prod_{B\t} x, Z_t = prod_{B\t} A
 * Input j of B gets Xb_h dpi/dx_j + Zb_h dZ/dA_j from the head output and
 * sum_{t != j} (Xb_t prod_{B\{t,j}} x + Zb_t prod_{B\{t,j}} A) from the
 * body outputs; the double leave-outs come from prefix products P, suffix
 * products Q and the running sums W (forward) and U (backward), as in
 * var_vjp_one.  The head slot gets the sum of the body outputs' head
 * adjoints.  Reads of m(j) precede out(j), so out may overwrite m. */
template <class MIn, class AIn, class Out>
__device__ __forceinline__ double fac_vjp_general(const MIn &m, const AIn &a, const Out &out, double2 *row, int d,
                                                  int h, uint32_t type, bool clamped, double p, double q,
                                                  double mask0, double mask1, double keep) {
    double pb = 0.0;
    if (type == 0) {   /* (1-p, p) */
        const double2 ub = make_double2(keep * a(0).x, keep * a(0).y);
        const double u0 = 1 - p, u1 = p, S = u0 + u1;
        const double dot = ub.x * (u0 / S) + ub.y * (u1 / S);
        pb = ((ub.y - dot) - (ub.x - dot)) / S;
        for (int j = 0; j < d; j++) out(j, make_double2(0.0, 0.0));
        return pb;
    }
    const bool is_and = (type == 1);
    const double mk0 = clamped ? mask0 : 1.0, mk1 = clamped ? mask1 : 1.0;
    auto xa = [&](double2 v, double &x, double &A) {
        x = is_and ? v.y : v.x;
        A = v.x + v.y;
    };
    const double2 mh = m(h);
    /* pass 1: prefix products over B */
    double px = 1.0, pa = 1.0;
    for (int j = 0; j < d; j++) {
        if (j == h) continue;
        double x, A;
        xa(m(j), x, A);
        row[4 * j] = make_double2(px, pa);
        px *= x;
        pa *= A;
    }
    /* the head output */
    double Xbh = 0.0, Zbh = 0.0;
    {
        const double2 ub = make_double2(keep * a(h).x, keep * a(h).y);
        const double Mu = (p - q) * px + q * pa, Mw = (q - p) * px + (1 - q) * pa;
        const double M0 = (is_and ? Mw : Mu) * mk0, M1 = (is_and ? Mu : Mw) * mk1;
        const double S = M0 + M1;
        if (S != 0.0) {
            const double dot = ub.x * (M0 / S) + ub.y * (M1 / S);
            const double Mb0 = (ub.x - dot) / S * mk0, Mb1 = (ub.y - dot) / S * mk1;
            const double Ub = is_and ? Mb1 : Mb0, Wb = is_and ? Mb0 : Mb1;
            Xbh = (p - q) * (Ub - Wb);
            Zbh = q * Ub + (1 - q) * Wb;
            pb += (Ub - Wb) * px;
        }
    }
    /* pass 2: suffix products, the body outputs' adjoints, running sums U */
    const double hx = is_and ? mh.y * mk1 : mh.x * mk0;
    const double hy = is_and ? mh.x * mk0 : mh.y * mk1;
    const double B = (p - q) * (hx - hy), Aq = q * hx + (1 - q) * hy;
    double hxb = 0.0, hyb = 0.0;
    double qx = 1.0, qa = 1.0, ux = 0.0, ua = 0.0;
    for (int j = d - 1; j >= 0; j--) {
        if (j == h) continue;
        double x, A;
        xa(m(j), x, A);
        const double2 P = row[4 * j];
        const double pit = P.x * qx, Zt = P.y * qa;
        const double2 ub = make_double2(keep * a(j).x, keep * a(j).y);
        const double Ma = Aq * Zt, Mb = Ma + B * pit;
        const double M0 = is_and ? Ma : Mb, M1 = is_and ? Mb : Ma;
        const double S = M0 + M1;
        double Xb = 0.0, Zb = 0.0;
        if (S != 0.0) {
            const double dot = ub.x * (M0 / S) + ub.y * (M1 / S);
            const double Mb0 = (ub.x - dot) / S, Mb1 = (ub.y - dot) / S;
            const double Mab = is_and ? Mb0 : Mb1, Mbb = is_and ? Mb1 : Mb0;
            Zb = (Mab + Mbb) * Aq;
            const double Ab = (Mab + Mbb) * Zt;
            Xb = Mbb * B;
            const double Bbar = Mbb * pit;
            pb += Bbar * (hx - hy);
            hxb += Bbar * (p - q) + Ab * q;
            hyb += -Bbar * (p - q) + Ab * (1 - q);
        }
        row[4 * j + 1] = make_double2(qx, qa);
        row[4 * j + 2] = make_double2(Xb, Zb);
        row[4 * j + 3] = make_double2(ux, ua);
        ux = ux * x + Xb * qx;
        ua = ua * A + Zb * qa;
        qx *= x;
        qa *= A;
    }
    /* pass 3: running sums W and the input adjoints */
    double wx = 0.0, wa = 0.0;
    for (int k = 0; k < d; k++) {
        if (k == h) continue;
        double x, A;
        xa(m(k), x, A);
        const double2 P = row[4 * k], Q = row[4 * k + 1], g = row[4 * k + 2], U = row[4 * k + 3];
        const double Sx = wx * Q.x + P.x * U.x, SA = wa * Q.y + P.y * U.y;
        const double xb = Xbh * (P.x * Q.x) + Zbh * (P.y * Q.y) + Sx + SA;
        const double yb = Zbh * (P.y * Q.y) + SA;
        out(k, is_and ? make_double2(yb, xb) : make_double2(xb, yb));
        wx = wx * x + g.x * P.x;
        wa = wa * A + g.y * P.y;
    }
    out(h, is_and ? make_double2(hyb * mk0, hxb * mk1) : make_double2(hxb * mk0, hyb * mk1));
    return pb;
}

/* Slots per block of the staged reverse kernels. */
constexpr int kAdjStage = 512;

/* Reverse of the variable pass, LDS-staged: a block owns a contiguous
 * range of at most kBlock variables and kAdjStage var-major slots, gathers
 * the sweep's f->v messages and v->f adjoints and reads the carried f->v
 * adjoints into LDS, then one lane per variable (work order: by degree)
 * runs the prefix / suffix passes of var_vjp_one with its scratch in LDS,
 * and the block writes the f->v adjoints back coalesced.  A variable with
 * more than kAdjStage slots gets a block of its own and runs var_vjp_one
 * with the global scratch row. */
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void var_vjp_staged(
    const uint32_t *__restrict__ blk, const uint32_t *__restrict__ var_off,
    const uint32_t *__restrict__ var_edge, const uint16_t *__restrict__ work,
    const double2 *__restrict__ m_fv, const double2 *__restrict__ avf, const double *__restrict__ bbar,
    double2 *__restrict__ afv, double2 *__restrict__ scratch, double damp) {
    __shared__ double2 s_m[kAdjStage];   /* the variable's f->v messages */
    __shared__ double2 s_a[kAdjStage];   /* v->f adjoint, then g_j * P_j */
    __shared__ double2 s_c[kAdjStage];   /* carried f->v adjoint, then the result */
    __shared__ double2 s_x[kAdjStage];   /* prefix P_j, then P_j U_j + G P_j Q_j */
    __shared__ double2 s_y[kAdjStage];   /* suffix Q_j */
    __shared__ uint32_t s_start[kBlock + 1];
    __shared__ double s_bb[kBlock];
    const uint32_t v0 = blk[blockIdx.x], v1 = blk[blockIdx.x + 1];
    const uint32_t base = var_off[v0];
    const uint32_t n = var_off[v1] - base;
    const int tid = threadIdx.x;
    if (n > (uint32_t)kAdjStage) {
        if (tid == 0) var_vjp_one(var_off, var_edge, m_fv, avf, bbar, afv, scratch, damp, v0);
        return;
    }
    const uint32_t nv = v1 - v0;
    uint32_t lv = 0;
    if ((uint32_t)tid < nv) {
        lv = work[v0 + tid];
        s_start[tid] = var_off[v0 + tid] - base;
        s_bb[tid] = bbar ? bbar[v0 + tid] : 0.0;
    }
    if (tid == 0) s_start[nv] = n;
    for (uint32_t e = tid; e < n; e += kBlock) {
        const uint32_t sl = var_edge[base + e];
        s_m[e] = m_fv[sl];
        s_a[e] = avf[sl];
        s_c[e] = afv[base + e];
    }
    __syncthreads();
    if ((uint32_t)tid < nv) {
        const uint32_t s = s_start[lv];
        const int d = (int)(s_start[lv + 1] - s);
        double a0 = 1.0, a1 = 1.0;
        for (int j = 0; j < d; j++) {
            const double2 m = s_m[s + j];
            s_x[s + j] = make_double2(a0, a1);
            a0 *= m.x;
            a1 *= m.y;
        }
        double G0, G1;
        belief_adj(s_bb[lv], a0, a1, G0, G1);
        double q0 = 1.0, q1 = 1.0, u0 = 0.0, u1 = 0.0;
        for (int j = d - 1; j >= 0; j--) {
            const double2 m = s_m[s + j], P = s_x[s + j];
            double g0, g1;
            gadj(P.x * q0, P.y * q1, s_a[s + j], g0, g1);
            s_x[s + j] = make_double2(P.x * u0 + G0 * (P.x * q0), P.y * u1 + G1 * (P.y * q1));
            s_y[s + j] = make_double2(q0, q1);
            s_a[s + j] = make_double2(g0 * P.x, g1 * P.y);
            u0 = u0 * m.x + g0 * q0;
            u1 = u1 * m.y + g1 * q1;
            q0 *= m.x;
            q1 *= m.y;
        }
        double w0 = 0.0, w1 = 0.0;
        for (int k = 0; k < d; k++) {
            const double2 m = s_m[s + k], Q = s_y[s + k], R = s_x[s + k], gP = s_a[s + k], c = s_c[s + k];
            s_c[s + k] = make_double2(damp * c.x + (w0 * Q.x + R.x), damp * c.y + (w1 * Q.y + R.y));
            w0 = w0 * m.x + gP.x;
            w1 = w1 * m.y + gP.y;
        }
    }
    __syncthreads();
    for (uint32_t e = tid; e < n; e += kBlock) afv[base + e] = s_c[e];
}

/* every output VJP of one factor of degree exactly D (operands in LDS):
 * accumulates the input adjoints in place of the inputs and returns dL/dp */
template <int D>
__device__ __forceinline__ double fac_vjp_exact(double2 *m, const double2 *a, int h, uint32_t type, bool clamped,
                                                double p, double q, double mask0, double mask1, double keep) {
    double m0[D], m1[D], acc0[D], acc1[D];
#pragma unroll
    for (int j = 0; j < D; j++) {
        const double2 x = m[j];
        m0[j] = x.x;
        m1[j] = x.y;
        acc0[j] = 0.0;
        acc1[j] = 0.0;
    }
    double pb = 0.0;
#pragma unroll
    for (int t = 0; t < D; t++) {
        const double2 x = a[t];
        fac_output_vjp<D>(m0, m1, D, t, h, type, clamped, p, q, mask0, mask1, make_double2(keep * x.x, keep * x.y),
                          acc0, acc1, pb);
    }
#pragma unroll
    for (int j = 0; j < D; j++) m[j] = make_double2(acc0[j], acc1[j]);
    return pb;
}

/* Reverse of the factor pass, LDS-staged: a block owns a contiguous range
 * of at most kBlock factors and kAdjStage slots, gathers the sweep's v->f
 * inputs and the f->v adjoints (both var-major) into LDS, one lane per
 * factor (work order: by type and degree, so a wave runs one degree's
 * code) forms the adjoints of its inputs in place, and the block writes
 * them back factor-major, coalesced.  pbar[f] += dL/dp_f by its lane. */
__global__ __launch_bounds__(kBlock) __attribute__((unused)) void fac_vjp_staged(
    const uint32_t *__restrict__ blk, const uint32_t *__restrict__ fac_off,
    const uint32_t *__restrict__ fac_meta, const double *__restrict__ fac_p,
    const double *__restrict__ fac_q, const double2 *__restrict__ fac_mask,
    const uint32_t *__restrict__ fac_vpos, const uint16_t *__restrict__ work,
    const double2 *__restrict__ m_vf, const double2 *__restrict__ afv, double2 *__restrict__ avf_out,
    double *__restrict__ pbar, double2 *__restrict__ scratch, double keep) {
    __shared__ double2 s_m[kAdjStage];
    __shared__ double2 s_a[kAdjStage];
    __shared__ uint32_t s_start[kBlock + 1];
    const uint32_t f0 = blk[blockIdx.x], f1 = blk[blockIdx.x + 1];
    const uint32_t base = fac_off[f0];
    const uint32_t n = fac_off[f1] - base;
    const int tid = threadIdx.x;
    if (n > (uint32_t)kAdjStage) {
        /* one factor wider than a staged block: lane 0, from global memory */
        if (tid == 0) {
            const uint32_t meta = fac_meta[f0];
            double mask0 = 1.0, mask1 = 1.0;
            if (meta & 4u) {
                const double2 mk = fac_mask[f0];
                mask0 = mk.x;
                mask1 = mk.y;
            }
            const uint32_t *vp = fac_vpos + base;
            auto m = [&](int j) { return m_vf[vp[j]]; };
            auto a = [&](int j) { return afv[vp[j]]; };
            auto out = [&](int j, double2 r) { avf_out[base + j] = r; };
            pbar[f0] += fac_vjp_general(m, a, out, scratch + 4ull * base, (int)n, (int)(meta >> 8), meta & 3u,
                                        (meta & 4u) != 0, fac_p[f0], fac_q[f0], mask0, mask1, keep);
        }
        return;
    }
    const uint32_t nf = f1 - f0;
    uint32_t lf = 0;
    if ((uint32_t)tid < nf) {
        lf = work[f0 + tid];
        s_start[tid] = fac_off[f0 + tid] - base;
    }
    if (tid == 0) s_start[nf] = n;
    for (uint32_t e = tid; e < n; e += kBlock) {
        const uint32_t k = fac_vpos[base + e];
        s_m[e] = m_vf[k];
        s_a[e] = afv[k];
    }
    __syncthreads();
    if ((uint32_t)tid < nf) {
        const uint32_t f = f0 + lf;
        const uint32_t s = s_start[lf];
        const int d = (int)(s_start[lf + 1] - s);
        const uint32_t meta = fac_meta[f];
        const uint32_t type = meta & 3u;
        const bool clamped = (meta & 4u) != 0;
        const int h = (int)(meta >> 8);
        const double p = fac_p[f], q = fac_q[f];
        double mask0 = 1.0, mask1 = 1.0;
        if (clamped) {
            const double2 mk = fac_mask[f];
            mask0 = mk.x;
            mask1 = mk.y;
        }
        double pb = 0.0;
        double2 *m = s_m + s;
        const double2 *a = s_a + s;
        switch (d) {
        case 1: pb = fac_vjp_exact<1>(m, a, h, type, clamped, p, q, mask0, mask1, keep); break;
        case 2: pb = fac_vjp_exact<2>(m, a, h, type, clamped, p, q, mask0, mask1, keep); break;
        case 3: pb = fac_vjp_exact<3>(m, a, h, type, clamped, p, q, mask0, mask1, keep); break;
        case 4: pb = fac_vjp_exact<4>(m, a, h, type, clamped, p, q, mask0, mask1, keep); break;
        case 5: pb = fac_vjp_exact<5>(m, a, h, type, clamped, p, q, mask0, mask1, keep); break;
        case 6: pb = fac_vjp_exact<6>(m, a, h, type, clamped, p, q, mask0, mask1, keep); break;
        default:
            if (d > kAdjFacDeg) {   /* wider factors: the O(d) form, scratch rows at the factor's slots */
                auto mi = [&](int j) { return m[j]; };
                auto ai = [&](int j) { return a[j]; };
                auto out = [&](int j, double2 r) { m[j] = r; };
                pb = fac_vjp_general(mi, ai, out, scratch + 4ull * (base + s), d, h, type, clamped, p, q, mask0,
                                     mask1, keep);
            }
            break;   /* degree 0 has no slots */
        }
        pbar[f] += pb;
    }
    __syncthreads();
    for (uint32_t e = tid; e < n; e += kBlock) avf_out[base + e] = s_m[e];
}

}  // namespace

#endif