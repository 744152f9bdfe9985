/*
 * oracle.cpp -- CPU FP64 restatement of the reference's CausalBP path
 * (PARALL, SEQFIX and TOPO updates).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * It follows the reference line by line where the floating-point operation
 * order matters (citations are paths relative to the reference repository):
 *   - graph ordering      csrc/ext/causal_fg.cpp:28-68, csrc/lbp/bipgraph.cpp:24-42
 *   - CSR + descriptors   csrc/ext/causal_bp.cu:310-407, 556-581, 724-817
 *   - f->v messages       csrc/kernel/causal_bp_parall.cu:14-18, 94-398
 *   - products            csrc/kernel/causal_bp_parall.cu:20-53
 *   - v->f messages       csrc/kernel/causal_bp_parall.cu:55-92
 *   - beliefs             csrc/kernel/causal_bp_seq.cu:102-111, csrc/ext/causal_bp.cu:118-128, 483-515
 *   - run loops           csrc/ext/causal_bp.cu:824-1004
 *   - clamp + rebuild     csrc/ext/causal_fg.cpp:17-26, csrc/ext/causal_factor.cpp:63-100,
 *                         examples/fl_wrapper.cu:26-45
 *   - SEQFIX / TOPO       csrc/ext/causal_bp.cu:186-304 (orders), 583-722 (waves),
 *                         446-477, 868-872 (per-wave launches); csrc/kernel/
 *                         causal_bp_seq.cu:13-23, 34-83, 113-147, 149-459
 *     The SEQFIX order comes from std::shuffle with std::mt19937, restated
 *     from libstdc++ 11/12 (the reference pins gxx 12.4.0,
 *     environment.yml:74) and checked against this compiler's std::shuffle.
 * Damping is not in the reference; it is defined here as
 *   m_fv <- (1-d)*m_new + d*m_fv_old   (both states, after normalization)
 * and is skipped entirely when d == 0, so d == 0 is the reference.
 *
 * Parity pins (tests/): every f->v code against brute-force sum-product
 * enumeration; products/v->f against brute-force leave-one-out; whole-graph
 * BP on trees against exact joint enumeration; graph ordering and clamping
 * against the reference's own graph model compiled from its sources
 * (oracle/_ref).  The adjoint (orc_grad) is pinned by central finite
 * differences of orc_loss.  The sequential schedules are pinned by tree
 * exactness, by agreement with the PARALL fixed point on loopy graphs, by
 * the maintained products matching the messages' products, and by wave
 * splits worked by hand from the reference's conflict rule.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <queue>
#include <random>
#include <set>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

int fail(const std::string &msg) {
    g_err = msg;
    return -1;
}

/* csrc/kernel/causal_bp_parall.cu:14-18 */
inline void scale_parall(double &x, double &y) {
    double m = std::fmax(x, y);
    x /= m;
    y /= m;
}

/* csrc/kernel/causal_bp_seq.cu:13-23: the sequential kernels' scale */
inline void scale_seq(double &x, double &y) {
    if (x < y) {
        x = x / y;
        y = 1.0;
    } else if (y < x) {
        y = y / x;
        x = 1.0;
    } else {
        x = y = 1.0;
    }
}

template <bool kSeq>
inline void scale(double &x, double &y) {
    if (kSeq) scale_seq(x, y);
    else scale_parall(x, y);
}

struct VarMajorDesc {       /* Index6Real4 of causal_bp.cu:724-805 */
    int code;
    uint64_t src, start, end, head;
    double q, p, mask0, mask1;
};

struct FacMajorDesc {       /* Index3 of causal_bp.cu:806-813 */
    uint64_t ind, v, src;
};

struct SeqFvItem {          /* Index7Real4 of causal_bp.cu:650-700 */
    uint64_t ind, v;        /* var-major position written, its variable */
    VarMajorDesc d;
};

}  // namespace

struct orc_graph {
    uint64_t V = 0, F = 0, E = 0;
    std::vector<uint64_t> var_label;
    std::vector<uint64_t> row_ptr_vf;    /* F+1 */
    std::vector<uint64_t> edge_var;      /* E, factor-major */
    std::vector<uint64_t> row_ptr_fv;    /* V+1 */
    std::vector<uint64_t> head_slot;     /* F */
    std::vector<VarMajorDesc> fv_desc;   /* E, var-major */
    std::vector<FacMajorDesc> vf_desc;   /* E, factor-major */
    /* state */
    std::vector<double> m_fv0, m_fv1, m_vf0, m_vf1;
    std::vector<double> prod0, prod1;
    std::vector<uint64_t> nz0, nz1;
    std::vector<double> old_beliefs;     /* value-initialised to 0.0, causal_bp.cu:371 */
    uint64_t iters = 0;
    /* nbV(i)[k] of var-major edge e = row_ptr_fv[i] + k: its factor and variable */
    std::vector<uint64_t> vm_fac, vm_var;
    /* sequential schedule (empty waves = PARALL) */
    bool seq = false;
    std::vector<uint64_t> seq_order;                 /* _updateSeq, var-major positions */
    std::vector<std::vector<FacMajorDesc>> wave_vf;  /* _updateSeqVF */
    std::vector<std::vector<SeqFvItem>> wave_fv;     /* _updateSeqFV */
};

struct orc_flist {
    std::vector<int32_t> type;
    std::vector<uint64_t> head, body_off, body;
    std::vector<double> p, q, mask0, mask1;
    std::vector<uint8_t> head_clamped;
};

namespace {

/* vars() of a factor: VarSet(body) + head, sorted and unique
 * (csrc/ext/causal_factor.cpp:12-29, csrc/lbp/smallset.cpp:14-21). */
void factor_vars(const orc_factor_list *fl, uint64_t I, std::vector<uint64_t> &out) {
    out.clear();
    for (uint64_t k = fl->body_off[I]; k < fl->body_off[I + 1]; k++) out.push_back(fl->body[k]);
    out.push_back(fl->head[I]);
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}

/* f->v message for one var-major edge: the body of
 * calcMessageFVFusedAllKernel, csrc/kernel/causal_bp_parall.cu:102-397
 * (kSeq = false), or of calcMessageFVFusedKernel,
 * csrc/kernel/causal_bp_seq.cu:160-452 (kSeq = true), which is the same
 * arithmetic with the sequential scale().  vf0/vf1 point at the
 * factor-major message arrays. */
template <bool kSeq = false>
void fv_edge(const VarMajorDesc &d, const double *vf0, const double *vf1,
             double &out0, double &out1) {
    const uint64_t src_ind = d.src, h = d.head;
    const uint64_t start_ind = d.start, end_ind = d.end;
    const double p0 = d.q, p1 = d.p, mask0 = d.mask0, mask1 = d.mask1;
    const int type = d.code;
    double marg0 = 0, marg1 = 0;
    if (type == 0) {                                    /* 110-112 */
        marg0 = 1 - p1;
        marg1 = p1;
    } else if (type == 1 || type == 3) {                /* 113-151, 191-239 */
        double prod_1 = 1.0, prod_10 = 1.0;
        if (end_ind - start_ind == 1) {
            marg0 = 1 - p1;
            marg1 = p1;
        } else if (end_ind - start_ind == 2) {
            uint64_t j = start_ind;
            if (start_ind == src_ind) j++;
            double m_0 = vf0[j], m_1 = vf1[j];
            marg0 = (1 - p1) * m_1 + (1 - p0) * m_0;
            marg1 = p1 * m_1 + p0 * m_0;
        } else {
            double e1 = 0;
            for (uint64_t j = start_ind; j < end_ind; j++) {
                if (j != src_ind) {
                    double m_0 = vf0[j], m_1 = vf1[j];
                    double a0 = m_0 + m_1, a1 = m_1;
                    prod_1 *= a1;
                    prod_10 *= a0;
                    scale<kSeq>(prod_1, prod_10);
                    double delta = m_0;
                    if (a1 != 0 && a0 == a1 && delta != 0) e1 += delta / a1;
                }
            }
            marg0 = (1 - p1) * prod_1 + (1 - p0) * (e1 * prod_10 + (prod_10 - prod_1));
            marg1 = p1 * prod_1 + p0 * (e1 * prod_10 + (prod_10 - prod_1));
        }
        if (type == 3) {
            marg0 *= mask0;
            marg1 *= mask1;
        }
        double marg = marg0 + marg1;
        if (marg != 0) {
            marg0 = marg0 / marg;
            marg1 = marg1 / marg;
        } else if (type == 3 && mask0 == 0) {
            marg0 = 0.0; marg1 = 1.0;
        } else if (type == 3 && mask1 == 0) {
            marg0 = 1.0; marg1 = 0.0;
        } else {
            marg0 = 0.5; marg1 = 0.5;
        }
    } else if (type == 2 || type == 4) {                /* 152-190, 240-288 */
        double prod_0 = 1.0, prod_10 = 1.0;
        if (end_ind - start_ind == 1) {
            marg0 = p1;
            marg1 = 1 - p1;
        } else if (end_ind - start_ind == 2) {
            uint64_t j = start_ind;
            if (start_ind == src_ind) j++;
            double m_0 = vf0[j], m_1 = vf1[j];
            marg0 = p1 * m_0 + p0 * m_1;
            marg1 = (1 - p1) * m_0 + (1 - p0) * m_1;
        } else {
            double e1 = 0;
            for (uint64_t j = start_ind; j < end_ind; j++) {
                if (j != src_ind) {
                    double m_0 = vf0[j], m_1 = vf1[j];
                    double a0 = m_0 + m_1, a1 = m_0;
                    prod_0 *= a1;
                    prod_10 *= a0;
                    scale<kSeq>(prod_0, prod_10);
                    double delta = m_1;
                    if (a1 != 0 && a0 == a1 && delta != 0) e1 += delta / a1;
                }
            }
            marg0 = p1 * prod_0 + p0 * (e1 * prod_10 + (prod_10 - prod_0));
            marg1 = (1 - p1) * prod_0 + (1 - p0) * (e1 * prod_10 + (prod_10 - prod_0));
        }
        if (type == 4) {
            marg0 *= mask0;
            marg1 *= mask1;
        }
        double marg = marg0 + marg1;
        if (marg != 0) {
            marg0 = marg0 / marg;
            marg1 = marg1 / marg;
        } else if (type == 4 && mask0 == 0) {
            marg0 = 0.0; marg1 = 1.0;
        } else if (type == 4 && mask1 == 0) {
            marg0 = 1.0; marg1 = 0.0;
        } else {
            marg0 = 0.5; marg1 = 0.5;
        }
    } else if (type == 5 || type == 7) {                /* 289-314, 341-366 */
        double m0h = vf0[h], m1h = vf1[h];
        double prod_1, prod_10;
        if (type == 5) {
            prod_1 = (p1 - p0) * (m1h - m0h);
            prod_10 = p0 * m1h + (1 - p0) * m0h;
        } else {
            prod_1 = (p1 - p0) * (m1h * mask1 - m0h * mask0);
            prod_10 = p0 * m1h * mask1 + (1 - p0) * m0h * mask0;
        }
        scale<kSeq>(prod_1, prod_10);
        if (prod_10 != 0) {
            for (uint64_t j = start_ind; j < end_ind; j++) {
                if ((j != src_ind) && (j != h)) {
                    double m0 = vf0[j], m1 = vf1[j], m = m0 + m1;
                    prod_10 *= m;
                    prod_1 *= m1;
                    scale<kSeq>(prod_1, prod_10);
                }
            }
        }
        marg0 = prod_10;
        marg1 = prod_10 + prod_1;
        double marg = marg0 + marg1;
        if (marg != 0) {
            marg0 = marg0 / marg;
            marg1 = marg1 / marg;
        } else {
            marg0 = 0.5; marg1 = 0.5;
        }
    } else if (type == 6 || type == 8) {                /* 315-340, 367-393 */
        double m0h = vf0[h], m1h = vf1[h];
        double prod_0, prod_10;
        if (type == 6) {
            prod_0 = (p1 - p0) * (m0h - m1h);
            prod_10 = p0 * m0h + (1 - p0) * m1h;
        } else {
            prod_0 = (p1 - p0) * (m0h * mask0 - m1h * mask1);
            prod_10 = p0 * m0h * mask0 + (1 - p0) * m1h * mask1;
        }
        scale<kSeq>(prod_0, prod_10);
        if (prod_10 != 0) {
            for (uint64_t j = start_ind; j < end_ind; j++) {
                if ((j != src_ind) && (j != h)) {
                    double m0 = vf0[j], m1 = vf1[j], m = m0 + m1;
                    prod_10 *= m;
                    prod_0 *= m0;
                    scale<kSeq>(prod_0, prod_10);
                }
            }
        }
        marg0 = prod_0 + prod_10;
        marg1 = prod_10;
        double marg = marg0 + marg1;
        if (marg != 0) {
            marg0 = marg0 / marg;
            marg1 = marg1 / marg;
        } else {
            marg0 = 0.5; marg1 = 0.5;
        }
    }
    out0 = marg0;
    out1 = marg1;
}

/* calcProdAllKernel body, csrc/kernel/causal_bp_parall.cu:30-52 */
void prod_var(const double *fv0, const double *fv1, uint64_t start_ind, uint64_t end_ind,
              double &p0_out, double &p1_out, uint64_t &z0_out, uint64_t &z1_out) {
    double prod_0 = 1.0, prod_1 = 1.0;
    uint64_t zero_0 = 0, zero_1 = 0;
    for (uint64_t j = start_ind; j < end_ind; j++) {
        double m0 = fv0[j], m1 = fv1[j];
        if (m0 == 0) zero_0++; else prod_0 *= m0;
        if (m1 == 0) zero_1++; else prod_1 *= m1;
        scale_parall(prod_0, prod_1);
    }
    p0_out = prod_0;
    p1_out = prod_1;
    z0_out = zero_0;
    z1_out = zero_1;
}

/* calcMessageVFAllKernel body, csrc/kernel/causal_bp_parall.cu:68-90 */
void vf_edge(double prod_0, double prod_1, uint64_t zero_cnt_0, uint64_t zero_cnt_1,
             double m_src_0, double m_src_1, double &out0, double &out1) {
    if ((zero_cnt_0 >= 2) || (zero_cnt_0 == 1 && m_src_0 != 0)) {
        prod_0 = 0.0;
    } else if (m_src_0 != 0) {
        prod_0 = prod_0 / m_src_0;
    }
    if ((zero_cnt_1 >= 2) || (zero_cnt_1 == 1 && m_src_1 != 0)) {
        prod_1 = 0.0;
    } else if (m_src_1 != 0) {
        prod_1 = prod_1 / m_src_1;
    }
    if (prod_0 == 0 && prod_1 == 0) prod_0 = prod_1 = 0.5;
    double prod = prod_0 + prod_1;
    out0 = prod_0 / prod;
    out1 = prod_1 / prod;
}

/* getBeliefsVKernel (causal_bp_seq.cu:102-111) on both states, then
 * norm_functor (causal_bp.cu:124-128). */
inline double belief_of(double prod_0, double prod_1, uint64_t nz0, uint64_t nz1) {
    double b0 = prod_0 * (double)(!nz0);
    double b1 = prod_1 * (double)(!nz1);
    return b1 / (b0 + b1);
}

/* max with NaN propagation: a NaN difference stops the loop (NaN > tol is false). */
inline double nanmax(double a, double b) {
    if (std::isnan(a) || std::isnan(b)) return std::numeric_limits<double>::quiet_NaN();
    return a < b ? b : a;
}

void set_threads(int threads) {
#ifdef _OPENMP
    omp_set_num_threads(threads > 0 ? threads : 1);
#else
    (void)threads;
#endif
}

double belief_step(orc_graph *g);
double seq_sweep(orc_graph *g);

/* One PARALL sweep (calcNewMessageFusedAll, causal_bp.cu:409-444) plus the
 * belief/max-diff step of run() (causal_bp.cu:874-889); with a sequential
 * schedule set, one pass over its waves instead (causal_bp.cu:868-872). */
double sweep(orc_graph *g, double damping) {
    if (g->seq) return seq_sweep(g);
    const int64_t E = (int64_t)g->E, V = (int64_t)g->V;
    /* f -> v over var-major edges */
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; e++) {
        double o0, o1;
        fv_edge(g->fv_desc[e], g->m_vf0.data(), g->m_vf1.data(), o0, o1);
        if (damping != 0.0) {
            o0 = (1 - damping) * o0 + damping * g->m_fv0[e];
            o1 = (1 - damping) * o1 + damping * g->m_fv1[e];
        }
        g->m_fv0[e] = o0;
        g->m_fv1[e] = o1;
    }
    /* products over var-major edges */
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < V; v++) {
        prod_var(g->m_fv0.data(), g->m_fv1.data(), g->row_ptr_fv[v], g->row_ptr_fv[v + 1],
                 g->prod0[v], g->prod1[v], g->nz0[v], g->nz1[v]);
    }
    /* v -> f over factor-major edges */
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; e++) {
        const FacMajorDesc &d = g->vf_desc[e];
        double o0, o1;
        vf_edge(g->prod0[d.v], g->prod1[d.v], g->nz0[d.v], g->nz1[d.v],
                g->m_fv0[d.src], g->m_fv1[d.src], o0, o1);
        g->m_vf0[d.ind] = o0;
        g->m_vf1[d.ind] = o1;
    }
    (void)V;
    return belief_step(g);
}

/* calcBeliefsV (causal_bp.cu:483-515) from the products and zero counts,
 * then max |old - new| (dist_functor + maximum reduce with init 0.0,
 * causal_bp.cu:878-889); the new beliefs become the old ones. */
double belief_step(orc_graph *g) {
    const int64_t V = (int64_t)g->V;
    double maxdiff = 0.0;
#pragma omp parallel
    {
        double local = 0.0;
#pragma omp for schedule(static)
        for (int64_t v = 0; v < V; v++) {
            double b = belief_of(g->prod0[v], g->prod1[v], g->nz0[v], g->nz1[v]);
            double diff = std::fabs(g->old_beliefs[v] - b);
            local = nanmax(local, diff);
            g->old_beliefs[v] = b;
        }
#pragma omp critical
        maxdiff = nanmax(maxdiff, local);
    }
    return maxdiff;
}

/* ---- sequential schedules (updates = SEQFIX / TOPO) ---------------------- */

/* calcMessageVFKernel body, csrc/kernel/causal_bp_seq.cu:124-145: the
 * PARALL v->f message without its (0, 0) guard. */
void vf_edge_seq(double prod_0, double prod_1, uint64_t zero_cnt_0, uint64_t zero_cnt_1,
                 double m_src_0, double m_src_1, double &out0, double &out1) {
    if ((zero_cnt_0 >= 2) || (zero_cnt_0 == 1 && m_src_0 != 0)) {
        prod_0 = 0.0;
    } else if (m_src_0 != 0) {
        prod_0 = prod_0 / m_src_0;
    }
    if ((zero_cnt_1 >= 2) || (zero_cnt_1 == 1 && m_src_1 != 0)) {
        prod_1 = 0.0;
    } else if (m_src_1 != 0) {
        prod_1 = prod_1 / m_src_1;
    }
    double prod = prod_0 + prod_1;
    out0 = prod_0 / prod;
    out1 = prod_1 / prod;
}

/* update_message, csrc/kernel/causal_bp_seq.cu:34-83: divides the old f->v
 * message out of the variable's product, multiplies the new one in (zeros
 * counted, not multiplied; unsigned counts wrap as size_t does), rescales
 * and stores the new message. */
void update_message(orc_graph *g, uint64_t ind, uint64_t v, double marg0, double marg1) {
    double old_msg_0 = g->m_fv0[ind], old_msg_1 = g->m_fv1[ind];
    double prod_0 = g->prod0[v], prod_1 = g->prod1[v];
    uint64_t zero_cnt_0 = g->nz0[v], zero_cnt_1 = g->nz1[v];
    if (old_msg_0 == 0) zero_cnt_0--; else prod_0 /= old_msg_0;
    if (marg0 == 0) zero_cnt_0++; else prod_0 *= marg0;
    if (old_msg_1 == 0) zero_cnt_1--; else prod_1 /= old_msg_1;
    if (marg1 == 0) zero_cnt_1++; else prod_1 *= marg1;
    scale_seq(prod_0, prod_1);
    g->prod0[v] = prod_0;
    g->prod1[v] = prod_1;
    g->nz0[v] = zero_cnt_0;
    g->nz1[v] = zero_cnt_1;
    g->m_fv0[ind] = marg0;
    g->m_fv1[ind] = marg1;
}

/* One pass over the waves (causal_bp.cu:868-872 with calcNewMessageFused,
 * 446-477): per wave the v->f kernel, then the fused f->v kernel, then the
 * belief step.  Items of one kernel touch disjoint outputs and read nothing
 * another item of the same kernel writes (that is what the wave split
 * guarantees), so running them one after another here gives the parallel
 * kernels' results. */
double seq_sweep(orc_graph *g) {
    for (size_t w = 0; w < g->wave_vf.size(); w++) {
        for (const FacMajorDesc &d : g->wave_vf[w]) {
            double o0, o1;
            vf_edge_seq(g->prod0[d.v], g->prod1[d.v], g->nz0[d.v], g->nz1[d.v], g->m_fv0[d.src],
                        g->m_fv1[d.src], o0, o1);
            g->m_vf0[d.ind] = o0;
            g->m_vf1[d.ind] = o1;
        }
        for (const SeqFvItem &it : g->wave_fv[w]) {
            double o0, o1;
            fv_edge<true>(it.d, g->m_vf0.data(), g->m_vf1.data(), o0, o1);
            update_message(g, it.ind, it.v, o0, o1);
        }
    }
    return belief_step(g);
}

/* libstdc++'s uniform_int_distribution<unsigned long>::operator() over
 * [a, b] driven by std::mt19937 (a 32-bit engine): Lemire's nearly
 * divisionless method ("Fast Random Integer Generation in an Interval",
 * ACM TOMACS 29(1), 2019) as libstdc++ 11/12 apply it to a 32-bit
 * generator (bits/uniform_int_dist.h).  Ranges of 2^32 or more are not
 * needed here (factor counts are below 2^32). */
uint64_t uniform_between(std::mt19937 &g, uint64_t a, uint64_t b) {
    const uint64_t urange = b - a;
    if (urange >= 0xffffffffull) return a + (uint64_t)g();   /* not reached for < 2^32 - 1 items */
    const uint32_t range = (uint32_t)(urange + 1);
    uint64_t product = (uint64_t)(uint32_t)g() * (uint64_t)range;
    uint32_t low = (uint32_t)product;
    if (low < range) {
        const uint32_t threshold = (uint32_t)(0u - range) % range;
        while (low < threshold) {
            product = (uint64_t)(uint32_t)g() * (uint64_t)range;
            low = (uint32_t)product;
        }
    }
    return a + (product >> 32);
}

/* std::shuffle as libstdc++ 11/12 implement it (bits/stl_algo.h): a
 * Fisher-Yates pass from the front; while n^2 fits the generator's range,
 * two swap positions come from one draw over (i+1)(i+2) values. */
void shuffle_like_libstdcxx(std::vector<uint64_t> &v, std::mt19937 &g) {
    const uint64_t n = v.size();
    if (n == 0) return;
    const uint64_t urngrange = 0xffffffffull;
    if (urngrange / n >= n) {
        uint64_t i = 1;
        if (n % 2 == 0) {
            std::swap(v[i], v[uniform_between(g, 0, 1)]);
            i++;
        }
        while (i != n) {
            const uint64_t r = i + 1;
            const uint64_t x = uniform_between(g, 0, r * (r + 1) - 1);
            std::swap(v[i], v[x / (r + 1)]);
            i++;
            std::swap(v[i], v[x % (r + 1)]);
            i++;
        }
        return;
    }
    for (uint64_t i = 1; i < n; i++) std::swap(v[i], v[uniform_between(g, 0, i)]);
}

}  // namespace

struct orc_rng {
    std::mt19937 g;
};

namespace {

/* putUpdateSeqToKernelFused, csrc/ext/causal_bp.cu:583-722: walks the order
 * and closes the current wave when the next f->v update I -> i (a) needs a
 * v->f message j -> I (j != i) that this wave has not scheduled while this
 * wave already updates some f->v message J -> j (J != I) -- with
 * respect_preorder, only one ranked before I -> i -- or (b) targets a
 * variable this wave already updates.  Each update then adds the v->f
 * messages j -> I of its factor that the wave does not have yet. */
int build_waves(orc_graph *g, const uint64_t *order, uint64_t n, bool respect_preorder) {
    for (uint64_t k = 0; k < n; k++)
        if (order[k] >= g->E) return fail("update order names an edge out of range");
    /* _edgePreorderRank (causal_bp.cu:256-302): 0 for edges not in the
     * order, else the last position holding the edge */
    std::vector<uint64_t> rank(g->E, 0);
    for (uint64_t k = 0; k < n; k++) rank[order[k]] = k;
    g->seq_order.assign(order, order + n);
    g->wave_vf.clear();
    g->wave_fv.clear();
    std::vector<FacMajorDesc> seq_vf;
    std::vector<SeqFvItem> seq_fv;
    std::set<std::pair<uint64_t, uint64_t>> fv_set, vf_set;
    std::set<uint64_t> v_set;
    for (uint64_t k = 0; k < n; k++) {
        const uint64_t e = order[k];
        const uint64_t i = g->vm_var[e], I = g->vm_fac[e];
        const uint64_t dual = e - g->row_ptr_fv[i];
        bool conflict = false;
        for (uint64_t s = g->row_ptr_vf[I]; s < g->row_ptr_vf[I + 1] && !conflict; s++) {
            const uint64_t j = g->edge_var[s];
            if (j == i) continue;
            if (vf_set.count({I, s - g->row_ptr_vf[I]})) continue;
            for (uint64_t e2 = g->row_ptr_fv[j]; e2 < g->row_ptr_fv[j + 1]; e2++) {
                if (g->vm_fac[e2] == I) continue;
                if (fv_set.count({j, e2 - g->row_ptr_fv[j]}) && (!respect_preorder || rank[e2] < rank[e])) {
                    conflict = true;
                    break;
                }
            }
        }
        if (!conflict && v_set.count(i)) conflict = true;
        if (conflict) {
            g->wave_vf.push_back(seq_vf);
            g->wave_fv.push_back(seq_fv);
            v_set.clear();
            fv_set.clear();
            vf_set.clear();
            seq_vf.clear();
            seq_fv.clear();
        }
        v_set.insert(i);
        fv_set.insert({i, dual});
        seq_fv.push_back({e, i, g->fv_desc[e]});
        for (uint64_t s = g->row_ptr_vf[I]; s < g->row_ptr_vf[I + 1]; s++) {
            const uint64_t j = g->edge_var[s];
            if (j == i) continue;
            if (vf_set.insert({I, s - g->row_ptr_vf[I]}).second) seq_vf.push_back(g->vf_desc[s]);
        }
    }
    g->wave_vf.push_back(seq_vf);
    g->wave_fv.push_back(seq_fv);
    g->seq = true;
    return 0;
}

/* the var-major position of Edge(i, i.dual) for factor-major slot s */
inline uint64_t slot_edge(const orc_graph *g, uint64_t s) { return g->vf_desc[s].src; }

/* buildFixedUpdateSeq, csrc/ext/causal_bp.cu:186-201: the factor indices
 * shuffled by the generator, then each factor's edges in nbF order. */
void seqfix_order(const orc_graph *g, std::mt19937 &rng, std::vector<uint64_t> &out) {
    std::vector<uint64_t> numbers(g->F);
    for (uint64_t I = 0; I < g->F; I++) numbers[I] = I;
    shuffle_like_libstdcxx(numbers, rng);
    out.clear();
    for (uint64_t I : numbers)
        for (uint64_t s = g->row_ptr_vf[I]; s < g->row_ptr_vf[I + 1]; s++) out.push_back(slot_edge(g, s));
}

/* buildTopoUpdateSeq, csrc/ext/causal_bp.cu:203-304: breadth-first order of
 * the undirected factor graph (variables are nodes 0..V-1, factors V..V+F-1,
 * neighbours visited in ascending node order, roots in ascending order),
 * then the edges sorted by (later endpoint's rank, earlier endpoint's rank,
 * factor, variable, dual). */
void topo_order(const orc_graph *g, std::vector<uint64_t> &out) {
    const uint64_t V = g->V, F = g->F, nodes = V + F;
    std::vector<std::vector<uint64_t>> adj(nodes);
    for (uint64_t I = 0; I < F; I++) {
        for (uint64_t s = g->row_ptr_vf[I]; s < g->row_ptr_vf[I + 1]; s++) {
            adj[V + I].push_back(g->edge_var[s]);
            adj[g->edge_var[s]].push_back(V + I);
        }
    }
    for (auto &a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    std::vector<uint64_t> rank(nodes, nodes);
    uint64_t next = 0;
    std::queue<uint64_t> ready;
    for (uint64_t root = 0; root < nodes; root++) {
        if (rank[root] != nodes) continue;
        ready.push(root);
        rank[root] = next++;
        while (!ready.empty()) {
            const uint64_t x = ready.front();
            ready.pop();
            for (uint64_t y : adj[x]) {
                if (rank[y] != nodes) continue;
                rank[y] = next++;
                ready.push(y);
            }
        }
    }
    struct Ranked {
        uint64_t target, source, factor, var, dual, edge;
    };
    std::vector<Ranked> r;
    r.reserve(g->E);
    for (uint64_t I = 0; I < F; I++) {
        for (uint64_t s = g->row_ptr_vf[I]; s < g->row_ptr_vf[I + 1]; s++) {
            const uint64_t i = g->edge_var[s], e = slot_edge(g, s);
            const uint64_t a = rank[V + I], b = rank[i];
            r.push_back({std::max(a, b), std::min(a, b), I, i, e - g->row_ptr_fv[i], e});
        }
    }
    std::stable_sort(r.begin(), r.end(), [](const Ranked &x, const Ranked &y) {
        if (x.target != y.target) return x.target < y.target;
        if (x.source != y.source) return x.source < y.source;
        if (x.factor != y.factor) return x.factor < y.factor;
        if (x.var != y.var) return x.var < y.var;
        return x.dual < y.dual;
    });
    out.clear();
    for (const Ranked &x : r) out.push_back(x.edge);
}

}  // namespace

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

/* CausalFactorGraph(vector) (causal_fg.cpp:50-68) + constructGraph (28-48)
 * + BipartiteGraph::addEdge (bipgraph.cpp:24-42) + CausalBP::construct
 * (causal_bp.cu:310-407) + putParallUpdateSeqToKernelFused (724-817). */
int orc_build(const orc_factor_list *fl, orc_graph **out) {
    if (!fl || !out) return fail("null argument");
    const uint64_t F = fl->n_factors;
    std::vector<std::vector<uint64_t>> fvars(F);
    std::set<uint64_t> labels;
    uint64_t E = 0;
    for (uint64_t I = 0; I < F; I++) {
        if (fl->type[I] < 0 || fl->type[I] > 2) return fail("bad factor type");
        factor_vars(fl, I, fvars[I]);
        labels.insert(fvars[I].begin(), fvars[I].end());
        E += fvars[I].size();
    }
    orc_graph *g = new orc_graph();
    g->F = F;
    g->E = E;
    g->V = labels.size();
    g->var_label.assign(labels.begin(), labels.end());
    std::map<uint64_t, uint64_t> index;
    for (uint64_t i = 0; i < g->V; i++) index[g->var_label[i]] = i;

    /* nbF(I): vars of I in ascending label; nbV(i): factors in ascending
     * index, each with its dual position. */
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> nbV(g->V);  /* (I, dual) */
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> nbF(F);     /* (i, dual) */
    for (uint64_t I = 0; I < F; I++) {
        for (uint64_t lab : fvars[I]) {
            uint64_t i = index[lab];
            uint64_t it1 = nbV[i].size(), it2 = nbF[I].size();
            nbV[i].push_back({I, it2});
            nbF[I].push_back({i, it1});
        }
    }
    g->row_ptr_vf.assign(F + 1, 0);
    for (uint64_t I = 0; I < F; I++) g->row_ptr_vf[I + 1] = g->row_ptr_vf[I] + nbF[I].size();
    g->row_ptr_fv.assign(g->V + 1, 0);
    for (uint64_t i = 0; i < g->V; i++) g->row_ptr_fv[i + 1] = g->row_ptr_fv[i] + nbV[i].size();
    g->edge_var.resize(E);
    g->head_slot.resize(F);
    for (uint64_t I = 0; I < F; I++) {
        for (uint64_t k = 0; k < nbF[I].size(); k++) {
            uint64_t i = nbF[I][k].first;
            g->edge_var[g->row_ptr_vf[I] + k] = i;
            if (g->var_label[i] == fl->head[I]) g->head_slot[I] = g->row_ptr_vf[I] + k;
        }
    }
    g->fv_desc.resize(E);
    g->vm_fac.resize(E);
    g->vm_var.resize(E);
    for (uint64_t i = 0; i < g->V; i++) {
        for (uint64_t k = 0; k < nbV[i].size(); k++) {
            uint64_t I = nbV[i][k].first, dual = nbV[i][k].second;
            g->vm_fac[g->row_ptr_fv[i] + k] = I;
            g->vm_var[g->row_ptr_fv[i] + k] = i;
            VarMajorDesc &d = g->fv_desc[g->row_ptr_fv[i] + k];
            d.src = g->row_ptr_vf[I] + dual;
            d.start = g->row_ptr_vf[I];
            d.end = g->row_ptr_vf[I + 1];
            d.head = g->head_slot[I];
            d.q = fl->q[I];
            d.p = fl->p[I];
            d.mask0 = fl->mask0[I];
            d.mask1 = fl->mask1[I];
            bool is_head = fl->head[I] == g->var_label[i];
            bool clamped = fl->head_clamped[I] != 0;
            int t = fl->type[I];
            if (t == 0) d.code = 0;
            else if (t == 1) d.code = clamped ? (is_head ? 3 : 7) : (is_head ? 1 : 5);
            else d.code = clamped ? (is_head ? 4 : 8) : (is_head ? 2 : 6);
        }
    }
    g->vf_desc.resize(E);
    for (uint64_t I = 0; I < F; I++) {
        for (uint64_t k = 0; k < nbF[I].size(); k++) {
            uint64_t i = nbF[I][k].first, dual = nbF[I][k].second;
            g->vf_desc[g->row_ptr_vf[I] + k] = {g->row_ptr_vf[I] + k, i, g->row_ptr_fv[i] + dual};
        }
    }
    g->m_fv0.assign(E, 0.0); g->m_fv1.assign(E, 0.0);
    g->m_vf0.assign(E, 0.0); g->m_vf1.assign(E, 0.0);
    g->prod0.assign(g->V, 0.0); g->prod1.assign(g->V, 0.0);
    g->nz0.assign(g->V, 0); g->nz1.assign(g->V, 0);
    g->old_beliefs.assign(g->V, 0.0);
    *out = g;
    return 0;
}

void orc_free(orc_graph *g) { delete g; }
uint64_t orc_nvars(const orc_graph *g) { return g->V; }
uint64_t orc_nfactors(const orc_graph *g) { return g->F; }
uint64_t orc_nedges(const orc_graph *g) { return g->E; }
uint64_t orc_iters(const orc_graph *g) { return g->iters; }

int orc_dump(const orc_graph *g, uint64_t *var_label, uint64_t *fac_off, uint64_t *edge_var,
             uint64_t *var_off, uint64_t *var_src, uint64_t *head_slot, int32_t *code) {
    if (var_label) std::copy(g->var_label.begin(), g->var_label.end(), var_label);
    if (fac_off) std::copy(g->row_ptr_vf.begin(), g->row_ptr_vf.end(), fac_off);
    if (edge_var) std::copy(g->edge_var.begin(), g->edge_var.end(), edge_var);
    if (var_off) std::copy(g->row_ptr_fv.begin(), g->row_ptr_fv.end(), var_off);
    if (head_slot) std::copy(g->head_slot.begin(), g->head_slot.end(), head_slot);
    for (uint64_t e = 0; e < g->E; e++) {
        if (var_src) var_src[e] = g->fv_desc[e].src;
        if (code) code[e] = g->fv_desc[e].code;
    }
    return 0;
}

/* CausalBP::init, causal_bp.cu:824-856 (old beliefs are not reset). */
int orc_init(orc_graph *g, int logdomain) {
    double c = logdomain ? 0.0 : 1.0;
    std::fill(g->m_fv0.begin(), g->m_fv0.end(), c);
    std::fill(g->m_fv1.begin(), g->m_fv1.end(), c);
    std::fill(g->m_vf0.begin(), g->m_vf0.end(), c);
    std::fill(g->m_vf1.begin(), g->m_vf1.end(), c);
    std::fill(g->prod0.begin(), g->prod0.end(), c);
    std::fill(g->prod1.begin(), g->prod1.end(), c);
    std::fill(g->nz0.begin(), g->nz0.end(), 0);
    std::fill(g->nz1.begin(), g->nz1.end(), 0);
    return 0;
}

/* CausalBP::run(), causal_bp.cu:858-910 (maxtime = infinity). */
int orc_run(orc_graph *g, double tol, uint64_t maxiter, double damping, int threads,
            uint64_t *iters_out, double *maxdiff_out) {
    if (g->seq && damping != 0.0) return fail("damping is defined for PARALL updates only");
    set_threads(threads);
    double maxDiff = INFINITY;
    for (; g->iters < maxiter && maxDiff > tol; g->iters++) {
        maxDiff = sweep(g, damping);
    }
    if (iters_out) *iters_out = g->iters;
    if (maxdiff_out) *maxdiff_out = maxDiff;
    return 0;
}

/* CausalBP::run(tol, maxIters, histLength), causal_bp.cu:912-1004 for the
 * PARALL schedule: exactly maxIters sweeps, then beliefs once. */
int orc_run_fixed(orc_graph *g, uint64_t n_iters, double damping, int threads) {
    if (g->seq && damping != 0.0) return fail("damping is defined for PARALL updates only");
    set_threads(threads);
    for (uint64_t k = 0; k < n_iters; k++, g->iters++) {
        (void)sweep(g, damping);
    }
    return 0;
}

int orc_sweep(orc_graph *g, double damping, int threads, double *maxdiff_out) {
    if (g->seq && damping != 0.0) return fail("damping is defined for PARALL updates only");
    set_threads(threads);
    double d = sweep(g, damping);
    if (maxdiff_out) *maxdiff_out = d;
    return 0;
}

/* CausalBP::run(tol, minIters, maxIters, histLength), causal_bp.cu:1006-1143
 * with maxtime = infinity, and HostBeliefHistory (28-78): per sweep the
 * beliefs, the count of variables whose belief moved by more than tol and
 * a history push of every finite belief; the exit tests in the reference's
 * order at the top of each pass.  lowpass gets the beliefs on ALL_CONVERGED,
 * else the mean of each variable's last histLength finite beliefs, oldest
 * first, divided once.  reason: 0 all converged, 1 big fraction converged,
 * 2 diverged. */
int orc_run_hist(orc_graph *g, double tol, uint64_t min_iters, uint64_t max_iters, uint64_t hist_len,
                 double damping, double *yet_out, double *lowpass, int *reason_out, double *maxdiff_out) {
    if (hist_len == 0) return fail("belief history length must be positive");
    if (g->seq && damping != 0.0) return fail("damping is defined for PARALL updates only");
    const uint64_t V = g->V;
    std::vector<double> values(V * hist_len, 0.0);
    std::vector<uint64_t> next(V, 0), count(V, 0);
    uint64_t num_iters = 0;
    double max_diff = INFINITY, yet = 1.0, node_frac = 0.0;
    int reason = 2;
    for (;; num_iters++, g->iters++) {
        if (num_iters >= min_iters)
            node_frac = (double)(uint64_t)(num_iters - min_iters) / (double)(uint64_t)(max_iters - min_iters);
        if (max_diff <= tol) {
            reason = 0;
            break;
        } else if (num_iters > min_iters && yet < node_frac) {
            reason = 1;
            break;
        } else if (num_iters > max_iters) {
            reason = 2;
            break;
        }
        const std::vector<double> old = g->old_beliefs;
        const double diff = sweep(g, damping);
        /* maxDiff = std::max(diff, -INFINITY) */
        max_diff = (diff < -INFINITY) ? -INFINITY : diff;
        uint64_t non_converged = 0;
        for (uint64_t v = 0; v < V; v++) {
            const double b = g->old_beliefs[v];
            non_converged += std::fabs(old[v] - b) > tol;
            const int c = std::fpclassify(b);
            if (c != FP_NORMAL && c != FP_SUBNORMAL && c != FP_ZERO) continue;
            uint64_t slot = next[v];
            values[slot * V + v] = b;
            if (++slot == hist_len) slot = 0;
            next[v] = slot;
            if (count[v] < hist_len) ++count[v];
        }
        yet = (double)non_converged / (double)V;
    }
    if (lowpass) {
        for (uint64_t v = 0; v < V; v++) {
            if (reason == 0) {
                lowpass[v] = g->old_beliefs[v];
                continue;
            }
            const uint64_t c = count[v];
            uint64_t slot = c == hist_len ? next[v] : 0;
            double sum = 0.0;
            for (uint64_t k = 0; k < c; k++) {
                sum += values[slot * V + v];
                if (++slot == hist_len) slot = 0;
            }
            if (c > 0) sum /= (double)c;
            lowpass[v] = sum;
        }
    }
    if (yet_out) *yet_out = yet;
    if (reason_out) *reason_out = reason;
    if (maxdiff_out) *maxdiff_out = max_diff;
    return 0;
}

int orc_beliefs(const orc_graph *g, double *out) {
    std::copy(g->old_beliefs.begin(), g->old_beliefs.end(), out);
    return 0;
}

int orc_messages(const orc_graph *g, double *fv0, double *fv1, double *vf0, double *vf1) {
    if (fv0) std::copy(g->m_fv0.begin(), g->m_fv0.end(), fv0);
    if (fv1) std::copy(g->m_fv1.begin(), g->m_fv1.end(), fv1);
    if (vf0) std::copy(g->m_vf0.begin(), g->m_vf0.end(), vf0);
    if (vf1) std::copy(g->m_vf1.begin(), g->m_vf1.end(), vf1);
    return 0;
}

int orc_set_messages(orc_graph *g, const double *fv0, const double *fv1, const double *vf0,
                     const double *vf1) {
    if (fv0) std::copy(fv0, fv0 + g->E, g->m_fv0.begin());
    if (fv1) std::copy(fv1, fv1 + g->E, g->m_fv1.begin());
    if (vf0) std::copy(vf0, vf0 + g->E, g->m_vf0.begin());
    if (vf1) std::copy(vf1, vf1 + g->E, g->m_vf1.begin());
    return 0;
}


/* ---- sequential schedules ------------------------------------------------ */

orc_rng *orc_rng_new(uint32_t seed) {
    orc_rng *r = new orc_rng();
    r->g.seed(seed);
    return r;
}
void orc_rng_free(orc_rng *r) { delete r; }
uint32_t orc_rng_next(orc_rng *r) { return (uint32_t)r->g(); }

int orc_shuffle(orc_rng *r, uint64_t n, uint64_t *out) {
    std::vector<uint64_t> v(n);
    for (uint64_t k = 0; k < n; k++) v[k] = k;
    shuffle_like_libstdcxx(v, r->g);
    std::copy(v.begin(), v.end(), out);
    return 0;
}

/* std::shuffle itself (this compiler's libstdc++) on the vector<int> the
 * reference shuffles, for checking the restatement above */
int orc_std_shuffle(uint32_t seed, uint64_t n, uint64_t *out) {
    std::mt19937 g(seed);
    std::vector<int> v(n);
    for (uint64_t k = 0; k < n; k++) v[k] = (int)k;
    std::shuffle(v.begin(), v.end(), g);
    for (uint64_t k = 0; k < n; k++) out[k] = (uint64_t)v[k];
    return 0;
}

int orc_order_seqfix(const orc_graph *g, orc_rng *r, uint64_t *out) {
    std::vector<uint64_t> o;
    seqfix_order(g, r->g, o);
    std::copy(o.begin(), o.end(), out);
    return 0;
}

int orc_order_topo(const orc_graph *g, uint64_t *out) {
    std::vector<uint64_t> o;
    topo_order(g, o);
    std::copy(o.begin(), o.end(), out);
    return 0;
}

int orc_seq_set(orc_graph *g, const uint64_t *order, uint64_t n, int respect_preorder) {
    if (n && !order) return fail("null order");
    return build_waves(g, order, n, respect_preorder != 0);
}

int orc_seq_clear(orc_graph *g) {
    g->seq = false;
    g->seq_order.clear();
    g->wave_vf.clear();
    g->wave_fv.clear();
    return 0;
}

int orc_seq_info(const orc_graph *g, uint64_t *n_waves, uint64_t *n_fv, uint64_t *n_vf) {
    uint64_t a = 0, b = 0;
    for (const auto &w : g->wave_fv) a += w.size();
    for (const auto &w : g->wave_vf) b += w.size();
    if (n_waves) *n_waves = g->wave_fv.size();
    if (n_fv) *n_fv = a;
    if (n_vf) *n_vf = b;
    return 0;
}

int orc_seq_waves(const orc_graph *g, uint64_t *fv_off, uint64_t *fv_edge, uint64_t *vf_off,
                  uint64_t *vf_slot) {
    uint64_t a = 0, b = 0;
    for (size_t w = 0; w < g->wave_fv.size(); w++) {
        if (fv_off) fv_off[w] = a;
        if (vf_off) vf_off[w] = b;
        for (const SeqFvItem &it : g->wave_fv[w]) {
            if (fv_edge) fv_edge[a] = it.ind;
            a++;
        }
        for (const FacMajorDesc &d : g->wave_vf[w]) {
            if (vf_slot) vf_slot[b] = d.ind;
            b++;
        }
    }
    if (fv_off) fv_off[g->wave_fv.size()] = a;
    if (vf_off) vf_off[g->wave_vf.size()] = b;
    return 0;
}

/* Checker of a dependency-level split (not a restatement of the
 * reference): one sweep that runs the listed items one after another --
 * kind 0: the v->f message of factor-major slot item[x]; kind 1: the f->v
 * update of var-major edge item[x] -- then the belief step.  Listed level
 * by level in any order within a level, the items must give the wave
 * execution's bits. */
int orc_seq_sweep_items(orc_graph *g, uint64_t n, const uint8_t *kind, const uint64_t *item, double *maxdiff_out) {
    for (uint64_t x = 0; x < n; x++) {
        if (item[x] >= g->E) return fail("item out of range");
        double o0, o1;
        if (kind[x] == 0) {
            const FacMajorDesc &d = g->vf_desc[item[x]];
            vf_edge_seq(g->prod0[d.v], g->prod1[d.v], g->nz0[d.v], g->nz1[d.v], g->m_fv0[d.src],
                        g->m_fv1[d.src], o0, o1);
            g->m_vf0[d.ind] = o0;
            g->m_vf1[d.ind] = o1;
        } else {
            const uint64_t e = item[x];
            fv_edge<true>(g->fv_desc[e], g->m_vf0.data(), g->m_vf1.data(), o0, o1);
            update_message(g, e, g->vm_var[e], o0, o1);
        }
    }
    const double d = belief_step(g);
    if (maxdiff_out) *maxdiff_out = d;
    return 0;
}

/* Checker of the engine's execution items (not a restatement of the
 * reference): one sweep of items in the listed order, then the belief step.
 *   kind 0: the v->f message of factor-major slot a[x]
 *   kind 1: the f->v update of var-major edge a[x]
 *   kind 2: factor a[x]: first the v->f messages of its slots in mask c[x],
 *           from the current states, then the f->v updates of its slots in
 *           mask b[x] (bit k = the factor's k-th slot)
 * Listed level by level in any order within a level, the items must give
 * the wave execution's bits. */
int orc_seq_sweep_exec(orc_graph *g, uint64_t n, const uint8_t *kind, const uint32_t *a, const uint32_t *b,
                       const uint32_t *c, double *maxdiff_out) {
    for (uint64_t x = 0; x < n; x++) {
        double o0, o1;
        if (kind[x] == 0 || kind[x] == 1) {
            if (a[x] >= g->E) return fail("item out of range");
            if (kind[x] == 0) {
                const FacMajorDesc &d = g->vf_desc[a[x]];
                vf_edge_seq(g->prod0[d.v], g->prod1[d.v], g->nz0[d.v], g->nz1[d.v], g->m_fv0[d.src],
                            g->m_fv1[d.src], o0, o1);
                g->m_vf0[d.ind] = o0;
                g->m_vf1[d.ind] = o1;
            } else {
                fv_edge<true>(g->fv_desc[a[x]], g->m_vf0.data(), g->m_vf1.data(), o0, o1);
                update_message(g, a[x], g->vm_var[a[x]], o0, o1);
            }
            continue;
        }
        if (kind[x] != 2 || a[x] >= g->F) return fail("bad item");
        const uint64_t s0 = g->row_ptr_vf[a[x]], deg = g->row_ptr_vf[a[x] + 1] - s0;
        if (deg > 32) return fail("group item of a factor wider than 32");
        for (uint64_t k = 0; k < deg; k++) {
            if (!((c[x] >> k) & 1u)) continue;
            const FacMajorDesc &d = g->vf_desc[s0 + k];
            vf_edge_seq(g->prod0[d.v], g->prod1[d.v], g->nz0[d.v], g->nz1[d.v], g->m_fv0[d.src], g->m_fv1[d.src],
                        o0, o1);
            g->m_vf0[d.ind] = o0;
            g->m_vf1[d.ind] = o1;
        }
        for (uint64_t k = 0; k < deg; k++) {
            if (!((b[x] >> k) & 1u)) continue;
            const uint64_t e = g->vf_desc[s0 + k].src;
            fv_edge<true>(g->fv_desc[e], g->m_vf0.data(), g->m_vf1.data(), o0, o1);
            update_message(g, e, g->vm_var[e], o0, o1);
        }
    }
    const double d = belief_step(g);
    if (maxdiff_out) *maxdiff_out = d;
    return 0;
}

int orc_products(const orc_graph *g, double *prod0, double *prod1, uint64_t *nz0, uint64_t *nz1) {
    if (prod0) std::copy(g->prod0.begin(), g->prod0.end(), prod0);
    if (prod1) std::copy(g->prod1.begin(), g->prod1.end(), prod1);
    if (nz0) std::copy(g->nz0.begin(), g->nz0.end(), nz0);
    if (nz1) std::copy(g->nz1.begin(), g->nz1.end(), nz1);
    return 0;
}

void orc_fv_message(int code, uint64_t deg, uint64_t src, uint64_t head, double p, double q,
                    double mask0, double mask1, const double *m0, const double *m1,
                    double *out0, double *out1) {
    VarMajorDesc d;
    d.code = code;
    d.src = src;
    d.start = 0;
    d.end = deg;
    d.head = head;
    d.q = q;
    d.p = p;
    d.mask0 = mask0;
    d.mask1 = mask1;
    fv_edge(d, m0, m1, *out0, *out1);
}

void orc_fv_message_seq(int code, uint64_t deg, uint64_t src, uint64_t head, double p, double q,
                        double mask0, double mask1, const double *m0, const double *m1,
                        double *out0, double *out1) {
    VarMajorDesc d;
    d.code = code;
    d.src = src;
    d.start = 0;
    d.end = deg;
    d.head = head;
    d.q = q;
    d.p = p;
    d.mask0 = mask0;
    d.mask1 = mask1;
    fv_edge<true>(d, m0, m1, *out0, *out1);
}

void orc_vf_message(uint64_t deg, uint64_t src, const double *m0, const double *m1,
                    double *out0, double *out1) {
    double p0, p1;
    uint64_t z0, z1;
    prod_var(m0, m1, 0, deg, p0, p1, z0, z1);
    vf_edge(p0, p1, z0, z1, m0[src], m1[src], *out0, *out1);
}

/* CausalFactor::gen_clamped, csrc/ext/causal_factor.cpp:63-100, applied by
 * CausalFactorGraph::clamp (causal_fg.cpp:17-26) to every factor holding the
 * var, one clamp after another in ascending label order (fl_wrapper.cu:26-45:
 * std::map<int,bool> iteration order). */
int orc_clamp(const orc_factor_list *fl, uint64_t n_clamps, const uint64_t *labels,
              const uint8_t *values, orc_flist **out) {
    const uint64_t F = fl->n_factors;
    struct Fac {
        int32_t type;
        uint64_t head;
        std::vector<uint64_t> body;  /* sorted unique */
        double p, q, mask0, mask1;
        uint8_t clamped;
    };
    std::vector<Fac> facs(F);
    for (uint64_t I = 0; I < F; I++) {
        Fac &f = facs[I];
        f.type = fl->type[I];
        f.head = fl->head[I];
        f.body.assign(fl->body + fl->body_off[I], fl->body + fl->body_off[I + 1]);
        std::sort(f.body.begin(), f.body.end());
        f.body.erase(std::unique(f.body.begin(), f.body.end()), f.body.end());
        f.p = fl->p[I]; f.q = fl->q[I];
        f.mask0 = fl->mask0[I]; f.mask1 = fl->mask1[I];
        f.clamped = fl->head_clamped[I];
    }
    std::map<uint64_t, uint8_t> clamps;
    for (uint64_t k = 0; k < n_clamps; k++) clamps[labels[k]] = values[k] ? 1 : 0;
    /* the adjacency used is the original graph's (clamp does not rebuild _G) */
    std::map<uint64_t, std::vector<uint64_t>> holders;
    for (uint64_t I = 0; I < F; I++) {
        std::vector<uint64_t> vs;
        factor_vars(fl, I, vs);
        for (uint64_t lab : vs)
            if (clamps.count(lab)) holders[lab].push_back(I);
    }
    for (auto &c : clamps) {
        const uint64_t lab = c.first;
        const uint64_t x = c.second;
        for (uint64_t I : holders[lab]) {
            Fac &f = facs[I];
            if (f.type == 0) {
                if (lab == f.head) f.p = (double)x;
            } else if (lab == f.head) {
                double n0 = x == 0 ? 1.0 : 0.0, n1 = x == 1 ? 1.0 : 0.0;
                if (f.clamped) {
                    n0 *= f.mask0;
                    n1 *= f.mask1;
                }
                f.clamped = 1;
                f.mask0 = n0;
                f.mask1 = n1;
            } else {
                auto it = std::lower_bound(f.body.begin(), f.body.end(), lab);
                if (it != f.body.end() && *it == lab) {
                    f.body.erase(it);
                    if ((x == 0 && f.type == 1) || (x > 0 && f.type == 2)) {
                        f.p = f.q;
                        f.body.clear();
                    }
                }
            }
        }
    }
    orc_flist *l = new orc_flist();
    l->body_off.push_back(0);
    for (auto &f : facs) {
        l->type.push_back(f.type);
        l->head.push_back(f.head);
        l->body.insert(l->body.end(), f.body.begin(), f.body.end());
        l->body_off.push_back(l->body.size());
        l->p.push_back(f.p); l->q.push_back(f.q);
        l->mask0.push_back(f.mask0); l->mask1.push_back(f.mask1);
        l->head_clamped.push_back(f.clamped);
    }
    *out = l;
    return 0;
}

void orc_flist_view(const orc_flist *l, orc_factor_list *v) {
    v->n_factors = l->type.size();
    v->type = l->type.data();
    v->head = l->head.data();
    v->body_off = l->body_off.data();
    v->body = l->body.data();
    v->p = l->p.data();
    v->q = l->q.data();
    v->head_clamped = l->head_clamped.data();
    v->mask0 = l->mask0.data();
    v->mask1 = l->mask1.data();
}

void orc_flist_free(orc_flist *l) { delete l; }

}  // extern "C"

namespace {

/* p of every AND factor with a rule becomes sigmoid(theta[rule]). */
int tied_list(const orc_factor_list *fl, const int32_t *rule_of_factor, uint64_t n_rules,
              const double *theta, std::vector<double> &p_out) {
    p_out.assign(fl->p, fl->p + fl->n_factors);
    for (uint64_t I = 0; I < fl->n_factors; I++) {
        int32_t r = rule_of_factor[I];
        if (r < 0) continue;
        if ((uint64_t)r >= n_rules) return fail("rule index out of range");
        p_out[I] = 1.0 / (1.0 + std::exp(-theta[r]));
    }
    return 0;
}

int alarm_index(const orc_graph *g, uint64_t n_alarms, const uint64_t *alarm_labels,
                std::vector<uint64_t> &idx) {
    idx.resize(n_alarms);
    for (uint64_t a = 0; a < n_alarms; a++) {
        auto it = std::lower_bound(g->var_label.begin(), g->var_label.end(), alarm_labels[a]);
        if (it == g->var_label.end() || *it != alarm_labels[a]) return fail("alarm label not in graph");
        idx[a] = (uint64_t)(it - g->var_label.begin());
    }
    return 0;
}

}  // namespace

extern "C" {

int orc_loss(const orc_factor_list *fl, const int32_t *rule_of_factor, uint64_t n_rules,
             const double *theta, uint64_t n_alarms, const uint64_t *alarm_labels,
             const double *targets, uint64_t T, double damping, double *loss_out) {
    std::vector<double> p;
    if (tied_list(fl, rule_of_factor, n_rules, theta, p)) return -1;
    orc_factor_list tied = *fl;
    tied.p = p.data();
    orc_graph *g = nullptr;
    if (orc_build(&tied, &g)) return -1;
    std::vector<uint64_t> idx;
    if (alarm_index(g, n_alarms, alarm_labels, idx)) { orc_free(g); return -1; }
    orc_init(g, 0);
    orc_run_fixed(g, T, damping, 0);
    double L = 0.0;
    for (uint64_t a = 0; a < n_alarms; a++) {
        double r = g->old_beliefs[idx[a]] - targets[a];
        L += r * r;
    }
    *loss_out = L;
    orc_free(g);
    return 0;
}

int orc_grad(const orc_factor_list *fl, const int32_t *rule_of_factor, uint64_t n_rules,
             const double *theta, uint64_t n_alarms, const uint64_t *alarm_labels,
             const double *targets, uint64_t T, double damping, double *loss_out,
             double *grad_out) {
    std::vector<double> p;
    if (tied_list(fl, rule_of_factor, n_rules, theta, p)) return -1;
    orc_factor_list tied = *fl;
    tied.p = p.data();
    orc_graph *g = nullptr;
    if (orc_build(&tied, &g)) return -1;
    std::unique_ptr<orc_graph> hold(g);
    std::vector<uint64_t> idx;
    if (alarm_index(g, n_alarms, alarm_labels, idx)) return -1;
    const uint64_t E = g->E, V = g->V, F = g->F;
    orc_init(g, 0);
    /* forward, every state kept: vf[t] = m_vf after t sweeps (factor-major),
     * fv[t] = m_fv after t sweeps (var-major) */
    std::vector<std::vector<double>> vf0(T + 1), vf1(T + 1), fv0(T + 1), fv1(T + 1);
    vf0[0] = g->m_vf0; vf1[0] = g->m_vf1; fv0[0] = g->m_fv0; fv1[0] = g->m_fv1;
    for (uint64_t t = 1; t <= T; t++) {
        (void)sweep(g, damping);
        vf0[t] = g->m_vf0; vf1[t] = g->m_vf1; fv0[t] = g->m_fv0; fv1[t] = g->m_fv1;
    }
    double L = 0.0;
    std::vector<double> bbar(V, 0.0);
    for (uint64_t a = 0; a < n_alarms; a++) {
        double r = g->old_beliefs[idx[a]] - targets[a];
        L += r * r;
        bbar[idx[a]] += 2.0 * r;
    }
    *loss_out = L;
    /* factor of each factor-major slot, and each factor's var-major edges */
    std::vector<uint64_t> fac_of(E);
    for (uint64_t I = 0; I < F; I++)
        for (uint64_t s = g->row_ptr_vf[I]; s < g->row_ptr_vf[I + 1]; s++) fac_of[s] = I;
    std::vector<double> pbar(F, 0.0);
    std::vector<double> avf0(E, 0.0), avf1(E, 0.0);          /* adjoint of m_vf^t */
    std::vector<double> carry0(E, 0.0), carry1(E, 0.0);      /* adjoint of m_fv^t via damping */
    std::vector<double> afv0(E), afv1(E);
    std::vector<double> nvf0(E), nvf1(E);
    for (uint64_t t = T; t >= 1; t--) {
        /* variable pass of sweep t: inputs fv[t] (var-major), outputs vf[t] and beliefs */
        for (uint64_t v = 0; v < V; v++) {
            const uint64_t a = g->row_ptr_fv[v], b = g->row_ptr_fv[v + 1], d = b - a;
            const double *m0 = &fv0[t][a], *m1 = &fv1[t][a];
            double P0 = 1.0, P1 = 1.0;
            for (uint64_t k = 0; k < d; k++) { P0 *= m0[k]; P1 *= m1[k]; }
            double G0 = 0.0, G1 = 0.0;
            if (t == T && bbar[v] != 0.0 && P0 + P1 != 0.0) {
                const double den = (P0 + P1) * (P0 + P1);
                G1 = bbar[v] * P0 / den;
                G0 = -bbar[v] * P1 / den;
            }
            std::vector<double> g0(d, 0.0), g1(d, 0.0);
            for (uint64_t e = 0; e < d; e++) {
                double L0 = 1.0, L1 = 1.0;
                for (uint64_t j = 0; j < d; j++)
                    if (j != e) { L0 *= m0[j]; L1 *= m1[j]; }
                const double S = L0 + L1;
                if (S == 0.0) continue;
                const uint64_t slot = g->fv_desc[a + e].src;
                const double o0 = L0 / S, o1 = L1 / S;
                const double dot = avf0[slot] * o0 + avf1[slot] * o1;
                g0[e] = (avf0[slot] - dot) / S;
                g1[e] = (avf1[slot] - dot) / S;
            }
            for (uint64_t e2 = 0; e2 < d; e2++) {
                double s0 = 0.0, s1 = 0.0;
                for (uint64_t e = 0; e < d; e++) {
                    if (e == e2) continue;
                    double q0 = 1.0, q1 = 1.0;
                    for (uint64_t j = 0; j < d; j++)
                        if (j != e && j != e2) { q0 *= m0[j]; q1 *= m1[j]; }
                    s0 += g0[e] * q0;
                    s1 += g1[e] * q1;
                }
                double l0 = 1.0, l1 = 1.0;
                for (uint64_t j = 0; j < d; j++)
                    if (j != e2) { l0 *= m0[j]; l1 *= m1[j]; }
                afv0[a + e2] = carry0[a + e2] + s0 + G0 * l0;
                afv1[a + e2] = carry1[a + e2] + s1 + G1 * l1;
            }
        }
        /* factor pass of sweep t: m_fv^t = (1-d) u + d m_fv^{t-1}, u = FV(vf[t-1], p) */
        std::fill(nvf0.begin(), nvf0.end(), 0.0);
        std::fill(nvf1.begin(), nvf1.end(), 0.0);
        for (uint64_t e = 0; e < E; e++) {
            carry0[e] = damping * afv0[e];
            carry1[e] = damping * afv1[e];
            const double ub0 = (1 - damping) * afv0[e], ub1 = (1 - damping) * afv1[e];
            const VarMajorDesc &dsc = g->fv_desc[e];
            const uint64_t I = fac_of[dsc.src];
            const double *x0 = vf0[t - 1].data(), *x1 = vf1[t - 1].data();
            const double pp = dsc.p, qq = dsc.q;
            const bool clamped = dsc.code == 3 || dsc.code == 4 || dsc.code == 7 || dsc.code == 8;
            const double mk0 = clamped ? dsc.mask0 : 1.0, mk1 = clamped ? dsc.mask1 : 1.0;
            double M0, M1;
            if (dsc.code == 0) {
                M0 = 1 - pp; M1 = pp;
                const double S = M0 + M1;
                const double u0 = M0 / S, u1 = M1 / S, dot = ub0 * u0 + ub1 * u1;
                pbar[I] += ((ub1 - dot) - (ub0 - dot)) / S;
                continue;
            }
            const bool is_and = dsc.code == 1 || dsc.code == 3 || dsc.code == 5 || dsc.code == 7;
            const bool head_target = dsc.code >= 1 && dsc.code <= 4;
            /* x = the state the CPD singles out (1 for AND, 0 for OR), y = the other */
            auto X = [&](uint64_t j) { return is_and ? x1[j] : x0[j]; };
            auto Y = [&](uint64_t j) { return is_and ? x0[j] : x1[j]; };
            auto addX = [&](uint64_t j, double v) { if (is_and) nvf1[j] += v; else nvf0[j] += v; };
            auto addY = [&](uint64_t j, double v) { if (is_and) nvf0[j] += v; else nvf1[j] += v; };
            if (head_target) {
                double pi = 1.0, Z = 1.0;
                for (uint64_t j = dsc.start; j < dsc.end; j++)
                    if (j != dsc.src) { pi *= X(j); Z *= X(j) + Y(j); }
                const double Mu = (pp - qq) * pi + qq * Z, Mw = (qq - pp) * pi + (1 - qq) * Z;
                M0 = (is_and ? Mw : Mu) * mk0;
                M1 = (is_and ? Mu : Mw) * mk1;
                const double S = M0 + M1;
                if (S == 0.0) continue;
                const double u0 = M0 / S, u1 = M1 / S, dot = ub0 * u0 + ub1 * u1;
                const double Mb0 = (ub0 - dot) / S * mk0, Mb1 = (ub1 - dot) / S * mk1;
                const double Ub = is_and ? Mb1 : Mb0, Wb = is_and ? Mb0 : Mb1;
                const double pib = (pp - qq) * (Ub - Wb), Zb = qq * Ub + (1 - qq) * Wb;
                pbar[I] += (Ub - Wb) * pi;
                for (uint64_t j = dsc.start; j < dsc.end; j++) {
                    if (j == dsc.src) continue;
                    double px = 1.0, pz = 1.0;
                    for (uint64_t k = dsc.start; k < dsc.end; k++)
                        if (k != dsc.src && k != j) { px *= X(k); pz *= X(k) + Y(k); }
                    addX(j, pib * px + Zb * pz);
                    addY(j, Zb * pz);
                }
            } else {
                const uint64_t h = dsc.head;
                const double hx = is_and ? x1[h] * mk1 : x0[h] * mk0;
                const double hy = is_and ? x0[h] * mk0 : x1[h] * mk1;
                const double B = (pp - qq) * (hx - hy), A = qq * hx + (1 - qq) * hy;
                double pi = 1.0, Z = 1.0;
                for (uint64_t j = dsc.start; j < dsc.end; j++)
                    if (j != dsc.src && j != h) { pi *= X(j); Z *= X(j) + Y(j); }
                const double Ma = A * Z, Mb = A * Z + B * pi;
                M0 = is_and ? Ma : Mb;
                M1 = is_and ? Mb : Ma;
                const double S = M0 + M1;
                if (S == 0.0) continue;
                const double u0 = M0 / S, u1 = M1 / S, dot = ub0 * u0 + ub1 * u1;
                const double Mb0 = (ub0 - dot) / S, Mb1 = (ub1 - dot) / S;
                const double Mab = is_and ? Mb0 : Mb1, Mbb = is_and ? Mb1 : Mb0;
                const double Zb = (Mab + Mbb) * A, Ab = (Mab + Mbb) * Z, Bb = Mbb * pi, pib = Mbb * B;
                pbar[I] += Bb * (hx - hy);
                const double hxb = Bb * (pp - qq) + Ab * qq, hyb = -Bb * (pp - qq) + Ab * (1 - qq);
                if (is_and) { nvf1[h] += hxb * mk1; nvf0[h] += hyb * mk0; }
                else { nvf0[h] += hxb * mk0; nvf1[h] += hyb * mk1; }
                for (uint64_t j = dsc.start; j < dsc.end; j++) {
                    if (j == dsc.src || j == h) continue;
                    double px = 1.0, pz = 1.0;
                    for (uint64_t k = dsc.start; k < dsc.end; k++)
                        if (k != dsc.src && k != h && k != j) { px *= X(k); pz *= X(k) + Y(k); }
                    addX(j, pib * px + Zb * pz);
                    addY(j, Zb * pz);
                }
            }
        }
        avf0.swap(nvf0);
        avf1.swap(nvf1);
    }
    for (uint64_t r = 0; r < n_rules; r++) grad_out[r] = 0.0;
    for (uint64_t I = 0; I < F; I++) {
        const int32_t r = rule_of_factor[I];
        if (r >= 0) grad_out[r] += pbar[I] * p[I] * (1.0 - p[I]);
    }
    return 0;
}

}  // extern "C"
