/*
 * oracle.h -- CPU FP64 restatement of the reference's CausalBP path (PARALL,
 * SEQFIX and TOPO updates).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * engine: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (fastlbp_amd) never links or calls it.
 *
 * Every function restates one piece of the reference (cited per function in
 * oracle.cpp, paths relative to the reference repository root) with the same
 * floating-point operation order, compiled with -ffp-contract=off so that its
 * results are the reference's own FP64 results on x86-64.
 */
#ifndef FASTLBP_ORACLE_H
#define FASTLBP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A list of causal factors in list order (the reference's
 * std::vector<CausalFactor>, include/ext/causal_factor.h:14-76). */
typedef struct orc_factor_list {
    uint64_t n_factors;
    const int32_t *type;          /* [F] 0 Singleton, 1 DefiniteAnd, 2 DefiniteOr */
    const uint64_t *head;         /* [F] head label */
    const uint64_t *body_off;     /* [F+1] offsets into body */
    const uint64_t *body;         /* [body_off[F]] body labels */
    const double *p;              /* [F] prob() */
    const double *q;              /* [F] prob_default() */
    const uint8_t *head_clamped;  /* [F] */
    const double *mask0;          /* [F] head_mask[0] */
    const double *mask1;          /* [F] head_mask[1] */
} orc_factor_list;

typedef struct orc_graph orc_graph;
typedef struct orc_flist orc_flist;  /* an oracle-owned factor list */

/* Graph construction: CausalFactorGraph(vector) + CausalBP::construct +
 * putParallUpdateSeqToKernelFused.  Returns 0 on success. */
int orc_build(const orc_factor_list *fl, orc_graph **out);
void orc_free(orc_graph *g);
uint64_t orc_nvars(const orc_graph *g);
uint64_t orc_nfactors(const orc_graph *g);
uint64_t orc_nedges(const orc_graph *g);

/* Structure dump.  Any pointer may be NULL.
 *  var_label[V]; fac_off[F+1] (h_row_ptr_vf); edge_var[E] (var index of each
 *  factor-major slot); var_off[V+1] (h_row_ptr_fv); var_src[E] (factor-major
 *  slot of each var-major edge, Index6Real4.src_ind); head_slot[F] (h_head);
 *  code[E] (type code 0..8 of each var-major edge). */
int orc_dump(const orc_graph *g, uint64_t *var_label, uint64_t *fac_off,
             uint64_t *edge_var, uint64_t *var_off, uint64_t *var_src,
             uint64_t *head_slot, int32_t *code);

/* CausalBP::init: messages and products <- c (1.0, or 0.0 if logdomain). */
int orc_init(orc_graph *g, int logdomain);

/* CausalBP::run() with maxtime = infinity, plus build-defined damping.
 * Iteration counter persists across calls like the reference's _iters. */
int orc_run(orc_graph *g, double tol, uint64_t maxiter, double damping,
            int threads, uint64_t *iters_out, double *maxdiff_out);

/* CausalBP::run(tol, maxIters, histLength): exactly n_iters sweeps, beliefs
 * once at the end. */
int orc_run_fixed(orc_graph *g, uint64_t n_iters, double damping, int threads);

/* CausalBP::run(tol, minIters, maxIters, histLength) (maxtime = infinity)
 * with HostBeliefHistory: returns the yet-to-converge fraction in *yet,
 * the low-pass beliefs [V] (newBelief), the exit reason (0 all converged,
 * 1 big fraction converged, 2 diverged) and the last max diff. */
int orc_run_hist(orc_graph *g, double tol, uint64_t min_iters, uint64_t max_iters, uint64_t hist_len,
                 double damping, double *yet, double *lowpass, int *reason, double *maxdiff);

/* One sweep + beliefs without touching the iteration counter; writes the
 * max belief change.  Used to time a bounded sample. */
int orc_sweep(orc_graph *g, double damping, int threads, double *maxdiff_out);

int orc_beliefs(const orc_graph *g, double *out);   /* [V] P(x=1) */
uint64_t orc_iters(const orc_graph *g);
/* messages in the reference's storage orders: fv var-major, vf factor-major */
int orc_messages(const orc_graph *g, double *fv0, double *fv1, double *vf0,
                 double *vf1);
int orc_set_messages(orc_graph *g, const double *fv0, const double *fv1,
                     const double *vf0, const double *vf1);

/* A single f->v message (calcMessageFVFusedAllKernel body) for one edge of a
 * factor whose incoming v->f messages are m0[0..deg), m1[0..deg).
 * src = slot of the target var, head = slot of the head var. */
void orc_fv_message(int code, uint64_t deg, uint64_t src, uint64_t head,
                    double p, double q, double mask0, double mask1,
                    const double *m0, const double *m1, double *out0,
                    double *out1);

/* Product + leave-one-out v->f message (calcProdAllKernel +
 * calcMessageVFAllKernel) for target slot src of a variable with incoming
 * f->v messages m0[0..deg), m1[0..deg). */
void orc_vf_message(uint64_t deg, uint64_t src, const double *m0,
                    const double *m1, double *out0, double *out1);

/* Clamp + rebuild as examples/fl_wrapper.cu:26-45 does it: clamps applied in
 * ascending label order with CausalFactorGraph::clamp semantics. */
int orc_clamp(const orc_factor_list *fl, uint64_t n_clamps,
              const uint64_t *labels, const uint8_t *values, orc_flist **out);
void orc_flist_view(const orc_flist *l, orc_factor_list *view);
void orc_flist_free(orc_flist *l);

/* Adjoint reference: loss L = sum_a (b_a - y_a)^2 over alarm variables after
 * exactly T damped sweeps from init; p of every AND factor with rule r >= 0 is
 * sigmoid(theta[r]).  Writes L and dL/dtheta[n_rules] by an unrolled
 * reverse-mode sweep in FP64 that stores every iteration's state. */
int orc_grad(const orc_factor_list *fl, const int32_t *rule_of_factor,
             uint64_t n_rules, const double *theta, uint64_t n_alarms,
             const uint64_t *alarm_labels, const double *targets, uint64_t T,
             double damping, double *loss_out, double *grad_out);

/* Forward only: same setting, returns the loss (for finite differences). */
int orc_loss(const orc_factor_list *fl, const int32_t *rule_of_factor,
             uint64_t n_rules, const double *theta, uint64_t n_alarms,
             const uint64_t *alarm_labels, const double *targets, uint64_t T,
             double damping, double *loss_out);

/* ---- sequential schedules (updates = SEQFIX / TOPO) ----
 * Orders are lists of var-major edge positions (Edge(i, dual) ->
 * h_row_ptr_fv[i] + dual).  With a schedule set, orc_run / orc_run_fixed /
 * orc_sweep run its waves instead of the PARALL sweep; damping must be 0. */
typedef struct orc_rng orc_rng;          /* a std::mt19937 */
orc_rng *orc_rng_new(uint32_t seed);
void orc_rng_free(orc_rng *r);
uint32_t orc_rng_next(orc_rng *r);
/* libstdc++'s std::shuffle of 0..n-1, restated; and std::shuffle itself */
int orc_shuffle(orc_rng *r, uint64_t n, uint64_t *out);
int orc_std_shuffle(uint32_t seed, uint64_t n, uint64_t *out);
/* buildFixedUpdateSeq (advances r) / buildTopoUpdateSeq: E positions */
int orc_order_seqfix(const orc_graph *g, orc_rng *r, uint64_t *out);
int orc_order_topo(const orc_graph *g, uint64_t *out);
/* putUpdateSeqToKernelFused: split an order into waves */
int orc_seq_set(orc_graph *g, const uint64_t *order, uint64_t n, int respect_preorder);
int orc_seq_clear(orc_graph *g);
int orc_seq_info(const orc_graph *g, uint64_t *n_waves, uint64_t *n_fv, uint64_t *n_vf);
/* fv_off/vf_off [n_waves+1]; fv_edge [n_fv] var-major positions updated;
 * vf_slot [n_vf] factor-major slots recomputed */
int orc_seq_waves(const orc_graph *g, uint64_t *fv_off, uint64_t *fv_edge, uint64_t *vf_off,
                  uint64_t *vf_slot);
/* Checker of a dependency-level split (not a restatement): one sweep of the
 * listed items in the listed order (kind 0: v->f of factor-major slot
 * item[x]; kind 1: f->v update of var-major edge item[x]), then beliefs. */
int orc_seq_sweep_items(orc_graph *g, uint64_t n, const uint8_t *kind, const uint64_t *item,
                        double *maxdiff_out);
/* Checker of the engine's execution items: kind 0 v->f of slot a, kind 1
 * f->v of var-major edge a, kind 2 factor a: v->f of the slots in mask c
 * (from current states), then f->v of the slots in mask b. */
int orc_seq_sweep_exec(orc_graph *g, uint64_t n, const uint8_t *kind, const uint32_t *a, const uint32_t *b,
                       const uint32_t *c, double *maxdiff_out);
/* per-variable products and zero counts (maintained in place by SEQ sweeps) */
int orc_products(const orc_graph *g, double *prod0, double *prod1, uint64_t *nz0, uint64_t *nz1);
/* one f->v message with the sequential kernels' scale() */
void orc_fv_message_seq(int code, uint64_t deg, uint64_t src, uint64_t head,
                        double p, double q, double mask0, double mask1,
                        const double *m0, const double *m1, double *out0,
                        double *out1);

const char *orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
