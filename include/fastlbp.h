/*
 * fastlbp.h -- C ABI of the MI355X loopy-BP engine (libfastlbp.so).
 *
 * Plain pointers and sizes only.  Every entry point returns an int status:
 * 0 on success, a negative FLBP_E* code on failure, with a message available
 * from flbp_last_error() (per calling thread).  Nothing in the library calls
 * exit().
 *
 * The engine replaces the reference's device-side PARALL sweep, which is
 * reached through the C++ free functions
 *     lbp::kernel::calcMessageFVFusedAll  include/kernel/causal_bp_parall.h:37-44
 *     lbp::kernel::calcProdAll            include/kernel/causal_bp_parall.h:13-22
 *     lbp::kernel::calcMessageVFAll       include/kernel/causal_bp_parall.h:24-35
 *     lbp::kernel::getBeliefsV            include/kernel/causal_bp_seq.h:22-27
 * and the device-state management of lbp::CausalBP (csrc/ext/causal_bp.cu)
 * that drives them.  Paths are relative to the reference repository root.
 * The C++ drop-in classes (include/ext/causal_bp.h and friends) sit on top
 * of this ABI.
 */
#ifndef FASTLBP_H
#define FASTLBP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLBP_OK 0
#define FLBP_EINVAL (-1)     /* bad argument / malformed graph */
#define FLBP_ENOMEM (-2)     /* host or device allocation failed */
#define FLBP_EHIP (-3)       /* a HIP runtime call failed */
#define FLBP_ENODEV (-4)     /* no usable GPU */
#define FLBP_ERANGE (-5)     /* size beyond what the engine supports */
#define FLBP_EIO (-6)        /* file could not be read / parsed */

/* ------------------------------------------------------------------ */
/* Factor lists and graphs (host side, no GPU needed)                  */
/* ------------------------------------------------------------------ */

/* A list of causal factors in list order: the flat form of the reference's
 * std::vector<CausalFactor> (include/ext/causal_factor.h:14-76).
 * type: 0 Singleton, 1 DefiniteAnd, 2 DefiniteOr (CausalFactor::type()).
 * p = prob(), q = prob_default(); head_clamped/mask0/mask1 = the public
 * head_clamped/head_mask members. */
typedef struct flbp_factor_list {
    uint64_t n_factors;
    const int32_t *type;
    const uint64_t *head;
    const uint64_t *body_off;     /* [F+1] */
    const uint64_t *body;         /* [body_off[F]] labels */
    const double *p;
    const double *q;
    const uint8_t *head_clamped;
    const double *mask0;
    const double *mask1;
} flbp_factor_list;

/* The CSR form of a factor graph in the reference's orders, i.e. what
 * CausalBP::construct (csrc/ext/causal_bp.cu:310-407) and
 * putParallUpdateSeqToKernelFused (724-817) derive on the host:
 *   var i        = rank of its label (causal_fg.cpp:50-68)
 *   fac_off      = h_row_ptr_vf; slots of factor I in nbF(I) order, i.e.
 *                  ascending label
 *   edge_var[s]  = var index of factor-major slot s
 *   var_off      = h_row_ptr_fv; edges of var i in nbV(i) order, i.e.
 *                  ascending factor index
 *   var_edge[k]  = factor-major slot of var-major edge k (Index6Real4.src_ind)
 *   head_slot[I] = factor-major slot of I's head (h_head) */
typedef struct flbp_csr {
    uint64_t n_vars, n_factors, n_edges;
    const uint64_t *var_label;    /* [V] */
    const uint64_t *fac_off;      /* [F+1] */
    const uint64_t *edge_var;     /* [E] */
    const uint64_t *var_off;      /* [V+1] */
    const uint64_t *var_edge;     /* [E] */
    const int32_t *fac_type;      /* [F] */
    const uint64_t *head_slot;    /* [F] */
    const uint8_t *head_clamped;  /* [F] */
    const double *p, *q, *mask0, *mask1;  /* [F] */
} flbp_csr;

typedef struct flbp_graph flbp_graph;   /* owns a flbp_csr */
typedef struct flbp_flist flbp_flist;   /* owns a factor list */

/* CausalFactorGraph(const std::vector<CausalFactor>&) (csrc/ext/causal_fg.cpp:50-68)
 * followed by the CSR derivation of CausalBP::construct. */
int flbp_graph_build(const flbp_factor_list *fl, flbp_graph **out);
int flbp_graph_csr(const flbp_graph *g, flbp_csr *view);
void flbp_graph_free(flbp_graph *g);

/* The .causal_fg reader (csrc/ext/causal_fg.cpp:118-279), implicit
 * Singleton(0.5) factors included. */
int flbp_read_causal_fg(const char *path, flbp_flist **out);
/* Writes a factor list in the .causal_fg text format the reader accepts. */
int flbp_write_causal_fg(const flbp_factor_list *fl, const char *path);

/* Binary graph files (not in the reference): the CSR above with 32-bit
 * index arrays and a trailing 64-bit integrity check, little-endian, so a
 * graph of tens of millions of variables loads without parsing or
 * rebuilding.  Layout: "FLBPCSR1", u64 n_vars, n_factors, n_edges,
 * u64 var_label[V], u32 fac_off[F+1], u32 edge_var[E], u32 var_off[V+1],
 * u32 var_edge[E], u32 head_slot[F], u8 fac_type[F], u8 head_clamped[F],
 * f64 p[F], q[F], mask0[F], mask1[F], u64 check.  Reading validates the
 * check and the CSR's structure (FLBP_EIO / FLBP_EINVAL). */
int flbp_write_graph(const flbp_csr *csr, const char *path);
int flbp_read_graph(const char *path, flbp_graph **out);

/* CausalFactorGraph::clamp (causal_fg.cpp:17-26) for each (label, value),
 * applied in ascending label order as examples/fl_wrapper.cu:26-45 does,
 * returning the factor list to rebuild from (fl_wrapper.cu:45). */
int flbp_clamp(const flbp_factor_list *fl, uint64_t n_clamps, const uint64_t *labels,
               const uint8_t *values, flbp_flist **out);
int flbp_flist_view(const flbp_flist *l, flbp_factor_list *view);
void flbp_flist_free(flbp_flist *l);

/* Seeded synthetic causal graphs (not in the reference): a layered
 * derivation graph.  Layer 0 holds leaves with Singleton priors; every
 * head of layers 1..n_layers-1 is a deterministic OR over k clause
 * variables, each clause an AND factor over b body variables from earlier
 * layers, plus with probability back_frac one body variable from a later
 * layer (a loop).  On top of that one defining factor per variable,
 * extra_frac * n_vars further AND factors (further rule instances) each
 * take a uniformly chosen head of layers 1.. as their head, with a body
 * drawn as a clause's and q = extra_q: when its body holds the factor
 * favours its head by p : 1-p, otherwise (extra_q = 0.5) it is flat.  With
 * the defaults this gives F = 1.5 V and E = 4.5 V.  alarm_free = 1 keeps the
 * last layer out of every rule body (loops close through layers
 * 1..n_layers-2 only) and draws the alarms from the last layer's heads, so
 * every alarm clamp is a head clamp that keeps the graph's structure. */
typedef struct flbp_gen_params {
    uint64_t n_vars;
    uint64_t seed;
    uint32_t n_layers;           /* >= 2; layer 0 = leaves */
    uint32_t k_min, k_max;       /* clauses per head */
    uint32_t b_min, b_max;       /* body variables per clause */
    uint32_t n_rules;            /* 0: no tying; else rule ids 0..n_rules-1 */
    double leaf_frac;
    double p_and_lo, p_and_hi;
    double p_leaf_lo, p_leaf_hi;
    double back_frac;
    uint64_t n_alarms;
    double extra_frac;           /* further AND factors per variable (default 0.5) */
    double extra_q;              /* their q (default 0.5) */
    uint32_t alarm_free;         /* 1: alarms only among heads without body uses (default 1) */
} flbp_gen_params;

typedef struct flbp_generated flbp_generated;

void flbp_gen_defaults(flbp_gen_params *p);
int flbp_generate(const flbp_gen_params *params, flbp_generated **out);
int flbp_generated_view(const flbp_generated *g, flbp_factor_list *view);
/* alarm labels [n_alarms] and rule id per factor [F] (-1: none) */
int flbp_generated_extra(const flbp_generated *g, uint64_t *n_alarms,
                         const uint64_t **alarms, const int32_t **rule_of_factor);
void flbp_generated_free(flbp_generated *g);

/* Alarm ranking (not in the reference): indices of alarms sorted by
 * marginal descending, ties (bitwise-equal marginals) by label ascending. */
int flbp_rank_alarms(uint64_t n_alarms, const uint64_t *labels, const double *marginals,
                     uint64_t *order_out);

/* ------------------------------------------------------------------ */
/* Plain solve on the GPU                                              */
/* ------------------------------------------------------------------ */

typedef struct flbp_plan flbp_plan;

/* CausalBP::construct's device half (csrc/ext/causal_bp.cu:310-407, the H->D
 * copies at 368-369 and 815-816): copies the CSR into device memory owned by
 * the plan, in the engine's layout.  device = HIP device ordinal. */
int flbp_plan_create(const flbp_csr *csr, int device, flbp_plan **out);
void flbp_plan_destroy(flbp_plan *plan);

/* CausalBP::init (causal_bp.cu:824-856): every message (1,1) -- (0,0) when
 * logdomain is set -- and zero counts 0.  The belief array and the
 * iteration counter are left alone, as in the reference. */
int flbp_plan_init(flbp_plan *plan, int logdomain);

typedef struct flbp_run_opts {
    double tol;              /* stop when max |b_new - b_old| <= tol */
    uint64_t maxiter;        /* iteration cap (counted like CausalBP::_iters) */
    double maxtime;          /* seconds; checked between chunks of sweeps */
    double damping;          /* 0 = the reference */
    uint32_t chunk;          /* sweeps enqueued per host check; 0 = default */
} flbp_run_opts;

typedef struct flbp_run_result {
    uint64_t iters;          /* the plan's iteration counter after the call */
    uint64_t sweeps;         /* sweeps done by this call */
    double maxdiff;          /* last max belief change (INFINITY if none) */
    double seconds;          /* wall time of the call */
} flbp_run_result;

/* CausalBP::run() (causal_bp.cu:858-910): the three passes of
 * calcNewMessageFusedAll (409-444) + calcBeliefsV (483-515) + the max-diff
 * reduce (879-886) per iteration, fused into two kernels per sweep with a
 * device-side stop test. */
int flbp_plan_run(flbp_plan *plan, const flbp_run_opts *opts, flbp_run_result *res);

/* CausalBP::run(tol, maxIters, histLength) for PARALL (causal_bp.cu:912-1004):
 * exactly n_sweeps sweeps, no convergence test. */
int flbp_plan_run_fixed(flbp_plan *plan, uint64_t n_sweeps, double damping,
                        flbp_run_result *res);

/* CausalBP::run(tol, minIters, maxIters, histLength), the 4-argument run
 * (causal_bp.cu:1006-1143): one sweep at a time until max-diff <= tol
 * (reason 0), or, past min_iters, the fraction of variables whose belief
 * moved by more than tol falls below (iters - min_iters) / (max_iters -
 * min_iters) (reason 1), or iters > max_iters or maxtime seconds pass
 * (reason 2).  Each sweep's finite beliefs enter a per-variable ring of
 * hist_len values, kept on the device (the reference's HostBeliefHistory
 * keeps it on the host); low_pass[V] (NULL skips) receives the beliefs for
 * reason 0, else each variable's mean over its ring. */
typedef struct flbp_hist_result {
    double yet_to_converge;  /* the moved fraction of the last sweep (1.0 if none ran) */
    double maxdiff;          /* max belief change of the last sweep (INFINITY if none ran) */
    uint64_t iters;          /* sweeps run */
    int reason;              /* 0 converged, 1 the fraction test, 2 max_iters or maxtime */
} flbp_hist_result;
int flbp_plan_run_hist(flbp_plan *plan, double tol, uint64_t min_iters, uint64_t max_iters, uint64_t hist_len,
                       double maxtime, double damping, flbp_hist_result *res, double *low_pass);
/* flbp_plan_run_hist with no device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory during the
 * call (serving loops with a fixed memory budget, harnesses that forbid
 * hipMalloc).  Same arguments, results and low-pass bits as
 * flbp_plan_run_hist.  The belief history lives in a fixed 64 MiB array of
 * device memory per GPU, held from the library's first use of that GPU
 * until the process ends: ring hist_len x V doubles plus 24 bytes per
 * variable, each range rounded up to 256 bytes; a history beyond 64 MiB
 * (hist_len <= 5 at V = 1M, 416 at V = 20k) returns FLBP_ERANGE.
 * Every call on one GPU shares that array and clears what it needs at its
 * start: calls that overlap on one GPU (from several threads or plans of
 * the same device) race and must be serialized by the caller.  The sweeps
 * run as direct launches even where flbp_plan_use_graphs enabled graphs. */
int flbp_plan_run_hist_noalloc(flbp_plan *plan, double tol, uint64_t min_iters, uint64_t max_iters,
                               uint64_t hist_len, double maxtime, double damping, flbp_hist_result *res,
                               double *low_pass);

/* Back to the state of a freshly created plan: iteration counter 0 and old
 * beliefs 0.0 (the reference gets this only by constructing a new
 * CausalBP).  Messages are left as they are; call flbp_plan_init next. */
int flbp_plan_reset(flbp_plan *plan);

/* transferBeliefToHost (causal_bp.cu:552-554): P(x_i = 1) for i < n_vars. */
int flbp_plan_beliefs(flbp_plan *plan, double *out, uint64_t n);
/* out[i] = belief of variable vars[i] (variable order): n values gathered on
 * the device, so a caller that reads a few variables of a large graph does
 * not copy all V beliefs.  Not in the reference. */
int flbp_plan_beliefs_at(flbp_plan *plan, uint64_t n, const uint64_t *vars, double *out);

/* transferMessagesToHost (causal_bp.cu:1145-1161) and its inverse: f->v
 * messages in var-major (nbV) order, v->f messages in factor-major order,
 * each as interleaved (state0, state1) pairs [2*E].  NULL skips one. */
int flbp_plan_get_messages(flbp_plan *plan, double *fv_varmajor, double *vf_facmajor);
/* (Not in the reference.)  Under a sequential schedule only the v->f
 * messages may be set: the schedule keeps each variable's product of its
 * f->v messages up to date incrementally, so setting f->v messages fails. */
int flbp_plan_set_messages(flbp_plan *plan, const double *fv_varmajor,
                           const double *vf_facmajor);

/* Replaces p of factor I by p[I] for every I (used by the adjoint's
 * training loop; the reference rebuilds the whole CausalBP instead). */
int flbp_plan_set_p(flbp_plan *plan, const double *p, uint64_t n);

/* Head clamps applied in place: for each listed AND/OR factor, the head
 * rule of CausalFactor::gen_clamped (csrc/ext/causal_factor.cpp:75-83) --
 * head_clamped set, mask (1, 0) for value 0 or (0, 1) for value 1,
 * multiplied into an existing mask.  The graph structure is unchanged.  A
 * factor whose head appears in some factor body (its clamp would also
 * erase or collapse those body uses, causal_factor.cpp:84-95), and a prior
 * (Singleton) factor, need the rebuild (flbp_clamp) and are rejected here
 * (FLBP_EINVAL).  CausalFactorGraph::clamp (causal_fg.cpp:17-26) clamps
 * every factor of the variable; flbp_plan_clamp_vars below does that.  Not
 * in the reference: it lets a training loop switch between clamped
 * variants without rebuilding the plan. */
int flbp_plan_set_head_clamps(flbp_plan *plan, uint64_t n, const uint64_t *factors, const uint8_t *values);
/* CausalFactorGraph::clamp(i, x) (causal_fg.cpp:17-26) in place, for
 * variables (indices) that appear in no factor body: every AND/OR factor
 * headed by vars[k] gets the head clamp of values[k], as
 * flbp_plan_set_head_clamps does.  This equals rebuilding the clamped graph
 * (fl_wrapper.cu:26-45).  A variable with body uses, one that heads no
 * factor or only a prior, is rejected (FLBP_EINVAL) and nothing changes. */
int flbp_plan_clamp_vars(flbp_plan *plan, uint64_t n, const uint64_t *vars, const uint8_t *values);
/* every head clamp and mask back to the plan's creation values */
int flbp_plan_reset_head_clamps(flbp_plan *plan);

/* Kernel timing of the last run call: sum of device time of the factor and
 * variable passes in milliseconds, and launches of each (0 if not
 * recorded).  Enabled with flbp_plan_profile(plan, 1). */
int flbp_plan_profile(flbp_plan *plan, int enable);
int flbp_plan_last_timing(flbp_plan *plan, double *fac_ms, double *var_ms,
                          uint64_t *fac_launches, uint64_t *var_launches);

/* Kernel shape and message layout of the sweep.  4 (the default above a
 * 64 MB device footprint, 3 below it): the state layout -- f->v messages factor-major, rewritten in place by the
 * factor pass, and per variable its product pair and zero-count code,
 * from which the factor pass derives each v->f input (the reference's
 * leave-one-out division); the variable pass gathers f->v messages and
 * writes no per-edge data.  3: LDS-staged blocks over balanced edge
 * ranges, v->f messages factor-major and f->v var-major, so both passes
 * read contiguously and scatter their outputs (plus a factor-major f->v
 * copy written by damped sweeps); 0: LDS-staged, v->f var-major and f->v
 * factor-major (both passes gather, all stores contiguous); 1: one lane per
 * factor / variable in layout 0; 2: LDS-staged, both arrays factor-major.
 * Results are identical; switching between layouts moves the stored
 * messages, so a solve may continue under another variant.
 * flbp_adjoint_create switches its plan to 0. */
int flbp_plan_set_variant(flbp_plan *plan, int variant);
/* Replay each chunk of sweeps from a hipGraph captured on first use (one
 * per chunk length, damping, tol and variant), or launch its kernels one by
 * one (0, the default: on MI355X the replay measured within 1% of direct
 * launches at 100k and 1M variables, the per-sweep cost being on the GPU
 * side).  Profiling (flbp_plan_profile) always launches one by one, since
 * its events sit between the kernels.  Results are the same either way. */
int flbp_plan_use_graphs(flbp_plan *plan, int enable);
int flbp_plan_get_variant(const flbp_plan *plan);

uint64_t flbp_plan_device_bytes(const flbp_plan *plan);

/* ------------------------------------------------------------------ */
/* Sequential update schedules (updates = SEQFIX / TOPO)               */
/* ------------------------------------------------------------------ */

/* CausalBP::Properties::UpdateType (include/ext/causal_bp.h:29-33). */
#define FLBP_UPDATES_SEQFIX 0
#define FLBP_UPDATES_PARALL 1
#define FLBP_UPDATES_TOPO 2

/* An update order split into the reference's waves
 * (putUpdateSeqToKernelFused, csrc/ext/causal_bp.cu:583-722), built on the
 * host from a graph's CSR; no GPU is needed.  Orders are var-major edge
 * positions: the reference's Edge(i, dual) is var_off[i] + dual. */
typedef struct flbp_schedule flbp_schedule;

typedef struct flbp_schedule_info {
    uint64_t n_order;        /* length of the update order (E for SEQFIX / TOPO) */
    uint64_t n_waves;        /* the reference's _updateSeqVF.size() */
    uint64_t n_fv_items;     /* f->v updates over all waves */
    uint64_t n_vf_items;     /* v->f recomputations over all waves */
    uint64_t n_levels;       /* launches per sweep on the GPU */
    int32_t updates;         /* FLBP_UPDATES_SEQFIX / _TOPO, or -1 for an explicit order */
    int32_t respect_preorder;
    uint64_t n_exec_items;   /* what runs per sweep (flbp_schedule_exec) */
} flbp_schedule_info;

/* The reference's schedule for `updates`: SEQFIX shuffles the factor
 * indices with std::shuffle by the calling thread's std::mt19937
 * (buildFixedUpdateSeq, causal_bp.cu:186-201; the generator starts at seed
 * 42 in each thread and every SEQFIX schedule advances it, as every
 * reference CausalBP construction does); TOPO takes the breadth-first
 * preorder (buildTopoUpdateSeq, 203-304) with respect_preorder set.
 * FLBP_UPDATES_PARALL is rejected (a plan without schedule is PARALL). */
int flbp_schedule_create(const flbp_csr *c, int updates, flbp_schedule **out);
/* An explicit order of n var-major edge positions (repeats allowed). */
int flbp_schedule_from_order(const flbp_csr *c, const uint64_t *order, uint64_t n, int respect_preorder,
                             flbp_schedule **out);
void flbp_schedule_destroy(flbp_schedule *s);
int flbp_schedule_get_info(const flbp_schedule *s, flbp_schedule_info *info);
/* the order [n_order] */
int flbp_schedule_order(const flbp_schedule *s, uint64_t *order);
/* wave w updates fv_edge[fv_off[w] .. fv_off[w+1]) (var-major positions)
 * after recomputing vf_slot[vf_off[w] .. vf_off[w+1]) (factor-major slots);
 * fv_off / vf_off have n_waves + 1 entries.  NULL skips an array. */
int flbp_schedule_waves(const flbp_schedule *s, uint64_t *fv_off, uint64_t *fv_edge, uint64_t *vf_off,
                        uint64_t *vf_slot);
/* the dependency level (1 .. n_levels) of every f->v item [n_fv_items] and
 * v->f item [n_vf_items], in the order of flbp_schedule_waves.  NULL skips. */
int flbp_schedule_levels(const flbp_schedule *s, uint32_t *fv_level, uint32_t *vf_level);
/* What runs on the device per sweep, n_exec_items entries, each with its
 * level: kind 0 = the v->f message of factor-major slot a; kind 1 = the
 * f->v update of var-major edge a (factors wider than 6 variables); kind 2
 * = factor a's updates of one wave as one item: the v->f messages of its
 * slots in mask c, recomputed from the variables' states before the item's
 * own updates, then the f->v updates of its slots in mask b (bit k = the
 * factor's k-th slot).  NULL skips an array. */
int flbp_schedule_exec(const flbp_schedule *s, uint8_t *kind, uint32_t *a, uint32_t *b, uint32_t *c,
                       uint32_t *level);
/* rnd_gen.seed(seed) on the calling thread, as examples/bingo_wrapper.cu:115
 * does before constructing its CausalBP. */
void flbp_seed(uint32_t seed);

/* Runs the plan's sweeps by a schedule (NULL: back to PARALL).  Its waves
 * run as dependency levels: every f->v and v->f update of the waves gets a
 * level one above the updates before it whose results it reads or whose
 * inputs it overwrites, and one launch runs one level, so the results are
 * the reference's per-wave results bit for bit with far fewer launches
 * (a SEQFIX sweep of a 1M-variable graph: thousands of waves, under a
 * hundred levels).  The plan switches to variant 3 (the reference's message
 * orders) and keeps per-variable products and zero counts, set to the
 * values flbp_plan_init leaves (c, no zeros) here and by every init.  The
 * schedule is copied; it may be destroyed afterwards.  Damping is defined
 * for PARALL only: run calls with damping != 0 fail under a schedule, and
 * so do flbp_plan_set_variant to another variant and flbp_adjoint_create. */
int flbp_plan_set_schedule(flbp_plan *plan, const flbp_schedule *s);
/* FLBP_UPDATES_PARALL without a schedule, else the schedule's updates
 * (-1 for an explicit order) */
int flbp_plan_get_updates(const flbp_plan *plan);
/* Execution only, results unchanged: the longest run of trailing levels
 * with at most max_items items each (0 .. 1024, default 1024) runs in one
 * single-workgroup launch with barriers between levels; 0 launches every
 * level.  No counterpart in the reference (its waves are launches); kept
 * for tests that drive both paths on small graphs.  Applies to the
 * current schedule and to later ones. */
int flbp_plan_set_seq_tail(flbp_plan *plan, uint32_t max_items);

/* ------------------------------------------------------------------ */
/* Batched solve (not in the reference)                                */
/* ------------------------------------------------------------------ */

/* Many graphs -- independent graphs and clamped variants of them -- solved
 * together in one device state.  The batch holds groups: a group is one
 * graph structure solved for K = 1, 2, 4, ..., 64 variants, whose messages,
 * products and beliefs are stored K-interleaved (an index load and a
 * gathered cache line serve all K variants).  A variant of a group differs
 * from the group's graph only by head clamps of variables that appear in no
 * factor body -- CausalFactorGraph::clamp (csrc/ext/causal_fg.cpp:17-26) of
 * such a variable keeps the structure -- so it is bit for bit the solve of
 * the graph rebuilt with those clamps (examples/fl_wrapper.cu:26-45).
 * Variants are numbered across groups in order (group 0's variants first);
 * an independent graph is a group of one variant.  Sweeps use the state
 * layout of the plain solve (variant 4); sweeps are captured into a
 * hipGraph and replayed, the per-variant stop test and the list of running
 * groups kept on the device (opts.chunk).  Each variant stops after
 * exactly the sweeps its plain solve runs, with bit-identical beliefs. */
typedef struct flbp_batch flbp_batch;

typedef struct flbp_batch_opts {
    uint32_t tile_factors;   /* edge slots per factor tile of a 1-variant group, 64..512 (K variants: 1/K
                                of it, 2/K for K >= 4) */
    uint32_t tile_vars;      /* edge slots per variable tile, 64..512 (same scaling) */
    uint32_t grid_blocks;    /* reserved: the sweep kernels launch one block per tile */
    uint32_t chunk;          /* sweeps per hipGraph launch (default 8): launches are enqueued back to back,
                                at most two ahead of a finished one, and stop once no variant runs; 0: the
                                whole solve up to maxiter in one launch (the sweeps after the last variant
                                stopped still dispatch their blocks) */
    uint32_t use_graph;      /* 0: launch the sweeps' kernels directly */
    uint32_t layout;         /* reserved, 0 */
} flbp_batch_opts;

typedef struct flbp_batch_group {
    const flbp_csr *graph;       /* the structure (copied at creation) */
    uint32_t n_variants;         /* K */
    const uint64_t *clamp_off;   /* [K+1], NULL for no clamps: variant k clamps clamp_var/clamp_val
                                    [clamp_off[k] .. clamp_off[k+1]) */
    const uint64_t *clamp_var;   /* variable indices of the graph, without body uses */
    const uint8_t *clamp_val;    /* 0 or 1 */
} flbp_batch_group;

void flbp_batch_default_opts(flbp_batch_opts *o);
int flbp_batch_create_groups(uint64_t n_groups, const flbp_batch_group *groups, int device,
                             const flbp_batch_opts *opts, flbp_batch **out);
/* independent graphs: one group of one variant per graph */
int flbp_batch_create(uint64_t n_graphs, const flbp_csr *const *graphs, int device,
                      const flbp_batch_opts *opts, flbp_batch **out);
void flbp_batch_destroy(flbp_batch *b);
uint64_t flbp_batch_n_variants(const flbp_batch *b);
/* CausalBP::init for every variant */
int flbp_batch_init(flbp_batch *b, int logdomain);
/* CausalBP::run() for every variant: iters_out[NV] / maxdiff_out[NV] per
 * variant (NULL skips); total->sweeps = sweeps of the longest-running one */
int flbp_batch_run(flbp_batch *b, const flbp_run_opts *opts, uint64_t *iters_out,
                   double *maxdiff_out, flbp_run_result *total);
/* beliefs of variant v, in its graph's variable order */
int flbp_batch_beliefs(flbp_batch *b, uint64_t v, double *out, uint64_t n);
/* out[i] = belief of variable vars[i] of variant v, gathered on the device */
int flbp_batch_beliefs_at(flbp_batch *b, uint64_t v, uint64_t n, const uint64_t *vars, double *out);
/* p of every factor of group g's structure (shared by its variants) */
int flbp_batch_set_p(flbp_batch *b, uint64_t g, const double *p, uint64_t n);
/* new options; new tile bounds re-cut the tiles, and the messages a run
 * has set move with them, so a solve may continue under the new bounds */
int flbp_batch_set_opts(flbp_batch *b, const flbp_batch_opts *opts);
int flbp_batch_get_opts(flbp_batch *b, flbp_batch_opts *opts);
/* iteration counters and beliefs back to their creation values */
int flbp_batch_reset_counters(flbp_batch *b);
/* times trial_sweeps sweeps for each pair of candidate factor and variable
 * tile bounds (128, 256, 512 slots), keeps the fastest, then resets
 * counters and re-inits */
int flbp_batch_autotune(flbp_batch *b, double damping, uint32_t trial_sweeps,
                        flbp_batch_opts *best);
uint64_t flbp_batch_device_bytes(const flbp_batch *b);

/* ------------------------------------------------------------------ */
/* Adjoint (not in the reference)                                      */
/* ------------------------------------------------------------------ */

/* Reverse-mode gradient through n_sweeps damped PARALL sweeps from init:
 * the forward keeps the message state every checkpoint_every sweeps, the
 * backward recomputes each segment from its checkpoint and runs the
 * hand-written reverse kernels of the factor and variable passes.  The
 * adjoint works on its plan's device buffers (the forward overwrites the
 * plan's messages and beliefs; creation switches the plan to kernel
 * variant 0, whose message layout the reverse kernels read, and the plan
 * must stay there while the adjoint is used).  Factors of any degree are
 * differentiated (degrees up to 6 by exact-degree code, wider ones in
 * O(degree) with prefix / suffix products).  Destroy an adjoint before its
 * plan. */
typedef struct flbp_adjoint flbp_adjoint;

int flbp_adjoint_create(flbp_plan *plan, uint64_t n_sweeps, uint32_t checkpoint_every,
                        double damping, flbp_adjoint **out);
void flbp_adjoint_destroy(flbp_adjoint *adj);
/* init + n_sweeps sweeps with the plan's current p; beliefs[V] (NULL skips;
 * flbp_plan_beliefs_at then reads selected beliefs) */
int flbp_adjoint_forward(flbp_adjoint *adj, double *beliefs);
/* dl_db[V] = dL/d(belief) of every variable -> dl_dp[F] = dL/dp of every
 * factor, for the last forward */
int flbp_adjoint_backward(flbp_adjoint *adj, const double *dl_db, double *dl_dp);
/* device time of the last forward and backward, milliseconds */
/* The backward with a sparse loss gradient: dL/db is dl_db[i] at variable
 * vars[i] (variable order; repeated variables add) and 0 elsewhere.  dL/dp
 * stays on the device and is ADDED to the adjoint's accumulator, so a
 * training loop can sum over variants without moving V or F values; the
 * accumulator is cleared by flbp_adjoint_clear_grad and overwritten by
 * flbp_adjoint_backward. */
int flbp_adjoint_backward_sparse(flbp_adjoint *adj, uint64_t n, const uint64_t *vars, const double *dl_db);
int flbp_adjoint_clear_grad(flbp_adjoint *adj);
/* the accumulated dL/dp[F] */
int flbp_adjoint_get_dl_dp(flbp_adjoint *adj, double *dl_dp);
/* rule[F] (-1 for untied factors, else < n_rules): the ties of
 * flbp_tied_p, kept on the device for flbp_adjoint_rule_grad */
int flbp_adjoint_set_rules(flbp_adjoint *adj, const int32_t *rule, uint64_t n_rules);
/* dL/dtheta[n_rules] from the accumulated dL/dp and the plan's current p, as
 * flbp_rule_grad computes it, reduced on the device (one deterministic
 * reduction per rule) */
int flbp_adjoint_rule_grad(flbp_adjoint *adj, double *dl_dtheta);
int flbp_adjoint_last_timing(flbp_adjoint *adj, double *forward_ms, double *backward_ms);
uint64_t flbp_adjoint_device_bytes(const flbp_adjoint *adj);

/* p_out[I] = sigmoid(theta[rule_of_factor[I]]) = 1 / (1 + exp(-theta)) for
 * rule >= 0, else base_p[I] */
int flbp_tied_p(uint64_t n_factors, const double *base_p, const int32_t *rule_of_factor,
                uint64_t n_rules, const double *theta, double *p_out);

/* dL/dtheta[r] = sum over factors I with rule_of_factor[I] == r of
 * dl_dp[I] * p[I] * (1 - p[I])  (p = sigmoid(theta[rule])), in factor order */
int flbp_rule_grad(uint64_t n_factors, const double *dl_dp, const double *p,
                   const int32_t *rule_of_factor, uint64_t n_rules, double *dl_dtheta);

const char *flbp_last_error(void);
const char *flbp_version(void);

#ifdef __cplusplus
}
#endif
#endif
