/*
 * plan_internal.h -- the plain-solve plan (device state of one graph).
 */
#ifndef FLBP_PLAN_INTERNAL_H
#define FLBP_PLAN_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "fastlbp.h"
#include "flbp_internal.h"

namespace flbp {

namespace dev {
/* prod_fv_0/1 and num_zeros_fv_0/1 of one variable under a sequential
 * schedule (seq_kernels.h) */
struct __align__(32) VarState {
    double p0, p1;
    unsigned long long z0, z1;
};
/* a factor's numbers for the sequential kernels in one 32-byte record:
 * prob, prob_default (0 for priors), head mask ((1, 1) unless clamped) */
struct __align__(32) FacRec {
    double p, q, m0, m1;
};
/* the state layout (kernel variant 4, state_kernels.h) keeps each
 * variable's num_zeros_fv_0/1 (calcProdAllKernel) in one code byte,
 * min(z0, 2) | min(z1, 2) << 2: the leave-one-out and the belief only ask
 * whether a count is 0, 1 or more */
__host__ __device__ inline uint8_t pack_zc(uint32_t z0, uint32_t z1) {
    return (uint8_t)((z0 < 2 ? z0 : 2u) | ((z1 < 2 ? z1 : 2u) << 2));
}
/* ... and its prod_fv_0/1 in 8 bytes.  calcProdAllKernel rescales the pair
 * after every factor (causal_bp_parall.cu:14-18, 37-47; scale_pair): the
 * larger value becomes the unit -- exactly 1.0, or the quiet NaN
 * 0x7ff8000000000000 when the larger value was 0 or not finite -- and only
 * the other value carries information.  A variable without edges keeps
 * (1, 1).  So the record stores the other value, and two more bits of the
 * code byte say which component is the unit and whether it is the NaN;
 * unpack_prod returns the pair bit for bit. */
constexpr uint32_t kProdUnitX = 16u, kProdUnitNan = 32u;
constexpr uint64_t kUnitOneBits = 0x3ff0000000000000ull, kUnitNanBits = 0x7ff8000000000000ull;
__host__ __device__ inline uint8_t pack_prod(double x, double y, uint32_t z0, uint32_t z1, double &rest) {
    const uint64_t bx = __builtin_bit_cast(uint64_t, x), by = __builtin_bit_cast(uint64_t, y);
    uint32_t f;
    if (bx == kUnitOneBits) {
        f = kProdUnitX;
        rest = y;
    } else if (by == kUnitOneBits) {
        f = 0;
        rest = x;
    } else if (bx == kUnitNanBits) {
        f = kProdUnitX | kProdUnitNan;
        rest = y;
    } else {
        f = kProdUnitNan;
        rest = x;
    }
    return (uint8_t)(pack_zc(z0, z1) | f);
}
__host__ __device__ inline double2 unpack_prod(double rest, uint32_t code) {
    const double unit = (code & kProdUnitNan) ? __builtin_bit_cast(double, kUnitNanBits) : 1.0;
    return (code & kProdUnitX) ? make_double2(unit, rest) : make_double2(rest, unit);
}
/* where the state layout's next factor pass takes its v->f inputs from:
 * the product records, the init value, or a var-major array */
enum : int { kVfState = 0, kVfConst = 1, kVfArray = 2 };
}  // namespace dev

constexpr int kMaxChunk = 64;   /* sweeps enqueued between host checks */
constexpr uint32_t kStageItems = 256;    /* factors / variables per staged block (== kBlock) */
constexpr uint32_t kStageSlots = 512;    /* edge slots per staged block (== kStage) */
/* edge slots per variable tile of the state layout: its var pass stages no
 * outputs, so a tile can gather more messages at once (on the 1M-variable
 * graph 36.0 us per launch against 38.1 with 512-slot tiles; 640: 37.2,
 * 896: 38.2, 1,024: 40.4) */
constexpr uint32_t kStateVarSlots = 768;
/* a plan's default layout: the scatter layout up to this footprint, the
 * state layout above it (plan_host.hip has the measurements) */
constexpr uint64_t kStateMinBytes = 64ull << 20;

inline unsigned grid_for(uint64_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

int hip_fail(hipError_t e, const char *what);
int validate_csr(const flbp_csr &c);
/* device order of variables: perm[device index] = variable index;
 * dev_off/dev_edge = var_off/var_edge in that order; vpos[slot] = the slot's
 * var-major position in that order.  kOrderLabel keeps label order (each
 * variable near the factors it heads); kOrderDegree sorts by degree,
 * heaviest first; kOrderHotFirst puts variables of degree >= kHotDegree
 * (those many factors read: premises shared by many rule bodies) ahead of
 * the rest, label order within each class.  The state layout's factor pass
 * reads each slot's variable record at random, and the hot class packed
 * together keeps those reads in fewer lines: on a 1M-variable graph (E =
 * 4.5M) its factor pass took 69.3 us against 79.8 us in label order (full
 * degree order: 69.8 us, with a 5 us slower var pass). */
enum : int { kOrderLabel = 0, kOrderDegree = 1, kOrderHotFirst = 2 };
constexpr uint64_t kHotDegree = 3;
/* ranges [blk[b], blk[b+1]) of at most max_items items and max_slots
 * slots each (an item above max_slots alone) */
void cut_blocks(const uint32_t *off, uint64_t n_items, uint32_t max_items, uint32_t max_slots,
                std::vector<uint32_t> &blk);
/* work orders of the staged kernels (see staged_kernels.h): for each block
 * of at most kStageSlots slots, fac_work lists its slots (local slot | local
 * factor << 10) grouped by output kind (prior, head, body) and factor
 * degree; var_work lists its variables (local index) by degree.  Given
 * slot_var (device variable of each slot), the body outputs of factors of
 * degree <= reg_deg (which all run the same register path) form one group,
 * ordered by variable: the state layout stores messages in work order, and
 * bodies of nearby variables then share lines. */
void stage_work_orders(const uint32_t *fac_off, const uint32_t *fac_meta, const std::vector<uint32_t> &fb,
                       const uint32_t *var_off, const std::vector<uint32_t> &vb, std::vector<uint32_t> &fac_work,
                       std::vector<uint16_t> &var_work, const uint32_t *slot_var = nullptr, uint32_t reg_deg = 0);
/* the variable half of stage_work_orders: per block, its variables by degree */
void var_work_order(const uint32_t *var_off, const std::vector<uint32_t> &vb, std::vector<uint16_t> &var_work);
void var_device_order(const flbp_csr &c, int order, std::vector<uint32_t> &perm,
                      std::vector<uint32_t> &dev_off, std::vector<uint32_t> &dev_edge,
                      std::vector<uint32_t> &vpos);
/* A CSR with its factors in device order: factors sorted (stably) by the
 * device index of their head variable, so that a head's factors sit
 * together and the factor pass walks head variables in device order.  Each
 * factor keeps its slots in their order and each variable its edges in
 * their nbV order, so every message and product is computed exactly as in
 * the reference's order.  The plan works in this order throughout; the
 * entry points translate factor indices and factor-major slot positions. */
struct DeviceCsr {
    std::vector<uint64_t> fac_off, edge_var, var_edge, head_slot;
    std::vector<int32_t> fac_type;
    std::vector<uint8_t> head_clamped;
    std::vector<double> p, q, mask0, mask1;
    flbp_csr view{};
};
/* fperm[d] = the reference factor at device position d, finv its inverse;
 * slot_dev[s] = device slot of reference slot s, slot_ref its inverse */
void factor_device_order(const flbp_csr &c, const std::vector<uint32_t> &var_inv, DeviceCsr &out,
                         std::vector<uint32_t> &fperm, std::vector<uint32_t> &finv, std::vector<uint32_t> &slot_dev,
                         std::vector<uint32_t> &slot_ref);

struct Plan;
int seq_reset_state(Plan *P, double c);   /* seq.hip */
int seq_refresh_factors(Plan *P);         /* seq.hip: factor records after p / clamp changes */

struct Plan {
    int device = -1;
    hipStream_t stream = nullptr;
    uint64_t V = 0, F = 0, E = 0;

    uint32_t *fac_off = nullptr, *fac_meta = nullptr, *var_off = nullptr, *var_edge = nullptr;
    uint32_t *fac_vpos = nullptr;
    uint32_t *fac_blk = nullptr, *var_blk = nullptr;   /* staged block ranges */
    uint32_t *fac_work = nullptr;   /* per staged factor block: its output slots in work order */
    uint16_t *var_work = nullptr;   /* per staged variable block: its variables in work order */
    uint32_t *var_blk4 = nullptr;   /* the state layout's variable tiles (kStateVarSlots) */
    uint16_t *var_work4 = nullptr;  /* ... and their work orders */
    uint64_t n_var_blocks4 = 0;
    uint64_t n_fac_blocks = 0, n_var_blocks = 0;
    int variant = 3;   /* 0 staged pull, 1 one lane per item, 2 staged push, 3 staged scatter, 4 staged state */
    double *vprod = nullptr;        /* variant 4: per-variable products (device order, pack_prod) */
    uint8_t *vzc = nullptr;         /* variant 4: per-variable codes (pack_prod) */
    uint32_t *fac_var = nullptr;    /* variant 4: device variable at each storage position */
    uint32_t *var_src = nullptr;    /* variant 4: storage position of each var-major edge */
    int vf_src = dev::kVfConst;     /* variant 4: v->f inputs of the next factor pass */
    int materialize_vf();           /* variant 4: the v->f messages into m_vf, storage order (plan.hip) */
    /* position in m_vf of the v->f message of device slot s: var-major
     * (variants 0, 1), factor-major (2, 3) or in storage order (4) */
    uint64_t vf_pos(uint64_t s) const {
        return variant == 4 ? h_store[s] : (variant == 2 || variant == 3) ? s : h_fac_vpos[s];
    }
    double *fac_p = nullptr, *fac_q = nullptr, *belief = nullptr;
    double2 *fac_mask = nullptr, *m_vf = nullptr, *m_fv = nullptr;
    double2 *m_fv_vm = nullptr;    /* var-major f->v messages (variant 3) */
    bool fm_stale = false;         /* variant 3: factor-major f->v copy out of date */
    bool vm_stale = false;         /* under a schedule: var-major f->v copy out of date (its sweeps
                                      update the factor-major array only) */
    int fm_current();              /* bring either copy up to date from the other (plan.hip) */
    int vm_current();
    unsigned long long *ctl = nullptr;             /* partial maxima, see sweep_kernels.h */
    unsigned long long *ctl_out = nullptr;         /* per-sweep maxima of the last chunk */
    unsigned long long *host_ctl = nullptr;        /* pinned */
    std::vector<void *> allocs;
    uint64_t device_bytes = 0;
    uint64_t max_fac_deg = 0, max_var_deg = 0;

    std::vector<uint64_t> h_var_edge;              /* for message export/import */
    std::vector<uint32_t> h_perm;                  /* device variable -> variable */
    uint32_t *var_perm = nullptr;                  /* the same on the device */
    std::vector<uint32_t> h_inv_perm;              /* variable -> device variable */
    std::vector<uint32_t> h_fperm, h_finv;         /* device factor <-> factor (factor_device_order) */
    std::vector<uint32_t> h_slot_dev, h_slot_ref;  /* factor-major slot <-> device slot */
    uint32_t *stage_idx = nullptr;                 /* small index / value uploads (grown on demand) */
    double *stage_val = nullptr;
    uint64_t stage_cap = 0;
    int stage(uint64_t n);
    std::vector<uint32_t> h_fac_vpos;              /* factor-major slot -> var-major position */
    std::vector<uint32_t> h_store;                 /* factor-major slot -> variant 4 storage position */
    std::vector<uint32_t> h_meta, h_meta0;         /* factor meta words: current and at creation */
    std::vector<double2> h_mask, h_mask0;          /* head masks: current and at creation */
    std::vector<uint64_t> clamp_dirty;             /* factors whose meta / mask differ from creation */
    std::vector<uint32_t> h_fac_head;              /* head variable of each factor */
    std::vector<uint8_t> h_var_body;               /* 1: the variable appears in some factor body */
    std::vector<uint32_t> h_head_off, h_head_fac;  /* factors headed by each variable (CSR, factor order) */

    /* sequential schedule (updates = SEQFIX / TOPO, seq.hip); none = PARALL */
    bool seq = false;
    int seq_updates = -1;                          /* FLBP_UPDATES_* or -1 (explicit order) */
    uint4 *seq_vf = nullptr, *seq_fv = nullptr;    /* items grouped by level */
    uint4 *seq_fv_aux = nullptr;                   /* f->v items: degree | type << 24, target slot, head slot */
    uint4 *seq_grp = nullptr;                      /* group items, 2 words per lane */
    dev::FacRec *seq_fac = nullptr;                /* per-factor records */
    dev::VarState *seq_state = nullptr;            /* per-variable products and zero counts */
    std::vector<uint32_t> seq_lvl_vf, seq_lvl_fv, seq_lvl_grp;  /* [levels+1] item offsets per level */
    uint32_t *seq_lvl_vf_dev = nullptr, *seq_lvl_fv_dev = nullptr, *seq_lvl_grp_dev = nullptr;
    size_t seq_tail_from = 0;                      /* levels from here run in the tail kernel */
    uint32_t seq_tail_max = kSeqTailItemsMax;      /* tail levels hold at most this many items */
    void seq_place_tail();
    uint64_t seq_waves = 0;
    std::vector<uint64_t> h_fac_off;               /* host copy of fac_off */
    int launch_seq_sweep(int k, double tol);
    void free_seq_items();

    uint64_t iters = 0;                            /* CausalBP::_iters */
    double last_maxdiff = INFINITY;
    double init_value = 1.0;
    bool swept_since_init = false;

    bool profiling = false;
    /* captured chunks: one hipGraph per (sweeps, damping, tol, variant),
     * replayed when profiling is off */
    struct ChunkGraph {
        int n, variant, vf_src;
        double damping, tol;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<ChunkGraph> graphs;
    bool use_graphs = false;   /* measured: no gain over direct launches on MI355X */
    void drop_graphs();
    int chunk_graph(int n, double tol, double damping, hipGraphExec_t &exec);
    hipEvent_t ev[2 * kMaxChunk + 1] = {};
    double prof_fac_ms = 0, prof_var_ms = 0;
    uint64_t prof_fac_n = 0, prof_var_n = 0;

    void reset_profile() {
        prof_fac_ms = prof_var_ms = 0;
        prof_fac_n = prof_var_n = 0;
    }
    int upload(const flbp_csr &c);
    int launch_sweep(int k, double tol, double damping);
    int launch_sweep_on(int k, double tol, double damping, const double2 *vf_in,
                        const double2 *fv_old, double2 *fv_out, double2 *vf_out, double *bel);
    int run_chunk(int n, double tol, double damping, int &ran, double &last_diff);
    double prev_diff = INFINITY;   /* max diff of the sweep before the last one run_chunk saw */
    ~Plan();

    template <class T>
    int alloc(T *&p, uint64_t n);
};

}  // namespace flbp

#endif
