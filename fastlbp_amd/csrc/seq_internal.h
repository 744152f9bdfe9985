/*
 * seq_internal.h -- the sequential update schedules (SEQFIX / TOPO) on the
 * host: orders, waves and dependency levels (seq_host.cpp).
 */
#ifndef FLBP_SEQ_INTERNAL_H
#define FLBP_SEQ_INTERNAL_H

#include <cstdint>
#include <random>
#include <vector>

#include "fastlbp.h"

namespace lbp {
extern thread_local std::mt19937 rnd_gen;
}

namespace flbp {

/* A split update order.  Orders and fv_edge hold var-major edge positions
 * (the reference's Edge(i, dual) -> h_row_ptr_fv[i] + dual); vf_slot holds
 * factor-major slots.  Wave w owns fv_edge[fv_wave_off[w] ..
 * fv_wave_off[w+1]) and vf_slot[vf_wave_off[w] .. vf_wave_off[w+1]); every
 * item also has a dependency level 1..n_levels. */
struct Schedule {
    uint64_t V = 0, F = 0, E = 0;
    int updates = -1;               /* FLBP_UPDATES_* it was built for, -1 = explicit order */
    bool respect_preorder = false;
    std::vector<uint64_t> order;
    uint64_t n_waves = 0;
    std::vector<uint64_t> fv_wave_off, vf_wave_off;
    std::vector<uint32_t> fv_edge, vf_slot;
    std::vector<uint32_t> fv_level, vf_level;
    uint32_t n_levels = 0;

    /* What runs on the device, each with its level:
     *   kExecVF    a = factor-major slot: one v->f message
     *   kExecFV    a = var-major edge: one f->v update (factors wider than
     *              kSeqGroupDeg)
     *   kExecGroup a = factor, b = mask of its slots updated (f->v), c = mask
     *              of its slots whose v->f message it recomputes first, from
     *              the variables' states before its own updates
     * fv_level / vf_level give the level at which each reference item is
     * computed (a v->f message folded into a group takes the group's). */
    std::vector<uint8_t> ex_kind;
    std::vector<uint32_t> ex_a, ex_b, ex_c, ex_level;
    std::vector<uint32_t> slot_var, slot_pos;   /* variable and var-major position of each slot */
};

enum : uint8_t { kExecVF = 0, kExecFV = 1, kExecGroup = 2 };

void seqfix_order(const flbp_csr &c, std::mt19937 &g, std::vector<uint64_t> &out);
void topo_order(const flbp_csr &c, std::vector<uint64_t> &out);
int build_schedule(const flbp_csr &c, const uint64_t *order, uint64_t n, bool respect_preorder, Schedule &S);
const Schedule *schedule_of(const flbp_schedule *h);

}  // namespace flbp

#endif
