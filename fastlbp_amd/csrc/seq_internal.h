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
    /* per item, what the kernels need: v->f item -> var-major position of
     * its slot and its variable; f->v item -> its slot, factor, variable */
    std::vector<uint32_t> vf_pos, vf_var;
    std::vector<uint32_t> fv_slot, fv_fac, fv_var;
    uint32_t n_levels = 0;
};

void seqfix_order(const flbp_csr &c, std::mt19937 &g, std::vector<uint64_t> &out);
void topo_order(const flbp_csr &c, std::vector<uint64_t> &out);
int build_schedule(const flbp_csr &c, const uint64_t *order, uint64_t n, bool respect_preorder, Schedule &S);
const Schedule *schedule_of(const flbp_schedule *h);

}  // namespace flbp

#endif
