/*
 * seq_host.cpp -- host side of the sequential update schedules
 * (updates = SEQFIX / TOPO): the update orders, their split into waves,
 * and the dependency levels the GPU runs them by.
 *
 * Reference behaviour (csrc/ext/causal_bp.cu):
 *   - SEQFIX order, buildFixedUpdateSeq 186-201: the factor indices shuffled
 *     with std::shuffle by the thread's std::mt19937 rnd_gen (seeded 42 at
 *     thread start, causal_bp.cu:116, never reseeded by construction), then
 *     each factor's edges in nbF order.
 *   - TOPO order, buildTopoUpdateSeq 203-304: BFS order of the undirected
 *     factor graph, edges sorted by (later endpoint's rank, earlier
 *     endpoint's rank, factor, variable, dual).
 *   - waves, putUpdateSeqToKernelFused 583-722: the order is cut into waves;
 *     a wave runs its v->f messages first (each from the state before the
 *     wave), then its f->v messages, each of which updates its target's
 *     product in place.  A new wave starts when the next f->v update I -> i
 *     targets a variable the wave already updates, or needs a v->f message
 *     j -> I the wave has not scheduled while the wave already updates an
 *     f->v message into j from another factor (with respect_preorder: one
 *     ranked before I -> i).
 *
 * The reference launches two kernels per wave.  Its results are fixed by
 * the waves alone (the items of one kernel touch disjoint outputs), so any
 * execution that gives every item the same inputs gives the same bits.
 * Here each item gets a level: one more than the highest level of the
 * earlier items (in wave order) that write what it reads (read after
 * write), read or write what it writes (write after read / write).  Items
 * of one level are independent and run in one launch; SEQFIX on a
 * 1M-variable graph has a few thousand waves but under a hundred levels.
 *
 * Because an f->v update is the only writer of its target's state and a
 * wave updates each variable at most once, "does the wave update an f->v
 * message into j from a factor other than I" is one lookup of the factor
 * that updates j in this wave, not a scan of j's edges.
 */
#include <algorithm>
#include <cstdint>
#include <queue>
#include <random>
#include <string>
#include <vector>

#include "fastlbp.h"
#include "flbp_internal.h"
#include "seq_internal.h"

namespace lbp {
/* the reference's global generator (include/utils/utils.h:133-136,
 * csrc/ext/causal_bp.cu:116) */
thread_local std::mt19937 rnd_gen(42U);
}  // namespace lbp

namespace flbp {

namespace {

std::vector<uint32_t> slot_factors(const flbp_csr &c) {
    std::vector<uint32_t> f(c.n_edges);
    for (uint64_t I = 0; I < c.n_factors; I++)
        for (uint64_t s = c.fac_off[I]; s < c.fac_off[I + 1]; s++) f[s] = (uint32_t)I;
    return f;
}

/* var-major position of each factor-major slot */
std::vector<uint64_t> slot_positions(const flbp_csr &c) {
    std::vector<uint64_t> pos(c.n_edges);
    for (uint64_t k = 0; k < c.n_edges; k++) pos[c.var_edge[k]] = k;
    return pos;
}

}  // namespace

void seqfix_order(const flbp_csr &c, std::mt19937 &g, std::vector<uint64_t> &out) {
    /* the reference shuffles a std::vector<int> of factor indices */
    std::vector<int> numbers(c.n_factors);
    for (uint64_t I = 0; I < c.n_factors; I++) numbers[I] = (int)I;
    std::shuffle(numbers.begin(), numbers.end(), g);
    const std::vector<uint64_t> pos = slot_positions(c);
    out.clear();
    out.reserve(c.n_edges);
    for (int I : numbers)
        for (uint64_t s = c.fac_off[I]; s < c.fac_off[I + 1]; s++) out.push_back(pos[s]);
}

void topo_order(const flbp_csr &c, std::vector<uint64_t> &out) {
    const uint64_t V = c.n_vars, F = c.n_factors, N = V + F;
    /* adjacency of the bipartite graph in CSR form, each list ascending:
     * a variable's factors are ascending in var_edge order (nbV order), a
     * factor's variables ascending in slot order (nbF order) */
    std::vector<uint64_t> rank(N, N);
    std::vector<uint32_t> sf = slot_factors(c);
    std::vector<uint64_t> nbr;
    uint64_t next = 0;
    std::vector<uint64_t> q;
    q.reserve(N);
    for (uint64_t root = 0; root < N; root++) {
        if (rank[root] != N) continue;
        q.clear();
        q.push_back(root);
        rank[root] = next++;
        for (size_t h = 0; h < q.size(); h++) {
            const uint64_t x = q[h];
            nbr.clear();
            if (x < V) {
                for (uint64_t k = c.var_off[x]; k < c.var_off[x + 1]; k++) nbr.push_back(V + sf[c.var_edge[k]]);
            } else {
                const uint64_t I = x - V;
                for (uint64_t s = c.fac_off[I]; s < c.fac_off[I + 1]; s++) nbr.push_back(c.edge_var[s]);
            }
            std::sort(nbr.begin(), nbr.end());
            nbr.erase(std::unique(nbr.begin(), nbr.end()), nbr.end());
            for (uint64_t y : nbr) {
                if (rank[y] != N) continue;
                rank[y] = next++;
                q.push_back(y);
            }
        }
    }
    const std::vector<uint64_t> pos = slot_positions(c);
    /* every (factor, variable) pair is distinct, so the 5-key order is
     * total and a plain sort equals the reference's stable sort */
    struct Key {
        uint64_t target, source, factor, var, dual, edge;
        bool operator<(const Key &o) const {
            if (target != o.target) return target < o.target;
            if (source != o.source) return source < o.source;
            if (factor != o.factor) return factor < o.factor;
            if (var != o.var) return var < o.var;
            return dual < o.dual;
        }
    };
    std::vector<Key> keys;
    keys.reserve(c.n_edges);
    for (uint64_t I = 0; I < F; I++) {
        for (uint64_t s = c.fac_off[I]; s < c.fac_off[I + 1]; s++) {
            const uint64_t i = c.edge_var[s], e = pos[s];
            const uint64_t a = rank[V + I], b = rank[i];
            keys.push_back({std::max(a, b), std::min(a, b), I, i, e - c.var_off[i], e});
        }
    }
    std::sort(keys.begin(), keys.end());
    out.clear();
    out.reserve(keys.size());
    for (const Key &k : keys) out.push_back(k.edge);
}

int build_schedule(const flbp_csr &c, const uint64_t *order, uint64_t n, bool respect_preorder, Schedule &S) {
    const uint64_t V = c.n_vars, E = c.n_edges;
    if (n && !order) return fail(FLBP_EINVAL, "schedule: null order");
    for (uint64_t k = 0; k < n; k++)
        if (order[k] >= E) return fail(FLBP_EINVAL, "schedule: order names an edge out of range");
    if (n >= (1ull << 32) - 1) return fail(FLBP_ERANGE, "schedule: order longer than 2^32-2");
    S = Schedule();
    S.V = V;
    S.F = c.n_factors;
    S.E = E;
    S.respect_preorder = respect_preorder;
    S.order.assign(order, order + n);

    const std::vector<uint32_t> sf = slot_factors(c);
    std::vector<uint64_t> vm_var(E);
    for (uint64_t i = 0; i < V; i++)
        for (uint64_t k = c.var_off[i]; k < c.var_off[i + 1]; k++) vm_var[k] = i;
    std::vector<uint64_t> rank(E, 0);
    if (respect_preorder)
        for (uint64_t k = 0; k < n; k++) rank[order[k]] = k;

    /* membership stamps: equal to the current wave number = in the wave */
    std::vector<uint32_t> v_stamp(V, 0), vf_stamp(E, 0);
    std::vector<uint32_t> v_fac(V, 0);      /* factor updating the variable in the wave */
    std::vector<uint64_t> v_edge(V, 0);     /* the edge doing it */
    const std::vector<uint64_t> pos = slot_positions(c);
    uint32_t wave = 1;
    S.fv_wave_off.push_back(0);
    S.vf_wave_off.push_back(0);
    for (uint64_t k = 0; k < n; k++) {
        const uint64_t e = order[k];
        const uint64_t i = vm_var[e];
        const uint32_t I = sf[c.var_edge[e]];
        bool conflict = v_stamp[i] == wave;
        for (uint64_t s = c.fac_off[I]; s < c.fac_off[I + 1] && !conflict; s++) {
            const uint64_t j = c.edge_var[s];
            if (j == i || vf_stamp[s] == wave) continue;
            if (v_stamp[j] == wave && v_fac[j] != I && (!respect_preorder || rank[v_edge[j]] < rank[e]))
                conflict = true;
        }
        if (conflict) {
            S.fv_wave_off.push_back(S.fv_edge.size());
            S.vf_wave_off.push_back(S.vf_slot.size());
            if (++wave == 0) return fail(FLBP_ERANGE, "schedule: more than 2^32-1 waves");
        }
        v_stamp[i] = wave;
        v_fac[i] = I;
        v_edge[i] = e;
        if (S.vf_slot.size() + S.fv_edge.size() + (c.fac_off[I + 1] - c.fac_off[I]) >= (1ull << 32) - 1)
            return fail(FLBP_ERANGE, "schedule: more than 2^32-2 items");
        S.fv_edge.push_back((uint32_t)e);
        for (uint64_t s = c.fac_off[I]; s < c.fac_off[I + 1]; s++) {
            if (c.edge_var[s] == i || vf_stamp[s] == wave) continue;
            vf_stamp[s] = wave;
            S.vf_slot.push_back((uint32_t)s);
        }
    }
    S.fv_wave_off.push_back(S.fv_edge.size());
    S.vf_wave_off.push_back(S.vf_slot.size());
    S.n_waves = S.fv_wave_off.size() - 1;

    /* Execution items and their dependency levels.  Per wave, the f->v
     * updates of one factor (of degree <= kSeqGroupDeg) form one group
     * item, which also recomputes the wave's v->f messages of that factor
     * -- from the variables' states as they stand before the wave, since it
     * reads them before its own updates -- unless another factor updates
     * that variable in the same wave: such a v->f message stays a separate
     * item.  Wider factors keep one item per v->f and per f->v message.
     *
     * Levels are assigned in wave order (a wave's separate v->f items
     * before the rest).  Locations: a variable's state (product, zero
     * counts, the f->v messages into it) and a factor-major v->f slot.  An
     * item's level is one above every earlier item that writes what it
     * reads or reads or writes what it writes.  Within a wave the groups
     * and wide f->v items touch disjoint locations: each variable is
     * updated at most once, a group reads only states no other item of the
     * wave writes, and each factor's slots belong to one item. */
    S.slot_var.assign(c.edge_var, c.edge_var + E);
    S.slot_pos.assign(pos.begin(), pos.end());
    std::vector<uint32_t> w_var(V, 0), r_var(V, 0), w_slot(E, 0), r_slot(E, 0);
    S.vf_level.assign(S.vf_slot.size(), 0);
    S.fv_level.assign(S.fv_edge.size(), 0);
    std::vector<uint32_t> tgt_stamp(V, 0), tgt_fac(V, 0);   /* factor updating the variable in the wave */
    std::vector<uint32_t> grp_stamp(c.n_factors, 0), grp_of(c.n_factors, 0);
    struct Group {
        uint32_t I, tmask, imask;
        std::vector<uint64_t> fv_x, vf_x;   /* reference items folded in */
    };
    std::vector<Group> groups;
    std::vector<uint64_t> wide_fv;
    uint32_t top = 0;
    auto emit = [&](uint8_t kind, uint32_t a, uint32_t b, uint32_t cc, uint32_t L) {
        S.ex_kind.push_back(kind);
        S.ex_a.push_back(a);
        S.ex_b.push_back(b);
        S.ex_c.push_back(cc);
        S.ex_level.push_back(L);
        top = std::max(top, L);
    };
    for (uint64_t w = 0; w < S.n_waves; w++) {
        const uint32_t stamp = (uint32_t)w + 1;
        groups.clear();
        wide_fv.clear();
        for (uint64_t x = S.fv_wave_off[w]; x < S.fv_wave_off[w + 1]; x++) {
            const uint32_t e = S.fv_edge[x];
            const uint32_t own = (uint32_t)c.var_edge[e], I = sf[own];
            tgt_stamp[vm_var[e]] = stamp;
            tgt_fac[vm_var[e]] = I;
            if (c.fac_off[I + 1] - c.fac_off[I] > kSeqGroupDeg) {
                wide_fv.push_back(x);
                continue;
            }
            if (grp_stamp[I] != stamp) {
                grp_stamp[I] = stamp;
                grp_of[I] = (uint32_t)groups.size();
                groups.push_back({I, 0u, 0u, {}, {}});
            }
            Group &G = groups[grp_of[I]];
            G.tmask |= 1u << (own - c.fac_off[I]);
            G.fv_x.push_back(x);
        }
        /* the wave's v->f messages: folded into their factor's group when
         * safe, else separate items (levels first, as they run first) */
        for (uint64_t x = S.vf_wave_off[w]; x < S.vf_wave_off[w + 1]; x++) {
            const uint32_t s = S.vf_slot[x], I = sf[s];
            const uint64_t j = c.edge_var[s];
            const bool wide = c.fac_off[I + 1] - c.fac_off[I] > kSeqGroupDeg;
            if (!wide && !(tgt_stamp[j] == stamp && tgt_fac[j] != I)) {
                Group &G = groups[grp_of[I]];
                G.imask |= 1u << (s - c.fac_off[I]);
                G.vf_x.push_back(x);
                continue;
            }
            const uint32_t L = std::max(w_var[j], std::max(w_slot[s], r_slot[s])) + 1;
            r_var[j] = std::max(r_var[j], L);
            w_slot[s] = L;
            S.vf_level[x] = L;
            emit(kExecVF, s, 0u, 0u, L);
        }
        for (const Group &G : groups) {
            const uint64_t a = c.fac_off[G.I], d = c.fac_off[G.I + 1] - a;
            const bool multi = (G.tmask & (G.tmask - 1)) != 0;
            uint32_t L = 0;
            for (uint64_t k = 0; k < d; k++) {
                const uint64_t s = a + k, j = c.edge_var[s];
                const bool inl = (G.imask >> k) & 1u, tgt = (G.tmask >> k) & 1u;
                if (inl) L = std::max(L, std::max(w_var[j], std::max(w_slot[s], r_slot[s])));
                if (tgt) L = std::max(L, std::max(w_var[j], r_var[j]));
                if (!inl && (multi || !tgt)) L = std::max(L, w_slot[s]);   /* read from m_vf */
            }
            L += 1;
            for (uint64_t k = 0; k < d; k++) {
                const uint64_t s = a + k, j = c.edge_var[s];
                const bool inl = (G.imask >> k) & 1u, tgt = (G.tmask >> k) & 1u;
                if (inl) {
                    r_var[j] = std::max(r_var[j], L);
                    w_slot[s] = L;
                }
                if (!inl && (multi || !tgt)) r_slot[s] = std::max(r_slot[s], L);
                if (tgt) w_var[j] = L;
            }
            for (uint64_t x : G.fv_x) S.fv_level[x] = L;
            for (uint64_t x : G.vf_x) S.vf_level[x] = L;
            emit(kExecGroup, G.I, G.tmask, G.imask, L);
        }
        for (uint64_t x : wide_fv) {
            const uint32_t e = S.fv_edge[x];
            const uint64_t i = vm_var[e], own = c.var_edge[e];
            const uint32_t I = sf[own];
            uint32_t L = std::max(w_var[i], r_var[i]);
            for (uint64_t s = c.fac_off[I]; s < c.fac_off[I + 1]; s++)
                if (s != own) L = std::max(L, w_slot[s]);
            L += 1;
            for (uint64_t s = c.fac_off[I]; s < c.fac_off[I + 1]; s++)
                if (s != own) r_slot[s] = std::max(r_slot[s], L);
            w_var[i] = L;
            S.fv_level[x] = L;
            emit(kExecFV, e, 0u, 0u, L);
        }
    }
    S.n_levels = top;
    return FLBP_OK;
}

}  // namespace flbp

/* ---- C ABI: schedules ---------------------------------------------------- */

struct flbp_schedule {
    flbp::Schedule s;
};

extern "C" {

int flbp_schedule_create(const flbp_csr *c, int updates, flbp_schedule **out) {
    if (!c || !out) return flbp::fail(FLBP_EINVAL, "flbp_schedule_create: null argument");
    if (updates != FLBP_UPDATES_SEQFIX && updates != FLBP_UPDATES_TOPO)
        return flbp::fail(FLBP_EINVAL, "flbp_schedule_create: updates must be SEQFIX or TOPO");
    if (int rc = flbp::validate_csr(*c)) return rc;
    if (c->n_factors > (uint64_t)INT32_MAX || c->n_edges >= (1ull << 32) - 1 || c->n_vars >= (1ull << 32) - 1)
        return flbp::fail(FLBP_ERANGE, "flbp_schedule_create: graph too large");
    std::vector<uint64_t> order;
    if (updates == FLBP_UPDATES_SEQFIX) flbp::seqfix_order(*c, lbp::rnd_gen, order);
    else flbp::topo_order(*c, order);
    auto *h = new flbp_schedule();
    if (int rc = flbp::build_schedule(*c, order.data(), order.size(), updates == FLBP_UPDATES_TOPO, h->s)) {
        delete h;
        return rc;
    }
    h->s.updates = updates;
    *out = h;
    return FLBP_OK;
}

int flbp_schedule_from_order(const flbp_csr *c, const uint64_t *order, uint64_t n, int respect_preorder,
                             flbp_schedule **out) {
    if (!c || !out) return flbp::fail(FLBP_EINVAL, "flbp_schedule_from_order: null argument");
    if (int rc = flbp::validate_csr(*c)) return rc;
    if (c->n_edges >= (1ull << 32) - 1 || c->n_vars >= (1ull << 32) - 1)
        return flbp::fail(FLBP_ERANGE, "flbp_schedule_from_order: graph too large");
    auto *h = new flbp_schedule();
    if (int rc = flbp::build_schedule(*c, order, n, respect_preorder != 0, h->s)) {
        delete h;
        return rc;
    }
    h->s.updates = -1;
    *out = h;
    return FLBP_OK;
}

void flbp_schedule_destroy(flbp_schedule *s) { delete s; }

int flbp_schedule_get_info(const flbp_schedule *h, flbp_schedule_info *info) {
    if (!h || !info) return flbp::fail(FLBP_EINVAL, "flbp_schedule_get_info: null argument");
    info->n_order = h->s.order.size();
    info->n_waves = h->s.n_waves;
    info->n_fv_items = h->s.fv_edge.size();
    info->n_vf_items = h->s.vf_slot.size();
    info->n_levels = h->s.n_levels;
    info->updates = h->s.updates;
    info->respect_preorder = h->s.respect_preorder ? 1 : 0;
    info->n_exec_items = h->s.ex_kind.size();
    return FLBP_OK;
}

int flbp_schedule_exec(const flbp_schedule *h, uint8_t *kind, uint32_t *a, uint32_t *b, uint32_t *c,
                       uint32_t *level) {
    if (!h) return flbp::fail(FLBP_EINVAL, "flbp_schedule_exec: null schedule");
    const flbp::Schedule &S = h->s;
    if (kind) std::copy(S.ex_kind.begin(), S.ex_kind.end(), kind);
    if (a) std::copy(S.ex_a.begin(), S.ex_a.end(), a);
    if (b) std::copy(S.ex_b.begin(), S.ex_b.end(), b);
    if (c) std::copy(S.ex_c.begin(), S.ex_c.end(), c);
    if (level) std::copy(S.ex_level.begin(), S.ex_level.end(), level);
    return FLBP_OK;
}

int flbp_schedule_order(const flbp_schedule *h, uint64_t *order) {
    if (!h || (!order && !h->s.order.empty())) return flbp::fail(FLBP_EINVAL, "flbp_schedule_order: null argument");
    std::copy(h->s.order.begin(), h->s.order.end(), order);
    return FLBP_OK;
}

int flbp_schedule_waves(const flbp_schedule *h, uint64_t *fv_off, uint64_t *fv_edge, uint64_t *vf_off,
                        uint64_t *vf_slot) {
    if (!h) return flbp::fail(FLBP_EINVAL, "flbp_schedule_waves: null schedule");
    const flbp::Schedule &S = h->s;
    if (fv_off) std::copy(S.fv_wave_off.begin(), S.fv_wave_off.end(), fv_off);
    if (vf_off) std::copy(S.vf_wave_off.begin(), S.vf_wave_off.end(), vf_off);
    if (fv_edge) std::copy(S.fv_edge.begin(), S.fv_edge.end(), fv_edge);
    if (vf_slot) std::copy(S.vf_slot.begin(), S.vf_slot.end(), vf_slot);
    return FLBP_OK;
}

int flbp_schedule_levels(const flbp_schedule *h, uint32_t *fv_level, uint32_t *vf_level) {
    if (!h) return flbp::fail(FLBP_EINVAL, "flbp_schedule_levels: null schedule");
    if (fv_level) std::copy(h->s.fv_level.begin(), h->s.fv_level.end(), fv_level);
    if (vf_level) std::copy(h->s.vf_level.begin(), h->s.vf_level.end(), vf_level);
    return FLBP_OK;
}

void flbp_seed(uint32_t seed) { lbp::rnd_gen.seed(seed); }

}  // extern "C"

namespace flbp {
const Schedule *schedule_of(const flbp_schedule *h) { return h ? &h->s : nullptr; }
}  // namespace flbp
