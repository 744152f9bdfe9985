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
            /* nbV and nbF lists are ascending already in a graph built from
             * a factor list; other CSRs get sorted */
            if (!std::is_sorted(nbr.begin(), nbr.end())) std::sort(nbr.begin(), nbr.end());
            nbr.erase(std::unique(nbr.begin(), nbr.end()), nbr.end());
            for (uint64_t y : nbr) {
                if (rank[y] != N) continue;
                rank[y] = next++;
                q.push_back(y);
            }
        }
    }
    const std::vector<uint64_t> pos = slot_positions(c);
    /* Ranks are distinct, so (later rank, earlier rank) names one pair of
     * nodes and so one edge: the first two keys already order the edges
     * totally, and a plain sort on them equals the reference's stable sort
     * on all five.  With fewer than 2^32 nodes both ranks pack into one
     * 64-bit key. */
    if (N < (1ull << 32)) {
        std::vector<std::pair<uint64_t, uint64_t>> kv;
        kv.reserve(c.n_edges);
        for (uint64_t I = 0; I < F; I++) {
            for (uint64_t s = c.fac_off[I]; s < c.fac_off[I + 1]; s++) {
                const uint64_t a = rank[V + I], b = rank[c.edge_var[s]];
                kv.push_back({(std::max(a, b) << 32) | std::min(a, b), pos[s]});
            }
        }
        /* keys are distinct: an LSD radix sort over 16-bit digits, as many
         * passes as the largest key needs */
        uint64_t kmax = 0;
        for (const auto &x : kv) kmax = std::max(kmax, x.first);
        std::vector<std::pair<uint64_t, uint64_t>> tmp(kv.size());
        std::vector<uint64_t> cnt(65537);
        for (int shift = 0; shift < 64 && (kmax >> shift) != 0; shift += 16) {
            std::fill(cnt.begin(), cnt.end(), 0);
            for (const auto &x : kv) cnt[((x.first >> shift) & 0xffffu) + 1]++;
            for (size_t d = 0; d < 65536; d++) cnt[d + 1] += cnt[d];
            for (const auto &x : kv) tmp[cnt[(x.first >> shift) & 0xffffu]++] = x;
            kv.swap(tmp);
        }
        out.clear();
        out.reserve(kv.size());
        for (const auto &x : kv) out.push_back(x.second);
        return;
    }
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

    /* 32-bit copies of what the loops below read per edge, and one record
     * per variable / slot / factor for what they keep, so that each visit
     * of a random variable costs one cache line */
    const uint64_t F = c.n_factors;
    std::vector<uint32_t> fo(F + 1), sf(E), vm_var(E), vm_own(E), sv(E);
    for (uint64_t I = 0; I <= F; I++) fo[I] = (uint32_t)c.fac_off[I];
    for (uint64_t I = 0; I < F; I++)
        for (uint64_t s = c.fac_off[I]; s < c.fac_off[I + 1]; s++) sf[s] = (uint32_t)I;
    for (uint64_t s = 0; s < E; s++) sv[s] = (uint32_t)c.edge_var[s];
    for (uint64_t i = 0; i < V; i++)
        for (uint64_t k = c.var_off[i]; k < c.var_off[i + 1]; k++) {
            vm_var[k] = (uint32_t)i;
            vm_own[k] = (uint32_t)c.var_edge[k];
        }
    std::vector<uint32_t> rank;
    if (respect_preorder) {
        rank.assign(E, 0);
        for (uint64_t k = 0; k < n; k++) rank[order[k]] = (uint32_t)k;
    }

    /* the wave split; stamps equal to the current wave number = in the wave */
    struct VarW {
        uint32_t stamp, fac, edge;   /* updated in wave stamp by factor fac through edge */
    };
    std::vector<VarW> vw(V, VarW{0u, 0u, 0u});
    std::vector<uint32_t> vf_stamp(E, 0);
    uint32_t wave = 1;
    S.fv_wave_off.push_back(0);
    S.vf_wave_off.push_back(0);
    S.fv_edge.reserve(n);
    S.vf_slot.reserve(2 * n);
    /* Each step's variable and factor slot range, gathered in order first
     * (independent loads the core overlaps); the loop then prefetches the
     * records of the variables a few steps ahead will visit. */
    std::vector<uint32_t> ov(n), oI(n), os0(n), os1(n);
    for (uint64_t k = 0; k < n; k++) {
        const uint32_t e = (uint32_t)order[k];
        const uint32_t I = sf[vm_own[e]];
        ov[k] = vm_var[e];
        oI[k] = I;
        os0[k] = fo[I];
        os1[k] = fo[I + 1];
    }
    constexpr uint64_t kAhead = 8;
    for (uint64_t k = 0; k < n; k++) {
        if (k + 2 * kAhead < n) {
            __builtin_prefetch(&sv[os0[k + 2 * kAhead]]);
            __builtin_prefetch(&vf_stamp[os0[k + 2 * kAhead]]);
        }
        if (k + kAhead < n) {
            __builtin_prefetch(&vw[ov[k + kAhead]]);
            for (uint32_t s = os0[k + kAhead]; s < os1[k + kAhead]; s++) __builtin_prefetch(&vw[sv[s]]);
        }
        const uint32_t e = (uint32_t)order[k];
        const uint32_t i = ov[k];
        const uint32_t I = oI[k];
        const uint32_t s0 = os0[k], s1 = os1[k];
        bool conflict = vw[i].stamp == wave;
        for (uint32_t s = s0; s < s1 && !conflict; s++) {
            const uint32_t j = sv[s];
            if (j == i || vf_stamp[s] == wave) continue;
            const VarW &x = vw[j];
            if (x.stamp == wave && x.fac != I && (!respect_preorder || rank[x.edge] < rank[e])) conflict = true;
        }
        if (conflict) {
            S.fv_wave_off.push_back(S.fv_edge.size());
            S.vf_wave_off.push_back(S.vf_slot.size());
            if (++wave == 0) return fail(FLBP_ERANGE, "schedule: more than 2^32-1 waves");
        }
        vw[i] = VarW{wave, I, e};
        if (S.vf_slot.size() + S.fv_edge.size() + (s1 - s0) >= (1ull << 32) - 1)
            return fail(FLBP_ERANGE, "schedule: more than 2^32-2 items");
        S.fv_edge.push_back(e);
        for (uint32_t s = s0; s < s1; s++) {
            if (sv[s] == i || vf_stamp[s] == wave) continue;
            vf_stamp[s] = wave;
            S.vf_slot.push_back(s);
        }
    }
    S.fv_wave_off.push_back(S.fv_edge.size());
    S.vf_wave_off.push_back(S.vf_slot.size());
    S.n_waves = S.fv_wave_off.size() - 1;
    std::vector<uint32_t>().swap(vf_stamp);
    std::vector<VarW>().swap(vw);
    std::vector<uint32_t>().swap(os0);
    std::vector<uint32_t>().swap(os1);

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
    S.slot_var.assign(sv.begin(), sv.end());
    S.slot_pos.resize(E);
    for (uint64_t k = 0; k < E; k++) S.slot_pos[vm_own[k]] = (uint32_t)k;
    struct VarL {
        uint32_t w, r;             /* last level writing / reading the state */
        uint32_t tgt_stamp, tgt_fac;   /* factor updating the variable in wave tgt_stamp */
    };
    struct SlotL {
        uint32_t w, r;             /* last level writing / reading the v->f slot */
    };
    struct FacG {
        uint32_t stamp, g;         /* group index of the factor in wave stamp */
    };
    std::vector<VarL> vl(V, VarL{0u, 0u, 0u, 0u});
    std::vector<SlotL> sl(E, SlotL{0u, 0u});
    std::vector<FacG> fg(F, FacG{0u, 0u});
    S.vf_level.assign(S.vf_slot.size(), 0);
    S.fv_level.assign(S.fv_edge.size(), 0);
    struct Group {
        uint32_t I, tmask, imask, L;
    };
    std::vector<Group> groups;
    std::vector<uint32_t> fv_g, vf_g;   /* per item of the wave: its group, or ~0u */
    std::vector<uint64_t> wide_fv;
    uint32_t top = 0;
    S.ex_kind.reserve(n);
    S.ex_a.reserve(n);
    S.ex_b.reserve(n);
    S.ex_c.reserve(n);
    S.ex_level.reserve(n);
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
        const uint64_t f0 = S.fv_wave_off[w], f1 = S.fv_wave_off[w + 1];
        const uint64_t v0 = S.vf_wave_off[w], v1 = S.vf_wave_off[w + 1];
        groups.clear();
        wide_fv.clear();
        fv_g.assign(f1 - f0, ~0u);
        vf_g.assign(v1 - v0, ~0u);
        if (w + 1 < S.n_waves) {   /* the next wave's variable and factor records */
            for (uint64_t x = f1; x < S.fv_wave_off[w + 2] && x < f1 + 16; x++) {
                __builtin_prefetch(&vl[ov[x]]);
                __builtin_prefetch(&fg[oI[x]]);
            }
        }
        for (uint64_t x = f0; x < f1; x++) {
            /* fv_edge[x] is order[x]: each order step is one f->v item */
            const uint32_t e = S.fv_edge[x];
            const uint32_t own = vm_own[e], I = oI[x];
            VarL &t = vl[ov[x]];
            t.tgt_stamp = stamp;
            t.tgt_fac = I;
            if (fo[I + 1] - fo[I] > kSeqGroupDeg) {
                wide_fv.push_back(x);
                continue;
            }
            FacG &q = fg[I];
            if (q.stamp != stamp) {
                q.stamp = stamp;
                q.g = (uint32_t)groups.size();
                groups.push_back(Group{I, 0u, 0u, 0u});
            }
            groups[q.g].tmask |= 1u << (own - fo[I]);
            fv_g[x - f0] = q.g;
        }
        /* the wave's v->f messages: folded into their factor's group when
         * safe, else separate items (levels first, as they run first) */
        for (uint64_t x = v0; x < v1; x++) {
            if (x + 16 < v1) __builtin_prefetch(&sv[S.vf_slot[x + 16]]);
            if (x + 8 < v1) __builtin_prefetch(&vl[sv[S.vf_slot[x + 8]]]);
            const uint32_t s = S.vf_slot[x], I = sf[s];
            const uint32_t j = sv[s];
            const bool wide = fo[I + 1] - fo[I] > kSeqGroupDeg;
            VarL &v = vl[j];
            if (!wide && !(v.tgt_stamp == stamp && v.tgt_fac != I)) {
                const uint32_t g = fg[I].g;
                groups[g].imask |= 1u << (s - fo[I]);
                vf_g[x - v0] = g;
                continue;
            }
            SlotL &o = sl[s];
            const uint32_t L = std::max(v.w, std::max(o.w, o.r)) + 1;
            v.r = std::max(v.r, L);
            o.w = L;
            S.vf_level[x] = L;
            emit(kExecVF, s, 0u, 0u, L);
        }
        for (size_t gi = 0; gi < groups.size(); gi++) {
            Group &G = groups[gi];
            if (gi + 8 < groups.size()) {
                const uint32_t I8 = groups[gi + 8].I;
                for (uint32_t sp = fo[I8]; sp < fo[I8 + 1]; sp++) __builtin_prefetch(&vl[sv[sp]]);
            }
            const uint32_t a = fo[G.I], d = fo[G.I + 1] - a;
            const bool multi = (G.tmask & (G.tmask - 1)) != 0;
            uint32_t L = 0;
            for (uint32_t k = 0; k < d; k++) {
                const uint32_t s = a + k;
                const VarL &v = vl[sv[s]];
                const SlotL &o = sl[s];
                const bool inl = (G.imask >> k) & 1u, tgt = (G.tmask >> k) & 1u;
                if (inl) L = std::max(L, std::max(v.w, std::max(o.w, o.r)));
                if (tgt) L = std::max(L, std::max(v.w, v.r));
                if (!inl && (multi || !tgt)) L = std::max(L, o.w);   /* read from m_vf */
            }
            L += 1;
            for (uint32_t k = 0; k < d; k++) {
                const uint32_t s = a + k;
                VarL &v = vl[sv[s]];
                SlotL &o = sl[s];
                const bool inl = (G.imask >> k) & 1u, tgt = (G.tmask >> k) & 1u;
                if (inl) {
                    v.r = std::max(v.r, L);
                    o.w = L;
                }
                if (!inl && (multi || !tgt)) o.r = std::max(o.r, L);
                if (tgt) v.w = L;
            }
            G.L = L;
            emit(kExecGroup, G.I, G.tmask, G.imask, L);
        }
        for (uint64_t x = f0; x < f1; x++)
            if (fv_g[x - f0] != ~0u) S.fv_level[x] = groups[fv_g[x - f0]].L;
        for (uint64_t x = v0; x < v1; x++)
            if (vf_g[x - v0] != ~0u) S.vf_level[x] = groups[vf_g[x - v0]].L;
        for (uint64_t x : wide_fv) {
            const uint32_t e = S.fv_edge[x];
            const uint32_t own = vm_own[e], I = oI[x];
            VarL &t = vl[ov[x]];
            uint32_t L = std::max(t.w, t.r);
            for (uint32_t s = fo[I]; s < fo[I + 1]; s++)
                if (s != own) L = std::max(L, sl[s].w);
            L += 1;
            for (uint32_t s = fo[I]; s < fo[I + 1]; s++)
                if (s != own) sl[s].r = std::max(sl[s].r, L);
            t.w = L;
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
