/*
 * host_graph.cpp -- host-side graph handling of libfastlbp.so (no GPU).
 *
 *  - flbp_graph_build: factor list -> CSR in the reference's orders
 *    (CausalFactorGraph ctor, csrc/ext/causal_fg.cpp:28-68; BipartiteGraph
 *    ::addEdge, csrc/lbp/bipgraph.cpp:24-42; CausalBP::construct,
 *    csrc/ext/causal_bp.cu:310-407).  Built with flat arrays and a counting
 *    sort so that 100M-edge graphs take seconds.
 *  - flbp_read_causal_fg / flbp_write_causal_fg: the .causal_fg text format
 *    (reader semantics of csrc/ext/causal_fg.cpp:118-279).
 *  - flbp_clamp: CausalFactor::gen_clamped (csrc/ext/causal_factor.cpp:63-100)
 *    applied as examples/fl_wrapper.cu:26-45 applies it.
 *  - flbp_generate: seeded synthetic layered derivation graphs.
 *  - flbp_rank_alarms.
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <numeric>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "fastlbp.h"
#include "flbp_internal.h"

/* ------------------------------------------------------------------ */
/* owned containers                                                    */
/* ------------------------------------------------------------------ */

struct flbp_flist {
    std::vector<int32_t> type;
    std::vector<uint64_t> head, body_off{0}, body;
    std::vector<double> p, q, mask0, mask1;
    std::vector<uint8_t> head_clamped;

    void push(int32_t t, uint64_t h, const uint64_t *b, size_t nb, double pp, double qq,
              uint8_t clamped, double m0, double m1) {
        type.push_back(t);
        head.push_back(h);
        body.insert(body.end(), b, b + nb);
        body_off.push_back(body.size());
        p.push_back(pp);
        q.push_back(qq);
        head_clamped.push_back(clamped);
        mask0.push_back(m0);
        mask1.push_back(m1);
    }
    void view(flbp_factor_list *v) const {
        v->n_factors = type.size();
        v->type = type.data();
        v->head = head.data();
        v->body_off = body_off.data();
        v->body = body.data();
        v->p = p.data();
        v->q = q.data();
        v->head_clamped = head_clamped.data();
        v->mask0 = mask0.data();
        v->mask1 = mask1.data();
    }
};

struct flbp_graph {
    std::vector<uint64_t> var_label, fac_off, edge_var, var_off, var_edge, head_slot;
    std::vector<int32_t> fac_type;
    std::vector<uint8_t> head_clamped;
    std::vector<double> p, q, mask0, mask1;
};

struct flbp_generated {
    flbp_flist list;
    std::vector<uint64_t> alarms;
    std::vector<int32_t> rule;
};

namespace {

bool valid_list(const flbp_factor_list *fl) {
    if (!fl) return false;
    if (fl->n_factors == 0) return true;
    return fl->type && fl->head && fl->body_off && fl->p && fl->q && fl->head_clamped &&
           fl->mask0 && fl->mask1 && (fl->body || fl->body_off[fl->n_factors] == 0);
}

/* splitmix64 + xoshiro256** */
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; i++) {
            x += 0x9E3779B97F4A7C15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    /* integer in [0, n) */
    uint64_t below(uint64_t n) { return (uint64_t)(((__uint128_t)next() * n) >> 64); }
    uint64_t range(uint64_t lo, uint64_t hi) { return lo + below(hi - lo + 1); }
};

}  // namespace

extern "C" {

/* ------------------------------------------------------------------ */
/* CSR builder                                                          */
/* ------------------------------------------------------------------ */

int flbp_graph_build(const flbp_factor_list *fl, flbp_graph **out) {
    if (!valid_list(fl) || !out) return flbp::fail(FLBP_EINVAL, "flbp_graph_build: bad arguments");
    const uint64_t F = fl->n_factors;
    auto *g = new (std::nothrow) flbp_graph();
    if (!g) return flbp::fail(FLBP_ENOMEM, "flbp_graph_build: out of memory");
    try {
        /* factor-major slots: vars(I) = VarSet(body) + head, ascending label */
        g->fac_off.resize(F + 1);
        g->fac_off[0] = 0;
        std::vector<uint64_t> slot_label;
        slot_label.reserve(fl->body_off[F] + F);
        g->head_slot.resize(F);
        std::vector<uint64_t> tmp;
        for (uint64_t I = 0; I < F; I++) {
            const int32_t t = fl->type[I];
            if (t < 0 || t > 2) {
                delete g;
                return flbp::fail(FLBP_EINVAL, "flbp_graph_build: factor type must be 0, 1 or 2");
            }
            tmp.assign(fl->body + fl->body_off[I], fl->body + fl->body_off[I + 1]);
            tmp.push_back(fl->head[I]);
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            const uint64_t base = slot_label.size();
            for (size_t k = 0; k < tmp.size(); k++) {
                if (tmp[k] == fl->head[I]) g->head_slot[I] = base + k;
                slot_label.push_back(tmp[k]);
            }
            g->fac_off[I + 1] = slot_label.size();
        }
        const uint64_t E = slot_label.size();
        /* var index = rank of label */
        g->var_label = slot_label;
        std::sort(g->var_label.begin(), g->var_label.end());
        g->var_label.erase(std::unique(g->var_label.begin(), g->var_label.end()),
                           g->var_label.end());
        const uint64_t V = g->var_label.size();
        g->edge_var.resize(E);
        const uint64_t maxlab = V ? g->var_label.back() : 0;
        if (V && maxlab < 4 * (E + V) + 1024) {
            std::vector<uint64_t> lut(maxlab + 1, 0);
            for (uint64_t i = 0; i < V; i++) lut[g->var_label[i]] = i;
            for (uint64_t s = 0; s < E; s++) g->edge_var[s] = lut[slot_label[s]];
        } else {
            for (uint64_t s = 0; s < E; s++)
                g->edge_var[s] = (uint64_t)(std::lower_bound(g->var_label.begin(),
                                                             g->var_label.end(), slot_label[s]) -
                                            g->var_label.begin());
        }
        /* var-major: stable counting sort of factor-major slots by var, so
         * nbV(i) lists factors in ascending index */
        g->var_off.assign(V + 1, 0);
        for (uint64_t s = 0; s < E; s++) g->var_off[g->edge_var[s] + 1]++;
        for (uint64_t i = 0; i < V; i++) g->var_off[i + 1] += g->var_off[i];
        g->var_edge.resize(E);
        std::vector<uint64_t> cur(g->var_off.begin(), g->var_off.end() - 1);
        for (uint64_t s = 0; s < E; s++) g->var_edge[cur[g->edge_var[s]]++] = s;
        g->fac_type.assign(fl->type, fl->type + F);
        g->head_clamped.assign(fl->head_clamped, fl->head_clamped + F);
        g->p.assign(fl->p, fl->p + F);
        g->q.assign(fl->q, fl->q + F);
        g->mask0.assign(fl->mask0, fl->mask0 + F);
        g->mask1.assign(fl->mask1, fl->mask1 + F);
    } catch (const std::bad_alloc &) {
        delete g;
        return flbp::fail(FLBP_ENOMEM, "flbp_graph_build: out of memory");
    }
    *out = g;
    return FLBP_OK;
}

int flbp_graph_csr(const flbp_graph *g, flbp_csr *v) {
    if (!g || !v) return flbp::fail(FLBP_EINVAL, "flbp_graph_csr: null argument");
    v->n_vars = g->var_label.size();
    v->n_factors = g->fac_type.size();
    v->n_edges = g->edge_var.size();
    v->var_label = g->var_label.data();
    v->fac_off = g->fac_off.data();
    v->edge_var = g->edge_var.data();
    v->var_off = g->var_off.data();
    v->var_edge = g->var_edge.data();
    v->fac_type = g->fac_type.data();
    v->head_slot = g->head_slot.data();
    v->head_clamped = g->head_clamped.data();
    v->p = g->p.data();
    v->q = g->q.data();
    v->mask0 = g->mask0.data();
    v->mask1 = g->mask1.data();
    return FLBP_OK;
}

void flbp_graph_free(flbp_graph *g) { delete g; }

/* ------------------------------------------------------------------ */
/* .causal_fg reader / writer                                           */
/* ------------------------------------------------------------------ */

int flbp_read_causal_fg(const char *path, flbp_flist **out) {
    if (!path || !out) return flbp::fail(FLBP_EINVAL, "flbp_read_causal_fg: null argument");
    std::ifstream in(path);
    if (!in.is_open()) return flbp::fail(FLBP_EIO, std::string("Cannot read from file ") + path);
    auto *l = new flbp_flist();
    std::vector<flbp::ParsedFactor> facs;
    std::string err;
    if (!flbp::parse_causal_fg(in, facs, err)) {
        delete l;
        return flbp::fail(FLBP_EIO, err);
    }
    for (const auto &f : facs)
        l->push(f.type, f.head, f.body.data(), f.body.size(), f.p, f.q, 0, 1.0, 1.0);
    *out = l;
    return FLBP_OK;
}

int flbp_write_causal_fg(const flbp_factor_list *fl, const char *path) {
    if (!valid_list(fl) || !path) return flbp::fail(FLBP_EINVAL, "flbp_write_causal_fg: bad arguments");
    FILE *f = std::fopen(path, "w");
    if (!f) return flbp::fail(FLBP_EIO, std::string("cannot write ") + path);
    std::fprintf(f, "%llu\n", (unsigned long long)fl->n_factors);
    for (uint64_t I = 0; I < fl->n_factors; I++) {
        std::fprintf(f, "\n%llu\n", (unsigned long long)fl->head[I]);
        if (fl->type[I] == 0) {
            std::fprintf(f, "I\n%.17g\n", fl->p[I]);
        } else {
            std::vector<uint64_t> b(fl->body + fl->body_off[I], fl->body + fl->body_off[I + 1]);
            std::sort(b.begin(), b.end());
            b.erase(std::unique(b.begin(), b.end()), b.end());
            std::fprintf(f, "%c%.17g;%.17g\n%zu\n", fl->type[I] == 1 ? '*' : '+', fl->p[I],
                         fl->q[I], b.size());
            for (size_t k = 0; k < b.size(); k++)
                std::fprintf(f, k ? " %llu" : "%llu", (unsigned long long)b[k]);
            std::fprintf(f, "\n");
        }
    }
    bool ok = std::ferror(f) == 0;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? FLBP_OK : flbp::fail(FLBP_EIO, std::string("write failed: ") + path);
}

/* ------------------------------------------------------------------ */
/* clamp                                                                */
/* ------------------------------------------------------------------ */

int flbp_clamp(const flbp_factor_list *fl, uint64_t n_clamps, const uint64_t *labels,
               const uint8_t *values, flbp_flist **out) {
    if (!valid_list(fl) || !out || (n_clamps && (!labels || !values)))
        return flbp::fail(FLBP_EINVAL, "flbp_clamp: bad arguments");
    std::vector<flbp::ClampFactor> facs(fl->n_factors);
    for (uint64_t I = 0; I < fl->n_factors; I++) {
        flbp::ClampFactor &f = facs[I];
        f.type = fl->type[I];
        f.head = fl->head[I];
        f.body.assign(fl->body + fl->body_off[I], fl->body + fl->body_off[I + 1]);
        std::sort(f.body.begin(), f.body.end());
        f.body.erase(std::unique(f.body.begin(), f.body.end()), f.body.end());
        f.p = fl->p[I];
        f.q = fl->q[I];
        f.clamped = fl->head_clamped[I] != 0;
        f.mask0 = fl->mask0[I];
        f.mask1 = fl->mask1[I];
    }
    /* fl_wrapper keeps the clamps in a std::map<int, bool>: one value per
     * label (the last one given), applied in ascending label order */
    std::map<uint64_t, unsigned> clamps;
    for (uint64_t k = 0; k < n_clamps; k++) clamps[labels[k]] = values[k] ? 1u : 0u;
    /* which factors hold each clamped label (the unrebuilt graph's nbV) */
    std::map<uint64_t, std::vector<uint64_t>> holders;
    for (uint64_t I = 0; I < fl->n_factors; I++) {
        const flbp::ClampFactor &f = facs[I];
        if (clamps.count(f.head)) holders[f.head].push_back(I);
        for (uint64_t b : f.body)
            if (b != f.head && clamps.count(b)) holders[b].push_back(I);
    }
    for (const auto &c : clamps) {
        auto it = holders.find(c.first);
        if (it == holders.end()) continue;
        for (uint64_t I : it->second) flbp::gen_clamped(facs[I], c.first, c.second);
    }
    auto *l = new flbp_flist();
    for (const auto &f : facs)
        l->push(f.type, f.head, f.body.data(), f.body.size(), f.p, f.q, f.clamped ? 1 : 0,
                f.mask0, f.mask1);
    *out = l;
    return FLBP_OK;
}

int flbp_flist_view(const flbp_flist *l, flbp_factor_list *view) {
    if (!l || !view) return flbp::fail(FLBP_EINVAL, "flbp_flist_view: null argument");
    l->view(view);
    return FLBP_OK;
}

void flbp_flist_free(flbp_flist *l) { delete l; }

/* ------------------------------------------------------------------ */
/* generator                                                            */
/* ------------------------------------------------------------------ */

void flbp_gen_defaults(flbp_gen_params *p) {
    if (!p) return;
    p->n_vars = 2000;
    p->seed = 1;
    p->n_layers = 5;
    p->k_min = 1;
    p->k_max = 4;
    p->b_min = 1;
    p->b_max = 3;
    p->n_rules = 0;
    p->leaf_frac = 0.1;
    p->p_and_lo = 0.90;
    p->p_and_hi = 0.99;
    p->p_leaf_lo = 0.5;
    p->p_leaf_hi = 0.99;
    p->back_frac = 0.10;
    p->n_alarms = 100;
    p->extra_frac = 0.5;
    p->extra_q = 0.5;
    p->alarm_free = 1;
}

int flbp_generate(const flbp_gen_params *prm, flbp_generated **out) {
    if (!prm || !out || prm->n_vars < 1 || prm->n_layers < 2 || prm->k_min < 1 ||
        prm->k_max < prm->k_min || prm->b_min < 1 || prm->b_max < prm->b_min ||
        !(prm->leaf_frac > 0 && prm->leaf_frac <= 1) || !(prm->extra_frac >= 0 && prm->extra_frac <= 64) ||
        !(prm->extra_q >= 0 && prm->extra_q <= 1))
        return flbp::fail(FLBP_EINVAL, "flbp_generate: bad parameters");
    if (prm->n_vars >= (1ull << 40)) return flbp::fail(FLBP_ERANGE, "flbp_generate: too many variables");
    Rng rng(prm->seed);
    const uint64_t V = prm->n_vars;
    uint64_t L = (uint64_t)std::llround(prm->leaf_frac * (double)V);
    L = std::max<uint64_t>(1, std::min(L, V));
    const uint32_t NL = prm->n_layers;
    struct Head {
        uint64_t label;
        uint32_t layer;
        uint64_t first_clause;
        uint32_t k;
    };
    std::vector<Head> heads;
    std::vector<uint64_t> extra_leaves;
    uint64_t next = L;
    const uint64_t budget_total = V - L;
    /* tuples (leaves + heads) in layer order; layer_start[l] = first index of
     * layer l in `tuples` */
    std::vector<uint64_t> tuples;
    tuples.reserve(L + budget_total / 2 + 1);
    for (uint64_t i = 0; i < L; i++) tuples.push_back(i);
    std::vector<uint64_t> layer_start(NL + 1, 0);
    layer_start[1] = L;
    uint64_t carry = 0;
    for (uint32_t l = 1; l < NL; l++) {
        uint64_t budget = budget_total / (NL - 1) + (l <= budget_total % (NL - 1) ? 1 : 0) + carry;
        carry = 0;
        while (budget >= 2) {
            uint64_t k = rng.range(prm->k_min, prm->k_max);
            k = std::min<uint64_t>(k, budget - 1);
            heads.push_back({next, l, next + 1, (uint32_t)k});
            tuples.push_back(next);
            next += 1 + k;
            budget -= 1 + k;
        }
        if (budget == 1) {
            if (l + 1 < NL) carry = 1;
            else extra_leaves.push_back(next++);
        }
        layer_start[l + 1] = tuples.size();
    }
    auto *g = new flbp_generated();
    flbp_flist &fl = g->list;
    /* leaf priors */
    std::vector<uint64_t> leaves(L);
    std::iota(leaves.begin(), leaves.end(), 0);
    leaves.insert(leaves.end(), extra_leaves.begin(), extra_leaves.end());
    for (uint64_t lab : leaves) {
        double p = rng.uniform(prm->p_leaf_lo, prm->p_leaf_hi);
        fl.push(0, lab, nullptr, 0, p, 0.0, 0, 1.0, 1.0);
        g->rule.push_back(-1);
    }
    /* heads: OR over clauses, each clause an AND over earlier tuples */
    std::vector<uint8_t> in_body(prm->alarm_free ? V : 0, 0);
    std::vector<uint64_t> body, clauses;
    /* a rule body for a head of layer `layer`: b tuples of earlier layers,
     * plus with probability back_frac one of a later layer (a loop); with
     * alarm_free the last layer (the alarms' layer) feeds no rule */
    const uint64_t later_end = prm->alarm_free ? layer_start[NL - 1] : tuples.size();
    auto draw_body = [&](uint32_t layer) {
        const uint64_t avail = layer_start[layer];               /* tuples of earlier layers */
        const uint64_t later_lo = layer_start[layer + 1], later_hi = later_end;
        uint64_t b = rng.range(prm->b_min, prm->b_max);
        b = std::min<uint64_t>(b, avail);
        body.clear();
        while (body.size() < b) {
            uint64_t cand = tuples[rng.below(avail)];
            if (std::find(body.begin(), body.end(), cand) == body.end()) body.push_back(cand);
        }
        if (rng.uniform() < prm->back_frac && later_hi > later_lo) {
            uint64_t cand = tuples[later_lo + rng.below(later_hi - later_lo)];
            if (std::find(body.begin(), body.end(), cand) == body.end()) body.push_back(cand);
        }
        std::sort(body.begin(), body.end());
        if (prm->alarm_free)
            for (uint64_t x : body) in_body[x] = 1;
    };
    for (const Head &h : heads) {
        clauses.clear();
        for (uint32_t c = 0; c < h.k; c++) clauses.push_back(h.first_clause + c);
        fl.push(2, h.label, clauses.data(), clauses.size(), 1.0, 0.0, 0, 1.0, 1.0);
        g->rule.push_back(-1);
        for (uint32_t c = 0; c < h.k; c++) {
            draw_body(h.layer);
            double p = rng.uniform(prm->p_and_lo, prm->p_and_hi);
            int32_t r = prm->n_rules ? (int32_t)rng.below(prm->n_rules) : -1;
            fl.push(1, h.first_clause + c, body.data(), body.size(), p, 0.0, 0, 1.0, 1.0);
            g->rule.push_back(r);
        }
    }
    /* further rule instances: AND factors on existing heads */
    const uint64_t n_extra = heads.empty() ? 0 : (uint64_t)std::llround(prm->extra_frac * (double)V);
    for (uint64_t e = 0; e < n_extra; e++) {
        const Head &h = heads[rng.below(heads.size())];
        draw_body(h.layer);
        double p = rng.uniform(prm->p_and_lo, prm->p_and_hi);
        int32_t r = prm->n_rules ? (int32_t)rng.below(prm->n_rules) : -1;
        fl.push(1, h.label, body.data(), body.size(), p, prm->extra_q, 0, 1.0, 1.0);
        g->rule.push_back(r);
    }
    /* alarms: distinct heads of the last layer (without body uses when
     * alarm_free is set) */
    std::vector<uint64_t> last;
    for (const Head &h : heads)
        if (h.layer == NL - 1 && !(prm->alarm_free && in_body[h.label])) last.push_back(h.label);
    const uint64_t na = std::min<uint64_t>(prm->n_alarms, last.size());
    for (uint64_t a = 0; a < na; a++) {      /* partial Fisher-Yates */
        uint64_t j = a + rng.below(last.size() - a);
        std::swap(last[a], last[j]);
        g->alarms.push_back(last[a]);
    }
    *out = g;
    return FLBP_OK;
}

int flbp_generated_view(const flbp_generated *g, flbp_factor_list *view) {
    if (!g || !view) return flbp::fail(FLBP_EINVAL, "flbp_generated_view: null argument");
    g->list.view(view);
    return FLBP_OK;
}

int flbp_generated_extra(const flbp_generated *g, uint64_t *n_alarms, const uint64_t **alarms,
                         const int32_t **rule_of_factor) {
    if (!g) return flbp::fail(FLBP_EINVAL, "flbp_generated_extra: null argument");
    if (n_alarms) *n_alarms = g->alarms.size();
    if (alarms) *alarms = g->alarms.data();
    if (rule_of_factor) *rule_of_factor = g->rule.data();
    return FLBP_OK;
}

void flbp_generated_free(flbp_generated *g) { delete g; }

/* ------------------------------------------------------------------ */
/* ranking                                                              */
/* ------------------------------------------------------------------ */

int flbp_rank_alarms(uint64_t n, const uint64_t *labels, const double *marg, uint64_t *order) {
    if (n && (!labels || !marg || !order)) return flbp::fail(FLBP_EINVAL, "flbp_rank_alarms: null argument");
    std::vector<uint64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
        const bool na = std::isnan(marg[a]), nb = std::isnan(marg[b]);
        if (na != nb) return nb;              /* NaN marginals rank last */
        if (!na && marg[a] != marg[b]) return marg[a] > marg[b];
        return labels[a] < labels[b];
    });
    std::copy(idx.begin(), idx.end(), order);
    return FLBP_OK;
}

}  // extern "C"

/* ------------------------------------------------------------------ */
/* Binary graph files (not in the reference)                            */
/* ------------------------------------------------------------------ */

namespace {

constexpr char kGraphMagic[8] = {'F', 'L', 'B', 'P', 'C', 'S', 'R', '1'};

/* Integrity check over the file's 64-bit words (the tail zero-padded):
 * h = rotl(h ^ w, 29) * K per word.  Detects truncation and corruption; it
 * is not a cryptographic hash. */
struct WordHash {
    uint64_t h = 0x6a09e667f3bcc909ull;
    uint64_t pend = 0;
    unsigned npend = 0;
    void word(uint64_t w) {
        h ^= w;
        h = (h << 29) | (h >> 35);
        h *= 0x9e3779b97f4a7c15ull;
    }
    void bytes(const void *p, size_t n) {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        size_t i = 0;
        while (npend && i < n) {
            pend |= (uint64_t)b[i++] << (8 * npend);
            if (++npend == 8) {
                word(pend);
                pend = 0;
                npend = 0;
            }
        }
        for (; i + 8 <= n; i += 8) {
            uint64_t w;
            std::memcpy(&w, b + i, 8);
            word(w);
        }
        for (; i < n; i++) {
            pend |= (uint64_t)b[i] << (8 * npend);
            npend++;
        }
    }
    uint64_t digest() {
        if (npend) word(pend);
        word(0x8000000000000000ull | npend);
        return h;
    }
};

struct Out {
    std::FILE *f;
    WordHash hash;
    bool ok = true;
    void put(const void *p, size_t n) {
        if (!n) return;
        hash.bytes(p, n);
        ok = ok && std::fwrite(p, 1, n, f) == n;
    }
    template <class T, class S>
    void narrow(const S *src, uint64_t n) {   /* src[n] as T, in blocks */
        std::vector<T> buf(std::min<uint64_t>(n, 1 << 20));
        for (uint64_t i = 0; i < n; i += buf.size()) {
            const uint64_t m = std::min<uint64_t>(buf.size(), n - i);
            for (uint64_t k = 0; k < m; k++) buf[k] = (T)src[i + k];
            put(buf.data(), m * sizeof(T));
        }
    }
};

struct In {
    std::FILE *f;
    WordHash hash;
    bool ok = true;
    void get(void *p, size_t n) {
        if (!n || !ok) return;
        ok = std::fread(p, 1, n, f) == n;
        if (ok) hash.bytes(p, n);
    }
    template <class T, class D>
    void widen(std::vector<D> &dst, uint64_t n) {   /* n values of T into dst */
        dst.resize(n);
        std::vector<T> buf(std::min<uint64_t>(n, 1 << 20));
        for (uint64_t i = 0; i < n && ok; i += buf.size()) {
            const uint64_t m = std::min<uint64_t>(buf.size(), n - i);
            get(buf.data(), m * sizeof(T));
            for (uint64_t k = 0; k < m; k++) dst[i + k] = (D)buf[k];
        }
    }
};

}  // namespace

int flbp_write_graph(const flbp_csr *c, const char *path) {
    if (!c || !path) return flbp::fail(FLBP_EINVAL, "flbp_write_graph: null argument");
    if (int rc = flbp::validate_csr(*c)) return rc;
    const uint64_t V = c->n_vars, F = c->n_factors, E = c->n_edges;
    if (V >= (1ull << 32) - 1 || F >= (1ull << 32) - 1 || E >= (1ull << 32) - 1)
        return flbp::fail(FLBP_ERANGE, "flbp_write_graph: V, F and E must each be below 2^32-1");
    std::FILE *f = std::fopen(path, "wb");
    if (!f) return flbp::fail(FLBP_EIO, std::string("flbp_write_graph: cannot open ") + path);
    Out o{f};
    const uint64_t hdr[3] = {V, F, E};
    o.put(kGraphMagic, 8);
    o.put(hdr, sizeof hdr);
    o.put(c->var_label, V * 8);
    o.narrow<uint32_t>(c->fac_off, F + 1);
    o.narrow<uint32_t>(c->edge_var, E);
    o.narrow<uint32_t>(c->var_off, V + 1);
    o.narrow<uint32_t>(c->var_edge, E);
    o.narrow<uint32_t>(c->head_slot, F);
    o.narrow<uint8_t>(c->fac_type, F);
    o.put(c->head_clamped, F);
    o.put(c->p, F * 8);
    o.put(c->q, F * 8);
    o.put(c->mask0, F * 8);
    o.put(c->mask1, F * 8);
    const uint64_t digest = o.hash.digest();
    o.ok = o.ok && std::fwrite(&digest, 1, 8, f) == 8;
    const bool closed = std::fclose(f) == 0;
    if (!o.ok || !closed) return flbp::fail(FLBP_EIO, std::string("flbp_write_graph: write to ") + path + " failed");
    return FLBP_OK;
}

int flbp_read_graph(const char *path, flbp_graph **out) {
    if (!path || !out) return flbp::fail(FLBP_EINVAL, "flbp_read_graph: null argument");
    std::FILE *f = std::fopen(path, "rb");
    if (!f) return flbp::fail(FLBP_EIO, std::string("flbp_read_graph: cannot open ") + path);
    In in{f};
    char magic[8];
    uint64_t hdr[3];
    in.get(magic, 8);
    in.get(hdr, sizeof hdr);
    if (!in.ok || std::memcmp(magic, kGraphMagic, 8) != 0) {
        std::fclose(f);
        return flbp::fail(FLBP_EIO, std::string("flbp_read_graph: ") + path + " is not a graph file");
    }
    const uint64_t V = hdr[0], F = hdr[1], E = hdr[2];
    if (V >= (1ull << 32) - 1 || F >= (1ull << 32) - 1 || E >= (1ull << 32) - 1) {
        std::fclose(f);
        return flbp::fail(FLBP_EIO, "flbp_read_graph: header sizes out of range");
    }
    /* the file's length follows from the header: a damaged size field is
     * refused here, before anything of that size is allocated */
    const uint64_t want_len = 32 + 8 * V + 4 * (F + 1) + 4 * E + 4 * (V + 1) + 4 * E + 4 * F + F + F + 32 * F + 8;
    const long here = std::ftell(f);
    const bool sized = here >= 0 && std::fseek(f, 0, SEEK_END) == 0 && std::ftell(f) >= 0 &&
                       (uint64_t)std::ftell(f) == want_len && std::fseek(f, here, SEEK_SET) == 0;
    if (!sized) {
        std::fclose(f);
        return flbp::fail(FLBP_EIO, std::string("flbp_read_graph: ") + path +
                                        " is truncated or corrupt (length does not match its header)");
    }
    auto *g = new (std::nothrow) flbp_graph();
    if (!g) {
        std::fclose(f);
        return flbp::fail(FLBP_ENOMEM, "flbp_read_graph: out of memory");
    }
    try {
        g->var_label.resize(V);
        in.get(g->var_label.data(), V * 8);
        in.widen<uint32_t>(g->fac_off, F + 1);
        in.widen<uint32_t>(g->edge_var, E);
        in.widen<uint32_t>(g->var_off, V + 1);
        in.widen<uint32_t>(g->var_edge, E);
        in.widen<uint32_t>(g->head_slot, F);
        in.widen<uint8_t>(g->fac_type, F);
        g->head_clamped.resize(F);
        in.get(g->head_clamped.data(), F);
        for (auto *v : {&g->p, &g->q, &g->mask0, &g->mask1}) {
            v->resize(F);
            in.get(v->data(), F * 8);
        }
    } catch (const std::bad_alloc &) {
        std::fclose(f);
        delete g;
        return flbp::fail(FLBP_ENOMEM, "flbp_read_graph: out of memory");
    }
    uint64_t digest = 0;
    const uint64_t want = in.hash.digest();
    const bool got = in.ok && std::fread(&digest, 1, 8, f) == 8;
    const bool extra = got && std::fgetc(f) != EOF;
    std::fclose(f);
    if (!got || digest != want || extra) {
        delete g;
        return flbp::fail(FLBP_EIO, std::string("flbp_read_graph: ") + path + " is truncated or corrupt");
    }
    flbp_csr view;
    flbp_graph_csr(g, &view);
    if (int rc = flbp::validate_csr(view)) {
        delete g;
        return rc;
    }
    *out = g;
    return FLBP_OK;
}
