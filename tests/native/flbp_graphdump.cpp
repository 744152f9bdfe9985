/*
 * flbp_graphdump -- prints a .causal_fg graph as built by the drop-in C++
 * classes, after fl_wrapper-style clamps (examples/fl_wrapper.cu:26-45:
 * label -> index, CausalFactorGraph::clamp, rebuild from factors()).
 * Output format: line 1 "V F E"; V labels; per factor "type p q clamped
 * mask0 mask1 head n node:dual..."; per variable "n node:dual...".
 * Also checks PropertySet's exact-type lookup (properties.h:44-53).
 *
 * usage: flbp_graphdump <file.causal_fg> [observations]
 *        flbp_graphdump <file.fg>      (generic factor graph: line 1 "V F";
 *                                       per variable "label states"; per
 *                                       factor "n labels... | table...")
 *        flbp_graphdump --properties
 *        flbp_graphdump --index        (IndexFor / Permute on fixed inputs)
 *        flbp_graphdump --graph <file> (binary graph file: flbp_read_graph;
 *                                       prints "V F E" or the error)
 *        flbp_graphdump --schedules <file.causal_fg>
 *                                      (C-ABI reader, graph build and the
 *                                       SEQFIX and TOPO schedules; prints
 *                                       sizes or the error)
 *
 * A test tool, not part of the library: the CPU tests compare its dumps
 * with the reference's graph model, and the sanitizer target of the
 * Makefile beside it replays damaged input files through it.
 */
#include <cstdio>
#include <fstream>
#include <map>
#include <string>

#include <fastlbp.h>
#include <ext/causal_fg.h>
#include <lbp/factorgraph.h>
#include <lbp/index.h>
#include <lbp/properties.h>

using namespace lbp;

static int properties_check() {
    PropertySet opts;
    opts.set("maxiter", (size_t)1000);
    opts.set("maxiter_int", 1000);
    opts.set("tol", Real(1e-6));
    opts.set("updates", std::string("PARALL"));
    size_t a = opts.getStringAs<size_t>("maxiter");
    size_t b = opts.getStringAs<size_t>("maxiter_int");   /* stored as int: T() */
    Real t = opts.getStringAs<Real>("tol");
    std::string u = opts.getStringAs<std::string>("updates");
    std::printf("%zu %zu %.17g %s %d %d\n", a, b, t, u.c_str(), (int)opts.hasKey("tol"),
                (int)opts.hasKey("verbose"));
    return 0;
}

/* IndexFor of {x1(2), x3(3)} over {x1(2), x2(2), x3(3)}; Permute of
 * (x7(2), x3(3), x5(2)) forward and reversed, for linear indices 0..11 */
static int index_check() {
    std::vector<Var> for_vars{Var(1, 2), Var(2, 2), Var(3, 3)}, idx_vars{Var(1, 2), Var(3, 3)};
    IndexFor it(VarSet(idx_vars.begin(), idx_vars.end(), 2), VarSet(for_vars.begin(), for_vars.end(), 3));
    for (; it.valid(); ++it) std::printf("%zu ", (size_t)it);
    std::printf("\n");
    const std::vector<Var> listed{Var(7, 2), Var(3, 3), Var(5, 2)};
    for (int rev = 0; rev < 2; rev++) {
        const Permute p(listed, rev != 0);
        for (size_t li = 0; li < 12; li++) std::printf("%zu ", p.convertLinearIndex(li));
        std::printf("\n");
    }
    return 0;
}

static int dump_fg(const char *path) {
    FactorGraph fg;
    fg.ReadFromFile(path);
    std::printf("%zu %zu\n", fg.nrVars(), fg.nrFactors());
    for (size_t i = 0; i < fg.nrVars(); i++) std::printf("%zu %zu\n", fg.var(i).label(), fg.var(i).states());
    for (size_t I = 0; I < fg.nrFactors(); I++) {
        const Factor &f = fg.factor(I);
        std::printf("%zu", f.vars().size());
        for (const Var &v : f.vars()) std::printf(" %zu", v.label());
        std::printf(" |");
        for (size_t k = 0; k < f.nrStates(); k++) std::printf(" %.17g", f[k]);
        std::printf("\n");
    }
    return 0;
}

/* the binary graph reader: prints the sizes, or the status and message */
static int read_graph(const char *path) {
    flbp_graph *g = nullptr;
    const int st = flbp_read_graph(path, &g);
    if (st != FLBP_OK) {
        std::printf("error %d %s\n", st, flbp_last_error());
        return 1;
    }
    flbp_csr c;
    flbp_graph_csr(g, &c);
    std::printf("%llu %llu %llu\n", (unsigned long long)c.n_vars, (unsigned long long)c.n_factors,
                (unsigned long long)c.n_edges);
    flbp_graph_free(g);
    return 0;
}

/* the C-ABI .causal_fg reader, the graph build and both sequential schedules */
static int schedules(const char *path) {
    flbp_flist *l = nullptr;
    int st = flbp_read_causal_fg(path, &l);
    if (st != FLBP_OK) {
        std::printf("error %d %s\n", st, flbp_last_error());
        return 1;
    }
    flbp_factor_list view;
    flbp_flist_view(l, &view);
    flbp_graph *g = nullptr;
    st = flbp_graph_build(&view, &g);
    flbp_flist_free(l);
    if (st != FLBP_OK) {
        std::printf("error %d %s\n", st, flbp_last_error());
        return 1;
    }
    flbp_csr c;
    flbp_graph_csr(g, &c);
    std::printf("%llu %llu %llu", (unsigned long long)c.n_vars, (unsigned long long)c.n_factors,
                (unsigned long long)c.n_edges);
    for (int upd : {FLBP_UPDATES_SEQFIX, FLBP_UPDATES_TOPO}) {
        flbp_schedule *s = nullptr;
        st = flbp_schedule_create(&c, upd, &s);
        if (st != FLBP_OK) {
            std::printf(" error %d", st);
            continue;
        }
        flbp_schedule_info info;
        flbp_schedule_get_info(s, &info);
        std::printf(" %llu/%llu", (unsigned long long)info.n_order, (unsigned long long)info.n_levels);
        flbp_schedule_destroy(s);
    }
    std::printf("\n");
    flbp_graph_free(g);
    return 0;
}

int main(int argc, char **argv) {
    if (argc >= 2 && std::string(argv[1]) == "--properties") return properties_check();
    if (argc >= 2 && std::string(argv[1]) == "--index") return index_check();
    if (argc >= 3 && std::string(argv[1]) == "--graph") return read_graph(argv[2]);
    if (argc >= 3 && std::string(argv[1]) == "--schedules") return schedules(argv[2]);
    if (argc >= 2) {
        const std::string f = argv[1];
        if (f.size() > 3 && f.compare(f.size() - 3, 3, ".fg") == 0) return dump_fg(argv[1]);
    }
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <file.causal_fg> [observations]\n", argv[0]);
        return 2;
    }
    CausalFactorGraph fg;
    fg.ReadFromFile(argv[1]);
    std::map<long, bool> clamps;
    if (argc >= 3) {
        std::ifstream obs(argv[2]);
        long id;
        std::string val;
        while (obs >> id >> val) clamps[id] = (val == "true");
    }
    for (const auto &c : clamps)
        for (size_t i = 0; i < fg.nrVars(); i++)
            if (fg.var(i).label() == (size_t)c.first) fg.clamp(i, c.second ? 1 : 0);
    CausalFactorGraph g(fg.factors());
    std::printf("%zu %zu %zu\n", g.nrVars(), g.nrFactors(), g.nrEdges());
    for (size_t i = 0; i < g.nrVars(); i++) std::printf("%zu\n", g.var(i).label());
    for (size_t I = 0; I < g.nrFactors(); I++) {
        const CausalFactor &f = g.factor(I);
        std::printf("%d %.17g %.17g %d %.17g %.17g %zu %zu", f.type(), f.prob(), f.prob_default(),
                    f.head_clamped ? 1 : 0, f.head_mask[0], f.head_mask[1], f.head().label(),
                    g.nbF(I).size());
        for (const auto &n : g.nbF(I)) std::printf(" %zu:%zu", n.node, n.dual);
        std::printf("\n");
    }
    for (size_t i = 0; i < g.nrVars(); i++) {
        std::printf("%zu", g.nbV(i).size());
        for (const auto &n : g.nbV(i)) std::printf(" %zu:%zu", n.node, n.dual);
        std::printf("\n");
    }
    return 0;
}
