/*
 * flbp_graphdump -- prints a .causal_fg graph as built by the drop-in C++
 * classes, after fl_wrapper-style clamps (examples/fl_wrapper.cu:26-45:
 * label -> index, CausalFactorGraph::clamp, rebuild from factors()).
 * Output format: line 1 "V F E"; V labels; per factor "type p q clamped
 * mask0 mask1 head n node:dual..."; per variable "n node:dual...".
 * Also checks PropertySet's exact-type lookup (properties.h:44-53).
 *
 * usage: flbp_graphdump <file.causal_fg> [observations]
 *        flbp_graphdump --properties
 */
#include <cstdio>
#include <fstream>
#include <map>
#include <string>

#include <ext/causal_fg.h>
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

int main(int argc, char **argv) {
    if (argc >= 2 && std::string(argv[1]) == "--properties") return properties_check();
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
