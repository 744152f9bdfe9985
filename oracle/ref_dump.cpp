/*
 * ref_dump.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * A driver linked against the reference's own graph-model sources
 * (csrc/lbp/*.cpp, csrc/ext/causal_factor.cpp, csrc/ext/causal_fg.cpp),
 * compiled by oracle/Makefile into oracle/_ref/ref_dump.  It reads a
 * .causal_fg file with the reference's reader (causal_fg.cpp:118-279),
 * applies clamps the way examples/fl_wrapper.cu:26-45 does (label -> index,
 * CausalFactorGraph::clamp, rebuild from factors()), and prints the rebuilt
 * graph: variables, factors (type, p, q, clamp mask, nbF order with duals)
 * and nbV order with duals.  The tests compare this against the oracle's and
 * the engine's graph builders.
 *
 * usage: ref_dump <file.causal_fg> [<observation file: "label true|false" per line>]
 */
#include <cstdio>
#include <fstream>
#include <map>
#include <string>
#include <ext/causal_fg.h>

using namespace lbp;

int main(int argc, char **argv) {
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
    for (const auto &c : clamps) {
        for (size_t i = 0; i < fg.nrVars(); i++) {
            if (fg.var(i).label() == (size_t)c.first) fg.clamp(i, c.second ? 1 : 0);
        }
    }
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
