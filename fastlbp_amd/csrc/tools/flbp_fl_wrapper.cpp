/*
 * flbp_fl_wrapper -- the clamp-and-solve CLI of examples/fl_wrapper.cu:26-136
 * on the MI355X engine: reads a .causal_fg graph and an observation file
 * ("label true|false" per line), clamps, rebuilds, runs exactly 500 PARALL
 * sweeps (CausalBP::run(1e-6, 500, 1)) and prints "label marginal" for every
 * label up to the largest one, %.17g; clamped labels that left the graph
 * print their clamp value, absent labels 0.5.  The solve time goes to stderr.
 *
 * A file whose extension is not .causal_fg is read as a generic libDAI .fg
 * graph, as the reference does (fl_wrapper.cu:106-115); the reference runs
 * no inference on it: an observation ends the program with "We don't
 * support clamp for BP" (fl_wrapper.cu:39-42), and without one the empty
 * causal graph's report is the single line "0 0.5" (fl_wrapper.cu:57-88).
 *
 * usage: flbp_fl_wrapper <file.causal_fg | file.fg> <observations> [sweeps]
 */
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include <ext/causal_bp.h>
#include <lbp/factorgraph.h>
#include <utils/log.h>

using namespace lbp;

int main(int argc, char **argv) {
    if (argc < 3) {
        std::cerr << "Incorrect number of arguments." << std::endl;
        return 1;
    }
    const std::string file = argv[1];
    const size_t dot = file.find_last_of('.');
    const bool is_causal = dot != std::string::npos && file.substr(dot + 1) == "causal_fg";
    CausalFactorGraph causal_fg;
    FactorGraph fg;
    if (is_causal) causal_fg.ReadFromFile(argv[1]);
    else fg.ReadFromFile(argv[1]);
    std::ifstream observeFile(argv[2]);
    if (!observeFile.is_open()) {
        std::cerr << "Cannot open observation file!\n";
        return 1;
    }
    std::map<int, bool> clamps;
    int id;
    std::string boolStr;
    while (observeFile >> id >> boolStr) clamps[id] = (boolStr == "true");
    const size_t sweeps = argc >= 4 ? std::strtoull(argv[3], nullptr, 10) : 500;

    PropertySet opts;
    opts.set("maxiter", static_cast<size_t>(10000000));
    opts.set("maxtime", Real(57600));
    opts.set("tol", Real(1e-6));
    opts.set("verb", static_cast<size_t>(1));
    opts.set("updates", std::string("PARALL"));
    opts.set("logdomain", false);

    for (const auto &c : clamps) {
        if (!is_causal) LOG_FATAL("We don't support clamp for BP");
        for (size_t i = 0; i < causal_fg.nrVars(); i++)
            if (causal_fg.var(i).label() == (size_t)c.first) causal_fg.clamp(i, c.second ? 1 : 0);
    }
    CausalFactorGraph new_causal_fg(causal_fg.factors());
    Real tic1 = toc();
    std::unique_ptr<CausalBP> bp;
    if (is_causal) {
        bp.reset(new CausalBP(new_causal_fg, opts));
        bp->init();
        bp->run(1e-6, sweeps, 1);
    }
    std::clog << toc() - tic1 << std::endl;

    size_t maxVarLabel = 0;
    for (size_t i = 0; i < new_causal_fg.nrVars(); i++)
        if (new_causal_fg.var(i).label() > maxVarLabel) maxVarLabel = new_causal_fg.var(i).label();
    std::vector<Real> marginals(maxVarLabel + 1, -1);
    for (size_t i = 0; i < new_causal_fg.nrVars(); i++)
        marginals[new_causal_fg.var(i).label()] = bp->newBelief(i);
    for (size_t i = 0; i < marginals.size(); i++) {
        if (marginals[i] >= 0 && marginals[i] <= 1) std::printf("%zu %.17g\n", i, marginals[i]);
        else if (clamps.count((int)i)) std::printf("%zu %g\n", i, clamps[(int)i] ? 1.0 : 0.0);
        else std::printf("%zu %g\n", i, 0.5);
    }
    return 0;
}
