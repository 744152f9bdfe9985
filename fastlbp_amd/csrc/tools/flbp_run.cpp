/*
 * flbp_run -- solves a .causal_fg graph with CausalBP::run() on the GPU
 * through the drop-in classes and prints "iters N", "maxdiff X" and one
 * "label marginal" line per variable (%.17g).
 *
 * usage: flbp_run <file.causal_fg> [tol] [maxiter] [damping] [updates] [skip] [hist]
 *   updates: PARALL (default), SEQFIX or TOPO
 *   skip:    CausalBP instances constructed and dropped first (each SEQFIX
 *            construction advances the thread's rnd_gen)
 *   hist:    "min,max,len": a 4-argument run(tol, min, max, len) after init
 *            and before run()
 * stderr gets "maxDiff <x>" (the member the 4-argument run raises) before
 * and after run().
 */
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <string>

#include <ext/causal_bp.h>

using namespace lbp;

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <file.causal_fg> [tol] [maxiter] [damping] [updates] [skip]\n", argv[0]);
        return 2;
    }
    CausalFactorGraph fg;
    fg.ReadFromFile(argv[1]);
    PropertySet opts;
    opts.set("tol", Real(argc > 2 ? std::atof(argv[2]) : 1e-6));
    opts.set("maxiter", static_cast<size_t>(argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 1000));
    opts.set("damping", Real(argc > 4 ? std::atof(argv[4]) : 0.0));
    opts.set("maxtime", Real(1e30));
    opts.set("updates", std::string(argc > 5 ? argv[5] : "PARALL"));
    opts.set("logdomain", false);
    const size_t skip = argc > 6 ? std::strtoull(argv[6], nullptr, 10) : 0;
    for (size_t k = 0; k < skip; k++) CausalBP dropped(fg, opts);
    CausalBP bp(fg, opts);
    bp.init();
    if (argc > 7) {
        size_t mn = 0, mx = 0, len = 0;
        if (std::sscanf(argv[7], "%zu,%zu,%zu", &mn, &mx, &len) != 3) {
            std::fprintf(stderr, "hist must be min,max,len\n");
            return 2;
        }
        const Real frac = bp.run(argc > 2 ? std::atof(argv[2]) : 1e-6, mn, mx, len);
        std::fprintf(stderr, "hist iters %zu frac %.17g\n", bp.Iterations(), frac);
    }
    std::fprintf(stderr, "maxDiff %.17g\n", bp.maxDiff());
    bp.run();
    std::fprintf(stderr, "maxDiff %.17g\n", bp.maxDiff());
    std::printf("iters %zu\nmaxdiff %.17g\n", bp.Iterations(), bp.lastMaxDiff());
    for (size_t i = 0; i < fg.nrVars(); i++)
        std::printf("%zu %.17g\n", fg.var(i).label(), bp.beliefV(i)[1]);
    return 0;
}
