/*
 * flbp_bingo -- the Bingo inference server protocol (the reference's
 * examples/bingo_wrapper.cu:25-195) on the MI355X engine, through the
 * drop-in CausalFactorGraph / CausalBP classes.
 *
 * usage: flbp_bingo <file.causal_fg | file.fg> [seed] [--auto | --manual id] [--digits N]
 *
 * Commands, whitespace-separated tokens on stdin:
 *   O <var index> true|false                 clamp; echoes "O <index> <value>"
 *   UC <var index>                           drop a clamp; echoes "UC <index>"
 *   BP <tol> <minIters> <maxIters> <histLen> a new CausalBP on the clamped
 *                                            graph (updates SEQFIX), init(),
 *                                            run(tol, minIters, maxIters, histLen);
 *                                            prints the yet-to-converge fraction
 *   Q <var index>                            prints newBelief(index) of the last BP
 *   NL                                       prints an empty line
 *   BYE                                      ends the session
 * Variable indices are label ranks (CausalFactorGraph's var order).  Logs
 * go to stderr, among them each BP's sweeps and its construct + init and
 * run times.
 *
 * As in the reference: seed reseeds the thread's rnd_gen before anything
 * else (each BP draws the next SEQFIX shuffle from it); clamps are applied
 * to a copy of the graph in ascending index order; properties maxiter
 * 10000000, maxtime 57600 s, tol 1e-6, updates SEQFIX; numbers print with
 * the stream's default 6 significant digits (--digits changes that, not in
 * the reference).  Not as in the reference: when a clamped variable
 * appears in a factor body, the clamped copy is rebuilt from its factors
 * before inference -- a body clamp erases the variable from the factor,
 * and the reference's un-rebuilt copy then no longer matches its adjacency
 * (its construct writes past its buffers).  With head clamps only the copy
 * is used as it is, as in the reference.
 * --auto takes device 0 (one process per GPU); --manual id takes device id.
 *
 * A file whose extension is not .causal_fg is read as a generic libDAI .fg
 * graph, as the reference does (bingo_wrapper.cu:156-165), which then runs
 * no inference: BP prints 1 (or, with a clamp set, ends the program with
 * "We don't support clamp for BP", bingo_wrapper.cu:25-41) and Q prints 0
 * (bingo_wrapper.cu:43-53).
 */
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <map>
#include <memory>
#include <string>

#include <ext/causal_bp.h>
#include <lbp/factorgraph.h>
#include <utils/utils.h>

using namespace lbp;

namespace {

CausalFactorGraph g_fg;
FactorGraph g_generic;
bool g_is_causal = true;
PropertySet g_opts;
std::unique_ptr<CausalBP> g_bp;
std::map<int, bool> g_clamps;

/* does clamping variable i change a factor's variable set (i in a body)? */
bool used_in_a_body(size_t i) {
    for (const Neighbor &I : g_fg.nbV(i))
        if (!(g_fg.factor(I.node).head() == g_fg.var(i))) return true;
    return false;
}

void init_bp() {
    CausalFactorGraph clamped(g_fg);
    bool body = false;
    for (const auto &c : g_clamps) {
        clamped.clamp((size_t)c.first, c.second ? 1 : 0);
        body = body || used_in_a_body((size_t)c.first);
    }
    g_bp.reset();
    if (body) {   /* the adjacency no longer matches the factors: rebuild */
        g_bp.reset(new CausalBP(CausalFactorGraph(clamped.factors()), g_opts));
    } else {      /* head clamps only: the clamped copy is consistent as it is */
        g_bp.reset(new CausalBP(std::move(clamped), g_opts));
    }
    g_bp->init();
}

bool valid_index(int i) { return i >= 0 && (size_t)i < (g_is_causal ? g_fg.nrVars() : g_generic.nrVars()); }

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " <file.causal_fg | file.fg> [seed] [--auto | --manual id] [--digits N]\n";
        return 1;
    }
    int digits = 6;
    size_t device = 0;
    int a = 2;
    if (argc > 2 && argv[2][0] != '-') {
        const int seed = std::stoi(argv[2]);
        std::clog << "flbp_bingo: setting random seed " << seed << std::endl;
        rnd_gen.seed(static_cast<unsigned int>(seed));
        a = 3;
    }
    for (; a < argc; a++) {
        const std::string arg = argv[a];
        if (arg == "--auto") {
            device = 0;
        } else if (arg == "--manual" && a + 1 < argc) {
            device = std::strtoull(argv[++a], nullptr, 10);
        } else if (arg == "--digits" && a + 1 < argc) {
            digits = std::atoi(argv[++a]);
        } else {
            std::cerr << "flbp_bingo: wrong arguments\n";
            return 1;
        }
    }
    const std::string file = argv[1];
    const size_t dot = file.find_last_of('.');
    g_is_causal = dot != std::string::npos && file.substr(dot + 1) == "causal_fg";
    if (g_is_causal) {
        g_fg.ReadFromFile(file.c_str());
        std::clog << "flbp_bingo: read " << g_fg.nrVars() << " variables, " << g_fg.nrFactors() << " factors"
                  << std::endl;
    } else {
        g_generic.ReadFromFile(file.c_str());
        std::clog << "flbp_bingo: read a generic graph of " << g_generic.nrVars() << " variables, "
                  << g_generic.nrFactors() << " factors (no inference runs on it)" << std::endl;
    }
    g_opts.set("maxiter", static_cast<size_t>(10000000));
    g_opts.set("maxtime", Real(57600));
    g_opts.set("tol", Real(1e-6));
    g_opts.set("updates", std::string("SEQFIX"));
    g_opts.set("logdomain", false);
    g_opts.set("device", device);
    std::cout.precision(digits);

    std::string cmd;
    while (std::cin >> cmd) {
        if (cmd == "Q") {
            int i;
            std::cin >> i;
            if (!g_is_causal) {
                std::cout << Real(0) << std::endl;
                continue;
            }
            if (!g_bp) LOG_FATAL("Q before any BP");
            if (!valid_index(i)) LOG_FATAL("Q: no variable %d", i);
            std::cout << g_bp->newBelief((size_t)i) << std::endl;
        } else if (cmd == "BP") {
            Real tol;
            size_t min_iters, max_iters, hist;
            std::cin >> tol >> min_iters >> max_iters >> hist;
            std::clog << "flbp_bingo: BP " << tol << " " << min_iters << " " << max_iters << " " << hist << std::endl;
            if (!g_is_causal) {
                if (!g_clamps.empty()) LOG_FATAL("We don't support clamp for BP");
                std::cout << Real(1) << std::endl;
                continue;
            }
            const auto t0 = std::chrono::steady_clock::now();
            init_bp();
            const auto t1 = std::chrono::steady_clock::now();
            const Real yet = g_bp->run(tol, min_iters, max_iters, hist);
            const auto t2 = std::chrono::steady_clock::now();
            std::cout << yet << std::endl;
            std::clog << "flbp_bingo: BP done: " << g_bp->Iterations() << " sweeps over " << g_bp->nrEdges()
                      << " edges; construct + init "
                      << std::chrono::duration<double, std::milli>(t1 - t0).count() << " ms, run "
                      << std::chrono::duration<double, std::milli>(t2 - t1).count() << " ms" << std::endl;
        } else if (cmd == "O") {
            int i;
            std::string v;
            std::cin >> i >> v;
            if (v != "true" && v != "false") LOG_FATAL("O: value must be true or false, got %s", v.c_str());
            if (!valid_index(i)) LOG_FATAL("O: no variable %d", i);
            g_clamps[i] = (v == "true");
            std::cout << "O " << i << " " << v << std::endl;
        } else if (cmd == "UC") {
            int i;
            std::cin >> i;
            if (g_clamps.erase(i) == 0) LOG_FATAL("UC: variable %d is not clamped", i);
            std::cout << "UC " << i << std::endl;
        } else if (cmd == "BYE") {
            break;
        } else if (cmd == "NL") {
            std::cout << std::endl;
        } else {
            LOG_FATAL("unknown command %s", cmd.c_str());
        }
    }
    std::clog << "flbp_bingo: bye" << std::endl;
    return 0;
}
