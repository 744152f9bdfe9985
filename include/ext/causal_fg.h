/* ext/causal_fg.h -- causal factor graph (interface of the reference's
 * include/ext/causal_fg.h:15-74). */
#pragma once

#include <istream>
#include <map>
#include <vector>

#include <ext/causal_factor.h>
#include <lbp/bipgraph.h>
#include <utils/log.h>

namespace lbp {

class CausalFactorGraph {
    BipartiteGraph _G;
    std::vector<Var> _vars;
    std::vector<CausalFactor> _factors;

    void setFactor(size_t I, const CausalFactor &f) { _factors[I] = f; }
    void setFactors(const std::map<size_t, CausalFactor> &facs) {
        for (const auto &kv : facs) setFactor(kv.first, kv.second);
    }
    void constructGraph(size_t nrEdges);

public:
    CausalFactorGraph() = default;
    /* variables = sorted label set of all factors; var index = label rank */
    CausalFactorGraph(const std::vector<CausalFactor> &P);

    const Var &var(size_t i) const { return _vars[i]; }
    const CausalFactor &factor(size_t I) const { return _factors[I]; }
    const std::vector<Var> &vars() const { return _vars; }
    const std::vector<CausalFactor> &factors() const { return _factors; }
    const Neighbors &nbV(size_t i) const { return _G.nb1(i); }
    const Neighbor &nbV(size_t i, size_t _I) const { return _G.nb1(i)[_I]; }
    const Neighbors &nbF(size_t I) const { return _G.nb2(I); }
    size_t nrVars() const { return _vars.size(); }
    size_t nrFactors() const { return _factors.size(); }
    size_t nrEdges() const { return _G.nrEdges(); }
    size_t findVar(const Var &n) const;

    void ReadFromFile(const char *filename);
    /* clamps variable index i to x in every factor holding it; like the
     * reference, the adjacency is not rebuilt: construct a new graph from
     * factors() before inference */
    void clamp(size_t i, size_t x);

    friend std::istream &operator>>(std::istream &is, CausalFactorGraph &fg);
};

}  // namespace lbp
