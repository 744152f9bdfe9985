/* ext/causal_bp.h -- loopy belief propagation on a CausalFactorGraph,
 * running on one MI355X through libfastlbp.so's C ABI (include/fastlbp.h).
 *
 * Interface of the reference's include/ext/causal_bp.h:19-193.  Properties
 * (csrc/ext/causal_bp.cu:144-184 semantics, exact stored types):
 *   tol (Real, 1e-9), logdomain (bool, false), maxiter (size_t, 10000),
 *   maxtime (Real, 10 s), updates (std::string, "SEQFIX"), verbose (size_t, 0)
 * plus, not in the reference:
 *   damping (Real, 0), device (size_t, 0).
 * updates: "PARALL" (flooding), "SEQFIX" (the default, and any unknown
 * string: factors shuffled by the thread's rnd_gen, declared in
 * utils/utils.h) or "TOPO" (breadth-first preorder).  The sequential
 * schedules run as dependency levels of the reference's waves with the
 * reference's per-wave results; damping applies to PARALL only. */
#pragma once

#include <cstdint>
#include <vector>

#include <ext/causal_fg.h>
#include <lbp/daialg.h>
#include <lbp/factor.h>
#include <lbp/properties.h>
#include <utils/utils.h>

struct flbp_plan;

namespace lbp {

class CausalBP : public DAIAlg<CausalFactorGraph> {
    struct Properties {
        enum class UpdateType { SEQFIX, PARALL, TOPO };
        size_t verbose;
        size_t maxiter;
        Real maxtime;
        Real tol;
        bool logdomain;
        UpdateType updates;
        Real damping;
        size_t device;
    } props;
    Real _maxdiff = 0;        /* the reference leaves it uninitialised; only the 4-argument run raises it */
    Real _lastdiff = INFINITY; /* last max belief change of run() (not in the reference) */
    size_t _iters = 0;
    struct EdgeProp {
        Prob message;
    };
    std::vector<std::vector<EdgeProp>> _edges;
    std::vector<Real> _beliefsV;
    std::vector<Real> _lowPassBeliefs;
    flbp_plan *_plan = nullptr;

    void setProperties(const PropertySet &opts);
    void construct();
    void transferBeliefToHost();
    void calcBeliefV(size_t i, Prob &p) const;

public:
    CausalBP(const CausalFactorGraph &fg, const PropertySet &opts);
    /* Not in the reference: builds on a graph the caller gives up, without
     * copying it (a server building one CausalBP per request). */
    CausalBP(CausalFactorGraph &&fg, const PropertySet &opts);
    CausalBP(const CausalBP &) = delete;
    CausalBP &operator=(const CausalBP &) = delete;
    ~CausalBP();

    Factor belief(const VarSet &vs) const;
    Factor beliefV(size_t i) const;
    void init();
    void run();
    Real run(Real tolerance, size_t maxIters, size_t histLength);
    Real run(Real tolerance, size_t minIters, size_t maxIters, size_t histLength);
    Real newBelief(size_t varIndex) const { return _lowPassBeliefs[varIndex]; }
    void transferMessagesToHost();
    const Prob &newMessage(size_t i, size_t _I) const { return _edges[i][_I].message; }

    /* not in the reference */
    size_t Iterations() const { return _iters; }
    /* the running maximum the 4-argument run keeps (causal_bp.cu:1101-1102);
     * run() leaves it alone, as in the reference */
    Real maxDiff() const { return _maxdiff; }
    /* the max belief change of the last sweep of the last run() */
    Real lastMaxDiff() const { return _lastdiff; }
    const std::vector<Real> &marginals() const { return _beliefsV; }
};

}  // namespace lbp
