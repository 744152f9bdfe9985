/*
 * dropin.cpp -- the C++ drop-in classes (the headers under include/ext and
 * include/lbp) on top of the C ABI.
 *
 * CausalFactor / CausalFactorGraph keep the reference's semantics
 * (csrc/ext/causal_factor.cpp, csrc/ext/causal_fg.cpp); CausalBP keeps the
 * interface and loop semantics of csrc/ext/causal_bp.cu but owns a
 * flbp_plan instead of thrust device vectors and kernel wrappers.  Engine
 * errors end the process through LOG_FATAL, as the reference's
 * CUDA_CHECK_ERROR does (include/utils/cuda_utils.h:16-21).
 */
#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <limits>
#include <memory>
#include <set>
#include <stdexcept>
#include <unordered_map>

#include <ext/causal_bp.h>
#include <ext/causal_fg.h>

#include "fastlbp.h"
#include "flbp_internal.h"

namespace lbp {

/* ------------------------------------------------------------------ */
/* CausalFactor                                                         */
/* ------------------------------------------------------------------ */

CausalFactor::CausalFactor(const Var &v, Real p)
    : head_clamped(false), head_mask(2, 1.0), _type(Singleton), _head(v), _body(), _p(p), _vs(v) {}

CausalFactor::CausalFactor(const Var &head, const VarSet &body, bool isAnd, Real p, Real q)
    : head_clamped(false), head_mask(2, 1.0), _type(isAnd ? DefiniteAnd : DefiniteOr), _head(head),
      _body(body), _p(p), _q(q), _vs(body) {
    _vs.insert(_head);
}

CausalFactor CausalFactor::gen_clamped(Var i, size_t x) const {
    CausalFactor f(*this);
    if (_type == Singleton) {
        if (i == _head) f._p = (Real)x;
        else std::cerr << "Clamp unrelated var " << i << " in factor " << *this << std::endl;
        return f;
    }
    if (i == _head) {
        f.head_clamped = true;
        f.head_mask = Prob(2, 0.0);
        f.head_mask.set(x, 1.0);
        if (head_clamped) {
            std::cerr << "Duplicate clamp head of factor " << *this << "with val " << x << std::endl;
            f.head_mask *= head_mask;
        }
        return f;
    }
    if (_body.contains(i)) {
        f._body.erase(i);
        f._vs.erase(i);
        const bool falsified_and = (x == 0 && _type == DefiniteAnd);
        const bool satisfied_or = (x > 0 && _type == DefiniteOr);
        if (falsified_and || satisfied_or) {
            f._p = _q;
            f._body = VarSet();
            f._vs = VarSet(_head);
        }
    }
    return f;
}

std::ostream &operator<<(std::ostream &os, const CausalFactor &f) {
    os << f.head() << std::endl << static_cast<char>(f._type);
    if (f._type == CausalFactor::Singleton) {
        os << std::endl << std::setw(os.precision() + 4) << f.prob() << std::endl;
    } else {
        os << std::setw(os.precision() + 4) << f.prob() << ";" << std::setw(os.precision() + 4)
           << f.prob_default() << std::endl
           << f.body().size() << std::endl;
        for (const Var &v : f.body()) os << v.label() << " ";
        os << std::endl;
    }
    return os << std::endl;
}

/* ------------------------------------------------------------------ */
/* CausalFactorGraph                                                    */
/* ------------------------------------------------------------------ */

CausalFactorGraph::CausalFactorGraph(const std::vector<CausalFactor> &P) : _G(), _factors(P) {
    /* the variables ordered by label, the first one met of each label kept
     * (the reference collects them in a std::set<Var>): a stable sort
     * keeps the order of equal labels, unique keeps the first */
    size_t nrEdges = 0;
    for (const CausalFactor &f : _factors) nrEdges += f.vars().size();
    _vars.clear();
    _vars.reserve(nrEdges);
    for (const CausalFactor &f : _factors) _vars.insert(_vars.end(), f.vars().begin(), f.vars().end());
    std::stable_sort(_vars.begin(), _vars.end());
    _vars.erase(std::unique(_vars.begin(), _vars.end()), _vars.end());
    _vars.shrink_to_fit();
    constructGraph(nrEdges);
}

void CausalFactorGraph::constructGraph(size_t nrEdges) {
    /* label -> variable index: a table when labels are dense enough, else
     * a hash map */
    const size_t V = _vars.size();
    const size_t max_label = V ? _vars.back().label() : 0;
    std::vector<size_t> table;
    std::unordered_map<size_t, size_t> index;
    const bool dense = V && max_label < 8 * V + 1024;
    if (dense) {
        table.assign(max_label + 1, 0);
        for (size_t i = 0; i < V; i++) table[_vars[i].label()] = i;
    } else {
        index.reserve(V * 2 + 1);
        for (size_t i = 0; i < V; i++) index[_vars[i].label()] = i;
    }
    std::vector<Edge> edges;
    edges.reserve(nrEdges);
    for (size_t I = 0; I < _factors.size(); I++)
        for (const Var &v : _factors[I].vars()) edges.emplace_back(dense ? table[v.label()] : index[v.label()], I);
    _G.construct(_vars.size(), _factors.size(), edges.begin(), edges.end(), false);
}

size_t CausalFactorGraph::findVar(const Var &n) const {
    size_t i = 0;
    while (i < _vars.size() && !(_vars[i] == n)) i++;
    if (i == _vars.size()) std::cerr << "var not found \n";
    return i;
}

void CausalFactorGraph::clamp(size_t i, size_t x) {
    std::map<size_t, CausalFactor> changed;
    for (const Neighbor &I : nbV(i)) changed.emplace(I.node, factor(I.node).gen_clamped(var(i), x));
    setFactors(changed);
}

void CausalFactorGraph::ReadFromFile(const char *filename) {
    std::ifstream in(filename);
    if (!in.is_open()) LOG_FATAL("Cannot read from file %s.", filename);
    in >> *this;
}

std::istream &operator>>(std::istream &is, CausalFactorGraph &fg) {
    std::vector<flbp::ParsedFactor> parsed;
    std::string err;
    if (!flbp::parse_causal_fg(is, parsed, err)) LOG_FATAL("%s", err.c_str());
    std::vector<CausalFactor> facs;
    facs.reserve(parsed.size());
    for (const auto &f : parsed) {
        Var head(f.head, 2);
        if (f.type == 0) {
            facs.emplace_back(head, f.p);
        } else {
            VarSet body;
            for (uint64_t b : f.body) body.insert(Var(b, 2));
            facs.emplace_back(head, body, f.type == 1, f.p, f.q);
        }
    }
    fg = CausalFactorGraph(facs);
    return is;
}

/* ------------------------------------------------------------------ */
/* CausalBP                                                             */
/* ------------------------------------------------------------------ */

namespace {

void check(int rc, const char *what) {
    if (rc != FLBP_OK) LOG_FATAL("%s: %s", what, flbp_last_error());
}

/* Per-variable ring of the last `capacity` finite beliefs, averaged oldest
 * to newest (the 4-argument run's low-pass result, causal_bp.cu:28-78). */

}  // namespace

CausalBP::CausalBP(const CausalFactorGraph &fg, const PropertySet &opts) : DAIAlg<CausalFactorGraph>(fg) {
    setProperties(opts);
    construct();
}

CausalBP::CausalBP(CausalFactorGraph &&fg, const PropertySet &opts) : DAIAlg<CausalFactorGraph>(std::move(fg)) {
    setProperties(opts);
    construct();
}

CausalBP::~CausalBP() {
    if (_plan) flbp_plan_destroy(_plan);
}

void CausalBP::setProperties(const PropertySet &opts) {
    props.tol = opts.hasKey("tol") ? opts.getStringAs<Real>("tol") : 1e-9;
    props.logdomain = opts.hasKey("logdomain") ? opts.getStringAs<bool>("logdomain") : false;
    props.maxiter = opts.hasKey("maxiter") ? opts.getStringAs<size_t>("maxiter") : 10000;
    props.maxtime = opts.hasKey("maxtime") ? opts.getStringAs<Real>("maxtime") : 10;
    props.updates = Properties::UpdateType::SEQFIX;
    if (opts.hasKey("updates")) {
        const std::string u = opts.getStringAs<std::string>("updates");
        /* anything else is SEQFIX (causal_bp.cu:165-177) */
        if (u == "PARALL") props.updates = Properties::UpdateType::PARALL;
        else if (u == "TOPO") props.updates = Properties::UpdateType::TOPO;
    }
    props.verbose = opts.hasKey("verbose") ? opts.getStringAs<size_t>("verbose") : 0;
    props.damping = opts.hasKey("damping") ? opts.getStringAs<Real>("damping") : 0.0;
    props.device = opts.hasKey("device") ? opts.getStringAs<size_t>("device") : 0;
}

void CausalBP::construct() {
    const size_t V = nrVars(), F = nrFactors();
    /* factor-major slots in nbF order, var-major edges in nbV order
     * (causal_bp.cu:336-369, 724-817) */
    std::vector<uint64_t> fac_off(F + 1, 0), head_slot(F), var_label(V), var_off(V + 1, 0);
    for (size_t I = 0; I < F; I++) fac_off[I + 1] = fac_off[I] + nbF(I).size();
    const uint64_t E = fac_off[F];
    std::vector<uint64_t> edge_var(E), var_edge(E);
    std::vector<int32_t> type(F);
    std::vector<uint8_t> clamped(F);
    std::vector<double> p(F), q(F), mask0(F), mask1(F);
    for (size_t I = 0; I < F; I++) {
        const CausalFactor &f = factor(I);
        const size_t head_label = f.head().label();
        for (size_t k = 0; k < nbF(I).size(); k++) {
            const size_t i = nbF(I)[k].node;
            edge_var[fac_off[I] + k] = i;
            if (var(i).label() == head_label) head_slot[I] = fac_off[I] + k;
        }
        type[I] = f.type();
        clamped[I] = f.head_clamped ? 1 : 0;
        p[I] = f.prob();
        q[I] = f.prob_default();
        mask0[I] = f.head_mask[0];
        mask1[I] = f.head_mask[1];
    }
    for (size_t i = 0; i < V; i++) {
        var_label[i] = var(i).label();
        var_off[i + 1] = var_off[i] + nbV(i).size();
        for (size_t k = 0; k < nbV(i).size(); k++) {
            const Neighbor &n = nbV(i)[k];
            var_edge[var_off[i] + k] = fac_off[n.node] + n.dual;
        }
    }
    if (var_off[V] != E) LOG_FATAL("graph adjacency is inconsistent (clamped without a rebuild?)");
    flbp_csr c;
    c.n_vars = V;
    c.n_factors = F;
    c.n_edges = E;
    c.var_label = var_label.data();
    c.fac_off = fac_off.data();
    c.edge_var = edge_var.data();
    c.var_off = var_off.data();
    c.var_edge = var_edge.data();
    c.fac_type = type.data();
    c.head_slot = head_slot.data();
    c.head_clamped = clamped.data();
    c.p = p.data();
    c.q = q.data();
    c.mask0 = mask0.data();
    c.mask1 = mask1.data();
    check(flbp_plan_create(&c, (int)props.device, &_plan), "flbp_plan_create");
    if (props.updates != Properties::UpdateType::PARALL) {
        /* buildFixedUpdateSeq / buildTopoUpdateSeq + putUpdateSeqToKernelFused
         * (causal_bp.cu:396-406); SEQFIX draws from this thread's rnd_gen */
        flbp_schedule *s = nullptr;
        const int u = props.updates == Properties::UpdateType::TOPO ? FLBP_UPDATES_TOPO : FLBP_UPDATES_SEQFIX;
        check(flbp_schedule_create(&c, u, &s), "flbp_schedule_create");
        const int rc = flbp_plan_set_schedule(_plan, s);
        flbp_schedule_destroy(s);
        check(rc, "flbp_plan_set_schedule");
    }
    _beliefsV.assign(V, 0.0);
}

void CausalBP::init() { check(flbp_plan_init(_plan, props.logdomain ? 1 : 0), "flbp_plan_init"); }

void CausalBP::transferBeliefToHost() {
    check(flbp_plan_beliefs(_plan, _beliefsV.data(), _beliefsV.size()), "flbp_plan_beliefs");
}

void CausalBP::run() {
    if (props.verbose >= 1) std::cerr << "Starting ..." << std::endl;
    flbp_run_opts o;
    o.tol = props.tol;
    o.maxiter = props.maxiter;
    o.maxtime = props.maxtime;
    o.damping = props.damping;
    o.chunk = 0;
    flbp_run_result r;
    check(flbp_plan_run(_plan, &o, &r), "flbp_plan_run");
    _iters = r.iters;
    _lastdiff = r.maxdiff;   /* the reference's run() keeps its maxDiff local (causal_bp.cu:864-890) */
    if (props.verbose >= 1) {
        if (r.maxdiff > props.tol) {
            if (props.verbose == 1) std::cerr << std::endl;
            std::cerr << "CausalBP::run:  WARNING: not converged after " << _iters << " passes ("
                      << r.seconds << " seconds)...final maxdiff:" << r.maxdiff << std::endl;
        } else {
            if (props.verbose >= 3) std::cerr << "CausalBP::run:  ";
            std::cerr << "converged in " << _iters << " passes (" << r.seconds << " seconds)." << std::endl;
        }
    }
    transferBeliefToHost();
}

Real CausalBP::run(Real tolerance, size_t maxIters, size_t histLength) {
    (void)tolerance;
    (void)histLength;
    flbp_run_result r;
    check(flbp_plan_run_fixed(_plan, maxIters, props.damping, &r), "flbp_plan_run_fixed");
    _iters = r.iters;
    transferBeliefToHost();
    _lowPassBeliefs = _beliefsV;
    return 1.0;
}

Real CausalBP::run(Real tolerance, size_t minIters, size_t maxIters, size_t histLength) {
    /* the history's own checks (HostBeliefHistory, causal_bp.cu:28-78);
     * the loop and the history run in flbp_plan_run_hist */
    const size_t V = nrVars();
    if (histLength == 0) throw std::invalid_argument("belief history length must be positive");
    if (V > std::numeric_limits<size_t>::max() / sizeof(Real) / histLength)
        throw std::length_error("belief history allocation size overflow");
    std::vector<Real> low(V);
    flbp_hist_result r;
    check(flbp_plan_run_hist(_plan, tolerance, minIters, maxIters, histLength, props.maxtime, props.damping, &r,
                             low.data()),
          "flbp_plan_run_hist");
    _iters += r.iters;
    if (r.maxdiff > _maxdiff) _maxdiff = r.maxdiff;
    transferBeliefToHost();
    _lowPassBeliefs = std::move(low);
    return r.yet_to_converge;
}

void CausalBP::calcBeliefV(size_t i, Prob &p) const { p = Prob(std::vector<Real>{1 - _beliefsV[i], _beliefsV[i]}); }

Factor CausalBP::beliefV(size_t i) const {
    Prob p;
    calcBeliefV(i, p);
    return Factor(var(i), p);
}

Factor CausalBP::belief(const VarSet &vs) const {
    if (vs.size() == 0) return Factor();
    if (vs.size() == 1) return beliefV(findVar(*vs.begin()));
    LOG_FATAL("unreachable");
}

void CausalBP::transferMessagesToHost() {
    const size_t E = nrEdges();
    std::vector<double> fv(2 * (E ? E : 1));
    check(flbp_plan_get_messages(_plan, fv.data(), nullptr), "flbp_plan_get_messages");
    _edges.assign(nrVars(), {});
    size_t idx = 0;
    for (size_t i = 0; i < nrVars(); i++) {
        _edges[i].reserve(nbV(i).size());
        for (size_t k = 0; k < nbV(i).size(); k++, idx++)
            _edges[i].push_back({Prob(std::vector<Real>{fv[2 * idx], fv[2 * idx + 1]})});
    }
}

}  // namespace lbp
