/* lbp/factor.h -- a factor over a VarSet (the part of the reference's
 * include/lbp/factor.h:16-90 that CausalBP::belief returns). */
#pragma once

#include <iostream>

#include <lbp/prob.h>
#include <lbp/varset.h>

namespace lbp {

template <typename T>
class TFactor {
    VarSet _vs;
    TProb<T> _p;

public:
    TFactor() = default;
    TFactor(const Var &v) : _vs(v), _p(v.states()) {}
    TFactor(const VarSet &vars) : _vs(vars), _p(vars.nrStates()) {}
    TFactor(const VarSet &vars, T p) : _vs(vars), _p(vars.nrStates(), p) {}
    TFactor(const VarSet &vars, const TProb<T> &p) : _vs(vars), _p(p) {}

    void set(size_t i, T val) { _p.set(i, val); }
    const std::vector<T> &toVec() const { return _p.toVec(); }
    TProb<T> &p() { return _p; }
    T operator[](size_t i) const { return _p[i]; }
    const VarSet &vars() const { return _vs; }
    size_t nrStates() const { return _p.size(); }
    T normalize() { return _p.normalize(); }
};

template <typename T>
std::ostream &operator<<(std::ostream &os, const TFactor<T> &f) {
    os << "(" << f.vars() << ", (";
    for (size_t i = 0; i < f.nrStates(); i++) os << (i ? ", " : "") << f[i];
    return os << "))";
}

typedef TFactor<Real> Factor;

}  // namespace lbp
