/* ext/causal_factor.h -- Singleton / DefiniteAnd / DefiniteOr factors
 * (interface of the reference's include/ext/causal_factor.h:14-76). */
#pragma once

#include <iostream>

#include <lbp/prob.h>
#include <lbp/varset.h>

namespace lbp {

class CausalFactor {
public:
    enum CausalType { Singleton = 'I', DefiniteAnd = '*', DefiniteOr = '+' };

    bool head_clamped;
    Prob head_mask;

private:
    CausalType _type;
    Var _head;
    VarSet _body;
    Real _p;
    Real _q = 0;
    VarSet _vs;

public:
    /* a Singleton prior P(v = 1) = p */
    CausalFactor(const Var &v, Real p);
    /* P(head = 1 | every body var 1) = p, else q (AND); for OR,
     * P(head = 0 | every body var 0) = p, else q */
    CausalFactor(const Var &head, const VarSet &body, bool isAnd, Real p, Real q);

    Real prob() const { return _p; }
    Real prob_default() const { return _q; }
    /* 0 Singleton, 1 DefiniteAnd, 2 DefiniteOr */
    int type() const { return _type == Singleton ? 0 : (_type == DefiniteAnd ? 1 : 2); }
    const Var &head() const { return _head; }
    const VarSet &body() const { return _body; }
    const VarSet &vars() const { return _vs; }

    /* the factor with variable i clamped to x (csrc/ext/causal_factor.cpp:63-100 semantics) */
    CausalFactor gen_clamped(Var i, size_t x) const;

    friend std::ostream &operator<<(std::ostream &, const CausalFactor &);
};

}  // namespace lbp
