/* lbp/index.h -- index helpers for factor tables (interface of the
 * reference's include/lbp/index.h:14-70).
 *
 * Tables are stored with the variables of their VarSet in ascending label
 * order, the first variable changing fastest. */
#pragma once

#include <cstddef>
#include <vector>

#include <lbp/var.h>
#include <lbp/varset.h>

namespace lbp {

/* Walks the joint states of forVars (first fastest) and yields, for each,
 * the linear index of the matching state of indexVars (variables of
 * indexVars not in forVars held at 0); valid() turns false after the last
 * state. */
class IndexFor {
    long _index = -1;
    std::vector<size_t> _sum;      /* stride in indexVars of each forVars variable (0 if absent) */
    std::vector<size_t> _state;    /* current state of each forVars variable */
    std::vector<size_t> _ranges;   /* state counts of forVars */

public:
    IndexFor() = default;
    IndexFor(const VarSet &indexVars, const VarSet &forVars);
    operator size_t() const;
    IndexFor &operator++();
    void operator++(int);
    bool valid() const;
};

/* Maps a linear index over vars in the order given (first fastest; last
 * fastest with reverse) to the linear index of the same joint state in the
 * table order (ascending label). */
class Permute {
    std::vector<size_t> _ranges;   /* state counts in the given order */
    std::vector<size_t> _sigma;    /* for each table position, its position in the given order */

public:
    Permute() = default;
    Permute(const std::vector<Var> &vars, bool reverse = false);
    size_t convertLinearIndex(size_t li) const;
};

}  // namespace lbp
