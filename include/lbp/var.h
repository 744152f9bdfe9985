/* lbp/var.h -- a discrete random variable, ordered by label
 * (interface of the reference's include/lbp/var.h:14-52). */
#pragma once

#include <cstddef>
#include <iostream>

namespace lbp {

class Var {
    size_t _label;
    size_t _nstates;

public:
    Var(size_t label = 0, size_t nstates = 0) : _label(label), _nstates(nstates) {}
    size_t label() const { return _label; }
    size_t states() const { return _nstates; }
    bool operator<(const Var &n) const { return _label < n._label; }
    bool operator<=(const Var &n) const { return _label <= n._label; }
    bool operator!=(const Var &n) const { return _label != n._label; }
    bool operator==(const Var &n) const { return _label == n._label; }
    friend std::ostream &operator<<(std::ostream &os, const Var &n) { return os << "x" << n._label; }
};

}  // namespace lbp
