/* lbp/prob.h -- probability vectors (interface of the reference's
 * include/lbp/prob.h:20-150: TProb<T>, Prob). */
#pragma once

#include <cstddef>
#include <iostream>
#include <vector>

#include <utils/utils.h>

namespace lbp {

template <typename T>
class TProb {
    std::vector<T> _p;

public:
    typedef typename std::vector<T>::const_iterator const_iterator;

    TProb() = default;
    /* n entries of 1/n */
    explicit TProb(size_t n) : _p(n, n ? T(1) / (T)n : T(0)) {}
    TProb(size_t n, T p) : _p(n, p) {}
    template <typename S>
    TProb(const std::vector<S> &v) : _p(v.begin(), v.end()) {}

    const_iterator begin() const { return _p.begin(); }
    const_iterator end() const { return _p.end(); }
    T get(size_t i) const { return _p.at(i); }
    void set(size_t i, T val) { _p.at(i) = val; }
    const std::vector<T> &toVec() const { return _p; }
    T operator[](size_t i) const { return _p[i]; }
    size_t size() const { return _p.size(); }

    T sum() const {
        T t = 0;
        for (const T &x : _p) t += x;
        return t;
    }
    /* divides by the sum (0/0 entries stay 0); prints a warning when the sum is 0 */
    T normalize() {
        T Z = sum();
        if (Z == (T)0) std::cerr << "Not normalizable!" << std::endl;
        else *this /= Z;
        return Z;
    }
    TProb &operator/=(T x) {
        if (x != 1)
            for (T &v : _p) v = (x == (T)0) ? (T)0 : v / x;
        return *this;
    }
    TProb &operator*=(const TProb &q) {
        if (size() != q.size())
            std::cerr << "The sizes of two vectors performing pointwise operation are not the same!"
                      << std::endl;
        for (size_t i = 0; i < _p.size() && i < q._p.size(); i++) _p[i] = _p[i] * q._p[i];
        return *this;
    }
};

template <typename T>
std::ostream &operator<<(std::ostream &os, const TProb<T> &p) {
    os << "(";
    for (size_t i = 0; i < p.size(); i++) os << (i ? ", " : "") << p.get(i);
    return os << ")";
}

typedef TProb<Real> Prob;

}  // namespace lbp
