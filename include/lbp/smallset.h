/* lbp/smallset.h -- a sorted vector set (interface of the reference's
 * include/lbp/smallset.h:14-100). */
#pragma once

#include <algorithm>
#include <cstddef>
#include <iterator>
#include <vector>

namespace lbp {

template <typename T>
class SmallSet {
    std::vector<T> _elements;

public:
    typedef typename std::vector<T>::const_iterator const_iterator;
    typedef typename std::vector<T>::iterator iterator;

    SmallSet() = default;
    SmallSet(const T &t) : _elements(1, t) {}
    template <typename TIterator>
    SmallSet(TIterator begin, TIterator end, size_t sizeHint) {
        _elements.reserve(sizeHint);
        _elements.assign(begin, end);
        std::sort(_elements.begin(), _elements.end());
        _elements.erase(std::unique(_elements.begin(), _elements.end()), _elements.end());
    }

    SmallSet &insert(const T &t) {
        auto it = std::lower_bound(_elements.begin(), _elements.end(), t);
        if (it == _elements.end() || *it != t) _elements.insert(it, t);
        return *this;
    }
    SmallSet &operator/=(const T &t) {
        auto it = std::lower_bound(_elements.begin(), _elements.end(), t);
        if (it != _elements.end() && *it == t) _elements.erase(it);
        return *this;
    }
    SmallSet &erase(const T &t) { return *this /= t; }
    SmallSet operator&(const SmallSet &x) const {
        SmallSet r;
        std::set_intersection(_elements.begin(), _elements.end(), x._elements.begin(),
                              x._elements.end(), std::back_inserter(r._elements));
        return r;
    }
    /* true if x is a subset of *this */
    bool operator>>(const SmallSet &x) const {
        return std::includes(_elements.begin(), _elements.end(), x._elements.begin(), x._elements.end());
    }
    bool contains(const T &t) const { return std::binary_search(_elements.begin(), _elements.end(), t); }
    const_iterator begin() const { return _elements.begin(); }
    const_iterator end() const { return _elements.end(); }
    typename std::vector<T>::size_type size() const { return _elements.size(); }
};

}  // namespace lbp
