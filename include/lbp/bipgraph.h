/* lbp/bipgraph.h -- bipartite adjacency (interface of the reference's
 * include/lbp/bipgraph.h:14-69, addEdge semantics of csrc/lbp/bipgraph.cpp:24-42:
 * neighbours are appended in edge order). */
#pragma once

#include <iterator>
#include <type_traits>
#include <vector>

#include <lbp/graph.h>

namespace lbp {

class BipartiteGraph {
    std::vector<Neighbors> _nb1;   /* variables */
    std::vector<Neighbors> _nb2;   /* factors */

public:
    BipartiteGraph() = default;
    const Neighbors &nb1(size_t i1) const { return _nb1[i1]; }
    const Neighbors &nb2(size_t i2) const { return _nb2[i2]; }

    template <typename EdgeInputIterator>
    void construct(size_t nrNodes1, size_t nrNodes2, EdgeInputIterator begin, EdgeInputIterator end,
                   bool check = true) {
        _nb1.assign(nrNodes1, Neighbors());
        _nb2.assign(nrNodes2, Neighbors());
        using Cat = typename std::iterator_traits<EdgeInputIterator>::iterator_category;
        if constexpr (std::is_base_of<std::forward_iterator_tag, Cat>::value) {
            /* room for every node's edges first: one allocation per node
             * (a second pass over the edges, so forward iterators only) */
            std::vector<size_t> d1(nrNodes1, 0), d2(nrNodes2, 0);
            for (EdgeInputIterator e = begin; e != end; ++e) {
                d1[e->first]++;
                d2[e->second]++;
            }
            for (size_t i = 0; i < nrNodes1; i++) _nb1[i].reserve(d1[i]);
            for (size_t i = 0; i < nrNodes2; i++) _nb2[i].reserve(d2[i]);
        }
        for (EdgeInputIterator e = begin; e != end; ++e) addEdge(e->first, e->second, check);
    }

    BipartiteGraph &addEdge(size_t n1, size_t n2, bool check = true) {
        if (check)
            for (const Neighbor &nb : _nb1[n1])
                if (nb.node == n2) return *this;
        const size_t it1 = _nb1[n1].size(), it2 = _nb2[n2].size();
        _nb1[n1].emplace_back(it1, n2, it2);
        _nb2[n2].emplace_back(it2, n1, it1);
        return *this;
    }

    size_t nrNodes1() const { return _nb1.size(); }
    size_t nrNodes2() const { return _nb2.size(); }
    size_t nrEdges() const {
        size_t s = 0;
        for (const auto &n : _nb1) s += n.size();
        return s;
    }
};

}  // namespace lbp
