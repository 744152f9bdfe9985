/* lbp/graph.h -- neighbour records (interface of the reference's
 * include/lbp/graph.h:13-38). */
#pragma once

#include <cstddef>
#include <utility>
#include <vector>

namespace lbp {

struct Neighbor {
    size_t iter;   /* position in this node's neighbour list */
    size_t node;   /* index of the neighbouring node */
    size_t dual;   /* position of this node in the neighbour's list */
    Neighbor(size_t iter, size_t node, size_t dual) : iter(iter), node(node), dual(dual) {}
    operator size_t() const { return node; }
};

typedef std::vector<Neighbor> Neighbors;
typedef std::pair<size_t, size_t> Edge;

}  // namespace lbp
