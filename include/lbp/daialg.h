/* lbp/daialg.h -- inference-algorithm mixin (interface of the reference's
 * include/lbp/daialg.h:14-33): DAIAlg<GRM> is-a GRM holding a copy. */
#pragma once

#include <utility>

namespace lbp {

class InfAlg {
public:
    InfAlg() = default;
};

template <class GRM>
class DAIAlg : public InfAlg, public GRM {
public:
    DAIAlg() = default;
    DAIAlg(const GRM &grm) : InfAlg(), GRM(grm) {}
    DAIAlg(GRM &&grm) : InfAlg(), GRM(std::move(grm)) {}   /* not in the reference: takes the graph over */
};

}  // namespace lbp
