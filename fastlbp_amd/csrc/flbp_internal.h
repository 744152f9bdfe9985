/*
 * flbp_internal.h -- shared internals of libfastlbp.so (not installed).
 */
#ifndef FLBP_INTERNAL_H
#define FLBP_INTERNAL_H

#include <cstdint>
#include <istream>
#include <string>
#include <vector>

#include "fastlbp.h"

namespace flbp {

/* factors up to this degree run a sequential wave's updates as one group
 * item (seq_host.cpp, seq_kernels.h) */
constexpr uint32_t kSeqGroupDeg = 6;
/* the single-workgroup tail kernel of a sequential sweep runs levels of at
 * most this many items (seq.hip) */
constexpr uint32_t kSeqTailItemsMax = 1024;

/* structural checks of a CSR (plan_host.hip) */
int validate_csr(const flbp_csr &c);

/* Records msg as the calling thread's last error and returns code. */
int fail(int code, const std::string &msg);
void clear_error();

/* One factor as the .causal_fg reader produces it. */
struct ParsedFactor {
    int32_t type;                 /* 0 Singleton, 1 AND, 2 OR */
    uint64_t head;
    std::vector<uint64_t> body;   /* sorted, unique */
    double p, q;
};

/* The reader of csrc/ext/causal_fg.cpp:129-279: factor records, then one
 * implicit Singleton(0.5) per label that heads no factor, in ascending
 * label order.  Returns false with err set on malformed input. */
bool parse_causal_fg(std::istream &is, std::vector<ParsedFactor> &out, std::string &err);

/* A factor being clamped (csrc/ext/causal_factor.cpp:63-100). */
struct ClampFactor {
    int32_t type;
    uint64_t head;
    std::vector<uint64_t> body;   /* sorted, unique */
    double p, q, mask0, mask1;
    bool clamped;
};

/* CausalFactor::gen_clamped(label, x) in place. */
void gen_clamped(ClampFactor &f, uint64_t label, unsigned x);

}  // namespace flbp

#endif
