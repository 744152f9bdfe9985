/* utils/utils.h -- basic types and helpers of the drop-in C++ API.
 * Same names as the reference's include/utils/utils.h:21-22, 109-122,
 * 133-136. */
#pragma once

#include <sys/time.h>

#include <cstddef>
#include <iostream>
#include <istream>
#include <map>
#include <random>
#include <set>
#include <string>
#include <vector>

namespace lbp {

typedef double Real;
typedef size_t Size;

/* the global generator behind SEQFIX update orders: one per thread, seeded
 * 42 at thread start (include/utils/utils.h:133-136, causal_bp.cu:116) */
typedef std::mt19937 rnd_gen_type;
extern thread_local rnd_gen_type rnd_gen;

/* skips the '#' lines that start at the stream's position
 * (include/utils/utils.h:109-113) */
inline void read_annotation(std::istream &is) {
    std::string line;
    while (is.peek() == '#') std::getline(is, line);
}
inline void read_annotation(std::istream &is, std::string line) { read_annotation(is); }

/* wall clock in seconds */
inline Real toc() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return (Real)tv.tv_sec + (Real)tv.tv_usec / 1000000.0;
}

template <class T>
std::ostream &operator<<(std::ostream &os, const std::vector<T> &x) {
    os << "(";
    for (size_t i = 0; i < x.size(); i++) os << (i ? ", " : "") << x[i];
    return os << ")";
}

template <class T>
std::ostream &operator<<(std::ostream &os, const std::set<T> &x) {
    os << "{";
    bool first = true;
    for (const auto &v : x) {
        os << (first ? "" : ", ") << v;
        first = false;
    }
    return os << "}";
}

}  // namespace lbp
