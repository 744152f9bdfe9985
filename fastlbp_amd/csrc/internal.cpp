/*
 * internal.cpp -- error state, the .causal_fg parser and the clamp rule,
 * shared by the C ABI (host_graph.cpp) and the C++ drop-in classes.
 */
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "fastlbp.h"
#include "flbp_internal.h"

namespace flbp {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

void clear_error() { g_last_error.clear(); }

const char *last_error() { return g_last_error.c_str(); }

namespace {

/* utils.h:109-113 read_annotation: skips following lines that begin with '#' */
void skip_annotation(std::istream &is) {
    std::string line;
    while (is.peek() == '#') std::getline(is, line);
}

bool next_line(std::istream &is, std::string &line) {
    std::getline(is, line);
    return !is.fail();
}

/* std::stoi semantics: leading blanks, sign, digits; throws on no digits */
bool to_int(const std::string &s, long long &v) {
    try {
        v = std::stoi(s);
    } catch (...) {
        return false;
    }
    return true;
}

/* `is >> long` on the rest of a line: blanks, an optional sign, decimal
 * digits; fails with no digits or out of range (strtol in the C locale
 * accepts exactly that grammar) */
bool next_long(const char *&cur, long &v) {
    char *end = nullptr;
    errno = 0;
    const long x = std::strtol(cur, &end, 10);
    if (end == cur || errno == ERANGE) return false;
    cur = end;
    v = x;
    return true;
}

bool to_real(const std::string &s, double &v) {
    try {
        v = static_cast<double>(std::stold(s));
    } catch (...) {
        return false;
    }
    return true;
}

}  // namespace

bool parse_causal_fg(std::istream &is, std::vector<ParsedFactor> &out, std::string &err) {
    out.clear();
    std::string line;
    long long n = 0;
    next_line(is, line);
    skip_annotation(is);
    if (!to_int(line, n) || n < 0) {
        err = "Cannot read number of causal factors";
        return false;
    }
    /* every head and every label met, sorted and deduplicated at the end
     * (the reference keeps two std::sets) */
    std::vector<uint64_t> heads, all_vars;
    heads.reserve((size_t)std::min<long long>(n, 1ll << 26));
    all_vars.reserve((size_t)std::min<long long>(3 * n, 1ll << 27));
    for (long long I = 0; I < n; I++) {
        if (!next_line(is, line) || !line.empty())
            std::fprintf(stderr, "[ERROR] (causal_fg reader) Expecting empty line\n");
        long long head_id;
        next_line(is, line);
        skip_annotation(is);
        if (!to_int(line, head_id)) {
            err = "bad head id in factor " + std::to_string(I);
            return false;
        }
        ParsedFactor f;
        f.head = (uint64_t)head_id;
        heads.push_back(f.head);
        all_vars.push_back(f.head);
        next_line(is, line);
        skip_annotation(is);
        if (line.empty()) {
            err = "missing type in factor " + std::to_string(I);
            return false;
        }
        const char type_ch = line.front();
        if (type_ch == 'I') {
            f.type = 0;
        } else if (type_ch == '*') {
            f.type = 1;
        } else if (type_ch == '+') {
            f.type = 2;
        } else {
            err = "Unreachable";   /* causal_fg.cpp:195-197 */
            return false;
        }
        if (f.type == 0) {
            next_line(is, line);
            skip_annotation(is);
            if (!to_real(line, f.p)) {
                err = "bad probability in factor " + std::to_string(I);
                return false;
            }
            f.q = 0;
        } else {
            f.p = 1;
            f.q = 0;
            if (line.length() > 1) {
                const size_t split_pos = line.find(';');
                if (split_pos > 1 && !to_real(line.substr(1, split_pos - 1), f.p)) {
                    err = "bad probability in factor " + std::to_string(I);
                    return false;
                }
                if (split_pos != std::string::npos && split_pos + 1 != line.length() &&
                    !to_real(line.substr(split_pos + 1), f.q)) {
                    err = "bad default probability in factor " + std::to_string(I);
                    return false;
                }
            }
            long long body_len;
            next_line(is, line);
            skip_annotation(is);
            if (!to_int(line, body_len) || body_len < 0) {
                err = "bad body length in factor " + std::to_string(I);
                return false;
            }
            next_line(is, line);
            skip_annotation(is);
            const char *cur = line.c_str();
            for (long long b = 0; b < body_len; b++) {
                long body_id;
                if (!next_long(cur, body_id)) {
                    err = "short body line in factor " + std::to_string(I);
                    return false;
                }
                f.body.push_back((uint64_t)body_id);
                all_vars.push_back((uint64_t)body_id);
            }
            std::sort(f.body.begin(), f.body.end());
            f.body.erase(std::unique(f.body.begin(), f.body.end()), f.body.end());
        }
        out.push_back(std::move(f));
    }
    std::sort(heads.begin(), heads.end());
    heads.erase(std::unique(heads.begin(), heads.end()), heads.end());
    std::sort(all_vars.begin(), all_vars.end());
    all_vars.erase(std::unique(all_vars.begin(), all_vars.end()), all_vars.end());
    /* labels that head no factor get Singleton(0.5), in ascending order */
    size_t h = 0;
    for (uint64_t lab : all_vars) {
        while (h < heads.size() && heads[h] < lab) h++;
        if (h == heads.size() || heads[h] != lab) out.push_back(ParsedFactor{0, lab, {}, 0.5, 0.0});
    }
    return true;
}

void gen_clamped(ClampFactor &f, uint64_t label, unsigned x) {
    if (f.type == 0) {
        if (label == f.head) f.p = (double)x;
        else std::fprintf(stderr, "Clamp unrelated var x%llu\n", (unsigned long long)label);
        return;
    }
    if (label == f.head) {
        double m0 = x == 0 ? 1.0 : 0.0, m1 = x == 1 ? 1.0 : 0.0;
        if (f.clamped) {                       /* duplicate head clamp multiplies */
            m0 *= f.mask0;
            m1 *= f.mask1;
        }
        f.clamped = true;
        f.mask0 = m0;
        f.mask1 = m1;
        return;
    }
    auto it = std::lower_bound(f.body.begin(), f.body.end(), label);
    if (it == f.body.end() || *it != label) return;
    f.body.erase(it);
    if ((x == 0 && f.type == 1) || (x > 0 && f.type == 2)) {
        f.p = f.q;                            /* falsified AND / satisfied OR */
        f.body.clear();
    }
}

}  // namespace flbp

extern "C" const char *flbp_last_error(void) { return flbp::last_error(); }
extern "C" const char *flbp_version(void) { return "fastlbp-mi355x 0.1"; }
