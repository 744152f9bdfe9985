/*
 * dropin_fg.cpp -- the generic factor graph of the drop-in API: index
 * helpers for factor tables and libDAI's .fg text format (the reference's
 * csrc/lbp/index.cpp, csrc/lbp/factorgraph.cpp:19-212).  A data structure
 * only, as in the reference: no inference runs on these graphs.
 */
#include <algorithm>
#include <fstream>
#include <map>
#include <set>
#include <string>
#include <unordered_map>

#include <lbp/factorgraph.h>
#include <lbp/index.h>
#include <utils/log.h>
#include <utils/utils.h>

namespace lbp {

/* ---------------------------------------------------------------- IndexFor */

IndexFor::IndexFor(const VarSet &indexVars, const VarSet &forVars) : _state(forVars.size(), 0) {
    /* stride of each variable of indexVars in its table */
    std::map<size_t, size_t> stride;
    size_t s = 1;
    for (const Var &v : indexVars) {
        stride[v.label()] = s;
        s *= v.states();
    }
    _ranges.reserve(forVars.size());
    _sum.reserve(forVars.size());
    for (const Var &v : forVars) {
        _ranges.push_back(v.states());
        auto it = stride.find(v.label());
        _sum.push_back(it == stride.end() ? 0 : it->second);
    }
    _index = 0;
}

IndexFor::operator size_t() const { return (size_t)_index; }

IndexFor &IndexFor::operator++() {
    if (_index < 0) return *this;
    size_t k = 0;
    for (; k < _state.size(); k++) {
        _index += (long)_sum[k];
        if (++_state[k] < _ranges[k]) break;
        _index -= (long)(_sum[k] * _ranges[k]);
        _state[k] = 0;
    }
    if (k == _state.size()) _index = -1;
    return *this;
}

void IndexFor::operator++(int) { ++*this; }

bool IndexFor::valid() const { return _index >= 0; }

/* ----------------------------------------------------------------- Permute */

Permute::Permute(const std::vector<Var> &vars, bool reverse) {
    const size_t N = vars.size();
    _ranges.resize(N);
    for (size_t k = 0; k < N; k++) _ranges[k] = vars[reverse ? N - 1 - k : k].states();
    const VarSet sorted(vars.begin(), vars.end(), N);
    for (const Var &v : sorted) {
        const size_t at = (size_t)(std::find(vars.begin(), vars.end(), v) - vars.begin());
        _sigma.push_back(reverse ? N - 1 - at : at);
    }
}

size_t Permute::convertLinearIndex(size_t li) const {
    /* digits of li in the given order's mixed radix, first fastest */
    std::vector<size_t> digit(_ranges.size());
    for (size_t k = 0; k < _ranges.size(); k++) {
        digit[k] = li % _ranges[k];
        li /= _ranges[k];
    }
    /* reassembled in table order */
    size_t out = 0, radix = 1;
    for (size_t k = 0; k < _sigma.size(); k++) {
        out += digit[_sigma[k]] * radix;
        radix *= _ranges[_sigma[k]];
    }
    return out;
}

/* ------------------------------------------------------------- FactorGraph */

FactorGraph::FactorGraph(const std::vector<Factor> &P) : _G(), _factors(P) {
    std::set<Var> all;
    size_t nrEdges = 0;
    for (const Factor &f : _factors) {
        all.insert(f.vars().begin(), f.vars().end());
        nrEdges += f.vars().size();
    }
    _vars.assign(all.begin(), all.end());
    constructGraph(nrEdges);
}

void FactorGraph::constructGraph(size_t nrEdges) {
    std::unordered_map<size_t, size_t> index;
    for (size_t i = 0; i < _vars.size(); i++) index[_vars[i].label()] = i;
    std::vector<Edge> edges;
    edges.reserve(nrEdges);
    for (size_t I = 0; I < _factors.size(); I++)
        for (const Var &v : _factors[I].vars()) edges.emplace_back(index[v.label()], I);
    _G.construct(_vars.size(), _factors.size(), edges.begin(), edges.end());
}

size_t FactorGraph::findVar(const Var &n) const {
    const size_t i = (size_t)(std::find(_vars.begin(), _vars.end(), n) - _vars.begin());
    if (i == nrVars()) std::cerr << "var not found \n";
    return i;
}

void FactorGraph::ReadFromFile(const char *filename) {
    std::ifstream in(filename);
    if (!in.is_open()) LOG_FATAL("Cannot read from file %s.", filename);
    in >> *this;
}

/* libDAI's .fg format (factorgraph.cpp:106-212): the factor count on a line
 * of its own; per factor the member count, the member labels, their state
 * counts, the count of listed entries and (linear index, value) pairs, the
 * linear index running over the members as listed, first fastest (an
 * index past the table keeps only the digits that fit, as the reference's
 * Permute does).  '#' lines are skipped only where one starts right at the
 * read position (read_annotation), so a comment after a blank line is a
 * parse error.  Unlisted entries are 0.  Where the reference reads on from
 * a failed extraction, this reader stops with LOG_FATAL. */
constexpr size_t kMaxTable = size_t(1) << 28;

std::istream &operator>>(std::istream &is, FactorGraph &fg) {
    std::string line;
    size_t n_factors = 0;
    read_annotation(is);
    is >> n_factors;
    if (is.fail()) LOG_FATAL(".fg: cannot read the number of factors");
    std::getline(is, line);
    if (is.fail() || !line.empty()) LOG_FATAL(".fg: expected only the number of factors on the first line");
    std::map<long, size_t> dims_of;
    std::vector<Factor> facs;
    /* counts come from the file: storage grows as entries are read, so a
     * damaged count ends in a failed read, not in a huge allocation */
    for (size_t I = 0; I < n_factors; I++) {
        size_t n_members = 0;
        read_annotation(is);
        is >> n_members;
        if (is.fail()) LOG_FATAL(".fg: factor %zu: bad member count", I);
        std::vector<long> labels;
        for (size_t m = 0; m < n_members; m++) {
            long l = 0;
            read_annotation(is);
            is >> l;
            if (is.fail()) LOG_FATAL(".fg: factor %zu: bad member list", I);
            labels.push_back(l);
        }
        std::vector<size_t> dims;
        size_t table = 1;
        for (size_t m = 0; m < n_members; m++) {
            size_t d = 0;
            read_annotation(is);
            is >> d;
            if (is.fail()) LOG_FATAL(".fg: factor %zu: bad member list", I);
            /* a variable has at least one state; a table beyond 2^28
             * entries (2 GiB of doubles) is refused */
            if (d == 0 || d > kMaxTable / table)
                LOG_FATAL(".fg: factor %zu: %zu states for member %zu (table limit %zu entries)", I, d, m,
                          kMaxTable);
            table *= d;
            dims.push_back(d);
        }
        {   /* a label listed twice would index past the (merged) table */
            std::vector<long> sorted_labels(labels);
            std::sort(sorted_labels.begin(), sorted_labels.end());
            if (std::adjacent_find(sorted_labels.begin(), sorted_labels.end()) != sorted_labels.end())
                LOG_FATAL(".fg: factor %zu lists a variable twice", I);
        }
        std::vector<Var> members;
        members.reserve(n_members);
        for (size_t m = 0; m < n_members; m++) {
            auto it = dims_of.find(labels[m]);
            if (it == dims_of.end()) dims_of[labels[m]] = dims[m];
            else if (it->second != dims[m])
                LOG_FATAL(".fg: variable %ld has %zu states in factor %zu and %zu before", labels[m], dims[m], I,
                          it->second);
            members.emplace_back((size_t)labels[m], dims[m]);
        }
        facs.emplace_back(VarSet(members.begin(), members.end(), members.size()), Real(0.0));
        const Permute perm(members);
        size_t n_entries = 0;
        read_annotation(is);
        is >> n_entries;
        for (size_t k = 0; k < n_entries; k++) {
            size_t li;
            Real val;
            read_annotation(is);
            is >> li;
            read_annotation(is);
            is >> val;
            if (is.fail()) LOG_FATAL(".fg: factor %zu: bad entry %zu", I, k);
            facs.back().set(perm.convertLinearIndex(li), val);
        }
    }
    fg = FactorGraph(facs);
    return is;
}

}  // namespace lbp
