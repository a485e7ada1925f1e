// Host aggregation for the device AMG hierarchy. The transpose is built once
// per operator with, for every entry, the index of its mirror entry, so edge
// strengths and canonical entry ids need no per-entry binary search.
#include "amg_setup.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace kutta_amd {

float add_up(float a, float b) {
  float s = a + b;
  float part = s - a;
  float err = (a - (s - part)) + (b - part);
  return err > 0.0f ? std::nextafter(s, std::numeric_limits<float>::infinity()) : s;
}
void UpperSum::add(float b) {
  float s = sum + b;
  float part = s - sum;
  float err = (sum - (s - part)) + (b - part);
  sum = s;
  error = add_up(error, err);
}
float UpperSum::value() const { return add_up(sum, error); }

namespace {

// Incoming edges of every row, in increasing source order, with the position
// of the stored entry (source -> row).
struct Transpose {
  std::vector<Index> offsets, sources, entries;
};
Transpose transpose(const CsrHost& a) {
  Transpose t;
  t.offsets.assign(std::size_t(a.n_rows) + 1, 0);
  t.sources.resize(a.nnz());
  t.entries.resize(a.nnz());
  for (auto c : a.col) ++t.offsets[c + 1];
  std::partial_sum(t.offsets.begin(), t.offsets.end(), t.offsets.begin());
  std::vector<Index> cursor(t.offsets.begin(), t.offsets.end() - 1);
  for (Index i = 0; i < a.n_rows; ++i)
    for (Index j = a.row_ptr[i]; j < a.row_ptr[i + 1]; ++j) {
      const Index k = cursor[a.col[j]]++;
      t.sources[k] = i;
      t.entries[k] = j;
    }
  return t;
}
// mirror[j] = index of entry (col, row) for entry j = (row, col), or -1.
std::vector<Index> mirrors(const CsrHost& a, const Transpose& t) {
  std::vector<Index> mirror(a.nnz(), -1);
  for (Index i = 0; i < a.n_rows; ++i) {
    Index out = a.row_ptr[i], in = t.offsets[i];
    while (out < a.row_ptr[i + 1] && in < t.offsets[i + 1]) {
      const Index oc = a.col[out], ic = t.sources[in];
      if (oc == ic) {
        mirror[out] = t.entries[in];
        ++out;
        ++in;
      } else if (oc < ic) {
        ++out;
      } else {
        ++in;
      }
    }
  }
  return mirror;
}

}  // namespace

void validate_host(const CsrHost& p) {
  KA_REQUIRE(p.n_rows >= 0, "CSR: negative row count");
  KA_REQUIRE(p.row_ptr.size() == std::size_t(p.n_rows) + 1 && p.diag.size() == std::size_t(p.n_rows),
             "CSR: invalid array sizes");
  KA_REQUIRE(p.row_ptr.front() == 0 && p.row_ptr.back() == p.nnz(), "CSR: invalid row offsets");
  for (Index i = 0; i < p.n_rows; ++i) {
    KA_REQUIRE(p.row_ptr[i] >= 0 && p.row_ptr[i] <= p.row_ptr[i + 1] && p.row_ptr[i + 1] <= p.nnz(),
               "CSR: invalid row range");
    Index previous = -1;
    for (Index j = p.row_ptr[i]; j < p.row_ptr[i + 1]; ++j) {
      KA_REQUIRE(p.col[j] > previous && p.col[j] < p.n_rows,
                 "CSR: columns must be in range, unique and sorted");
      previous = p.col[j];
    }
    KA_REQUIRE(p.diag[i] >= p.row_ptr[i] && p.diag[i] < p.row_ptr[i + 1] && p.col[p.diag[i]] == i,
               "CSR: missing or incorrect diagonal");
  }
  KA_REQUIRE(p.val.size() == p.col.size(), "CSR: value count mismatch");
  for (double v : p.val)
    KA_REQUIRE(std::isfinite(v) && std::abs(v) <= std::numeric_limits<float>::max(),
               "CSR: nonfinite or out-of-range value");
}

std::vector<Index> pair_rows(const CsrHost& a, double threshold, bool exclude_isolated) {
  const Transpose t = transpose(a);
  const Index n = a.n_rows;
  // Visits each distinct neighbour col != i with weight max(|a_ic|, |a_ci|) in
  // increasing column order, skipping zero weights.
  auto neighbours = [&](Index i, auto&& visit) {
    Index out = a.row_ptr[i], in = t.offsets[i];
    const Index out_end = a.row_ptr[i + 1], in_end = t.offsets[i + 1];
    while (out < out_end || in < in_end) {
      const Index oc = out < out_end ? a.col[out] : n;
      const Index ic = in < in_end ? t.sources[in] : n;
      const Index col = std::min(oc, ic);
      double w = 0;
      if (oc == col) w = std::abs(a.val[out++]);
      if (ic == col) w = std::max(w, std::abs(a.val[t.entries[in++]]));
      if (col != i && w != 0) visit(col, w);
    }
  };
  std::vector<Index> map(n, -1);
  Index next = 0;
  for (Index i = 0; i < n; ++i) {
    if (map[i] != -1) continue;
    double strongest = 0;
    neighbours(i, [&](Index, double w) { strongest = std::max(strongest, w); });
    if (exclude_isolated && strongest == 0) continue;
    Index partner = -1;
    double best = 0;
    const double floor = threshold * strongest;
    neighbours(i, [&](Index col, double w) {
      if (map[col] == -1 && w > best && w >= floor) {
        partner = col;
        best = w;
      }
    });
    map[i] = next;
    if (partner != -1) map[partner] = next;
    ++next;
  }
  return map;
}

Coarsening coarsen(const CsrHost& fine, std::vector<Index> aggregate, bool entry_map, bool m_matrix) {
  Coarsening c;
  c.aggregate = std::move(aggregate);
  const Index nc = c.aggregate.empty() ? 0
      : *std::max_element(c.aggregate.begin(), c.aggregate.end()) + 1;
  std::vector<Index> mirror;
  if (m_matrix) mirror = mirrors(fine, transpose(fine));
  auto& cp = c.matrix;
  cp.n_rows = nc;
  cp.row_ptr.assign(std::size_t(nc) + 1, 0);
  cp.diag.assign(nc, 0);
  c.member_offsets.assign(std::size_t(nc) + 1, 0);
  for (auto g : c.aggregate) if (g >= 0) ++c.member_offsets[g + 1];
  std::partial_sum(c.member_offsets.begin(), c.member_offsets.end(), c.member_offsets.begin());
  c.members.resize(c.member_offsets.back());
  {
    std::vector<Index> cursor(c.member_offsets.begin(), c.member_offsets.end() - 1);
    for (Index i = 0; i < fine.n_rows; ++i)
      if (c.aggregate[i] >= 0) c.members[cursor[c.aggregate[i]]++] = i;
  }
  if (entry_map) {
    c.entries.reserve(fine.nnz());
    c.entry_offsets.reserve(fine.nnz() + 1);
    c.entry_offsets.push_back(0);
  }
  cp.col.reserve(fine.nnz());
  cp.val.reserve(fine.nnz());
  // Each contributing fine entry sorts by (coarse column, canonical id, entry):
  // opposite coarse entries then sum the same float edges in the same order.
  struct Term {
    Index col, canonical, source;
    bool operator<(const Term& o) const {
      if (col != o.col) return col < o.col;
      if (canonical != o.canonical) return canonical < o.canonical;
      return source < o.source;
    }
  };
  std::vector<Term> row;
  for (Index g = 0; g < nc; ++g) {
    row.clear();
    for (Index m = c.member_offsets[g]; m < c.member_offsets[g + 1]; ++m) {
      const Index i = c.members[m];
      for (Index j = fine.row_ptr[i]; j < fine.row_ptr[i + 1]; ++j) {
        const Index target = c.aggregate[fine.col[j]];
        if (target < 0) continue;
        Index canonical = j;
        if (m_matrix) {
          KA_REQUIRE(mirror[j] >= 0, "AMG M-matrix: expected a symmetric pattern");
          canonical = std::min(j, mirror[j]);
        }
        row.push_back({target, canonical, j});
      }
    }
    std::sort(row.begin(), row.end());
    Index previous = -1;
    float s = 0, e = 0;
    for (const auto& term : row) {
      if (term.col != previous) {
        if (previous != -1) cp.val.back() = s + e;
        s = e = 0;
        if (entry_map && previous != -1) c.entry_offsets.push_back(Index(c.entries.size()));
        if (term.col == g) cp.diag[g] = Index(cp.col.size());
        cp.col.push_back(term.col);
        cp.val.push_back(0);
        previous = term.col;
      }
      if (entry_map) c.entries.push_back(term.source);
      // Neumaier (linalg/amg.cpp:135-138).
      const float t = static_cast<float>(fine.val[term.source]);
      const float nx = s + t;
      e += std::abs(s) >= std::abs(t) ? (s - nx) + t : (t - nx) + s;
      s = nx;
    }
    if (previous != -1) cp.val.back() = s + e;
    if (entry_map) c.entry_offsets.push_back(Index(c.entries.size()));
    cp.row_ptr[g + 1] = Index(cp.col.size());
    if (m_matrix) {
      // linalg/amg.cpp:143-155: aggregated row surpluses minus coarse edges.
      UpperSum d{0, 0};
      for (Index m = c.member_offsets[g]; m < c.member_offsets[g + 1]; ++m) {
        const Index i = c.members[m];
        UpperSum surplus{float(fine.val[fine.diag[i]]), 0};
        for (Index j = fine.row_ptr[i]; j < fine.row_ptr[i + 1]; ++j)
          if (j != fine.diag[i]) surplus.add(float(fine.val[j]));
        d.add(surplus.value());
      }
      for (Index j = cp.row_ptr[g]; j < cp.row_ptr[g + 1]; ++j)
        if (j != cp.diag[g]) d.add(-float(cp.val[j]));
      cp.val[cp.diag[g]] = d.value();
    }
  }
  return c;
}

Coarsening aggregate(const CsrHost& a, const AmgOptions& options) {
  const bool m = options.structure == Structure::SymmetricMMatrix;
  auto map = pair_rows(a, options.strength_threshold, true);
  for (int pass = 1; pass < options.aggregation_passes; ++pass) {
    auto intermediate = coarsen(a, std::move(map), false, m);
    const auto next = pair_rows(intermediate.matrix, options.strength_threshold, false);
    map = std::move(intermediate.aggregate);
    for (auto& g : map) if (g >= 0) g = next[g];
  }
  return coarsen(a, std::move(map), true, m);
}

std::vector<float> diagonal_values(const CsrHost& a, bool l1, bool nullspace) {
  std::vector<float> inverse(a.n_rows);
  for (Index i = 0; i < a.n_rows; ++i) {
    const float d = static_cast<float>(a.val[a.diag[i]]);
    KA_REQUIRE(!l1 || d >= 0, "AMG: requires positive diagonals");
    KA_REQUIRE(d != 0 || (nullspace && a.n_rows == 1), "solver: invalid values or zero diagonal");
    float off = 0;
    if (l1)
      for (Index j = a.row_ptr[i]; j < a.row_ptr[i + 1]; ++j)
        if (j != a.diag[i]) off += std::abs(static_cast<float>(a.val[j]));
    const float scale = l1 ? std::max(std::abs(d), 0.5f * (std::abs(d) + off)) : d;
    inverse[i] = scale != 0 ? 1.0f / scale : 0;
    KA_REQUIRE(std::isfinite(scale) && std::isfinite(inverse[i]),
               "solver: invalid values or diagonal inverse");
  }
  return inverse;
}

void check_m_matrix(const CsrHost& a) {
  const auto t = transpose(a);
  const auto mirror = mirrors(a, t);
  for (Index i = 0; i < a.n_rows; ++i) {
    long double total = 0;
    for (Index j = a.row_ptr[i]; j < a.row_ptr[i + 1]; ++j) {
      const float value = float(a.val[j]);
      total += value;
      KA_REQUIRE(mirror[j] >= 0 && value == float(a.val[mirror[j]]) && (a.col[j] == i || value <= 0),
                 "AMG M-matrix: expected symmetric nonpositive edges");
    }
    KA_REQUIRE(total >= 0, "AMG M-matrix: expected diagonally dominant rows");
  }
}

}  // namespace kutta_amd
