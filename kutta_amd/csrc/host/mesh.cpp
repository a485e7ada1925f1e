// Host mesh setup (see mesh.hpp). Faces of a generated box are deduplicated
// through a hash of their sorted node ids rather than an ordered map of node
// vectors; the resulting mesh is the reference's, face for face.
#include "mesh.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <numeric>
#include <queue>
#include <random>
#include <tuple>
#include <limits>
#include <unordered_map>

namespace kutta_amd {
namespace {

// Vector arithmetic of mesh/detail.hpp:145-164; a quotient is a product with
// the reciprocal there, and is here too.
inline Vec3 add(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 sub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 scale(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 divide(Vec3 a, double s) { return scale(a, 1 / s); }
inline double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double length(Vec3 a) { return std::sqrt(dot(a, a)); }

// One face record during generation.
struct FaceRec {
  std::array<Index, 4> nodes{};
  int count = 0;
  Index owner = -1, neighbour = -1, patch = -1, order = 0;
};
struct FaceKey {
  std::array<Index, 4> sorted{};
  bool operator==(const FaceKey& o) const { return sorted == o.sorted; }
};
struct FaceKeyHash {
  std::size_t operator()(const FaceKey& k) const {
    std::uint64_t h = 1469598103934665603ull;
    for (Index v : k.sorted) h = (h ^ std::uint64_t(std::uint32_t(v))) * 1099511628211ull;
    return std::size_t(h ^ (h >> 29));
  }
};

// Face centroid and area vector (mesh/mesh.cpp:13-36).
struct FaceGeo {
  Vec3 centroid, area;
  bool ok;
};
FaceGeo face_geometry(const PolyMesh& m, Index f) {
  const Index b = m.face_offsets[f], s = m.face_offsets[f + 1] - b;
  const Index* ns = m.face_nodes.data() + b;
  const Vec3 origin = m.points[ns[0]];
  Vec3 d{};
  for (Index k = 0; k < s; ++k) d = add(d, sub(m.points[ns[k]], origin));
  const Vec3 center = add(origin, divide(d, double(s)));
  Vec3 av{};
  for (Index k = 0; k < s; ++k)
    av = add(av, scale(cross(sub(m.points[ns[k]], center), sub(m.points[ns[(k + 1) % s]], center)), 0.5));
  const double magnitude = length(av);
  const Vec3 normal = divide(av, magnitude);
  Vec3 moment{};
  double area = 0;
  for (Index k = 0; k < s; ++k) {
    const Vec3 a = sub(m.points[ns[k]], center), c = sub(m.points[ns[(k + 1) % s]], center);
    const double w = dot(scale(cross(a, c), 0.5), normal);
    area += w;
    moment = add(moment, scale(add(a, c), w / 3));
  }
  const bool ok = magnitude > 0 && std::isfinite(magnitude) && area > 0 && std::isfinite(area);
  return {add(center, divide(moment, area)), av, ok};
}

}  // namespace

MeshGeometry compute_geometry(const PolyMesh& m, bool strict) {
  MeshGeometry g;
  const Index nf = m.n_faces(), ni = m.n_internal_faces(), nc = m.n_cells;
  g.face_centroid.resize(nf);
  g.face_area.resize(nf);
  for (Index f = 0; f < nf; ++f) {
    const FaceGeo fg = face_geometry(m, f);
    KA_REQUIRE(!strict || fg.ok, "face " + std::to_string(f) + " has zero/invalid area");
    g.face_centroid[f] = fg.centroid;
    g.face_area[f] = fg.area;
  }
  // mesh/mesh.cpp:101-153: anchored mean of face centroids, then pyramids.
  std::vector<Vec3> anchor(nc), estimate(nc);
  std::vector<Index> count(nc, 0);
  auto first_pass = [&](Index c, Vec3 fc) {
    if (count[c] == 0) anchor[c] = fc;
    estimate[c] = add(estimate[c], sub(fc, anchor[c]));
    ++count[c];
  };
  for (Index f = 0; f < nf; ++f) {
    first_pass(m.owner[f], g.face_centroid[f]);
    if (f < ni) first_pass(m.neighbour[f], g.face_centroid[f]);
  }
  for (Index c = 0; c < nc; ++c) {
    KA_REQUIRE(!strict || count[c] >= 4, "cell " + std::to_string(c) + " has fewer than four faces");
    estimate[c] = add(anchor[c], divide(estimate[c], double(count[c])));
  }
  g.cell_centroid.assign(nc, {});
  g.cell_volume.assign(nc, 0.0);
  auto pyramid = [&](Index c, double sign, Index f) {
    const Vec3 r = sub(g.face_centroid[f], estimate[c]);
    const double v = sign * dot(g.face_area[f], r) / 3;
    g.cell_volume[c] += v;
    g.cell_centroid[c] = add(g.cell_centroid[c], scale(r, 0.75 * v));
  };
  for (Index f = 0; f < nf; ++f) {
    pyramid(m.owner[f], 1, f);
    if (f < ni) pyramid(m.neighbour[f], -1, f);
  }
  for (Index c = 0; c < nc; ++c) {
    KA_REQUIRE(!strict || (g.cell_volume[c] > 0 && std::isfinite(g.cell_volume[c])),
               "cell " + std::to_string(c) + " has non-positive/invalid volume");
    g.cell_centroid[c] = add(estimate[c], divide(g.cell_centroid[c], g.cell_volume[c]));
  }
  return g;
}

void validate(const PolyMesh& m, const MeshGeometry& g) {
  std::vector<Vec3> closure(m.n_cells);
  std::vector<double> areas(m.n_cells, 0.0);
  for (Index f = 0; f < m.n_faces(); ++f) {
    const Index o = m.owner[f];
    closure[o] = add(closure[o], g.face_area[f]);
    areas[o] += length(g.face_area[f]);
    KA_REQUIRE(dot(g.face_area[f], sub(g.face_centroid[f], g.cell_centroid[o])) > 0,
               "face " + std::to_string(f) + " normal does not point out of owner");
    if (f < m.n_internal_faces()) {
      const Index n = m.neighbour[f];
      closure[n] = add(closure[n], scale(g.face_area[f], -1.0));
      areas[n] += length(g.face_area[f]);
      KA_REQUIRE(dot(g.face_area[f], sub(g.cell_centroid[n], g.face_centroid[f])) > 0,
                 "face " + std::to_string(f) + " normal does not point into neighbour");
    }
  }
  for (Index c = 0; c < m.n_cells; ++c)
    KA_REQUIRE(length(closure[c]) <= 1e-10 * areas[c], "cell " + std::to_string(c) + " is not closed");
}

PolyMesh make_box(int nx, int ny, int nz, Vec3 lo, Vec3 hi, bool tet) {
  KA_REQUIRE(nx > 0 && ny > 0 && nz > 0, "box dimensions must be positive");
  KA_REQUIRE(std::isfinite(lo.x) && std::isfinite(hi.x) && hi.x > lo.x && hi.y > lo.y && hi.z > lo.z,
             "box requires finite lo < hi in every coordinate");
  const std::uint64_t bricks = std::uint64_t(nx) * ny * nz;
  KA_REQUIRE(bricks * (tet ? 36 : 6) < (1ull << 31), "box exceeds 32-bit mesh capacity");
  const Index sx = nx + 1, sy = ny + 1;
  auto node = [&](int i, int j, int k) { return Index((std::int64_t(k) * sy + j) * sx + i); };
  PolyMesh m;
  m.points.reserve(std::size_t(sx) * sy * (nz + 1));
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i)
        m.points.push_back({lo.x + (hi.x - lo.x) * i / nx, lo.y + (hi.y - lo.y) * j / ny,
                            lo.z + (hi.z - lo.z) * k / nz});
  m.n_cells = Index(bricks * (tet ? 6 : 1));
  std::vector<FaceRec> faces;
  faces.reserve(std::size_t(m.n_cells) * (tet ? 2 : 3) + 16);
  std::unordered_map<FaceKey, Index, FaceKeyHash> lookup;
  lookup.reserve(faces.capacity());
  // Adds the faces of one cell: the first visit of a face owns it, with nodes
  // oriented out of the owner; the second visit is the neighbour.
  auto add_cell = [&](Index c, const Index* const* cell_faces, const int* sizes, int nfaces) {
    // Every face-node occurrence of the cell (hex: 6 faces x 4 nodes) before
    // deduplication to its vertex set.
    std::array<Index, 24> verts{};
    int nv = 0;
    for (int q = 0; q < nfaces; ++q)
      for (int s = 0; s < sizes[q]; ++s) verts[nv++] = cell_faces[q][s];
    std::sort(verts.begin(), verts.begin() + nv);
    nv = int(std::unique(verts.begin(), verts.begin() + nv) - verts.begin());
    const Vec3 origin = m.points[verts[0]];
    Vec3 center{};
    for (int q = 0; q < nv; ++q) center = add(center, sub(m.points[verts[q]], origin));
    center = add(origin, divide(center, double(nv)));
    for (int q = 0; q < nfaces; ++q) {
      FaceKey key;
      key.sorted.fill(-1);
      std::copy(cell_faces[q], cell_faces[q] + sizes[q], key.sorted.begin());
      std::sort(key.sorted.begin(), key.sorted.begin() + sizes[q]);
      const auto [it, added] = lookup.emplace(key, Index(faces.size()));
      if (!added) {
        FaceRec& f = faces[it->second];
        KA_REQUIRE(f.neighbour < 0, "construction: more than two cells share a face");
        f.neighbour = c;
        continue;
      }
      FaceRec f;
      f.count = sizes[q];
      std::copy(cell_faces[q], cell_faces[q] + sizes[q], f.nodes.begin());
      const Vec3 r = m.points[f.nodes[0]];
      Vec3 fc{}, av{};
      for (int s = 0; s < f.count; ++s) fc = add(fc, sub(m.points[f.nodes[s]], r));
      fc = add(r, divide(fc, double(f.count)));
      for (int s = 0; s < f.count; ++s)
        av = add(av, cross(sub(m.points[f.nodes[s]], fc), sub(m.points[f.nodes[(s + 1) % f.count]], fc)));
      if (dot(av, sub(fc, center)) < 0) std::reverse(f.nodes.begin(), f.nodes.begin() + f.count);
      f.owner = c;
      f.order = Index(faces.size());
      faces.push_back(f);
    }
  };
  Index c = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const std::array<Index, 8> v{node(i, j, k),         node(i + 1, j, k),
                                     node(i + 1, j + 1, k), node(i, j + 1, k),
                                     node(i, j, k + 1),     node(i + 1, j, k + 1),
                                     node(i + 1, j + 1, k + 1), node(i, j + 1, k + 1)};
        if (tet) {
          static constexpr int ring[6] = {1, 2, 3, 7, 4, 5};
          for (int t = 0; t < 6; ++t) {
            const Index a = v[0], b = v[ring[t]], cc = v[ring[(t + 1) % 6]], d = v[6];
            const Index f0[3] = {a, b, cc}, f1[3] = {a, d, b}, f2[3] = {b, d, cc}, f3[3] = {cc, d, a};
            const Index* fs[4] = {f0, f1, f2, f3};
            const int sizes[4] = {3, 3, 3, 3};
            add_cell(c++, fs, sizes, 4);
          }
        } else {
          const Index f0[4] = {v[0], v[3], v[2], v[1]}, f1[4] = {v[4], v[5], v[6], v[7]};
          const Index f2[4] = {v[0], v[1], v[5], v[4]}, f3[4] = {v[1], v[2], v[6], v[5]};
          const Index f4[4] = {v[2], v[3], v[7], v[6]}, f5[4] = {v[3], v[0], v[4], v[7]};
          const Index* fs[6] = {f0, f1, f2, f3, f4, f5};
          const int sizes[6] = {4, 4, 4, 4, 4, 4};
          add_cell(c++, fs, sizes, 6);
        }
      }
  // Boundary classification: the first box side holding every node.
  for (auto& f : faces) {
    if (f.neighbour >= 0) continue;
    for (Index side = 0; side < 6 && f.patch < 0; ++side) {
      bool all = true;
      for (int s = 0; s < f.count; ++s) {
        const Index n = f.nodes[s], ii = n % sx, jj = (n / sx) % sy, kk = n / (sx * sy);
        all = all && (side == 0 ? ii == 0 : side == 1 ? ii == nx : side == 2 ? jj == 0
                      : side == 3 ? jj == ny : side == 4 ? kk == 0 : kk == nz);
      }
      if (all) f.patch = side;
    }
    KA_REQUIRE(f.patch >= 0, "box boundary classification failed");
  }
  // Public face order (mesh/topology.cpp:27-33): internal faces by (owner,
  // neighbour), then boundary faces by patch in creation order.
  std::vector<Index> order(faces.size());
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](Index a, Index b) {
    const FaceRec &fa = faces[a], &fb = faces[b];
    const bool ba = fa.neighbour < 0, bb = fb.neighbour < 0;
    return std::tuple(ba, ba ? fa.patch : fa.owner, fa.neighbour, fa.order) <
           std::tuple(bb, bb ? fb.patch : fb.owner, fb.neighbour, fb.order);
  });
  const char* names[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
  for (int p = 0; p < 6; ++p) m.patches.push_back({names[p], 0, 0});
  for (Index idx : order) {
    const FaceRec& f = faces[idx];
    m.owner.push_back(f.owner);
    if (f.neighbour >= 0) m.neighbour.push_back(f.neighbour);
    else ++m.patches[f.patch].size;
    m.face_nodes.insert(m.face_nodes.end(), f.nodes.begin(), f.nodes.begin() + f.count);
    m.face_offsets.push_back(Index(m.face_nodes.size()));
  }
  Index start = m.n_internal_faces();
  for (auto& p : m.patches) {
    p.start = start;
    start += p.size;
  }
  validate(m, compute_geometry(m));
  return m;
}

CellFaces build_cell_faces(const PolyMesh& m) {
  CellFaces a;
  a.offsets.assign(std::size_t(m.n_cells) + 1, 0);
  for (Index c : m.owner) ++a.offsets[c + 1];
  for (Index c : m.neighbour) ++a.offsets[c + 1];
  std::partial_sum(a.offsets.begin(), a.offsets.end(), a.offsets.begin());
  a.faces.resize(a.offsets.back());
  a.signs.resize(a.faces.size());
  std::vector<Index> cursor(a.offsets.begin(), a.offsets.end() - 1);
  for (Index f = 0; f < m.n_faces(); ++f) {
    Index k = cursor[m.owner[f]]++;
    a.faces[k] = f;
    a.signs[k] = 1;
    if (f < m.n_internal_faces()) {
      k = cursor[m.neighbour[f]]++;
      a.faces[k] = f;
      a.signs[k] = -1;
    }
  }
  return a;
}

CellNeighbours build_cell_neighbours(const PolyMesh& m) {
  const CellFaces cf = build_cell_faces(m);
  CellNeighbours a;
  a.offsets.reserve(std::size_t(m.n_cells) + 1);
  a.offsets.push_back(0);
  std::vector<Index> row;
  for (Index c = 0; c < m.n_cells; ++c) {
    row.clear();
    for (Index k = cf.offsets[c]; k < cf.offsets[c + 1]; ++k) {
      const Index f = cf.faces[k];
      if (f < m.n_internal_faces()) row.push_back(m.owner[f] == c ? m.neighbour[f] : m.owner[f]);
    }
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    a.cells.insert(a.cells.end(), row.begin(), row.end());
    a.offsets.push_back(Index(a.cells.size()));
  }
  return a;
}

MeshPermutation renumber_cells_rcm(PolyMesh& m) {
  const CellNeighbours adj = build_cell_neighbours(m);
  const Index n = m.n_cells;
  auto degree = [&](Index c) { return adj.offsets[c + 1] - adj.offsets[c]; };
  // Breadth-first from the lowest-degree unseen seed, neighbours by degree
  // (stable), each component reversed (mesh/topology.cpp:148-176).
  std::vector<Index> seeds(n);
  std::iota(seeds.begin(), seeds.end(), 0);
  std::stable_sort(seeds.begin(), seeds.end(), [&](Index a, Index b) { return degree(a) < degree(b); });
  MeshPermutation p;
  p.cells_new_to_old.reserve(n);
  std::vector<char> seen(n, 0);
  std::vector<Index> component, row;
  for (Index seed : seeds) {
    if (seen[seed]) continue;
    component.clear();
    std::size_t head = 0;
    component.push_back(seed);
    seen[seed] = 1;
    while (head < component.size()) {
      const Index c = component[head++];
      row.clear();
      for (Index k = adj.offsets[c]; k < adj.offsets[c + 1]; ++k)
        if (!seen[adj.cells[k]]) row.push_back(adj.cells[k]);
      std::stable_sort(row.begin(), row.end(), [&](Index a, Index b) { return degree(a) < degree(b); });
      for (Index nb : row)
        if (!seen[nb]) {
          seen[nb] = 1;
          component.push_back(nb);
        }
    }
    p.cells_new_to_old.insert(p.cells_new_to_old.end(), component.rbegin(), component.rend());
  }
  std::vector<Index> old_to_new(n);
  for (Index c = 0; c < n; ++c) old_to_new[p.cells_new_to_old[c]] = c;
  // Faces: renumbered owner < neighbour (winding reversed on a swap), then
  // (internal?, patch, owner, neighbour) order (mesh/topology.cpp:177-248).
  const Index nf = m.n_faces(), ni = m.n_internal_faces();
  std::vector<Index> patch(nf, -1), owner(nf), neighbour(nf, -1);
  std::vector<char> flipped(nf, 0);
  for (Index q = 0; q < Index(m.patches.size()); ++q)
    for (Index f = m.patches[q].start; f < m.patches[q].start + m.patches[q].size; ++f) patch[f] = q;
  for (Index f = 0; f < nf; ++f) {
    owner[f] = old_to_new[m.owner[f]];
    if (f < ni) neighbour[f] = old_to_new[m.neighbour[f]];
    if (neighbour[f] >= 0 && owner[f] > neighbour[f]) {
      std::swap(owner[f], neighbour[f]);
      flipped[f] = 1;
    }
  }
  p.faces_new_to_old.resize(nf);
  std::iota(p.faces_new_to_old.begin(), p.faces_new_to_old.end(), 0);
  std::stable_sort(p.faces_new_to_old.begin(), p.faces_new_to_old.end(), [&](Index a, Index b) {
    return std::tuple(neighbour[a] < 0, patch[a], owner[a], neighbour[a]) <
           std::tuple(neighbour[b] < 0, patch[b], owner[b], neighbour[b]);
  });
  PolyMesh r;
  r.points = std::move(m.points);
  r.n_cells = n;
  r.patches = std::move(m.patches);
  r.owner.resize(nf);
  r.neighbour.resize(ni);
  r.face_offsets.assign(std::size_t(nf) + 1, 0);
  r.face_nodes.resize(m.face_nodes.size());
  Index cursor = 0;
  for (Index i = 0; i < nf; ++i) {
    const Index f = p.faces_new_to_old[i];
    r.owner[i] = owner[f];
    if (i < ni) r.neighbour[i] = neighbour[f];
    const auto first = m.face_nodes.begin() + m.face_offsets[f];
    const auto last = m.face_nodes.begin() + m.face_offsets[f + 1];
    if (flipped[f]) std::reverse_copy(first, last, r.face_nodes.begin() + cursor);
    else std::copy(first, last, r.face_nodes.begin() + cursor);
    cursor += Index(last - first);
    r.face_offsets[i + 1] = cursor;
  }
  m = std::move(r);
  validate(m, compute_geometry(m));
  return p;
}

int perturb_box(PolyMesh& m, int nx, int ny, int nz, Vec3 lo, Vec3 hi, double amplitude,
                std::uint64_t seed) {
  // Each point (i, j, k) moves by (2u - 1) * amplitude * h per axis, u from
  // mt19937_64(seed) as (next >> 11) * 2^-53, three draws per point in point
  // order; not along an axis whose boundary plane holds the point. Points of
  // cells or faces that fail the FV geometry checks are redrawn, in point
  // order, from the continuing stream until none fails.
  const double hx = (hi.x - lo.x) / nx, hy = (hi.y - lo.y) / ny, hz = (hi.z - lo.z) / nz;
  const std::vector<Vec3> base = m.points;
  std::mt19937_64 gen(seed);
  auto uniform = [&] { return double(gen() >> 11) * 0x1.0p-53; };
  const Index sx = nx + 1, sy = ny + 1;
  auto draw = [&](Index v) {
    const Index i = v % sx, j = (v / sx) % sy, k = v / (sx * sy);
    double dx = (2 * uniform() - 1) * amplitude * hx;
    double dy = (2 * uniform() - 1) * amplitude * hy;
    double dz = (2 * uniform() - 1) * amplitude * hz;
    if (i == 0 || i == nx) dx = 0;
    if (j == 0 || j == ny) dy = 0;
    if (k == 0 || k == nz) dz = 0;
    m.points[v] = {base[v].x + dx, base[v].y + dy, base[v].z + dz};
  };
  for (Index v = 0; v < Index(m.points.size()); ++v) draw(v);
  for (int round = 0;; ++round) {
    const MeshGeometry g = compute_geometry(m, false);
    std::vector<char> bad_cell(m.n_cells, 0), bad_point(m.points.size(), 0);
    for (Index c = 0; c < m.n_cells; ++c)
      bad_cell[c] = !(g.cell_volume[c] > 0) || !std::isfinite(g.cell_volume[c]);
    bool any = false;
    for (Index f = 0; f < m.n_faces(); ++f) {
      const Index o = m.owner[f];
      const bool internal = f < m.n_internal_faces();
      const Vec3 link = sub(internal ? g.cell_centroid[m.neighbour[f]] : g.face_centroid[f], g.cell_centroid[o]);
      bool ok = dot(g.face_area[f], link) > 0 &&
                dot(g.face_area[f], sub(g.face_centroid[f], g.cell_centroid[o])) > 0;
      if (internal) ok = ok && dot(g.face_area[f], sub(g.cell_centroid[m.neighbour[f]], g.face_centroid[f])) > 0;
      if (ok && !bad_cell[o] && !(internal && bad_cell[m.neighbour[f]])) continue;
      any = true;
      for (Index q = m.face_offsets[f]; q < m.face_offsets[f + 1]; ++q) bad_point[m.face_nodes[q]] = 1;
    }
    if (!any) return round;
    KA_REQUIRE(round < 64, "perturbation: no valid mesh after 64 redraw rounds");
    for (Index v = 0; v < Index(m.points.size()); ++v)
      if (bad_point[v]) draw(v);
  }
}

Index locate_cell(const PolyMesh& m, const MeshGeometry& g, Vec3 point) {
  const CellFaces cf = build_cell_faces(m);
  Index nearest = 0;
  double best = std::numeric_limits<double>::infinity();
  for (Index c = 0; c < m.n_cells; ++c) {
    bool inside = true;
    for (Index k = cf.offsets[c]; k < cf.offsets[c + 1] && inside; ++k) {
      const Index f = cf.faces[k];
      const Vec3 s = scale(g.face_area[f], double(cf.signs[k]));
      inside = dot(s, sub(point, g.face_centroid[f])) <= 1e-12 * length(s) * std::cbrt(g.cell_volume[c]);
    }
    if (inside) return c;
    const Vec3 d = sub(point, g.cell_centroid[c]);
    if (dot(d, d) < best) {
      best = dot(d, d);
      nearest = c;
    }
  }
  return nearest;
}

Stencil make_stencil(const PolyMesh& m, const MeshGeometry& g) {
  Stencil s;
  const CellFaces cf = build_cell_faces(m);
  const CellNeighbours cn = build_cell_neighbours(m);
  const Index nc = m.n_cells, ni = m.n_internal_faces();
  s.lsq_inverse.assign(6 * std::size_t(nc), 0.0f);
  s.row_ptr.reserve(std::size_t(nc) + 1);
  s.col.reserve(cn.cells.size() + nc);
  s.diag.reserve(nc);
  for (Index c = 0; c < nc; ++c) {
    // Row c: sorted neighbours with the diagonal inserted in place.
    const Index row_begin = Index(s.col.size());
    bool placed = false;
    for (Index k = cn.offsets[c]; k < cn.offsets[c + 1]; ++k) {
      if (!placed && cn.cells[k] > c) {
        s.diag.push_back(Index(s.col.size()));
        s.col.push_back(c);
        placed = true;
      }
      s.col.push_back(cn.cells[k]);
    }
    if (!placed) {
      s.diag.push_back(Index(s.col.size()));
      s.col.push_back(c);
    }
    s.row_ptr.push_back(Index(s.col.size()));
    std::array<double, 6> a{};
    for (Index k = cf.offsets[c]; k < cf.offsets[c + 1]; ++k) {
      const Index f = cf.faces[k];
      const Index n = f < ni ? (m.owner[f] == c ? m.neighbour[f] : m.owner[f]) : -(f - ni + 1);
      s.adjacent.push_back(n);
      s.matrix_entry.push_back(
          n < 0 ? -1 : Index(std::lower_bound(s.col.begin() + row_begin, s.col.end(), n) - s.col.begin()));
      const Vec3 d = sub(n < 0 ? g.face_centroid[f] : g.cell_centroid[n], g.cell_centroid[c]);
      const double w = 1 / dot(d, d);
      const Vec3 lsq = scale(d, w);
      s.lsq_x.push_back(float(lsq.x));
      s.lsq_y.push_back(float(lsq.y));
      s.lsq_z.push_back(float(lsq.z));
      a[0] += w * d.x * d.x; a[1] += w * d.x * d.y; a[2] += w * d.x * d.z;
      a[3] += w * d.y * d.y; a[4] += w * d.y * d.z; a[5] += w * d.z * d.z;
    }
    const double det = a[0] * (a[3] * a[5] - a[4] * a[4]) - a[1] * (a[1] * a[5] - a[4] * a[2]) +
                       a[2] * (a[1] * a[4] - a[3] * a[2]);
    KA_REQUIRE(det > 1e-14, "FV LSQ stencil is rank deficient");
    const double inv[6] = {(a[3] * a[5] - a[4] * a[4]) / det, (a[2] * a[4] - a[1] * a[5]) / det,
                           (a[1] * a[4] - a[2] * a[3]) / det, (a[0] * a[5] - a[2] * a[2]) / det,
                           (a[1] * a[2] - a[0] * a[4]) / det, (a[0] * a[3] - a[1] * a[1]) / det};
    for (int k = 0; k < 6; ++k) s.lsq_inverse[c + std::size_t(k) * nc] = float(inv[k]);
  }
  s.diffusion.reserve(m.n_faces());
  for (Index f = 0; f < m.n_faces(); ++f) {
    const Vec3 d = sub(f < ni ? g.cell_centroid[m.neighbour[f]] : g.face_centroid[f], g.cell_centroid[m.owner[f]]);
    const Vec3 area = g.face_area[f];
    const double projection = dot(area, d);
    KA_REQUIRE(projection > 0, "FV requires positive face/cell projection");
    const double coeff = dot(area, area) / projection;
    s.diffusion.push_back(float(coeff));
    const Vec3 no = sub(area, scale(d, coeff));
    s.nonorth_x.push_back(float(no.x));
    s.nonorth_y.push_back(float(no.y));
    s.nonorth_z.push_back(float(no.z));
  }
  return s;
}

}  // namespace kutta_amd
