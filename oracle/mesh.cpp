// ORACLE — test infrastructure only (see oracle.hpp). Restatement of the
// reference's mesh construction (io/generators.cpp, mesh/topology.cpp),
// FP64 geometry (mesh/mesh.cpp), RCM renumbering, FV stencil
// (fv/operators.cpp) and pressure assembly (kernels/fv_flow.kern.h).
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <numeric>
#include <queue>
#include <random>
#include <tuple>

#include "oracle_mesh.hpp"

namespace oracle {
namespace {

// mesh/detail.hpp:145-164 (division is multiplication by the reciprocal)
Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec3 operator*(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
Vec3 operator/(Vec3 a, double s) { return a * (1 / s); }
double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3 cross(Vec3 a, Vec3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
double norm(Vec3 a) { return std::sqrt(dot(a, a)); }

struct FaceRecord {
  std::vector<Index> nodes;
  Index owner, neighbour = -1, patch = -1;
};

// mesh/detail.hpp:175-181
Vec3 face_center(const PolyMesh& m, Index f) {
  const Index b = m.face_offsets[f], e = m.face_offsets[f + 1];
  const Vec3 origin = m.points[m.face_nodes[b]];
  Vec3 d{};
  for (Index k = b; k < e; ++k) d = d + (m.points[m.face_nodes[k]] - origin);
  return origin + d / static_cast<double>(e - b);
}

// mesh/mesh.cpp:13-36
struct FaceGeo {
  Vec3 centroid, area;
  bool ok;
};
FaceGeo face_geometry(const PolyMesh& m, Index f) {
  const Vec3 center = face_center(m, f);
  const Index b = m.face_offsets[f], e = m.face_offsets[f + 1], s = e - b;
  Vec3 moment{}, av{};
  double area = 0;
  for (Index j = 0; j < s; ++j)
    av = av + cross(m.points[m.face_nodes[b + j]] - center,
                    m.points[m.face_nodes[b + (j + 1) % s]] - center) * 0.5;
  const double magnitude = norm(av);
  bool ok = magnitude > 0 && std::isfinite(magnitude);
  const Vec3 normal = av / magnitude;
  for (Index j = 0; j < s; ++j) {
    const Vec3 a = m.points[m.face_nodes[b + j]] - center;
    const Vec3 c = m.points[m.face_nodes[b + (j + 1) % s]] - center;
    const double w = dot(cross(a, c) * 0.5, normal);
    area += w;
    moment = moment + (a + c) * (w / 3);
  }
  ok = ok && area > 0 && std::isfinite(area);
  return {center + moment / area, av, ok};
}

// mesh/mesh.cpp:101-166
Geometry geometry(const PolyMesh& m, bool strict) {
  Geometry g;
  const Index nf = m.n_faces();
  g.face_centroid.resize(nf);
  g.face_area.resize(nf);
  for (Index f = 0; f < nf; ++f) {
    const auto fg = face_geometry(m, f);
    ORACLE_REQUIRE(!strict || fg.ok, "face " + std::to_string(f) + " has invalid area");
    g.face_centroid[f] = fg.centroid;
    g.face_area[f] = fg.area;
  }
  g.cell_centroid.assign(m.n_cells, {});
  g.cell_volume.assign(m.n_cells, 0);
  std::vector<Vec3> estimate(m.n_cells), anchor(m.n_cells);
  std::vector<Index> count(m.n_cells, 0);
  for (Index f = 0; f < nf; ++f) {
    const Vec3 fc = g.face_centroid[f];
    auto add = [&](Index c) {
      if (!count[c]) anchor[c] = fc;
      estimate[c] = estimate[c] + (fc - anchor[c]);
      ++count[c];
    };
    add(m.owner[f]);
    if (f < m.n_internal_faces()) add(m.neighbour[f]);
  }
  for (Index c = 0; c < m.n_cells; ++c) {
    ORACLE_REQUIRE(!strict || count[c] >= 4, "cell has fewer than four faces");
    estimate[c] = anchor[c] + estimate[c] / count[c];
  }
  for (Index f = 0; f < nf; ++f) {
    auto add = [&](Index c, double sign) {
      const Vec3 r = g.face_centroid[f] - estimate[c];
      const double v = sign * dot(g.face_area[f], r) / 3;
      g.cell_volume[c] += v;
      g.cell_centroid[c] = g.cell_centroid[c] + r * (0.75 * v);
    };
    add(m.owner[f], 1);
    if (f < m.n_internal_faces()) add(m.neighbour[f], -1);
  }
  for (Index c = 0; c < m.n_cells; ++c) {
    ORACLE_REQUIRE(!strict || (g.cell_volume[c] > 0 && std::isfinite(g.cell_volume[c])),
                   "cell " + std::to_string(c) + " has non-positive volume");
    g.cell_centroid[c] = estimate[c] + g.cell_centroid[c] / g.cell_volume[c];
  }
  return g;
}

// mesh/topology.cpp:8-55
PolyMesh pack(std::vector<Vec3> points, Index cells, std::vector<FaceRecord> faces,
              std::vector<Patch> patches) {
  for (auto& f : faces)
    if (f.neighbour >= 0 && f.owner > f.neighbour) {
      std::swap(f.owner, f.neighbour);
      std::reverse(f.nodes.begin(), f.nodes.end());
    }
  std::stable_sort(faces.begin(), faces.end(), [](const FaceRecord& a, const FaceRecord& b) {
    return std::tuple(a.neighbour < 0, a.neighbour < 0 ? a.patch : a.owner, a.neighbour) <
           std::tuple(b.neighbour < 0, b.neighbour < 0 ? b.patch : b.owner, b.neighbour);
  });
  PolyMesh m;
  m.points = std::move(points);
  m.n_cells = cells;
  for (auto& p : patches) p.size = 0;
  for (const auto& f : faces) {
    m.owner.push_back(f.owner);
    if (f.neighbour >= 0) m.neighbour.push_back(f.neighbour);
    else ++patches[f.patch].size;
    m.face_nodes.insert(m.face_nodes.end(), f.nodes.begin(), f.nodes.end());
    m.face_offsets.push_back(Index(m.face_nodes.size()));
  }
  Index start = m.n_internal_faces();
  for (auto& p : patches) {
    p.start = start;
    start += p.size;
  }
  m.patches = std::move(patches);
  return m;
}

}  // namespace

Geometry compute_geometry(const PolyMesh& m) { return geometry(m, true); }
Geometry geometry_unchecked(const PolyMesh& m) { return geometry(m, false); }

// mesh/mesh.cpp:188-218 (orientation and closure; the edge-manifold check
// changes no output and is not restated).
void validate(const PolyMesh& m) {
  const Geometry g = geometry(m, true);
  std::vector<Vec3> closure(m.n_cells);
  std::vector<double> areas(m.n_cells, 0);
  for (Index f = 0; f < m.n_faces(); ++f) {
    auto add = [&](Index c, int sign) {
      closure[c] = closure[c] + g.face_area[f] * double(sign);
      areas[c] += norm(g.face_area[f]);
    };
    add(m.owner[f], 1);
    if (f < m.n_internal_faces()) add(m.neighbour[f], -1);
    ORACLE_REQUIRE(dot(g.face_area[f], g.face_centroid[f] - g.cell_centroid[m.owner[f]]) > 0,
                   "face normal does not point out of owner");
    if (f < m.n_internal_faces())
      ORACLE_REQUIRE(dot(g.face_area[f], g.cell_centroid[m.neighbour[f]] - g.face_centroid[f]) > 0,
                     "face normal does not point into neighbour");
  }
  for (Index c = 0; c < m.n_cells; ++c)
    ORACLE_REQUIRE(norm(closure[c]) <= 1e-10 * areas[c], "cell is not closed");
}

// io/generators.cpp:10-96 with mesh/topology.cpp:56-100 (from_cells)
PolyMesh make_box(int nx, int ny, int nz, Vec3 lo, Vec3 hi, bool tet) {
  ORACLE_REQUIRE(nx > 0 && ny > 0 && nz > 0, "box dimensions must be positive");
  auto node = [&](int i, int j, int k) {
    return Index((std::uint64_t(k) * (ny + 1) + j) * (nx + 1) + i);
  };
  std::vector<Vec3> points;
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i)
        points.push_back({lo.x + (hi.x - lo.x) * i / nx, lo.y + (hi.y - lo.y) * j / ny,
                          lo.z + (hi.z - lo.z) * k / nz});
  using Cell = std::vector<std::vector<Index>>;
  std::vector<Cell> cells;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const std::array<Index, 8> v{node(i, j, k),         node(i + 1, j, k),
                                     node(i + 1, j + 1, k), node(i, j + 1, k),
                                     node(i, j, k + 1),     node(i + 1, j, k + 1),
                                     node(i + 1, j + 1, k + 1), node(i, j + 1, k + 1)};
        if (!tet) {
          cells.push_back({{v[0], v[3], v[2], v[1]}, {v[4], v[5], v[6], v[7]},
                           {v[0], v[1], v[5], v[4]}, {v[1], v[2], v[6], v[5]},
                           {v[2], v[3], v[7], v[6]}, {v[3], v[0], v[4], v[7]}});
        } else {
          constexpr int ring[6] = {1, 2, 3, 7, 4, 5};
          for (int t = 0; t < 6; ++t) {
            const Index a = v[0], b = v[ring[t]], c = v[ring[(t + 1) % 6]], d = v[6];
            cells.push_back({{a, b, c}, {a, d, b}, {b, d, c}, {c, d, a}});
          }
        }
      }
  std::vector<Patch> patches;
  for (auto name : {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"}) patches.push_back({name, 0, 0});
  auto classify = [&](const std::vector<Index>& ns) -> Index {
    for (Index side = 0; side < 6; ++side) {
      bool all = true;
      for (auto n : ns) {
        const auto ii = n % (nx + 1), jj = (n / (nx + 1)) % (ny + 1), kk = n / ((nx + 1) * (ny + 1));
        all = all && (side == 0 ? ii == 0 : side == 1 ? ii == nx : side == 2 ? jj == 0
                      : side == 3 ? jj == ny : side == 4 ? kk == 0 : kk == nz);
      }
      if (all) return side;
    }
    throw Error("box boundary classification failed");
  };
  std::map<std::vector<Index>, std::size_t> lookup;
  std::vector<FaceRecord> faces;
  for (Index c = 0; c < Index(cells.size()); ++c) {
    std::vector<Index> vertices;
    for (const auto& ns : cells[c]) vertices.insert(vertices.end(), ns.begin(), ns.end());
    std::sort(vertices.begin(), vertices.end());
    vertices.erase(std::unique(vertices.begin(), vertices.end()), vertices.end());
    const Vec3 origin = points[vertices[0]];
    Vec3 center{};
    for (auto n : vertices) center = center + (points[n] - origin);
    center = origin + center / static_cast<double>(vertices.size());
    for (auto ns : cells[c]) {
      auto key = ns;
      std::sort(key.begin(), key.end());
      auto [it, added] = lookup.emplace(key, faces.size());
      if (added) {
        Vec3 fc{}, av{};
        const Vec3 r = points[ns[0]];
        for (auto n : ns) fc = fc + (points[n] - r);
        fc = r + fc / static_cast<double>(ns.size());
        for (std::size_t q = 0; q < ns.size(); ++q)
          av = av + cross(points[ns[q]] - fc, points[ns[(q + 1) % ns.size()]] - fc);
        if (dot(av, fc - center) < 0) std::reverse(ns.begin(), ns.end());
        faces.push_back({std::move(ns), c, -1, -1});
      } else {
        auto& f = faces[it->second];
        ORACLE_REQUIRE(f.neighbour < 0, "more than two cells share a face");
        f.neighbour = c;
      }
    }
  }
  for (auto& f : faces)
    if (f.neighbour < 0) f.patch = classify(f.nodes);
  auto m = pack(std::move(points), Index(cells.size()), std::move(faces), std::move(patches));
  validate(m);
  return m;
}

// mesh/topology.cpp:108-129
CellFaces build_cell_faces(const PolyMesh& m) {
  CellFaces a;
  a.offsets.assign(std::size_t(m.n_cells) + 1, 0);
  for (auto c : m.owner) ++a.offsets[c + 1];
  for (auto c : m.neighbour) ++a.offsets[c + 1];
  for (Index c = 0; c < m.n_cells; ++c) a.offsets[c + 1] += a.offsets[c];
  a.faces.resize(a.offsets.back());
  a.signs.resize(a.faces.size());
  auto cursor = a.offsets;
  for (Index f = 0; f < m.n_faces(); ++f) {
    Index i = cursor[m.owner[f]]++;
    a.faces[i] = f;
    a.signs[i] = 1;
    if (f < m.n_internal_faces()) {
      i = cursor[m.neighbour[f]]++;
      a.faces[i] = f;
      a.signs[i] = -1;
    }
  }
  return a;
}

// mesh/topology.cpp:130-147
CellNeighbours build_cell_neighbours(const PolyMesh& m) {
  const auto faces = build_cell_faces(m);
  CellNeighbours a;
  a.offsets.push_back(0);
  std::vector<Index> row;
  for (Index c = 0; c < m.n_cells; ++c) {
    row.clear();
    for (auto i = faces.offsets[c]; i < faces.offsets[c + 1]; ++i) {
      const auto f = faces.faces[i];
      if (f < m.n_internal_faces()) row.push_back(m.owner[f] == c ? m.neighbour[f] : m.owner[f]);
    }
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    a.cells.insert(a.cells.end(), row.begin(), row.end());
    a.offsets.push_back(Index(a.cells.size()));
  }
  return a;
}

// mesh/topology.cpp:148-248
MeshPermutation renumber_cells_rcm(PolyMesh& m) {
  validate(m);
  const auto adj = build_cell_neighbours(m);
  MeshPermutation p;
  std::vector<bool> seen(m.n_cells);
  auto degree = [&](Index c) { return adj.offsets[c + 1] - adj.offsets[c]; };
  std::vector<Index> seeds(m.n_cells);
  std::iota(seeds.begin(), seeds.end(), 0);
  std::stable_sort(seeds.begin(), seeds.end(), [&](Index a, Index b) { return degree(a) < degree(b); });
  for (auto seed : seeds) {
    if (seen[seed]) continue;
    std::vector<Index> component;
    std::queue<Index> q;
    q.push(seed);
    seen[seed] = true;
    std::vector<Index> row;
    while (!q.empty()) {
      const auto c = q.front();
      q.pop();
      component.push_back(c);
      row.clear();
      for (auto i = adj.offsets[c]; i < adj.offsets[c + 1]; ++i)
        if (!seen[adj.cells[i]]) row.push_back(adj.cells[i]);
      std::stable_sort(row.begin(), row.end(), [&](Index a, Index b) { return degree(a) < degree(b); });
      for (auto n : row)
        if (!seen[n]) {
          seen[n] = true;
          q.push(n);
        }
    }
    p.cells_new_to_old.insert(p.cells_new_to_old.end(), component.rbegin(), component.rend());
  }
  std::vector<Index> old_to_new(m.n_cells);
  for (Index c = 0; c < m.n_cells; ++c) old_to_new[p.cells_new_to_old[c]] = c;
  const Index faces = m.n_faces(), internal = m.n_internal_faces();
  std::vector<Index> patch(faces, -1), owner(faces), neighbour(faces, -1);
  std::vector<bool> flipped(faces);
  for (Index i = 0; i < Index(m.patches.size()); ++i)
    for (Index f = m.patches[i].start; f < m.patches[i].start + m.patches[i].size; ++f) patch[f] = i;
  for (Index f = 0; f < faces; ++f) {
    owner[f] = old_to_new[m.owner[f]];
    if (f < internal) neighbour[f] = old_to_new[m.neighbour[f]];
    if (neighbour[f] >= 0 && owner[f] > neighbour[f]) {
      std::swap(owner[f], neighbour[f]);
      flipped[f] = true;
    }
  }
  p.faces_new_to_old.resize(faces);
  std::iota(p.faces_new_to_old.begin(), p.faces_new_to_old.end(), 0);
  std::stable_sort(p.faces_new_to_old.begin(), p.faces_new_to_old.end(), [&](Index a, Index b) {
    return std::tuple(neighbour[a] < 0, patch[a], owner[a], neighbour[a]) <
           std::tuple(neighbour[b] < 0, patch[b], owner[b], neighbour[b]);
  });
  PolyMesh r;
  r.points = std::move(m.points);
  r.n_cells = m.n_cells;
  r.owner.resize(faces);
  r.neighbour.resize(internal);
  r.face_offsets.resize(std::size_t(faces) + 1);
  r.face_offsets[0] = 0;
  r.face_nodes.resize(m.face_nodes.size());
  Index cursor = 0;
  for (Index i = 0; i < faces; ++i) {
    const auto f = p.faces_new_to_old[i];
    r.owner[i] = owner[f];
    if (i < internal) r.neighbour[i] = neighbour[f];
    const Index b = m.face_offsets[f], e = m.face_offsets[f + 1];
    if (flipped[f]) std::reverse_copy(m.face_nodes.begin() + b, m.face_nodes.begin() + e, r.face_nodes.begin() + cursor);
    else std::copy(m.face_nodes.begin() + b, m.face_nodes.begin() + e, r.face_nodes.begin() + cursor);
    cursor += e - b;
    r.face_offsets[i + 1] = cursor;
  }
  r.patches = std::move(m.patches);
  m = std::move(r);
  validate(m);
  return p;
}

// Seeded perturbation of a box mesh's points (no reference counterpart). Point
// v = (i, j, k) draws three uniforms u in [0, 1) from mt19937_64(seed), each
// (next() >> 11) * 2^-53, and moves by (2u - 1) * amplitude * h along x, y, z,
// except along an axis whose boundary plane it lies on, so the six patches
// stay planar. Any cell of non-positive volume, and any face whose area does
// not point out of its owner, into its neighbour, or along the centroid link
// (fv/operators.cpp:81), has its points redrawn from the continuing stream in
// increasing point order until none remains. Returns the redraw rounds.
int perturb_box(PolyMesh& m, int nx, int ny, int nz, Vec3 lo, Vec3 hi, double amplitude,
                std::uint64_t seed) {
  const double hx = (hi.x - lo.x) / nx, hy = (hi.y - lo.y) / ny, hz = (hi.z - lo.z) / nz;
  const std::vector<Vec3> base = m.points;
  std::mt19937_64 gen(seed);
  auto uniform = [&] { return double(gen() >> 11) * 0x1.0p-53; };
  auto draw = [&](Index v) {
    const Index i = v % (nx + 1), j = (v / (nx + 1)) % (ny + 1), k = v / ((nx + 1) * (ny + 1));
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
    const Geometry g = geometry(m, false);
    std::vector<char> bad(m.points.size(), 0);
    bool any = false;
    auto mark_face = [&](Index f) {
      for (Index q = m.face_offsets[f]; q < m.face_offsets[f + 1]; ++q) bad[m.face_nodes[q]] = 1;
      any = true;
    };
    std::vector<char> bad_cell(m.n_cells, 0);
    for (Index c = 0; c < m.n_cells; ++c)
      if (!(g.cell_volume[c] > 0) || !std::isfinite(g.cell_volume[c])) bad_cell[c] = 1;
    for (Index f = 0; f < m.n_faces(); ++f) {
      const Index o = m.owner[f];
      const bool internal = f < m.n_internal_faces();
      const Vec3 link = (internal ? g.cell_centroid[m.neighbour[f]] : g.face_centroid[f]) - g.cell_centroid[o];
      bool ok = dot(g.face_area[f], link) > 0 &&
                dot(g.face_area[f], g.face_centroid[f] - g.cell_centroid[o]) > 0;
      if (internal) ok = ok && dot(g.face_area[f], g.cell_centroid[m.neighbour[f]] - g.face_centroid[f]) > 0;
      if (!ok || bad_cell[o] || (internal && bad_cell[m.neighbour[f]])) mark_face(f);
    }
    if (!any) return round;
    ORACLE_REQUIRE(round < 64, "perturbation: no valid mesh after 64 redraw rounds");
    for (Index v = 0; v < Index(m.points.size()); ++v)
      if (bad[v]) draw(v);
  }
}

// fv/operators.cpp:43-91
Stencil make_stencil(const PolyMesh& m, const Geometry& g) {
  Stencil s;
  const auto cf = build_cell_faces(m);
  const auto cn = build_cell_neighbours(m);
  auto& p = s.pattern;
  p.n_rows = m.n_cells;
  s.lsq_inverse.assign(6 * std::size_t(m.n_cells), 0.0f);
  std::vector<Index> cols;
  for (Index c = 0; c < m.n_cells; ++c) {
    cols.assign(cn.cells.begin() + cn.offsets[c], cn.cells.begin() + cn.offsets[c + 1]);
    cols.insert(std::lower_bound(cols.begin(), cols.end(), c), c);
    for (auto j : cols) {
      if (j == c) p.diag.push_back(Index(p.col.size()));
      p.col.push_back(j);
    }
    p.row_ptr.push_back(Index(p.col.size()));
    std::array<double, 6> a{};
    for (auto j = cf.offsets[c]; j < cf.offsets[c + 1]; ++j) {
      const Index f = cf.faces[j];
      const Index n = f < m.n_internal_faces() ? (m.owner[f] == c ? m.neighbour[f] : m.owner[f])
                                               : -(f - m.n_internal_faces() + 1);
      s.adjacent.push_back(n);
      s.matrix_entry.push_back(n < 0 ? -1 : Index(std::lower_bound(p.col.begin() + p.row_ptr[c], p.col.end(), n) - p.col.begin()));
      const Vec3 d = (n < 0 ? g.face_centroid[f] : g.cell_centroid[n]) - g.cell_centroid[c];
      const double w = 1 / dot(d, d);
      const Vec3 row = d * w;
      s.lsq_x.push_back(float(row.x));
      s.lsq_y.push_back(float(row.y));
      s.lsq_z.push_back(float(row.z));
      a[0] += w * d.x * d.x; a[1] += w * d.x * d.y; a[2] += w * d.x * d.z;
      a[3] += w * d.y * d.y; a[4] += w * d.y * d.z; a[5] += w * d.z * d.z;
    }
    const double det = a[0] * (a[3] * a[5] - a[4] * a[4]) - a[1] * (a[1] * a[5] - a[4] * a[2]) +
                       a[2] * (a[1] * a[4] - a[3] * a[2]);
    ORACLE_REQUIRE(det > 1e-14, "FV LSQ stencil is rank deficient");
    const std::array<double, 6> inv{(a[3] * a[5] - a[4] * a[4]) / det, (a[2] * a[4] - a[1] * a[5]) / det,
                                    (a[1] * a[4] - a[2] * a[3]) / det, (a[0] * a[5] - a[2] * a[2]) / det,
                                    (a[1] * a[2] - a[0] * a[4]) / det, (a[0] * a[3] - a[1] * a[1]) / det};
    for (int k = 0; k < 6; ++k) s.lsq_inverse[c + std::size_t(k) * m.n_cells] = float(inv[k]);
  }
  for (Index f = 0; f < m.n_faces(); ++f) {
    const Vec3 d = (f < m.n_internal_faces() ? g.cell_centroid[m.neighbour[f]] : g.face_centroid[f]) -
                   g.cell_centroid[m.owner[f]];
    const Vec3 area = g.face_area[f];
    const double projection = dot(area, d);
    ORACLE_REQUIRE(projection > 0, "FV requires positive face/cell projection");
    const double coeff = dot(area, area) / projection;
    s.diffusion.push_back(float(coeff));
    const Vec3 no = area - d * coeff;
    s.nonorth_x.push_back(float(no.x));
    s.nonorth_y.push_back(float(no.y));
    s.nonorth_z.push_back(float(no.z));
  }
  p.val.assign(p.col.size(), 0.0);
  return s;
}

// fv/steady_solver.cpp:59-75
Index locate_cell(const PolyMesh& m, const Geometry& g, Vec3 point) {
  const auto cf = build_cell_faces(m);
  Index nearest = 0;
  double best = INFINITY;
  for (Index c = 0; c < m.n_cells; ++c) {
    bool inside = true;
    for (Index j = cf.offsets[c]; j < cf.offsets[c + 1] && inside; ++j) {
      const auto f = cf.faces[j];
      const Vec3 s = g.face_area[f] * double(cf.signs[j]);
      inside = dot(s, point - g.face_centroid[f]) <= 1e-12 * norm(s) * std::cbrt(g.cell_volume[c]);
    }
    if (inside) return c;
    const Vec3 d = point - g.cell_centroid[c];
    if (dot(d, d) < best) {
      best = dot(d, d);
      nearest = c;
    }
  }
  return nearest;
}

// kernels/fv_flow.kern.h:164-183
void fv_pressure_assemble(int n, const int* offsets, const int* faces, const float* sign,
                          const int* adjacent, const int* entry, const int* rows, const int* diag,
                          const float* coefficient, const float* flux, float* values, float* rhs,
                          int reference) {
  for (int i = 0; i < n; ++i) {
    UpperSum ap{0.0f, 0.0f};
    float b = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j)
      if (adjacent[j] >= 0) values[entry[j]] = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      const int f = faces[j], nb = adjacent[j];
      const float d = coefficient[f];
      b -= sign[j] * flux[f];
      if (nb < 0 || nb == reference) ap = kupper_add(ap, d);
      if (nb >= 0 && nb != reference && i != reference) values[entry[j]] -= d;
    }
    for (int j = rows[i]; j < rows[i + 1]; ++j)
      if (j != diag[i]) ap = kupper_add(ap, -values[j]);
    values[diag[i]] = i == reference ? 1.0f : kupper_value(ap);
    rhs[i] = i == reference ? 0.0f : b;
  }
}

}  // namespace oracle
