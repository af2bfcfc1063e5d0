// Voronoi tessellation of seeded points in a box, flattened to per-cell
// face tables (host C++, no GPU code).
//
// The reference runs IMC on MadVoro's Voronoi meshes through the Grid
// concept (GetCellFaces, GetFaceNeighbors, Normal, FaceCM, GetArea,
// GetVolume, GetCellCM, GetMeshPoint, GetPointsInFace / GetFacePoints;
// particle/Particle.hpp:188-279, radiation/RadiationIMC.hpp:1364-1404,
// utils/RandomInCell.hpp:10-67, utils/RandomOnFace.hpp:14-66).  MadVoro is
// not available, so the cells are built here, each on its own: the box is
// clipped by the bisector half-space of every generator close enough to
// matter, nearest first, found through a uniform grid of buckets searched in
// shells of growing radius.  A shell at index distance s holds points at
// least (s - 1) bucket widths away; once every vertex of the cell is within
// half of that of its generator, no farther bisector can cut it.
//
// Output (struct Mesh): cells renumbered in Z order of their generators (so
// that cells close in space are close in memory); per cell its volume,
// centroid and generator; per cell a contiguous range of faces ("half
// faces": a face shared by two cells appears once in each, with opposite
// normals), each with the neighbour cell (or -1 - box side, sides x-, x+,
// y-, y+, z-, z+), the outward unit normal, the area, the centroid and the
// polygon's vertices in counter-clockwise order seen from outside.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace storm {
namespace voronoi {

struct V3 {
  double x = 0, y = 0, z = 0;
  V3() = default;
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  V3 operator+(const V3 &o) const { return V3(x + o.x, y + o.y, z + o.z); }
  V3 operator-(const V3 &o) const { return V3(x - o.x, y - o.y, z - o.z); }
  V3 operator*(double s) const { return V3(x * s, y * s, z * s); }
};
inline double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3 &a, const V3 &b) {
  return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double norm(const V3 &a) { return std::sqrt(dot(a, a)); }

struct Mesh {
  uint32_t ncells = 0;
  double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  std::vector<double> generator;      // 3 per cell (GetMeshPoint)
  std::vector<double> volume;         // per cell
  std::vector<double> centroid;       // 3 per cell (GetCellCM)
  std::vector<uint32_t> face_start;   // ncells + 1
  std::vector<int32_t> face_neighbor; // per face: cell, or -1 - box side
  std::vector<double> face_normal;    // 3 per face, outward unit
  std::vector<double> face_area;      // per face
  std::vector<double> face_centroid;  // 3 per face
  std::vector<uint32_t> face_vstart;  // faces + 1
  std::vector<double> vertex;         // 3 per polygon vertex
  uint32_t nfaces() const { return static_cast<uint32_t>(face_neighbor.size()); }
};

namespace detail {

struct Face {
  V3 n;          // outward unit normal
  double d;      // plane n . x = d
  int32_t nbr;   // neighbour cell or -1 - box side
  std::vector<V3> poly;
};

// Keep the part of the convex polyhedron `faces` with n . x <= d; the cut
// becomes a new face towards `nbr`.  eps: plane-distance tolerance.
inline void clip(std::vector<Face> &faces, const V3 &n, double d, int32_t nbr, double eps) {
  std::vector<V3> cut;
  std::vector<Face> out;
  out.reserve(faces.size() + 1);
  bool any_out = false;
  for (const Face &f : faces)
    for (const V3 &v : f.poly)
      if (dot(n, v) - d > eps) any_out = true;
  if (!any_out) return;
  for (Face &f : faces) {
    const size_t m = f.poly.size();
    std::vector<V3> np;
    np.reserve(m + 2);
    for (size_t i = 0; i < m; ++i) {
      const V3 &a = f.poly[i], &b = f.poly[(i + 1) % m];
      const double sa = dot(n, a) - d, sb = dot(n, b) - d;
      const bool ina = sa <= eps;
      if (ina) {
        np.push_back(a);
        if (sa >= -eps) cut.push_back(a);
      }
      if ((sa < -eps && sb > eps) || (sa > eps && sb < -eps)) {
        const double t = sa / (sa - sb);
        const V3 p = a + (b - a) * t;
        np.push_back(p);
        cut.push_back(p);
      }
    }
    // drop repeated vertices
    std::vector<V3> q;
    for (const V3 &v : np)
      if (q.empty() || norm(v - q.back()) > eps) q.push_back(v);
    while (q.size() > 1 && norm(q.front() - q.back()) <= eps) q.pop_back();
    if (q.size() >= 3) {
      f.poly = std::move(q);
      out.push_back(std::move(f));
    }
  }
  // the new face: the cut points ordered counter-clockwise about n
  std::vector<V3> u;
  for (const V3 &p : cut) {
    bool dup = false;
    for (const V3 &o : u)
      if (norm(p - o) <= 4 * eps) {
        dup = true;
        break;
      }
    if (!dup) u.push_back(p);
  }
  if (u.size() >= 3) {
    V3 c(0, 0, 0);
    for (const V3 &p : u) c = c + p;
    c = c * (1.0 / static_cast<double>(u.size()));
    V3 e1 = u[0] - c;
    e1 = e1 * (1.0 / norm(e1));
    const V3 e2 = cross(n, e1);
    std::vector<std::pair<double, V3>> ang;
    for (const V3 &p : u) ang.push_back({std::atan2(dot(p - c, e2), dot(p - c, e1)), p});
    std::sort(ang.begin(), ang.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    Face nf;
    nf.n = n;
    nf.d = d;
    nf.nbr = nbr;
    for (const auto &a : ang) nf.poly.push_back(a.second);
    out.push_back(std::move(nf));
  }
  faces = std::move(out);
}

inline uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
  v &= 0x1fffffull;
  v = (v | v << 32) & 0x1f00000000ffffull;
  v = (v | v << 16) & 0x1f0000ff0000ffull;
  v = (v | v << 8) & 0x100f00f00f00f00full;
  v = (v | v << 4) & 0x10c30c30c30c30c3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}

}  // namespace detail

// Builds the tessellation of `n` generators (x, y, z interleaved) inside the
// box [lo, hi].  Generators must lie strictly inside the box and be
// distinct.
inline Mesh build(const double *points, uint32_t n, const double lo[3], const double hi[3]) {
  using detail::Face;
  if (n == 0) throw std::invalid_argument("voronoi: no generators");
  Mesh M;
  for (int a = 0; a < 3; ++a) {
    M.lo[a] = lo[a];
    M.hi[a] = hi[a];
    if (!(hi[a] > lo[a])) throw std::invalid_argument("voronoi: empty box");
  }
  const double L[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  const double scale = std::max(L[0], std::max(L[1], L[2]));
  const double eps = 1e-12 * scale;
  // Z order of the generators (21 bits per axis)
  std::vector<std::pair<uint64_t, uint32_t>> key(n);
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t k = 0;
    for (int a = 0; a < 3; ++a) {
      const double u = (points[3 * i + a] - lo[a]) / L[a];
      if (!(u > 0.0 && u < 1.0)) throw std::invalid_argument("voronoi: generator outside the box");
      const uint64_t q = static_cast<uint64_t>(u * 2097152.0);
      k |= detail::spread3(q < 2097151ull ? q : 2097151ull) << a;
    }
    key[i] = {k, i};
  }
  std::sort(key.begin(), key.end());
  std::vector<V3> P(n);
  M.generator.resize(3 * static_cast<size_t>(n));
  for (uint32_t c = 0; c < n; ++c) {
    const uint32_t i = key[c].second;
    P[c] = V3(points[3 * i], points[3 * i + 1], points[3 * i + 2]);
    M.generator[3 * c] = P[c].x;
    M.generator[3 * c + 1] = P[c].y;
    M.generator[3 * c + 2] = P[c].z;
  }
  // buckets: about one generator each
  const int G = std::max(1, static_cast<int>(std::cbrt(static_cast<double>(n))));
  const double h[3] = {L[0] / G, L[1] / G, L[2] / G};
  const double hmin = std::min(h[0], std::min(h[1], h[2]));
  std::vector<uint32_t> bstart(static_cast<size_t>(G) * G * G + 1, 0), bitems(n);
  auto bucket_of = [&](const V3 &p, int b[3]) {
    const double v[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) b[a] = std::min(G - 1, std::max(0, static_cast<int>((v[a] - lo[a]) / h[a])));
  };
  auto bidx = [&](int x, int y, int z) { return (static_cast<size_t>(x) * G + y) * G + z; };
  for (uint32_t c = 0; c < n; ++c) {
    int b[3];
    bucket_of(P[c], b);
    ++bstart[bidx(b[0], b[1], b[2]) + 1];
  }
  for (size_t q = 1; q < bstart.size(); ++q) bstart[q] += bstart[q - 1];
  {
    std::vector<uint32_t> fill(bstart.begin(), bstart.end() - 1);
    for (uint32_t c = 0; c < n; ++c) {
      int b[3];
      bucket_of(P[c], b);
      bitems[fill[bidx(b[0], b[1], b[2])]++] = c;
    }
  }
  M.ncells = n;
  M.volume.resize(n);
  M.centroid.resize(3 * static_cast<size_t>(n));
  // cells are independent: built in parallel into per-cell records, then
  // flattened in cell order (the result does not depend on the thread count)
  struct CellOut {
    double vol;
    V3 cm;
    std::vector<int32_t> nbr;
    std::vector<double> nrm, area, fc, vert;
    std::vector<uint32_t> nv;
  };
  std::vector<CellOut> cells(n);
  auto build_cell = [&](uint32_t c, std::vector<Face> &faces, std::vector<std::pair<double, uint32_t>> &cand) {
    const V3 p = P[c];
    // the box, faces x-, x+, y-, y+, z-, z+ (outward normals, counter-clockwise)
    faces.clear();
    {
      const V3 v000(lo[0], lo[1], lo[2]), v100(hi[0], lo[1], lo[2]), v010(lo[0], hi[1], lo[2]),
          v110(hi[0], hi[1], lo[2]), v001(lo[0], lo[1], hi[2]), v101(hi[0], lo[1], hi[2]), v011(lo[0], hi[1], hi[2]),
          v111(hi[0], hi[1], hi[2]);
      faces.push_back(Face{V3(-1, 0, 0), -lo[0], -1, {v000, v001, v011, v010}});
      faces.push_back(Face{V3(1, 0, 0), hi[0], -2, {v100, v110, v111, v101}});
      faces.push_back(Face{V3(0, -1, 0), -lo[1], -3, {v000, v100, v101, v001}});
      faces.push_back(Face{V3(0, 1, 0), hi[1], -4, {v010, v011, v111, v110}});
      faces.push_back(Face{V3(0, 0, -1), -lo[2], -5, {v000, v010, v110, v100}});
      faces.push_back(Face{V3(0, 0, 1), hi[2], -6, {v001, v101, v111, v011}});
    }
    int b[3];
    bucket_of(p, b);
    auto rmax = [&]() {
      double r = 0.0;
      for (const Face &f : faces)
        for (const V3 &v : f.poly) r = std::max(r, norm(v - p));
      return r;
    };
    double R = rmax();
    for (int s = 0;; ++s) {
      // generators of shell s, nearest first
      cand.clear();
      for (int x = b[0] - s; x <= b[0] + s; ++x)
        for (int y = b[1] - s; y <= b[1] + s; ++y)
          for (int z = b[2] - s; z <= b[2] + s; ++z) {
            if (std::max(std::abs(x - b[0]), std::max(std::abs(y - b[1]), std::abs(z - b[2]))) != s) continue;
            if (x < 0 || y < 0 || z < 0 || x >= G || y >= G || z >= G) continue;
            const size_t q = bidx(x, y, z);
            for (uint32_t t = bstart[q]; t < bstart[q + 1]; ++t) {
              const uint32_t o = bitems[t];
              if (o != c) cand.push_back({norm(P[o] - p), o});
            }
          }
      std::sort(cand.begin(), cand.end());
      for (const auto &co : cand) {
        if (co.first >= 2.0 * R) break;  // its bisector lies beyond every vertex
        const V3 q = P[co.second];
        V3 nn = q - p;
        const double len = norm(nn);
        if (!(len > 0.0)) throw std::invalid_argument("voronoi: repeated generator");
        nn = nn * (1.0 / len);
        detail::clip(faces, nn, dot(nn, (p + q) * 0.5), static_cast<int32_t>(co.second), eps);
        R = rmax();
      }
      // every generator not yet seen is at least s bucket widths away
      const bool covered = b[0] - s <= 0 && b[1] - s <= 0 && b[2] - s <= 0 && b[0] + s >= G - 1 &&
                           b[1] + s >= G - 1 && b[2] + s >= G - 1;
      if (covered || 2.0 * R <= static_cast<double>(s) * hmin) break;
    }
    // volume, centroid (tetrahedra from the generator), faces
    CellOut &o = cells[c];
    double vol = 0.0;
    V3 cm(0, 0, 0);
    for (const Face &f : faces) {
      V3 fc(0, 0, 0);
      double area = 0.0;
      const V3 &a = f.poly[0];
      for (size_t i = 1; i + 1 < f.poly.size(); ++i) {
        const V3 &bb = f.poly[i], &cc = f.poly[i + 1];
        const double ta = 0.5 * dot(cross(bb - a, cc - a), f.n);
        area += ta;
        fc = fc + (a + bb + cc) * (ta / 3.0);
        const double tv = dot(cross(a - p, bb - p), cc - p) / 6.0;
        vol += tv;
        cm = cm + (p + a + bb + cc) * (tv / 4.0);
      }
      fc = fc * (1.0 / area);
      o.nbr.push_back(f.nbr);
      o.nrm.insert(o.nrm.end(), {f.n.x, f.n.y, f.n.z});
      o.area.push_back(area);
      o.fc.insert(o.fc.end(), {fc.x, fc.y, fc.z});
      for (const V3 &v : f.poly) o.vert.insert(o.vert.end(), {v.x, v.y, v.z});
      o.nv.push_back(static_cast<uint32_t>(f.poly.size()));
    }
    o.vol = vol;
    o.cm = cm * (1.0 / vol);
  };
  {
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned T = n < 4096 ? 1u : std::max(1u, std::min(16u, hw ? hw : 1u));
    std::vector<std::thread> pool;
    std::vector<std::string> errs(T);
    for (unsigned t = 0; t < T; ++t)
      pool.emplace_back([&, t] {
        std::vector<Face> faces;
        std::vector<std::pair<double, uint32_t>> cand;
        try {
          for (uint32_t c = t; c < n; c += T) build_cell(c, faces, cand);
        } catch (const std::exception &e) {
          errs[t] = e.what();
        }
      });
    for (auto &th : pool) th.join();
    for (const std::string &e : errs)
      if (!e.empty()) throw std::invalid_argument(e);
  }
  // Near-degenerate vertices (four or more generators almost on a sphere)
  // can leave a sliver face on one side of a pair only; such a face is
  // dropped when its area is negligible (below 1e-12 of the box's scale
  // squared), and any larger unmatched face is an error.
  {
    auto has_face_to = [&](uint32_t a, int32_t b) {
      for (int32_t x : cells[a].nbr)
        if (x == b) return true;
      return false;
    };
    const double tiny = 1e-12 * scale * scale;
    for (uint32_t c = 0; c < n; ++c) {
      CellOut &o = cells[c];
      for (size_t q = 0; q < o.nbr.size();) {
        const int32_t nb = o.nbr[q];
        if (nb < 0 || has_face_to(static_cast<uint32_t>(nb), static_cast<int32_t>(c))) {
          ++q;
          continue;
        }
        if (!(o.area[q] < tiny)) throw std::runtime_error("voronoi: a face is missing its neighbour's side");
        size_t v0 = 0;
        for (size_t k = 0; k < q; ++k) v0 += o.nv[k];
        o.vert.erase(o.vert.begin() + 3 * v0, o.vert.begin() + 3 * (v0 + o.nv[q]));
        o.nbr.erase(o.nbr.begin() + q);
        o.nrm.erase(o.nrm.begin() + 3 * q, o.nrm.begin() + 3 * q + 3);
        o.area.erase(o.area.begin() + q);
        o.fc.erase(o.fc.begin() + 3 * q, o.fc.begin() + 3 * q + 3);
        o.nv.erase(o.nv.begin() + q);
      }
    }
  }
  M.face_start.assign(1, 0);
  M.face_vstart.assign(1, 0);
  for (uint32_t c = 0; c < n; ++c) {
    CellOut &o = cells[c];
    M.volume[c] = o.vol;
    M.centroid[3 * c] = o.cm.x;
    M.centroid[3 * c + 1] = o.cm.y;
    M.centroid[3 * c + 2] = o.cm.z;
    M.face_neighbor.insert(M.face_neighbor.end(), o.nbr.begin(), o.nbr.end());
    M.face_normal.insert(M.face_normal.end(), o.nrm.begin(), o.nrm.end());
    M.face_area.insert(M.face_area.end(), o.area.begin(), o.area.end());
    M.face_centroid.insert(M.face_centroid.end(), o.fc.begin(), o.fc.end());
    M.vertex.insert(M.vertex.end(), o.vert.begin(), o.vert.end());
    for (uint32_t k : o.nv) M.face_vstart.push_back(M.face_vstart.back() + k);
    M.face_start.push_back(static_cast<uint32_t>(M.face_neighbor.size()));
    o = CellOut();  // release the record
  }
  return M;
}

}  // namespace voronoi
}  // namespace storm
