// Host construction of the device tables of a polyhedral mesh
// (storm_unstructured_mesh -> UFace, tetrahedra, source triangles), shared
// by the engine and the CPU replay so that both track against the same
// numbers.
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/storm_problem.h"
#include "unstructured.h"

namespace storm {
namespace umesh {

template <typename Real>
struct Tables {
  uint32_t ncells = 0;
  std::vector<uint32_t> face_start;  // ncells + 1
  std::vector<UFace<Real>> face;
  std::vector<Real> face_centroid;   // 3 per face, in the cell's frame
  std::vector<double> face_geo;      // 2 per face: area, distance of the plane from the mesh point
  std::vector<uint32_t> tet_start;   // ncells + 1
  std::vector<UTet<Real>> tets;
  std::vector<double> chord;         // per cell: mean chord 4 V / S
  std::vector<double> surface;       // per cell: S
  // boundary sources: faces of side x- (and x+), in cell order
  std::vector<USource> src;
  std::vector<UTri<Real>> tris;
  std::vector<double> src_area;
  std::vector<uint32_t> src_instance;
};

inline void check_mesh(const storm_unstructured_mesh &m) {
  if (m.ncells == 0) throw std::invalid_argument("unstructured mesh: no cells");
  if (!m.volume || !m.mesh_point || !m.face_start || !m.face_neighbor || !m.face_normal || !m.face_centroid ||
      !m.face_area || !m.face_vertex_start || !m.vertices)
    throw std::invalid_argument("unstructured mesh: a table pointer is null");
  if (m.face_start[0] != 0 || m.face_start[m.ncells] != m.nfaces)
    throw std::invalid_argument("unstructured mesh: face_start must run from 0 to nfaces");
  for (uint32_t c = 0; c < m.ncells; ++c) {
    if (m.face_start[c + 1] < m.face_start[c] + 4)
      throw std::invalid_argument("unstructured mesh: a cell has fewer than 4 faces");
    if (!(m.volume[c] > 0.0)) throw std::invalid_argument("unstructured mesh: a cell has no volume");
  }
  for (uint32_t f = 0; f < m.nfaces; ++f) {
    const int32_t nb = m.face_neighbor[f];
    if (nb >= static_cast<int64_t>(m.ncells) || nb < -6)
      throw std::invalid_argument("unstructured mesh: face neighbour out of range");
    if (m.face_vertex_start[f + 1] < m.face_vertex_start[f] + 3)
      throw std::invalid_argument("unstructured mesh: a face has fewer than 3 vertices");
  }
}

// src_xlo / src_xhi: build the source faces of those sides; npc(face's
// cell) gives each face's packets per step (the instance's Npercell).
template <typename Real, typename NpcOf>
Tables<Real> build(const storm_unstructured_mesh &m, bool src_xlo, bool src_xhi, NpcOf npc_of) {
  check_mesh(m);
  Tables<Real> T;
  const uint32_t N = m.ncells;
  T.ncells = N;
  T.face_start.assign(m.face_start, m.face_start + N + 1);
  T.face.resize(m.nfaces);
  T.face_centroid.resize(3 * static_cast<size_t>(m.nfaces));
  T.face_geo.resize(2 * static_cast<size_t>(m.nfaces));
  T.chord.resize(N);
  T.surface.resize(N);
  T.tet_start.assign(1, 0);
  uint32_t first = 0;
  for (uint32_t c = 0; c < N; ++c) {
    const double *g = m.mesh_point + 3 * static_cast<size_t>(c);
    double S = 0.0;
    std::vector<double> vols;
    const size_t t0 = T.tets.size();
    for (uint32_t f = m.face_start[c]; f < m.face_start[c + 1]; ++f) {
      const double *n = m.face_normal + 3 * static_cast<size_t>(f);
      const double *fc = m.face_centroid + 3 * static_cast<size_t>(f);
      const double rel[3] = {fc[0] - g[0], fc[1] - g[1], fc[2] - g[2]};
      const int32_t nb = m.face_neighbor[f];
      UFace<Real> &F = T.face[f];
      for (int a = 0; a < 3; ++a) F.n[a] = static_cast<Real>(n[a]);
      F.d = static_cast<Real>(n[0] * rel[0] + n[1] * rel[1] + n[2] * rel[2]);
      F.nbr = nb;
      for (int a = 0; a < 3; ++a) {
        F.sh[a] = nb >= 0 ? static_cast<Real>(g[a] - m.mesh_point[3 * static_cast<size_t>(nb) + a]) : Real(0);
        T.face_centroid[3 * static_cast<size_t>(f) + a] = static_cast<Real>(rel[a]);
      }
      // DDMC distance: mesh point to the face plane along the normal, else
      // half the distance to the neighbour's (RadiationIMC.hpp:1379-1384)
      double dist = std::fabs(n[0] * rel[0] + n[1] * rel[1] + n[2] * rel[2]);
      if (!(dist > 0.0) && nb >= 0) {
        const double *h = m.mesh_point + 3 * static_cast<size_t>(nb);
        dist = 0.5 * std::fabs(n[0] * (h[0] - g[0]) + n[1] * (h[1] - g[1]) + n[2] * (h[2] - g[2]));
      }
      T.face_geo[2 * static_cast<size_t>(f)] = m.face_area[f];
      T.face_geo[2 * static_cast<size_t>(f) + 1] = dist;
      S += m.face_area[f];
      // tetrahedra of the face's fan about the mesh point (RandomInCell.hpp:24-42)
      const uint32_t v0 = m.face_vertex_start[f], v1 = m.face_vertex_start[f + 1];
      auto vr = [&](uint32_t v, int a) { return m.vertices[3 * static_cast<size_t>(v) + a] - g[a]; };
      for (uint32_t i = v0 + 1; i + 1 < v1; ++i) {
        const double A[3] = {vr(v0, 0), vr(v0, 1), vr(v0, 2)};
        const double B[3] = {vr(i, 0), vr(i, 1), vr(i, 2)};
        const double C[3] = {vr(i + 1, 0), vr(i + 1, 1), vr(i + 1, 2)};
        const double det = A[0] * (B[1] * C[2] - B[2] * C[1]) - A[1] * (B[0] * C[2] - B[2] * C[0]) +
                           A[2] * (B[0] * C[1] - B[1] * C[0]);
        UTet<Real> t{};
        for (int a = 0; a < 3; ++a) {
          t.a[a] = static_cast<Real>(A[a]);
          t.b[a] = static_cast<Real>(B[a]);
          t.c[a] = static_cast<Real>(C[a]);
        }
        vols.push_back(std::fabs(det) / 6.0);
        T.tets.push_back(t);
      }
      // source faces of side x- / x+
      if ((nb == -1 - STORM_XLO && src_xlo) || (nb == -1 - STORM_XHI && src_xhi)) {
        USource s{};
        s.cell = c;
        s.tri_start = static_cast<uint32_t>(T.tris.size());
        s.xhi = nb == -1 - STORM_XHI ? 1 : 0;
        std::vector<double> areas;
        for (uint32_t i = v0 + 1; i + 1 < v1; ++i) {
          double e1[3], e2[3];
          for (int a = 0; a < 3; ++a) {
            e1[a] = vr(i, a) - vr(v0, a);
            e2[a] = vr(i + 1, a) - vr(v0, a);
          }
          const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                       cz = e1[0] * e2[1] - e1[1] * e2[0];
          areas.push_back(0.5 * std::sqrt(cx * cx + cy * cy + cz * cz));
          UTri<Real> tr{};
          for (int a = 0; a < 3; ++a) {
            tr.a[a] = static_cast<Real>(vr(v0, a));
            tr.b[a] = static_cast<Real>(vr(i, a));
            tr.c[a] = static_cast<Real>(vr(i + 1, a));
          }
          T.tris.push_back(tr);
        }
        double tot = 0.0;
        for (double a : areas) tot += a;
        double run = 0.0;
        for (size_t q = 0; q < areas.size(); ++q) {
          run += areas[q];
          T.tris[s.tri_start + q].cum = q + 1 == areas.size() ? Real(1) : static_cast<Real>(run / tot);
        }
        s.tri_count = static_cast<uint32_t>(areas.size());
        const uint64_t npc = npc_of(c);
        if (npc > 0) {
          s.first = first;
          s.count = static_cast<uint32_t>(npc);
          if (static_cast<uint64_t>(first) + npc >= (1ull << 32)) throw std::invalid_argument("too many source packets");
          first += static_cast<uint32_t>(npc);
          T.src.push_back(s);
          T.src_area.push_back(m.face_area[f]);
          T.src_instance.push_back(m.cell_instance ? m.cell_instance[c] : 0u);
        }
      }
    }
    double vt = 0.0;
    for (double v : vols) vt += v;
    double run = 0.0;
    for (size_t q = 0; q < vols.size(); ++q) {
      run += vols[q];
      T.tets[t0 + q].cum = q + 1 == vols.size() ? Real(1) : static_cast<Real>(run / vt);
    }
    T.tet_start.push_back(static_cast<uint32_t>(T.tets.size()));
    T.surface[c] = S;
    T.chord[c] = 4.0 * m.volume[c] / S;
  }
  return T;
}

}  // namespace umesh
}  // namespace storm
