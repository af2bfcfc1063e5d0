// FP64 CPU oracle: a serial restatement of the reference's gray IMC hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (storm_amd/, include/)
// includes, links or calls this file; only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg use it, as the checker.
//
// It restates, in double precision and on one host thread:
//   manager    manager/MonteCarloManagerSerial.hpp:385-469 (step), 266-383 (HandleAll)
//   physics    physics/MonteCarloPhysics.hpp:60-99, particle/Particle.hpp:188-279
//   IMC        radiation/RadiationIMC.hpp:1618-1694 (preStep), 1700-1881 (step),
//              1887-2029 (postStep), 2035-2553 (generateParticles), 771-866
//              (negative-energy guard, material sync), 956-1612 (RW and DDMC)
//   RW tables  radiation/RandomWalk.hpp:20-242
//   boundary   boundary/BoundaryCondition.hpp:56-108, boundary/SideTemperature.hpp:84-163,
//              boundary/Vacuum.hpp:18-22
//   comb       population/CombPopulationControl.hpp:79-253
//   samplers   utils/RandomInCell.hpp:10-67, utils/RandomOnFace.hpp:14-66,
//              radiation/RadiationOpacityModel.hpp:139-150
// on a Cartesian mesh that follows the MadCart Grid concept the reference
// calls (cell (i*ny+j)*nz+k, six faces per cell, outside "mirror" points).
//
// RNG: std::mt19937_64 + std::uniform_real_distribution<double> exactly like
// the reference (radiation/RadiationIMC.hpp:519-520), but every stream is
// seeded from the problem seed (the reference seeds the boundary, face and
// comb streams with fixed constants, SideTemperature.hpp:119,
// RandomOnFace.hpp:17, CombPopulationControl.hpp:94); results are compared
// statistically, never bit by bit.
//
// Parity status of each part is recorded in tests/test_oracle_*.py (what pins
// it) -- see the function comments below.
#pragma once

#include <algorithm>
#include <array>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../include/storm_problem.h"
#include "../storm_amd/csrc/common/constants.h"
#include "../storm_amd/csrc/common/planck.h"

namespace storm_oracle {

namespace units = storm::units;

constexpr double kEpsilonParticle = 1e-12;     // particle/Particle.hpp:17
constexpr double kMonteCarloEpsilon = 1e-8;    // manager/MonteCarloConfig.hpp:11

struct OracleError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ------------------------------------------------------------------ points
struct Vec3 {
  double x = 0, y = 0, z = 0;
  Vec3() = default;
  Vec3(double v) : x(v), y(v), z(v) {}
  Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  double &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline Vec3 operator+(const Vec3 &a, const Vec3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(const Vec3 &a, const Vec3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(const Vec3 &v, double s) { return {v.x * s, v.y * s, v.z * s}; }
inline Vec3 operator*(double s, const Vec3 &v) { return {v.x * s, v.y * s, v.z * s}; }
inline Vec3 operator/(const Vec3 &v, double s) { return {v.x / s, v.y / s, v.z / s}; }
inline Vec3 &operator+=(Vec3 &a, const Vec3 &b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline Vec3 &operator-=(Vec3 &a, const Vec3 &b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
inline Vec3 &operator*=(Vec3 &v, double s) { v.x *= s; v.y *= s; v.z *= s; return v; }
inline double dot(const Vec3 &a, const Vec3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(const Vec3 &a, const Vec3 &b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(const Vec3 &v) { return std::sqrt(dot(v, v)); }
inline Vec3 normalize(const Vec3 &v) {
  double l = norm(v);
  return l > 0 ? v / l : v;
}

// ------------------------------------------------------------------ grid
// Cartesian mesh with the Grid methods the hot path calls (GetPointNo,
// GetCellFaces, GetFaceNeighbors, Normal, FaceCM, GetArea, GetVolume,
// GetMeshPoint, IsPointOutsideBox; MonteCarloManagerSerial.hpp:134-135,
// MonteCarloPhysics.hpp:75-97, RadiationIMC.hpp:1001, 1368-1378).  Outside
// "mirror" points have indices N + face, GetMeshPoint() reflects the inside
// centre across the face.
class CartesianGrid {
 public:
  struct Face {
    std::vector<Vec3> vertices;
  };

  CartesianGrid(std::vector<double> xe, std::vector<double> ye, std::vector<double> ze)
      : xe_(std::move(xe)), ye_(std::move(ye)), ze_(std::move(ze)) {
    nx_ = static_cast<int>(xe_.size()) - 1;
    ny_ = static_cast<int>(ye_.size()) - 1;
    nz_ = static_cast<int>(ze_.size()) - 1;
    if (nx_ < 1 || ny_ < 1 || nz_ < 1) throw OracleError("CartesianGrid: empty axis");
    N_ = static_cast<size_t>(nx_) * ny_ * nz_;
    Fx_ = static_cast<size_t>(nx_ + 1) * ny_ * nz_;
    Fy_ = static_cast<size_t>(nx_) * (ny_ + 1) * nz_;
    Fz_ = static_cast<size_t>(nx_) * ny_ * (nz_ + 1);
    ll_ = Vec3(xe_.front(), ye_.front(), ze_.front());
    ur_ = Vec3(xe_.back(), ye_.back(), ze_.back());
    // vertices
    verts_.resize(static_cast<size_t>(nx_ + 1) * (ny_ + 1) * (nz_ + 1));
    for (int i = 0; i <= nx_; ++i)
      for (int j = 0; j <= ny_; ++j)
        for (int k = 0; k <= nz_; ++k) verts_[vid(i, j, k)] = Vec3(xe_[i], ye_[j], ze_[k]);
    // per-cell face lists x-, x+, y-, y+, z-, z+
    cellFaces_.resize(N_);
    for (int i = 0; i < nx_; ++i)
      for (int j = 0; j < ny_; ++j)
        for (int k = 0; k < nz_; ++k) {
          size_t c = cell(i, j, k);
          cellFaces_[c] = {fx(i, j, k), fx(i + 1, j, k), fy(i, j, k),
                           fy(i, j + 1, k), fz(i, j, k), fz(i, j, k + 1)};
        }
    // box faces (vertices[0..2] span the plane; RigidBoundary/SideTemperature loop over them)
    boxFaces_.resize(6);
    boxFaces_[STORM_XLO].vertices = {Vec3(ll_.x, ll_.y, ll_.z), Vec3(ll_.x, ur_.y, ll_.z),
                                     Vec3(ll_.x, ll_.y, ur_.z), Vec3(ll_.x, ur_.y, ur_.z)};
    boxFaces_[STORM_XHI].vertices = {Vec3(ur_.x, ll_.y, ll_.z), Vec3(ur_.x, ur_.y, ll_.z),
                                     Vec3(ur_.x, ll_.y, ur_.z), Vec3(ur_.x, ur_.y, ur_.z)};
    boxFaces_[STORM_YLO].vertices = {Vec3(ll_.x, ll_.y, ll_.z), Vec3(ur_.x, ll_.y, ll_.z),
                                     Vec3(ll_.x, ll_.y, ur_.z), Vec3(ur_.x, ll_.y, ur_.z)};
    boxFaces_[STORM_YHI].vertices = {Vec3(ll_.x, ur_.y, ll_.z), Vec3(ur_.x, ur_.y, ll_.z),
                                     Vec3(ll_.x, ur_.y, ur_.z), Vec3(ur_.x, ur_.y, ur_.z)};
    boxFaces_[STORM_ZLO].vertices = {Vec3(ll_.x, ll_.y, ll_.z), Vec3(ur_.x, ll_.y, ll_.z),
                                     Vec3(ll_.x, ur_.y, ll_.z), Vec3(ur_.x, ur_.y, ll_.z)};
    boxFaces_[STORM_ZHI].vertices = {Vec3(ll_.x, ll_.y, ur_.z), Vec3(ur_.x, ll_.y, ur_.z),
                                     Vec3(ll_.x, ur_.y, ur_.z), Vec3(ur_.x, ur_.y, ur_.z)};
  }

  size_t GetPointNo() const { return N_; }
  std::pair<Vec3, Vec3> GetBoxCoordinates() const { return {ll_, ur_}; }
  int nx() const { return nx_; }
  int ny() const { return ny_; }
  int nz() const { return nz_; }
  const std::vector<double> &xEdges() const { return xe_; }

  size_t cell(int i, int j, int k) const {
    return (static_cast<size_t>(i) * ny_ + j) * nz_ + k;
  }
  void ijk(size_t c, int &i, int &j, int &k) const {
    k = static_cast<int>(c % nz_);
    j = static_cast<int>((c / nz_) % ny_);
    i = static_cast<int>(c / (static_cast<size_t>(ny_) * nz_));
  }
  size_t GetTotalFacesNumber() const { return Fx_ + Fy_ + Fz_; }

  Vec3 GetCellCM(size_t c) const {
    int i, j, k;
    ijk(c, i, j, k);
    return Vec3(0.5 * (xe_[i] + xe_[i + 1]), 0.5 * (ye_[j] + ye_[j + 1]), 0.5 * (ze_[k] + ze_[k + 1]));
  }
  double GetVolume(size_t c) const {
    int i, j, k;
    ijk(c, i, j, k);
    return (xe_[i + 1] - xe_[i]) * (ye_[j + 1] - ye_[j]) * (ze_[k + 1] - ze_[k]);
  }
  Vec3 GetMeshPoint(size_t p) const {
    if (p < N_) return GetCellCM(p);
    size_t f = p - N_;
    auto nb = GetFaceNeighbors(f);
    size_t inside = (nb.first < N_) ? nb.first : nb.second;
    Vec3 c = GetCellCM(inside);
    Vec3 fc = FaceCM(f);
    return fc + (fc - c);  // mirror across the face
  }
  bool IsPointOutsideBox(size_t p) const { return p >= N_; }
  bool IsPointOutsideBox(const Vec3 &p) const {
    return p.x < ll_.x || p.x > ur_.x || p.y < ll_.y || p.y > ur_.y || p.z < ll_.z || p.z > ur_.z;
  }
  bool IsPointInCell(const Vec3 &p, size_t c) const {
    int i, j, k;
    ijk(c, i, j, k);
    return p.x >= xe_[i] && p.x <= xe_[i + 1] && p.y >= ye_[j] && p.y <= ye_[j + 1] &&
           p.z >= ze_[k] && p.z <= ze_[k + 1];
  }
  size_t GetContainingCell(const Vec3 &p) const {
    auto loc = [](const std::vector<double> &e, double v) {
      auto it = std::upper_bound(e.begin(), e.end(), v);
      int idx = static_cast<int>(it - e.begin()) - 1;
      return std::clamp(idx, 0, static_cast<int>(e.size()) - 2);
    };
    return cell(loc(xe_, p.x), loc(ye_, p.y), loc(ze_, p.z));
  }

  const std::array<size_t, 6> &GetCellFaces(size_t c) const { return cellFaces_[c]; }

  // (first, second): lower-side cell (or outside point), upper-side cell (or outside point)
  std::pair<size_t, size_t> GetFaceNeighbors(size_t f) const {
    int axis, i, j, k;
    faceIJK(f, axis, i, j, k);
    int n[3] = {nx_, ny_, nz_};
    int idx[3] = {i, j, k};
    int lo = idx[axis] - 1, hi = idx[axis];
    size_t first, second;
    if (lo < 0) {
      first = N_ + f;
    } else {
      int a[3] = {i, j, k};
      a[axis] = lo;
      first = cell(a[0], a[1], a[2]);
    }
    if (hi >= n[axis]) {
      second = N_ + f;
    } else {
      int a[3] = {i, j, k};
      a[axis] = hi;
      second = cell(a[0], a[1], a[2]);
    }
    return {first, second};
  }
  Vec3 Normal(size_t f) const {
    int axis, i, j, k;
    faceIJK(f, axis, i, j, k);
    Vec3 n(0.0);
    n[axis] = GetArea(f);
    return n;
  }
  Vec3 FaceCM(size_t f) const {
    const auto &v = GetPointsInFace(f);
    Vec3 s(0.0);
    for (size_t q : v) s += verts_[q];
    return s * 0.25;
  }
  double GetArea(size_t f) const {
    int axis, i, j, k;
    faceIJK(f, axis, i, j, k);
    double dx = (axis == 0) ? 1.0 : (xe_[i + 1] - xe_[i]);
    double dy = (axis == 1) ? 1.0 : (ye_[j + 1] - ye_[j]);
    double dz = (axis == 2) ? 1.0 : (ze_[k + 1] - ze_[k]);
    return dx * dy * dz;
  }
  std::array<size_t, 4> GetPointsInFace(size_t f) const {
    int axis, i, j, k;
    faceIJK(f, axis, i, j, k);
    if (axis == 0) return {vid(i, j, k), vid(i, j + 1, k), vid(i, j + 1, k + 1), vid(i, j, k + 1)};
    if (axis == 1) return {vid(i, j, k), vid(i + 1, j, k), vid(i + 1, j, k + 1), vid(i, j, k + 1)};
    return {vid(i, j, k), vid(i + 1, j, k), vid(i + 1, j + 1, k), vid(i, j + 1, k)};
  }
  const std::vector<Vec3> &GetFacePoints() const { return verts_; }
  const std::vector<Face> &GetBoxFaces() const { return boxFaces_; }

  // which box face a boundary face lies on (-1 for interior faces)
  int BoxFaceOf(size_t f) const {
    int axis, i, j, k;
    faceIJK(f, axis, i, j, k);
    int n[3] = {nx_, ny_, nz_};
    int idx[3] = {i, j, k};
    if (idx[axis] == 0) return 2 * axis;
    if (idx[axis] == n[axis]) return 2 * axis + 1;
    return -1;
  }

 private:
  size_t vid(int i, int j, int k) const {
    return (static_cast<size_t>(i) * (ny_ + 1) + j) * (nz_ + 1) + k;
  }
  size_t fx(int i, int j, int k) const { return (static_cast<size_t>(i) * ny_ + j) * nz_ + k; }
  size_t fy(int i, int j, int k) const {
    return Fx_ + (static_cast<size_t>(i) * (ny_ + 1) + j) * nz_ + k;
  }
  size_t fz(int i, int j, int k) const {
    return Fx_ + Fy_ + (static_cast<size_t>(i) * ny_ + j) * (nz_ + 1) + k;
  }
  void faceIJK(size_t f, int &axis, int &i, int &j, int &k) const {
    if (f < Fx_) {
      axis = 0;
      k = static_cast<int>(f % nz_);
      j = static_cast<int>((f / nz_) % ny_);
      i = static_cast<int>(f / (static_cast<size_t>(ny_) * nz_));
    } else if (f < Fx_ + Fy_) {
      f -= Fx_;
      axis = 1;
      k = static_cast<int>(f % nz_);
      j = static_cast<int>((f / nz_) % (ny_ + 1));
      i = static_cast<int>(f / (static_cast<size_t>(ny_ + 1) * nz_));
    } else {
      f -= Fx_ + Fy_;
      axis = 2;
      k = static_cast<int>(f % (nz_ + 1));
      j = static_cast<int>((f / (nz_ + 1)) % ny_);
      i = static_cast<int>(f / (static_cast<size_t>(ny_) * (nz_ + 1)));
    }
  }

  std::vector<double> xe_, ye_, ze_;
  int nx_, ny_, nz_;
  size_t N_, Fx_, Fy_, Fz_;
  Vec3 ll_, ur_;
  std::vector<Vec3> verts_;
  std::vector<std::array<size_t, 6>> cellFaces_;
  std::vector<Face> boxFaces_;
};

// A mesh of convex polyhedral cells (storm_unstructured_mesh: a Voronoi
// tessellation, say) behind the same Grid interface: global faces are built
// from the per-cell half faces (a face shared by cells a < b takes a's
// normal, centroid, area and polygon; its neighbours are (a, b)); a face on
// the box is shared with the outside point N + face, whose GetMeshPoint is
// the cell's mesh point mirrored across the face (MadVoro's outside points).
class UnstructuredGrid {
 public:
  struct Face {
    std::vector<Vec3> vertices;
  };

  explicit UnstructuredGrid(const storm_unstructured_mesh &m) : N_(m.ncells) {
    if (!m.volume || !m.mesh_point || !m.face_start || !m.face_neighbor || !m.face_normal || !m.face_centroid ||
        !m.face_area || !m.face_vertex_start || !m.vertices)
      throw OracleError("UnstructuredGrid: null table");
    ll_ = Vec3(m.box_lo[0], m.box_lo[1], m.box_lo[2]);
    ur_ = Vec3(m.box_hi[0], m.box_hi[1], m.box_hi[2]);
    vol_.assign(m.volume, m.volume + N_);
    mp_.resize(N_);
    for (size_t c = 0; c < N_; ++c) mp_[c] = Vec3(m.mesh_point[3 * c], m.mesh_point[3 * c + 1], m.mesh_point[3 * c + 2]);
    cellFaces_.resize(N_);
    cm_.assign(N_, Vec3(0.0));
    std::unordered_map<uint64_t, size_t> shared;
    for (size_t c = 0; c < N_; ++c) {
      double vsum = 0.0;
      for (uint32_t h = m.face_start[c]; h < m.face_start[c + 1]; ++h) {
        const int32_t nb = m.face_neighbor[h];
        // centroid of the cell: tetrahedra of the face fans about the mesh point
        for (uint32_t v = m.face_vertex_start[h] + 1; v + 1 < m.face_vertex_start[h + 1]; ++v) {
          const uint32_t v0 = m.face_vertex_start[h];
          auto P = [&](uint32_t q) { return Vec3(m.vertices[3 * q], m.vertices[3 * q + 1], m.vertices[3 * q + 2]); };
          const Vec3 a = P(v0) - mp_[c], b = P(v) - mp_[c], d = P(v + 1) - mp_[c];
          const double tv = std::abs(dot(a, cross(b, d))) / 6.0;
          vsum += tv;
          cm_[c] += tv * (mp_[c] + 0.25 * (a + b + d));
        }
        size_t g;
        if (nb >= 0 && static_cast<size_t>(nb) < c) {
          auto it = shared.find((static_cast<uint64_t>(nb) << 32) | c);
          if (it == shared.end()) throw OracleError("UnstructuredGrid: a face is not shared by its neighbour");
          g = it->second;
        } else {
          if (nb >= 0 && static_cast<size_t>(nb) == c) throw OracleError("UnstructuredGrid: a face to its own cell");
          g = nbr_.size();
          const Vec3 n(m.face_normal[3 * h], m.face_normal[3 * h + 1], m.face_normal[3 * h + 2]);
          nbr_.push_back({c, nb >= 0 ? static_cast<size_t>(nb) : N_ + g});
          normal_.push_back(m.face_area[h] * n);
          fc_.push_back(Vec3(m.face_centroid[3 * h], m.face_centroid[3 * h + 1], m.face_centroid[3 * h + 2]));
          area_.push_back(m.face_area[h]);
          std::vector<size_t> idx;
          for (uint32_t v = m.face_vertex_start[h]; v < m.face_vertex_start[h + 1]; ++v) {
            idx.push_back(verts_.size());
            verts_.push_back(Vec3(m.vertices[3 * v], m.vertices[3 * v + 1], m.vertices[3 * v + 2]));
          }
          points_.push_back(std::move(idx));
          if (nb >= 0) shared[(static_cast<uint64_t>(c) << 32) | static_cast<uint32_t>(nb)] = g;
        }
        cellFaces_[c].push_back(g);
      }
      cm_[c] *= 1.0 / vsum;
    }
    const Vec3 ll = ll_, ur = ur_;
    boxFaces_.resize(6);
    boxFaces_[STORM_XLO].vertices = {Vec3(ll.x, ll.y, ll.z), Vec3(ll.x, ur.y, ll.z), Vec3(ll.x, ll.y, ur.z)};
    boxFaces_[STORM_XHI].vertices = {Vec3(ur.x, ll.y, ll.z), Vec3(ur.x, ur.y, ll.z), Vec3(ur.x, ll.y, ur.z)};
    boxFaces_[STORM_YLO].vertices = {Vec3(ll.x, ll.y, ll.z), Vec3(ur.x, ll.y, ll.z), Vec3(ll.x, ll.y, ur.z)};
    boxFaces_[STORM_YHI].vertices = {Vec3(ll.x, ur.y, ll.z), Vec3(ur.x, ur.y, ll.z), Vec3(ll.x, ur.y, ur.z)};
    boxFaces_[STORM_ZLO].vertices = {Vec3(ll.x, ll.y, ll.z), Vec3(ur.x, ll.y, ll.z), Vec3(ll.x, ur.y, ll.z)};
    boxFaces_[STORM_ZHI].vertices = {Vec3(ll.x, ll.y, ur.z), Vec3(ur.x, ll.y, ur.z), Vec3(ll.x, ur.y, ur.z)};
  }

  size_t GetPointNo() const { return N_; }
  std::pair<Vec3, Vec3> GetBoxCoordinates() const { return {ll_, ur_}; }
  Vec3 GetCellCM(size_t c) const { return cm_[c]; }
  double GetVolume(size_t c) const { return vol_[c]; }
  Vec3 GetMeshPoint(size_t p) const {
    if (p < N_) return mp_[p];
    const size_t g = p - N_;
    const Vec3 n = normalize(normal_[g]);
    const Vec3 &in = mp_[nbr_[g].first];
    return in + (2.0 * dot(fc_[g] - in, n)) * n;
  }
  bool IsPointOutsideBox(size_t p) const { return p >= N_; }
  bool IsPointOutsideBox(const Vec3 &p) const {
    return p.x < ll_.x || p.x > ur_.x || p.y < ll_.y || p.y > ur_.y || p.z < ll_.z || p.z > ur_.z;
  }
  const std::vector<size_t> &GetCellFaces(size_t c) const { return cellFaces_[c]; }
  std::pair<size_t, size_t> GetFaceNeighbors(size_t f) const { return nbr_[f]; }
  Vec3 Normal(size_t f) const { return normal_[f]; }
  Vec3 FaceCM(size_t f) const { return fc_[f]; }
  double GetArea(size_t f) const { return area_[f]; }
  const std::vector<size_t> &GetPointsInFace(size_t f) const { return points_[f]; }
  const std::vector<Vec3> &GetFacePoints() const { return verts_; }
  const std::vector<Face> &GetBoxFaces() const { return boxFaces_; }
  size_t GetTotalFacesNumber() const { return nbr_.size(); }

 private:
  size_t N_;
  Vec3 ll_, ur_;
  std::vector<double> vol_;
  std::vector<Vec3> mp_, cm_;
  std::vector<std::vector<size_t>> cellFaces_;
  std::vector<std::pair<size_t, size_t>> nbr_;
  std::vector<Vec3> normal_, fc_;
  std::vector<double> area_;
  std::vector<std::vector<size_t>> points_;
  std::vector<Vec3> verts_;
  std::vector<Face> boxFaces_;
};

// ------------------------------------------------------------------ particle
enum Status { NO_CELL_MOVE = 0, CELL_MOVE = 1, DONE = 2, REMOVE = 3, REFLECT = 4 };  // particle/ParticleStatus.hpp:8-15

struct StepResult {  // particle/StepResult.hpp:9-14
  Status change = NO_CELL_MOVE;
  size_t nextCellIndex = std::numeric_limits<size_t>::max();
};

struct Particle {  // particle/Particle.hpp:67-85 (non-MPI build)
  size_t id = std::numeric_limits<size_t>::max();
  size_t cellID = std::numeric_limits<size_t>::max();
  size_t sourceCellID = std::numeric_limits<size_t>::max();
  Vec3 location = Vec3(std::numeric_limits<double>::max());
  Vec3 velocity = Vec3(std::numeric_limits<double>::max());
  size_t cellIndex = std::numeric_limits<size_t>::max();
  double timeLeft = std::numeric_limits<double>::max();
  double frequency = std::numeric_limits<double>::max();
  double weight = 0.0;
  double initialWeight = 0.0;
  size_t steps = 0;
  bool on_track = false;
  bool sent = false;
};

// ------------------------------------------------------------------ RNG
struct Rng {
  std::mt19937_64 gen;
  std::uniform_real_distribution<double> dist{0.0, 1.0};
  explicit Rng(uint64_t seed) : gen(seed) {}
  double unit() { return dist(gen); }
  // RadiationIMC::randomUnitOpen (radiation/RadiationIMC.hpp:701-713)
  double unitOpen() {
    double v = dist(gen);
    if (v <= 0.0) return std::numeric_limits<double>::min();
    if (v >= 1.0) return std::nextafter(1.0, 0.0);
    return v;
  }
};

inline uint64_t mixSeed(uint64_t seed, uint64_t tag) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (tag + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// isotropicVelocity (radiation/RadiationOpacityModel.hpp:139-150)
inline Vec3 isotropicVelocity(Rng &rng) {
  constexpr double twoPi = 6.283185307179586476925286766559;
  double theta = twoPi * rng.unit();
  double cosTheta = 2.0 * rng.unit() - 1.0;
  double sinTheta = std::sqrt(std::max(0.0, 1.0 - cosTheta * cosTheta));
  return Vec3(sinTheta * std::cos(theta), sinTheta * std::sin(theta), cosTheta) * units::clight;
}

// ------------------------------------------------------------------ samplers
// RandomInCellPositionSampler (utils/RandomInCell.hpp:13-66): fan the cell's
// faces into tetrahedra around the cell centre, pick one by volume, then a
// uniform point from sorted uniforms.  The per-cell triangulation is cached
// (the reference rebuilds it each call; the draws and result are identical).
template <typename Grid>
class RandomInCellSampler {
 public:
  explicit RandomInCellSampler(const Grid &g) : grid_(g), cache_(g.GetPointNo()) {}
  Vec3 operator()(size_t c, Rng &rng) {
    Entry &e = cache_[c];
    if (e.tris.empty()) build(c, e);
    const auto &verts = grid_.GetFacePoints();
    double r = rng.unit() * e.total;
    size_t idx = static_cast<size_t>(std::lower_bound(e.cum.begin(), e.cum.end(), r) - e.cum.begin());
    if (idx >= e.tris.size()) idx = e.tris.size() - 1;
    double s = rng.unit(), t = rng.unit(), u = rng.unit();
    if (s > t) std::swap(s, t);
    if (t > u) std::swap(t, u);
    if (s > t) std::swap(s, t);
    const auto &tv = e.tris[idx];
    return s * verts[tv[0]] + (t - s) * verts[tv[1]] + (u - t) * verts[tv[2]] + (1 - u) * e.center;
  }

 private:
  struct Entry {
    Vec3 center;
    double total = 0;
    std::vector<double> cum;
    std::vector<std::array<size_t, 3>> tris;
  };
  void build(size_t c, Entry &e) {
    e.center = grid_.GetMeshPoint(c);
    const auto &verts = grid_.GetFacePoints();
    for (size_t f : grid_.GetCellFaces(c)) {
      auto fv = grid_.GetPointsInFace(f);
      for (size_t i = 1; i + 1 < fv.size(); ++i) {
        Vec3 a = verts[fv[0]] - e.center, b = verts[fv[i]] - e.center, d = verts[fv[i + 1]] - e.center;
        double vol = std::abs(dot(a, cross(b, d)));
        e.total += vol;
        e.cum.push_back(e.total);
        e.tris.push_back({fv[0], fv[i], fv[i + 1]});
      }
    }
  }
  const Grid &grid_;
  std::vector<Entry> cache_;
};

// RandomPointOnFace (utils/RandomOnFace.hpp:14-66): triangle fan, area-weighted
// triangle choice, reflected (r1, r2) inside the triangle.
template <typename Grid>
Vec3 RandomPointOnFace(const Grid &grid, size_t f, Rng &rng) {
  auto fv = grid.GetPointsInFace(f);
  const auto &verts = grid.GetFacePoints();
  std::vector<double> cumA;
  std::vector<std::array<size_t, 3>> tris;
  double total = 0;
  for (size_t i = 1; i + 1 < fv.size(); i++) {
    double area = norm(cross(verts[fv[i]] - verts[fv[0]], verts[fv[i + 1]] - verts[fv[0]]));
    total += area;
    cumA.push_back(total);
    tris.push_back({fv[0], fv[i], fv[i + 1]});
  }
  double r = rng.unit() * total;
  size_t idx = static_cast<size_t>(std::lower_bound(cumA.begin(), cumA.end(), r) - cumA.begin());
  if (idx >= tris.size()) idx = tris.size() - 1;
  double r1 = rng.unit(), r2 = rng.unit();
  if (r1 + r2 > 1) {
    r1 = 1 - r1;
    r2 = 1 - r2;
  }
  const auto &tv = tris[idx];
  return (1 - r1 - r2) * verts[tv[0]] + r1 * verts[tv[1]] + r2 * verts[tv[2]];
}

// ------------------------------------------------------------------ cells
struct Cell {          // radiation/RadiationCell.hpp:6-12
  double temperature = 0;
  double cv = 0;
  double internalEnergy = 0;
  double Erad = 0;
};
struct Extensives {    // radiation/RadiationCell.hpp:14-19
  double internal_energy = 0;
  double mass = 0;
  double Erad = 0;
};

// Power-law opacity of storm_material_desc (the MarshakOpacity family,
// examples/marshak_wave/MarshakOpacity.hpp:34-49).
struct Material {
  std::vector<double> sa0, ss0, rho;
  double Tref = 1, Tfloor = 0, apow = 0, arho = 0, spow = 0, srho = 0;
  double eos_f = 1, eos_beta = 1, eos_mu = 0, cvTfloor = 0;

  double density(size_t c) const { return rho.empty() ? 1.0 : rho[c]; }
  double tfac(double T, double p) const {
    if (p == 0.0) return 1.0;
    return std::pow(std::max(T, Tfloor) / Tref, -p);
  }
  double rfac(size_t c, double p) const {
    if (p == 0.0 || rho.empty()) return 1.0;
    return std::pow(rho[c], p);
  }
  // gray absorption opacity sigma_a(c, T)
  double grayOpacity(size_t c, double T) const { return sa0[c] * tfac(T, apow) * rfac(c, arho); }

  // Multigroup (storm_group_desc): group boundaries in erg, sigma_a(c, T, E)
  // = grayOpacity * (max(E, E_0) / nu_ref)^(-nu_pow)
  // (DensmoreOpacity::CalcAbsorptionOpacity, examples/densmore2012/DensmoreOpacity.hpp:56-64).
  std::vector<double> bounds;
  double nu_pow = 0.0, nu_ref = 1.0;
  size_t groups() const { return bounds.size() > 2 ? bounds.size() - 1 : 1; }
  bool multigroup() const { return groups() > 1; }
  double CalcAbsorptionOpacity(size_t c, double T, double E) const {
    return grayOpacity(c, T) * std::pow(std::max(E, bounds[0]) / nu_ref, -nu_pow);
  }
  // Cumulative sigma_a(Ec_g) B_g over groups (computeCumulativePlanck,
  // DensmoreOpacity.hpp:128-148) and the Planck weights B_g.
  void cumulativePlanck(size_t c, double T, std::vector<double> &cdf, std::vector<double> &Bg) const {
    const size_t G = groups();
    cdf.assign(G + 1, 0.0);
    Bg.assign(G, 0.0);
    const double kT = units::k_boltz * std::max(T, Tfloor);
    for (size_t g = 0; g < G; ++g) {
      const double a = bounds[g] / kT, b = bounds[g + 1] / kT;
      const double bg = (a >= 0.0 && b > a && std::isfinite(b)) ? storm::planck::integral(a, b) : 0.0;
      const double sg = CalcAbsorptionOpacity(c, T, 0.5 * (bounds[g] + bounds[g + 1]));
      Bg[g] = bg;
      const double w = (sg > 0.0 && std::isfinite(sg) && std::isfinite(bg)) ? sg * bg : 0.0;
      cdf[g + 1] = cdf[g] + w;
    }
  }
  // CalcPlanckOpacity (DensmoreOpacity.hpp:38-55 for multigroup; gray: sigma_a(c, T))
  double CalcPlanckOpacity(size_t c, double T) const {
    if (!multigroup()) return grayOpacity(c, T);
    std::vector<double> cdf, Bg;
    cumulativePlanck(c, T, cdf, Bg);
    double weighted = cdf.back(), total = 0.0;
    for (double b : Bg) total += b;
    return total > 0.0 ? weighted / total : 1e-20;
  }
  // GetThermalEnergy (DensmoreOpacity.hpp:66-77, base class
  // radiation/RadiationOpacityModel.hpp:88-92 when the weights vanish)
  double GetThermalEnergy(size_t c, double T, double random) const {
    std::vector<double> cdf, Bg;
    cumulativePlanck(c, T, cdf, Bg);
    return thermalEnergyFromCumulative(cdf, random);
  }
  // the same from a cumulative computed once per step (the cell temperature
  // is constant during transport)
  double thermalEnergyFromCumulative(const std::vector<double> &cdf, double random) const {
    const size_t G = groups();
    const double total = cdf[G];
    if (!(total > 0.0) || !std::isfinite(total)) return bounds[0] + random * (bounds[G] - bounds[0]);
    const double r = std::clamp(random, 0.0, std::nextafter(1.0, 0.0));
    return linearInterpolation(cdf, bounds, r * total);
  }
  // LinearInterpolation (utils/LinearInterpolation.hpp:17-42)
  static double linearInterpolation(const std::vector<double> &x, const std::vector<double> &y, double xi) {
    auto it = std::upper_bound(x.begin(), x.end(), xi);
    if (it == x.end()) return y.back();
    if (it == x.begin()) return y.front();
    const size_t i = static_cast<size_t>(it - x.begin());
    if (std::abs(*it - xi) < 1e-12 * std::max(1.0, std::abs(xi))) return y[i];
    return y[i] + (xi - *it) * (y[i - 1] - y[i]) / (*(it - 1) - *it);
  }

  double CalcScatteringOpacity(size_t c, double T) const {
    if (ss0.empty()) return 0.0;
    return ss0[c] * tfac(T, spow) * rfac(c, srho);
  }
  // EOS duck type (MarshakEOS, examples/marshak_wave/MarshakCommon.hpp:151-172)
  double dT2cv(double density, double T) const {
    double cvVol = eos_f * eos_beta * std::pow(std::max(T, cvTfloor), eos_beta - 1.0) *
                   std::pow(density, 1.0 - eos_mu);
    return cvVol / density;
  }
  double de2T(double density, double specificEnergy) const {
    double eVol = specificEnergy * density;
    double ratio = std::max(eVol, 0.0) / (eos_f * std::pow(density, 1.0 - eos_mu));
    return std::pow(std::max(ratio, 0.0), 1.0 / eos_beta);
  }
  double eFromT(double T, double density) const {
    return eos_f * std::pow(T, eos_beta) * std::pow(density, 1.0 - eos_mu);
  }
};

// ------------------------------------------------------------------ boundary
// A per-box-face table of REFLECT/REMOVE plus an optional Planck surface
// source.  apply() follows SideTemperature::apply (boundary/SideTemperature.hpp:84-112)
// generalised to a face table: the face the packet sits on is found with
// getInwardBoxFaceNormalIfClose (boundary/BoundaryCondition.hpp:56-85); a
// REMOVE face removes, every other face reflects
// (reflectParticleOnBoxFace, boundary/BoundaryCondition.hpp:87-108).
template <typename Grid>
class FaceTableBoundary {
 public:
  FaceTableBoundary(const Grid &g, const storm_boundary_desc &d, uint64_t seed)
      : grid_(g), desc_(d), re_(mixSeed(seed, 11)), faceRe_(mixSeed(seed, 12)) {}

  double escapedEnergy = 0.0;
  double sourceEnergy = 0.0;

  // SetTemperature (boundary/SideTemperature.hpp:48); ignored once T^4 is frozen
  void setSourceTemperature(double T) { desc_.source_temperature = T; }
  // multigroup source: group boundaries in erg
  void setGroupBounds(const std::vector<double> &b) { bounds_ = b; }

  Status apply(Particle &p) {
    Status status = DONE;
    const auto &faces = grid_.GetBoxFaces();
    for (int f = 0; f < 6; ++f) {
      Vec3 normal;
      double faceScale = 0.0;
      if (getInwardBoxFaceNormalIfClose(faces[f], p.location, normal, faceScale)) {
        if (desc_.face_kind[f] == STORM_FACE_ESCAPE) {
          escapedEnergy += p.weight;
          return REMOVE;
        }
        if (reflectParticleOnBoxFace(p, faces[f])) status = REFLECT;
      }
    }
    if (status == REFLECT) return status;
    throw OracleError("FaceTableBoundary: particle is not on any boundary");
  }

  // SideTemperature::generateNewBoundaryParticles (boundary/SideTemperature.hpp:114-163)
  // and, with source_xhi, the x+ side of TwoSidesTemperature
  // (boundary/TwoSidesTemperature.hpp:156-199: T^4 of each side, mu = sqrt(U)
  // along the inward normal).
  std::vector<Particle> generateNewBoundaryParticles(double fullDt) {
    std::vector<Particle> out;
    const bool lo = desc_.source_face == STORM_XLO, hi = desc_.source_xhi != 0;
    if ((!lo && !hi) || desc_.source_packets_per_face == 0) return out;
    const double Ts[2] = {desc_.source_temperature, desc_.source_temperature_xhi};
    double T4[2] = {Ts[0] * Ts[0] * Ts[0] * Ts[0], Ts[1] * Ts[1] * Ts[1] * Ts[1]};
    if (desc_.freeze_source_T4) {
      if (!frozen_) {
        frozenT4_[0] = T4[0];
        frozenT4_[1] = T4[1];
        frozen_ = true;
      }
      T4[0] = frozenT4_[0];
      T4[1] = frozenT4_[1];
    }
    size_t N = grid_.GetPointNo();
    size_t Npc = desc_.source_packets_per_face;
    std::vector<double> Bcdf[2];
    if (!bounds_.empty()) {
      const size_t G = bounds_.size() - 1;
      for (int s = 0; s < 2; ++s) {
        const double kT = units::k_boltz * Ts[s];
        Bcdf[s].assign(G + 1, 0.0);
        for (size_t g = 0; g < G; ++g)
          Bcdf[s][g + 1] = Bcdf[s][g] + storm::planck::integral(bounds_[g] / kT, bounds_[g + 1] / kT);
      }
    }
    for (size_t i = 0; i < N; i++) {
      const Vec3 point = grid_.GetMeshPoint(i);
      for (size_t faceIdx : grid_.GetCellFaces(i)) {
        auto nb = grid_.GetFaceNeighbors(faceIdx);
        size_t neighborIdx = (nb.first == i) ? nb.second : nb.first;
        if (neighborIdx >= N && grid_.IsPointOutsideBox(neighborIdx)) {
          Vec3 normal = normalize(grid_.GetMeshPoint(neighborIdx) - point);
          int side = -1;
          if (normal.x < -0.99 && lo) side = 0;
          if (normal.x > 0.99 && hi) side = 1;
          if (side >= 0) {
            double energyToProduce = units::sigma_sb * T4[side] * grid_.GetArea(faceIdx) * fullDt / Npc;
            for (size_t j = 0; j < Npc; j++) {
              Particle np;
              np.location = RandomPointOnFace(grid_, faceIdx, faceRe_);
              double mu = std::sqrt(re_.unit());
              np.velocity.x = side == 0 ? mu : -mu;
              double _1mmu = std::sqrt(1 - mu * mu);
              double theta = 2 * M_PI * re_.unit();
              np.velocity.y = _1mmu * std::cos(theta);
              np.velocity.z = _1mmu * std::sin(theta);
              np.velocity *= units::clight;
              np.frequency = 0;
              if (!bounds_.empty()) {
                // DensmoreBoundary::generateNewBoundaryParticles
                // (examples/densmore2012/DensmoreBoundary.hpp:96-130): group by
                // the Planck fractions at T, energy uniform in the group
                const size_t G = bounds_.size() - 1;
                double r = re_.unit() * Bcdf[side].back();
                size_t g = 0;
                while (g + 1 < G && Bcdf[side][g + 1] < r) ++g;
                np.frequency = bounds_[g] + re_.unit() * (bounds_[g + 1] - bounds_[g]);
              }
              np.weight = energyToProduce;
              np.initialWeight = np.weight;
              np.timeLeft = fullDt * re_.unit();
              np.cellIndex = i;
              sourceEnergy += np.weight;
              out.push_back(np);
            }
          }
        }
      }
    }
    return out;
  }

 private:
  bool getInwardBoxFaceNormalIfClose(const typename Grid::Face &face, const Vec3 &location,
                                     Vec3 &normal, double &faceScale) const {
    const Vec3 &onFace = face.vertices[0];
    Vec3 u = face.vertices[1] - face.vertices[0];
    Vec3 v = face.vertices[2] - face.vertices[0];
    normal = cross(u, v);
    faceScale = std::min(norm(u), norm(v));
    double normalNorm = norm(normal);
    if (!(normalNorm > 0.0) || !std::isfinite(normalNorm) || !(faceScale > 0.0) || !std::isfinite(faceScale))
      return false;
    double planeDistance = dot(normal, location - onFace);
    if (std::fabs(planeDistance) >= kEpsilonParticle * faceScale * faceScale * faceScale) return false;
    normal *= 1.0 / normalNorm;
    auto box = grid_.GetBoxCoordinates();
    Vec3 boxCenter = 0.5 * (box.first + box.second);
    if (dot(normal, boxCenter - onFace) < 0.0) normal *= -1.0;
    return true;
  }
  bool reflectParticleOnBoxFace(Particle &p, const typename Grid::Face &face) const {
    Vec3 normal;
    double faceScale = 0.0;
    if (!getInwardBoxFaceNormalIfClose(face, p.location, normal, faceScale)) return false;
    const Vec3 &onFace = face.vertices[0];
    double signedDistance = dot(p.location - onFace, normal);
    p.location -= signedDistance * normal;
    constexpr double nudge = 1e-6;
    p.location += nudge * faceScale * normal;
    double vn = dot(p.velocity, normal);
    if (vn < 0.0) p.velocity -= 2.0 * vn * normal;
    return true;
  }

  const Grid &grid_;
  storm_boundary_desc desc_;
  std::vector<double> bounds_;
  Rng re_, faceRe_;
  bool frozen_ = false;
  double frozenT4_[2] = {0.0, 0.0};
};

// ------------------------------------------------------------------ random walk tables
// RandomWalk (radiation/RandomWalk.hpp:20-242): survival and radius CDF
// tables of diffusion in a sphere from the eigenfunction series.
class RandomWalk {
 public:
  static constexpr double PI_ = 3.14159265358979323846;
  static constexpr int eigenTerms_ = 500;
  static constexpr double tauFreeSpace_ = 0.01;
  static constexpr double tauMin_ = 1e-8;
  static constexpr double tauMax_ = 64.0;
  static constexpr size_t tableSize_ = 1024;
  static constexpr size_t radiusTableSize_ = 256;

  RandomWalk() {
    tauTable_.resize(tableSize_);
    survivalTable_.resize(tableSize_);
    double logMin = std::log(tauMin_), logMax = std::log(tauMax_);
    for (size_t i = 0; i < tableSize_; ++i) {
      double frac = static_cast<double>(i) / static_cast<double>(tableSize_ - 1);
      double tau = std::exp(logMin + frac * (logMax - logMin));
      tauTable_[i] = tau;
      survivalTable_[i] = computeSurvival(tau);
    }
    constexpr size_t xGridSize = 512;
    radiusTable_.resize(tableSize_ * radiusTableSize_);
    std::vector<double> gGrid(xGridSize);
    for (size_t i = 0; i < tableSize_; ++i) {
      double tau = tauTable_[i];
      for (size_t k = 0; k < xGridSize; ++k) {
        double x = static_cast<double>(k) / static_cast<double>(xGridSize - 1);
        gGrid[k] = computeRadialCDF(x, tau);
      }
      for (size_t j = 0; j < radiusTableSize_; ++j) {
        double xi = static_cast<double>(j) / static_cast<double>(radiusTableSize_ - 1);
        double &out = radiusTable_[i * radiusTableSize_ + j];
        if (xi <= gGrid[0]) {
          out = 0.0;
        } else if (xi >= gGrid[xGridSize - 1]) {
          out = 1.0;
        } else {
          auto it = std::lower_bound(gGrid.begin(), gGrid.end(), xi);
          size_t idx = static_cast<size_t>(it - gGrid.begin());
          if (idx == 0) idx = 1;
          double g0 = gGrid[idx - 1], g1 = gGrid[idx];
          double x0 = static_cast<double>(idx - 1) / static_cast<double>(xGridSize - 1);
          double x1 = static_cast<double>(idx) / static_cast<double>(xGridSize - 1);
          double f = (g1 > g0) ? (xi - g0) / (g1 - g0) : 0.5;
          out = x0 + f * (x1 - x0);
        }
      }
    }
  }

  double sampleLeakTime(double xi) const {  // RandomWalk.hpp:81-112
    double target = 1.0 - xi;
    if (target >= survivalTable_.front()) return tauTable_.front();
    if (target <= survivalTable_.back()) return tauTable_.back();
    auto it = std::lower_bound(survivalTable_.begin(), survivalTable_.end(), target,
                               [](double a, double b) { return a > b; });
    if (it == survivalTable_.begin()) return tauTable_.front();
    if (it == survivalTable_.end()) return tauTable_.back();
    size_t idx = static_cast<size_t>(it - survivalTable_.begin());
    double S0 = survivalTable_[idx - 1], S1 = survivalTable_[idx];
    double t0 = tauTable_[idx - 1], t1 = tauTable_[idx];
    double frac = (S0 - target) / (S0 - S1);
    double logT = std::log(t0) + frac * (std::log(t1) - std::log(t0));
    return std::exp(logT);
  }

  double sampleRadius(double tau, double xi) const {  // RandomWalk.hpp:114-137
    double logTau = std::log(std::clamp(tau, tauMin_, tauMax_));
    double logMin = std::log(tauMin_), logMax = std::log(tauMax_);
    double tauPos = (logTau - logMin) / (logMax - logMin) * static_cast<double>(tableSize_ - 1);
    size_t tauIdx = std::min(static_cast<size_t>(tauPos), tableSize_ - 2);
    double tauFrac = tauPos - static_cast<double>(tauIdx);
    double xiClamped = std::clamp(xi, 0.0, 1.0);
    double xiPos = xiClamped * static_cast<double>(radiusTableSize_ - 1);
    size_t xiIdx = std::min(static_cast<size_t>(xiPos), radiusTableSize_ - 2);
    double xiFrac = xiPos - static_cast<double>(xiIdx);
    double r00 = radiusTable_[tauIdx * radiusTableSize_ + xiIdx];
    double r01 = radiusTable_[tauIdx * radiusTableSize_ + xiIdx + 1];
    double r10 = radiusTable_[(tauIdx + 1) * radiusTableSize_ + xiIdx];
    double r11 = radiusTable_[(tauIdx + 1) * radiusTableSize_ + xiIdx + 1];
    double r0 = r00 + xiFrac * (r01 - r00);
    double r1 = r10 + xiFrac * (r11 - r10);
    return std::clamp(r0 + tauFrac * (r1 - r0), 0.0, 1.0);
  }

  double computeSurvival(double tau) const {  // RandomWalk.hpp:139-150
    if (tau <= 0.0) return 1.0;
    if (tau < 1e-4) return 1.0;
    return survivalEigen(tau);
  }
  double computeRadialCDF(double x, double tau) const {  // RandomWalk.hpp:152-167
    if (tau <= 0.0) return (x > 0.0) ? 1.0 : 0.0;
    if (x <= 0.0) return 0.0;
    if (x >= 1.0) return 1.0;
    return (tau < tauFreeSpace_) ? radialCdfFreeSpace(x, tau) : radialCdfEigen(x, tau);
  }

  const std::vector<double> &tauTable() const { return tauTable_; }
  const std::vector<double> &survivalTable() const { return survivalTable_; }
  const std::vector<double> &radiusTable() const { return radiusTable_; }

 private:
  double survivalEigen(double tau) const {
    double sum = 0.0;
    for (int n = 1; n <= eigenTerms_; ++n) {
      double arg = -static_cast<double>(n * n) * PI_ * PI_ * tau;
      if (arg < -700.0) break;
      double term = std::exp(arg);
      sum += (n % 2 == 1) ? term : -term;
    }
    return std::clamp(2.0 * sum, 0.0, 1.0);
  }
  static double radialCdfFreeSpace(double x, double tau) {
    if (x <= 0.0) return 0.0;
    if (x >= 1.0) return 1.0;
    double a = x / (2.0 * std::sqrt(tau));
    return std::clamp(std::erf(a) - 2.0 * a * std::exp(-a * a) / std::sqrt(PI_), 0.0, 1.0);
  }
  double radialCdfEigen(double x, double tau) const {
    if (x <= 0.0) return 0.0;
    if (x >= 1.0) return 1.0;
    double S = computeSurvival(tau);
    if (S < 1e-30) return 1.0;
    double sum = 0.0;
    for (int n = 1; n <= eigenTerms_; ++n) {
      double arg = -static_cast<double>(n * n) * PI_ * PI_ * tau;
      if (arg < -700.0) break;
      double npi = static_cast<double>(n) * PI_;
      double npx = npi * x;
      sum += std::exp(arg) * (std::sin(npx) / npi - x * std::cos(npx));
    }
    return std::clamp(2.0 * sum / S, 0.0, 1.0);
  }

  std::vector<double> tauTable_, survivalTable_, radiusTable_;
};

// ------------------------------------------------------------------ comb
// PopulationControl (population/PopulationControl.hpp): the manager's hook.
struct PopulationControl {
  virtual ~PopulationControl() = default;
  virtual std::vector<Particle> activate(const std::vector<Particle> &particles) = 0;
};

// CombPopulationControl::activate (population/CombPopulationControl.hpp:79-253)
template <typename Grid>
class Comb : public PopulationControl {
 public:
  Comb(const Grid &g, size_t Nmin, double factor, uint64_t seed)
      : grid_(g), Nmin_(Nmin), factor_(factor), gen_(mixSeed(seed, 13)) {}

  std::vector<Particle> activate(const std::vector<Particle> &particles) override {
    std::vector<Particle> result;
    size_t Ncells = grid_.GetPointNo();
    size_t Ntotal = Ncells;
    std::vector<double> weights(Ncells, 0);
    std::vector<std::vector<const Particle *>> inCell(Ncells);
    double totalWeight = 0.0;
    for (const Particle &p : particles) {
      if (grid_.IsPointOutsideBox(p.location)) throw OracleError("Comb: input particle is outside the box");
      if (!std::isfinite(p.weight)) throw OracleError("Comb: particle weight is not finite");
      if (p.weight < 0.0) throw OracleError("Comb: particle weight is negative");
      if (p.weight == 0.0) continue;
      weights[p.cellIndex] += p.weight;
      totalWeight += p.weight;
      inCell[p.cellIndex].push_back(&p);
    }
    if (!std::isfinite(totalWeight)) throw OracleError("Comb: total particle weight is not finite");
    if (totalWeight == 0.0) return result;
    Ntotal = static_cast<size_t>(Ntotal * Nmin_ * factor_);
    std::uniform_real_distribution<double> dist(0, 1);
    for (size_t i = 0; i < Ncells; i++) {
      size_t NinCell = std::min(Nmin_ * 20, std::max<size_t>(Nmin_, static_cast<size_t>(Ntotal * weights[i] / totalWeight)));
      if (inCell[i].size() <= NinCell) {
        double weight_ideal = weights[i] / NinCell;
        for (const Particle *p : inCell[i]) {
          if (p->weight > 2 * weight_ideal) {
            Particle cpy = *p;
            size_t Nsplit = static_cast<size_t>(std::ceil(p->weight / weight_ideal));
            double weight_split = p->weight / Nsplit;
            cpy.weight = weight_split;
            cpy.initialWeight = weight_split;
            cpy.id = std::numeric_limits<size_t>::max();
            cpy.cellIndex = i;
            cpy.timeLeft = 0;
            cpy.steps = 0;
            for (size_t j = 0; j < Nsplit; j++) result.push_back(cpy);
          } else {
            result.push_back(*p);
          }
        }
        continue;
      }
      std::sort(inCell[i].begin(), inCell[i].end(), [](const Particle *a, const Particle *b) { return a->id < b->id; });
      std::shuffle(inCell[i].begin(), inCell[i].end(), gen_);
      double new_energy = weights[i] / NinCell;
      double r = dist(gen_);
      size_t comb_index = 0;
      double cum_sum_w = 0;
      for (const Particle *p : inCell[i]) {
        while ((cum_sum_w + p->weight) > (comb_index + r) * new_energy) {
          ++comb_index;
          result.push_back(*p);
          result.back().id = std::numeric_limits<size_t>::max();
          result.back().cellIndex = i;
          result.back().timeLeft = 0;
          result.back().weight = new_energy;
          result.back().initialWeight = new_energy;
        }
        cum_sum_w += p->weight;
      }
    }
    for (Particle &p : result) {
      if (grid_.IsPointOutsideBox(p.location)) {
        Vec3 original = p.location;
        Vec3 direction = grid_.GetMeshPoint(p.cellIndex) - original;
        double t = 1e-6;
        while (grid_.IsPointOutsideBox(p.location) && t < 1.0) {
          p.location = original + t * direction;
          t *= 2;
        }
        if (grid_.IsPointOutsideBox(p.location)) p.location = grid_.GetMeshPoint(p.cellIndex);
      }
    }
    return result;
  }

 private:
  const Grid &grid_;
  size_t Nmin_;
  double factor_;
  std::mt19937_64 gen_;
};

// StratifiedCombPopulationControl::activate
// (population/CombPopulationControl.hpp:256-511): the census binned by cell
// and class (here the photon group of the packet's energy, findGroup over
// `bounds`); each cell's target (as in Comb) is shared among its occupied
// bins -- minPerGroup each when they all fit, else one each for the heaviest
// `target` bins and none for the rest, whose packets are dropped (the
// reference discards them, :409-413 with :348-349) -- then the remainder in
// proportion to the bin weights with the largest fractional parts rounded up
// (:421-451); each bin is then combed or split like a Comb cell (:346-404).
// `dropped` sums the weight of the dropped bins.
template <typename Grid>
class StratifiedComb : public PopulationControl {
 public:
  StratifiedComb(const Grid &g, std::vector<double> bounds, size_t Nmin, double factor, size_t minPerGroup,
                 uint64_t seed)
      : grid_(g), bounds_(std::move(bounds)), G_(std::max<size_t>(1, bounds_.size() > 1 ? bounds_.size() - 1 : 1)),
        Nmin_(Nmin), factor_(factor), minPer_(minPerGroup), gen_(mixSeed(seed, 17)) {}

  double dropped = 0.0;

  size_t classify(const Particle &p) const {
    if (bounds_.size() < 2) return 0;
    for (size_t g = 0; g < G_; ++g)
      if (p.frequency < bounds_[g + 1]) return g;
    return G_ - 1;
  }

  std::vector<Particle> activate(const std::vector<Particle> &particles) override {
    std::vector<Particle> result;
    dropped = 0.0;
    const size_t N = grid_.GetPointNo();
    std::vector<double> cellW(N, 0.0), binW(N * G_, 0.0);
    std::vector<std::vector<const Particle *>> bins(N * G_);
    double total = 0.0;
    for (const Particle &p : particles) {
      if (grid_.IsPointOutsideBox(p.location)) throw OracleError("Stratified comb: input particle is outside the box");
      if (!std::isfinite(p.weight)) throw OracleError("Stratified comb: particle weight is not finite");
      if (p.weight < 0.0) throw OracleError("Stratified comb: particle weight is negative");
      if (p.weight == 0.0) continue;
      if (p.cellIndex >= N) throw OracleError("Stratified comb: invalid cell index");
      const size_t b = p.cellIndex * G_ + std::min(classify(p), G_ - 1);
      cellW[p.cellIndex] += p.weight;
      binW[b] += p.weight;
      total += p.weight;
      bins[b].push_back(&p);
    }
    if (!std::isfinite(total)) throw OracleError("Stratified comb: total particle weight is not finite");
    if (total == 0.0) return result;
    const size_t Ntotal = static_cast<size_t>(N * Nmin_ * factor_);
    std::uniform_real_distribution<double> dist(0, 1);
    for (size_t c = 0; c < N; ++c) {
      if (cellW[c] <= 0.0) continue;
      const size_t target = std::min(Nmin_ * 20, std::max(Nmin_, static_cast<size_t>(Ntotal * cellW[c] / total)));
      std::vector<size_t> active;
      for (size_t g = 0; g < G_; ++g)
        if (!bins[c * G_ + g].empty()) active.push_back(g);
      if (active.empty()) continue;
      std::vector<size_t> tg(G_, 0);
      size_t given = 0;
      if (active.size() * minPer_ <= target) {
        for (size_t g : active) {
          tg[g] = minPer_;
          given += minPer_;
        }
      } else {
        std::sort(active.begin(), active.end(), [&](size_t a, size_t b) { return binW[c * G_ + a] > binW[c * G_ + b]; });
        const size_t keep = std::min(active.size(), target);
        for (size_t j = 0; j < keep; ++j) {
          tg[active[j]] = 1;
          ++given;
        }
      }
      const size_t rest = target > given ? target - given : 0;
      size_t extraSum = 0;
      std::vector<double> frac(G_, 0.0);
      for (size_t g : active) {
        if (tg[g] == 0) continue;
        const double x = static_cast<double>(rest) * binW[c * G_ + g] / cellW[c];
        const size_t e = static_cast<size_t>(std::floor(x));
        tg[g] += e;
        extraSum += e;
        frac[g] = x - static_cast<double>(e);
      }
      while (extraSum < rest) {
        size_t best = G_;
        double bf = -1.0;
        for (size_t g : active)
          if (tg[g] > 0 && frac[g] > bf) {
            best = g;
            bf = frac[g];
          }
        if (best == G_) break;
        ++tg[best];
        frac[best] = 0.0;
        ++extraSum;
      }
      for (size_t g : active) combBin(bins[c * G_ + g], binW[c * G_ + g], tg[g], c, result, dist);
    }
    // packets outside the box move toward their cell's mesh point (:487-506)
    for (Particle &p : result) {
      if (!grid_.IsPointOutsideBox(p.location)) continue;
      const Vec3 original = p.location, direction = grid_.GetMeshPoint(p.cellIndex) - original;
      double t = 1e-6;
      while (grid_.IsPointOutsideBox(p.location) && t < 1.0) {
        p.location = original + t * direction;
        t *= 2;
      }
      if (grid_.IsPointOutsideBox(p.location)) p.location = grid_.GetMeshPoint(p.cellIndex);
    }
    return result;
  }

 private:
  void combBin(std::vector<const Particle *> &bin, double W, size_t target, size_t c, std::vector<Particle> &out,
               std::uniform_real_distribution<double> &dist) {
    if (bin.empty() || W <= 0.0) return;
    if (target == 0) {
      dropped += W;
      return;
    }
    if (bin.size() <= target) {
      const double ideal = W / static_cast<double>(target);
      for (const Particle *p : bin) {
        if (p->weight > 2.0 * ideal) {
          const size_t k = static_cast<size_t>(std::ceil(p->weight / ideal));
          Particle cp = *p;
          cp.weight = cp.initialWeight = p->weight / static_cast<double>(k);
          cp.id = std::numeric_limits<size_t>::max();
          cp.cellIndex = c;
          cp.timeLeft = 0;
          cp.steps = 0;
          for (size_t j = 0; j < k; ++j) out.push_back(cp);
        } else {
          out.push_back(*p);
        }
      }
      return;
    }
    std::sort(bin.begin(), bin.end(), [](const Particle *a, const Particle *b) { return a->id < b->id; });
    std::shuffle(bin.begin(), bin.end(), gen_);
    const double w = W / static_cast<double>(target);
    const double r = dist(gen_);
    size_t k = 0;
    double cum = 0.0;
    for (const Particle *p : bin) {
      while (cum + p->weight > (static_cast<double>(k) + r) * w && k < target) {
        ++k;
        out.push_back(*p);
        out.back().id = std::numeric_limits<size_t>::max();
        out.back().cellIndex = c;
        out.back().timeLeft = 0;
        out.back().weight = out.back().initialWeight = w;
      }
      cum += p->weight;
    }
  }

  const Grid &grid_;
  std::vector<double> bounds_;
  size_t G_, Nmin_;
  double factor_;
  size_t minPer_;
  std::mt19937_64 gen_;
};

// ------------------------------------------------------------------ IMC physics
struct IMCParams {  // radiation/RadiationIMCParameters.hpp:12-56 (gray subset)
  size_t newPhotonsPerCell = 0;
  bool withRandomWalk = false;
  double rwMinCellOpticalDepth = 25.0;
  double rwMinParticleOpticalDepth = 5.0;
  bool withDDMC = false;
  double ddmcMinCellOpticalDepth = 15.0;
  double ddmcMinParticleOpticalDepth = 5.0;
  bool ddmcUseMultigroupPGRW = false;
  bool noHydroFeedback = false;
  bool withEgTimeAvg = false;  // per-group Erad_time_avg (multigroup only)
};

struct EnergyLedger {  // diagnostics for the per-step balance (not in the reference)
  double emitted = 0;     // sum of E_i removed from the material
  double source = 0;      // boundary source packets
  double escaped = 0;     // REMOVE at escape faces
  double clamp = 0;       // energy added by the negative-energy guard
  double censusIn = 0;    // census weight entering the step
  double censusOut = 0;   // census weight before population control
  double censusComb = 0;  // census weight after population control
  double materialBefore = 0, materialAfter = 0;
};

// Photons per cell of generateParticles (radiation/RadiationIMC.hpp:2076-2087):
// a share of N * n * 10 packets proportional to the cell's emission energy,
// clamped to [n, 20 n], and none where the energy is 0 (n everywhere when the
// total is 0).
inline std::vector<size_t> photonAllocation(const std::vector<double> &E, double total, size_t n) {
  const size_t N = E.size();
  const size_t totalParticles = N * n * 10;
  std::vector<size_t> out(N, 0);
  for (size_t i = 0; i < N; ++i) {
    size_t share = (total > 0) ? static_cast<size_t>(E[i] / total * totalParticles) : n;
    out[i] = E[i] > 0.0 ? std::max(n, std::min(share, n * 20)) : 0;
  }
  return out;
}

template <typename Grid>
class RadiationIMC {
 public:
  struct DDMCFaceLeak {
    size_t faceIndex;
    size_t nextCellIndex;
    double rate;
  };
  struct DDMCCellData {
    bool eligible = false;
    double sigmaA = 0.0, sigmaT = 0.0, diffusionCoefficient = 0.0, gamma = 1.0, totalLeakRate = 0.0;
    size_t groupCutoff = 0;
    std::vector<DDMCFaceLeak> faceLeaks;
  };

  RadiationIMC(const Grid &grid, FaceTableBoundary<Grid> *boundary, std::vector<Cell> &cells,
               std::vector<Extensives> &ext, const Material &mat, IMCParams params, uint64_t seed)
      : grid_(grid), boundary_(boundary), cells_(cells), ext_(ext), mat_(mat), params_(params),
        rng_(seed), sampler_(grid) {
    if (params_.newPhotonsPerCell == 0) throw OracleError("RadiationIMC requires newPhotonsPerCell > 0");
    if (params_.withRandomWalk) rw_ = std::make_unique<RandomWalk>();
    Erad_time_avg_.assign(grid.GetPointNo(), 0.0);
    updateGridData();
  }

  // MonteCarloPhysics::updateGridData (physics/MonteCarloPhysics.hpp:60-84)
  void updateGridData() {
    size_t N = grid_.GetPointNo();
    normals_.assign(N, {});
    facePts_.assign(N, {});
    for (size_t i = 0; i < N; i++) {
      const auto &faces = grid_.GetCellFaces(i);
      normals_[i].resize(faces.size());
      facePts_[i].resize(faces.size());
      for (size_t q = 0; q < faces.size(); ++q) {
        size_t f = faces[q];
        Vec3 n = grid_.Normal(f);
        if (dot(n, grid_.GetMeshPoint(i) - grid_.FaceCM(f)) < 0) n *= -1;
        normals_[i][q] = normalize(n);
        facePts_[i][q] = grid_.FaceCM(f);
      }
    }
  }

  EnergyLedger ledger;
  size_t rwStepCount = 0, ddmcStepCount = 0, ddmcLeakCount = 0;

  const std::vector<double> &factorFleck() const { return factorFleck_; }
  const std::vector<double> &planckOpacities() const { return planck_; }
  const std::vector<double> &eradTimeAvg() const { return Erad_time_avg_; }
  // getEgTimeAvg (RadiationIMC.hpp:433-434): [cell * G + g], empty unless
  // withEgTimeAvg and multigroup
  const std::vector<double> &egTimeAvg() const { return Eg_time_avg_; }
  const std::vector<size_t> &photonsPerCell() const { return nPhotons_; }
  const std::vector<double> &emissionEnergy() const { return Eemit_; }
  const RandomWalk *randomWalk() const { return rw_.get(); }
  const std::vector<DDMCCellData> &ddmcData() const { return ddmc_; }
  const std::vector<char> &rwEligible() const { return rwEligible_; }
  // one DDMC attempt on a packet (tryDDMCStep), for sampling its laws
  bool ddmcStep(Particle &p, StepResult &fn) { return tryDDMCStep(p, fn); }

  // preStep (radiation/RadiationIMC.hpp:1618-1694), gray, no hydro
  std::vector<Particle> preStep(double fullDt) {
    size_t N = grid_.GetPointNo();
    planck_.assign(N, 0.0);
    factorFleck_.assign(N, 1.0);
    Erad_time_avg_.assign(N, 0.0);
    if (egOn()) Eg_time_avg_.assign(N * (mat_.bounds.size() - 1), 0.0);  // :1626-1630
    for (size_t i = 0; i < N; ++i) {
      Cell &cell = cells_[i];
      planck_[i] = mat_.CalcPlanckOpacity(i, cell.temperature);
      double cv = mat_.dT2cv(mat_.density(i), cell.temperature);
      double T = cell.temperature;
      factorFleck_[i] = 1.0 / (1.0 + (4.0 * units::arad * (T * T * T) * planck_[i] * units::clight * fullDt) / cv);
      if (factorFleck_[i] < 0.0 || factorFleck_[i] > 1.0) throw OracleError("Invalid factor fleck in RadiationIMC::preStep");
    }
    if (mat_.multigroup()) {
      thermalCdf_.assign(N, {});
      std::vector<double> Bg;
      for (size_t i = 0; i < N; ++i) mat_.cumulativePlanck(i, cells_[i].temperature, thermalCdf_[i], Bg);
    }
    if (params_.withRandomWalk || params_.withDDMC) updateGridData();
    if (params_.withRandomWalk) {
      precomputeRandomWalkData();
      rwStepCount = 0;
    }
    if (params_.withDDMC) {
      precomputeDDMCData();
      ddmcStepCount = ddmcLeakCount = 0;
    }
    std::vector<Particle> np = generateParticles(fullDt);
    if (boundary_) {
      std::vector<Particle> bp = boundary_->generateNewBoundaryParticles(fullDt);
      for (Particle &p : bp) p.initialWeight = std::abs(p.weight);
      np.insert(np.end(), bp.begin(), bp.end());
    }
    return np;
  }

  // generateParticles (radiation/RadiationIMC.hpp:2035-2510) with the
  // adaptive / learned-score options off (their defaults).
  std::vector<Particle> generateParticles(double fullDt) {
    std::vector<Particle> out;
    size_t N = grid_.GetPointNo();
    Eemit_.assign(N, 0.0);
    double total = 0.0;
    for (size_t i = 0; i < N; ++i) {
      double T = cells_[i].temperature;
      Eemit_[i] = factorFleck_[i] * grid_.GetVolume(i) * units::arad * (T * T * T * T) * planck_[i] * fullDt * units::clight;
      total += Eemit_[i];
    }
    nPhotons_ = photonAllocation(Eemit_, total, params_.newPhotonsPerCell);
    for (size_t i = 0; i < N; ++i) {
      size_t n = nPhotons_[i];
      if (n == 0) continue;
      double E = Eemit_[i];
      if (!params_.noHydroFeedback) ext_[i].internal_energy -= E;
      ledger.emitted += E;
      double perPhoton = E / n;
      for (size_t j = 0; j < n; ++j) {
        Particle p = generateSingleParticle(i);
        p.timeLeft = fullDt * rng_.unitOpen();
        if (mat_.multigroup()) p.frequency = mat_.thermalEnergyFromCumulative(thermalCdf_[i], rng_.unitOpen());
        p.weight = perPhoton;
        p.initialWeight = std::abs(p.weight);
        out.push_back(p);
      }
    }
    return out;
  }

  // generateSingleParticle (radiation/RadiationIMC.hpp:2516-2553)
  Particle generateSingleParticle(size_t i) {
    Particle p;
    p.id = std::numeric_limits<size_t>::max();
    p.cellIndex = i;
    p.frequency = 0.0;
    p.location = sampler_(i, rng_);
    if (grid_.IsPointOutsideBox(p.location)) {
      Vec3 mp = grid_.GetMeshPoint(i), original = p.location, dir = mp - original;
      double t = 1e-8;
      while (grid_.IsPointOutsideBox(p.location) && t < 1.0) {
        p.location = original + t * dir;
        t *= 2;
      }
      p.location = p.location + 1e-8 * (mp - p.location);
    }
    p.velocity = isotropicVelocity(rng_);
    p.timeLeft = 0.0;
    p.steps = 0;
    return p;
  }

  // Particle::distanceToNearestFace (particle/Particle.hpp:188-279) via
  // getIntersectionDetails (physics/MonteCarloPhysics.hpp:86-99)
  void intersection(const Particle &p, size_t &face, double &time, size_t &next) const {
    double minAlpha = std::numeric_limits<double>::max();
    size_t minFace = std::numeric_limits<size_t>::max();
    double velAbs = kEpsilonParticle * norm(p.velocity);
    const auto &faces = grid_.GetCellFaces(p.cellIndex);
    for (size_t q = 0; q < normals_[p.cellIndex].size(); ++q) {
      const Vec3 &n = normals_[p.cellIndex][q];
      double nv = dot(n, p.velocity);
      if (nv >= -velAbs) continue;
      double alpha = dot(facePts_[p.cellIndex][q] - p.location, n) / nv;
      if (alpha < minAlpha && alpha > 0) {
        minAlpha = alpha;
        minFace = faces[q];
      }
    }
    if (minAlpha == std::numeric_limits<double>::max()) {
      if (grid_.IsPointOutsideBox(p.location)) throw OracleError("Particle is outside the domain, but still considered");
      throw OracleError("distanceToNearestFace: no face intersection found");
    }
    auto nb = grid_.GetFaceNeighbors(minFace);
    face = minFace;
    time = minAlpha;
    next = (nb.first == p.cellIndex) ? nb.second : nb.first;
  }

  // RadiationIMC::step (radiation/RadiationIMC.hpp:1700-1881), gray, no hydro
  StepResult step(Particle &p) {
    StepResult fn;
    size_t c = p.cellIndex;
    size_t faceI, nextC;
    double tFace;
    intersection(p, faceI, tFace, nextC);
    if (params_.withRandomWalk && rw_ && rwEligible_[c]) {
      if (tryRandomWalkStep(p, fn)) {
        ++rwStepCount;
        return fn;
      }
    }
    if (params_.withDDMC && ddmc_[c].eligible) {
      if (tryDDMCStep(p, fn)) return fn;
    }
    double sa = planck_[c];
    if (mat_.multigroup()) {
      // clampFrequencyToBounds (RadiationIMC.hpp:868-872), CalcAbsorptionOpacity (:1753)
      double E = std::clamp(p.frequency, mat_.bounds.front(), mat_.bounds.back());
      sa = mat_.CalcAbsorptionOpacity(c, cells_[c].temperature, E);
    }
    double ss = mat_.CalcScatteringOpacity(c, cells_[c].temperature);
    double eff = (1.0 - factorFleck_[c]) * sa;
    double eventOp = ss + eff;
    double scatLen = (eventOp > 0.0) ? 1.0 / eventOp : std::numeric_limits<double>::infinity();
    double l1p = -std::log1p(rng_.unitOpen() - 1.0);
    double scatDist = scatLen * l1p;
    double tScat = std::isfinite(scatDist) ? scatDist / norm(p.velocity) : std::numeric_limits<double>::infinity();
    double tLeft = p.timeLeft;
    int ev = 0;
    double dt = tFace;
    if (tScat < dt) { ev = 1; dt = tScat; }
    if (tLeft < dt) { ev = 2; dt = tLeft; }
    p.timeLeft -= dt;
    double tmp2 = sa * factorFleck_[c] * units::clight;
    double tmp = -dt * tmp2;
    double e1 = std::expm1(tmp);
    double e2 = std::expm1(tmp);
    double integrated = p.weight * dt;
    if (std::abs(tmp2 * dt) >= 1e-12) integrated = p.weight * e2 * (-1.0 / tmp2);
    p.location += p.velocity * dt;
    if (!params_.noHydroFeedback) ext_[c].internal_energy += -e2 * p.weight;
    Erad_time_avg_[c] += integrated;
    egTally(c, p.frequency, integrated);  // :1802-1809
    p.weight *= 1.0 + e1;
    if (std::abs(p.weight) < p.initialWeight * 1e-3) {
      fn.change = REMOVE;
      if (!params_.noHydroFeedback) ext_[c].internal_energy += p.weight;
      return fn;
    }
    if (ev == 0) {
      fn.change = CELL_MOVE;
      fn.nextCellIndex = nextC;
    } else if (ev == 1) {
      double eventRandom = rng_.unitOpen() * eventOp;
      // both branches redirect isotropically (RadiationIMC.hpp:1834-1842); an
      // effective scatter re-emits at a thermal energy (:1843-1852)
      const bool effective = !(eventRandom < ss);
      p.velocity = isotropicVelocity(rng_);
      if (mat_.multigroup()) {
        p.frequency = std::clamp(p.frequency, mat_.bounds.front(), mat_.bounds.back());
        if (effective) p.frequency = mat_.thermalEnergyFromCumulative(thermalCdf_[c], rng_.unitOpen());
      }
      fn.change = NO_CELL_MOVE;
    } else {
      fn.change = DONE;
    }
    return fn;
  }

  // postStep (radiation/RadiationIMC.hpp:1887-2029), no hydro
  void postStep(const std::vector<Particle> &census, double fullDt) {
    size_t N = grid_.GetPointNo();
    for (size_t i = 0; i < N; ++i) Erad_time_avg_[i] /= (fullDt * grid_.GetVolume(i));
    if (egOn()) {  // :1896-1903
      const size_t G = mat_.bounds.size() - 1;
      for (size_t i = 0; i < N; ++i)
        for (size_t g = 0; g < G; ++g) Eg_time_avg_[i * G + g] /= fullDt * grid_.GetVolume(i);
    }
    if (!params_.noHydroFeedback) {
      for (size_t i = 0; i < N; ++i) {
        throwIfNegativeInternalEnergy(i);
        synchronizeMaterialCell(i);
      }
    }
    for (size_t i = 0; i < N; ++i) ext_[i].Erad = 0.0;
    for (const Particle &p : census) ext_[p.cellIndex].Erad += p.weight;
    for (size_t i = 0; i < N; ++i) {
      double m = ext_[i].mass;
      cells_[i].Erad = (m > 0.0) ? ext_[i].Erad / m : 0.0;
    }
  }

 private:
  // throwIfNegativeInternalEnergy (radiation/RadiationIMC.hpp:771-817) with
  // RadiationCell (which has internalEnergy, the previous step's value)
  void throwIfNegativeInternalEnergy(size_t i) {
    double &E = ext_[i].internal_energy;
    if (E >= 0.0) return;
    double T = cells_[i].temperature;
    double thermal = units::arad * (T * T * T * T) * grid_.GetVolume(i);
    if (thermal < 1e-30) thermal = 1e-30;
    if (std::abs(E) / thermal < 0.1) {
      double newE = std::max(0.0, cells_[i].internalEnergy);
      ledger.clamp += newE - E;
      E = newE;
      return;
    }
    throw OracleError("Negative material internal energy in RadiationIMC");
  }
  // synchronizeMaterialCell (radiation/RadiationIMC.hpp:830-866)
  void synchronizeMaterialCell(size_t i) {
    Cell &cell = cells_[i];
    double rho = mat_.density(i);
    double specific = ext_[i].internal_energy / ext_[i].mass;
    cell.internalEnergy = ext_[i].internal_energy;
    cell.temperature = mat_.de2T(rho, specific);
  }

  double minDistanceToFaces(size_t c, const Vec3 &x) const {  // RadiationIMC.hpp:956-969
    double Ro = std::numeric_limits<double>::max();
    for (size_t q = 0; q < normals_[c].size(); ++q) Ro = std::min(Ro, dot(x - facePts_[c][q], normals_[c][q]));
    return (Ro > 0.0) ? Ro : 0.0;
  }
  double surfaceArea(size_t c) const {
    double s = 0.0;
    for (size_t f : grid_.GetCellFaces(c)) s += grid_.GetArea(f);
    return s;
  }

  // precomputeRandomWalkData (radiation/RadiationIMC.hpp:982-1063): gray, and
  // the partial-gray random walk of multigroup runs (:1008-1060): the groups
  // from the lowest up, as long as each is optically thick over the cell's
  // mean chord, form the diffusive part; their Planck-weighted opacities
  // drive the walk.
  void precomputeRandomWalkData() {
    size_t N = grid_.GetPointNo();
    rwEligible_.assign(N, 0);
    rwSigmaT_.assign(N, 0.0);
    if (mat_.multigroup()) pgrw_.assign(N, PGRWCellData{});
    for (size_t i = 0; i < N; ++i) {
      double ss = mat_.CalcScatteringOpacity(i, cells_[i].temperature);
      double sT = planck_[i] + ss;
      rwSigmaT_[i] = sT;
      double S = surfaceArea(i), V = grid_.GetVolume(i);
      double chord = (S > 0.0) ? 4.0 * V / S : 0.0;
      if (!mat_.multigroup()) {
        rwEligible_[i] = (sT * chord >= params_.rwMinCellOpticalDepth) ? 1 : 0;
        continue;
      }
      const size_t G = mat_.groups();
      const double kT = units::k_boltz * cells_[i].temperature;
      double allAB = 0.0, sB = 0.0, sBA = 0.0, sBT = 0.0, sBoT = 0.0;
      size_t cutoff = 0;
      bool nonDiffusive = false;
      for (size_t g = 0; g < G; ++g) {
        const double a = mat_.bounds[g] / kT, b = mat_.bounds[g + 1] / kT;
        const double Bg = (a > 0.0 && b > a) ? storm::planck::integral(a, b) : 1.0 / static_cast<double>(G);
        const double sA = mat_.CalcAbsorptionOpacity(i, cells_[i].temperature,
                                                     0.5 * (mat_.bounds[g] + mat_.bounds[g + 1]));
        const double sTg = sA + ss;
        allAB += sA * Bg;
        if (!nonDiffusive && sTg * chord >= params_.rwMinCellOpticalDepth) {
          cutoff = g + 1;
          sB += Bg;
          sBA += Bg * sA;
          sBT += Bg * sTg;
          if (sTg > 0.0) sBoT += Bg / sTg;
        } else {
          nonDiffusive = true;
        }
      }
      if (cutoff > 0 && sB > 0.0) {
        PGRWCellData &d = pgrw_[i];
        d.groupCutoff = cutoff;
        d.sigmaA = sBA / sB;
        d.sigmaT = sBT / sB;
        d.D = (units::clight / 3.0) * sBoT / sB;
        d.gamma = (allAB > 0.0) ? sBA / allAB : 1.0;
        rwSigmaT_[i] = d.sigmaT;
        rwEligible_[i] = 1;
      }
    }
  }

  // tryRandomWalkStep (radiation/RadiationIMC.hpp:1065-1276), gray, no hydro
  bool tryRandomWalkStep(Particle &p, StepResult &fn) {
    size_t c = p.cellIndex;
    double Ro = minDistanceToFaces(c, p.location);
    const bool pg = mat_.multigroup();
    double sigmaT = rwSigmaT_[c];
    double sa = planck_[c];
    double D = (sigmaT > 0.0) ? units::clight / (3.0 * sigmaT) : 0.0;
    double gamma = 1.0;
    size_t cutoff = 0;
    if (pg) {
      const PGRWCellData &d = pgrw_[c];
      sigmaT = d.sigmaT;
      sa = d.sigmaA;
      D = d.D;
      gamma = d.gamma;
      cutoff = d.groupCutoff;
    }
    bool doRW = (Ro > 0.0 && sigmaT > 0.0 && D > 0.0 && Ro * sigmaT >= params_.rwMinParticleOpticalDepth);
    if (doRW && pg) {
      // only packets in the diffusive groups walk (:1094-1114)
      const double E = std::clamp(p.frequency, mat_.bounds.front(), mat_.bounds.back());
      if (E >= mat_.bounds[cutoff]) doRW = false;
    }
    if (!doRW) return false;
    double f = factorFleck_[c];
    double tauLeak = rw_->sampleLeakTime(rng_.unitOpen());
    double tLeak = tauLeak * Ro * Ro / D;
    double tCensus = p.timeLeft;
    double tUp = std::numeric_limits<double>::max();  // gray: no upscatter
    if (pg && gamma < 1.0 && sa > 0.0 && f > 0.0) {
      // absorption re-emitted above the diffusive groups ends the walk (:1126-1131)
      tUp = -std::log(rng_.unitOpen()) / (units::clight * (1.0 - f) * sa * (1.0 - gamma));
    }
    int ev;
    double dt;
    if (tLeak <= tCensus && tLeak <= tUp) { ev = 0; dt = tLeak; }
    else if (tCensus <= tUp) { ev = 1; dt = tCensus; }
    else { ev = 2; dt = tUp; }
    double rate = sa * f * units::clight;
    double ex = std::expm1(-dt * rate);
    if (!params_.noHydroFeedback) ext_[c].internal_energy += -ex * p.weight;
    if (rate > 0.0) {
      Erad_time_avg_[c] += p.weight * ex * (-1.0 / rate);
      egTally(c, p.frequency, p.weight * ex * (-1.0 / rate));  // :1166-1173
    }
    p.weight *= 1.0 + ex;
    p.timeLeft -= dt;
    if (std::abs(p.weight) < p.initialWeight * 1e-4) {
      fn.change = REMOVE;
      if (!params_.noHydroFeedback) ext_[c].internal_energy += p.weight;
      return true;
    }
    constexpr double RW_PI = 3.14159265358979323846;
    double cosT = 2.0 * rng_.unitOpen() - 1.0;
    double sinT = std::sqrt(std::max(0.0, 1.0 - cosT * cosT));
    double phi = 2.0 * RW_PI * rng_.unitOpen();
    Vec3 posDir(sinT * std::cos(phi), sinT * std::sin(phi), cosT);
    double disp;
    if (ev == 0) {
      disp = Ro;
    } else {
      double tauPos = D * dt / (Ro * Ro);
      disp = Ro * rw_->sampleRadius(tauPos, rng_.unitOpen());
    }
    if (disp > Ro * (1.0 + 1e-12)) disp = Ro;
    Vec3 center = p.location;
    p.location = center + disp * posDir;
    constexpr double nudge = 1e-6;
    p.location = p.location * (1.0 - nudge) + nudge * grid_.GetMeshPoint(c);
    for (size_t q = 0; q < normals_[c].size(); ++q) {
      double d = dot(p.location - facePts_[c][q], normals_[c][q]);
      if (d < 0.0) {
        disp *= 0.99;
        p.location = center + disp * posDir;
        q = static_cast<size_t>(-1);
      }
    }
    p.velocity = isotropicVelocity(rng_);
    if (ev == 2) {
      // re-emission above the cutoff (:1237-1252): the cell's emission
      // spectrum restricted to the non-diffusive groups
      const std::vector<double> &cdf = thermalCdf_[c];
      const size_t G = mat_.groups();
      const double atCut = cdf[cutoff], total = cdf[G];
      if (total > atCut) {
        const double lo = atCut / total;
        p.frequency = mat_.thermalEnergyFromCumulative(cdf, lo + rng_.unitOpen() * (1.0 - lo));
      } else {
        p.frequency = std::nextafter(mat_.bounds[cutoff], std::numeric_limits<double>::max());
      }
      p.frequency = std::clamp(p.frequency, mat_.bounds.front(), mat_.bounds.back());
    }
    if (ev == 1) fn.change = DONE;
    return true;
  }

  // precomputeDDMCData (radiation/RadiationIMC.hpp:1282-1413) without
  // ddmcUseMultigroupPGRW (its default, RadiationIMCParameters.hpp:26): the
  // Planck mean opacity, also in a multigroup problem (:1343-1350)
  void precomputeDDMCData() {
    size_t N = grid_.GetPointNo();
    ddmc_.assign(N, DDMCCellData{});
    for (size_t i = 0; i < N; ++i) {
      DDMCCellData &d = ddmc_[i];
      double ss = mat_.CalcScatteringOpacity(i, cells_[i].temperature);
      double V = grid_.GetVolume(i), S = surfaceArea(i);
      if (V <= 0.0 || S <= 0.0) continue;
      double chord = 4.0 * V / S;
      if (params_.ddmcUseMultigroupPGRW && mat_.multigroup()) {
        // (:1301-1342) the diffusive groups from the lowest up, equal
        // weights 1/G; D the harmonic mean over them
        const double kT = units::k_boltz * cells_[i].temperature;
        if (kT <= 0.0) continue;
        const size_t G = mat_.groups();
        double sB = 0.0, sBA = 0.0, sBT = 0.0, sBoT = 0.0;
        size_t cutoff = 0;
        bool nonDiffusive = false;
        for (size_t g = 0; g < G; ++g) {
          const double sA = mat_.CalcAbsorptionOpacity(i, cells_[i].temperature,
                                                       0.5 * (mat_.bounds[g] + mat_.bounds[g + 1]));
          const double sTg = sA + ss;
          const double Bg = 1.0 / static_cast<double>(G);
          if (!nonDiffusive && sTg * chord >= params_.ddmcMinCellOpticalDepth) {
            cutoff = g + 1;
            sB += Bg;
            sBA += Bg * sA;
            sBT += Bg * sTg;
            if (sTg > 0.0) sBoT += Bg / sTg;
          } else {
            nonDiffusive = true;
          }
        }
        if (cutoff > 0 && sB > 0.0) {
          d.groupCutoff = cutoff;
          d.sigmaA = sBA / sB;
          d.sigmaT = sBT / sB;
          d.diffusionCoefficient = (units::clight / 3.0) * sBoT / sB;
          d.gamma = 1.0;
          d.eligible = (d.sigmaT > 0.0 && d.diffusionCoefficient > 0.0);
        }
        continue;
      }
      d.sigmaA = planck_[i];
      d.sigmaT = d.sigmaA + ss;
      d.diffusionCoefficient = (d.sigmaT > 0.0) ? units::clight / (3.0 * d.sigmaT) : 0.0;
      d.gamma = 1.0;
      d.eligible = (d.sigmaT * chord >= params_.ddmcMinCellOpticalDepth && d.diffusionCoefficient > 0.0);
    }
    for (size_t i = 0; i < N; ++i) {
      DDMCCellData &d = ddmc_[i];
      if (!d.eligible) continue;
      double V = grid_.GetVolume(i);
      Vec3 cc = grid_.GetMeshPoint(i);
      for (size_t f : grid_.GetCellFaces(i)) {
        auto nb = grid_.GetFaceNeighbors(f);
        size_t next = (nb.first == i) ? nb.second : nb.first;
        if (grid_.IsPointOutsideBox(next)) continue;
        Vec3 n = grid_.Normal(f);
        double nm = norm(n);
        if (nm <= 0.0) continue;
        n = n / nm;
        Vec3 fc = grid_.FaceCM(f);
        double fd = std::abs(dot(fc - cc, n));
        if (fd <= 0.0) fd = 0.5 * std::abs(dot(grid_.GetMeshPoint(next) - cc, n));
        if (fd <= 0.0) continue;
        double Df = d.diffusionCoefficient;
        if (next < N && ddmc_[next].diffusionCoefficient > 0.0) {
          double D1 = d.diffusionCoefficient, D2 = ddmc_[next].diffusionCoefficient;
          Df = (D1 > 0.0 && D2 > 0.0) ? (2.0 * D1 * D2 / (D1 + D2)) : std::max(D1, D2);
        }
        double rate = Df * grid_.GetArea(f) / (V * fd);
        if (rate > 0.0 && std::isfinite(rate)) {
          d.faceLeaks.push_back({f, next, rate});
          d.totalLeakRate += rate;
        }
      }
      if (d.totalLeakRate <= 0.0) {
        d.eligible = false;
        d.faceLeaks.clear();
      }
    }
  }

  // tryDDMCStep (radiation/RadiationIMC.hpp:1415-1612) without
  // ddmcUseMultigroupPGRW: Planck-mean absorption, the photon energy is kept
  bool tryDDMCStep(Particle &p, StepResult &fn) {
    size_t c = p.cellIndex;
    const DDMCCellData &d = ddmc_[c];
    if (!d.eligible || d.totalLeakRate <= 0.0 || d.faceLeaks.empty()) return false;
    double Ro = minDistanceToFaces(c, p.location);
    if (Ro * d.sigmaT < params_.ddmcMinParticleOpticalDepth) return false;
    if (params_.ddmcUseMultigroupPGRW && mat_.multigroup()) {
      // (:1437-1450) only packets below the cutoff group
      if (d.groupCutoff == 0 || d.groupCutoff > mat_.groups()) return false;
      const double coFreq = std::clamp(p.frequency, mat_.bounds.front(), mat_.bounds.back());
      if (coFreq >= mat_.bounds[d.groupCutoff]) return false;
    }
    double f = factorFleck_[c];
    // gamma = 1 in both forms of the data, so the upscatter rate
    // c (1 - f) sigma_A (1 - gamma) of :1455-1458 is zero
    double eventRate = d.totalLeakRate;
    double tEvent = -std::log(rng_.unitOpen()) / eventRate;
    double tCensus = p.timeLeft;
    double dt = std::min(tEvent, tCensus);
    bool census = (tCensus <= tEvent);
    double absRate = d.sigmaA * f * units::clight;
    double oldW = p.weight;
    double ex = std::expm1(-dt * absRate);
    if (!params_.noHydroFeedback) ext_[c].internal_energy += -ex * oldW;
    double integ = (absRate > 0.0) ? oldW * ex * (-1.0 / absRate) : oldW * dt;
    Erad_time_avg_[c] += integ;
    egTally(c, p.frequency, integ);  // :1485-1492
    p.weight *= 1.0 + ex;
    p.timeLeft -= dt;
    if (std::abs(p.weight) < p.initialWeight * 1e-3) {
      fn.change = REMOVE;
      if (!params_.noHydroFeedback) ext_[c].internal_energy += p.weight;
      ++ddmcStepCount;
      return true;
    }
    ++ddmcStepCount;
    if (census) {
      fn.change = DONE;
      return true;
    }
    double eventPick = rng_.unitOpen() * eventRate;
    if (eventPick <= d.totalLeakRate) {
      double facePick = rng_.unitOpen() * d.totalLeakRate;
      const DDMCFaceLeak *chosen = nullptr;
      for (const DDMCFaceLeak &fl : d.faceLeaks) {
        facePick -= fl.rate;
        if (facePick <= 0.0) {
          chosen = &fl;
          break;
        }
      }
      if (!chosen) chosen = &d.faceLeaks.back();
      Vec3 lc = grid_.FaceCM(chosen->faceIndex);
      Vec3 nOut = grid_.Normal(chosen->faceIndex);
      nOut = nOut / norm(nOut);
      Vec3 toward = grid_.GetMeshPoint(chosen->nextCellIndex) - grid_.GetMeshPoint(c);
      if (dot(nOut, toward) < 0.0) nOut = nOut * (-1.0);
      constexpr double DDMC_PI = 3.14159265358979323846;
      double mu = std::sqrt(std::min(1.0, rng_.unitOpen()));
      double phi = 2.0 * DDMC_PI * rng_.unitOpen();
      double sinT = std::sqrt(std::max(0.0, 1.0 - mu * mu));
      Vec3 helper = (std::abs(nOut.x) < 0.9) ? Vec3(1.0, 0.0, 0.0) : Vec3(0.0, 1.0, 0.0);
      Vec3 e1 = helper - dot(helper, nOut) * nOut;
      double e1m = norm(e1);
      if (e1m > 0.0) e1 = e1 / e1m;
      Vec3 e2 = cross(nOut, e1);
      double e2m = norm(e2);
      if (e2m > 0.0) e2 = e2 / e2m;
      Vec3 dir = mu * nOut + sinT * std::cos(phi) * e1 + sinT * std::sin(phi) * e2;
      double dm = norm(dir);
      if (dm > 0.0) dir = dir / dm;
      p.location = lc;
      p.velocity = dir * units::clight;
      fn.change = CELL_MOVE;
      fn.nextCellIndex = chosen->nextCellIndex;
      ++ddmcLeakCount;
    } else {
      fn.change = DONE;
    }
    return true;
  }

  const Grid &grid_;
  FaceTableBoundary<Grid> *boundary_;
  std::vector<Cell> &cells_;
  std::vector<Extensives> &ext_;
  const Material &mat_;
  IMCParams params_;
  Rng rng_;
  RandomInCellSampler<Grid> sampler_;
  std::vector<std::vector<Vec3>> normals_, facePts_;
  std::vector<double> factorFleck_, planck_, Erad_time_avg_, Eemit_, Eg_time_avg_;
  bool egOn() const { return params_.withEgTimeAvg && mat_.multigroup(); }
  // RadiationOpacityModel::findGroup (RadiationOpacityModel.hpp:76-86): the
  // first group whose upper bound lies above the energy, else the last
  size_t findGroup(double E) const {
    const size_t G = mat_.bounds.size() - 1;
    for (size_t g = 0; g < G; ++g)
      if (E < mat_.bounds[g + 1]) return g;
    return G - 1;
  }
  void egTally(size_t c, double E, double v) {
    if (egOn()) Eg_time_avg_[c * (mat_.bounds.size() - 1) + findGroup(E)] += v;
  }
  std::vector<size_t> nPhotons_;
  std::unique_ptr<RandomWalk> rw_;
  std::vector<char> rwEligible_;
  std::vector<double> rwSigmaT_;
  // partial-gray random walk data (PGRWCellData, RadiationIMC.hpp:1044-1051)
  struct PGRWCellData {
    size_t groupCutoff = 0;
    double sigmaA = 0.0, sigmaT = 0.0, D = 0.0, gamma = 1.0;
  };
  std::vector<PGRWCellData> pgrw_;
  std::vector<DDMCCellData> ddmc_;
  std::vector<std::vector<double>> thermalCdf_;
};

// ------------------------------------------------------------------ manager
// MonteCarloManagerSerial (manager/MonteCarloManagerSerial.hpp:25-469).  The
// free-list buffers (av/th) only decide memory reuse; the history loop is
// restated over a plain vector in the same order.
template <typename Grid>
class Manager {
 public:
  Manager(const Grid &g, RadiationIMC<Grid> &phys, PopulationControl *comb, FaceTableBoundary<Grid> &bc)
      : grid_(g), phys_(phys), comb_(comb), bc_(bc) {}

  std::vector<size_t> cellsStepsCounters;
  size_t startParticleCount = 0, initialParticleCount = 0, preStepParticleCount = 0, endParticleCount = 0;
  size_t leavingCount = 0;

  std::vector<Particle> step(std::vector<Particle> &&census, double dt) {
    size_t N = grid_.GetPointNo();
    std::vector<Particle> active = std::move(census);
    for (Particle &p : active) {
      if (p.id == std::numeric_limits<size_t>::max()) p.id = idCounter_++;
      p.timeLeft = dt;
      p.initialWeight = std::abs(p.weight);
      p.steps = 0;
    }
    initialParticleCount = active.size();
    std::vector<Particle> np = phys_.preStep(dt);
    for (Particle &p : np) p.id = idCounter_++;
    preStepParticleCount = np.size();
    active.insert(active.end(), np.begin(), np.end());
    startParticleCount = active.size();
    cellsStepsCounters.assign(N, 0);
    leavingCount = 0;
    std::vector<Particle> remaining;
    for (Particle &p : active) {  // HandleAll (MonteCarloManagerSerial.hpp:266-383)
      while (true) {
        p.steps++;
        cellsStepsCounters[p.cellIndex]++;
        StepResult r = phys_.step(p);
        if (r.change == CELL_MOVE) {
          size_t next = r.nextCellIndex;
          if (next < N) {
            p.location = (1 - kMonteCarloEpsilon) * p.location + kMonteCarloEpsilon * grid_.GetMeshPoint(next);
            p.cellIndex = next;
          } else {
            Status s = bc_.apply(p);
            if (s == REFLECT) {
              p.location = (1 - kMonteCarloEpsilon) * p.location + kMonteCarloEpsilon * grid_.GetMeshPoint(p.cellIndex);
            } else if (s == REMOVE) {
              leavingCount++;
              break;
            } else {
              throw OracleError("Unknown boundary condition for particle");
            }
            continue;
          }
        } else if (r.change == REMOVE) {
          break;
        } else if (r.change == DONE) {
          remaining.push_back(p);
          break;
        }
      }
    }
    std::vector<Particle> out = comb_ ? comb_->activate(remaining) : remaining;
    endParticleCount = out.size();
    lastRemaining_ = std::move(remaining);
    phys_.postStep(out, dt);
    return out;
  }

  const std::vector<Particle> &lastPreCombCensus() const { return lastRemaining_; }

 private:
  const Grid &grid_;
  RadiationIMC<Grid> &phys_;
  PopulationControl *comb_;
  FaceTableBoundary<Grid> &bc_;
  size_t idCounter_ = 0;
  std::vector<Particle> lastRemaining_;
};

}  // namespace storm_oracle
