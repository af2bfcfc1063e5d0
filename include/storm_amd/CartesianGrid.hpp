// A Cartesian mesh with the Grid methods the IMC path calls.
//
// The reference takes its mesh from the external MadCart library
// (MadCart::CartesianMesh3D, examples/serial_cartesian/main.cpp:13-16) which
// is not vendored; this class provides the same Grid concept for callers
// that have no MadCart: cells (i*ny + j)*nz + k
// (examples/cartesian_parallel_check/main.cpp:54), six faces per cell,
// outside "mirror" points with indices >= N, box faces with vertices.
#pragma once

#include <algorithm>
#include <array>
#include <cstddef>
#include <stdexcept>
#include <utility>
#include <vector>

namespace STORM {

template <typename PointT>
class CartesianGrid {
 public:
  using Point_T = PointT;
  struct Face_T {
    std::vector<PointT> vertices;
  };

  // CartesianMesh3D(ll, ur, nx, ny, nz)
  CartesianGrid(const PointT &ll, const PointT &ur, size_t nx, size_t ny, size_t nz)
      : CartesianGrid(edges(ll.x, ur.x, nx), edges(ll.y, ur.y, ny), edges(ll.z, ur.z, nz)) {}

  // CartesianMesh3D(xEdges, yLo, yHi, ny, zLo, zHi, nz)
  CartesianGrid(const std::vector<double> &xEdges, double yLo, double yHi, size_t ny, double zLo, double zHi,
                size_t nz)
      : CartesianGrid(xEdges, edges(yLo, yHi, ny), edges(zLo, zHi, nz)) {}

  CartesianGrid(std::vector<double> xe, std::vector<double> ye, std::vector<double> ze)
      : xe_(std::move(xe)), ye_(std::move(ye)), ze_(std::move(ze)) {
    nx_ = xe_.size() - 1;
    ny_ = ye_.size() - 1;
    nz_ = ze_.size() - 1;
    if (xe_.size() < 2 || ye_.size() < 2 || ze_.size() < 2) throw std::invalid_argument("CartesianGrid: empty axis");
    N_ = nx_ * ny_ * nz_;
    Fx_ = (nx_ + 1) * ny_ * nz_;
    Fy_ = nx_ * (ny_ + 1) * nz_;
    Fz_ = nx_ * ny_ * (nz_ + 1);
    ll_ = PointT(xe_.front(), ye_.front(), ze_.front());
    ur_ = PointT(xe_.back(), ye_.back(), ze_.back());
    verts_.resize((nx_ + 1) * (ny_ + 1) * (nz_ + 1));
    for (size_t i = 0; i <= nx_; ++i)
      for (size_t j = 0; j <= ny_; ++j)
        for (size_t k = 0; k <= nz_; ++k) verts_[vid(i, j, k)] = PointT(xe_[i], ye_[j], ze_[k]);
    cellFaces_.resize(N_);
    for (size_t i = 0; i < nx_; ++i)
      for (size_t j = 0; j < ny_; ++j)
        for (size_t k = 0; k < nz_; ++k)
          cellFaces_[cell(i, j, k)] = {fx(i, j, k), fx(i + 1, j, k), fy(i, j, k), fy(i, j + 1, k), fz(i, j, k),
                                       fz(i, j, k + 1)};
    const double X[2] = {ll_.x, ur_.x}, Y[2] = {ll_.y, ur_.y}, Z[2] = {ll_.z, ur_.z};
    boxFaces_.resize(6);
    for (int s = 0; s < 2; ++s) {
      boxFaces_[s].vertices = {PointT(X[s], Y[0], Z[0]), PointT(X[s], Y[1], Z[0]), PointT(X[s], Y[0], Z[1]),
                               PointT(X[s], Y[1], Z[1])};
      boxFaces_[2 + s].vertices = {PointT(X[0], Y[s], Z[0]), PointT(X[1], Y[s], Z[0]), PointT(X[0], Y[s], Z[1]),
                                   PointT(X[1], Y[s], Z[1])};
      boxFaces_[4 + s].vertices = {PointT(X[0], Y[0], Z[s]), PointT(X[1], Y[0], Z[s]), PointT(X[0], Y[1], Z[s]),
                                   PointT(X[1], Y[1], Z[s])};
    }
  }

  size_t GetPointNo() const { return N_; }
  std::pair<PointT, PointT> GetBoxCoordinates() const { return {ll_, ur_}; }
  size_t GetTotalFacesNumber() const { return Fx_ + Fy_ + Fz_; }
  size_t GetBuildGeneration() const { return 0; }
  const std::vector<double> &xEdges() const { return xe_; }
  const std::vector<double> &yEdges() const { return ye_; }
  const std::vector<double> &zEdges() const { return ze_; }

  PointT GetCellCM(size_t c) const {
    size_t i, j, k;
    ijk(c, i, j, k);
    return PointT(0.5 * (xe_[i] + xe_[i + 1]), 0.5 * (ye_[j] + ye_[j + 1]), 0.5 * (ze_[k] + ze_[k + 1]));
  }
  double GetVolume(size_t c) const {
    size_t i, j, k;
    ijk(c, i, j, k);
    return (xe_[i + 1] - xe_[i]) * (ye_[j + 1] - ye_[j]) * (ze_[k + 1] - ze_[k]);
  }
  PointT GetMeshPoint(size_t p) const {
    if (p < N_) return GetCellCM(p);
    const size_t f = p - N_;
    auto nb = GetFaceNeighbors(f);
    const PointT c = GetCellCM(nb.first < N_ ? nb.first : nb.second);
    const PointT fc = FaceCM(f);
    return PointT(2 * fc.x - c.x, 2 * fc.y - c.y, 2 * fc.z - c.z);
  }
  bool IsPointOutsideBox(size_t p) const { return p >= N_; }
  bool IsPointOutsideBox(const PointT &p) const {
    return p.x < ll_.x || p.x > ur_.x || p.y < ll_.y || p.y > ur_.y || p.z < ll_.z || p.z > ur_.z;
  }
  bool IsPointInCell(const PointT &p, size_t c) const {
    size_t i, j, k;
    ijk(c, i, j, k);
    return p.x >= xe_[i] && p.x <= xe_[i + 1] && p.y >= ye_[j] && p.y <= ye_[j + 1] && p.z >= ze_[k] &&
           p.z <= ze_[k + 1];
  }
  size_t GetContainingCell(const PointT &p) const {
    return cell(locate(xe_, p.x), locate(ye_, p.y), locate(ze_, p.z));
  }
  const std::array<size_t, 6> &GetCellFaces(size_t c) const { return cellFaces_[c]; }
  std::pair<size_t, size_t> GetFaceNeighbors(size_t f) const {
    int ax;
    size_t a[3];
    faceIJK(f, ax, a);
    const size_t n[3] = {nx_, ny_, nz_};
    size_t first = N_ + f, second = N_ + f;
    if (a[ax] > 0) {
      size_t b[3] = {a[0], a[1], a[2]};
      b[ax] -= 1;
      first = cell(b[0], b[1], b[2]);
    }
    if (a[ax] < n[ax]) second = cell(a[0], a[1], a[2]);
    return {first, second};
  }
  PointT Normal(size_t f) const {
    int ax;
    size_t a[3];
    faceIJK(f, ax, a);
    const double A = GetArea(f);
    return PointT(ax == 0 ? A : 0.0, ax == 1 ? A : 0.0, ax == 2 ? A : 0.0);
  }
  PointT FaceCM(size_t f) const {
    const auto v = GetPointsInFace(f);
    double x = 0, y = 0, z = 0;
    for (size_t q : v) {
      x += verts_[q].x;
      y += verts_[q].y;
      z += verts_[q].z;
    }
    return PointT(0.25 * x, 0.25 * y, 0.25 * z);
  }
  double GetArea(size_t f) const {
    int ax;
    size_t a[3];
    faceIJK(f, ax, a);
    const double dx = ax == 0 ? 1.0 : xe_[a[0] + 1] - xe_[a[0]];
    const double dy = ax == 1 ? 1.0 : ye_[a[1] + 1] - ye_[a[1]];
    const double dz = ax == 2 ? 1.0 : ze_[a[2] + 1] - ze_[a[2]];
    return dx * dy * dz;
  }
  std::array<size_t, 4> GetPointsInFace(size_t f) const {
    int ax;
    size_t a[3];
    faceIJK(f, ax, a);
    const size_t i = a[0], j = a[1], k = a[2];
    if (ax == 0) return {vid(i, j, k), vid(i, j + 1, k), vid(i, j + 1, k + 1), vid(i, j, k + 1)};
    if (ax == 1) return {vid(i, j, k), vid(i + 1, j, k), vid(i + 1, j, k + 1), vid(i, j, k + 1)};
    return {vid(i, j, k), vid(i + 1, j, k), vid(i + 1, j + 1, k), vid(i, j + 1, k)};
  }
  const std::vector<PointT> &GetFacePoints() const { return verts_; }
  const std::vector<Face_T> &GetBoxFaces() const { return boxFaces_; }

 private:
  static std::vector<double> edges(double lo, double hi, size_t n) {
    std::vector<double> e(n + 1);
    for (size_t i = 0; i <= n; ++i) e[i] = lo + (hi - lo) * static_cast<double>(i) / static_cast<double>(n);
    e[n] = hi;
    return e;
  }
  static size_t locate(const std::vector<double> &e, double v) {
    auto it = std::upper_bound(e.begin(), e.end(), v);
    long idx = static_cast<long>(it - e.begin()) - 1;
    return static_cast<size_t>(std::clamp(idx, 0L, static_cast<long>(e.size()) - 2));
  }
  size_t cell(size_t i, size_t j, size_t k) const { return (i * ny_ + j) * nz_ + k; }
  void ijk(size_t c, size_t &i, size_t &j, size_t &k) const {
    k = c % nz_;
    j = (c / nz_) % ny_;
    i = c / (ny_ * nz_);
  }
  size_t vid(size_t i, size_t j, size_t k) const { return (i * (ny_ + 1) + j) * (nz_ + 1) + k; }
  size_t fx(size_t i, size_t j, size_t k) const { return (i * ny_ + j) * nz_ + k; }
  size_t fy(size_t i, size_t j, size_t k) const { return Fx_ + (i * (ny_ + 1) + j) * nz_ + k; }
  size_t fz(size_t i, size_t j, size_t k) const { return Fx_ + Fy_ + (i * ny_ + j) * (nz_ + 1) + k; }
  void faceIJK(size_t f, int &ax, size_t a[3]) const {
    if (f < Fx_) {
      ax = 0;
      a[2] = f % nz_;
      a[1] = (f / nz_) % ny_;
      a[0] = f / (ny_ * nz_);
    } else if (f < Fx_ + Fy_) {
      f -= Fx_;
      ax = 1;
      a[2] = f % nz_;
      a[1] = (f / nz_) % (ny_ + 1);
      a[0] = f / ((ny_ + 1) * nz_);
    } else {
      f -= Fx_ + Fy_;
      ax = 2;
      a[2] = f % (nz_ + 1);
      a[1] = (f / (nz_ + 1)) % ny_;
      a[0] = f / (ny_ * (nz_ + 1));
    }
  }

  std::vector<double> xe_, ye_, ze_;
  size_t nx_, ny_, nz_, N_, Fx_, Fy_, Fz_;
  PointT ll_, ur_;
  std::vector<PointT> verts_;
  std::vector<std::array<size_t, 6>> cellFaces_;
  std::vector<Face_T> boxFaces_;
};

}  // namespace STORM
