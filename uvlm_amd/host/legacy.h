// legacy.h -- the container-based API of the reference's older loop
// (vortex_container.h, uvlm_vortex_ring.h, linear.h, shed.h, calc_impulse.h)
// on the GPU kernels.  Same names and argument meaning; iterator templates
// are replaced by std::vector<VortexRing> iterators.
#pragma once

#include <memory>
#include <utility>
#include <vector>

#include "morphing.h"
#include "vortex.h"

namespace UVLM {

// vortex_container.h:22-135: a window (id, rows, cols) onto a shared vector
// of rings.
class VortexContainer {
  typedef std::vector<VortexRing> container_t;
  typedef std::shared_ptr<container_t> vortices_ptr_t;

 public:
  VortexContainer() : vortices_(nullptr), rows_(0), cols_(0), id_(0), chord_(0), span_(0) {}
  VortexContainer(const vortices_ptr_t& vortices, std::size_t rows, std::size_t cols,
                  std::size_t id, double chord, double span)
      : vortices_(vortices), rows_(rows), cols_(cols), id_(id), chord_(chord), span_(span) {}

  std::size_t Offset() const { return id_ * rows_ * cols_; }
  std::size_t Index(std::size_t i) const { return Offset() + i; }
  std::size_t Index(std::size_t i, std::size_t j) const { return Offset() + j + i * cols_; }
  VortexRing& operator[](std::size_t i) { return (*vortices_)[Index(i)]; }
  const VortexRing& operator[](std::size_t i) const { return (*vortices_)[Index(i)]; }
  VortexRing& at(std::size_t i, std::size_t j) { return (*vortices_)[Index(i, j)]; }
  const VortexRing& at(std::size_t i, std::size_t j) const { return (*vortices_)[Index(i, j)]; }
  bool ShapeEquals(const VortexContainer& c) const {
    return rows_ == c.rows() && cols_ == c.cols() && id_ == c.id();
  }
  void alloc(std::size_t sz) {
    if (vortices_->size() < sz) vortices_->resize(sz);
  }
  // vortex_container.h:87-96
  double DeltaGamma(std::size_t i, std::size_t j) const {
    const double g = at(i, j).gamma();
    const double g_i = i == 0 ? 0 : at(i - 1, j).gamma();
    return g - g_i;
  }
  Vector3d Grad(std::size_t i, std::size_t j) const {
    const auto& v = at(i, j);
    return v.Tangent() * DeltaGamma(i, j) / v.CalcC();
  }
  const vortices_ptr_t& vortices() const { return vortices_; }
  std::size_t cols() const { return cols_; }
  std::size_t rows() const { return rows_; }
  std::size_t id() const { return id_; }
  double chord() const { return chord_; }
  double span() const { return span_; }
  void set_vortices(const vortices_ptr_t& v, std::size_t r, std::size_t c, std::size_t i,
                    double chord, double span) {
    vortices_ = v;
    rows_ = r;
    cols_ = c;
    id_ = i;
    chord_ = chord;
    span_ = span;
  }
  std::size_t size() const { return cols_ * rows_; }
  container_t::iterator begin() { return vortices_->begin() + Index(0); }
  container_t::iterator end() { return vortices_->begin() + Index(rows_ * cols_); }
  container_t::const_iterator begin() const { return vortices_->cbegin() + Index(0); }
  container_t::const_iterator end() const { return vortices_->cbegin() + Index(rows_ * cols_); }
  container_t::const_iterator cbegin() const { return begin(); }
  container_t::const_iterator cend() const { return end(); }
  container_t::iterator edge_begin() { return vortices_->begin() + Index(rows_ - 1, 0); }
  container_t::iterator edge_end() { return end(); }
  container_t::const_iterator edge_begin() const {
    return vortices_->cbegin() + Index(rows_ - 1, 0);
  }
  container_t::const_iterator edge_end() const { return end(); }

 private:
  vortices_ptr_t vortices_;
  std::size_t rows_, cols_, id_;
  double chord_, span_;
};

// vortex_container.h:140-185
std::size_t CountTotalSize(std::vector<VortexContainer>::const_iterator first,
                           std::vector<VortexContainer>::const_iterator last);
// Deep copy: the shared rings are copied into one new vector.
void CopyContainers(std::vector<VortexContainer>::const_iterator first,
                    std::vector<VortexContainer>::const_iterator last,
                    std::vector<VortexContainer>::iterator result);
std::pair<std::vector<VortexRing>::iterator, std::vector<VortexRing>::iterator> GetWake(
    const std::vector<VortexContainer>& containers);

// uvlm_vortex_ring.h:45-118: bound rings (y >= 0 rows mirrored) + wake rings.
class UVLMVortexRing {
 public:
  UVLMVortexRing() : cols_(0), rows_(0), origin_(0, 0, 0) {}
  // pos: (rows+1) x (cols+1) nodes of the half wing y >= 0, y fastest.
  void InitWing(const std::vector<Vector3d>& pos, std::size_t cols);
  void InducedVelocity(Vector3d* result, const Vector3d& pos) const;
  void InducedVelocityByBound(Vector3d* result, const Vector3d& pos) const;
  void InducedVelocityByWake(Vector3d* result, const Vector3d& pos) const;
  void InducedVelocityByChordwiseBound(Vector3d* result, const Vector3d& pos) const;
  void PlaneSymmetry();
  // Rings of the trailing-edge row: [first, last) indices into bound_vortices().
  std::pair<std::size_t, std::size_t> TrailingEdgeRange() const {
    return {(rows_ - 1) * 2 * cols_, rows_ * 2 * cols_};
  }
  void AppendWake(std::vector<VortexRing>::const_iterator first,
                  std::vector<VortexRing>::const_iterator last) {
    wake_vortices_.insert(wake_vortices_.end(), first, last);
  }
  std::vector<VortexRing>& bound_vortices() { return bound_vortices_; }
  const std::vector<VortexRing>& bound_vortices() const { return bound_vortices_; }
  std::vector<VortexRing>& wake_vortices() { return wake_vortices_; }
  const std::vector<VortexRing>& wake_vortices() const { return wake_vortices_; }
  const Vector3d& origin() const { return origin_; }
  std::size_t rows() const { return rows_; }
  std::size_t cols() const { return cols_; }
  void set_origin(const Vector3d& pos) { origin_ = pos; }

 private:
  std::vector<VortexRing> bound_vortices_;
  std::vector<VortexRing> wake_vortices_;
  std::size_t cols_, rows_;
  Vector3d origin_;
};

namespace internal {
// linear.inl:14-109
std::vector<double> CalcMatrix(std::vector<VortexRing>::const_iterator first,
                               std::vector<VortexRing>::const_iterator last);
std::vector<double> CalcRhsFreeStream(const Vector3d& freestream,
                                      std::vector<VortexRing>::const_iterator first,
                                      std::vector<VortexRing>::const_iterator last);
std::vector<double> CalcRhsWake(std::vector<VortexRing>::const_iterator bound_first,
                                std::vector<VortexRing>::const_iterator bound_last,
                                std::vector<VortexRing>::const_iterator wake_first,
                                std::vector<VortexRing>::const_iterator wake_last);
std::vector<double> CalcRhsMorphing(std::vector<VortexRing>::const_iterator first,
                                    std::vector<VortexRing>::const_iterator last,
                                    const Morphing& morphing, double t);
}  // namespace internal

// linear.inl:113-130: Gamma of the bound rings.  As in the reference the
// wake term is computed and not added to the right-hand side (the statement
// after the stray ';' at linear.inl:124).  Aborts on a singular matrix.
std::vector<double> SolveLinearProblem(std::vector<VortexRing>::const_iterator bound_first,
                                       std::vector<VortexRing>::const_iterator bound_last,
                                       std::vector<VortexRing>::const_iterator wake_first,
                                       std::vector<VortexRing>::const_iterator wake_last,
                                       const Vector3d& freestream, const Morphing& morphing,
                                       const double t);
// linear.inl:132-157 (containers sharing one ring vector, one morphing each).
std::vector<double> SolveLinearProblem(const std::vector<VortexContainer>& containers,
                                       const std::vector<Morphing>& morphings,
                                       const Vector3d& freestream, const double t);

// shed.inl:43-64: each node of every wake ring moves by
// (InducedVelocity(node; [vortex_first, vortex_last)) + freestream) * dt.
// All velocities are evaluated on the positions before the update (the
// reference's parallel variant races on shared rings; this is its
// race-free reading).
void Advect(std::vector<VortexRing>::const_iterator vortex_first,
            std::vector<VortexRing>::const_iterator vortex_last,
            std::vector<VortexRing>::iterator wake_first,
            std::vector<VortexRing>::iterator wake_last, const Vector3d& freestream,
            const double dt);
void AdvectParallel(std::vector<VortexRing>::const_iterator vortex_first,
                    std::vector<VortexRing>::const_iterator vortex_last,
                    std::vector<VortexRing>::iterator wake_first,
                    std::vector<VortexRing>::iterator wake_last, const Vector3d& freestream,
                    const double dt);
// shed.inl:79-86: wake ring k takes nodes 1 and 2 of edge ring k as its
// nodes 0 and 3.
void ConnectTrailingEdge(std::vector<VortexRing>::const_iterator edge_first,
                         std::vector<VortexRing>::const_iterator edge_last,
                         std::vector<VortexRing>::iterator wake);

namespace calc_impulse {
// calc_impulse.h:14-30: sum over the ring's segments of
// (u_m + u(mid) - U) x dl * Gamma, u(mid) from `vortices`.
Vector3d CalcLambVectorOnPanel(const VortexRing& v, const std::vector<Vector3d>& v_nodes,
                               const Vector3d& freestream,
                               const std::vector<VortexRing>& vortices);
}  // namespace calc_impulse

}  // namespace UVLM
