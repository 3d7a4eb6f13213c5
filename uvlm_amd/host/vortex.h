// vortex.h -- Biot-Savart entry points and the VortexRing type of the host API.
//
// Same names and argument meaning as the reference (vortex.h:17-187,
// velocity.h:14-24, shed.h:24-45).  The sums run on the GPU through the
// C-ABI; batched overloads (vectors of targets) are the efficient form.
#pragma once

#include <functional>
#include <vector>

#include "multiple_sheet.h"
#include "vector3d.h"

namespace UVLM {

// vortex.cpp:11-19
double DistanceLineAndPoint(const Vector3d& start, const Vector3d& end,
                            const Vector3d& pos);

// vortex.cpp:21-39 (one segment).
void BiotSavartLaw(Vector3d* result, const Vector3d& start, const Vector3d& end,
                   const Vector3d& pos);

// vortex.h:36-171: a ring of (usually 4) nodes with a circulation.
class VortexRing {
 public:
  static const std::size_t DEFAULT_NODE_SIZE = 4;

  VortexRing() : gamma_(0.0) {}

  // vortex.cpp:41-50: sum of the segments node[(k+1)%n] -> node[k], x Gamma.
  void BiotSavartLaw(Vector3d* result, const Vector3d& pos, double gamma) const;
  // vortex.cpp:52-65: segments k in {0, 2} only.
  void ChordwiseBiotSavartLaw(Vector3d* result, const Vector3d& pos,
                              double gamma) const;
  void BiotSavartLaw(Vector3d* result, const Vector3d& pos) const {
    BiotSavartLaw(result, pos, gamma_);
  }
  void ChordwiseBiotSavartLaw(Vector3d* result, const Vector3d& pos) const {
    ChordwiseBiotSavartLaw(result, pos, gamma_);
  }

  VortexRing& PushNode(const Vector3d& pos) {
    nodes_.push_back(pos);
    return *this;
  }
  VortexRing& PushNode(double x, double y, double z) {
    nodes_.emplace_back(x, y, z);
    return *this;
  }
  void SaveReferenceNode() { nodes0_ = nodes_; }

  Vector3d Normal() const;     // vortex.cpp:82-87
  Vector3d Tangent() const;    // vortex.cpp:89-94
  Vector3d Tangent2() const;   // vortex.cpp:96
  Vector3d Centroid() const;   // vortex.cpp:98-100
  Vector3d ReferenceCentroid() const;
  Vector3d TanVecChord() const;
  Vector3d TanVecSpan() const;
  double CalcC() const;
  double CalcB() const;
  double Area() const;         // vortex.cpp:132-138
  Vector3d Impulse() const;    // vortex.cpp:140-143
  double AngleOfAttack(const Vector3d& Q) const;

  double gamma() const { return gamma_; }
  void set_gamma(double g) { gamma_ = g; }
  std::vector<Vector3d>& nodes() { return nodes_; }
  const std::vector<Vector3d>& nodes() const { return nodes_; }
  const std::vector<Vector3d>& nodes0() const { return nodes0_; }
  void Clear() { nodes_.clear(); }

  // vortex.h:132-160
  void ForEachSegment(
      std::function<void(const Vector3d&, const Vector3d&)> func) const {
    for (std::size_t i = 0; i < nodes_.size(); i++)
      func(nodes_[(i + 1) % nodes_.size()], nodes_[i]);
  }
  void ForEachSegment(std::function<void(const Vector3d&, const Vector3d&,
                                         const Vector3d&, const Vector3d&)>
                          func) const {
    for (std::size_t i = 0; i < nodes_.size(); i++)
      func(nodes_[(i + 1) % nodes_.size()], nodes_[i],
           nodes0_[(i + 1) % nodes_.size()], nodes0_[i]);
  }

 private:
  double gamma_;
  std::vector<Vector3d> nodes_;
  std::vector<Vector3d> nodes0_;
};

// vortex.h:173-182
Vector3d VORTEX(const Vector3d& x, const Vector3d& x1, const Vector3d& x2,
                double gamma);

// vortex.cpp:145-160
Vector3d VORING(const Vector3d& x, const MultipleSheet<Vector3d>& pos,
                const MultipleSheet<double>& gamma, std::size_t n, std::size_t i,
                std::size_t j);

// velocity.h:14-24 (one target) and its batched form.
Vector3d InducedVelocity(const Vector3d& x, const MultipleSheet<Vector3d>& pos,
                         const MultipleSheet<double>& gamma);
std::vector<Vector3d> InducedVelocity(const std::vector<Vector3d>& targets,
                                      const MultipleSheet<Vector3d>& pos,
                                      const MultipleSheet<double>& gamma);

// shed.inl:22-33 and :66-77: ring ranges (rings of 4 nodes, own Gamma).
void InducedVelocity(Vector3d* result, const Vector3d& pos,
                     std::vector<VortexRing>::const_iterator first,
                     std::vector<VortexRing>::const_iterator last);
void ChordwiseInducedVelocity(Vector3d* result, const Vector3d& pos,
                              std::vector<VortexRing>::const_iterator first,
                              std::vector<VortexRing>::const_iterator last);
std::vector<Vector3d> InducedVelocity(const std::vector<Vector3d>& targets,
                                      std::vector<VortexRing>::const_iterator first,
                                      std::vector<VortexRing>::const_iterator last,
                                      bool chordwise = false);

}  // namespace UVLM
