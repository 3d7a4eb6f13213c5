// vortex.cpp -- Biot-Savart entry points of the host API (GPU via the C-ABI)
// and VortexRing geometry (vortex.cpp:11-160 of the reference).
#include "vortex.h"

#include <cmath>
#include <cstring>

#include "device.h"

namespace UVLM {

double DistanceLineAndPoint(const Vector3d& start, const Vector3d& end,
                            const Vector3d& pos) {
  Vector3d direction = end - start;
  direction.normalize();
  Vector3d diff = pos - start;
  return (diff - direction * direction.dot(diff)).norm();
}

void BiotSavartLaw(Vector3d* result, const Vector3d& start, const Vector3d& end,
                   const Vector3d& pos) {
  uvlm_ctx* c = gpu::Context();
  gpu::Check(uvlm_biot_savart(c, start.data(), end.data(), pos.data(), 1,
                              gpu::DefaultFlags(), result->data()),
             "BiotSavartLaw", c);
}

namespace {
std::vector<double> pack_rings(std::vector<VortexRing>::const_iterator first,
                               std::vector<VortexRing>::const_iterator last,
                               std::vector<double>* gamma) {
  std::vector<double> rings;
  for (auto it = first; it != last; ++it) {
    if (it->nodes().size() != 4) {
      std::fprintf(stderr, "UVLM: rings must have 4 nodes\n");
      std::abort();
    }
    for (const auto& p : it->nodes()) rings.insert(rings.end(), p.v, p.v + 3);
    gamma->push_back(it->gamma());
  }
  return rings;
}

void ring_sum(const std::vector<Vector3d>& targets, const double* rings,
              const double* gamma, std::size_t n, unsigned flags,
              std::vector<Vector3d>* out) {
  out->assign(targets.size(), Vector3d());
  if (n == 0 || targets.empty()) return;  // nothing to sum: no device work
  uvlm_ctx* c = gpu::Context();
  gpu::Check(uvlm_ring_induced_velocity(c, targets.empty() ? nullptr : targets[0].data(),
                                        targets.size(), rings, gamma, n,
                                        flags | gpu::DefaultFlags(),
                                        out->empty() ? nullptr : (*out)[0].data()),
             "ring InducedVelocity", c);
}
}  // namespace

void VortexRing::BiotSavartLaw(Vector3d* result, const Vector3d& pos,
                               double gamma) const {
  VortexRing tmp(*this);
  tmp.set_gamma(gamma);
  std::vector<VortexRing> one{tmp};
  std::vector<Vector3d> out;
  std::vector<double> g;
  auto rings = pack_rings(one.cbegin(), one.cend(), &g);
  ring_sum({pos}, rings.data(), g.data(), 1, UVLM_SCALE_PER_RING, &out);
  *result = out[0];
}

void VortexRing::ChordwiseBiotSavartLaw(Vector3d* result, const Vector3d& pos,
                                        double gamma) const {
  VortexRing tmp(*this);
  tmp.set_gamma(gamma);
  std::vector<VortexRing> one{tmp};
  std::vector<Vector3d> out;
  std::vector<double> g;
  auto rings = pack_rings(one.cbegin(), one.cend(), &g);
  ring_sum({pos}, rings.data(), g.data(), 1,
           UVLM_SCALE_PER_RING | UVLM_CHORDWISE_ONLY, &out);
  *result = out[0];
}

Vector3d VortexRing::Normal() const {
  Vector3d res = (nodes_[1] - nodes_[3]).cross(nodes_[2] - nodes_[0]);
  res.normalize();
  return res;
}
Vector3d VortexRing::Tangent() const {
  Vector3d res = (nodes_[1] - nodes_[3]) + (nodes_[2] - nodes_[0]);
  res.normalize();
  return res;
}
Vector3d VortexRing::Tangent2() const { return Normal().cross(Tangent()); }
Vector3d VortexRing::Centroid() const {
  return (((nodes_[0] + nodes_[1]) + nodes_[2]) + nodes_[3]) / 4;
}
Vector3d VortexRing::ReferenceCentroid() const {
  return (((nodes0_[0] + nodes0_[1]) + nodes0_[2]) + nodes0_[3]) / 4;
}
Vector3d VortexRing::TanVecChord() const { return (nodes_[1] - nodes_[0]).normalized(); }
Vector3d VortexRing::TanVecSpan() const { return (nodes_[3] - nodes_[0]).normalized(); }
double VortexRing::CalcC() const { return (nodes_[0] - nodes_[1]).norm(); }
double VortexRing::CalcB() const { return (nodes_[0] - nodes_[3]).norm(); }
double VortexRing::Area() const {
  const Vector3d u1 = nodes_[1] - nodes_[0], u2 = nodes_[3] - nodes_[0];
  const Vector3d v1 = nodes_[3] - nodes_[2], v2 = nodes_[1] - nodes_[2];
  return (u1.cross(u2).norm() + v1.cross(v2).norm()) / 2;
}
Vector3d VortexRing::Impulse() const { return Normal() * gamma_ * Area() * (-1); }
double VortexRing::AngleOfAttack(const Vector3d& Q) const {
  return std::atan2(Q.dot(Normal()), Q.dot(Tangent()));
}

Vector3d VORTEX(const Vector3d& x, const Vector3d& x1, const Vector3d& x2,
                double gamma) {
  Vector3d res;
  UVLM::BiotSavartLaw(&res, x1, x2, x);
  res *= gamma;
  return res;
}

Vector3d VORING(const Vector3d& x, const MultipleSheet<Vector3d>& pos,
                const MultipleSheet<double>& gamma, std::size_t n, std::size_t i,
                std::size_t j) {
  // one ring as a 1x1 lattice
  const double nodes[12] = {pos.at(n, i, j)[0],         pos.at(n, i, j)[1],
                            pos.at(n, i, j)[2],         pos.at(n, i, j + 1)[0],
                            pos.at(n, i, j + 1)[1],     pos.at(n, i, j + 1)[2],
                            pos.at(n, i + 1, j)[0],     pos.at(n, i + 1, j)[1],
                            pos.at(n, i + 1, j)[2],     pos.at(n, i + 1, j + 1)[0],
                            pos.at(n, i + 1, j + 1)[1], pos.at(n, i + 1, j + 1)[2]};
  const double g = gamma.at(n, i, j);
  Vector3d out;
  uvlm_ctx* c = gpu::Context();
  gpu::Check(uvlm_induced_velocity(c, x.data(), 1, nodes, &g, 1, 1, 1,
                                   gpu::DefaultFlags() | UVLM_SEQUENTIAL, out.data()),
             "VORING", c);
  return out;
}

std::vector<Vector3d> InducedVelocity(const std::vector<Vector3d>& targets,
                                      const MultipleSheet<Vector3d>& pos,
                                      const MultipleSheet<double>& gamma) {
  std::vector<Vector3d> out(targets.size());
  if (targets.empty()) return out;
  if (gamma.size() == 0) return out;
  if (pos.num() != gamma.num() || pos.rows() != gamma.rows() + 1 ||
      pos.cols() != gamma.cols() + 1) {
    std::fprintf(stderr, "UVLM: InducedVelocity: node sheet must be (num, rows+1, cols+1)\n");
    std::abort();
  }
  uvlm_ctx* c = gpu::Context();
  gpu::Check(uvlm_induced_velocity(c, targets[0].data(), targets.size(), pos.data()->data(),
                                   gamma.data(), gamma.num(), gamma.rows(), gamma.cols(),
                                   gpu::DefaultFlags(), out[0].data()),
             "InducedVelocity", c);
  return out;
}

Vector3d InducedVelocity(const Vector3d& x, const MultipleSheet<Vector3d>& pos,
                         const MultipleSheet<double>& gamma) {
  return InducedVelocity(std::vector<Vector3d>{x}, pos, gamma)[0];
}

std::vector<Vector3d> InducedVelocity(const std::vector<Vector3d>& targets,
                                      std::vector<VortexRing>::const_iterator first,
                                      std::vector<VortexRing>::const_iterator last,
                                      bool chordwise) {
  std::vector<double> g;
  auto rings = pack_rings(first, last, &g);
  std::vector<Vector3d> out;
  ring_sum(targets, rings.data(), g.data(), g.size(),
           UVLM_SCALE_PER_RING | (chordwise ? UVLM_CHORDWISE_ONLY : 0u), &out);
  return out;
}

void InducedVelocity(Vector3d* result, const Vector3d& pos,
                     std::vector<VortexRing>::const_iterator first,
                     std::vector<VortexRing>::const_iterator last) {
  *result = InducedVelocity(std::vector<Vector3d>{pos}, first, last, false)[0];
}

void ChordwiseInducedVelocity(Vector3d* result, const Vector3d& pos,
                              std::vector<VortexRing>::const_iterator first,
                              std::vector<VortexRing>::const_iterator last) {
  *result = InducedVelocity(std::vector<Vector3d>{pos}, first, last, true)[0];
}

}  // namespace UVLM
