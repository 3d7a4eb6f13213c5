// legacy.cpp -- container-based API of the reference's older loop on the
// GPU kernels (vortex_container.h, uvlm_vortex_ring.cpp, linear.inl,
// shed.inl, calc_impulse.h).
#include "legacy.h"

#include <cstdio>
#include <cstdlib>

#include "device.h"

namespace UVLM {

std::size_t CountTotalSize(std::vector<VortexContainer>::const_iterator first,
                           std::vector<VortexContainer>::const_iterator last) {
  std::size_t res = 0;
  for (; first != last; ++first) res += first->size();
  return res;
}

void CopyContainers(std::vector<VortexContainer>::const_iterator first,
                    std::vector<VortexContainer>::const_iterator last,
                    std::vector<VortexContainer>::iterator result) {
  if (first == last) return;
  auto vortices = first->vortices();
  const std::size_t total = CountTotalSize(first, last);
  auto copied = std::make_shared<std::vector<VortexRing>>(vortices->begin(),
                                                          vortices->begin() + total);
  for (; first != last; ++first, ++result)
    *result = VortexContainer(copied, first->rows(), first->cols(), first->id(),
                              first->chord(), first->span());
}

std::pair<std::vector<VortexRing>::iterator, std::vector<VortexRing>::iterator> GetWake(
    const std::vector<VortexContainer>& containers) {
  const std::size_t off = CountTotalSize(containers.cbegin(), containers.cend());
  auto v = containers.begin()->vortices();
  return {v->begin() + off, v->end()};
}

// ---------------------------------------------------------------------------
void UVLMVortexRing::InitWing(const std::vector<Vector3d>& pos, std::size_t cols) {
  const std::size_t cols_pos = cols + 1;
  cols_ = cols;
  rows_ = pos.size() / cols_pos - 1;
  for (std::size_t i = 0; i < rows_; i++) {
    for (std::size_t j = 0; j < cols_; j++) {  // y >= 0
      const std::size_t idx[4] = {j + i * cols_pos, j + (i + 1) * cols_pos,
                                  j + 1 + (i + 1) * cols_pos, j + 1 + i * cols_pos};
      VortexRing v;
      for (auto k : idx) v.nodes().push_back(pos[k]);
      bound_vortices_.push_back(v);
    }
    for (std::size_t j = 0; j < cols_; j++) {  // y <= 0, filled by PlaneSymmetry
      bound_vortices_.emplace_back();
      bound_vortices_.back().nodes().resize(4);
    }
  }
  PlaneSymmetry();
  for (auto& v : bound_vortices_)
    for (auto& node : v.nodes()) node += origin_;
  for (auto& v : bound_vortices_) v.SaveReferenceNode();
}

void UVLMVortexRing::PlaneSymmetry() {
  auto mirror = [](const Vector3d& p) { return Vector3d(p.x(), -p.y(), p.z()); };
  for (std::size_t i = 0; i < rows_; i++)
    for (std::size_t j = 0; j < cols_; j++) {
      const std::size_t idx = j + i * 2 * cols_;
      auto& to = bound_vortices_[idx + cols_].nodes();
      const auto& from = bound_vortices_[idx].nodes();
      to[0] = mirror(from[3]);
      to[1] = mirror(from[2]);
      to[2] = mirror(from[1]);
      to[3] = mirror(from[0]);
    }
}

void UVLMVortexRing::InducedVelocityByBound(Vector3d* result, const Vector3d& pos) const {
  UVLM::InducedVelocity(result, pos, bound_vortices_.cbegin(), bound_vortices_.cend());
}
void UVLMVortexRing::InducedVelocityByWake(Vector3d* result, const Vector3d& pos) const {
  UVLM::InducedVelocity(result, pos, wake_vortices_.cbegin(), wake_vortices_.cend());
}
void UVLMVortexRing::InducedVelocityByChordwiseBound(Vector3d* result,
                                                     const Vector3d& pos) const {
  ChordwiseInducedVelocity(result, pos, bound_vortices_.cbegin(), bound_vortices_.cend());
}
void UVLMVortexRing::InducedVelocity(Vector3d* result, const Vector3d& pos) const {
  Vector3d v, w;
  InducedVelocityByBound(&v, pos);
  InducedVelocityByWake(&w, pos);
  *result = v + w;
}

// ---------------------------------------------------------------------------
namespace internal {

std::vector<double> CalcMatrix(std::vector<VortexRing>::const_iterator first,
                               std::vector<VortexRing>::const_iterator last) {
  const std::size_t n = last - first;
  std::vector<double> rings, cent, nrm, A(n * n);
  for (auto it = first; it != last; ++it) {
    if (it->nodes().size() != 4) {
      std::fprintf(stderr, "UVLM: rings must have 4 nodes\n");
      std::abort();
    }
    for (const auto& p : it->nodes()) rings.insert(rings.end(), p.v, p.v + 3);
    const Vector3d c = it->Centroid(), nl = it->Normal();
    cent.insert(cent.end(), c.v, c.v + 3);
    nrm.insert(nrm.end(), nl.v, nl.v + 3);
  }
  if (n == 0) return A;
  uvlm_ctx* ctx = gpu::Context();
  gpu::Check(uvlm_ring_influence_matrix(ctx, rings.data(), n, cent.data(), nrm.data(),
                                        A.data()),
             "CalcMatrix", ctx);
  return A;
}

std::vector<double> CalcRhsFreeStream(const Vector3d& freestream,
                                      std::vector<VortexRing>::const_iterator first,
                                      std::vector<VortexRing>::const_iterator last) {
  std::vector<double> res;
  for (auto it = first; it != last; ++it) res.push_back(freestream.dot(it->Normal()));
  return res;
}

std::vector<double> CalcRhsWake(std::vector<VortexRing>::const_iterator bound_first,
                                std::vector<VortexRing>::const_iterator bound_last,
                                std::vector<VortexRing>::const_iterator wake_first,
                                std::vector<VortexRing>::const_iterator wake_last) {
  std::vector<Vector3d> cent;
  for (auto it = bound_first; it != bound_last; ++it) cent.push_back(it->Centroid());
  const auto vel = UVLM::InducedVelocity(cent, wake_first, wake_last);
  std::vector<double> res;
  std::size_t k = 0;
  for (auto it = bound_first; it != bound_last; ++it, ++k)
    res.push_back(vel[k].dot(it->Normal()));
  return res;
}

std::vector<double> CalcRhsMorphing(std::vector<VortexRing>::const_iterator first,
                                    std::vector<VortexRing>::const_iterator last,
                                    const Morphing& morphing, double t) {
  std::vector<double> res;
  for (auto it = first; it != last; ++it) {
    Vector3d vel;
    morphing.Velocity(&vel, it->ReferenceCentroid(), t);
    res.push_back(vel.dot(it->Normal()));
  }
  return res;
}

}  // namespace internal

namespace {
std::vector<double> solve_dense(std::vector<double> A, const std::vector<double>& rhs) {
  const std::size_t n = rhs.size();
  std::vector<double> x(n);
  if (n == 0) return x;
  uvlm_ctx* ctx = gpu::Context();
  const int rc = uvlm_solve(ctx, n, A.data(), rhs.data(), x.data(), UVLM_SOLVE_F64, 0, 0,
                            nullptr);
  if (rc == UVLM_ERR_SINGULAR) {
    std::fprintf(stderr, "Matrix is not invertible.\n");
    std::abort();
  }
  gpu::Check(rc, "SolveLinearProblem", ctx);
  return x;
}
}  // namespace

std::vector<double> SolveLinearProblem(std::vector<VortexRing>::const_iterator bound_first,
                                       std::vector<VortexRing>::const_iterator bound_last,
                                       std::vector<VortexRing>::const_iterator wake_first,
                                       std::vector<VortexRing>::const_iterator wake_last,
                                       const Vector3d& freestream, const Morphing& morphing,
                                       const double t) {
  (void)wake_first;
  (void)wake_last;
  auto A = internal::CalcMatrix(bound_first, bound_last);
  auto rhs = internal::CalcRhsMorphing(bound_first, bound_last, morphing, t);
  const auto fs = internal::CalcRhsFreeStream(freestream, bound_first, bound_last);
  for (std::size_t k = 0; k < rhs.size(); ++k) rhs[k] = rhs[k] - fs[k];
  return solve_dense(std::move(A), rhs);
}

std::vector<double> SolveLinearProblem(const std::vector<VortexContainer>& containers,
                                       const std::vector<Morphing>& morphings,
                                       const Vector3d& freestream, const double t) {
  if (containers.size() != morphings.size()) {
    std::fprintf(stderr, "containers and morphings have to be same size\n");
    std::abort();
  }
  const auto& vortices = containers.begin()->vortices();
  const std::size_t off = CountTotalSize(containers.cbegin(), containers.cend());
  auto bf = vortices->cbegin(), bl = vortices->cbegin() + off;
  auto A = internal::CalcMatrix(bf, bl);
  std::vector<double> rhs;
  for (std::size_t i = 0; i < containers.size(); ++i) {
    auto part = internal::CalcRhsMorphing(containers[i].cbegin(), containers[i].cend(),
                                          morphings[i], t);
    rhs.insert(rhs.end(), part.begin(), part.end());
  }
  const auto fs = internal::CalcRhsFreeStream(freestream, bf, bl);
  for (std::size_t k = 0; k < rhs.size(); ++k) rhs[k] = rhs[k] - fs[k];
  return solve_dense(std::move(A), rhs);
}

// ---------------------------------------------------------------------------
void Advect(std::vector<VortexRing>::const_iterator vortex_first,
            std::vector<VortexRing>::const_iterator vortex_last,
            std::vector<VortexRing>::iterator wake_first,
            std::vector<VortexRing>::iterator wake_last, const Vector3d& freestream,
            const double dt) {
  std::vector<Vector3d> targets;
  for (auto w = wake_first; w != wake_last; ++w)
    for (const auto& node : w->nodes()) targets.push_back(node);
  if (targets.empty()) return;
  const auto vel = UVLM::InducedVelocity(targets, vortex_first, vortex_last);
  std::size_t k = 0;
  for (auto w = wake_first; w != wake_last; ++w)
    for (auto& node : w->nodes()) {
      Vector3d v = vel[k++];
      v += freestream;
      node += v * dt;
    }
}

void AdvectParallel(std::vector<VortexRing>::const_iterator vortex_first,
                    std::vector<VortexRing>::const_iterator vortex_last,
                    std::vector<VortexRing>::iterator wake_first,
                    std::vector<VortexRing>::iterator wake_last, const Vector3d& freestream,
                    const double dt) {
  Advect(vortex_first, vortex_last, wake_first, wake_last, freestream, dt);
}

void ConnectTrailingEdge(std::vector<VortexRing>::const_iterator edge_first,
                         std::vector<VortexRing>::const_iterator edge_last,
                         std::vector<VortexRing>::iterator wake) {
  for (auto e = edge_first; e != edge_last; ++e, ++wake) {
    wake->nodes()[0] = e->nodes()[1];
    wake->nodes()[3] = e->nodes()[2];
  }
}

namespace calc_impulse {

Vector3d CalcLambVectorOnPanel(const VortexRing& v, const std::vector<Vector3d>& v_nodes,
                               const Vector3d& freestream,
                               const std::vector<VortexRing>& vortices) {
  std::vector<Vector3d> mids, dls;
  v.ForEachSegment([&](const Vector3d& start, const Vector3d& end) {
    dls.push_back(end - start);
    mids.push_back((start + end) / 2);
  });
  const auto u = UVLM::InducedVelocity(mids, vortices.cbegin(), vortices.cend());
  Vector3d res(0, 0, 0);
  for (std::size_t k = 0; k < mids.size(); ++k) {
    const Vector3d ue = v_nodes[k] + u[k] - freestream;
    res += ue.cross(dls[k]) * v.gamma();
  }
  return res;
}

}  // namespace calc_impulse
}  // namespace UVLM
