// output.cpp -- snapshot message conversions (declarations in output.h).
#include "output.h"

namespace UVLM {

Vector3d PointToVector3d(const proto::Point& p) { return Vector3d(p.x(), p.y(), p.z()); }

proto::Point Vector3dToPoint(const Vector3d& v) {
  proto::Point p;
  p.set_x(v.x());
  p.set_y(v.y());
  p.set_z(v.z());
  return p;
}

proto::VortexRing VortexRingToProto(const VortexRing& v) {
  proto::VortexRing r;
  r.has_gamma = true;
  r.gamma = v.gamma();
  for (const auto& p : v.nodes()) r.nodes.push_back(Vector3dToPoint(p));
  for (const auto& p : v.nodes0()) r.nodes0.push_back(Vector3dToPoint(p));
  return r;
}

VortexRing ProtoToVortexRing(const proto::VortexRing& v) {
  VortexRing r;
  r.set_gamma(v.gamma);
  for (const auto& p : v.nodes) r.PushNode(PointToVector3d(p));
  return r;
}

std::shared_ptr<std::vector<VortexRing>> Snapshot2ToContainers(
    std::vector<VortexContainer>* containers, const proto::Snapshot2& snapshot) {
  auto vortices = std::make_shared<std::vector<VortexRing>>();
  vortices->reserve(snapshot.vortices.size());
  for (const auto& v : snapshot.vortices) vortices->push_back(ProtoToVortexRing(v));
  containers->clear();
  for (const auto& s : snapshot.container_shapes)
    containers->emplace_back(vortices, s.rows, s.cols, s.id, s.chord, s.span);
  return vortices;
}

void Snapshot2ToMorphingVelocities(std::vector<Vector3d>* centers,
                                   std::vector<std::vector<Vector3d>>* v_nodes,
                                   std::vector<Vector3d>* freestreams,
                                   const proto::Snapshot2& snapshot) {
  centers->clear();
  v_nodes->clear();
  for (const auto& mv : snapshot.morphing_velocities) {
    centers->push_back(PointToVector3d(mv.center));
    freestreams->push_back(PointToVector3d(mv.freestream));
    std::vector<Vector3d> a;
    for (const auto& p : mv.nodes) a.push_back(PointToVector3d(p));
    v_nodes->push_back(a);
  }
}

namespace output {
void SheetToSnapshot(proto::Snapshot2* snapshot, const MultipleSheet<Vector3d>& pos,
                     const MultipleSheet<double>& gamma) {
  for (std::size_t n = 0; n < gamma.num(); ++n)
    for (std::size_t i = 0; i < gamma.rows(); ++i)
      for (std::size_t j = 0; j < gamma.cols(); ++j) {
        VortexRing v;
        v.PushNode(pos.at(n, i + 1, j))
            .PushNode(pos.at(n, i + 1, j + 1))
            .PushNode(pos.at(n, i, j + 1))
            .PushNode(pos.at(n, i, j));
        v.set_gamma(gamma.at(n, i, j));
        snapshot->vortices.push_back(VortexRingToProto(v));
      }
}
}  // namespace output

namespace proto_adaptor {
proto::VortexSheet ToVortexSheet(const MultipleSheet<Vector3d>& pos,
                                 const MultipleSheet<double>& gamma) {
  proto::VortexSheet s;
  s.num = static_cast<uint32_t>(gamma.num());
  s.rows = static_cast<uint32_t>(gamma.rows());
  s.cols = static_cast<uint32_t>(gamma.cols());
  for (std::size_t n = 0; n < gamma.num(); ++n)
    for (std::size_t i = 0; i < gamma.rows(); ++i)
      for (std::size_t j = 0; j < gamma.cols(); ++j) {
        VortexRing v;
        v.PushNode(pos.at(n, i, j))
            .PushNode(pos.at(n, i, j + 1))
            .PushNode(pos.at(n, i + 1, j + 1))
            .PushNode(pos.at(n, i + 1, j));
        v.set_gamma(gamma.at(n, i, j));
        s.vortices.push_back(VortexRingToProto(v));
      }
  return s;
}

MultipleSheet<VortexRing> FromVortexSheet(const proto::VortexSheet& sheet) {
  MultipleSheet<VortexRing> res(sheet.num, sheet.rows, sheet.cols);
  const std::size_t n = std::min<std::size_t>(res.size(), sheet.vortices.size());
  for (std::size_t k = 0; k < n; ++k) res[k] = ProtoToVortexRing(sheet.vortices[k]);
  return res;
}
}  // namespace proto_adaptor

}  // namespace UVLM
