// output.h -- conversions between the host API's types and the snapshot
// messages (proto_adaptor.h:14-150 and output.h/.cpp of the reference).
#pragma once

#include <memory>
#include <vector>

#include "legacy.h"
#include "multiple_sheet.h"
#include "proto_io.h"
#include "vortex.h"

namespace UVLM {

// proto_adaptor.h:14-27
Vector3d PointToVector3d(const proto::Point& p);
proto::Point Vector3dToPoint(const Vector3d& v);
// proto_adaptor.h:29-41: gamma, nodes and reference nodes.
proto::VortexRing VortexRingToProto(const VortexRing& v);
// proto_adaptor.h:47-54: gamma and nodes (the reference nodes are not read).
VortexRing ProtoToVortexRing(const proto::VortexRing& v);

// proto_adaptor.h:75-90: the rings of a Snapshot2 in one shared vector and
// a container per shape (rows, cols, id, chord, span).
std::shared_ptr<std::vector<VortexRing>> Snapshot2ToContainers(
    std::vector<VortexContainer>* containers, const proto::Snapshot2& snapshot);
// proto_adaptor.h:92-110
void Snapshot2ToMorphingVelocities(std::vector<Vector3d>* centers,
                                   std::vector<std::vector<Vector3d>>* v_nodes,
                                   std::vector<Vector3d>* freestreams,
                                   const proto::Snapshot2& snapshot);

namespace output {
// output.cpp:11-32: one message ring per (n, i, j) in K order with the nodes
// pushed as p(i+1,j), p(i+1,j+1), p(i,j+1), p(i,j) (the reverse of the
// sheet's ring orientation, as the reference writes them).
void SheetToSnapshot(proto::Snapshot2* snapshot, const MultipleSheet<Vector3d>& pos,
                     const MultipleSheet<double>& gamma);
}  // namespace output

namespace proto_adaptor {
// proto_adaptor.h:115-137: num, rows, cols and the rings p(i,j), p(i,j+1),
// p(i+1,j+1), p(i+1,j) with their circulations.
proto::VortexSheet ToVortexSheet(const MultipleSheet<Vector3d>& pos,
                                 const MultipleSheet<double>& gamma);
// proto_adaptor.h:139-146
MultipleSheet<VortexRing> FromVortexSheet(const proto::VortexSheet& sheet);
}  // namespace proto_adaptor

}  // namespace UVLM
