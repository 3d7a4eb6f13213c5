// tools.h -- the reference's post-processing drivers as library functions:
// velocity at a point over a sheet stream (tools/velocity_at_main.cpp),
// velocity on a plane and on a 3-D grid from a snapshot
// (tools/2d_velocity_main.cpp, tools/3d_velocity_main.cpp) and the vortex
// impulse with the Lamb-vector term (calc_impulse_main.cpp).  Every target
// set goes to the GPU ring sums in one call.  Command-line parsing and file
// globbing stay with the caller.
#pragma once

#include <string>
#include <vector>

#include "legacy.h"
#include "proto_io.h"
#include "vector3d.h"

namespace UVLM {
namespace tools {

// velocity_at_main.cpp:54-97: for every AllVortexSheets record of `input`,
// the velocity at pos induced by the rings of sheet `id` of the wing sheet
// (wing) plus those of the wake sheet (wake); a TSV "step t x y z" to
// `output`.  Returns the number of records (aborts when neither wing nor
// wake is chosen, id < 0, or a file cannot be opened).
std::size_t VelocityAt(const std::string& input, const Vector3d& pos, int id, bool wing,
                       bool wake, const std::string& output);

// 2d_velocity_main.cpp:59-89: resolution x resolution points of the plane
// `plane` ("xz", "yx", ...; the first letter varies along [min0, max0], the
// second along [min1, max1]) at the third coordinate `at`, x slowest.
std::vector<Vector3d> PlanePoints(const std::string& plane, double min0, double max0,
                                  double min1, double max1, double at,
                                  std::size_t resolution);
// 3d_velocity_main.cpp:34-48: the grid's points, x fastest.
std::vector<Vector3d> GridPoints(double xmin, double xmax, std::size_t nx, double ymin,
                                 double ymax, std::size_t ny, double zmin, double zmax,
                                 std::size_t nz);
// Velocity at `points` of every ring of a Snapshot2 (each ring's segments
// summed, then times its circulation; shed.inl:22-33).
std::vector<Vector3d> SnapshotVelocity(const proto::Snapshot2& snapshot,
                                       const std::vector<Vector3d>& points);
// 2d_velocity_main.cpp:109-127, 129-155: one plane of one snapshot as csv
// or tsv ("x y z u v w").
void WritePlane(const proto::Snapshot2& snapshot, const std::string& plane, double min0,
                double max0, double min1, double max1, double at, std::size_t resolution,
                const std::string& output, const std::string& filetype);
// 3d_velocity_main.cpp:78-132: the velocity plus the unit freestream (1, 0, 0)
// on a resolution^3 grid as a legacy-VTK structured-points file; returns
// the number of points.
std::size_t WriteGrid(const proto::Snapshot2& snapshot, double xmin, double xmax, double ymin,
                      double ymax, double zmin, double zmax, std::size_t resolution,
                      const std::string& output);

// calc_impulse_main.cpp:55-85: t, the impulse of the first container's
// rings, the sum of their Lamb vectors (calc_impulse.h:14-30) with every
// ring of the snapshot as the source, and the impulse of the wake rings.
struct ImpulseTerms {
  double t;
  Vector3d impulse, other_term, wake_impulse;
};
ImpulseTerms CalcImpulse(const proto::Snapshot2& snapshot);

// Read one serialised Snapshot2 file (aborts when unreadable).
proto::Snapshot2 ReadSnapshot2(const std::string& path);

}  // namespace tools
}  // namespace UVLM
