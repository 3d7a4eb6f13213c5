// tools.cpp -- post-processing functions (declarations in tools.h).
#include "tools.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <map>
#include <set>

#include "output.h"
#include "wing.h"

namespace UVLM {
namespace tools {
namespace {

[[noreturn]] void fatal(const std::string& msg) {
  std::fprintf(stderr, "UVLM: %s\n", msg.c_str());
  std::abort();
}

std::vector<VortexRing> Rings(const proto::Snapshot2& snapshot) {
  std::vector<VortexRing> res;
  res.reserve(snapshot.vortices.size());
  for (const auto& v : snapshot.vortices) res.push_back(ProtoToVortexRing(v));
  return res;
}

}  // namespace

std::size_t VelocityAt(const std::string& input, const Vector3d& pos, int id, bool wing,
                       bool wake, const std::string& output) {
  if (!(wing || wake)) fatal("At least one of --wing or --wake must be true.");
  if (id < 0) fatal("CHECK failed: FLAGS_id >= 0");
  std::ifstream ifs(input, std::ios::binary);
  if (!ifs) fatal("Unable to open " + input);
  std::ofstream ofs(output);
  if (!ofs) fatal("Unable to open " + output);
  ofs << "step\tt\tx\ty\tz\n";
  RecordReader reader(&ifs);
  proto::AllVortexSheets sheets;
  std::size_t count = 0;
  const std::vector<Vector3d> target{pos};
  auto velocity = [&](const proto::VortexSheet& s) {
    const auto sheet = proto_adaptor::FromVortexSheet(s);
    if (static_cast<std::size_t>(id) + 1 > sheet.num() && sheet.size())
      fatal("wing id out of range");
    if (!sheet.size()) return Vector3d(0, 0, 0);
    const auto first = sheet.begin() + sheet.index(id, 0, 0);
    const auto last = first + sheet.rows() * sheet.cols();
    std::vector<VortexRing> rings(first, last);
    return UVLM::InducedVelocity(target, rings.cbegin(), rings.cend(), false)[0];
  };
  while (reader.ReadProtocolMessage(&sheets)) {
    const Vector3d v_wing = !wing ? Vector3d(0, 0, 0) : velocity(sheets.wing);
    const Vector3d v_wake = !wake ? Vector3d(0, 0, 0) : velocity(sheets.wake);
    const Vector3d r = v_wing + v_wake;
    ofs << count << "\t" << sheets.t << "\t" << r.x() << "\t" << r.y() << "\t" << r.z()
        << "\n";
    ++count;
  }
  return count;
}

std::vector<Vector3d> PlanePoints(const std::string& plane, double min0, double max0,
                                  double min1, double max1, double at,
                                  std::size_t resolution) {
  if (plane.size() != 2) fatal("FLAGS_plane must be size 2");
  if (plane[0] == plane[1]) fatal("FLAGS_plane must have different charactors");
  for (char ch : plane)
    if (std::string("xyz").find(ch) == std::string::npos)
      fatal("Charactors in FLAGS_plane must be one of 'xyz'");
  std::map<char, std::vector<double>> pos;
  std::set<char> keys{'x', 'y', 'z'};
  pos[plane[0]] = linspace(min0, max0, resolution);
  keys.erase(plane[0]);
  pos[plane[1]] = linspace(min1, max1, resolution);
  keys.erase(plane[1]);
  pos[*keys.begin()].push_back(at);
  std::vector<Vector3d> points;
  for (double x : pos['x'])
    for (double y : pos['y'])
      for (double z : pos['z']) points.emplace_back(x, y, z);
  return points;
}

std::vector<Vector3d> GridPoints(double xmin, double xmax, std::size_t nx, double ymin,
                                 double ymax, std::size_t ny, double zmin, double zmax,
                                 std::size_t nz) {
  const auto xs = linspace(xmin, xmax, nx);
  const auto ys = linspace(ymin, ymax, ny);
  const auto zs = linspace(zmin, zmax, nz);
  std::vector<Vector3d> res;
  res.reserve(nx * ny * nz);
  for (double z : zs)
    for (double y : ys)
      for (double x : xs) res.emplace_back(x, y, z);
  return res;
}

std::vector<Vector3d> SnapshotVelocity(const proto::Snapshot2& snapshot,
                                       const std::vector<Vector3d>& points) {
  const auto rings = Rings(snapshot);
  return UVLM::InducedVelocity(points, rings.cbegin(), rings.cend(), false);
}

void WritePlane(const proto::Snapshot2& snapshot, const std::string& plane, double min0,
                double max0, double min1, double max1, double at, std::size_t resolution,
                const std::string& output, const std::string& filetype) {
  std::string sep;
  if (filetype == "csv")
    sep = ",";
  else if (filetype == "tsv")
    sep = "\t";
  else
    fatal("Invalid filetype: " + filetype);
  const auto points = PlanePoints(plane, min0, max0, min1, max1, at, resolution);
  const auto vel = SnapshotVelocity(snapshot, points);
  std::ofstream ofs(output);
  if (!ofs) fatal("Unable to open " + output);
  ofs << "x" << sep << "y" << sep << "z" << sep << "u" << sep << "v" << sep << "w" << "\n";
  for (std::size_t k = 0; k < points.size(); ++k)
    ofs << points[k].x() << sep << points[k].y() << sep << points[k].z() << sep << vel[k].x()
        << sep << vel[k].y() << sep << vel[k].z() << "\n";
}

std::size_t WriteGrid(const proto::Snapshot2& snapshot, double xmin, double xmax, double ymin,
                      double ymax, double zmin, double zmax, std::size_t resolution,
                      const std::string& output) {
  const std::size_t N = resolution;
  const double dx = (xmax - xmin) / (N - 1);
  const double dy = (ymax - ymin) / (N - 1);
  const double dz = (zmax - zmin) / (N - 1);
  const auto points = GridPoints(xmin, xmax, N, ymin, ymax, N, zmin, zmax, N);
  auto vel = SnapshotVelocity(snapshot, points);
  for (auto& v : vel) v += Vector3d(1, 0, 0);  // freestream
  FILE* fp = std::fopen(output.c_str(), "w");
  if (!fp) fatal("Unable to open " + output);
  std::fprintf(fp, "# vtk DataFile Version 2.0\n");
  std::fprintf(fp, "HOGE FLOW\n");
  std::fprintf(fp, "ASCII\n");
  std::fprintf(fp, "DATASET STRUCTURED_POINTS\n");
  std::fprintf(fp, "DIMENSIONS %lu %lu %lu\n", (unsigned long)N, (unsigned long)N,
               (unsigned long)N);
  std::fprintf(fp, "ORIGIN %f %f %f\n", xmin, ymin, zmin);
  std::fprintf(fp, "SPACING %f %f %f\n", dx, dy, dz);
  std::fprintf(fp, "POINT_DATA %lu\n", (unsigned long)(N * N * N));
  std::fprintf(fp, "VECTORS volume_vectors double\n");
  for (const auto& u : vel) std::fprintf(fp, "%lf %lf %lf\n", u.x(), u.y(), u.z());
  std::fclose(fp);
  return vel.size();
}

ImpulseTerms CalcImpulse(const proto::Snapshot2& snapshot) {
  std::vector<VortexContainer> containers;
  auto vortices = Snapshot2ToContainers(&containers, snapshot);
  std::vector<Vector3d> centers, freestreams;
  std::vector<std::vector<Vector3d>> v_nodes;
  Snapshot2ToMorphingVelocities(&centers, &v_nodes, &freestreams, snapshot);
  if (containers.empty()) fatal("snapshot without containers");
  const VortexContainer& c0 = containers[0];
  ImpulseTerms r{snapshot.t, Vector3d(0, 0, 0), Vector3d(0, 0, 0), Vector3d(0, 0, 0)};
  for (auto it = c0.begin(); it != c0.end(); ++it) r.impulse += it->Impulse();
  std::size_t total = 0;
  for (const auto& c : containers) total += c.rows() * c.cols();
  for (auto it = vortices->cbegin() + total; it < vortices->cend(); ++it)
    r.wake_impulse += it->Impulse();
  // Lamb vectors of the first container's rings (calc_impulse.h:14-30), the
  // segment midpoints of all of them in one ring sum
  if (v_nodes.size() < c0.size() || freestreams.size() < c0.size())
    fatal("snapshot without morphing velocities for every ring");
  std::vector<Vector3d> mids, dls;
  for (std::size_t i = 0; i < c0.size(); i++)
    c0[i].ForEachSegment([&](const Vector3d& start, const Vector3d& end) {
      dls.push_back(end - start);
      mids.push_back((start + end) / 2);
    });
  const auto u = UVLM::InducedVelocity(mids, vortices->cbegin(), vortices->cend(), false);
  std::size_t q = 0;
  for (std::size_t i = 0; i < c0.size(); i++) {
    Vector3d res(0, 0, 0);
    const std::size_t m = c0[i].nodes().size();
    if (v_nodes[i].size() < m) fatal("morphing velocities missing segments");
    for (std::size_t k = 0; k < m; ++k, ++q) {
      const Vector3d ue = v_nodes[i][k] + u[q] - freestreams[i];
      res += ue.cross(dls[q]) * c0[i].gamma();
    }
    r.other_term += res;
  }
  return r;
}

proto::Snapshot2 ReadSnapshot2(const std::string& path) {
  std::ifstream ifs(path, std::ios::binary);
  if (!ifs) fatal("Unable to open " + path);
  const std::string bytes((std::istreambuf_iterator<char>(ifs)),
                          std::istreambuf_iterator<char>());
  proto::Snapshot2 s;
  if (!proto::Parse(bytes, &s)) fatal("cannot parse " + path);
  return s;
}

}  // namespace tools
}  // namespace UVLM
