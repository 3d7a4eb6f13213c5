// host_gpu_test.cpp -- host-API tests that run the GPU kernels: the
// reference's known-answer tests of vortex_test.cpp, shed_test.cpp,
// linear_test.cpp through the UVLM C++ API, and a SimpleSimulator run with
// a load file.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>

#include "../../uvlm_amd/host/advect.h"
#include "../../uvlm_amd/host/legacy.h"
#include "../../uvlm_amd/host/simple_simulator.h"
#include "../../uvlm_amd/host/vortex.h"
#include "check.h"

using namespace UVLM;

static VortexRing square(double l, double x0, double y0) {
  VortexRing r;
  x0 -= l / 2;
  y0 -= l / 2;
  r.PushNode(x0, y0, 0).PushNode(x0 + l, y0, 0).PushNode(x0 + l, y0 + l, 0).PushNode(x0, y0 + l, 0);
  return r;
}

// vortex_test.cpp:34-79
TEST(biot_savart_law) {
  Vector3d r;
  BiotSavartLaw(&r, Vector3d(0, 0, -1e10), Vector3d(0, 0, 1e10), Vector3d(1, 0, 0));
  EXPECT_NEAR(r.y(), 1. / 2. / M_PI, 1e-10);
  BiotSavartLaw(&r, Vector3d(0, 0, -1), Vector3d(0, 0, 1), Vector3d(1, 0, 0));
  EXPECT_NEAR(r.y(), M_SQRT2 / 4. / M_PI, 1e-10);
  const Vector3d a(M_PI, M_SQRT2, -1), b(0, 0, 1);
  BiotSavartLaw(&r, a, b, a);
  EXPECT_NEAR(r.norm(), 0, 1e-10);
}

// vortex_test.cpp:104-130
TEST(vortex_ring_biot_savart) {
  VortexRing ring;
  ring.set_gamma(4 * M_PI);
  ring.PushNode(-1, -1, 0).PushNode(1, -1, 0).PushNode(1, 1, 0).PushNode(-1, 1, 0);
  Vector3d r;
  ring.BiotSavartLaw(&r, Vector3d(0, 0, 0));
  EXPECT_DOUBLE_EQ(0, r.x());
  EXPECT_DOUBLE_EQ(0, r.y());
  EXPECT_DOUBLE_EQ(-4 * M_SQRT2, r.z());
  ring.ChordwiseBiotSavartLaw(&r, Vector3d(0, 0, 0));
  EXPECT_DOUBLE_EQ(-2 * M_SQRT2, r.z());
}

// shed_test.cpp:21-50, 67-91
TEST(induced_velocity_many_rings_and_advect) {
  VortexRing v;
  v.set_gamma(4 * M_PI);
  v.PushNode(1, 1, 0).PushNode(-1, 1, 0).PushNode(-1, -1, 0).PushNode(1, -1, 0);
  std::vector<VortexRing> rings(20, v);
  Vector3d res;
  InducedVelocity(&res, Vector3d(0, 0, 0), rings.cbegin(), rings.cend());
  EXPECT_DOUBLE_EQ(-4 * M_SQRT2 * 20, res.z());
  ChordwiseInducedVelocity(&res, Vector3d(0, 0, 0), rings.cbegin(), rings.cend());
  EXPECT_DOUBLE_EQ(-4 * M_SQRT2 / 2 * 20, res.z());
  std::vector<VortexRing> wake{square(2, 0, 0)}, none;
  Advect(none.cbegin(), none.cend(), wake.begin(), wake.end(), Vector3d(1, 0, 0), 0.1);
  EXPECT_DOUBLE_EQ(-1 + 0.1, wake[0].nodes()[0].x());
  EXPECT_DOUBLE_EQ(1 + 0.1, wake[0].nodes()[2].x());
  std::vector<VortexRing> edge{square(2, 1, 1), square(2, 1, 3)};
  std::vector<VortexRing> w2{square(2, 99, 1), square(2, 99, 3)};
  ConnectTrailingEdge(edge.cbegin(), edge.cend(), w2.begin());
  EXPECT_EQ(w2[0].nodes()[0], Vector3d(2, 0, 0));
  EXPECT_EQ(w2[1].nodes()[3], Vector3d(2, 4, 0));
}

// linear_test.cpp:61-116
TEST(solve_linear_problem) {
  const double alpha = M_PI / 6, U = 2, a = 0.1;
  VortexRing v;
  v.PushNode(0, 0, 0)
      .PushNode(a * std::cos(alpha), 0, -a * std::sin(alpha))
      .PushNode(a * std::cos(alpha), a, -a * std::sin(alpha))
      .PushNode(0, a, 0);
  v.SaveReferenceNode();
  std::vector<VortexRing> vs{v}, dummy;
  Morphing m;
  auto g = SolveLinearProblem(vs.cbegin(), vs.cend(), dummy.cbegin(), dummy.cend(),
                              Vector3d(U, 0, 0), m, 0);
  EXPECT_NEAR(g[0], M_SQRT2 * M_PI * a * U * std::sin(alpha) / 4, 1e-10);

  std::vector<VortexRing> two{square(2, -1, 0), square(2, 1, 0)};
  for (auto& r : two) r.SaveReferenceNode();
  Morphing plug;
  plug.set_plug([](double t) { return 0.1 * std::sin(t); });
  auto g2 = SolveLinearProblem(two.cbegin(), two.cend(), dummy.cbegin(), dummy.cend(),
                               Vector3d(0, 0, 1), plug, 0);
  two[0].set_gamma(g2[0]);
  two[1].set_gamma(g2[1]);
  Vector3d v0, v1;
  InducedVelocity(&v0, Vector3d(-1, 0, 0), two.cbegin(), two.cend());
  InducedVelocity(&v1, Vector3d(1, 0, 0), two.cbegin(), two.cend());
  EXPECT_NEAR(v0.z(), -1 + 0.1, 1e-10);
  EXPECT_NEAR(v1.z(), -1 + 0.1, 1e-10);
}

// UVLMVortexRing velocity = bound + wake
TEST(uvlm_vortex_ring_velocity) {
  std::vector<Vector3d> pos;
  for (int i = 0; i <= 2; i++)
    for (int j = 0; j <= 3; j++) pos.emplace_back(0.3 * i, 0.5 * j, 0);
  UVLMVortexRing rings;
  rings.InitWing(pos, 3);
  for (auto& v : rings.bound_vortices()) v.set_gamma(1.0);
  Vector3d u;
  rings.InducedVelocity(&u, Vector3d(0.3, 0.0, 0.5));
  EXPECT_NEAR(u.y(), 0.0, 1e-12);  // mirror symmetry about y = 0
  EXPECT_TRUE(u.norm() > 0);
}

// SimpleSimulator::Run with a load file (simple_simulator.cpp:288-345)
TEST(simple_simulator_run_writes_loads) {
  simulator::SimpleSimulator sim;
  Morphing m;
  m.set_alpha(5 * M_PI / 180);
  m.set_plug([](double t) { return 0.1 * std::cos(t); });
  sim.AddWing(new wing::RectGenerator, m, 1.0, 4.0, 3, 6, Vector3d(0, 0, 0));
  sim.set_forward_flight(Vector3d(-1, 0, 0));
  sim.set_advection(advect::AdvectFactory("RK2"));
  const std::string path = "host_gpu_test_loads.tsv";
  sim.set_load_path(path);
  sim.Run(5, 0.125);
  std::ifstream ifs(path);
  std::string header;
  std::getline(ifs, header);
  EXPECT_EQ(header, std::string("t\tCD0\tCL0"));
  int lines = 0;
  std::string line;
  double t = -1, cd = 0, cl = 0;
  while (std::getline(ifs, line)) {
    std::istringstream ss(line);
    ss >> t >> cd >> cl;
    ++lines;
  }
  EXPECT_EQ(lines, 6);
  EXPECT_NEAR(t, 5 * 0.125, 1e-15);
  EXPECT_TRUE(cl > 0.1 && cl < 1.0);
  std::remove(path.c_str());
}

// the host-sheet advection (advect.cpp) equals the device-resident one
TEST(standalone_euler_matches_device_state) {
  simulator::SimpleSimulator sim;
  Morphing m;
  m.set_alpha(5 * M_PI / 180);
  sim.AddWing(new wing::RectGenerator, m, 1.0, 4.0, 2, 4, Vector3d(0, 0, 0));
  sim.set_forward_flight(Vector3d(-1, 0, 0));
  sim.set_advection(advect::AdvectFactory("euler"));
  sim.Prepare(3, 0.125);
  for (std::size_t s = 0; s < 3; ++s) sim.Step(s);
  MultipleSheet<Vector3d> wpos;
  MultipleSheet<double> wg;
  sim.wake(&wpos, &wg);
  // one more step, also advected on the host from the pre-advection state
  simulator::SimpleSimulator sim2;
  sim2.AddWing(new wing::RectGenerator, m, 1.0, 4.0, 2, 4, Vector3d(0, 0, 0));
  sim2.set_forward_flight(Vector3d(-1, 0, 0));
  sim2.Prepare(3, 0.125);
  for (std::size_t s = 0; s < 3; ++s) sim2.Step(s);
  sim.Sync();
  sim2.Sync();
  MultipleSheet<Vector3d> p2;
  MultipleSheet<double> g2;
  sim2.wake(&p2, &g2);
  for (std::size_t k = 0; k < wpos.size(); ++k)
    for (int d = 0; d < 3; ++d) EXPECT_EQ(wpos[k][d], p2[k][d]);
}

int main(int argc, char** argv) { return check::run(argc, argv); }
