// host_cpu_test.cpp -- host-API tests that need no GPU: containers, the
// kinematics and the wing generators (bitwise against the oracle), and the
// reference's container tests (vortex_container_test.cpp,
// uvlm_vortex_ring_test.cpp, wing_builder_test.cpp, proto_adaptor_test.cpp
// index order).
#include <cmath>
#include <memory>

#include "../../oracle/uvlm_oracle.h"
#include "../../uvlm_amd/host/advect.h"
#include "../../uvlm_amd/host/legacy.h"
#include "../../uvlm_amd/host/legacy_simulator.h"
#include "../../uvlm_amd/host/morphing.h"
#include "../../uvlm_amd/host/multiple_sheet.h"
#include "../../uvlm_amd/host/wing.h"
#include "check.h"

using namespace UVLM;

TEST(vector3d_conventions) {
  const Vector3d a(1e16, 1.0, -1e16), b(1.0, 1.0, 1.0);
  EXPECT_EQ(a.dot(b), (1e16 + 1.0) + -1e16);  // (x+y)+z
  EXPECT_EQ(a.squaredNorm(), (1e32 + 1.0) + 1e32);
  Vector3d c(3, 4, 0);
  c.normalize();
  EXPECT_EQ(c, Vector3d(3.0 / 5.0, 4.0 / 5.0, 0.0));
  Vector3d z(0, 0, 0);
  z.normalize();
  EXPECT_EQ(z, Vector3d(0, 0, 0));
  EXPECT_EQ(Vector3d(1, 0, 0).cross(Vector3d(0, 1, 0)), Vector3d(0, 0, 1));
}

// proto_adaptor_test.cpp:52-79 (n,i,j) order; App. B contract
TEST(multiple_sheet_layout) {
  MultipleSheet<double> s(2, 3, 4);
  EXPECT_EQ(s.size(), 24u);
  EXPECT_EQ(s.index(1, 2, 3), 3u + 4u * (2u + 3u * 1u));
  std::size_t k = 0;
  for (const auto& t : s.list_index()) {
    EXPECT_EQ(std::get<0>(t), k);
    EXPECT_EQ(std::get<0>(t), s.index(std::get<1>(t), std::get<2>(t), std::get<3>(t)));
    ++k;
  }
  for (std::size_t K = 0; K < s.size(); ++K) s[K] = (double)K;
  std::vector<double> row = {-1, -2, -3, -4, -5, -6, -7, -8};
  s.prepend_row(row.begin(), row.end());
  EXPECT_EQ(s.rows(), 4u);
  EXPECT_EQ(s.at(0, 0, 0), -1.0);
  EXPECT_EQ(s.at(1, 0, 3), -8.0);
  EXPECT_EQ(s.at(0, 1, 0), 0.0);
  EXPECT_EQ(s.at(1, 1, 0), 12.0);
  s.erase_last_row();
  EXPECT_EQ(s.rows(), 3u);
  EXPECT_EQ(s.at(1, 2, 3), 7.0 + 12.0);
}

static orc_wing_spec spec_of(double alpha, double plug_a, double plug_w, double flap_a,
                             double flap_w, double flap_p, double span) {
  orc_wing_spec w;
  std::memset(&w, 0, sizeof(w));
  w.chord = 1.0;
  w.span = span;
  w.alpha = alpha;
  w.plug_amp = plug_a;
  w.plug_omega = plug_w;
  w.flap_amp = flap_a;
  w.flap_omega = flap_w;
  w.flap_phase = flap_p;
  w.fourier_omega = 1.0;
  for (int m = 0; m < 4; m++)
    for (int q = 0; q < 2; q++) {
      w.twist_amp[m][q] = 0.01 * (m + 1) * (q ? -1 : 1);
      w.twist_phase[m][q] = 0.3 * m + q;
      w.bend_amp[m][q] = 0.005 * (4 - m) * (q ? 1 : -1);
      w.bend_phase[m][q] = 0.7 * q + 0.1 * m;
    }
  return w;
}

static Morphing morph_of(const orc_wing_spec& w) {
  Morphing m;
  m.set_alpha(w.alpha);
  m.set_plug([w](double t) { return w.plug_amp * std::cos(w.plug_omega * t + w.plug_phase); });
  m.set_flap([w](double t) { return w.flap_amp * std::cos(w.flap_omega * t + w.flap_phase); });
  m.set_twist([w](const Vector3d& x, double t) {
    const double eta = x.y() / (w.span / 2);
    double s = 0;
    for (int mm = 0; mm < 4; mm++)
      for (int q = 0; q < 2; q++)
        s = s + w.twist_amp[mm][q] * std::sin((mm + 1) * M_PI_2 * eta) *
                    std::cos((q + 1) * w.fourier_omega * t + w.twist_phase[mm][q]);
    return s;
  });
  m.set_bend([w](const Vector3d& x, double t) {
    const double eta = x.y() / (w.span / 2);
    const double xi = x.x() / w.chord;
    double s = 0;
    for (int mm = 0; mm < 4; mm++)
      for (int q = 0; q < 2; q++)
        s = s + w.bend_amp[mm][q] * std::sin(M_PI * xi) * std::sin((mm + 1) * M_PI_2 * eta) *
                    std::cos((q + 1) * w.fourier_omega * t + w.bend_phase[mm][q]);
    return s;
  });
  return m;
}

// morphing.cpp:23-52: host kinematics bitwise equal to the oracle's
TEST(morphing_bitwise_vs_oracle) {
  const orc_wing_spec specs[] = {spec_of(5 * M_PI / 180, 0.1, 1.0, 0, 0, 0, 4),
                                 spec_of(5 * M_PI / 180, 0, 0, M_PI / 6, 0.6, 0, 6),
                                 spec_of(0.0, 0.2, 0.3, -M_PI / 4, 0.2, M_PI / 2, 6)};
  for (const auto& w : specs) {
    Morphing m = morph_of(w);
    for (double x : {0.0, 0.3, 1.0})
      for (double y : {-3.0, -0.7, 0.0, 0.4, 2.9})
        for (double t : {0.0, 0.125, 3.7, 100.25}) {
          const double x0[3] = {x, y, 0.01 * x};
          double o[3], ov[3];
          orc_morph_perform(&w, x0, t, o);
          orc_morph_velocity(&w, x0, t, ov);
          const Vector3d p = m.Perfome(Vector3d(x0[0], x0[1], x0[2]), t);
          const Vector3d v = m.Velocity(Vector3d(x0[0], x0[1], x0[2]), t);
          for (int d = 0; d < 3; ++d) {
            EXPECT_EQ(p[d], o[d]);
            EXPECT_EQ(v[d], ov[d]);
          }
        }
  }
}

// morphing_test.cpp: plug / flap closed forms
TEST(morphing_closed_forms) {
  Morphing m;
  m.set_plug([](double t) { return std::sin(t); });
  EXPECT_NEAR(m.Perfome(Vector3d(1, 2, 0), M_PI / 2).z(), 1.0, 1e-15);
  EXPECT_NEAR(m.Velocity(Vector3d(1, 2, 0), 0.0).z(), 1.0, 1e-9);
  Morphing f;
  f.set_flap([](double t) { return t; });
  const Vector3d p = f.Perfome(Vector3d(0, 1, 0), M_PI / 2);
  EXPECT_NEAR(p.y(), 0.0, 1e-15);
  EXPECT_NEAR(p.z(), 1.0, 1e-15);
  const Vector3d q = f.Perfome(Vector3d(0, -1, 0), M_PI / 2);  // mirrored side
  EXPECT_NEAR(q.y(), 0.0, 1e-15);
  EXPECT_NEAR(q.z(), 1.0, 1e-15);
}

// wing.cpp / rect.cpp / naca4digit.cpp vs the oracle's lattice, bitwise
TEST(wing_generators_bitwise_vs_oracle) {
  struct G {
    int gen, digit;
    double chord, span;
    std::size_t rows, cols;
    double origin[3];
  };
  const G cases[] = {{ORC_GEN_RECT, 0, 1.0, 4.0, 8, 16, {0, 0, 0}},
                     {ORC_GEN_NACA4, 44, 1.0, 6.0, 20, 80, {0, 0, 0}},
                     {ORC_GEN_NACA4, 83, 1.0, 6.0, 16, 64, {2.5, 0, 0}},
                     {ORC_GEN_NACA00, 12, 2.0, 3.0, 5, 6, {0.1, -0.2, 0.3}}};
  for (const auto& c : cases) {
    std::unique_ptr<wing::WingGenerator> gen;
    if (c.gen == ORC_GEN_NACA4)
      gen.reset(new wing::NACA4digitGenerator(c.digit));
    else if (c.gen == ORC_GEN_NACA00)
      gen.reset(new wing::NACA00XXGenerator(c.digit));
    else
      gen.reset(new wing::RectGenerator());
    proto::Wing half, whole;
    gen->Generate(&half, c.chord, c.span / 2, c.rows, c.cols / 2);
    wing::WholeWing(&whole, half);
    auto pts = PointsToVector(whole.points());
    const Vector3d o(c.origin[0], c.origin[1], c.origin[2]);
    std::vector<double> ref(3 * (c.rows + 1) * (c.cols + 1));
    orc_build_wing(c.gen, c.digit, c.chord, c.span, c.rows, c.cols, c.origin, ref.data());
    EXPECT_EQ(pts.size(), (c.rows + 1) * (c.cols + 1));
    for (std::size_t k = 0; k < pts.size(); ++k) {
      const Vector3d p = pts[k] + o;
      for (int d = 0; d < 3; ++d) EXPECT_EQ(p[d], ref[3 * k + d]);
    }
  }
}

// rect_test.cpp:6-24
TEST(transform_chebyshev) {
  std::vector<double> x;
  wing::TransformChebyshev(&x, 5, 0, 2);
  EXPECT_DOUBLE_EQ(0, x[0]);
  EXPECT_DOUBLE_EQ(2 * std::cos(-M_PI / 8 * 3), x[1]);
  EXPECT_DOUBLE_EQ(2 * std::cos(-M_PI / 8 * 2), x[2]);
  EXPECT_DOUBLE_EQ(2 * std::cos(-M_PI / 8 * 1), x[3]);
  EXPECT_DOUBLE_EQ(2, x[4]);
}

// advect_factory.cpp:22-62
TEST(advect_factory_names) {
  EXPECT_EQ(advect::GetNames(), std::string("AB2, RK2, RK4, RK4classic, euler"));
  std::unique_ptr<advect::Advection> e(advect::AdvectFactory("euler"));
  std::unique_ptr<advect::Advection> r2(advect::AdvectFactory("RK2"));
  std::unique_ptr<advect::Advection> r4(advect::AdvectFactory("RK4"));
  std::unique_ptr<advect::Advection> ab(advect::AdvectFactory("AB2"));
  EXPECT_EQ(e->scheme(), 0);
  EXPECT_EQ(r2->scheme(), 1);
  EXPECT_EQ(r4->scheme(), 2);
  EXPECT_EQ(ab->scheme(), 3);
  EXPECT_EQ(r2->ignore_steps(), 5u);
}

// vortex_container_test.cpp:51-186
TEST(vortex_container) {
  auto v = std::make_shared<std::vector<VortexRing>>(24);
  for (std::size_t k = 0; k < 24; ++k) (*v)[k].set_gamma((double)k);
  VortexContainer c(v, 2, 3, 1, 1, 1);
  EXPECT_EQ(c.Index(0), 6u);
  EXPECT_EQ(c.Index(1, 2), 11u);
  EXPECT_EQ(c.at(1, 1).gamma(), 10.0);
  VortexContainer c2(v, 2, 3, 2, 1, 1);
  EXPECT_EQ(c2.Offset(), 12u);
  EXPECT_EQ(c2[1].gamma(), 13.0);
  c[0].set_gamma(100);
  EXPECT_EQ((*v)[6].gamma(), 100.0);
  std::vector<VortexContainer> cs = {VortexContainer(v, 2, 3, 0, 1, 1), c, c2};
  EXPECT_EQ(CountTotalSize(cs.cbegin(), cs.cend()), 18u);
  std::vector<VortexContainer> copy(3);
  CopyContainers(cs.cbegin(), cs.cend(), copy.begin());
  EXPECT_EQ(copy[1].at(0, 0).gamma(), 100.0);
  copy[1].at(0, 0).set_gamma(5);
  EXPECT_EQ(c.at(0, 0).gamma(), 100.0);  // deep copy
  EXPECT_EQ(copy[0].vortices()->size(), 18u);
  auto w = GetWake(cs);
  EXPECT_EQ((std::size_t)(w.second - w.first), 6u);
  EXPECT_EQ(c.DeltaGamma(1, 0), 9.0 - 100.0);
}

// uvlm_vortex_ring_test.cpp:25-85
TEST(uvlm_vortex_ring_init_wing) {
  const std::size_t rows = 5, cols = 10;
  const double dx = 0.1;
  std::vector<Vector3d> pos;
  for (std::size_t i = 0; i <= rows; i++)
    for (std::size_t j = 0; j <= cols; j++) pos.emplace_back(i * dx, j * dx, 0);
  UVLMVortexRing rings;
  rings.InitWing(pos, cols);
  const auto& v = rings.bound_vortices();
  EXPECT_EQ(v.size(), 2 * rows * cols);
  for (std::size_t i = 0; i < rows; i++)
    for (std::size_t j = 0; j < cols; j++) {
      const std::size_t idx = j + i * (2 * cols);
      EXPECT_DOUBLE_EQ(i * dx, v[idx].nodes()[0].x());
      EXPECT_DOUBLE_EQ(j * dx, v[idx].nodes()[0].y());
      EXPECT_DOUBLE_EQ((i + 1) * dx, v[idx].nodes()[2].x());
      EXPECT_DOUBLE_EQ((j + 1) * dx, v[idx].nodes()[2].y());
      const std::size_t m = j + cols + i * (2 * cols);
      EXPECT_DOUBLE_EQ(i * dx, v[m].nodes()[3].x());
      EXPECT_DOUBLE_EQ(-1. * j * dx, v[m].nodes()[3].y());
      EXPECT_DOUBLE_EQ(-1. * (j + 1) * dx, v[m].nodes()[0].y());
      EXPECT_DOUBLE_EQ((i + 1) * dx, v[m].nodes()[1].x());
    }
  const auto te = rings.TrailingEdgeRange();
  EXPECT_EQ(te.second - te.first, 2 * cols);
}

// wing_builder_test.cpp: a half wing of 2 x 3 rings with node (i, j) at
// (i, j, 0) becomes 2 x 6 rings; TransformRow mirrors in reverse order.
namespace {
proto::Wing half_wing_2x3(double oz) {
  proto::Wing w;
  w.set_rows(2);
  w.set_cols(3);
  for (int i = 0; i <= 2; i++)
    for (int j = 0; j <= 3; j++) {
      auto* p = w.add_points();
      p->set_x(i);
      p->set_y(j);
    }
  w.mutable_origin()->set_z(oz);
  return w;
}
}  // namespace

TEST(wing_builder_transform_row) {
  const std::vector<Vector3d> row{{1, 0, 2}, {1, 1, 2}, {1, 2, 2}};
  const auto r = internal::TransformRow(row.cbegin(), row.cend());
  EXPECT_EQ(r.size(), 5u);
  const double ys[5] = {-2, -1, 0, 1, 2};
  for (int k = 0; k < 5; k++) {
    EXPECT_DOUBLE_EQ(ys[k], r[k].y());
    EXPECT_DOUBLE_EQ(1, r[k].x());
    EXPECT_DOUBLE_EQ(2, r[k].z());
  }
  std::vector<internal::WingHolder> h(3);
  for (std::size_t k = 0; k < 3; k++) h[k].rows = h[k].cols = k + 1;
  EXPECT_EQ(WingBuilder::CountTotalSize(h), 14u);
}

TEST(wing_builder_build) {
  auto vortices = std::make_shared<std::vector<VortexRing>>();
  std::vector<VortexContainer> containers;
  WingBuilder builder(&containers, vortices);
  EXPECT_EQ(vortices->size(), 0u);
  builder.AddWing(half_wing_2x3(0)).AddWing(half_wing_2x3(1));
  const auto& h = builder.holders();
  EXPECT_EQ(h.size(), 2u);
  EXPECT_EQ(h[1].rows, 2u);
  EXPECT_EQ(h[1].cols, 6u);
  EXPECT_DOUBLE_EQ(0, h[1].points[0].x());
  EXPECT_DOUBLE_EQ(-3, h[1].points[0].y());
  for (std::size_t k : {0u, 1u, 6u}) EXPECT_DOUBLE_EQ(1, h[1].points[k].z());
  EXPECT_DOUBLE_EQ(0, h[0].points[0].z());
  builder.Build();
  EXPECT_EQ(vortices->size(), 24u);
  EXPECT_EQ(containers.size(), 2u);
  EXPECT_EQ(containers[0].rows(), 2u);
  EXPECT_EQ(containers[0].cols(), 6u);
  const VortexRing& v = containers[0].at(0, 0);
  EXPECT_EQ(v.nodes()[0], Vector3d(0, -3, 0));
  EXPECT_EQ(v.nodes0()[0], Vector3d(0, -3, 0));
  // ring (1, 5) of wing 1: nodes (1,5), (2,5), (2,6), (1,6) of the 7-column lattice
  const VortexRing& w = containers[1].at(1, 5);
  EXPECT_EQ(w.nodes()[0], Vector3d(1, 2, 1));
  EXPECT_EQ(w.nodes()[1], Vector3d(2, 2, 1));
  EXPECT_EQ(w.nodes()[2], Vector3d(2, 3, 1));
  EXPECT_EQ(w.nodes()[3], Vector3d(1, 3, 1));
  EXPECT_EQ(&(*vortices)[12 + 6 + 5], &w);  // id R C + j + i C
}

int main(int argc, char** argv) { return check::run(argc, argv); }
