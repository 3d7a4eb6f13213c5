// host_capi.cpp -- C entry points (include/uvlm_host.h) over
// UVLM::simulator::SimpleSimulator.
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/uvlm_host.h"
#include "legacy_simulator.h"
#include "simple_simulator.h"

using UVLM::Morphing;
using UVLM::Vector3d;
using UVLM::simulator::SimpleSimulator;

struct uvlm_host_sim {
  SimpleSimulator sim;
  std::size_t num = 0, R = 0, C = 0;
};

namespace {
thread_local std::string g_err;

int err(const char* msg) {
  g_err = msg;
  return UVLM_ERR_ARG;
}

// Motion laws of the case spec (see uvlm_host.h).
Morphing make_morphing(const uvlm_wing_spec& w) {
  Morphing m;
  m.set_alpha(w.alpha);
  const double pa = w.plug_amp, po = w.plug_omega, pp = w.plug_phase;
  m.set_plug([pa, po, pp](double t) { return pa * std::cos(po * t + pp); });
  const double fa = w.flap_amp, fo = w.flap_omega, fp = w.flap_phase;
  m.set_flap([fa, fo, fp](double t) { return fa * std::cos(fo * t + fp); });
  struct Fourier {
    double amp[4][2], ph[4][2];
  };
  Fourier tw, be;
  std::memcpy(tw.amp, w.twist_amp, sizeof(tw.amp));
  std::memcpy(tw.ph, w.twist_phase, sizeof(tw.ph));
  std::memcpy(be.amp, w.bend_amp, sizeof(be.amp));
  std::memcpy(be.ph, w.bend_phase, sizeof(be.ph));
  const double span = w.span, chord = w.chord, om = w.fourier_omega;
  m.set_twist([tw, span, om](const Vector3d& x, double t) {
    const double eta = x.y() / (span / 2);
    double s = 0;
    for (int mm = 0; mm < 4; mm++)
      for (int q = 0; q < 2; q++)
        s = s + tw.amp[mm][q] * std::sin((mm + 1) * M_PI_2 * eta) *
                    std::cos((q + 1) * om * t + tw.ph[mm][q]);
    return s;
  });
  m.set_bend([be, span, chord, om](const Vector3d& x, double t) {
    const double eta = x.y() / (span / 2);
    const double xi = x.x() / chord;
    double s = 0;
    for (int mm = 0; mm < 4; mm++)
      for (int q = 0; q < 2; q++)
        s = s + be.amp[mm][q] * std::sin(M_PI * xi) * std::sin((mm + 1) * M_PI_2 * eta) *
                    std::cos((q + 1) * om * t + be.ph[mm][q]);
    return s;
  });
  return m;
}
}  // namespace

extern "C" {

const char* uvlm_host_last_error(void) { return g_err.c_str(); }

int uvlm_host_build_wing(const uvlm_wing_spec* w, size_t rows, size_t cols, double* nodes) {
  if (!w || !nodes) return err("null argument");
  if (rows == 0 || cols == 0 || (cols & 1)) return err("columns must be even and non-zero");
  std::unique_ptr<UVLM::wing::WingGenerator> gen;
  if (w->generator == UVLM_GEN_NACA4)
    gen.reset(new UVLM::wing::NACA4digitGenerator(w->digit));
  else if (w->generator == UVLM_GEN_NACA00)
    gen.reset(new UVLM::wing::NACA00XXGenerator(w->digit));
  else
    gen.reset(new UVLM::wing::RectGenerator());
  // SimpleSimulator::BuildWing for one wing (simulator/simple_simulator.cpp:142-169)
  UVLM::proto::Wing whole, half;
  gen->Generate(&half, w->chord, w->span / 2, rows, cols / 2);
  UVLM::wing::WholeWing(&whole, half);
  const auto pts = UVLM::PointsToVector(whole.points());
  if (pts.size() != (rows + 1) * (cols + 1)) return err("generator returned a wrong lattice");
  const Vector3d o(w->origin[0], w->origin[1], w->origin[2]);
  for (std::size_t k = 0; k < pts.size(); ++k) {
    const Vector3d p = pts[k] + o;
    nodes[3 * k + 0] = p.x();
    nodes[3 * k + 1] = p.y();
    nodes[3 * k + 2] = p.z();
  }
  return UVLM_OK;
}

int uvlm_host_sim_create(const uvlm_case_spec* sp, size_t max_steps, uvlm_host_sim** out) {
  if (!sp || !out) return err("null argument");
  *out = nullptr;
  if (sp->num_wings < 1 || sp->num_wings > UVLM_HOST_MAX_WINGS) return err("num_wings out of range");
  if (sp->cols & 1) return err("columns must be even.");
  if (sp->rows == 0 || sp->cols == 0) return err("empty lattice");
  if (sp->scheme < 0 || sp->scheme > 4) return err("unknown advection scheme");
  auto* h = new uvlm_host_sim();
  for (int n = 0; n < sp->num_wings; ++n) {
    const uvlm_wing_spec& w = sp->wings[n];
    UVLM::wing::WingGenerator* gen = nullptr;
    if (w.generator == UVLM_GEN_NACA4)
      gen = new UVLM::wing::NACA4digitGenerator(w.digit);
    else if (w.generator == UVLM_GEN_NACA00)
      gen = new UVLM::wing::NACA00XXGenerator(w.digit);
    else
      gen = new UVLM::wing::RectGenerator();
    h->sim.AddWing(gen, make_morphing(w), w.chord, w.span, sp->rows, sp->cols,
                   Vector3d(w.origin[0], w.origin[1], w.origin[2]));
  }
  h->sim.set_forward_flight(Vector3d(sp->ff[0], sp->ff[1], sp->ff[2]));
  UVLM::advect::Advection* adv = nullptr;
  switch (sp->scheme) {
    case UVLM_ADVECT_EULER: adv = new UVLM::advect::Euler; break;
    case UVLM_ADVECT_RK2: adv = new UVLM::advect::RungeKutta2(sp->ignore_steps); break;
    case UVLM_ADVECT_RK4: adv = new UVLM::advect::RungeKutta4(sp->ignore_steps); break;
    case UVLM_ADVECT_AB2: adv = new UVLM::advect::AdamsBashforth2; break;
    default: adv = new UVLM::advect::RungeKutta4Classic(sp->ignore_steps); break;
  }
  h->sim.set_advection(adv);
  h->sim.set_precision(sp->precision);
  h->sim.set_solve_mode(sp->solve_mode);
  h->sim.set_erase_wake_after(sp->erase_wake_after);
  h->sim.set_compute_loads(sp->loads != 0);
  h->sim.Prepare(max_steps, sp->dt);
  h->num = (std::size_t)sp->num_wings;
  h->R = sp->rows;
  h->C = sp->cols;
  *out = h;
  return UVLM_OK;
}

void uvlm_host_sim_destroy(uvlm_host_sim* s) { delete s; }

int uvlm_host_sim_step(uvlm_host_sim* s, size_t step) {
  if (!s) return err("null sim");
  s->sim.Step(step);
  return UVLM_OK;
}

int uvlm_host_sim_sync(uvlm_host_sim* s) {
  if (!s) return err("null sim");
  s->sim.Sync();
  return UVLM_OK;
}

int uvlm_host_sim_wing_nodes(uvlm_host_sim* s, double* out) {
  if (!s || !out) return err("null argument");
  const auto& p = s->sim.wing_pos();
  std::memcpy(out, p.data()->data(), sizeof(double) * 3 * p.size());
  return UVLM_OK;
}

int uvlm_host_sim_gamma(uvlm_host_sim* s, double* out) {
  if (!s || !out) return err("null argument");
  const auto g = s->sim.wing_gamma();
  std::memcpy(out, g.data(), sizeof(double) * g.size());
  return UVLM_OK;
}

int uvlm_host_sim_wake_shape(uvlm_host_sim* s, size_t* node_rows, size_t* gamma_rows) {
  if (!s) return err("null sim");
  return uvlm_sim_wake_shape(s->sim.context(), node_rows, gamma_rows);
}

int uvlm_host_sim_wake(uvlm_host_sim* s, double* nodes, double* gamma) {
  if (!s) return err("null sim");
  return uvlm_sim_get_wake(s->sim.context(), nodes, gamma);
}

int uvlm_host_sim_set_wake(uvlm_host_sim* s, size_t node_rows, size_t gamma_rows,
                           const double* nodes, const double* gamma,
                           const double* wing_gamma) {
  if (!s) return err("null sim");
  return uvlm_sim_set_wake(s->sim.context(), node_rows, gamma_rows, nodes, gamma, wing_gamma);
}

int uvlm_host_sim_loads(uvlm_host_sim* s, double* out, size_t* count_inout) {
  if (!s || !count_inout) return err("null argument");
  return uvlm_sim_get_loads(s->sim.context(), out, count_inout);
}

int uvlm_host_sim_pairs(uvlm_host_sim* s, double* evaluated, double* reference) {
  if (!s) return err("null sim");
  return uvlm_sim_pair_counts(s->sim.context(), evaluated, reference);
}

uvlm_ctx* uvlm_host_sim_context(uvlm_host_sim* s) { return s ? s->sim.context() : nullptr; }


int uvlm_host_legacy_run(const uvlm_case_spec* sp, size_t steps, double* bound_nodes,
                         double* bound_gamma, double* wake_nodes, double* wake_gamma,
                         size_t* nw_inout) {
  namespace sim = UVLM::simulator;
  if (!sp) return err("null argument");
  if (sp->num_wings < 1 || sp->num_wings > UVLM_HOST_MAX_WINGS) return err("num_wings out of range");
  if (sp->cols & 1) return err("columns must be even.");
  if (sp->rows == 0 || sp->cols == 0) return err("empty lattice");
  if ((wake_nodes || wake_gamma) && !nw_inout) return err("wake capacity missing");
  sim::ResetSimulator();
  for (int n = 0; n < sp->num_wings; ++n) {
    const uvlm_wing_spec& w = sp->wings[n];
    std::unique_ptr<UVLM::wing::WingGenerator> gen;
    if (w.generator == UVLM_GEN_NACA4)
      gen.reset(new UVLM::wing::NACA4digitGenerator(w.digit));
    else if (w.generator == UVLM_GEN_NACA00)
      gen.reset(new UVLM::wing::NACA00XXGenerator(w.digit));
    else
      gen.reset(new UVLM::wing::RectGenerator());
    UVLM::proto::Wing half;
    gen->Generate(&half, w.chord, w.span / 2, sp->rows, sp->cols / 2);
    half.mutable_origin()->set_x(w.origin[0]);
    half.mutable_origin()->set_y(w.origin[1]);
    half.mutable_origin()->set_z(w.origin[2]);
    sim::AddWing(half, make_morphing(w));
  }
  sim::SetInlet(-sp->ff[0], -sp->ff[1], -sp->ff[2]);
  sim::Start(steps, sp->dt);
  const auto& v = sim::Vortices();
  const std::size_t nb = (std::size_t)sp->num_wings * sp->rows * sp->cols;
  if (v.size() < nb) return err("legacy loop lost its bound rings");
  const std::size_t nw = v.size() - nb;
  if (nw_inout) {
    if ((wake_nodes || wake_gamma) && *nw_inout < nw) return err("wake capacity too small");
    *nw_inout = nw;
  }
  auto put = [](const UVLM::VortexRing& r, double* nodes, double* g, std::size_t k) {
    if (nodes)
      for (int q = 0; q < 4; ++q) {
        nodes[12 * k + 3 * q + 0] = r.nodes()[q].x();
        nodes[12 * k + 3 * q + 1] = r.nodes()[q].y();
        nodes[12 * k + 3 * q + 2] = r.nodes()[q].z();
      }
    if (g) g[k] = r.gamma();
  };
  for (std::size_t k = 0; k < nb; ++k) put(v[k], bound_nodes, bound_gamma, k);
  for (std::size_t k = 0; k < nw; ++k) put(v[nb + k], wake_nodes, wake_gamma, k);
  return UVLM_OK;
}

}  // extern "C"
