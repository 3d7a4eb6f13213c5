// host_capi.cpp -- C entry points (include/uvlm_host.h) over
// UVLM::simulator::SimpleSimulator.
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>

#include "../../include/uvlm_host.h"
#include "legacy_simulator.h"
#include "output.h"
#include "simple_simulator.h"
#include "tools.h"

using UVLM::Morphing;
using UVLM::Vector3d;
using UVLM::simulator::SimpleSimulator;

struct uvlm_host_sim {
  SimpleSimulator sim;
  std::size_t num = 0, R = 0, C = 0;
};

namespace {
thread_local std::string g_err;
// output options of the next uvlm_host_legacy_run
struct LegacyOutputs {
  std::string output_path, load_path;
  bool use_joukowski = false;
};
LegacyOutputs g_legacy;

bool readable(const char* path) {
  std::ifstream f(path, std::ios::binary);
  return static_cast<bool>(f);
}
bool writable(const char* path) {
  std::ofstream f(path, std::ios::app);
  return static_cast<bool>(f);
}

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
  if (sp->cutoff > 0.0) h->sim.set_cutoff(sp->cutoff);
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
  sim::SetOutputPath(g_legacy.output_path);
  sim::SetOutputLoadPath(g_legacy.load_path);
  sim::SetUseJoukowski(g_legacy.use_joukowski);
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

int uvlm_host_sim_set_outputs(uvlm_host_sim* s, const char* result_path,
                              const char* sheet_path) {
  if (!s) return err("null simulation");
  if (result_path) {
    if (!*result_path || !writable(result_path)) return err("cannot open the result path");
    s->sim.set_result_path(result_path);
  }
  if (sheet_path) {
    if (!*sheet_path || !writable(sheet_path)) return err("cannot open the sheet path");
    s->sim.set_sheet_path(sheet_path);
  }
  return UVLM_OK;
}

int uvlm_host_legacy_set_outputs(const char* output_path, const char* load_path,
                                 int use_joukowski) {
  g_legacy.output_path = output_path ? output_path : "";
  g_legacy.load_path = load_path ? load_path : "";
  g_legacy.use_joukowski = use_joukowski != 0;
  return UVLM_OK;
}

namespace {
// Sheets of an explicit state in the MultipleSheet layout.
void sheets_of(const double* nodes, const double* gamma, size_t num, size_t rows, size_t cols,
               UVLM::MultipleSheet<Vector3d>* pos, UVLM::MultipleSheet<double>* g) {
  pos->resize(num, rows + 1, cols + 1);
  g->resize(num, rows, cols);
  for (size_t k = 0; k < pos->size(); ++k)
    (*pos)[k] = Vector3d(nodes[3 * k], nodes[3 * k + 1], nodes[3 * k + 2]);
  for (size_t k = 0; k < g->size(); ++k) (*g)[k] = gamma[k];
}
int put_message(const std::string& m, unsigned char* out, size_t* len_inout) {
  if (!len_inout) return err("null length");
  const size_t cap = *len_inout;
  *len_inout = m.size();
  if (!out) return UVLM_OK;
  if (cap < m.size()) return err("output buffer too small");
  std::memcpy(out, m.data(), m.size());
  return UVLM_OK;
}
}  // namespace

int uvlm_host_encode_panels(const double* wing_nodes, const double* wing_gamma, size_t num,
                            size_t rows, size_t cols, const double* wake_nodes,
                            const double* wake_gamma, size_t wake_rows, unsigned char* out,
                            size_t* len_inout) {
  if (!wing_nodes || !wing_gamma || (wake_rows && (!wake_nodes || !wake_gamma)))
    return err("null argument");
  UVLM::MultipleSheet<Vector3d> wp, kp;
  UVLM::MultipleSheet<double> wg, kg;
  sheets_of(wing_nodes, wing_gamma, num, rows, cols, &wp, &wg);
  UVLM::proto::Snapshot2 snapshot;
  UVLM::output::SheetToSnapshot(&snapshot, wp, wg);
  if (wake_rows) {
    sheets_of(wake_nodes, wake_gamma, num, wake_rows, cols, &kp, &kg);
    UVLM::output::SheetToSnapshot(&snapshot, kp, kg);
  }
  return put_message(UVLM::proto::Serialize(snapshot), out, len_inout);
}

int uvlm_host_encode_sheets(const double* wing_nodes, const double* wing_gamma, size_t num,
                            size_t rows, size_t cols, const double* wake_nodes,
                            const double* wake_gamma, size_t wake_rows, double t,
                            unsigned char* out, size_t* len_inout) {
  if (!wing_nodes || !wing_gamma || (wake_rows && (!wake_nodes || !wake_gamma)))
    return err("null argument");
  UVLM::MultipleSheet<Vector3d> wp, kp;
  UVLM::MultipleSheet<double> wg, kg;
  sheets_of(wing_nodes, wing_gamma, num, rows, cols, &wp, &wg);
  sheets_of(wake_rows ? wake_nodes : wing_nodes, wake_rows ? wake_gamma : wing_gamma, num,
            wake_rows, cols, &kp, &kg);
  UVLM::proto::AllVortexSheets sheets;
  sheets.has_t = sheets.has_wing = sheets.has_wake = true;
  sheets.t = t;
  sheets.wing = UVLM::proto_adaptor::ToVortexSheet(wp, wg);
  sheets.wake = UVLM::proto_adaptor::ToVortexSheet(kp, kg);
  return put_message(UVLM::proto::Serialize(sheets), out, len_inout);
}

int uvlm_host_velocity_at(const char* input, const double pos[3], int id, int wing, int wake,
                          const char* output, size_t* records) {
  if (!input || !pos || !output) return err("null argument");
  if (!(wing || wake)) return err("At least one of --wing or --wake must be true.");
  if (id < 0) return err("id must be >= 0");
  if (!readable(input)) return err("cannot open the input");
  if (!writable(output)) return err("cannot open the output");
  const std::size_t n = UVLM::tools::VelocityAt(input, Vector3d(pos[0], pos[1], pos[2]), id,
                                                wing != 0, wake != 0, output);
  if (records) *records = n;
  return UVLM_OK;
}

namespace {
// One Snapshot2: a file holding one message (recordio = 0) or record `index`
// of a record stream.
int load_snapshot(const char* path, int recordio, size_t index, UVLM::proto::Snapshot2* s) {
  if (!path) return err("null argument");
  if (!readable(path)) return err("cannot open the snapshot");
  if (!recordio) {
    *s = UVLM::tools::ReadSnapshot2(path);
    return UVLM_OK;
  }
  std::ifstream ifs(path, std::ios::binary);
  UVLM::RecordReader reader(&ifs);
  for (size_t k = 0;; ++k) {
    if (!reader.ReadProtocolMessage(s)) return err("no such record in the snapshot stream");
    if (k == index) return UVLM_OK;
  }
}
}  // namespace

int uvlm_host_snapshot_velocity(const char* snapshot, int recordio, size_t index,
                                const double* points, size_t n, double* out) {
  if (n && (!points || !out)) return err("null argument");
  UVLM::proto::Snapshot2 s;
  if (const int rc = load_snapshot(snapshot, recordio, index, &s)) return rc;
  std::vector<Vector3d> p(n);
  for (size_t k = 0; k < n; ++k) p[k] = Vector3d(points[3 * k], points[3 * k + 1], points[3 * k + 2]);
  const auto v = UVLM::tools::SnapshotVelocity(s, p);
  for (size_t k = 0; k < n; ++k)
    for (int d = 0; d < 3; ++d) out[3 * k + d] = v[k][d];
  return UVLM_OK;
}

int uvlm_host_plane_velocity(const char* snapshot, int recordio, size_t index,
                             const char* plane, const double range[4], double at,
                             size_t resolution, const char* output, const char* filetype) {
  if (!plane || !range || !output || !filetype) return err("null argument");
  const std::string pl(plane), ft(filetype);
  if (pl.size() != 2 || pl[0] == pl[1] || std::string("xyz").find(pl[0]) == std::string::npos ||
      std::string("xyz").find(pl[1]) == std::string::npos)
    return err("plane must be two different letters of 'xyz'");
  if (ft != "csv" && ft != "tsv") return err("filetype must be csv or tsv");
  if (resolution < 2) return err("resolution must be at least 2");
  if (!writable(output)) return err("cannot open the output");
  UVLM::proto::Snapshot2 s;
  if (const int rc = load_snapshot(snapshot, recordio, index, &s)) return rc;
  UVLM::tools::WritePlane(s, pl, range[0], range[1], range[2], range[3], at, resolution, output,
                          ft);
  return UVLM_OK;
}

int uvlm_host_grid_velocity(const char* snapshot, int recordio, size_t index,
                            const double lo[3], const double hi[3], size_t resolution,
                            const char* output, size_t* points) {
  if (!lo || !hi || !output) return err("null argument");
  if (resolution < 2) return err("resolution must be at least 2");
  if (!writable(output)) return err("cannot open the output");
  UVLM::proto::Snapshot2 s;
  if (const int rc = load_snapshot(snapshot, recordio, index, &s)) return rc;
  const size_t n = UVLM::tools::WriteGrid(s, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2],
                                          resolution, output);
  if (points) *points = n;
  return UVLM_OK;
}

int uvlm_host_calc_impulse(const char* snapshot, double* t, double impulse[3],
                           double other_term[3], double wake_impulse[3]) {
  UVLM::proto::Snapshot2 s;
  if (const int rc = load_snapshot(snapshot, 0, 0, &s)) return rc;
  if (s.container_shapes.empty()) return err("snapshot without containers");
  size_t first = (size_t)s.container_shapes[0].rows * s.container_shapes[0].cols;
  if (s.morphing_velocities.size() < first) return err("snapshot without morphing velocities");
  const auto r = UVLM::tools::CalcImpulse(s);
  if (t) *t = r.t;
  for (int d = 0; d < 3; ++d) {
    if (impulse) impulse[d] = r.impulse[d];
    if (other_term) other_term[d] = r.other_term[d];
    if (wake_impulse) wake_impulse[d] = r.wake_impulse[d];
  }
  return UVLM_OK;
}

int uvlm_host_lamb_vector(const double* ring, double gamma, const double* v_nodes,
                          const double freestream[3], const double* rings,
                          const double* ring_gamma, size_t n_rings, double out[3]) {
  if (!ring || !v_nodes || !freestream || !out || (n_rings && (!rings || !ring_gamma)))
    return err("null argument");
  UVLM::VortexRing v;
  for (int k = 0; k < 4; ++k) v.PushNode(ring[3 * k], ring[3 * k + 1], ring[3 * k + 2]);
  v.set_gamma(gamma);
  std::vector<Vector3d> vn(4);
  for (int k = 0; k < 4; ++k) vn[k] = Vector3d(v_nodes[3 * k], v_nodes[3 * k + 1], v_nodes[3 * k + 2]);
  std::vector<UVLM::VortexRing> src(n_rings);
  for (size_t q = 0; q < n_rings; ++q) {
    for (int k = 0; k < 4; ++k)
      src[q].PushNode(rings[12 * q + 3 * k], rings[12 * q + 3 * k + 1], rings[12 * q + 3 * k + 2]);
    src[q].set_gamma(ring_gamma[q]);
  }
  const Vector3d r = UVLM::calc_impulse::CalcLambVectorOnPanel(
      v, vn, Vector3d(freestream[0], freestream[1], freestream[2]), src);
  for (int d = 0; d < 3; ++d) out[d] = r[d];
  return UVLM_OK;
}

}  // extern "C"
