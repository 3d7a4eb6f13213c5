// simple_simulator.cpp -- UVLM time loop on the device-resident state
// (simulator/simple_simulator.cpp:19-345 of the reference).
#include "simple_simulator.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "device.h"
#include "output.h"

namespace UVLM {
namespace simulator {

std::vector<Vector3d> CollocationPoints(const MultipleSheet<Vector3d>& pos) {
  std::vector<Vector3d> res;
  for (std::size_t n = 0; n < pos.num(); n++)
    for (std::size_t i = 0; i < pos.rows() - 1; i++)
      for (std::size_t j = 0; j < pos.cols() - 1; j++)
        res.push_back((((pos.at(n, i, j) + pos.at(n, i + 1, j)) + pos.at(n, i + 1, j + 1)) +
                       pos.at(n, i, j + 1)) /
                      4);
  return res;
}

std::vector<Vector3d> Normals(const MultipleSheet<Vector3d>& pos) {
  std::vector<Vector3d> res;
  for (std::size_t n = 0; n < pos.num(); n++)
    for (std::size_t i = 0; i < pos.rows() - 1; ++i)
      for (std::size_t j = 0; j < pos.cols() - 1; ++j) {
        Vector3d nrml = (pos.at(n, i + 1, j + 1) - pos.at(n, i, j))
                            .cross(pos.at(n, i, j + 1) - pos.at(n, i + 1, j));
        nrml.normalize();
        res.push_back(nrml);
      }
  return res;
}

namespace {
void die(const char* msg) {
  std::fprintf(stderr, "UVLM: %s\n", msg);
  std::abort();
}
}  // namespace

SimpleSimulator::SimpleSimulator() : forward_flight_(0, 0, 0) {}

SimpleSimulator::~SimpleSimulator() {
  if (ctx_) uvlm_destroy(ctx_);
}

void SimpleSimulator::set_result_path(const std::string& path) {
  if (path.empty()) die("Invalid path");
  ofs_result_.reset(new std::ofstream(path, std::ios::binary));
  if (!*ofs_result_) die("cannot open the result path");
  result_path_ = path;
  writer_.reset(new RecordWriter(ofs_result_.get()));
}

void SimpleSimulator::set_sheet_path(const std::string& path) {
  if (path.empty()) die("Invalid path");
  ofs_sheet_.reset(new std::ofstream(path, std::ios::binary));
  if (!*ofs_sheet_) die("cannot open the sheet path");
  sheet_path_ = path;
  sheet_writer_.reset(new RecordWriter(ofs_sheet_.get()));
}

// simple_simulator.cpp:315-321: the rings of the wing, then of the wake (no
// time is set on this message in the reference).
void SimpleSimulator::OutputPanels(std::size_t step, double dt) {
  proto::Snapshot2 snapshot;
  MultipleSheet<Vector3d> wpos;
  MultipleSheet<double> wgam;
  wake(&wpos, &wgam);
  output::SheetToSnapshot(&snapshot, wing_pos_, wing_gamma());
  output::SheetToSnapshot(&snapshot, wpos, wgam);
  if (!writer_->WriteProtocolMessage(snapshot)) die("result write failed");
}

// simple_simulator.cpp:323-332
void SimpleSimulator::OutputSheet(std::size_t step, double dt) {
  proto::AllVortexSheets sheets;
  sheets.has_t = sheets.has_wing = sheets.has_wake = true;
  sheets.t = step * dt;
  MultipleSheet<Vector3d> wpos;
  MultipleSheet<double> wgam;
  wake(&wpos, &wgam);
  sheets.wing = proto_adaptor::ToVortexSheet(wing_pos_, wing_gamma());
  sheets.wake = proto_adaptor::ToVortexSheet(wpos, wgam);
  if (!sheet_writer_->WriteProtocolMessage(sheets)) die("sheet write failed");
}

void SimpleSimulator::AddWing(wing::WingGenerator* wing_generator,
                              const Morphing& morphing, const double chord,
                              const double span, const std::size_t rows,
                              const std::size_t cols, const Vector3d& origin) {
  if (!wing_info_.empty()) {
    if (wing_info_.back().rows != rows) die("CHECK failed: wing_info_.rbegin()->rows == rows");
    if (wing_info_.back().cols != cols) die("CHECK failed: wing_info_.rbegin()->cols == cols");
  }
  if (cols & 1) die("columns must be even.");
  wing_info_.push_back(WingInformation{std::unique_ptr<wing::WingGenerator>(wing_generator),
                                       morphing, chord, span, rows, cols, origin});
}

void SimpleSimulator::BuildWing() {
  if (wing_info_.empty()) die("call AddWing at least once");
  const std::size_t num = wing_info_.size();
  rows_ = wing_info_.back().rows;
  cols_ = wing_info_.back().cols;
  wing_pos_.resize(num, rows_ + 1, cols_ + 1);
  wing_pos_init_.resize(num, rows_ + 1, cols_ + 1);
  morphings_.clear();
  ref_area_.clear();
  std::size_t n = 0;
  for (const auto& info : wing_info_) {
    morphings_.push_back(info.morphing);
    morphings_.back().set_origin(info.origin);
    proto::Wing whole, half;
    info.generator->Generate(&half, info.chord, info.span / 2, info.rows, info.cols / 2);
    wing::WholeWing(&whole, half);
    auto points = PointsToVector(whole.points());
    for (auto& p : points) p = p + info.origin;
    std::copy(points.begin(), points.end(), wing_pos_init_.sheet_begin(n));
    ref_area_.push_back(info.chord * info.span);
    ++n;
  }
  cpos_init_ = CollocationPoints(wing_pos_init_);
  // unique non-trailing-edge edges of each wing (uvlm_sim_loads order)
  edge_mid0_.clear();
  for (std::size_t w = 0; w < num; ++w) {
    for (std::size_t i = 0; i < rows_; ++i)
      for (std::size_t j = 0; j < cols_; ++j)
        edge_mid0_.push_back((wing_pos_init_.at(w, i, j) + wing_pos_init_.at(w, i, j + 1)) / 2);
    for (std::size_t i = 0; i < rows_; ++i)
      for (std::size_t j = 0; j <= cols_; ++j)
        edge_mid0_.push_back((wing_pos_init_.at(w, i, j) + wing_pos_init_.at(w, i + 1, j)) / 2);
  }
  vm_cp_.assign(cpos_init_.size(), Vector3d());
  vm_edges_.assign(edge_mid0_.size(), Vector3d());
}

void SimpleSimulator::Prepare(std::size_t max_steps, double dt) {
  BuildWing();
  if (!advection_) {
    std::fprintf(stderr, "UVLM: Advection is not set. Euler is used.\n");
    advection_.reset(new advect::Euler);
  }
  if (!ctx_) {
    int dev = 0;
    if (const char* e = std::getenv("UVLM_DEVICE")) dev = std::atoi(e);
    const int rc = uvlm_create(dev, &ctx_);
    if (rc != UVLM_OK) {
      std::fprintf(stderr, "UVLM: no usable GPU (uvlm_create returned %d)\n", rc);
      std::abort();
    }
  }
  if (cutoff_ > 0.0) gpu::Check(uvlm_set_cutoff(ctx_, cutoff_), "uvlm_set_cutoff", ctx_);
  gpu::Check(uvlm_sim_init(ctx_, num(), rows_, cols_, max_steps, precision_),
             "uvlm_sim_init", ctx_);
  dt_ = dt;
  load_t_.clear();
  prepared_ = true;
}

void SimpleSimulator::MainLoop(const std::size_t step, const double dt) {
  const double t = step * dt;
  const std::size_t num = wing_info_.size();
  const std::size_t per = (rows_ + 1) * (cols_ + 1);
  // Morphing (simple_simulator.cpp:245-253)
#pragma omp parallel for schedule(static)
  for (std::size_t k = 0; k < num * per; ++k)
    wing_pos_[k] = morphings_[k / per].Perfome(wing_pos_init_[k], t);
  gpu::Check(uvlm_sim_set_wing(ctx_, wing_pos_.data()->data()), "uvlm_sim_set_wing", ctx_);
  // Shed (:254-255)
  gpu::Check(uvlm_sim_shed(ctx_, step), "uvlm_sim_shed", ctx_);
  // CalcMatrix + CalcRhs + solve (:256-268)
  const std::size_t Nper = rows_ * cols_;
#pragma omp parallel for schedule(static)
  for (std::size_t K = 0; K < vm_cp_.size(); ++K)
    vm_cp_[K] = morphings_[K / Nper].Velocity(cpos_init_[K], t);
  // the loads' kinematic velocities do not depend on the solve: computed
  // here, while the device works, since a solve may wait for the device
  const bool loads = compute_loads_ || !load_path_.empty();
  if (loads) {
    const std::size_t Eper = rows_ * (2 * cols_ + 1);
#pragma omp parallel for schedule(static)
    for (std::size_t e = 0; e < vm_edges_.size(); ++e)
      vm_edges_[e] = morphings_[e / Eper].Velocity(edge_mid0_[e], t);
  }
  gpu::Check(uvlm_sim_solve(ctx_, vm_cp_[0].data(), forward_flight_.data(), solve_mode_,
                            nullptr),
             "uvlm_sim_solve", ctx_);
  // output (:271-278)
  if (writer_) OutputPanels(step, dt);
  if (sheet_writer_) OutputSheet(step, dt);
  // CalcLoad (:279-282)
  if (loads) {
    gpu::Check(uvlm_sim_loads(ctx_, vm_edges_[0].data(), forward_flight_.data(), dt,
                              ref_area_.data()),
               "uvlm_sim_loads", ctx_);
    load_t_.push_back(t);
  }
  // Advect (:283-284)
  gpu::Check(uvlm_sim_advect(ctx_, advection_->scheme(), advection_->ignore_steps(),
                             forward_flight_.data(), dt),
             "uvlm_sim_advect", ctx_);
}

void SimpleSimulator::Step(std::size_t step) {
  if (!prepared_) die("Prepare() must be called before Step()");
  MainLoop(step, dt_);
  if (erase_wake_after_ > 0 && step >= erase_wake_after_)
    gpu::Check(uvlm_sim_erase_oldest_wake(ctx_), "uvlm_sim_erase_oldest_wake", ctx_);
}

void SimpleSimulator::Sync() {
  if (ctx_) gpu::Check(uvlm_sync(ctx_), "uvlm_sync", ctx_);
}

void SimpleSimulator::Run(const std::size_t steps, const double dt) {
  Prepare(steps, dt);
  for (std::size_t step = STEP_MIN; step <= STEP_MIN + steps; step++) Step(step);
  Sync();
  if (!load_path_.empty()) WriteLoads();
}

MultipleSheet<double> SimpleSimulator::wing_gamma() const {
  MultipleSheet<double> g(num(), rows_, cols_);
  gpu::Check(uvlm_sim_get_gamma(ctx_, g.data()), "uvlm_sim_get_gamma", ctx_);
  return g;
}

void SimpleSimulator::wake(MultipleSheet<Vector3d>* pos, MultipleSheet<double>* gamma) const {
  std::size_t nr = 0, gr = 0;
  gpu::Check(uvlm_sim_wake_shape(ctx_, &nr, &gr), "uvlm_sim_wake_shape", ctx_);
  pos->resize(num(), nr, cols_ + 1);
  gamma->resize(num(), gr, cols_);
  gpu::Check(uvlm_sim_get_wake(ctx_, pos->size() ? pos->data()->data() : nullptr,
                               gamma->size() ? gamma->data() : nullptr),
             "uvlm_sim_get_wake", ctx_);
}

void SimpleSimulator::set_wake(const MultipleSheet<Vector3d>& pos,
                               const MultipleSheet<double>& gamma,
                               const MultipleSheet<double>& wing_gamma) {
  gpu::Check(uvlm_sim_set_wake(ctx_, pos.rows(), gamma.rows(),
                               pos.size() ? pos.data()->data() : nullptr,
                               gamma.size() ? gamma.data() : nullptr,
                               wing_gamma.size() ? wing_gamma.data() : nullptr),
             "uvlm_sim_set_wake", ctx_);
}

std::vector<double> SimpleSimulator::load_history() const {
  std::size_t count = load_t_.size();
  std::vector<double> out(2 * num() * count);
  gpu::Check(uvlm_sim_get_loads(ctx_, out.data(), &count), "uvlm_sim_get_loads", ctx_);
  out.resize(2 * num() * std::min(count, load_t_.size()));
  return out;
}

double SimpleSimulator::pairs_evaluated() const {
  double e = 0;
  uvlm_sim_pair_counts(ctx_, &e, nullptr);
  return e;
}
double SimpleSimulator::pairs_reference() const {
  double r = 0;
  uvlm_sim_pair_counts(ctx_, nullptr, &r);
  return r;
}

void SimpleSimulator::WriteLoads() const {
  std::ofstream ofs(load_path_);
  if (!ofs) die("cannot open the load path");
  ofs << "t";
  for (std::size_t n = 0; n < num(); n++) ofs << "\tCD" << n << "\tCL" << n;
  ofs << "\n";
  const auto h = load_history();
  // values in the stream's default format, as the reference writes them
  // (util::join over a std::stringstream, simple_simulator.cpp:231-232)
  for (std::size_t s = 0; s < load_t_.size(); ++s) {
    ofs << load_t_[s];
    for (std::size_t n = 0; n < num(); n++)
      ofs << "\t" << h[2 * (s * num() + n)] << "\t" << h[2 * (s * num() + n) + 1];
    ofs << "\n";
  }
}

}  // namespace simulator
}  // namespace UVLM
