// simple_simulator.h -- the UVLM time loop of the host API
// (simulator/simple_simulator.h:20-110 of the reference).
//
// Same public interface (AddWing, Run, set_forward_flight, set_*_path,
// set_advection) plus stepping and accessors used by drivers and tests.  The
// wake lives in device memory; per step the host evaluates the kinematics
// (Morphing) and the GPU does shedding, matrix, right-hand side, solve,
// loads and advection (simple_simulator.cpp:240-286).
#pragma once

#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/uvlm_gpu.h"
#include "advect.h"
#include "morphing.h"
#include "multiple_sheet.h"
#include "proto_io.h"
#include "vector3d.h"
#include "wing.h"

namespace UVLM {
namespace simulator {

struct WingInformation {
  std::unique_ptr<wing::WingGenerator> generator;
  Morphing morphing;
  double chord, span;
  std::size_t rows, cols;
  Vector3d origin;
};

// simple_simulator.cpp:19-49 (host versions, used for the reference frame)
std::vector<Vector3d> CollocationPoints(const MultipleSheet<Vector3d>& pos);
std::vector<Vector3d> Normals(const MultipleSheet<Vector3d>& pos);

class SimpleSimulator {
 public:
  static const std::size_t STEP_MIN = 0;

  SimpleSimulator();
  ~SimpleSimulator();

  // simple_simulator.cpp:126-140: takes ownership of wing_generator.
  void AddWing(wing::WingGenerator* wing_generator, const Morphing& morphing,
               const double chord, const double span, const std::size_t rows,
               const std::size_t cols, const Vector3d& origin);
  // simple_simulator.cpp:288-313: steps+1 iterations (0..steps inclusive).
  void Run(const std::size_t steps, const double dt);

  void set_forward_flight(const Vector3d& v) { forward_flight_ = v; }
  // simple_simulator.cpp:52-66: every step, after the solve, a Snapshot2
  // record of the wing and wake rings (OutputPanels, :315-321) to the
  // result path and an AllVortexSheets record (OutputSheet, :323-332) to
  // the sheet path; the record framing is the one of proto_io.h.  Aborts on
  // an empty or unwritable path, as the reference does.
  void set_result_path(const std::string& path);
  void set_sheet_path(const std::string& path);
  // Load TSV "t CD0 CL0 CD1 CL1 ..." (simple_simulator.cpp:229-232,334-345).
  void set_load_path(const std::string& loadpath) { load_path_ = loadpath; }
  void set_advection(advect::Advection* advection) { advection_.reset(advection); }

  // --- extensions -----------------------------------------------------
  // UVLM_PREC_* of the Biot-Savart sums and UVLM_SOLVE_* of the solve.
  void set_precision(unsigned flags) { precision_ = flags; }
  void set_solve_mode(int mode) { solve_mode_ = mode; }
  // Core cut-off of the Biot-Savart sums (0: the reference's 1e-10;
  // uvlm_set_cutoff).
  void set_cutoff(double eps) { cutoff_ = eps; }
  // --erase_wake_after (simple_simulator.cpp:14,307-311); 0 disables.
  void set_erase_wake_after(std::size_t n) { erase_wake_after_ = n; }
  // Compute loads every step even without a load path.
  void set_compute_loads(bool on) { compute_loads_ = on; }
  // Build the wing and the device state for at most max_steps steps;
  // Run() calls it.  Then Step(step) runs one MainLoop (+ erase).
  void Prepare(std::size_t max_steps, double dt);
  void Step(std::size_t step);
  void Sync();

  std::size_t num() const { return wing_info_.size(); }
  std::size_t rows() const { return rows_; }
  std::size_t cols() const { return cols_; }
  const MultipleSheet<Vector3d>& wing_pos() const { return wing_pos_; }
  MultipleSheet<double> wing_gamma() const;
  void wake(MultipleSheet<Vector3d>* pos, MultipleSheet<double>* gamma) const;
  void set_wake(const MultipleSheet<Vector3d>& pos, const MultipleSheet<double>& gamma,
                const MultipleSheet<double>& wing_gamma);
  // (CD, CL) per wing of every step with loads: [steps][num][2]
  std::vector<double> load_history() const;
  double pairs_evaluated() const;
  double pairs_reference() const;
  uvlm_ctx* context() const { return ctx_; }

 private:
  void BuildWing();
  void MainLoop(const std::size_t step, const double dt);
  void WriteLoads() const;
  void OutputPanels(std::size_t step, double dt);
  void OutputSheet(std::size_t step, double dt);

  uvlm_ctx* ctx_ = nullptr;
  std::unique_ptr<advect::Advection> advection_;
  std::vector<WingInformation> wing_info_;
  std::vector<Morphing> morphings_;
  MultipleSheet<Vector3d> wing_pos_, wing_pos_init_;
  std::vector<Vector3d> cpos_init_;   // collocation points, reference frame
  std::vector<Vector3d> edge_mid0_;   // unique edge midpoints, reference frame
  std::vector<double> ref_area_;      // chord * span per wing
  std::vector<Vector3d> vm_cp_, vm_edges_;
  Vector3d forward_flight_;
  std::string result_path_, sheet_path_, load_path_;
  std::unique_ptr<std::ofstream> ofs_result_, ofs_sheet_;
  std::unique_ptr<RecordWriter> writer_, sheet_writer_;
  std::size_t rows_ = 0, cols_ = 0;
  unsigned precision_ = UVLM_PREC_F64_EXACT;
  int solve_mode_ = UVLM_SOLVE_F64;
  std::size_t erase_wake_after_ = 0;
  double cutoff_ = 0.0;
  bool compute_loads_ = false;
  bool prepared_ = false;
  double dt_ = 0;
  std::vector<double> load_t_;
};

}  // namespace simulator
}  // namespace UVLM
