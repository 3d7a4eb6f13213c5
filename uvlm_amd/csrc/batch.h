// batch.h -- the batched multi-case engine (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "internal.h"

namespace uvlm {

// Per-case constants in device memory.  Offsets are in points (3 doubles)
// for node/ring/edge/wake arrays and in elements otherwise.
struct BCase {
  int R, C, N, E;
  int cap;  // wake row capacity (node rows; Gamma rows use the same)
  int pad0;
  long node_off, ring_off, edge_off, A_off, wake_off, wgam_off, chb_off, chw_off;
  double dt, Q, ref_area, alpha;
  double ff[3], origin[3], plug[3], flap[3];
};

// Ring-buffer rows of a case's wake at step s (after the shed of step s the
// wake has s + 1 node rows and s circulation rows): node row i is physical
// row b_node_row0 + i, circulation row i is physical row b_gam_row0 + i.
// Both buffers hold cap = max_steps + 2 rows per case.
__host__ __device__ inline long b_node_row0(const BCase& c, int s) { return (long)c.cap - (s + 1); }
__host__ __device__ inline long b_gam_row0(const BCase& c, int s) { return (long)c.cap - s; }
// Load history: (CD, CL) of case cs at step `step` at hist[b_hist_index].
__host__ __device__ inline long b_hist_index(int step, int n_cases, int cs) {
  return 2 * ((long)step * n_cases + cs);
}
// Chain elements of a lattice of `rows` ring rows (node rows, then node
// columns; see k_b_pack_chains).
__host__ __device__ inline long b_chain_elems(int rows, int C) { return 2L * (rows + 1) * (C + 1); }

// Device-resident step counter and the work counters of the persistent
// N-body launches.
struct BState {
  int step;
  int pad;
  unsigned ctr[8];
};

// One case as handed in by the C-ABI (see uvlm_batch_case in uvlm_gpu.h).
struct BatchCaseSpec {
  size_t rows, cols;
  double chord, span, dt, alpha;
  double origin[3], ff[3], plug[3], flap[3];
};

// Offsets of every case in the concatenated arrays and the totals (host
// only, no device work): the plan BatchEngine::init allocates.
struct BatchPlan {
  std::vector<BCase> cases;
  long nodes = 0, rings = 0, edges = 0, A = 0, wake = 0, wgam = 0, cb = 0, cw = 0, te = 0;
  long hist = 0;  // doubles of the load history
  int maxN = 0;
};
int batch_plan(const BatchCaseSpec* specs, size_t n, size_t max_steps, BatchPlan* plan,
               std::string* err);
// Host replay of the store and load extents of every kernel of every step
// 0..max_steps of a plan (the launches of BatchEngine::step_once, with the
// index helpers above): 0 when all stay inside the allocations, else a
// description of the first one that does not.
int batch_check_extents(const BatchPlan& plan, size_t max_steps, std::string* err);

class BatchEngine {
 public:
  ~BatchEngine();
  int init(hipStream_t s, const BatchCaseSpec* specs, const double* wing_init, size_t n,
           size_t max_steps, int scheme, int ignore_steps, std::string* err);
  // Enqueue `steps` timesteps of every case; with use_graph the step is
  // captured once into a hipGraph and replayed.
  int run(size_t steps, bool use_graph, std::string* err);
  // Time every N-body launch shape on the current state and keep the best.
  int autotune(std::string* err, int* chosen, double* times_ms, int max_variants);
  int steps_done() const { return steps_done_; }
  int n_cases() const { return n_; }
  const BCase& case_info(int c) const { return cases_h_[c]; }
  int get_gamma(int c, double* out, std::string* err);
  int get_wake(int c, double* nodes, double* gamma, std::string* err);
  int get_loads(int c, double* out, size_t* count, std::string* err);
  int get_wing(int c, double* out, std::string* err);
  double pairs_per_step_at(int step) const;  // segment evaluations of one step
  // Run every case's solve through the partially pivoted path (test hook).
  void set_force_pivoted(bool on);

 private:
  void release();
  int build_items(int tile, std::string* err);
  int step_once(std::string* err);
  int nbody(int phase, const double* targets, double* out, int slot);
  int capture(std::string* err);

  hipStream_t stream_ = nullptr;
  int n_ = 0, max_steps_ = 0, scheme_ = 0, ign_ = 5, steps_done_ = 0, maxN_ = 0;
  int cus_ = 256, variant_ = 1, items_tile_ = 0;
  long tot_nodes_ = 0, tot_rings_ = 0, tot_edges_ = 0, tot_wake_ = 0, tot_cb_ = 0, tot_cw_ = 0,
       tot_te_ = 0;
  std::vector<BCase> cases_h_;
  std::vector<int> gemm_off_;
  std::vector<void*> allocs_;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  int n_rhs_ = 0, n_loads_ = 0, n_adv_ = 0, n_mat_ = 0;
  // device arrays
  BCase* d_cases_ = nullptr;
  BState* d_state_ = nullptr;
  int *d_node_case_ = nullptr, *d_ring_case_ = nullptr, *d_edge_case_ = nullptr,
      *d_wn_case_ = nullptr, *d_te_case_ = nullptr, *d_te_j_ = nullptr;
  long *d_chb_start_ = nullptr, *d_chw_start_ = nullptr;
  double *d_init_ = nullptr, *d_wing_ = nullptr, *d_cp0_ = nullptr, *d_mid0_ = nullptr,
         *d_vm_ = nullptr, *d_vme_ = nullptr, *d_gamma_ = nullptr, *d_gprev_ = nullptr,
         *d_rhs_ = nullptr, *d_uw_ = nullptr, *d_cp_ = nullptr, *d_nrm_ = nullptr,
         *d_area_ = nullptr, *d_mid_ = nullptr, *d_uflow_ = nullptr, *d_A_ = nullptr,
         *d_wake_ = nullptr, *d_wgam_ = nullptr, *d_pos_ = nullptr, *d_pst_ = nullptr,
         *d_k1_ = nullptr, *d_k2_ = nullptr, *d_k3_ = nullptr, *d_hist_ = nullptr;
  int* d_piv_ = nullptr;
  int* d_flag_ = nullptr;
  double* d_Asave_ = nullptr;
  double* d_dblk_ = nullptr;  // factored diagonal block of the current panel, per case
  long tot_A_ = 0;
  bool force_pivoted_ = false;
  RingRec* d_rings_ = nullptr;
  ChainRec *d_bch_ = nullptr, *d_wch_ = nullptr;
  int2 *d_it_rhs_ = nullptr, *d_it_loads_ = nullptr, *d_it_adv_ = nullptr;
  int4 *d_it_mat_ = nullptr, *d_it_gemm_ = nullptr;
  size_t cap_rhs_ = 0, cap_loads_ = 0, cap_adv_ = 0;
  std::vector<int> order_;
};

}  // namespace uvlm
