// internal.h -- shared declarations of the gfx950 UVLM kernels.
//
// Data layout in HBM
//   * lattices: AoS double3, sheet-major, row-major (the reference's
//     MultipleSheet order K = j + cols*(i + rows*n));
//   * wake: a row ring buffer per sheet of `cap` rows, logical row i at
//     physical row (cap - rows + i), so prepending a row is O(cols);
//   * source rings: packed 128-byte records (4 corners + Gamma), the unit the
//     N-body kernels stage through LDS;
//   * partial sums of the N-body kernels: SoA [chunk][3][ld];
//   * dense matrices: column-major, leading dimension = n.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace uvlm {

// Cut-off length of the reference's BiotSavartLaw (vortex.cpp:31).
constexpr double kCutoff = 1e-10;
// A run's core cut-off (the reference hard-codes 1e-10; another value keeps
// the rule's shape): a pair is zero when |r1 x r2|^2 < eps, |r1| < eps or
// |r2| < eps.  td: the smallest x with fl(fl(sqrt(x))^2) >= eps, so that the
// exact form tests |d|^2 < td instead of squaring the rounded norm; eps2 =
// eps^2 for the forms that test squared lengths; short2: segments with
// |r0|^2 below it imply the end-point tests through the |d|^2 test
// (|d| <= |r1| (|r1| + |r0|) < eps when |r0|^2 < 0.1 / eps).
struct CutOff {
  double eps, td, eps2, short2;
};
CutOff make_cutoff(double eps);
// 1. / 4. / M_PI as the reference folds it (vortex.cpp:36).
constexpr double kInv4Pi = 1. / 4. / 3.14159265358979323846;

// One source ring: corners p0..p3 (12 doubles), Gamma, 3 pad = 128 bytes.
struct alignas(16) RingRec {
  double p[12];
  double g;
  double pad[3];
};
static_assert(sizeof(RingRec) == 128, "RingRec must be 128 bytes");

// One element of a segment chain (the FP64 fast and F32C source format):
// node p of a polyline; the segment from the previous node to p carries the
// net circulation g (already scaled by 1/4pi) and the vector r0 = p - prev.
// A chain start has g = 0 and r0 = 0 (its "segment" is always cut off).
// A vortex lattice becomes (rows+1) spanwise chains of cols+1 nodes and
// (cols+1) chordwise chains of rows+1 nodes per sheet; an interior segment
// shared by two rings carries the difference of their circulations.
struct alignas(16) ChainRec {
  double p[3];
  double g;
  double r0[3];
  double r0sq;
};
static_assert(sizeof(ChainRec) == 64, "ChainRec must be 64 bytes");
// The same element for the F32C kernel: the node and g in FP64 (for the
// FP64 x - p and the near-core re-evaluation), r0, the rounding band
// 2^-42 |r0|^2 and g rounded to FP32 once when packed rather than in every
// (target, element) pair.
struct alignas(16) ChainRec32 {
  double p[3];
  double g;
  float r0[3];
  float band;
  float gf;
  float pad[3];
};
static_assert(sizeof(ChainRec32) == 64, "ChainRec32 must be 64 bytes");

// Element count of the chains of a lattice with `rows` ring rows.
inline int chain_elems(int num, int rows, int cols) {
  return rows > 0 ? 2 * num * (rows + 1) * (cols + 1) : 0;
}
// Segments of those chains (elements minus chain starts).
inline double chain_segments(int num, int rows, int cols) {
  return rows > 0 ? (double)num * ((double)(rows + 1) * cols + (double)rows * (cols + 1))
                  : 0.0;
}

// Segment order / scaling variants of a ring sum.
enum SegMode : int {
  kSegVoring = 0,      // p0->p1, p1->p2, p2->p3, p3->p0, Gamma per segment
  kSegRingPerRing = 1, // node[(k+1)%4] -> node[k], sum then x Gamma
  kSegRingChord = 2,   // legacy, segments k in {0,2}, sum then x Gamma
  kSegRingPerSeg = 3,  // legacy order, Gamma per segment
};

enum Prec : int { kF64Exact = 0, kF64Fast = 1, kF32C = 2 };

// Source ranges of one N-body launch: rings [0, nb) split into cb chunks
// and rings [nb, nb + nw) split into cw chunks.
struct ChunkPlan {
  int nb, nw;     // ring counts of the two ranges
  int cb, cw;     // chunks per range
  int szb, szw;   // rings per chunk
  int total() const { return cb + cw; }
};

// slots > 0: workgroups the device holds at once for the kernel; the grid
// (target blocks x chunks) is then sized to about three full rounds of them,
// so that the last round is not left mostly empty.
ChunkPlan plan_chunks(int n_targets, int nb, int nw, bool sequential, int slots = 0);
// Resident workgroups of the chain N-body kernel of precision prec on the
// current device (occupancy x compute units; cached per device).
int chain_slots(int prec);

// ---- launchers (stream-ordered, asynchronous) ----------------------------
// Threshold T with  fl(fl(sqrt(x))^2) < eps  <=>  x < T  (x >= 0).
double cutoff_dsq_threshold(double eps = kCutoff);

void launch_pack_lattice(hipStream_t s, const double* nodes, int sheet_rows,
                         int row_off, const double* gamma, int g_sheet_rows,
                         int g_row_off, int num, int rows, int cols,
                         RingRec* out);
void launch_pack_rings(hipStream_t s, const double* rings, const double* gamma,
                       int n, RingRec* out);
void launch_set_gamma(hipStream_t s, const double* gamma, int n, RingRec* out);

// Chains of a lattice (layout as launch_pack_lattice) -> out[0, chain_elems).
void launch_pack_chains(hipStream_t s, const double* nodes, int sheet_rows,
                        int row_off, const double* gamma, int g_sheet_rows,
                        int g_row_off, int num, int rows, int cols,
                        ChainRec* out, bool f32 = false);
// Chain N-body (prec = kF64Fast or kF32C); same partial-sum layout and chunk
// plan semantics as launch_nbody, in units of chain elements.
void launch_nbody_chain(hipStream_t s, int prec, const double* targets,
                        int n_targets, const ChainRec* recs,
                        const ChunkPlan& plan, double* partial, int ld,
                        const CutOff& cut);

// Partial sums partial[(c*3 + d)*ld + t] of target t over chunk c.
void launch_nbody(hipStream_t s, int prec, int seg, const double* targets,
                  int n_targets, const RingRec* recs, const ChunkPlan& plan,
                  double* partial, int ld, const CutOff& cut);

// Independent segment evaluations out[k] = BiotSavartLaw(a[k], b[k], x[k]).
void launch_segments(hipStream_t s, int prec, const double* a, const double* b,
                     const double* x, int n, const CutOff& cut, double* out);

// out[t] = sum of all chunks (in chunk order), AoS.
void launch_sum_chunks(hipStream_t s, const double* partial, int ld,
                       int n_targets, int chunks, double* out);
// out[t] = ((-ff + sum_bound) + sum_wake), AoS (advect.h:16-24).  When
// te_period > 0, targets with (t % te_period) < te_width are trailing-edge
// wake nodes and receive -ff (advect.h:41-45).  ff is a host array.
void launch_flow_velocity(hipStream_t s, const double* partial, int ld,
                          int n_targets, int cb, int cw, const double* ff,
                          double* out, int te_period, int te_width);

// Dense solve (column-major A, overwritten by its LU factors).
size_t dense_lu_work_bytes(int n);
int* dense_lu_info(void* work);
int dense_lu_factor(hipStream_t s, double* A, int n, void* work);
void dense_lu_solve(hipStream_t s, const double* LU, int n, void* work, const double* b,
                    double* x);
// Mixed precision solve with FP64 iterative refinement; A is not modified.
// mode: 1 = FP32 LU, 2 = BF16-input LU updates.
int dense_mixed_solve(hipStream_t s, const double* A, int n, const double* b,
                      double* x, int mode, double tol, int max_iter,
                      void* work, size_t work_bytes, int* iters);
size_t dense_mixed_work_bytes(int n);
size_t dense_fullpiv_work_bytes(int n);
int dense_fullpiv_solve(hipStream_t s, double* A, int n, const double* b, double* x,
                        void* work, size_t work_bytes, int* rank_dev);

// Iterative refinement on a kept approximate inverse X ~ A^-1 (inv_solve.hip):
// x += X (b - A x) in FP64 until max|dx| <= tol max|x|; X is refreshed by
// Newton-Schulz steps when the contraction slows and rebuilt when the
// refinement diverges.  A (column-major) is not modified.
class InvSolver {
 public:
  ~InvSolver();
  // warm: start from the x passed in (the previous solution of a run; after
  // two warm solves the start is extrapolated from the last two solutions)
  int solve(hipStream_t s, const double* A, int n, const double* b, double* x, bool warm,
            double tol, int max_sweeps, int* sweeps, std::string* err);
  void invalidate() { valid_ = false; }
  // the next warm start is not the continuation of the last solve
  void reset_history() { streak_ = 0; }
  long long sweeps_total() const { return sweeps_total_; }
  long long refreshes() const { return refreshes_; }
  long long rebuilds() const { return rebuilds_; }

 private:
  // buffers for n unknowns, X in FP64 (wide) or FP32
  int prepare(hipStream_t s, int n, bool wide, std::string* err);
  int build(hipStream_t s, const double* A, std::string* err);
  // one Newton-Schulz step (low: BF16 operands); wait: read the deviation back
  double newton_schulz(hipStream_t s, const double* A, bool low, bool wait);
  void enqueue_sweep(hipStream_t s, const double* A, const double* b, double* x, int k,
                     double prev, double tol);
  void release();

  void *X_ = nullptr, *T_ = nullptr, *X2_ = nullptr;  // FP32, or FP64 when wide_
  float* Af_ = nullptr;                               // A rounded to FP32
  unsigned short *Ab_ = nullptr, *Xb_ = nullptr, *Tb_ = nullptr;  // BF16 operands
  double *r_ = nullptr, *xold_ = nullptr, *part_ = nullptr;
  bool wide_ = false;
  unsigned long long* stat_ = nullptr;
  int n_ = 0, S_ = 1, cols_ = 1;
  bool valid_ = false;
  int streak_ = 0, last_sweeps_ = 0;
  double end_rho_ = 1.0;  // contraction observed in the last solve
  long long sweeps_total_ = 0, refreshes_ = 0, rebuilds_ = 0;
  std::string last_err_;
};

}  // namespace uvlm
