// farfield.h -- near/far split of the wake's Biot-Savart sum (farfield.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

namespace uvlm {

// Vorticity moments of a block of wake rings about its centre c:
// Q0[i], Q1[i][j], Q2[i][j][k], Q3[i][j][k][l] = sum over its segments of
// Gamma e_i int_0^1 (a + t e - c)^{(x)n} dt (e = segment vector, a = its
// start), n = 0..3; 3 + 9 + 27 + 81 doubles.  A block with circulation zero
// outside it is a set of closed loops: Q0 vanishes up to rounding.
constexpr int kFFMom = 120;
constexpr int kFFQ1 = 3, kFFQ2 = 12, kFFQ3 = 39;  // offsets of Q1, Q2, Q3
// Level-0 clusters: blocks of kFFRows ring rows x kFFCols ring columns of
// one sheet; level l merges 2 x 2 blocks of level l - 1.
constexpr int kFFRows = 4;
constexpr int kFFCols = 16;
// Network correction of the expansion, one evaluation per (target tile,
// far cluster): inputs, hidden width, outputs.  The outputs are the
// expansion's remainder about the tile centre c_T as a quadratic polynomial
// in delta = (x - c_T) / (kMlpTaylorUnit D), D = |c_T - c|, in units of the
// leading omitted term: value y[i] (3), gradient y[3 + 3i + j] (9), Hessian
// y[12 + 6i + m] for (j, k) = 00, 01, 02, 11, 12, 22 (18).  The unit, a
// typical tile radius over D, weighs the three parts of the training loss
// by what they add over a tile.
constexpr int kMlpIn = 32;
constexpr int kMlpHidden = 128;
constexpr int kMlpOut = 30;
// Highest cluster level the network corrects (and is trained on).
constexpr int kMlpMaxLevel = 2;
// A (tile, cluster) pair is corrected when the tile's radius is at most
// this fraction of D (the quadratic then represents the remainder over the
// tile) and c_T lies outside the cluster's sphere.
constexpr double kMlpTileRatio = 0.3;
// Far-list entries of a tile that can carry a correction (the first ones).
constexpr int kNetCap = 256;
// Offset of the finite differences that make the training targets, in
// units of D.
constexpr double kMlpFdStep = 0.02;
constexpr double kMlpTaylorUnit = 0.15;

struct alignas(16) FFCluster {
  double c[3];  // expansion centre
  double r;     // radius: largest distance of its nodes from c
  int n, r0, r1, c0;  // sheet, ring rows [r0, r1), ring columns [c0, c1)
  int c1, level, pad0, pad1;
};
static_assert(sizeof(FFCluster) == 64, "FFCluster must be 64 bytes");

// Layout of a lattice whose ring blocks form the cluster tree.
struct FFTree {
  int num = 0, rows = 0, cols = 0, levels = 0, total = 0;
  int nbr[32], nbc[32], off[32], hr[32], hc[32];
};

// Weights of the correction network: h1 = silu(W1 x + b1), h2, h3 likewise
// (W2, W3: 128 x 128), y = W4 h3 + b4 (kMlpOut outputs).  FP32 master copy;
// the inference kernel reads BF16-rounded weights.
struct MlpShape {
  static constexpr int n_w1 = kMlpHidden * kMlpIn;
  static constexpr int n_w2 = kMlpHidden * kMlpHidden;
  static constexpr int n_w4 = kMlpOut * kMlpHidden;
  static constexpr int off_w1 = 0;
  static constexpr int off_b1 = off_w1 + n_w1;
  static constexpr int off_w2 = off_b1 + kMlpHidden;
  static constexpr int off_b2 = off_w2 + n_w2;
  static constexpr int off_w3 = off_b2 + kMlpHidden;
  static constexpr int off_b3 = off_w3 + n_w2;
  static constexpr int off_w4 = off_b3 + kMlpHidden;
  static constexpr int off_b4 = off_w4 + n_w4;
  static constexpr int total = off_b4 + kMlpOut + 2;  // two pad words
};

// Graded criterion for targets away from the bodies.  The focus is the
// bounding sphere of the bodies (centre fc, radius fr, from the device box
// min[3], max[3]); a target tile whose gap g to it (distance of the spheres'
// surfaces) exceeds L = reach fr uses theta L / g instead of theta, but not
// below theta_min.  A velocity error at distance g moves wake that acts on
// the bodies like a dipole, ~1/g^3, so accuracy there is spent where it
// matters.  reach <= 0: the plain criterion.
struct FFFocus {
  const double* box = nullptr;
  double reach = 0.0, theta_min = 0.0;
};

// out = (-ff + bound chunks in order) + wake (AoS), trailing-edge rows -ff.
void launch_ff_combine(hipStream_t s, const double* bound, int ld, int cb, const double* wake,
                       int nt, const double* ff, int te_period, int te_width, double* out);
// res[0] <- bits of max |a - b|, res[1] <- bits of max |b| over n 3-vectors.
// With a focus (reach > 0), |a - b| at point x[i] is weighted by
// min(1, L / g)^3 (g, L as in FFFocus: the error the graded criterion allows).
// Largest |a - b| and |b| into res (doubles as bit patterns); a may hold
// the points at a_off + t a_stride with trailing-edge rows skipped.
void launch_ff_maxdiff(hipStream_t s, const double* a, const double* b, int n,
                       unsigned long long* res, const double* x = nullptr,
                       const FFFocus* focus = nullptr, int a_off = 0, int a_stride = 1,
                       int te_period = 0, int te_width = 0);
void launch_gather_strided(hipStream_t s, const double* src, int off, int stride, int n,
                           double* x);
// Bounding box min[3], max[3] of n points x[n][3] (device).
void launch_ff_bbox(hipStream_t s, const double* x, int n, double* box);
// Rows [row0, row0+rows) of a ring buffer [num][cap][w] -> compact [num][rows][w].
void launch_gather_rows_w(hipStream_t s, const double* rb, int cap, int row0, int num, int rows,
                          int w, double* out);

class FarField {
 public:
  ~FarField();
  // Sources: the wake lattice nodes [num][rows+1][cols+1][3] with ring
  // circulations gamma [num][rows][cols] (device, compact), cut into blocks
  // of kFFRows x kFFCols rings merged 2 x 2 up to one block per sheet;
  // moments of every block and the near-sum record of every level-0 block.
  int build(hipStream_t s, const double* nodes, const double* gamma, int num, int rows, int cols,
            std::string* err);
  // Velocity of the built sources at nt targets (device [nt][3]): targets
  // are tiled 64 at a time in Morton order; a cluster whose distance from a
  // tile exceeds theta times its radius (beyond the tile's radius) uses its
  // moment expansion (plus the network's correction when enabled), the
  // level-0 blocks of the others are summed exactly.  out [nt][3] overwritten.
  // focus (optional): the graded criterion of FFFocus.  per_target: a
  // cluster is also far when each target (not the tile's sphere) clears
  // theta times its radius.  With the network, k_ff_tilenet evaluates it
  // once per (tile, far cluster) and the far kernel adds the corrections.
  // order_tag != 0: the targets are a set the caller identifies by this tag
  // (e.g. the wings' collocation points and edge midpoints), whose
  // arrangement changes little between calls: their Morton order is kept
  // from a previous call with the same tag and count and recomputed every
  // kOrderReuse calls.  The order only groups targets into tiles; every
  // tile's centre, radius and lists are formed from the current positions.
  int velocity(hipStream_t s, const double* targets, int nt, double theta, double* out,
               std::string* err, const FFFocus* focus = nullptr, bool per_target = false,
               int order_tag = 0);
  // Targets tiled in Morton order of their bounding box (own stable radix
  // sort of 30-bit keys): map[ntiles * 64] gets the target indices, -1
  // beyond nt; keys_out (device, optional) receives the unsorted keys.
  int morton_order(hipStream_t s, const double* targets, int nt, int ntiles, int* map,
                   unsigned* keys_out, std::string* err);
  // build + velocity.
  int wake_velocity(hipStream_t s, const double* nodes, const double* gamma, int num, int rows,
                    int cols, const double* targets, int nt, double theta, double* out,
                    std::string* err);
  // Velocity of the built sources (each sheet's root cluster) at nt points
  // laid out as sheets of per_sheet points in rows of width points, SoA
  // out [3][nt]: points in the first near_rows rows of each sheet take the
  // sum of the cb SoA chunks of near_part (a compact copy of those rows,
  // [num][near_rows][width], chunk stride near_ld); the others the root
  // clusters' moment expansions where their distance exceeds theta times
  // the radius, else the sheets' exact sums.
  int root_velocity(hipStream_t s, const double* x, int nt, int per_sheet, int width,
                    int near_rows, double theta, const double* near_part, int near_ld, int cb,
                    double* out, std::string* err);
  // Level-0 blocks summed exactly and clusters expanded, summed over the
  // tiles of the last velocity call (reads the list counts back).
  int list_stats(hipStream_t s, double* near_blocks, double* far_clusters, std::string* err);
  // Network: FP32 weights (host, MlpShape::total floats) and the trained
  // output range ymax (outputs beyond it trip to the exact cluster sum);
  // enable = false switches the correction off.
  int set_mlp(hipStream_t s, const float* w, bool enable, float ymax, std::string* err);
  bool mlp_enabled() const { return mlp_on_; }
  // Measurement of the near sum (k_ff_near): with profiling on, every launch
  // is bracketed by HIP events on its stream and counts its (target,
  // segment) pairs on the device; profile_read returns the launches, their
  // summed duration and pairs since the last read.
  void set_profile(bool on) { prof_ = on; }
  int profile_read(hipStream_t s, size_t* launches, double* ms, double* pairs, std::string* err);
  // Near blocks in FP32 with FP64 compensation and near-core re-evaluation
  // (the F32C precision) instead of FP64.
  void set_f32c(bool on) { f32c_ = on; }
  // Far pairs corrected by the network and pairs tripped to the exact sum,
  // since the network was set.
  int mlp_stats(hipStream_t s, unsigned long long* pairs, unsigned long long* trips,
                std::string* err);
  // Training samples from the interaction lists of the last velocity call
  // (same targets): up to max_samples (tile, far cluster) pairs that the
  // network corrects, drawn with a seeded generator; features [k][kMlpIn]
  // at the tile centre and the remainder's Taylor coefficients [k][kMlpOut]
  // (exact - expansion, normalised) to host.
  int sample_residuals(hipStream_t s, const double* targets, int nt, int max_samples,
                       unsigned seed, std::vector<float>* feats, std::vector<float>* resid,
                       std::string* err);

 private:
  int ensure(void** p, size_t* cap, size_t bytes, std::string* err);


  FFTree tree_;
  bool built_ = false, mlp_on_ = false, f32c_ = false;
  float ymax_ = 0.0f;
  // device buffers
  void *d_rec_ = nullptr, *d_key_ = nullptr, *d_key2_ = nullptr,
       *d_idx_ = nullptr, *d_idx2_ = nullptr, *d_hist_ = nullptr, *d_box_ = nullptr,
       *d_cl_ = nullptr, *d_mom_ = nullptr, *d_map_ = nullptr, *d_tiles_ = nullptr,
       *d_near_ = nullptr, *d_far_ = nullptr, *d_cnt_ = nullptr, *d_front_ = nullptr,
       *d_part_ = nullptr, *d_w_ = nullptr, *d_wf_ = nullptr, *d_mstat_ = nullptr,
       *d_samp_ = nullptr, *d_corr_ = nullptr, *d_tr_ = nullptr, *d_nrow_ = nullptr;
  size_t c_rec_ = 0, c_key_ = 0, c_key2_ = 0, c_idx_ = 0, c_idx2_ = 0,
         c_hist_ = 0, c_box_ = 0, c_cl_ = 0, c_mom_ = 0, c_map_ = 0, c_tiles_ = 0,
         c_near_ = 0, c_far_ = 0, c_cnt_ = 0, c_front_ = 0, c_part_ = 0, c_w_ = 0, c_wf_ = 0,
         c_mstat_ = 0, c_samp_ = 0, c_corr_ = 0, c_tr_ = 0, c_nrow_ = 0;
  // kept Morton order of a tagged target set
  void* d_tmap_ = nullptr;
  size_t c_tmap_ = 0;
  int tmap_tag_ = 0, tmap_nt_ = -1, tmap_uses_ = 0;
  // near-sum measurement
  bool prof_ = false, np_live_ = false;
  void* d_np_ = nullptr;
  size_t c_np_ = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_used_, ev_free_;
  // state of the last velocity call (for sampling and statistics)
  int last_ntiles_ = 0, last_nt_ = 0;
};

}  // namespace uvlm
