// farfield.h -- near/far split of the wake's Biot-Savart sum (farfield.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace uvlm {

// Vorticity moments of a cluster about its centre c: Q1[i][j], Q2[i][j][k],
// Q3[i][j][k][l] = sum over its segments of Gamma e_i int_0^1 (a + t e - c)^{(x)n} dt
// (e = segment vector, a = its start), n = 1, 2, 3; 9 + 27 + 81 doubles.
constexpr int kFFMom = 117;
// Level-0 clusters: blocks of kFFRows ring rows x kFFCols ring columns.
constexpr int kFFRows = 4;
constexpr int kFFCols = 16;
// MLP correction of the expansion: inputs, hidden width, hidden layers.
constexpr int kMlpIn = 32;
constexpr int kMlpHidden = 128;
// Highest cluster level the network corrects (and is trained on).
constexpr int kMlpMaxLevel = 3;

struct alignas(16) FFCluster {
  double c[3];  // expansion centre
  double r;     // radius: largest distance of its nodes from c
  int n, r0, r1, c0;
  int c1, level, pad0, pad1;
};
static_assert(sizeof(FFCluster) == 64, "FFCluster must be 64 bytes");

// Layout of a lattice whose ring blocks form the cluster tree.
struct FFTree {
  int num = 0, rows = 0, cols = 0, levels = 0, total = 0;
  int nbr[32], nbc[32], off[32], hr[32], hc[32];
};

// Weights of the correction network: h1 = silu(W1 x + b1), h2, h3 likewise
// (W2, W3: 128 x 128), y = W4 h3 + b4 (3 outputs).  FP32 master copy; the
// inference kernel reads BF16-rounded weights.
struct MlpShape {
  static constexpr int n_w1 = kMlpHidden * kMlpIn;
  static constexpr int n_w2 = kMlpHidden * kMlpHidden;
  static constexpr int n_w4 = 3 * kMlpHidden;
  static constexpr int off_w1 = 0;
  static constexpr int off_b1 = off_w1 + n_w1;
  static constexpr int off_w2 = off_b1 + kMlpHidden;
  static constexpr int off_b2 = off_w2 + n_w2;
  static constexpr int off_w3 = off_b2 + kMlpHidden;
  static constexpr int off_b3 = off_w3 + n_w2;
  static constexpr int off_w4 = off_b3 + kMlpHidden;
  static constexpr int off_b4 = off_w4 + n_w4;
  static constexpr int total = off_b4 + 4;
};

// out = (-ff + bound chunks in order) + wake (AoS), trailing-edge rows -ff.
void launch_ff_combine(hipStream_t s, const double* bound, int ld, int cb, const double* wake,
                       int nt, const double* ff, int te_period, int te_width, double* out);
// res[0] <- bits of max |a - b|, res[1] <- bits of max |b| over n 3-vectors.
void launch_ff_maxdiff(hipStream_t s, const double* a, const double* b, int n,
                       unsigned long long* res);
// Rows [row0, row0+rows) of a ring buffer [num][cap][w] -> compact [num][rows][w].
void launch_gather_rows_w(hipStream_t s, const double* rb, int cap, int row0, int num, int rows,
                          int w, double* out);

class FarField {
 public:
  ~FarField();
  // Wake-induced velocity at nt targets of the lattice nodes [num][rows+1]
  // [cols+1][3], gamma [num][rows][cols] (device, compact).  Clusters whose
  // distance from a target tile exceeds theta times their radius (beyond the
  // tile's own radius) use the moment expansion (plus the network's
  // correction when one is loaded and enabled); all others are summed
  // exactly.  tile_targets: nullptr for tiles of 64 consecutive targets, or
  // ntiles x 64 target indices (-1 = empty slot).  out [nt][3] overwritten.
  int wake_velocity(hipStream_t s, const double* nodes, const double* gamma, int num,
                    int rows, int cols, const double* targets, int nt,
                    const int* tile_targets, int ntiles, double theta, double* out,
                    std::string* err);
  // Targets at the nodes of a wake lattice, tiled 4 node rows x 16 node
  // columns: builds and caches the tile map for (num, rows, cols).
  const int* lattice_tiles(hipStream_t s, int num, int rows, int cols, int* ntiles,
                           std::string* err);
  // Pair counts of the last call: exact segment evaluations, far (target,
  // cluster) evaluations.
  double last_near_pairs() const { return near_pairs_; }
  double last_far_pairs() const { return far_pairs_; }
  // Network: FP32 weights (host, MlpShape::total floats) and the trained
  // output range ymax (outputs beyond it trip to the exact cluster sum);
  // enable = false switches the correction off.
  int set_mlp(hipStream_t s, const float* w, bool enable, float ymax, std::string* err);
  bool mlp_enabled() const { return mlp_on_; }
  // Far pairs corrected by the network and pairs tripped to the exact sum,
  // since the network was set.
  int mlp_stats(hipStream_t s, unsigned long long* pairs, unsigned long long* trips,
                std::string* err);
  // Training samples from the interaction lists of the last wake_velocity
  // call (same targets): up to max_samples far (target, cluster) pairs of
  // clusters up to kMlpMaxLevel, drawn with a seeded generator; features
  // [k][kMlpIn] and normalised remainders [k][3] (exact - expansion) to host.
  int sample_residuals(hipStream_t s, const double* targets, int nt, int max_samples,
                       unsigned seed, std::vector<float>* feats, std::vector<float>* resid,
                       std::string* err);

 private:
  int ensure(void** p, size_t* cap, size_t bytes, std::string* err);
  void build_tree(int num, int rows, int cols);

  FFTree tree_;
  double near_pairs_ = 0.0, far_pairs_ = 0.0;
  bool mlp_on_ = false;
  float ymax_ = 0.0f;
  // device buffers
  void *d_cl_ = nullptr, *d_mom_ = nullptr, *d_tiles_ = nullptr, *d_near_ = nullptr,
       *d_far_ = nullptr, *d_cnt_ = nullptr, *d_part_ = nullptr, *d_map_ = nullptr,
       *d_w_ = nullptr, *d_samp_ = nullptr, *d_front_ = nullptr, *d_rec_ = nullptr,
       *d_wf_ = nullptr, *d_mstat_ = nullptr;
  size_t c_cl_ = 0, c_mom_ = 0, c_tiles_ = 0, c_near_ = 0, c_far_ = 0, c_cnt_ = 0,
         c_part_ = 0, c_map_ = 0, c_front_ = 0, c_rec_ = 0, c_w_ = 0, c_samp_ = 0, c_wf_ = 0,
         c_mstat_ = 0;
  int map_num_ = -1, map_rows_ = -1, map_cols_ = -1, map_tiles_ = 0;
  // state of the last wake_velocity call (for sampling)
  int last_ntiles_ = 0, last_nt_ = 0;
  const int* last_map_ = nullptr;
  double last_theta_ = 0.0;
};

}  // namespace uvlm
