// farfield.hip -- near/far split of the wake's Biot-Savart sum on gfx950.
//
// Sources: the wake lattice of every sheet is cut into blocks of kFFRows x
// kFFCols rings (level 0); level l merges 2 x 2 blocks of level l - 1
// (columns stop doubling at the sheet width).  A block with circulation zero
// outside it is a set of closed loops, so the blocks of any partition sum
// to the whole lattice and a block's field falls like a dipole's.  Inside a
// block, lattice neighbours share nodes and segments: the exact sum walks
// its segments once (circulation differences) as chains with 1/|x - p|
// carried from node to node.
//
// Targets (any points) are sorted in Morton order and tiled 64 at a time.
// A breadth-first traversal per tile classifies blocks: far when
// |c - c_tile| - r_tile > theta r, near at level 0, otherwise opened.  Far
// blocks use the multipole expansion of the vector potential
// A = 1/4pi sum Gamma int dl / |x - y| in the block's moments Q0..Q3:
//   u_a = 1/4pi sum_{n=0..3} (-1)^n / n! eps_abc Q_n[c, J] T_{n+1}[b, J],
// T_m = the m-th derivative tensor of 1/r.  Level-0 moments are integrated
// segment by segment (exact polynomials in the segment parameter); parents
// translate their children's.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/uvlm_gpu.h"
#include "devmath.h"
#include "farfield.h"
#include "mlp.h"

namespace uvlm {
namespace {

__host__ __device__ inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }
constexpr int kTileT = 64;        // targets per tile
constexpr int kGroups = 8;        // partial-sum groups per tile, far clusters
constexpr int kNearGroups = 64;   // partial-sum groups per tile, near blocks
// The near walk loads the next block's record into registers during the
// current walk (else it loads each record when it needs it, fewer registers).
constexpr bool kNearPrefetch = true;
constexpr int kMomThreads = 128;  // >= kFFMom: one thread per moment component

// --- tree layout ------------------------------------------------------------
__host__ __device__ inline int ff_id(const FFTree& t, int l, int n, int bi, int bj) {
  return t.off[l] + (n * t.nbr[l] + bi) * t.nbc[l] + bj;
}
// Level-0 block shapes: nodes and segments of a full block.
constexpr int kL0Nodes = (kFFRows + 1) * (kFFCols + 1);
constexpr int kL0Segs = (kFFRows + 1) * kFFCols + (kFFCols + 1) * kFFRows;
// Near-sum record of a level-0 block: its nodes (row-major, nc + 1 per row),
// its segment strengths / 4pi (spanwise segments by node row, then
// chordwise by node column) and, for the FP32 walk, each segment's vector
// and squared length rounded to FP32 (4 floats per segment, same order) and
// its strength rounded to FP32; contiguous so one wavefront loads it
// coalesced.  The FP64 walk stages nodes and strengths (the first kRecSeg
// doubles), the FP32 walk nodes, vectors and FP32 strengths (kRec32 doubles
// in LDS, the FP64 strengths skipped).
constexpr int kRecStr = 3 * kL0Nodes;
constexpr int kRecSeg = kRecStr + kL0Segs;  // FP32 segment vectors (as floats)
constexpr int kRecG = kRecSeg + 2 * kL0Segs;  // FP32 strengths (as floats)
constexpr int kRec = kRecG + kL0Segs / 2;
constexpr int kRec32 = kRec - kL0Segs;  // staged by the FP32 walk
static_assert(kL0Segs % 2 == 0, "FP32 strengths fill whole doubles");

// Bounding box of n points x[n][3]: box[0..2] min, box[3..5] max (one
// workgroup).
__global__ __launch_bounds__(1024) void k_bbox(const double* __restrict__ x, int n,
                                               double* __restrict__ box) {
  __shared__ double s[6][16];
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int q = threadIdx.x; q < n; q += 1024) {
    double p[3];
    for (int d = 0; d < 3; ++d) p[d] = x[3 * (long)q + d];
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], p[d]);
      hi[d] = fmax(hi[d], p[d]);
    }
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], __shfl_xor(lo[d], off));
      hi[d] = fmax(hi[d], __shfl_xor(hi[d], off));
    }
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 3; ++d) {
      s[d][threadIdx.x >> 6] = lo[d];
      s[3 + d][threadIdx.x >> 6] = hi[d];
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int d = threadIdx.x;
    double v = s[d][0];
    for (int w = 1; w < 16; ++w) v = d < 3 ? fmin(v, s[d][w]) : fmax(v, s[d][w]);
    box[d] = v;
  }
}

// Bounding box of many points in two launches: kBoxParts workgroups each
// reduce a contiguous range to part[g][6] (min x, y, z, max x, y, z), then
// one wavefront reduces the parts in order into box[6].
constexpr int kBoxParts = 64;
__global__ __launch_bounds__(256) void k_bbox_part(const double* __restrict__ x, int n,
                                                   double* __restrict__ part) {
  __shared__ double s[6][4];
  const int per = (n + kBoxParts - 1) / kBoxParts;
  const int q0 = blockIdx.x * per, q1 = min(n, q0 + per);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int q = q0 + threadIdx.x; q < q1; q += 256)
    for (int d = 0; d < 3; ++d) {
      const double v = x[3 * (long)q + d];
      lo[d] = fmin(lo[d], v);
      hi[d] = fmax(hi[d], v);
    }
  for (int off = 32; off > 0; off >>= 1)
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], __shfl_xor(lo[d], off));
      hi[d] = fmax(hi[d], __shfl_xor(hi[d], off));
    }
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 3; ++d) {
      s[d][threadIdx.x >> 6] = lo[d];
      s[3 + d][threadIdx.x >> 6] = hi[d];
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int d = threadIdx.x;
    double v = s[d][0];
    for (int w = 1; w < 4; ++w) v = d < 3 ? fmin(v, s[d][w]) : fmax(v, s[d][w]);
    part[blockIdx.x * 6 + d] = v;
  }
}

__global__ __launch_bounds__(64) void k_bbox_fin(const double* __restrict__ part,
                                                 double* __restrict__ box) {
  const int d = threadIdx.x;
  if (d >= 6) return;
  double v = part[d];
  for (int g = 1; g < kBoxParts; ++g) v = d < 3 ? fmin(v, part[g * 6 + d]) : fmax(v, part[g * 6 + d]);
  box[d] = v;
}

// --- stable radix sort of the targets' Morton keys ---------------------------
// LSD over the 30-bit keys (with their target indices): kRsPasses passes of
// 8-bit digits.  Large sets: the keys are cut into nb <= kRsMaxRows rows of
// `per` keys (a multiple of kRsChunk); per pass, k_rs_hist counts each
// row's digits (LDS atomics) and k_rs_pass places the keys of row b with
// digit d at
//   (keys of digits < d) + (keys of digit d in rows < b)
//   + (keys of digit d before it in row b),
// the last term from wavefront ballots over chunks of kRsChunk keys.  Small
// sets (<= kRsSmall keys, the wings' targets): one workgroup computes the
// keys and runs every pass in one launch (k_rs_small).  Stable, so the
// order equals any stable sort's of the same keys.
constexpr int kMortonBits = 10;  // per axis: cells of 1/1024 of the box
constexpr int kRsBits = 8, kRsBins = 1 << kRsBits, kRsPasses = 4;  // 32 >= 30 key bits
constexpr int kRsThreads = 256, kRsItems = 16, kRsChunk = kRsThreads * kRsItems;
constexpr int kRsMaxRows = 128;
constexpr int kOrderReuse = 50;  // calls a tagged target set keeps its order
constexpr int kRsSmallThreads = 1024, kRsSmallItems = 16;
constexpr int kRsSmall = kRsSmallThreads * kRsSmallItems;
struct RsPlan {
  int nb, per;
};
inline RsPlan rs_plan(int n) {
  const int per = cdiv(cdiv(n, kRsMaxRows), kRsChunk) * kRsChunk;
  return RsPlan{cdiv(n, per), per};
}

// Bits 0..9 of v spread to bits 0, 3, 6, ..., 27.
__device__ inline unsigned spread10(unsigned v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// 30-bit Morton key of point q in the box (10 bits per axis, x in bit 0 of
// each triple): cells of 1/1024 of the box, far finer than a tile of 64
// targets; points of one cell keep their index order.
__device__ inline unsigned morton_key(const double* __restrict__ x, long q,
                                      const double* __restrict__ box) {
  unsigned k = 0;
  for (int d = 0; d < 3; ++d) {
    const double p = x[3 * q + d];
    const double w = box[3 + d] - box[d];
    double t = w > 0.0 ? (p - box[d]) / w : 0.0;
    t = fmin(fmax(t, 0.0), 1.0);
    const unsigned c = (unsigned)(t * (double)((1 << kMortonBits) - 1));
    k |= spread10(c) << d;
  }
  return k;
}

__global__ void k_morton(const double* __restrict__ x, int n, const double* __restrict__ box,
                         unsigned* __restrict__ key, int* __restrict__ idx) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  key[q] = morton_key(x, q, box);
  if (idx) idx[q] = q;
}

// Lanes of this wavefront with the caller's digit (among lanes with ok).
__device__ __forceinline__ unsigned long long rs_peers(int d, bool ok) {
  unsigned long long m = __ballot(ok);
  for (int b = 0; b < kRsBits; ++b) {
    const bool on = (d >> b) & 1;
    const unsigned long long x = __ballot(on);
    m &= on ? x : ~x;
  }
  return m;
}

// hist[b][256]: digit counts of row b of the keys.
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const unsigned* __restrict__ kin, int n,
                                                         int per, int shift,
                                                         int* __restrict__ hist) {
  __shared__ int h[kRsBins];
  const int t = threadIdx.x, b = blockIdx.x;
  h[t] = 0;
  __syncthreads();
  const int lo = b * per, hi = min(n, lo + per);
  for (int e = lo + t; e < hi; e += kRsThreads) atomicAdd(&h[(kin[e] >> shift) & (kRsBins - 1)], 1);
  __syncthreads();
  hist[b * kRsBins + t] = h[t];
}

// One pass over the keys (kin, iin) -> (kout, iout): kout may be null (the
// last pass needs no keys); iout gets -1 at positions n..pad_to-1 (the tile
// map's padding).
__global__ __launch_bounds__(kRsThreads) void k_rs_pass(
    const unsigned* __restrict__ kin, const int* __restrict__ iin,
    unsigned* __restrict__ kout, int* __restrict__ iout, int n, int per, int nb,
    int shift, const int* __restrict__ hist, int pad_to) {
  __shared__ int cnt[kRsItems][4][kRsBins];  // per (item, wavefront, digit)
  __shared__ int base[kRsBins];
  __shared__ int wsum[4];
  const int t = threadIdx.x, b = blockIdx.x, lane = t & 63, w = t >> 6;
  const int lo = b * per, hi = min(n, lo + per);
  // the first chunk's keys are in flight while the offsets are formed
  unsigned key[kRsItems];
  int id[kRsItems], rk[kRsItems];
  for (int i = 0; i < kRsItems; ++i) {
    const int e = lo + i * kRsThreads + t;
    key[i] = e < hi ? kin[e] : 0u;
    id[i] = e < hi ? iin[e] : -1;
  }
  // thread t owns digit t: keys of digit t in all rows and in rows before b
  // (16 rows' counts loaded before they are summed)
  int total = 0, pre = 0;
  for (int r0 = 0; r0 < nb; r0 += 16) {
    int v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = r0 + q < nb ? hist[(r0 + q) * kRsBins + t] : 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      total += v[q];
      pre += r0 + q < b ? v[q] : 0;
    }
  }
  // exclusive prefix of the totals over the digits
  int inc = total;
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(inc, off);
    if (lane >= off) inc += u;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int before = 0;
  for (int q = 0; q < w; ++q) before += wsum[q];
  int run = before + inc - total + pre;  // next free position of digit t for this row
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int c0 = lo; c0 < hi; c0 += kRsChunk) {
    if (c0 > lo)
      for (int i = 0; i < kRsItems; ++i) {
        const int e = c0 + i * kRsThreads + t;
        key[i] = e < hi ? kin[e] : 0u;
        id[i] = e < hi ? iin[e] : -1;
      }
    for (int i = 0; i < kRsItems; ++i)
      for (int q = 0; q < 4; ++q) cnt[i][q][t] = 0;
    base[t] = run;
    __syncthreads();
    for (int i = 0; i < kRsItems; ++i) {
      const bool ok = c0 + i * kRsThreads + t < hi;
      const int d = (int)((key[i] >> shift) & (kRsBins - 1));
      const unsigned long long peers = rs_peers(d, ok);
      rk[i] = __popcll(peers & lt);
      if (ok && rk[i] == 0) cnt[i][w][d] = __popcll(peers);
    }
    __syncthreads();
    // digit t: exclusive prefix over (item, wavefront), the keys' order
    // (all counts read before the first is rewritten)
    int s = 0;
    {
      int c[kRsItems * 4];
#pragma unroll
      for (int i = 0; i < kRsItems; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) c[4 * i + q] = cnt[i][q][t];
#pragma unroll
      for (int i = 0; i < kRsItems; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          cnt[i][q][t] = s;
          s += c[4 * i + q];
        }
    }
    __syncthreads();
    for (int i = 0; i < kRsItems; ++i) {
      if (c0 + i * kRsThreads + t >= hi) continue;
      const int d = (int)((key[i] >> shift) & (kRsBins - 1));
      const int dst = base[d] + cnt[i][w][d] + rk[i];
      if (kout) kout[dst] = key[i];
      iout[dst] = id[i];
    }
    run += s;
    __syncthreads();  // cnt and base are rewritten by the next chunk
  }
  if (b == 0)
    for (int q = n + t; q < pad_to; q += kRsThreads) iout[q] = -1;
}

// The whole sort of n <= kRsSmall keys in one workgroup: keys of the points
// x in the box to (kA, iA), kRsPasses passes between (kA, iA) and (kB, iB)
// in global memory (the workgroup's barriers order its own global accesses),
// digit counts per (item, wavefront) in LDS; the last pass writes the tile
// map (-1 at n..pad_to-1).
__global__ __launch_bounds__(kRsSmallThreads) void k_rs_small(
    const double* __restrict__ x, int n, const double* __restrict__ box, unsigned* kA, int* iA,
    unsigned* kB, int* iB, int* __restrict__ map, int pad_to) {
  constexpr int kW = kRsSmallThreads / 64;  // wavefronts
  __shared__ unsigned short cnt[kRsSmallItems][kW][kRsBins];
  __shared__ int part[4][kRsBins];  // per quarter of the (item, wavefront) sequence
  __shared__ int base[kRsBins];
  __shared__ int wsum[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int dd = t & (kRsBins - 1), qt = t >> 8;  // digit and quarter this thread serves
  for (int e = t; e < n; e += kRsSmallThreads) {
    kA[e] = morton_key(x, e, box);
    iA[e] = e;
  }
  __syncthreads();
  const int items = (n + kRsSmallThreads - 1) / kRsSmallThreads;
  const int L = items * kW, f0 = qt * L / 4, f1 = (qt + 1) * L / 4;  // this quarter's entries
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int p = 0; p < kRsPasses; ++p) {
    const bool last = p + 1 == kRsPasses;
    const unsigned* kin = (p & 1) ? kB : kA;
    const int* iin = (p & 1) ? iB : iA;
    unsigned* kout = (p & 1) ? kA : kB;
    int* iout = last ? map : (p & 1) ? iA : iB;
    const int shift = p * kRsBits;
    for (int f = f0; f < f1; ++f) cnt[f / kW][f % kW][dd] = 0;
    unsigned key[kRsSmallItems];
    int id[kRsSmallItems], rk[kRsSmallItems];
#pragma unroll
    for (int i = 0; i < kRsSmallItems; ++i) {
      const int e = i * kRsSmallThreads + t;
      key[i] = i < items && e < n ? kin[e] : 0u;
      id[i] = i < items && e < n ? iin[e] : -1;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kRsSmallItems; ++i) {
      if (i >= items) continue;  // uniform
      const bool ok = i * kRsSmallThreads + t < n;
      const int d = (int)((key[i] >> shift) & (kRsBins - 1));
      const unsigned long long peers = rs_peers(d, ok);
      rk[i] = __popcll(peers & lt);
      if (ok && rk[i] == 0) cnt[i][w][d] = (unsigned short)__popcll(peers);
    }
    __syncthreads();
    // digit dd: sums of the quarters, totals, exclusive prefix over digits
    int s = 0;
    for (int f = f0; f < f1; ++f) s += cnt[f / kW][f % kW][dd];
    part[qt][dd] = s;
    __syncthreads();
    int qoff = 0, total = 0;
    for (int q = 0; q < 4; ++q) {
      const int v = part[q][dd];
      qoff += q < qt ? v : 0;
      total += v;
    }
    if (qt == 0) {  // the first four wavefronts scan the 256 totals
      int inc = total;
      for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(inc, off);
        if (lane >= off) inc += u;
      }
      if (lane == 63) wsum[w] = inc;
      base[dd] = inc - total;  // within this wavefront's 64 digits
    }
    __syncthreads();
    int run = base[dd] + qoff;
    for (int q = 0; q < (dd >> 6); ++q) run += wsum[q];
    for (int f = f0; f < f1; ++f) {
      const int c = cnt[f / kW][f % kW][dd];
      cnt[f / kW][f % kW][dd] = (unsigned short)run;
      run += c;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kRsSmallItems; ++i) {
      if (i >= items) continue;
      if (i * kRsSmallThreads + t >= n) continue;
      const int d = (int)((key[i] >> shift) & (kRsBins - 1));
      const int dst = cnt[i][w][d] + rk[i];
      if (!last) kout[dst] = key[i];
      iout[dst] = id[i];
    }
    __syncthreads();  // this pass's global writes before the next pass reads them
  }
  for (int q = n + t; q < pad_to; q += kRsSmallThreads) map[q] = -1;
}

// Component m of the moments: type 0 (m < 3), 1 (< 12), 2 (< 39), 3 (< 120)
// with indices i (of e), j, k, l.
struct MomIndex {
  int type, i, j, k, l;
};
__device__ inline MomIndex mom_index(int m) {
  MomIndex x{0, m, 0, 0, 0};
  if (m < kFFQ1) return x;
  if (m < kFFQ2) {
    const int r = m - kFFQ1;
    return MomIndex{1, r / 3, r % 3, 0, 0};
  }
  if (m < kFFQ3) {
    const int r = m - kFFQ2;
    return MomIndex{2, r / 9, (r / 3) % 3, r % 3, 0};
  }
  const int r = m - kFFQ3;
  return MomIndex{3, r / 27, (r / 9) % 3, (r / 3) % 3, r % 3};
}

// Component of the moments of one segment (A = a - c, E = e); the thirds
// are products with the rounded 1/3 (an FP64 division in this per-segment
// loop costs ~100 cycles of latency).
constexpr double kThird = 1.0 / 3.0;
__device__ __forceinline__ double segment_moment(const MomIndex& x, double g, const double* A,
                                                 const double* E) {
  const double ge = g * E[x.i];
  const int j = x.j, k = x.k, l = x.l;
  if (x.type == 0) return ge;
  if (x.type == 1) return ge * (A[j] + 0.5 * E[j]);
  if (x.type == 2) {
    const double p2 = A[j] * A[k] + 0.5 * (A[j] * E[k] + E[j] * A[k]) + E[j] * E[k] * kThird;
    return ge * p2;
  }
  const double p3 = A[j] * A[k] * A[l] +
                    0.5 * (A[j] * A[k] * E[l] + A[j] * E[k] * A[l] + E[j] * A[k] * A[l]) +
                    (A[j] * E[k] * E[l] + E[j] * A[k] * E[l] + E[j] * E[k] * A[l]) * kThird +
                    0.25 * E[j] * E[k] * E[l];
  return ge * p3;
}

// One workgroup per level-0 block: nodes staged in LDS, centre (node mean,
// in node order) and radius, the block's segments (spanwise node rows, then
// chordwise node columns; strength = difference of the in-block
// circulations either side), then one thread per moment component summing
// over the segments in order; the block's near record is written on the way.
__global__ __launch_bounds__(kMomThreads) void k_ff_level0(FFTree t,
                                                           const double* __restrict__ nodes,
                                                           const double* __restrict__ gamma,
                                                           FFCluster* __restrict__ cl,
                                                           double* __restrict__ mom,
                                                           double* __restrict__ rec) {
  __shared__ double sp[kL0Nodes * 3];
  __shared__ double sA[kL0Segs * 3], sE[kL0Segs * 3], sg[kL0Segs];
  __shared__ double sc[3], srad[kMomThreads / 64];
  const int q = blockIdx.x, tid = threadIdx.x;
  const int bj = q % t.nbc[0], bi = (q / t.nbc[0]) % t.nbr[0], n = q / (t.nbc[0] * t.nbr[0]);
  const int C1 = t.cols + 1;
  const int r0 = bi * kFFRows, r1 = min(t.rows, r0 + kFFRows);
  const int c0 = bj * kFFCols, c1 = min(t.cols, c0 + kFFCols);
  const int nr = r1 - r0, nc = c1 - c0, nn = (nr + 1) * (nc + 1);
  const double* P = nodes + 3 * (long)n * (t.rows + 1) * C1;
  const double* G = gamma + (long)n * t.rows * t.cols;
  double* R = rec + (long)q * kRec;
  for (int k = tid; k < nn; k += kMomThreads) {
    const double* p = P + 3 * ((long)(r0 + k / (nc + 1)) * C1 + c0 + k % (nc + 1));
    for (int d = 0; d < 3; ++d) {
      sp[3 * k + d] = p[d];
      R[3 * k + d] = p[d];
    }
  }
  __syncthreads();
  if (tid == 0) {
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (int k = 0; k < nn; ++k) {
      cx += sp[3 * k];
      cy += sp[3 * k + 1];
      cz += sp[3 * k + 2];
    }
    const double inv = 1.0 / nn;
    sc[0] = cx * inv;
    sc[1] = cy * inv;
    sc[2] = cz * inv;
  }
  __syncthreads();
  const double cx = sc[0], cy = sc[1], cz = sc[2];
  double rr = 0.0;
  for (int k = tid; k < nn; k += kMomThreads) {
    const double dx = sp[3 * k] - cx, dy = sp[3 * k + 1] - cy, dz = sp[3 * k + 2] - cz;
    rr = fmax(rr, dx * dx + dy * dy + dz * dz);
  }
  for (int off = 32; off > 0; off >>= 1) rr = fmax(rr, __shfl_xor(rr, off));
  if ((tid & 63) == 0) srad[tid >> 6] = rr;
  auto gl = [&](int i, int j) -> double {
    return (i >= 0 && i < nr && j >= 0 && j < nc) ? G[(long)(r0 + i) * t.cols + c0 + j] : 0.0;
  };
  const int s1 = (nr + 1) * nc, ns = s1 + (nc + 1) * nr;
  for (int s = tid; s < ns; s += kMomThreads) {
    int ia, ja, ib, jb;
    double ga, gb;
    if (s < s1) {
      ia = ib = s / nc;
      ja = s % nc;
      jb = ja + 1;
      ga = gl(ia, ja);
      gb = gl(ia - 1, ja);
    } else {
      const int u = s - s1;
      ja = jb = u / nr;
      ia = u % nr;
      ib = ia + 1;
      ga = gl(ia, ja - 1);
      gb = gl(ia, ja);
    }
    R[kRecStr + s] = ga * kInv4Pi - gb * kInv4Pi;
    reinterpret_cast<float*>(R + kRecG)[s] = (float)(ga * kInv4Pi - gb * kInv4Pi);
    const double* pa = sp + 3 * (ia * (nc + 1) + ja);
    const double* pb = sp + 3 * (ib * (nc + 1) + jb);
    sA[3 * s] = pa[0] - cx;
    sA[3 * s + 1] = pa[1] - cy;
    sA[3 * s + 2] = pa[2] - cz;
    sE[3 * s] = pb[0] - pa[0];
    sE[3 * s + 1] = pb[1] - pa[1];
    sE[3 * s + 2] = pb[2] - pa[2];
    {
      float* F = reinterpret_cast<float*>(R + kRecSeg) + 4 * s;
      const float ex = (float)sE[3 * s], ey = (float)sE[3 * s + 1], ez = (float)sE[3 * s + 2];
      F[0] = ex;
      F[1] = ey;
      F[2] = ez;
      F[3] = dev::kBand32 * fmaf(ez, ez, fmaf(ey, ey, ex * ex));  // the walk's rounding band
    }
    sg[s] = ga - gb;
  }
  __syncthreads();
  if (tid < kFFMom) {
    const MomIndex x = mom_index(tid);
    double acc = 0.0;
    for (int s = 0; s < ns; ++s) {
      const double g = sg[s];
      if (g == 0.0) continue;
      acc += segment_moment(x, g, sA + 3 * s, sE + 3 * s);
    }
    mom[(long)q * kFFMom + tid] = acc;
  }
  if (tid == 0) {
    FFCluster o;
    o.c[0] = cx;
    o.c[1] = cy;
    o.c[2] = cz;
    double r2 = srad[0];
    for (int w = 1; w < kMomThreads / 64; ++w) r2 = fmax(r2, srad[w]);
    o.r = sqrt(r2);
    o.n = n;
    o.r0 = r0;
    o.r1 = r1;
    o.c0 = c0;
    o.c1 = c1;
    o.level = 0;
    o.pad0 = o.pad1 = 0;
    cl[q] = o;
  }
}

// One workgroup per block of level l (at most 2 x 2 children): centre =
// ring-count weighted mean of the children's centres, radius bound from
// theirs; thread m translates component m of every child to the new centre
// (t = c_child - c):  Q1' = Q1 + Q0 t,  Q2' = Q2 + Q1 t + t Q1 + Q0 t t, ...
// and sums over the children in order.
__global__ __launch_bounds__(kMomThreads) void k_ff_parent(FFTree t, int l,
                                                           FFCluster* __restrict__ cl,
                                                           double* __restrict__ mom) {
  __shared__ int sid[4];
  __shared__ double stv[4][3];
  __shared__ int snch;
  const int q = blockIdx.x, tid = threadIdx.x;
  const int bj = q % t.nbc[l], bi = (q / t.nbc[l]) % t.nbr[l], n = q / (t.nbc[l] * t.nbr[l]);
  const int r0 = bi * t.hr[l], r1 = min(t.rows, r0 + t.hr[l]);
  const int c0 = bj * t.hc[l], c1 = min(t.cols, c0 + t.hc[l]);
  const int id = ff_id(t, l, n, bi, bj);
  if (tid == 0) {
    const int lc = l - 1;
    const int ci0 = r0 / t.hr[lc], ci1 = min(t.nbr[lc], cdiv(r1, t.hr[lc]));
    const int cj0 = c0 / t.hc[lc], cj1 = min(t.nbc[lc], cdiv(c1, t.hc[lc]));
    double cx = 0.0, cy = 0.0, cz = 0.0, w = 0.0;
    int nch = 0;
    for (int ci = ci0; ci < ci1; ++ci)
      for (int cj = cj0; cj < cj1; ++cj) {
        const int cid = ff_id(t, lc, n, ci, cj);
        const FFCluster& ch = cl[cid];
        const double wc = (double)(ch.r1 - ch.r0) * (ch.c1 - ch.c0);
        cx += wc * ch.c[0];
        cy += wc * ch.c[1];
        cz += wc * ch.c[2];
        w += wc;
        if (nch < 4) sid[nch++] = cid;
      }
    cx /= w;
    cy /= w;
    cz /= w;
    double rad = 0.0;
    for (int c = 0; c < nch; ++c) {
      const FFCluster& ch = cl[sid[c]];
      const double tv[3] = {ch.c[0] - cx, ch.c[1] - cy, ch.c[2] - cz};
      stv[c][0] = tv[0];
      stv[c][1] = tv[1];
      stv[c][2] = tv[2];
      rad = fmax(rad, sqrt(tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2]) + ch.r);
    }
    snch = nch;
    FFCluster o;
    o.c[0] = cx;
    o.c[1] = cy;
    o.c[2] = cz;
    o.r = rad;
    o.n = n;
    o.r0 = r0;
    o.r1 = r1;
    o.c0 = c0;
    o.c1 = c1;
    o.level = l;
    o.pad0 = o.pad1 = 0;
    cl[id] = o;
  }
  __syncthreads();
  if (tid >= kFFMom) return;
  const MomIndex x = mom_index(tid);
  const int i = x.i, j = x.j, k = x.k, m = x.l;
  double acc = 0.0;
  for (int c = 0; c < snch; ++c) {
    const double* q0 = mom + (long)sid[c] * kFFMom;
    const double* q1 = q0 + kFFQ1;
    const double* q2 = q0 + kFFQ2;
    const double* q3 = q0 + kFFQ3;
    const double* tv = stv[c];
    if (x.type == 0) {
      acc += q0[i];
    } else if (x.type == 1) {
      acc += q1[i * 3 + j] + q0[i] * tv[j];
    } else if (x.type == 2) {
      acc += q2[(i * 3 + j) * 3 + k] + q1[i * 3 + j] * tv[k] + q1[i * 3 + k] * tv[j] +
             q0[i] * tv[j] * tv[k];
    } else {
      acc += q3[((i * 3 + j) * 3 + k) * 3 + m] + q2[(i * 3 + j) * 3 + k] * tv[m] +
             q2[(i * 3 + j) * 3 + m] * tv[k] + q2[(i * 3 + k) * 3 + m] * tv[j] +
             q1[i * 3 + j] * tv[k] * tv[m] + q1[i * 3 + k] * tv[j] * tv[m] +
             q1[i * 3 + m] * tv[j] * tv[k] + q0[i] * tv[j] * tv[k] * tv[m];
    }
  }
  mom[(long)id * kFFMom + tid] = acc;
}

// --- targets -------------------------------------------------------------------
__device__ __forceinline__ int tile_target(const int* map, int tile, int lane) {
  return map[tile * kTileT + lane];
}

// One wavefront per tile: centre and radius of its targets.
__global__ __launch_bounds__(64) void k_ff_tiles(const double* __restrict__ x,
                                                 const int* __restrict__ map,
                                                 double4* __restrict__ tiles) {
  const int tile = blockIdx.x, lane = threadIdx.x;
  const int t = tile_target(map, tile, lane);
  const double px = t >= 0 ? x[3 * (long)t] : 0.0, py = t >= 0 ? x[3 * (long)t + 1] : 0.0,
               pz = t >= 0 ? x[3 * (long)t + 2] : 0.0;
  double sx = px, sy = py, sz = pz, cnt = t >= 0 ? 1.0 : 0.0;
  for (int off = 32; off > 0; off >>= 1) {
    sx += __shfl_xor(sx, off);
    sy += __shfl_xor(sy, off);
    sz += __shfl_xor(sz, off);
    cnt += __shfl_xor(cnt, off);
  }
  const double cx = sx / fmax(cnt, 1.0), cy = sy / fmax(cnt, 1.0), cz = sz / fmax(cnt, 1.0);
  double r2 = t >= 0 ? (px - cx) * (px - cx) + (py - cy) * (py - cy) + (pz - cz) * (pz - cz) : 0.0;
  for (int off = 32; off > 0; off >>= 1) r2 = fmax(r2, __shfl_xor(r2, off));
  if (lane == 0) tiles[tile] = double4{cx, cy, cz, sqrt(r2)};
}

// One wavefront per tile: breadth-first traversal of the block tree from the
// top level.  Each level's frontier is classified lane-parallel: far (the
// block's expansion serves the whole tile), near at level 0 (summed
// exactly), or opened (its children form the next frontier).  Lists are
// appended in frontier order (ballot prefix counts), so they are
// deterministic.  Capacities: near all level-0 blocks, far all blocks, each
// frontier the level-0 count (front: 2 of those per tile).
// Focus sphere (centre, radius) of a bounding box min[3], max[3].
__device__ inline double4 focus_sphere(const double* box) {
  const double hx = 0.5 * (box[3] - box[0]), hy = 0.5 * (box[4] - box[1]),
               hz = 0.5 * (box[5] - box[2]);
  return double4{box[0] + hx, box[1] + hy, box[2] + hz, sqrt(hx * hx + hy * hy + hz * hz)};
}

// A cluster is far when every target of the tile lies beyond theta times
// its radius: the tile's sphere decides when it clears the cluster or lies
// inside its reach, the targets one by one otherwise (per_target; without
// it, the sphere alone).
__global__ __launch_bounds__(64) void k_ff_lists(FFTree t, const FFCluster* __restrict__ cl,
                                                 const double4* __restrict__ tiles,
                                                 const double* __restrict__ x,
                                                 const int* __restrict__ map, bool per_target,
                                                 double theta,
                                                 const double* __restrict__ fbox, double reach,
                                                 double theta_min, int* __restrict__ near,
                                                 int* __restrict__ far, int2* __restrict__ cnt,
                                                 int* __restrict__ front) {
  const int tile = blockIdx.x, lane = threadIdx.x;
  const double4 tc = tiles[tile];
  __shared__ double tx[kTileT], ty[kTileT], tz[kTileT];
  double blo[3], bhi[3];  // the targets' box (a tighter test than the sphere)
  {  // the tile's targets (lane 0's stands in for the missing ones of the last tile)
    int tg = tile_target(map, tile, lane);
    if (tg < 0) tg = tile_target(map, tile, 0);
    const double p[3] = {x[3 * (long)tg], x[3 * (long)tg + 1], x[3 * (long)tg + 2]};
    tx[lane] = p[0];
    ty[lane] = p[1];
    tz[lane] = p[2];
    for (int d = 0; d < 3; ++d) {
      double lo = p[d], hi = p[d];
      for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off));
        hi = fmax(hi, __shfl_xor(hi, off));
      }
      blo[d] = lo;
      bhi[d] = hi;
    }
  }
  if (fbox) {  // graded criterion away from the focus
    const double4 f = focus_sphere(fbox);
    const double fx = tc.x - f.x, fy = tc.y - f.y, fz = tc.z - f.z;
    const double gap = sqrt(fx * fx + fy * fy + fz * fz) - tc.w - f.w, L = reach * f.w;
    if (gap > L) theta = fmax(theta_min, theta * (L / gap));
  }
  const int n0 = t.num * t.nbr[0] * t.nbc[0];
  int* my_near = near + (long)tile * n0;
  int* my_far = far + (long)tile * t.total;
  int* F = front + (long)tile * 2 * n0;
  int* Gn = F + n0;
  const int L = t.levels - 1;
  int nF = t.num * t.nbr[L] * t.nbc[L];
  for (int k = lane; k < nF; k += 64) F[k] = t.off[L] + k;
  const unsigned long long below = (1ull << lane) - 1ull;
  int nn = 0, nf = 0;
  __syncthreads();
  while (nF > 0) {
    int nG = 0;
    for (int base = 0; base < nF; base += 64) {
      const int k = base + lane;
      int id = -1, kind = 0, nch = 0, ci0 = 0, cj0 = 0, ncj = 1, lc = 0, cn = 0;
      if (k < nF) {
        id = F[k];
        const FFCluster c = cl[id];
        const double dx = c.c[0] - tc.x, dy = c.c[1] - tc.y, dz = c.c[2] - tc.z;
        const double dist = sqrt(dx * dx + dy * dy + dz * dz);
        const double reach_c = theta * c.r;
        bool is_far = dist - tc.w > reach_c;
        if (per_target && !is_far && dist + tc.w > reach_c) {
          const double lim2 = reach_c * reach_c;
          double b2 = 0.0;  // squared distance from the centre to the targets' box
          for (int d = 0; d < 3; ++d) {
            const double e = fmax(fmax(blo[d] - c.c[d], c.c[d] - bhi[d]), 0.0);
            b2 += e * e;
          }
          is_far = b2 > lim2;
          if (!is_far) {
            is_far = true;
            for (int q = 0; q < kTileT && is_far; ++q) {
              const double ex = c.c[0] - tx[q], ey = c.c[1] - ty[q], ez = c.c[2] - tz[q];
              is_far = ex * ex + ey * ey + ez * ez > lim2;
            }
          }
        }
        if (is_far) {
          kind = 1;
        } else if (c.level == 0) {
          kind = 2;
        } else {
          kind = 3;
          lc = c.level - 1;
          cn = c.n;
          ci0 = c.r0 / t.hr[lc];
          const int ci1 = min(t.nbr[lc], cdiv(c.r1, t.hr[lc]));
          cj0 = c.c0 / t.hc[lc];
          const int cj1 = min(t.nbc[lc], cdiv(c.c1, t.hc[lc]));
          ncj = cj1 - cj0;
          nch = (ci1 - ci0) * ncj;
        }
      }
      const unsigned long long mf = __ballot(kind == 1), mn = __ballot(kind == 2);
      if (kind == 1) my_far[nf + __popcll(mf & below)] = id;
      if (kind == 2) my_near[nn + __popcll(mn & below)] = id - t.off[0];
      nf += __popcll(mf);
      nn += __popcll(mn);
      int inc = nch;
      for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(inc, off);
        if (lane >= off) inc += v;
      }
      const int excl = inc - nch;
      for (int c = 0; c < nch; ++c)
        Gn[nG + excl + c] = ff_id(t, lc, cn, ci0 + c / ncj, cj0 + c % ncj);
      nG += __shfl(inc, 63);
    }
    __syncthreads();
    int* tmp = F;
    F = Gn;
    Gn = tmp;
    nF = nG;
  }
  if (lane == 0) cnt[tile] = int2{nn, nf};
}

// --- exact sums of blocks ---------------------------------------------------------
// One chain of a level-0 block's record at one target: count segments from
// node n0 in steps of stride nodes, strengths S[s0 ...]; 1/|x - p| carried
// along the chain (devmath.h); adds to u (strengths carry 1/4pi).  Ends as
// in dev::chain_step64.
template <bool Ends>
__device__ __forceinline__ void chain64(const double* R, int n0, int stride, int count,
                                        const double* S, int s0, double px, double py,
                                        double pz, double& ux, double& uy, double& uz) {
  const double* p = R + 3 * n0;
  dev::Carry64 cr;
  cr.rx = px - p[0];
  cr.ry = py - p[1];
  cr.rz = pz - p[2];
  cr.rs = __builtin_fma(cr.rz, cr.rz, __builtin_fma(cr.ry, cr.ry, cr.rx * cr.rx));
  cr.inv = dev::rsq64_1(cr.rs);
  double e[8];
  for (int k = 0; k < count; ++k) {
    const double* q = p + 3 * stride;
    e[0] = q[0];
    e[1] = q[1];
    e[2] = q[2];
    e[3] = S[s0 + k];
    e[4] = q[0] - p[0];
    e[5] = q[1] - p[1];
    e[6] = q[2] - p[2];
    e[7] = __builtin_fma(e[6], e[6], __builtin_fma(e[5], e[5], e[4] * e[4]));
    dev::chain_step64<Ends>(e, px, py, pz, cr, ux, uy, uz);
    p = q;
  }
}

// All chains of a record (nr x nc rings): node rows (spanwise segments),
// then node columns (chordwise).
template <bool Ends>
__device__ __forceinline__ void walk_record(const double* R, const double* S, int nr, int nc,
                                            double px, double py, double pz, double& ux,
                                            double& uy, double& uz) {
  for (int i = 0; i <= nr; ++i)
    chain64<Ends>(R, i * (nc + 1), 1, nc, S, i * nc, px, py, pz, ux, uy, uz);
  const int s1 = (nr + 1) * nc;
  for (int j = 0; j <= nc; ++j)
    chain64<Ends>(R, j, nc + 1, nr, S, s1 + j * nr, px, py, pz, ux, uy, uz);
}

// Every segment of a block lies within its radius r of the centre, so it is
// shorter than 2 r: below this radius the end-point cut-offs are implied.
constexpr double kShortRadius = 1.5e4;  // (2 r)^2 < dev::kShortSeg2

// The same walk in FP32 with compensation: x - p in FP64 (exact to FP32
// rounding relative to |x - p|), the segment law in FP32 with the hardware
// rsqrt / reciprocal, each chain (<= 17 elements) summed in FP32 and added
// to the FP64 sums; pairs within a band around the reference's cut-off
// (|r| < 1e-9, |d|^2 < 2e-9 plus the FP32 rounding bound of d) are
// re-evaluated in FP64 exactly as the reference does.  Two elements per
// trip, the end point alternating between two register sets.
template <bool Ends>
__device__ __forceinline__ void chain32(const double* R, int n0, int stride_nodes, int count,
                                        const double* __restrict__ S, const float* F,
                                        const float* G, int s0, double x, double y, double z,
                                        double& ux, double& uy, double& uz) {
  const double* p = R + 3 * n0;
  const int step = 3 * stride_nodes;
  dev::End32 A, B;
  dev::end32(p, x, y, z, A);
  float fx = 0.f, fy = 0.f, fz = 0.f;
  // lanes with an element of this chain to re-evaluate in FP64 (a wave mask:
  // scalar bookkeeping)
  unsigned long long redo = 0ull;
  int k = 0;
  for (; k + 2 <= count; k += 2) {
    dev::end32(p + step, x, y, z, B);
    redo |= __ballot(dev::elem32<Ends>(A, B, F + 4 * (s0 + k), G[s0 + k], fx, fy, fz));
    p += 2 * step;
    dev::end32(p, x, y, z, A);
    redo |= __ballot(dev::elem32<Ends>(B, A, F + 4 * (s0 + k + 1), G[s0 + k + 1], fx, fy, fz));
  }
  if (k < count) {
    dev::end32(p + step, x, y, z, B);
    redo |= __ballot(dev::elem32<Ends>(A, B, F + 4 * (s0 + k), G[s0 + k], fx, fy, fz));
  }
  if (redo) {  // rare: a second pass finds the flagged elements, in FP64
    const bool mine = (redo >> __lane_id()) & 1ull;
    p = R + 3 * n0;
    dev::end32(p, x, y, z, A);
    for (int e = 0; e < count; ++e) {
      const double* b = p + step;
      dev::end32(b, x, y, z, B);
      if (mine && dev::in32<Ends>(A, B, F + 4 * (s0 + e)))
        dev::seg_ref(p[0], p[1], p[2], b[0], b[1], b[2], x, y, z, S[s0 + e], ux, uy, uz);
      A = B;
      p = b;
    }
  }
  ux += (double)fx;
  uy += (double)fy;
  uz += (double)fz;
}

// R: the record as the FP32 walk stages it (nodes, segment vectors, FP32
// strengths); S: the FP64 strengths (in the global record, read only by the
// rare FP64 re-evaluation).
template <bool Ends>
__device__ __forceinline__ void walk_record32(const double* R, const double* S, int nr, int nc,
                                              double px, double py, double pz, double& ux,
                                              double& uy, double& uz) {
  const float* F = reinterpret_cast<const float*>(R + kRecStr);
  const float* G = reinterpret_cast<const float*>(R + kRecStr + 2 * kL0Segs);
  for (int i = 0; i <= nr; ++i)  // node rows: spanwise segments
    chain32<Ends>(R, i * (nc + 1), 1, nc, S, F, G, i * nc, px, py, pz, ux, uy, uz);
  const int s1 = (nr + 1) * nc;
  for (int j = 0; j <= nc; ++j)  // node columns: chordwise segments
    chain32<Ends>(R, j, nc + 1, nr, S, F, G, s1 + j * nr, px, py, pz, ux, uy, uz);
}

// Exact velocity (with 1/4pi) at one target of the level-0 blocks under
// cluster c, blocks in row order.
__device__ void exact_cluster(const FFTree& t, const FFCluster& c, const double* __restrict__ rec,
                              double px, double py, double pz, double u[3]) {
  double ux = 0.0, uy = 0.0, uz = 0.0;
  const int bi0 = c.r0 / kFFRows, bi1 = cdiv(c.r1, kFFRows);
  const int bj0 = c.c0 / kFFCols, bj1 = cdiv(c.c1, kFFCols);
  for (int bi = bi0; bi < bi1; ++bi)
    for (int bj = bj0; bj < bj1; ++bj) {
      const int nr = min(t.rows, (bi + 1) * kFFRows) - bi * kFFRows;
      const int nc = min(t.cols, (bj + 1) * kFFCols) - bj * kFFCols;
      const double* R = rec + (long)ff_id(t, 0, c.n, bi, bj) * kRec;
      walk_record<true>(R, R + kRecStr, nr, nc, px, py, pz, ux, uy, uz);
    }
  u[0] = ux;
  u[1] = uy;
  u[2] = uz;
}

// Near-count keys of the tiles: the near kernel takes tiles in decreasing
// order of their near lists (heaviest first), so that the longest
// workgroups do not start last and run alone at the end of the launch.
// One workgroup: a counting sort of the tiles by near count / 8 (1024
// buckets, the heaviest first; within a bucket in the order the atomics
// hand out, which changes no result: every (tile, group) workgroup writes
// its own partial sums).
constexpr int kOrderBuckets = 1024;
__global__ __launch_bounds__(1024) void k_order_heavy(const int2* __restrict__ cnt, int ntiles,
                                                      int* __restrict__ order) {
  __shared__ int hist[kOrderBuckets];
  const int t = threadIdx.x;
  hist[t] = 0;
  __syncthreads();
  auto bucket = [&](int q) {  // 0 = heaviest
    return kOrderBuckets - 1 - min(cnt[q].x >> 3, kOrderBuckets - 1);
  };
  for (int q = t; q < ntiles; q += 1024) atomicAdd(&hist[bucket(q)], 1);
  __syncthreads();
  // exclusive prefix sum over the buckets (Hillis-Steele on the inclusive sum)
  const int mine = hist[t];
  int v = mine;
  for (int off = 1; off < kOrderBuckets; off <<= 1) {
    __syncthreads();
    const int add = t >= off ? hist[t - off] : 0;
    __syncthreads();
    v += add;
    hist[t] = v;
  }
  __syncthreads();
  hist[t] = v - mine;
  __syncthreads();
  for (int q = t; q < ntiles; q += 1024) order[atomicAdd(&hist[bucket(q)], 1)] = q;
}

// Group g of a tile sums the tile's near blocks g, g + kNearGroups, ...: each
// block's record is staged in LDS (the next one is loaded into registers
// meanwhile) and walked for every lane's target (FP64, or FP32 compensated).
template <bool F32>
__global__ __launch_bounds__(64, 5) void k_ff_near(const FFCluster* __restrict__ cl,
                                                const double* __restrict__ rec,
                                                const double* __restrict__ x, int nt,
                                                const int* __restrict__ map, int n0,
                                                const int* __restrict__ near,
                                                const int2* __restrict__ cnt,
                                                const int* __restrict__ order,
                                                double* __restrict__ part,
                                                unsigned long long* __restrict__ npairs) {
  constexpr int kLen = F32 ? kRec32 : kRecSeg;  // doubles staged (see kRec32)
  constexpr int kPer = (kLen + 63) / 64;
  __shared__ double sr[kPer * 64];
  const int tile = order[blockIdx.x], g = blockIdx.y, lane = threadIdx.x;
  const int tg = tile_target(map, tile, lane);
  const double px = tg >= 0 ? x[3 * (long)tg] : 0.0, py = tg >= 0 ? x[3 * (long)tg + 1] : 0.0,
               pz = tg >= 0 ? x[3 * (long)tg + 2] : 0.0;
  const int count = cnt[tile].x;
  const int* list = near + (long)tile * n0;
  double ux = 0.0, uy = 0.0, uz = 0.0;
  double pre[kNearPrefetch ? kPer : 1];
  int id = g < count ? list[g] : 0;
  auto load = [&](int b) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int k = lane + 64 * u;
      const int src = F32 && k >= kRecStr ? k + kL0Segs : k;  // the FP32 walk skips S
      pre[kNearPrefetch ? u : 0] = k < kLen ? rec[(long)b * kRec + src] : 0.0;
      if (!kNearPrefetch) sr[lane + 64 * u] = pre[0];
    }
  };
  if (kNearPrefetch && g < count) load(id);
  unsigned segs = 0;  // segments walked by this group (for the pair count)
  for (int k = g; k < count; k += kNearGroups) {
    const FFCluster c = cl[id];
    const double* S = rec + (long)id * kRec + kRecStr;  // FP64 strengths, global
    const int nr = c.r1 - c.r0, nc = c.c1 - c.c0;
    segs += (unsigned)((nr + 1) * nc + (nc + 1) * nr);
    __syncthreads();
    if (kNearPrefetch) {
#pragma unroll
      for (int u = 0; u < kPer; ++u) sr[lane + 64 * u] = pre[u];
    } else {
      load(id);
    }
    __syncthreads();
    if (k + kNearGroups < count) {  // next block's record, in flight during the walk
      id = list[k + kNearGroups];
      if (kNearPrefetch) load(id);
    }
    // FP32 for blocks of short segments, whose end-point cut-offs the band
    // implies; the FP64 walk with the end-point tests otherwise (a block
    // 3e4 across: not met in practice, and one FP32 variant keeps the
    // kernel at 5 waves per SIMD)
    if (F32 && c.r < kShortRadius)
      walk_record32<false>(sr, S, nr, nc, px, py, pz, ux, uy, uz);
    else if (F32)
      walk_record<true>(sr, S, nr, nc, px, py, pz, ux, uy, uz);
    else if (c.r < kShortRadius)
      walk_record<false>(sr, sr + kRecStr, nr, nc, px, py, pz, ux, uy, uz);
    else
      walk_record<true>(sr, sr + kRecStr, nr, nc, px, py, pz, ux, uy, uz);
  }
  if (tg >= 0) {
    part[((long)g * 3 + 0) * nt + tg] = ux;
    part[((long)g * 3 + 1) * nt + tg] = uy;
    part[((long)g * 3 + 2) * nt + tg] = uz;
  }
  if (npairs) {  // (target, segment) pairs of the tile's valid targets
    const unsigned long long nv = (unsigned long long)__popcll(__ballot(tg >= 0));
    if (lane == 0 && segs) atomicAdd(npairs, nv * segs);
  }
}

// --- far clusters: moment expansion ------------------------------------------------
// Traces of the symmetric position indices: tr[c] = Q2[c,j,j] (c < 3),
// tr[3 + 3c + b] = Q3[c,b,k,k].
__device__ inline void moment_traces(const double* Q, double tr[12]) {
  const double* Q2 = Q + kFFQ2;
  const double* Q3 = Q + kFFQ3;
  for (int c = 0; c < 3; ++c) {
    tr[c] = Q2[(c * 3 + 0) * 3 + 0] + Q2[(c * 3 + 1) * 3 + 1] + Q2[(c * 3 + 2) * 3 + 2];
    for (int b = 0; b < 3; ++b) {
      const double* q = Q3 + ((c * 3 + b) * 3) * 3;
      tr[3 + 3 * c + b] = q[0] + q[4] + q[8];
    }
  }
}

// Traces of every cluster's moments (12 per cluster, one thread each),
// once per build: the far kernel reads them with the moments.
__global__ void k_ff_traces(const double* __restrict__ mom, int total, double* __restrict__ tr) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= total) return;
  double t[12];
  moment_traces(mom + (long)id * kFFMom, t);
  for (int k = 0; k < 12; ++k) tr[(long)id * 12 + k] = t[k];
}

// u (without 1/4pi) of the moments Q at d = x - c, orders 0..3.  With the
// position indices of Q2, Q3 symmetric, the derivative tensors of 1/r
// contract to
//   S[c][b] = d_b alpha[c] + beta[c][b],  u_a = eps_abc S[c][b],
//   alpha = -Q0/r^3 - 3 Q1d/r^5 - 15/2 Q2dd/r^7 + 3/2 tr2/r^5
//           - 35/2 Q3ddd/r^9 + 15/2 tr3 d/r^7,
//   beta  = Q1/r^3 + 3 Q2d/r^5 + 15/2 Q3dd/r^7 - 3/2 tr3/r^5
// (Qd = Q contracted with d over its last position index).
__device__ void expand_moments(const double* Q, const double* tr, double dx, double dy, double dz,
                               double u[3]) {
  // FMA-contracted by hand (this file is built without contraction, which
  // the reference-order kernels need): half the VALU instructions of the
  // separate products and sums; the expansion's truncation, ~(r/|d|)^4,
  // dwarfs the difference in rounding.
  const double d[3] = {dx, dy, dz};
  const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
  const double ir = dev::rsq64_1(r2);
  const double i2 = ir * ir, i3 = i2 * ir, i5 = i3 * i2, i7 = i5 * i2, i9 = i7 * i2;
  const double* Q1 = Q + kFFQ1;
  const double* Q2 = Q + kFFQ2;
  const double* Q3 = Q + kFFQ3;
  double al[3], be[3][3];
  for (int c = 0; c < 3; ++c) {
    double q1d = 0.0, q2dd = 0.0, q3ddd = 0.0, trd = 0.0;
    for (int b = 0; b < 3; ++b) {
      double q2d = 0.0, q3dd = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double* q3 = Q3 + ((c * 3 + b) * 3 + k) * 3;
        const double q3d = __builtin_fma(q3[2], d[2], __builtin_fma(q3[1], d[1], q3[0] * d[0]));
        q2d = __builtin_fma(Q2[(c * 3 + b) * 3 + k], d[k], q2d);
        q3dd = __builtin_fma(q3d, d[k], q3dd);
      }
      const double t = tr[3 + 3 * c + b];
      q1d = __builtin_fma(Q1[c * 3 + b], d[b], q1d);
      q2dd = __builtin_fma(q2d, d[b], q2dd);
      q3ddd = __builtin_fma(q3dd, d[b], q3ddd);
      trd = __builtin_fma(t, d[b], trd);
      // Q1 i3 + 3 q2d i5 + 7.5 q3dd i7 - 1.5 t i5
      be[c][b] = __builtin_fma(Q1[c * 3 + b], i3,
                               __builtin_fma(__builtin_fma(3.0, q2d, -1.5 * t), i5, 7.5 * q3dd * i7));
    }
    // -Q0 i3 - 3 q1d i5 - 7.5 q2dd i7 + 1.5 tr i5 - 17.5 q3ddd i9 + 7.5 trd i7
    al[c] = __builtin_fma(-Q[c], i3,
                          __builtin_fma(__builtin_fma(-3.0, q1d, 1.5 * tr[c]), i5,
                                        __builtin_fma(__builtin_fma(-7.5, q2dd, 7.5 * trd), i7,
                                                      -17.5 * q3ddd * i9)));
  }
  u[0] = __builtin_fma(d[1], al[2], -d[2] * al[1]) + (be[2][1] - be[1][2]);
  u[1] = __builtin_fma(d[2], al[0], -d[0] * al[2]) + (be[0][2] - be[2][0]);
  u[2] = __builtin_fma(d[0], al[1], -d[1] * al[0]) + (be[1][0] - be[0][1]);
}

// Network features of (target, cluster) from its moments Q, offset d = x - c
// and radius r; returns the correction per unit output, s rho^4 / (4pi D^2)
// (0 when the cluster carries no circulation),
// s = |Q0| + |Q1| / r + |Q2| / r^2 + |Q3| / r^3 (every feature is at most 1).
__device__ double ff_features(const double* Q, double dx, double dy, double dz, double r,
                              float f[kMlpIn]) {
  const double D = sqrt(dx * dx + dy * dy + dz * dz);
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int m = 0; m < kFFMom; ++m) {
    const int t = m < kFFQ1 ? 0 : m < kFFQ2 ? 1 : m < kFFQ3 ? 2 : 3;
    a[t] += Q[m] * Q[m];
  }
  const double s = sqrt(a[0]) + sqrt(a[1]) / r + sqrt(a[2]) / (r * r) + sqrt(a[3]) / (r * r * r);
  if (!(s > 0.0) || !(D > 0.0) || !(r > 0.0)) {
    for (int k = 0; k < kMlpIn; ++k) f[k] = 0.0f;
    return 0.0;
  }
  const double dh[3] = {dx / D, dy / D, dz / D};
  const double rho = r / D;
  const double* Q1 = Q + kFFQ1;
  const double* Q2 = Q + kFFQ2;
  const double* Q3 = Q + kFFQ3;
  f[0] = (float)dh[0];
  f[1] = (float)dh[1];
  f[2] = (float)dh[2];
  f[3] = (float)rho;
  for (int i = 0; i < 3; ++i) f[4 + i] = (float)(Q[i] / s);
  for (int m = 0; m < 9; ++m) f[7 + m] = (float)(Q1[m] / (s * r));
  const double s2 = s * r * r, s3 = s2 * r;
  for (int i = 0; i < 3; ++i) {
    double b2 = 0.0, b3 = 0.0;
    for (int j = 0; j < 3; ++j) {
      const double* q2 = Q2 + (i * 3 + j) * 3;
      const double v = q2[0] * dh[0] + q2[1] * dh[1] + q2[2] * dh[2];
      f[16 + i * 3 + j] = (float)(v / s2);
      b2 += v * dh[j];
      double w = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double* q3 = Q3 + ((i * 3 + j) * 3 + k) * 3;
        w += (q3[0] * dh[0] + q3[1] * dh[1] + q3[2] * dh[2]) * dh[k];
      }
      b3 += w * dh[j];
    }
    f[25 + i] = (float)(b2 / s2);
    f[28 + i] = (float)(b3 / s3);
  }
  f[31] = 1.0f;
  return s * rho * rho * rho * rho * kInv4Pi / (D * D);
}

// Taylor polynomial of a correction record at delta = (x - c_T) /
// (kMlpTaylorUnit D) (record: the 30 outputs times the scale, then
// 1 / (kMlpTaylorUnit D), status), FP32.
__device__ __forceinline__ void tile_correction(const float* __restrict__ rec, float dx, float dy,
                                                float dz, float v[3]) {
  const float q[6] = {0.5f * dx * dx, dx * dy, dx * dz, 0.5f * dy * dy, dy * dz, 0.5f * dz * dz};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float a = rec[i];
    a = fmaf(rec[3 + 3 * i], dx, a);
    a = fmaf(rec[4 + 3 * i], dy, a);
    a = fmaf(rec[5 + 3 * i], dz, a);
#pragma unroll
    for (int m = 0; m < 6; ++m) a = fmaf(rec[12 + 6 * i + m], q[m], a);
    v[i] = a;
  }
}

// Whether the network corrects the pair (tile centre, cluster): a cluster
// up to kMlpMaxLevel, c_T outside its sphere and the tile small against
// the distance D.  Returns D (0 when not corrected).
__device__ __forceinline__ double tile_pair_distance(const double4& tc, const FFCluster& c,
                                                     double* dx, double* dy, double* dz) {
  *dx = tc.x - c.c[0];
  *dy = tc.y - c.c[1];
  *dz = tc.z - c.c[2];
  const double D = sqrt(*dx * *dx + *dy * *dy + *dz * *dz);
  const bool ok = c.level <= kMlpMaxLevel && D > c.r && tc.w <= kMlpTileRatio * D;
  return ok ? D : 0.0;
}

// Pairs the network corrects, gathered: one thread per (tile, far-list
// entry e < kNetCap) decides (tile_pair_distance, circulation); an eligible
// pair takes a row of the compact batch (atomic counter: the row order
// varies from run to run, the records do not) with its features at the
// tile centre, its scale s rho^4 / (4pi D^2) and 1 / (kMlpTaylorUnit D);
// every other entry's correction record is marked 0 (none).
struct NetRow {
  int tile, e;
  float scale, inv;
};

__global__ __launch_bounds__(64) void k_ff_netrows(FFTree t, const FFCluster* __restrict__ cl,
                                                   const double* __restrict__ mom,
                                                   const double4* __restrict__ tiles,
                                                   const int* __restrict__ far,
                                                   const int2* __restrict__ cnt,
                                                   int* __restrict__ nrows,
                                                   NetRow* __restrict__ rows,
                                                   float* __restrict__ feats,
                                                   float* __restrict__ corr) {
  const int tile = blockIdx.x, e = blockIdx.y * 64 + threadIdx.x;
  const int count = min(cnt[tile].y, kNetCap);
  if (e >= count) return;
  const int id = far[(long)tile * t.total + e];
  const FFCluster c = cl[id];
  const double4 tc = tiles[tile];
  double dx, dy, dz;
  const double D = tile_pair_distance(tc, c, &dx, &dy, &dz);
  float f[kMlpIn];
  const double scale = D > 0.0 ? ff_features(mom + (long)id * kFFMom, dx, dy, dz, c.r, f) : 0.0;
  if (!(scale > 0.0)) {
    corr[((long)tile * kNetCap + e) * 32 + 31] = 0.0f;
    return;
  }
  const int r = atomicAdd(nrows, 1);
  rows[r] = NetRow{tile, e, (float)scale, (float)(1.0 / (kMlpTaylorUnit * D))};
  for (int k = 0; k < kMlpIn; ++k) feats[(long)r * kMlpIn + k] = f[k];
}

// The network on the gathered rows, one wavefront per batch of 64 rows
// (batches b = blockIdx.x, blockIdx.x + gridDim.x, ... below *nrows), on
// the matrix cores; per row a correction record of 32 floats at its (tile,
// entry): the outputs times the scale, 1 / (kMlpTaylorUnit D), and the
// status (1 corrected, -1 an output beyond ymax: the far kernel sums the
// cluster exactly).
__global__ __launch_bounds__(64) void k_ff_tilenet(const int* __restrict__ nrows,
                                                   const NetRow* __restrict__ rows,
                                                   const float* __restrict__ feats, MlpDev net,
                                                   float ymax, float* __restrict__ corr) {
  __shared__ unsigned short lds[2 * kMlpLdsHalfs];
  float* ylds = reinterpret_cast<float*>(lds);  // outputs over the first buffer
  const int lane = threadIdx.x;
  const int n = *nrows;
  for (int b = blockIdx.x; 64 * b < n; b += gridDim.x) {
    const int r = 64 * b + lane;
    float f[kMlpIn];
    for (int k = 0; k < kMlpIn; ++k) f[k] = r < n ? feats[(long)r * kMlpIn + k] : 0.0f;
    mlp_wave_bf16(net, f, lds, ylds);
    if (r < n) {
      const NetRow q = rows[r];
      const float* y = ylds + lane * kMlpYStride;
      float ym = 0.0f;
      for (int o = 0; o < kMlpOut; ++o) ym = fmaxf(ym, fabsf(y[o]));
      const bool ok = ym <= ymax;
      float* rec = corr + ((long)q.tile * kNetCap + q.e) * 32;
      for (int o = 0; o < kMlpOut; ++o) rec[o] = ok ? y[o] * q.scale : 0.0f;
      rec[30] = q.inv;
      rec[31] = ok ? 1.0f : -1.0f;
    }
    __syncthreads();  // the outputs' buffer takes the next batch's features
  }
}

__global__ void k_zero_int(int* p) { *p = 0; }

// Far clusters of a tile, group g: moment expansion; with the network
// (MLP = true) the first kNetCap entries add their correction records
// (k_ff_tilenet), and an entry whose network output left the trained range
// is summed exactly.  The group's clusters k = g, g + kGroups, ... are
// staged kFarBatch at a time (moments, traces, centre) between two
// barriers and expanded in order.
constexpr int kFarBatch = 4;
template <bool MLP>
__global__ __launch_bounds__(64) void k_ff_far(FFTree t, const FFCluster* __restrict__ cl,
                                               const double* __restrict__ mom,
                                               const double* __restrict__ trs,
                                               const double* __restrict__ rec,
                                               const double* __restrict__ x, int nt,
                                               const int* __restrict__ map,
                                               const int* __restrict__ far,
                                               const int2* __restrict__ cnt,
                                               const double4* __restrict__ tiles,
                                               const float* __restrict__ corr,
                                               unsigned long long* __restrict__ stats,
                                               double* __restrict__ part) {
  constexpr int kStage = kFFMom + 12 + 4;  // moments, traces, centre and radius
  __shared__ double sq[kFarBatch][kStage];
  __shared__ int sid[kFarBatch];
  __shared__ float srec[MLP ? kFarBatch : 1][32];  // correction records
  const int tile = blockIdx.x, g = blockIdx.y, lane = threadIdx.x;
  const int tg = tile_target(map, tile, lane);
  const double px = tg >= 0 ? x[3 * (long)tg] : 0.0, py = tg >= 0 ? x[3 * (long)tg + 1] : 0.0,
               pz = tg >= 0 ? x[3 * (long)tg + 2] : 0.0;
  float ox = 0.0f, oy = 0.0f, oz = 0.0f;  // x - c_T
  if (MLP) {
    const double4 tc = tiles[tile];
    ox = (float)(px - tc.x);
    oy = (float)(py - tc.y);
    oz = (float)(pz - tc.z);
  }
  const int count = cnt[tile].y;
  const int* list = far + (long)tile * t.total;
  double ux = 0.0, uy = 0.0, uz = 0.0;
  unsigned npairs = 0, ntrips = 0;
  for (int k0 = g; k0 < count; k0 += kFarBatch * kGroups) {
    const int nb = min(kFarBatch, (count - k0 + kGroups - 1) / kGroups);
    __syncthreads();
    for (int q = lane; q < nb * kStage; q += 64) {
      const int j = q / kStage, m = q % kStage;
      const int id = list[k0 + j * kGroups];
      double v;
      if (m < kFFMom)
        v = mom[(long)id * kFFMom + m];
      else if (m < kFFMom + 12)
        v = trs[(long)id * 12 + m - kFFMom];
      else if (m < kFFMom + 15)
        v = cl[id].c[m - kFFMom - 12];
      else
        v = cl[id].r;
      sq[j][m] = v;
      if (m == 0) sid[j] = id;
    }
    if (MLP)
      for (int q = lane; q < nb * 32; q += 64) {
        const int j = q / 32, m = q % 32, k = k0 + j * kGroups;
        srec[j][m] = k < kNetCap ? corr[((long)tile * kNetCap + k) * 32 + m] : 0.0f;
      }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
      const int k = k0 + j * kGroups;
      const double* Q = sq[j];
      const double dx = px - Q[kFFMom + 12], dy = py - Q[kFFMom + 13], dz = pz - Q[kFFMom + 14];
      double u[3];
      expand_moments(Q, Q + kFFMom, dx, dy, dz, u);
      u[0] *= kInv4Pi;
      u[1] *= kInv4Pi;
      u[2] *= kInv4Pi;
      if (MLP) {
        const float* r = srec[MLP ? j : 0];
        const float status = r[31];
        if (status != 0.0f && tg >= 0) {
          ++npairs;
          if (status > 0.0f) {
            const float inv = r[30];
            float v[3];
            tile_correction(r, ox * inv, oy * inv, oz * inv, v);
            u[0] += (double)v[0];
            u[1] += (double)v[1];
            u[2] += (double)v[2];
          } else {  // outside the trained range: the exact sum of the cluster
            ++ntrips;
            exact_cluster(t, cl[sid[j]], rec, px, py, pz, u);
          }
        }
      }
      ux += u[0];
      uy += u[1];
      uz += u[2];
    }
  }
  if (tg >= 0) {
    part[((long)(kNearGroups + g) * 3 + 0) * nt + tg] = ux;
    part[((long)(kNearGroups + g) * 3 + 1) * nt + tg] = uy;
    part[((long)(kNearGroups + g) * 3 + 2) * nt + tg] = uz;
  }
  if (MLP) {
    for (int off = 32; off > 0; off >>= 1) {
      npairs += __shfl_xor(npairs, off);
      ntrips += __shfl_xor(ntrips, off);
    }
    if (lane == 0 && npairs) {
      atomicAdd(stats, (unsigned long long)npairs);
      if (ntrips) atomicAdd(stats + 1, (unsigned long long)ntrips);
    }
  }
}

// Remainder of a cluster's expansion at x, exact - expansion, over scale.
__device__ void remainder_at(const FFTree& t, const FFCluster& c, const double* Q,
                             const double* tr, const double* __restrict__ rec, double x, double y,
                             double z, double scale, double out[3]) {
  double ue[3], ux[3];
  expand_moments(Q, tr, x - c.c[0], y - c.c[1], z - c.c[2], ue);
  exact_cluster(t, c, rec, x, y, z, ux);
  for (int d = 0; d < 3; ++d) out[d] = (ux[d] - ue[d] * kInv4Pi) / scale;
}

// Training samples: sample q = (tile, far-list entry); row q of feats (at
// the tile centre c_T) and of labels: the remainder's Taylor coefficients
// about c_T by central differences of step h = kMlpFdStep D over the 19
// points 0, +-e_j, +-e_j +-e_k (offsets in units of D):
//   value R(0); gradient (R(e_j) - R(-e_j)) / 2h;
//   Hessian jj (R(e_j) - 2 R(0) + R(-e_j)) / h^2,
//   jk (R(e_j + e_k) - R(e_j - e_k) - R(-e_j + e_k) + R(-e_j - e_k)) / 4h^2,
// gradient and Hessian then in the polynomial's unit kMlpTaylorUnit D
// (times kMlpTaylorUnit and its square).
// valid[q] = 0 for an entry the network does not correct (tile_pair_distance)
// or a cluster without circulation.
__global__ void k_ff_samples(FFTree t, const FFCluster* __restrict__ cl,
                             const double* __restrict__ mom, const double* __restrict__ rec,
                             const double4* __restrict__ tiles, const int* __restrict__ far,
                             const int4* __restrict__ pick, int n, float* __restrict__ feats,
                             float* __restrict__ labels, int* __restrict__ valid) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int4 pk = pick[q];
  const int id = far[(long)pk.x * t.total + pk.y];
  const FFCluster c = cl[id];
  const double4 tc = tiles[pk.x];
  valid[q] = 0;
  double dx, dy, dz;
  const double D = tile_pair_distance(tc, c, &dx, &dy, &dz);
  if (!(D > 0.0)) return;
  const double* Q = mom + (long)id * kFFMom;
  float f[kMlpIn];
  const double scale = ff_features(Q, dx, dy, dz, c.r, f);
  if (!(scale > 0.0)) return;
  double tr[12];
  moment_traces(Q, tr);
  const double hD = kMlpFdStep * D, h = kMlpFdStep;
  const double u1 = kMlpTaylorUnit, u2 = kMlpTaylorUnit * kMlpTaylorUnit;
  double R0[3];
  remainder_at(t, c, Q, tr, rec, tc.x, tc.y, tc.z, scale, R0);
  float* L = labels + (long)q * kMlpOut;
  for (int i = 0; i < 3; ++i) L[i] = (float)R0[i];
  double Rp[3][3], Rm[3][3];  // R(+h e_j), R(-h e_j)
  for (int j = 0; j < 3; ++j) {
    const double ej[3] = {j == 0 ? hD : 0.0, j == 1 ? hD : 0.0, j == 2 ? hD : 0.0};
    remainder_at(t, c, Q, tr, rec, tc.x + ej[0], tc.y + ej[1], tc.z + ej[2], scale, Rp[j]);
    remainder_at(t, c, Q, tr, rec, tc.x - ej[0], tc.y - ej[1], tc.z - ej[2], scale, Rm[j]);
    for (int i = 0; i < 3; ++i) L[3 + 3 * i + j] = (float)(u1 * (Rp[j][i] - Rm[j][i]) / (2.0 * h));
  }
  const int pj[6] = {0, 0, 0, 1, 1, 2}, pk2[6] = {0, 1, 2, 1, 2, 2};
  for (int m = 0; m < 6; ++m) {
    const int j = pj[m], k = pk2[m];
    if (j == k) {
      for (int i = 0; i < 3; ++i)
        L[12 + 6 * i + m] = (float)(u2 * (Rp[j][i] - 2.0 * R0[i] + Rm[j][i]) / (h * h));
    } else {
      double o[3] = {0.0, 0.0, 0.0}, a[3], b[3], cc[3], d[3];
      o[j] = hD;
      o[k] = hD;
      remainder_at(t, c, Q, tr, rec, tc.x + o[0], tc.y + o[1], tc.z + o[2], scale, a);
      remainder_at(t, c, Q, tr, rec, tc.x - o[0], tc.y - o[1], tc.z - o[2], scale, d);
      o[k] = -hD;
      remainder_at(t, c, Q, tr, rec, tc.x + o[0], tc.y + o[1], tc.z + o[2], scale, b);
      remainder_at(t, c, Q, tr, rec, tc.x - o[0], tc.y - o[1], tc.z - o[2], scale, cc);
      // a = R(e_j + e_k), b = R(e_j - e_k), cc = R(-e_j + e_k), d = R(-e_j - e_k)
      for (int i = 0; i < 3; ++i)
        L[12 + 6 * i + m] = (float)(u2 * (a[i] - b[i] - cc[i] + d[i]) / (4.0 * h * h));
    }
  }
  for (int k = 0; k < kMlpIn; ++k) feats[(long)q * kMlpIn + k] = f[k];
  valid[q] = 1;
}

// Bound-sheet velocity at the wake's nodes (sheets of per_sheet points,
// rows of width points): the first near_rows rows of each sheet take the
// exact chain sums of a compact copy ([num][near_rows][width], cb SoA
// chunks of near_ld, combined in chunk order); the others the moment
// expansion of each source sheet's root cluster where their distance from
// its centre exceeds theta times its radius, else that sheet's exact sum.
// out: SoA [3][nt].
__global__ void k_bound_far(FFTree t, const FFCluster* __restrict__ cl,
                            const double* __restrict__ mom, const double* __restrict__ trs,
                            const double* __restrict__ rec, const double* __restrict__ x, int nt,
                            int per_sheet, int width, int near_rows, double theta,
                            const double* __restrict__ near_part, int near_ld, int cb,
                            double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nt) return;
  const int n = p / per_sheet, q = p % per_sheet;
  double u[3] = {0.0, 0.0, 0.0};
  if (q / width < near_rows) {
    const long k = (long)n * near_rows * width + q;
    for (int d = 0; d < 3; ++d) {
      double b = near_part[(long)d * near_ld + k];
      for (int c = 1; c < cb; ++c) b = b + near_part[((long)c * 3 + d) * near_ld + k];
      u[d] = b;
    }
  } else {
    const double px = x[3 * (long)p], py = x[3 * (long)p + 1], pz = x[3 * (long)p + 2];
    const int L = t.levels - 1;
    for (int w = 0; w < t.num; ++w) {
      const int id = ff_id(t, L, w, 0, 0);
      const FFCluster c = cl[id];
      const double dx = px - c.c[0], dy = py - c.c[1], dz = pz - c.c[2];
      double v[3];
      if (dx * dx + dy * dy + dz * dz > theta * theta * c.r * c.r) {
        expand_moments(mom + (long)id * kFFMom, trs + (long)id * 12, dx, dy, dz, v);
        for (int d = 0; d < 3; ++d) v[d] *= kInv4Pi;
      } else {
        exact_cluster(t, c, rec, px, py, pz, v);
      }
      u[0] += v[0];
      u[1] += v[1];
      u[2] += v[2];
    }
  }
  for (int d = 0; d < 3; ++d) out[(long)d * nt + p] = u[d];
}

// out[t] = near groups (in order) + far groups (in order).
__global__ void k_ff_sum(const double* __restrict__ part, int nt, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  for (int d = 0; d < 3; ++d) {
    double a = 0.0, b = 0.0;
    for (int g = 0; g < kNearGroups; ++g) a += part[((long)g * 3 + d) * nt + t];
    for (int g = 0; g < kGroups; ++g) b += part[((long)(kNearGroups + g) * 3 + d) * nt + t];
    out[3 * (long)t + d] = a + b;
  }
}

// out = (-ff + sum_c bound[c]) + wake; trailing-edge rows -ff (advect.h:16-45).
__global__ void k_ff_combine(const double* __restrict__ bound, int ld, int cb,
                             const double* __restrict__ wake, int nt, double fx, double fy,
                             double fz, int te_period, int te_width, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const double mff[3] = {-fx, -fy, -fz};
  const bool te = te_period > 0 && (t % te_period) < te_width;
  for (int d = 0; d < 3; ++d) {
    double b = 0.0;
    if (cb > 0) {
      b = bound[(long)d * ld + t];
      for (int c = 1; c < cb; ++c) b = b + bound[((long)c * 3 + d) * ld + t];
    }
    out[3 * (long)t + d] = te ? mff[d] : (mff[d] + b) + wake[3 * (long)t + d];
  }
}

// Largest |a - b| and |b| over n vectors (atomic max on the bit patterns of
// non-negative doubles).
// b and x hold n points; a holds point t at a_off + t a_stride, and a's
// trailing-edge rows (te_period > 0) are skipped.
__global__ void k_ff_maxdiff(const double* __restrict__ a, const double* __restrict__ b, int n,
                             int a_off, int a_stride, int te_period, int te_width,
                             const double* __restrict__ x, const double* __restrict__ fbox,
                             double reach, unsigned long long* __restrict__ res) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0, m = 0.0;
  const long ta = a_off + (long)t * a_stride;
  if (t < n && !(te_period > 0 && (ta % te_period) < te_width)) {
    for (int d = 0; d < 3; ++d) {
      e = fmax(e, fabs(a[3 * ta + d] - b[3 * (long)t + d]));
      m = fmax(m, fabs(b[3 * (long)t + d]));
    }
    if (fbox) {  // the error the graded criterion allows at this point
      const double4 f = focus_sphere(fbox);
      const double fx = x[3 * (long)t] - f.x, fy = x[3 * (long)t + 1] - f.y,
                   fz = x[3 * (long)t + 2] - f.z;
      const double gap = sqrt(fx * fx + fy * fy + fz * fz) - f.w, L = reach * f.w;
      if (gap > L) {
        const double w = L / gap;
        e *= w * w * w;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    e = fmax(e, __shfl_xor(e, off));
    m = fmax(m, __shfl_xor(m, off));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(res, (unsigned long long)__double_as_longlong(e));
    atomicMax(res + 1, (unsigned long long)__double_as_longlong(m));
  }
}

__global__ void k_ff_zero2(unsigned long long* p) {
  p[0] = 0ull;
  p[1] = 0ull;
}

__global__ void k_gather_rows_w(const double* __restrict__ rb, int cap, int row0, int num,
                                int rows, int w, double* __restrict__ out) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)num * rows * w;
  if (k >= total) return;
  const long j = k % w, i = (k / w) % rows, n = k / ((long)w * rows);
  out[k] = rb[(n * cap + row0 + i) * w + j];
}

}  // namespace

void launch_ff_combine(hipStream_t s, const double* bound, int ld, int cb, const double* wake,
                       int nt, const double* ff, int te_period, int te_width, double* out) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_ff_combine, dim3(cdiv(nt, 256)), dim3(256), 0, s, bound, ld, cb, wake, nt,
                     ff[0], ff[1], ff[2], te_period, te_width, out);
}

// x[i] = src[off + i stride], i < n.
__global__ void k_gather_strided(const double* __restrict__ src, int off, int stride, int n,
                                 double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long k = off + (long)i * stride;
  x[3 * (long)i] = src[3 * k];
  x[3 * (long)i + 1] = src[3 * k + 1];
  x[3 * (long)i + 2] = src[3 * k + 2];
}

void launch_gather_strided(hipStream_t s, const double* src, int off, int stride, int n,
                           double* x) {
  if (n > 0)
    hipLaunchKernelGGL(k_gather_strided, dim3(cdiv(n, 256)), dim3(256), 0, s, src, off, stride,
                       n, x);
}

void launch_ff_maxdiff(hipStream_t s, const double* a, const double* b, int n,
                       unsigned long long* res, const double* x, const FFFocus* focus,
                       int a_off, int a_stride, int te_period, int te_width) {
  const bool graded = focus && focus->box && focus->reach > 0.0 && x;
  hipLaunchKernelGGL(k_ff_zero2, dim3(1), dim3(1), 0, s, res);
  if (n > 0)
    hipLaunchKernelGGL(k_ff_maxdiff, dim3(cdiv(n, 256)), dim3(256), 0, s, a, b, n, a_off,
                       a_stride, te_period, te_width, graded ? x : nullptr,
                       graded ? focus->box : nullptr,
                       graded ? focus->reach : 0.0, res);
}

void launch_ff_bbox(hipStream_t s, const double* x, int n, double* box) {
  hipLaunchKernelGGL(k_bbox, dim3(1), dim3(1024), 0, s, x, n, box);
}

void launch_gather_rows_w(hipStream_t s, const double* rb, int cap, int row0, int num, int rows,
                          int w, double* out) {
  const long total = (long)num * rows * w;
  if (total <= 0) return;
  hipLaunchKernelGGL(k_gather_rows_w, dim3(cdiv(total, 256)), dim3(256), 0, s, rb, cap, row0,
                     num, rows, w, out);
}

// ===========================================================================
FarField::~FarField() {
  void* bufs[] = {d_rec_,  d_key_,   d_key2_, d_idx_,   d_idx2_, d_hist_,
                  d_box_,  d_cl_,    d_mom_,  d_map_,  d_tmap_, d_tiles_, d_near_, d_far_,
                  d_cnt_,  d_front_, d_part_, d_w_,    d_wf_,    d_mstat_, d_samp_, d_np_, d_corr_,
                  d_tr_,   d_nrow_};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  for (auto& e : ev_used_) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  for (auto& e : ev_free_) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
}

int FarField::profile_read(hipStream_t s, size_t* launches, double* ms, double* pairs,
                           std::string* err) {
  *launches = 0;
  *ms = *pairs = 0.0;
  if (hipStreamSynchronize(s) != hipSuccess) {
    *err = "far-field profile: stream synchronisation failed";
    return UVLM_ERR_HIP;
  }
  for (auto& e : ev_used_) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, e.first, e.second) != hipSuccess) {
      *err = "far-field profile: event time failed";
      return UVLM_ERR_HIP;
    }
    *ms += t;
    ev_free_.push_back(e);
  }
  *launches = ev_used_.size();
  ev_used_.clear();
  if (np_live_) {
    unsigned long long h = 0;
    if (hipMemcpy(&h, d_np_, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) {
      *err = "far-field profile: pair count read failed";
      return UVLM_ERR_HIP;
    }
    *pairs = (double)h;
    np_live_ = false;  // the next profiled launch clears the count
  }
  return UVLM_OK;
}

int FarField::ensure(void** p, size_t* cap, size_t bytes, std::string* err) {
  if (*cap >= bytes && *p) return UVLM_OK;
  if (*p) {
    (void)hipDeviceSynchronize();  // queued kernels may still use the old buffer
    (void)hipFree(*p);
  }
  *p = nullptr;
  *cap = 0;
  // doubling: buffers that follow a growing wake are reallocated (a device
  // synchronisation each) O(log) times per run
  const size_t b = std::max<size_t>(2 * bytes, 256);
  if (hipMalloc(p, b) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    *err = "far-field buffer allocation failed";
    return UVLM_ERR_OOM;
  }
  *cap = b;
  return UVLM_OK;
}

// Targets in Morton order of their own bounding box: map[ntiles * kTileT]
// gets the target indices (stable in the index among equal keys), -1 beyond
// nt; keys_out (device, optional) the unsorted keys.
int FarField::morton_order(hipStream_t s, const double* targets, int nt, int ntiles, int* map,
                           unsigned* keys_out, std::string* err) {
  const RsPlan rp = rs_plan(nt);
  const size_t nk = (size_t)nt, nh = (size_t)rp.nb * kRsBins;
  int rc;
  if ((rc = ensure(&d_key_, &c_key_, sizeof(unsigned) * nk, err)) ||
      (rc = ensure(&d_key2_, &c_key2_, sizeof(unsigned) * nk, err)) ||
      (rc = ensure(&d_idx_, &c_idx_, sizeof(int) * nk, err)) ||
      (rc = ensure(&d_idx2_, &c_idx2_, sizeof(int) * nk, err)) ||
      (rc = ensure(&d_hist_, &c_hist_, sizeof(int) * nh, err)) ||
      (rc = ensure(&d_box_, &c_box_, sizeof(double) * (8 + 6 * kBoxParts), err)))
    return rc;
  double* tbox = (double*)d_box_;
  hipLaunchKernelGGL(k_bbox_part, dim3(kBoxParts), dim3(256), 0, s, targets, nt, tbox + 8);
  hipLaunchKernelGGL(k_bbox_fin, dim3(1), dim3(64), 0, s, (const double*)(tbox + 8), tbox);
  if (keys_out)
    hipLaunchKernelGGL(k_morton, dim3(cdiv(nt, 256)), dim3(256), 0, s, targets, nt,
                       (const double*)tbox, keys_out, (int*)nullptr);
  unsigned* K[2] = {(unsigned*)d_key_, (unsigned*)d_key2_};
  int* I[2] = {(int*)d_idx_, (int*)d_idx2_};
  if (nt <= kRsSmall) {
    hipLaunchKernelGGL(k_rs_small, dim3(1), dim3(kRsSmallThreads), 0, s, targets, nt,
                       (const double*)tbox, K[0], I[0], K[1], I[1], map, ntiles * kTileT);
  } else {
    hipLaunchKernelGGL(k_morton, dim3(cdiv(nt, 256)), dim3(256), 0, s, targets, nt,
                       (const double*)tbox, K[0], I[0]);
    int* hist = (int*)d_hist_;
    for (int p = 0; p < kRsPasses; ++p) {
      const bool last = p + 1 == kRsPasses;
      hipLaunchKernelGGL(k_rs_hist, dim3(rp.nb), dim3(kRsThreads), 0, s,
                         (const unsigned*)K[p & 1], nt, rp.per, p * kRsBits, hist);
      hipLaunchKernelGGL(k_rs_pass, dim3(rp.nb), dim3(kRsThreads), 0, s,
                         (const unsigned*)K[p & 1], (const int*)I[p & 1],
                         last ? nullptr : K[(p + 1) & 1], last ? map : I[(p + 1) & 1], nt, rp.per,
                         rp.nb, p * kRsBits, (const int*)hist, last ? ntiles * kTileT : 0);
    }
  }
  if (hipGetLastError() != hipSuccess) {
    *err = "Morton order launch failed";
    return UVLM_ERR_HIP;
  }
  return UVLM_OK;
}

int FarField::build(hipStream_t s, const double* nodes, const double* gamma, int num, int rows,
                    int cols, std::string* err) {
  built_ = false;
  if (num <= 0 || rows <= 0 || cols <= 0) {
    *err = "far field: empty source lattice";
    return UVLM_ERR_ARG;
  }
  if ((long)num * rows * cols > (1l << 28)) {
    *err = "far field: too many rings";
    return UVLM_ERR_ARG;
  }
  FFTree& t = tree_;
  t.num = num;
  t.rows = rows;
  t.cols = cols;
  int l = 0, off = 0;
  for (;;) {
    t.hr[l] = kFFRows << l;
    t.hc[l] = std::min(kFFCols << l, cols);
    t.nbr[l] = cdiv(rows, t.hr[l]);
    t.nbc[l] = cdiv(cols, t.hc[l]);
    t.off[l] = off;
    off += num * t.nbr[l] * t.nbc[l];
    if ((t.nbr[l] == 1 && t.nbc[l] == 1) || l == 31) break;
    ++l;
  }
  t.levels = l + 1;
  t.total = off;
  const int n0 = num * t.nbr[0] * t.nbc[0];
  int rc;
  if ((rc = ensure(&d_box_, &c_box_, sizeof(double) * (8 + 6 * kBoxParts), err)) ||
      (rc = ensure(&d_cl_, &c_cl_, sizeof(FFCluster) * (size_t)t.total, err)) ||
      (rc = ensure(&d_mom_, &c_mom_, sizeof(double) * kFFMom * (size_t)t.total, err)) ||
      (rc = ensure(&d_tr_, &c_tr_, sizeof(double) * 12 * (size_t)t.total, err)) ||
      (rc = ensure(&d_rec_, &c_rec_, sizeof(double) * kRec * (size_t)n0, err)))
    return rc;
  hipLaunchKernelGGL(k_ff_level0, dim3(n0), dim3(kMomThreads), 0, s, t, nodes, gamma,
                     (FFCluster*)d_cl_, (double*)d_mom_, (double*)d_rec_);
  for (int lv = 1; lv < t.levels; ++lv)
    hipLaunchKernelGGL(k_ff_parent, dim3(num * t.nbr[lv] * t.nbc[lv]), dim3(kMomThreads), 0, s,
                       t, lv, (FFCluster*)d_cl_, (double*)d_mom_);
  hipLaunchKernelGGL(k_ff_traces, dim3(cdiv(t.total, 256)), dim3(256), 0, s,
                     (const double*)d_mom_, t.total, (double*)d_tr_);
  if (hipGetLastError() != hipSuccess) {
    *err = "far-field build launch failed";
    return UVLM_ERR_HIP;
  }
  built_ = true;
  return UVLM_OK;
}

int FarField::velocity(hipStream_t s, const double* targets, int nt, double theta, double* out,
                       std::string* err, const FFFocus* focus, bool per_target, int order_tag) {
  if (nt <= 0) return UVLM_OK;
  if (!built_) {
    *err = "far field: no sources built";
    return UVLM_ERR_STATE;
  }
  const FFTree& t = tree_;
  const int ntiles = cdiv(nt, kTileT);
  int rc;
  const int n0 = t.num * t.nbr[0] * t.nbc[0];
  if ((rc = ensure(&d_map_, &c_map_, sizeof(int) * (size_t)ntiles * kTileT, err)) ||
      (rc = ensure(&d_tiles_, &c_tiles_, sizeof(double4) * (size_t)ntiles, err)) ||
      (rc = ensure(&d_near_, &c_near_, sizeof(int) * (size_t)ntiles * n0, err)) ||
      (rc = ensure(&d_far_, &c_far_, sizeof(int) * (size_t)ntiles * t.total, err)) ||
      (rc = ensure(&d_cnt_, &c_cnt_, sizeof(int2) * (size_t)ntiles, err)) ||
      (rc = ensure(&d_front_, &c_front_, sizeof(int) * 2 * (size_t)ntiles * n0, err)) ||
      (rc = ensure(&d_part_, &c_part_,
                   sizeof(double) * 3 * (kNearGroups + kGroups) * (size_t)nt, err)))
    return rc;
  // target tiles: the targets in Morton order of their own box (kept for a
  // tagged set)
  int* map = (int*)d_map_;
  if (order_tag != 0) {
    const void* before = d_tmap_;
    if ((rc = ensure(&d_tmap_, &c_tmap_, sizeof(int) * (size_t)ntiles * kTileT, err)) != UVLM_OK)
      return rc;
    map = (int*)d_tmap_;
    if (order_tag != tmap_tag_ || nt != tmap_nt_ || d_tmap_ != before ||
        tmap_uses_ >= kOrderReuse) {
      if ((rc = morton_order(s, targets, nt, ntiles, map, nullptr, err)) != UVLM_OK) return rc;
      tmap_tag_ = order_tag;
      tmap_nt_ = nt;
      tmap_uses_ = 0;
    }
    ++tmap_uses_;
  } else if ((rc = morton_order(s, targets, nt, ntiles, map, nullptr, err)) != UVLM_OK) {
    return rc;
  }
  hipLaunchKernelGGL(k_ff_tiles, dim3(ntiles), dim3(64), 0, s, targets, (const int*)map,
                     (double4*)d_tiles_);
  const bool graded = focus && focus->box && focus->reach > 0.0;
  hipLaunchKernelGGL(k_ff_lists, dim3(ntiles), dim3(64), 0, s, t, (const FFCluster*)d_cl_,
                     (const double4*)d_tiles_, targets, (const int*)map, per_target, theta,
                     graded ? focus->box : nullptr,
                     graded ? focus->reach : 0.0, graded ? focus->theta_min : theta,
                     (int*)d_near_, (int*)d_far_, (int2*)d_cnt_, (int*)d_front_);
  // heaviest tiles first (the Morton sort's buffers are free again)
  hipLaunchKernelGGL(k_order_heavy, dim3(1), dim3(1024), 0, s, (const int2*)d_cnt_, ntiles,
                     (int*)d_idx_);
  const int* order = (const int*)d_idx_;
  unsigned long long* npairs = nullptr;
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  if (prof_) {
    if ((rc = ensure(&d_np_, &c_np_, 2 * sizeof(unsigned long long), err)) != UVLM_OK) return rc;
    if (!np_live_) {
      hipLaunchKernelGGL(k_ff_zero2, dim3(1), dim3(1), 0, s, (unsigned long long*)d_np_);
      np_live_ = true;
    }
    npairs = (unsigned long long*)d_np_;
    if (!ev_free_.empty()) {
      ev = ev_free_.back();
      ev_free_.pop_back();
    } else if (hipEventCreate(&ev.first) != hipSuccess ||
               hipEventCreate(&ev.second) != hipSuccess) {
      *err = "hipEventCreate failed";
      return UVLM_ERR_HIP;
    }
    (void)hipEventRecord(ev.first, s);
  }
  if (f32c_)
    hipLaunchKernelGGL(k_ff_near<true>, dim3(ntiles, kNearGroups), dim3(64), 0, s,
                       (const FFCluster*)d_cl_, (const double*)d_rec_, targets, nt,
                       (const int*)map, n0, (const int*)d_near_, (const int2*)d_cnt_, order,
                       (double*)d_part_, npairs);
  else
    hipLaunchKernelGGL(k_ff_near<false>, dim3(ntiles, kNearGroups), dim3(64), 0, s,
                       (const FFCluster*)d_cl_, (const double*)d_rec_, targets, nt,
                       (const int*)map, n0, (const int*)d_near_, (const int2*)d_cnt_, order,
                       (double*)d_part_, npairs);
  if (ev.first) {
    (void)hipEventRecord(ev.second, s);
    ev_used_.push_back(ev);
  }
  if (mlp_on_) {
    const size_t cap = (size_t)ntiles * kNetCap;
    if ((rc = ensure(&d_corr_, &c_corr_, sizeof(float) * 32 * cap, err)) ||
        (rc = ensure(&d_nrow_, &c_nrow_, 256 + (sizeof(NetRow) + sizeof(float) * kMlpIn) * cap,
                     err)))
      return rc;
    int* nrows = (int*)d_nrow_;
    NetRow* rows = (NetRow*)((char*)d_nrow_ + 256);
    float* feats = (float*)(rows + cap);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, nrows);
    hipLaunchKernelGGL(k_ff_netrows, dim3(ntiles, kNetCap / 64), dim3(64), 0, s, t,
                       (const FFCluster*)d_cl_, (const double*)d_mom_,
                       (const double4*)d_tiles_, (const int*)d_far_, (const int2*)d_cnt_, nrows,
                       rows, feats, (float*)d_corr_);
    hipLaunchKernelGGL(k_ff_tilenet, dim3(std::min<size_t>(cdiv((long)cap, 64), 2048)), dim3(64),
                       0, s, (const int*)nrows, (const NetRow*)rows, (const float*)feats,
                       mlp_dev_view(d_w_), ymax_, (float*)d_corr_);
    hipLaunchKernelGGL(k_ff_far<true>, dim3(ntiles, kGroups), dim3(64), 0, s, t,
                       (const FFCluster*)d_cl_, (const double*)d_mom_, (const double*)d_tr_,
                       (const double*)d_rec_,
                       targets, nt, (const int*)map, (const int*)d_far_, (const int2*)d_cnt_,
                       (const double4*)d_tiles_, (const float*)d_corr_,
                       (unsigned long long*)d_mstat_, (double*)d_part_);
  } else {
    hipLaunchKernelGGL(k_ff_far<false>, dim3(ntiles, kGroups), dim3(64), 0, s, t,
                       (const FFCluster*)d_cl_, (const double*)d_mom_, (const double*)d_tr_,
                       (const double*)d_rec_,
                       targets, nt, (const int*)map, (const int*)d_far_, (const int2*)d_cnt_,
                       (const double4*)nullptr, (const float*)nullptr,
                       (unsigned long long*)nullptr, (double*)d_part_);
  }
  hipLaunchKernelGGL(k_ff_sum, dim3(cdiv(nt, 256)), dim3(256), 0, s, (const double*)d_part_, nt,
                     out);
  if (hipGetLastError() != hipSuccess) {
    *err = "far-field launch failed";
    return UVLM_ERR_HIP;
  }
  last_ntiles_ = ntiles;
  last_nt_ = nt;
  return UVLM_OK;
}

int FarField::wake_velocity(hipStream_t s, const double* nodes, const double* gamma, int num,
                            int rows, int cols, const double* targets, int nt, double theta,
                            double* out, std::string* err) {
  if (nt <= 0) return UVLM_OK;
  if (rows <= 0) {
    if (hipMemsetAsync(out, 0, sizeof(double) * 3 * (size_t)nt, s) != hipSuccess) {
      *err = "hipMemsetAsync failed";
      return UVLM_ERR_HIP;
    }
    return UVLM_OK;
  }
  int rc = build(s, nodes, gamma, num, rows, cols, err);
  if (rc != UVLM_OK) return rc;
  return velocity(s, targets, nt, theta, out, err);
}

int FarField::root_velocity(hipStream_t s, const double* x, int nt, int per_sheet, int width,
                            int near_rows, double theta, const double* near_part, int near_ld,
                            int cb, double* out, std::string* err) {
  if (nt <= 0) return UVLM_OK;
  if (!built_) {
    *err = "far field: no sources built";
    return UVLM_ERR_STATE;
  }
  hipLaunchKernelGGL(k_bound_far, dim3(cdiv(nt, 256)), dim3(256), 0, s, tree_,
                     (const FFCluster*)d_cl_, (const double*)d_mom_, (const double*)d_tr_,
                     (const double*)d_rec_, x, nt, per_sheet, width, near_rows, theta, near_part,
                     near_ld, cb, out);
  if (hipGetLastError() != hipSuccess) {
    *err = "bound far-field launch failed";
    return UVLM_ERR_HIP;
  }
  return UVLM_OK;
}

int FarField::list_stats(hipStream_t s, double* near_leaves, double* far_clusters,
                         std::string* err) {
  *near_leaves = *far_clusters = 0.0;
  if (last_ntiles_ <= 0) return UVLM_OK;
  std::vector<int2> cnt(last_ntiles_);
  if (hipMemcpyAsync(cnt.data(), d_cnt_, sizeof(int2) * cnt.size(), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    *err = "list counts read failed";
    return UVLM_ERR_HIP;
  }
  for (const int2& c : cnt) {
    *near_leaves += c.x;
    *far_clusters += c.y;
  }
  return UVLM_OK;
}

int FarField::set_mlp(hipStream_t s, const float* w, bool enable, float ymax,
                      std::string* err) {
  if (!enable) {
    mlp_on_ = false;
    return UVLM_OK;
  }
  if (!w || !(ymax > 0.0f)) {
    *err = "network weights and a positive output range are required";
    return UVLM_ERR_ARG;
  }
  int rc;
  if ((rc = ensure(&d_w_, &c_w_, mlp_dev_bytes(), err)) ||
      (rc = ensure(&d_wf_, &c_wf_, sizeof(float) * MlpShape::total, err)) ||
      (rc = ensure(&d_mstat_, &c_mstat_, 2 * sizeof(unsigned long long), err)))
    return rc;
  if (hipMemcpyAsync(d_wf_, w, sizeof(float) * MlpShape::total, hipMemcpyHostToDevice, s) !=
          hipSuccess ||
      hipMemsetAsync(d_mstat_, 0, 2 * sizeof(unsigned long long), s) != hipSuccess) {
    *err = "network upload failed";
    return UVLM_ERR_HIP;
  }
  launch_mlp_to_dev(s, (const float*)d_wf_, d_w_);
  if (hipStreamSynchronize(s) != hipSuccess) {
    *err = "network upload failed";
    return UVLM_ERR_HIP;
  }
  ymax_ = ymax;
  mlp_on_ = true;
  return UVLM_OK;
}

int FarField::mlp_stats(hipStream_t s, unsigned long long* pairs, unsigned long long* trips,
                        std::string* err) {
  unsigned long long h[2] = {0, 0};
  if (d_mstat_) {
    if (hipMemcpyAsync(h, d_mstat_, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      *err = "network statistics read failed";
      return UVLM_ERR_HIP;
    }
  }
  *pairs = h[0];
  *trips = h[1];
  return UVLM_OK;
}

int FarField::sample_residuals(hipStream_t s, const double* targets, int nt, int max_samples,
                               unsigned seed, std::vector<float>* feats,
                               std::vector<float>* resid, std::string* err) {
  feats->clear();
  resid->clear();
  if (last_ntiles_ <= 0 || last_nt_ != nt || max_samples <= 0) return UVLM_OK;
  const int ntiles = last_ntiles_;
  std::vector<int2> cnt(ntiles);
  if (hipMemcpyAsync(cnt.data(), d_cnt_, sizeof(int2) * ntiles, hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    *err = "list counts read failed";
    return UVLM_ERR_HIP;
  }
  std::vector<long> pre(ntiles + 1, 0);
  for (int i = 0; i < ntiles; ++i) pre[i + 1] = pre[i] + cnt[i].y;
  if (pre[ntiles] == 0) return UVLM_OK;
  // picks: far-list entry uniform over all tiles' entries, lane uniform
  std::vector<int4> pick(max_samples);
  unsigned long long st = 0x9e3779b97f4a7c15ull ^ seed;
  auto next = [&]() {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
  };
  for (int q = 0; q < max_samples; ++q) {
    const long e = (long)(next() % (unsigned long long)pre[ntiles]);
    const int tile = (int)(std::upper_bound(pre.begin(), pre.end(), e) - pre.begin()) - 1;
    pick[q] = int4{tile, (int)(e - pre[tile]), 0, 0};
  }
  int rc;
  const size_t sb = (size_t)max_samples;
  if ((rc = ensure(&d_samp_, &c_samp_, sb * (sizeof(int4) + sizeof(float) * (kMlpIn + kMlpOut) +
                                             sizeof(int)) + 256, err)))
    return rc;
  int4* dpick = (int4*)d_samp_;
  float* df = (float*)(dpick + sb);
  float* dr = df + sb * kMlpIn;
  int* dv = (int*)(dr + sb * kMlpOut);
  if (hipMemcpyAsync(dpick, pick.data(), sizeof(int4) * sb, hipMemcpyHostToDevice, s) !=
      hipSuccess) {
    *err = "sample upload failed";
    return UVLM_ERR_HIP;
  }
  hipLaunchKernelGGL(k_ff_samples, dim3(cdiv(max_samples, 64)), dim3(64), 0, s, tree_,
                     (const FFCluster*)d_cl_, (const double*)d_mom_, (const double*)d_rec_,
                     (const double4*)d_tiles_, (const int*)d_far_, (const int4*)dpick,
                     max_samples, df, dr, dv);
  std::vector<float> hf(sb * kMlpIn), hr(sb * kMlpOut);
  std::vector<int> hv(sb);
  if (hipMemcpyAsync(hf.data(), df, sizeof(float) * hf.size(), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipMemcpyAsync(hr.data(), dr, sizeof(float) * hr.size(), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipMemcpyAsync(hv.data(), dv, sizeof(int) * hv.size(), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    *err = "sample read-back failed";
    return UVLM_ERR_HIP;
  }
  for (size_t q = 0; q < sb; ++q)
    if (hv[q]) {
      feats->insert(feats->end(), hf.begin() + q * kMlpIn, hf.begin() + (q + 1) * kMlpIn);
      resid->insert(resid->end(), hr.begin() + q * kMlpOut, hr.begin() + (q + 1) * kMlpOut);
    }
  return UVLM_OK;
}

}  // namespace uvlm
