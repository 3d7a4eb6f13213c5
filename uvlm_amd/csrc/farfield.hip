// farfield.hip -- near/far split of the wake's Biot-Savart sum on gfx950.
//
// The wake lattice of every sheet is cut into blocks of kFFRows x kFFCols
// rings (level 0); level l merges 2 x 2 blocks of level l-1 (columns stop
// doubling at the sheet width).  A block's segments -- its sub-lattice
// with circulation zero outside it -- are closed loops, so the blocks of
// any partition of the rings sum to the whole lattice.
//
// For a target tile (64 targets, lattice tiles of 4 node rows x 16 node
// columns for wake nodes) a traversal from the top level classifies every
// block: far when |c - c_tile| - r_tile > theta r (relative to the block's
// field the expansion's error is ~ (r / |x - c|)^3 for every target of the
// tile), near
// otherwise down to level 0.  Near blocks are summed exactly (segment chains
// along the block's node rows and columns, 1/|x - p| carried along each
// chain); far blocks use the multipole expansion of the Biot-Savart law in
// the block's vorticity moments Q1..Q3:
//   A(x) = 1/4pi sum_n (-1)^n / n! Q_n[., J] d^n(1/|x - c|) / dx_J
//   u = curl A:  u_a = 1/4pi sum_n (-1)^n / n! eps_abc Q_n[c, J] T_{n+1}[b, J],
// T_m = the m-th derivative tensor of 1/r.  Level-0 moments are integrated
// segment by segment (exact polynomials in the segment parameter); parents
// are translated from their children.
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
constexpr int kTileT = 64;   // targets per tile
constexpr int kGroups = 8;   // partial-sum groups per tile (near and far each)

// --- tree layout ------------------------------------------------------------
__host__ __device__ inline int ff_id(const FFTree& t, int l, int n, int bi, int bj) {
  return t.off[l] + (n * t.nbr[l] + bi) * t.nbc[l] + bj;
}

// --- level 0: centre, radius, moments ---------------------------------------
// Circulation of ring (i, j) inside the block [r0,r1) x [c0,c1), zero outside.
__device__ __forceinline__ double gin(const double* G, int cols, int i, int j, int r0, int r1,
                                      int c0, int c1) {
  return (i >= r0 && i < r1 && j >= c0 && j < c1) ? G[(long)i * cols + j] : 0.0;
}

// Level-0 block shapes: nodes and segments of a full block.
constexpr int kL0Nodes = (kFFRows + 1) * (kFFCols + 1);
constexpr int kL0Segs = (kFFRows + 1) * kFFCols + (kFFCols + 1) * kFFRows;
constexpr int kMomThreads = 128;  // >= kFFMom: one thread per moment component
// Near-sum record of a level-0 block: its nodes (row-major, nc + 1 per row)
// and its segment strengths / 4pi (spanwise segments by node row, then
// chordwise by node column), contiguous so one wavefront loads it coalesced.
constexpr int kRecStr = 3 * kL0Nodes;
constexpr int kRec = kRecStr + kL0Segs;

// Component m of the moments of one segment a -> a + e of strength g,
// relative to the centre (A = a - c): type 1 (m < 9), 2 (< 36), 3 (< 117)
// with indices i (of e), j, k, l.
struct MomIndex {
  int type, i, j, k, l;
};
__device__ inline MomIndex mom_index(int m) {
  MomIndex x{0, 0, 0, 0, 0};
  if (m < 9) {
    x.type = 1;
    x.i = m / 3;
    x.j = m % 3;
  } else if (m < 36) {
    const int r = m - 9;
    x.type = 2;
    x.i = r / 9;
    x.j = (r / 3) % 3;
    x.k = r % 3;
  } else {
    const int r = m - 36;
    x.type = 3;
    x.i = r / 27;
    x.j = (r / 9) % 3;
    x.k = (r / 3) % 3;
    x.l = r % 3;
  }
  return x;
}

__device__ __forceinline__ double segment_moment(const MomIndex& x, double g, const double* A,
                                                 const double* E) {
  const double ge = g * E[x.i];
  const int j = x.j, k = x.k, l = x.l;
  if (x.type == 1) return ge * (A[j] + 0.5 * E[j]);
  if (x.type == 2) {
    const double p2 = A[j] * A[k] + 0.5 * (A[j] * E[k] + E[j] * A[k]) + E[j] * E[k] / 3.0;
    return ge * p2;
  }
  const double p3 = A[j] * A[k] * A[l] +
                    0.5 * (A[j] * A[k] * E[l] + A[j] * E[k] * A[l] + E[j] * A[k] * A[l]) +
                    (A[j] * E[k] * E[l] + E[j] * A[k] * E[l] + E[j] * E[k] * A[l]) / 3.0 +
                    0.25 * E[j] * E[k] * E[l];
  return ge * p3;
}

// One workgroup per level-0 block: nodes staged in LDS, centre (node mean,
// in node order) and radius, the block's segments (spanwise node rows, then
// chordwise node columns; strength = difference of the in-block
// circulations either side), then one thread per moment component summing
// over the segments in order.
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
    sp[3 * k] = p[0];
    sp[3 * k + 1] = p[1];
    sp[3 * k + 2] = p[2];
    R[3 * k] = p[0];
    R[3 * k + 1] = p[1];
    R[3 * k + 2] = p[2];
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
  // segments: spanwise (i, j) -> (i, j+1) with G(i, j) - G(i-1, j), then
  // chordwise (i, j) -> (i+1, j) with G(i, j-1) - G(i, j) (block-local i, j)
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
    const double g = ga - gb;
    R[kRecStr + s] = ga * kInv4Pi - gb * kInv4Pi;
    const double* pa = sp + 3 * (ia * (nc + 1) + ja);
    const double* pb = sp + 3 * (ib * (nc + 1) + jb);
    sA[3 * s] = pa[0] - cx;
    sA[3 * s + 1] = pa[1] - cy;
    sA[3 * s + 2] = pa[2] - cz;
    sE[3 * s] = pb[0] - pa[0];
    sE[3 * s + 1] = pb[1] - pa[1];
    sE[3 * s + 2] = pb[2] - pa[2];
    sg[s] = g;
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
//   Q2' = Q2 + Q1 (x) t (sym. in the position indices), Q3' likewise,
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
    const double* q1 = mom + (long)sid[c] * kFFMom;
    const double* q2 = q1 + 9;
    const double* q3 = q1 + 36;
    const double* tv = stv[c];
    const double a1 = q1[i * 3 + j];
    if (x.type == 1) {
      acc += a1;
    } else if (x.type == 2) {
      acc += q2[(i * 3 + j) * 3 + k] + a1 * tv[k] + q1[i * 3 + k] * tv[j];
    } else {
      acc += q3[((i * 3 + j) * 3 + k) * 3 + m] + q2[(i * 3 + j) * 3 + k] * tv[m] +
             q2[(i * 3 + j) * 3 + m] * tv[k] + q2[(i * 3 + k) * 3 + m] * tv[j] +
             a1 * tv[k] * tv[m] + q1[i * 3 + k] * tv[j] * tv[m] + q1[i * 3 + m] * tv[j] * tv[k];
    }
  }
  mom[(long)id * kFFMom + tid] = acc;
}

// --- tiles and interaction lists ----------------------------------------------
__device__ __forceinline__ int tile_target(const int* map, int tile, int lane, int nt) {
  if (map) return map[tile * kTileT + lane];
  const int t = tile * kTileT + lane;
  return t < nt ? t : -1;
}

// One wavefront per tile: centre and radius of its targets.
__global__ __launch_bounds__(64) void k_ff_tiles(const double* __restrict__ x, int nt,
                                                 const int* __restrict__ map, int ntiles,
                                                 double4* __restrict__ tiles) {
  const int tile = blockIdx.x, lane = threadIdx.x;
  const int t = tile_target(map, tile, lane, nt);
  const double px = t >= 0 ? x[3 * t] : 0.0, py = t >= 0 ? x[3 * t + 1] : 0.0,
               pz = t >= 0 ? x[3 * t + 2] : 0.0;
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
__global__ __launch_bounds__(64) void k_ff_lists(FFTree t, const FFCluster* __restrict__ cl,
                                                 const double4* __restrict__ tiles, int ntiles,
                                                 double theta, int* __restrict__ near,
                                                 int* __restrict__ far, int2* __restrict__ cnt,
                                                 int* __restrict__ front) {
  const int tile = blockIdx.x, lane = threadIdx.x;
  const double4 tc = tiles[tile];
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
      int id = -1, kind = 0, nch = 0, ci0 = 0, cj0 = 0, ncj = 0, lc = 0, cn = 0;
      if (k < nF) {
        id = F[k];
        const FFCluster c = cl[id];
        const double dx = c.c[0] - tc.x, dy = c.c[1] - tc.y, dz = c.c[2] - tc.z;
        const double dist = sqrt(dx * dx + dy * dy + dz * dz);
        if (dist - tc.w > theta * c.r) {
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
      // exclusive prefix of the children counts over the lanes
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

// --- exact sum of near blocks --------------------------------------------------
// Group g of a tile sums the tile's near blocks g, g + kGroups, ...: each
// block's record is staged in LDS (the next one is loaded into registers
// meanwhile) and walked as chains (node rows, then node columns) with the
// carried node of devmath.h.
__global__ __launch_bounds__(64) void k_ff_near(const FFCluster* __restrict__ cl,
                                                const double* __restrict__ rec,
                                                const double* __restrict__ x, int nt,
                                                const int* __restrict__ map, int n0,
                                                const int* __restrict__ near,
                                                const int2* __restrict__ cnt,
                                                double* __restrict__ part) {
  constexpr int kPer = (kRec + 63) / 64;
  __shared__ double sr[kPer * 64];
  const int tile = blockIdx.x, g = blockIdx.y, lane = threadIdx.x;
  const int tg = tile_target(map, tile, lane, nt);
  const double px = tg >= 0 ? x[3 * tg] : 0.0, py = tg >= 0 ? x[3 * tg + 1] : 0.0,
               pz = tg >= 0 ? x[3 * tg + 2] : 0.0;
  const int count = cnt[tile].x;
  const int* list = near + (long)tile * n0;
  double ux = 0.0, uy = 0.0, uz = 0.0;
  double pre[kPer];
  int id = g < count ? list[g] : 0;
  if (g < count) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int k = lane + 64 * u;
      pre[u] = k < kRec ? rec[(long)id * kRec + k] : 0.0;
    }
  }
  for (int k = g; k < count; k += kGroups) {
    const FFCluster c = cl[id];
    const int nr = c.r1 - c.r0, nc = c.c1 - c.c0;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kPer; ++u) sr[lane + 64 * u] = pre[u];
    __syncthreads();
    if (k + kGroups < count) {  // next block's record, in flight during the walk
      id = list[k + kGroups];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int q = lane + 64 * u;
        pre[u] = q < kRec ? rec[(long)id * kRec + q] : 0.0;
      }
    }
    const double* S = sr + kRecStr;
    double e[8];
    e[7] = 0.0;
    // spanwise chains: node rows 0..nr, segment strengths S[i nc + j - 1]
    for (int i = 0; i <= nr; ++i) {
      dev::Carry64 cr;
      dev::carry_reset(cr);
      const double* p = sr + 3 * i * (nc + 1);
      e[0] = p[0];
      e[1] = p[1];
      e[2] = p[2];
      e[3] = e[4] = e[5] = e[6] = 0.0;
      dev::chain_step64(e, px, py, pz, cr, ux, uy, uz);
      for (int j = 1; j <= nc; ++j) {
        p += 3;
        e[0] = p[0];
        e[1] = p[1];
        e[2] = p[2];
        e[3] = S[i * nc + j - 1];
        e[4] = p[0] - p[-3];
        e[5] = p[1] - p[-2];
        e[6] = p[2] - p[-1];
        dev::chain_step64(e, px, py, pz, cr, ux, uy, uz);
      }
    }
    // chordwise chains: node columns 0..nc, strengths S[s1 + j nr + i - 1]
    const int s1 = (nr + 1) * nc, st = 3 * (nc + 1);
    for (int j = 0; j <= nc; ++j) {
      dev::Carry64 cr;
      dev::carry_reset(cr);
      const double* p = sr + 3 * j;
      e[0] = p[0];
      e[1] = p[1];
      e[2] = p[2];
      e[3] = e[4] = e[5] = e[6] = 0.0;
      dev::chain_step64(e, px, py, pz, cr, ux, uy, uz);
      for (int i = 1; i <= nr; ++i) {
        p += st;
        e[0] = p[0];
        e[1] = p[1];
        e[2] = p[2];
        e[3] = S[s1 + j * nr + i - 1];
        e[4] = p[0] - p[-st];
        e[5] = p[1] - p[1 - st];
        e[6] = p[2] - p[2 - st];
        dev::chain_step64(e, px, py, pz, cr, ux, uy, uz);
      }
    }
  }
  if (tg >= 0) {
    part[((long)g * 3 + 0) * nt + tg] = ux;
    part[((long)g * 3 + 1) * nt + tg] = uy;
    part[((long)g * 3 + 2) * nt + tg] = uz;
  }
}

// --- far blocks: moment expansion -------------------------------------------
// u (without 1/4pi) of the moments Q at d = x - c, orders 1..3.
__device__ void expand_moments(const double* Q, double dx, double dy, double dz, double u[3]) {
  const double d[3] = {dx, dy, dz};
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double ir = rsqrt(r2) ;
  const double i2 = ir * ir, i3 = i2 * ir, i5 = i3 * i2, i7 = i5 * i2, i9 = i7 * i2;
  // S[c][b] = -Q1[c,j] T2[b,j] + 1/2 Q2[c,jk] T3[b,jk] - 1/6 Q3[c,jkl] T4[b,jkl]
  double S[3][3];
  for (int c = 0; c < 3; ++c)
    for (int b = 0; b < 3; ++b) S[c][b] = 0.0;
  for (int b = 0; b < 3; ++b)
    for (int j = 0; j < 3; ++j) {
      const double djb = (j == b) ? 1.0 : 0.0;
      const double t2 = 3.0 * d[b] * d[j] * i5 - djb * i3;
      for (int c = 0; c < 3; ++c) S[c][b] -= Q[c * 3 + j] * t2;
      for (int k = 0; k < 3; ++k) {
        const double dkb = (k == b) ? 1.0 : 0.0, djk = (j == k) ? 1.0 : 0.0;
        const double t3 = -15.0 * d[b] * d[j] * d[k] * i7 +
                          3.0 * (djb * d[k] + dkb * d[j] + djk * d[b]) * i5;
        for (int c = 0; c < 3; ++c) S[c][b] += 0.5 * Q[9 + (c * 3 + j) * 3 + k] * t3;
        for (int l = 0; l < 3; ++l) {
          const double dlb = (l == b) ? 1.0 : 0.0, djl = (j == l) ? 1.0 : 0.0,
                       dkl = (k == l) ? 1.0 : 0.0;
          const double t4 =
              105.0 * d[b] * d[j] * d[k] * d[l] * i9 -
              15.0 * (djb * d[k] * d[l] + dkb * d[j] * d[l] + dlb * d[j] * d[k] +
                      djk * d[b] * d[l] + djl * d[b] * d[k] + dkl * d[b] * d[j]) * i7 +
              3.0 * (djb * dkl + dkb * djl + dlb * djk) * i5;
          for (int c = 0; c < 3; ++c)
            S[c][b] -= (1.0 / 6.0) * Q[36 + ((c * 3 + j) * 3 + k) * 3 + l] * t4;
        }
      }
    }
  // u_a = eps_abc S[c][b]
  u[0] = S[2][1] - S[1][2];
  u[1] = S[0][2] - S[2][0];
  u[2] = S[1][0] - S[0][1];
}

// Exact velocity (with 1/4pi) at one target of the level-0 blocks under
// cluster id: every block's record walked as chains, blocks in row order.
__device__ void exact_cluster(const FFTree& t, const FFCluster& c, const double* __restrict__ rec,
                              double px, double py, double pz, double u[3]) {
  double ux = 0.0, uy = 0.0, uz = 0.0;
  const int bi0 = c.r0 / kFFRows, bi1 = cdiv(c.r1, kFFRows);
  const int bj0 = c.c0 / kFFCols, bj1 = cdiv(c.c1, kFFCols);
  double e[8];
  e[7] = 0.0;
  for (int bi = bi0; bi < bi1; ++bi)
    for (int bj = bj0; bj < bj1; ++bj) {
      const double* R = rec + (long)ff_id(t, 0, c.n, bi, bj) * kRec;
      const int nr = min(t.rows, (bi + 1) * kFFRows) - bi * kFFRows;
      const int nc = min(t.cols, (bj + 1) * kFFCols) - bj * kFFCols;
      const double* S = R + kRecStr;
      for (int i = 0; i <= nr; ++i) {
        dev::Carry64 cr;
        dev::carry_reset(cr);
        for (int j = 0; j <= nc; ++j) {
          const double* p = R + 3 * (i * (nc + 1) + j);
          e[0] = p[0];
          e[1] = p[1];
          e[2] = p[2];
          if (j > 0) {
            e[3] = S[i * nc + j - 1];
            e[4] = p[0] - p[-3];
            e[5] = p[1] - p[-2];
            e[6] = p[2] - p[-1];
          } else {
            e[3] = e[4] = e[5] = e[6] = 0.0;
          }
          dev::chain_step64(e, px, py, pz, cr, ux, uy, uz);
        }
      }
      const int s1 = (nr + 1) * nc, st = 3 * (nc + 1);
      for (int j = 0; j <= nc; ++j) {
        dev::Carry64 cr;
        dev::carry_reset(cr);
        for (int i = 0; i <= nr; ++i) {
          const double* p = R + 3 * (i * (nc + 1) + j);
          e[0] = p[0];
          e[1] = p[1];
          e[2] = p[2];
          if (i > 0) {
            e[3] = S[s1 + j * nr + i - 1];
            e[4] = p[0] - p[-st];
            e[5] = p[1] - p[1 - st];
            e[6] = p[2] - p[2 - st];
          } else {
            e[3] = e[4] = e[5] = e[6] = 0.0;
          }
          dev::chain_step64(e, px, py, pz, cr, ux, uy, uz);
        }
      }
    }
  u[0] = ux;
  u[1] = uy;
  u[2] = uz;
}

// Network features of (target, cluster) from its moments Q, offset d = x - c
// and radius r; returns the correction per unit output, s rho^3 / (4pi D^3)
// (0 when the cluster carries no circulation), s = |Q1| + |Q2|/r + |Q3|/r^2.
__device__ double ff_features(const double* Q, double dx, double dy, double dz, double r,
                              float f[kMlpIn]) {
  const double D = sqrt(dx * dx + dy * dy + dz * dz);
  const double dh[3] = {dx / D, dy / D, dz / D};
  // s = |Q1| + |Q2| / r + |Q3| / r^2 (Frobenius norms): every feature <= 1
  double a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int m = 0; m < 9; ++m) a1 += Q[m] * Q[m];
  for (int m = 9; m < 36; ++m) a2 += Q[m] * Q[m];
  for (int m = 36; m < kFFMom; ++m) a3 += Q[m] * Q[m];
  const double s = sqrt(a1) + sqrt(a2) / r + sqrt(a3) / (r * r);
  if (!(s > 0.0) || !(D > 0.0) || !(r > 0.0)) {
    for (int k = 0; k < kMlpIn; ++k) f[k] = 0.0f;
    return 0.0;
  }
  const double rho = r / D;
  f[0] = (float)dh[0];
  f[1] = (float)dh[1];
  f[2] = (float)dh[2];
  f[3] = (float)rho;
  for (int m = 0; m < 9; ++m) f[4 + m] = (float)(Q[m] / s);
  for (int ij = 0; ij < 9; ++ij) {
    const double* q2 = Q + 9 + ij * 3;
    f[13 + ij] = (float)((q2[0] * dh[0] + q2[1] * dh[1] + q2[2] * dh[2]) / (s * r));
    const double* q3 = Q + 36 + ij * 9;
    double a = 0.0;
    for (int k = 0; k < 3; ++k)
      a += (q3[3 * k] * dh[0] + q3[3 * k + 1] * dh[1] + q3[3 * k + 2] * dh[2]) * dh[k];
    f[22 + ij] = (float)(a / (s * r * r));
  }
  f[31] = 1.0f;
  return s * rho * rho * rho * kInv4Pi / (D * D * D);
}

// Far blocks of a tile, group g: moment expansion; with the network
// (MLP = true) clusters up to kMlpMaxLevel get its correction, and a pair
// whose output leaves the trained range (|y| > ymax) is summed exactly.
template <bool MLP>
__global__ __launch_bounds__(64) void k_ff_far(FFTree t, const FFCluster* __restrict__ cl,
                                               const double* __restrict__ mom,
                                               const double* __restrict__ rec,
                                               const double* __restrict__ x, int nt,
                                               const int* __restrict__ map,
                                               const int* __restrict__ far,
                                               const int2* __restrict__ cnt, MlpDev net,
                                               float ymax, unsigned long long* __restrict__ stats,
                                               double* __restrict__ part) {
  __shared__ double sq[kFFMom];
  __shared__ double sc[5];
  __shared__ unsigned short lds[MLP ? 2 * kMlpLdsHalfs : 1];
  __shared__ float ylds[MLP ? 64 * 4 : 1];
  const int tile = blockIdx.x, g = blockIdx.y, lane = threadIdx.x;
  const int tg = tile_target(map, tile, lane, nt);
  const double px = tg >= 0 ? x[3 * tg] : 0.0, py = tg >= 0 ? x[3 * tg + 1] : 0.0,
               pz = tg >= 0 ? x[3 * tg + 2] : 0.0;
  const int count = cnt[tile].y;
  const int* list = far + (long)tile * t.total;
  double ux = 0.0, uy = 0.0, uz = 0.0;
  unsigned npairs = 0, ntrips = 0;
  for (int k = g; k < count; k += kGroups) {
    const int id = list[k];
    __syncthreads();
    for (int m = lane; m < kFFMom; m += 64) sq[m] = mom[(long)id * kFFMom + m];
    if (lane < 3) sc[lane] = cl[id].c[lane];
    if (lane == 3) sc[3] = cl[id].r;
    if (lane == 4) sc[4] = (double)cl[id].level;
    __syncthreads();
    const double dx = px - sc[0], dy = py - sc[1], dz = pz - sc[2];
    double u[3];
    expand_moments(sq, dx, dy, dz, u);
    u[0] *= kInv4Pi;
    u[1] *= kInv4Pi;
    u[2] *= kInv4Pi;
    if (MLP && sc[4] <= (double)kMlpMaxLevel) {
      float f[kMlpIn];
      const double scale = ff_features(sq, dx, dy, dz, sc[3], f);
      float y[3];
      mlp_wave_bf16(net, f, lds, ylds, y);
      const float ym = fmaxf(fabsf(y[0]), fmaxf(fabsf(y[1]), fabsf(y[2])));
      if (tg >= 0 && scale > 0.0) {
        ++npairs;
        if (ym <= ymax) {
          u[0] += (double)y[0] * scale;
          u[1] += (double)y[1] * scale;
          u[2] += (double)y[2] * scale;
        } else {  // outside the trained range: the exact sum of the cluster
          ++ntrips;
          exact_cluster(t, cl[id], rec, px, py, pz, u);
        }
      }
    }
    ux += u[0];
    uy += u[1];
    uz += u[2];
  }
  if (tg >= 0) {
    part[((long)(kGroups + g) * 3 + 0) * nt + tg] = ux;
    part[((long)(kGroups + g) * 3 + 1) * nt + tg] = uy;
    part[((long)(kGroups + g) * 3 + 2) * nt + tg] = uz;
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

// Training samples: pair q = (tile, far-list entry, lane); row q of feats
// and resid (normalised remainder of the expansion, exact - expansion, in
// units of s rho^3 / (4pi D^3)); valid[q] = 0 for an empty lane, a cluster
// above kMlpMaxLevel or one without circulation.
__global__ void k_ff_samples(FFTree t, const FFCluster* __restrict__ cl,
                             const double* __restrict__ mom, const double* __restrict__ rec,
                             const double* __restrict__ x, int nt, const int* __restrict__ map,
                             const int* __restrict__ far, const int4* __restrict__ pick, int n,
                             float* __restrict__ feats, float* __restrict__ resid,
                             int* __restrict__ valid) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int4 pk = pick[q];
  const int tg = tile_target(map, pk.x, pk.z, nt);
  const int id = far[(long)pk.x * t.total + pk.y];
  const FFCluster c = cl[id];
  valid[q] = 0;
  if (tg < 0 || c.level > kMlpMaxLevel) return;
  const double px = x[3 * tg], py = x[3 * tg + 1], pz = x[3 * tg + 2];
  const double* Q = mom + (long)id * kFFMom;
  const double dx = px - c.c[0], dy = py - c.c[1], dz = pz - c.c[2];
  float f[kMlpIn];
  const double scale = ff_features(Q, dx, dy, dz, c.r, f);
  if (!(scale > 0.0)) return;
  double ue[3], ux[3];
  expand_moments(Q, dx, dy, dz, ue);
  exact_cluster(t, c, rec, px, py, pz, ux);
  for (int k = 0; k < kMlpIn; ++k) feats[(long)q * kMlpIn + k] = f[k];
  for (int d = 0; d < 3; ++d) resid[(long)q * 3 + d] = (float)((ux[d] - ue[d] * kInv4Pi) / scale);
  valid[q] = 1;
}

// out[t] = near groups (in order) + far groups (in order).
__global__ void k_ff_sum(const double* __restrict__ part, int nt, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  for (int d = 0; d < 3; ++d) {
    double a = 0.0, b = 0.0;
    for (int g = 0; g < kGroups; ++g) a += part[((long)g * 3 + d) * nt + t];
    for (int g = 0; g < kGroups; ++g) b += part[((long)(kGroups + g) * 3 + d) * nt + t];
    out[3 * (long)t + d] = a + b;
  }
}

// Lattice tiles: 4 node rows x 16 node columns per tile.
__global__ void k_ff_lattice_map(int num, int rows, int cols, int tr, int tc, int* map,
                                 int ntiles) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ntiles * kTileT) return;
  const int tile = k / kTileT, lane = k % kTileT;
  const int nbr = cdiv(rows + 1, tr), nbc = cdiv(cols + 1, tc);
  const int bj = tile % nbc, bi = (tile / nbc) % nbr, n = tile / (nbc * nbr);
  const int i = bi * tr + lane / tc, j = bj * tc + lane % tc;
  map[k] = (n < num && i <= rows && j <= cols) ? (n * (rows + 1) + i) * (cols + 1) + j : -1;
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
__global__ void k_ff_maxdiff(const double* __restrict__ a, const double* __restrict__ b, int n,
                             unsigned long long* __restrict__ res) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0, m = 0.0;
  if (t < n) {
    for (int d = 0; d < 3; ++d) {
      e = fmax(e, fabs(a[3 * (long)t + d] - b[3 * (long)t + d]));
      m = fmax(m, fabs(b[3 * (long)t + d]));
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

void launch_ff_maxdiff(hipStream_t s, const double* a, const double* b, int n,
                       unsigned long long* res) {
  hipLaunchKernelGGL(k_ff_zero2, dim3(1), dim3(1), 0, s, res);
  if (n > 0)
    hipLaunchKernelGGL(k_ff_maxdiff, dim3(cdiv(n, 256)), dim3(256), 0, s, a, b, n, res);
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
  void* bufs[] = {d_cl_, d_mom_, d_tiles_, d_near_, d_far_,  d_cnt_,  d_part_,
                  d_map_, d_w_,  d_samp_, d_front_, d_rec_, d_wf_, d_mstat_};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
}

int FarField::ensure(void** p, size_t* cap, size_t bytes, std::string* err) {
  if (*cap >= bytes && *p) return UVLM_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  const size_t b = std::max<size_t>(bytes + bytes / 4, 256);
  if (hipMalloc(p, b) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    *err = "far-field buffer allocation failed";
    return UVLM_ERR_OOM;
  }
  *cap = b;
  return UVLM_OK;
}

void FarField::build_tree(int num, int rows, int cols) {
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
}

const int* FarField::lattice_tiles(hipStream_t s, int num, int rows, int cols, int* ntiles,
                                   std::string* err) {
  const int tr = 4, tc = 16;
  const int n = num * cdiv(rows + 1, tr) * cdiv(cols + 1, tc);
  if (num != map_num_ || rows != map_rows_ || cols != map_cols_) {
    if (ensure(&d_map_, &c_map_, sizeof(int) * (size_t)n * kTileT, err) != UVLM_OK)
      return nullptr;
    hipLaunchKernelGGL(k_ff_lattice_map, dim3(cdiv((long)n * kTileT, 256)), dim3(256), 0, s, num,
                       rows, cols, tr, tc, (int*)d_map_, n);
    map_num_ = num;
    map_rows_ = rows;
    map_cols_ = cols;
    map_tiles_ = n;
  }
  *ntiles = n;
  return (const int*)d_map_;
}

int FarField::wake_velocity(hipStream_t s, const double* nodes, const double* gamma, int num,
                            int rows, int cols, const double* targets, int nt,
                            const int* tile_targets, int ntiles, double theta, double* out,
                            std::string* err) {
  near_pairs_ = far_pairs_ = 0.0;
  if (nt <= 0) return UVLM_OK;
  if (!tile_targets) ntiles = cdiv(nt, kTileT);
  if (rows <= 0) {
    if (hipMemsetAsync(out, 0, sizeof(double) * 3 * (size_t)nt, s) != hipSuccess) {
      *err = "hipMemsetAsync failed";
      return UVLM_ERR_HIP;
    }
    return UVLM_OK;
  }
  build_tree(num, rows, cols);
  const FFTree& t = tree_;
  const int n0 = num * t.nbr[0] * t.nbc[0];
  int rc;
  if ((rc = ensure(&d_cl_, &c_cl_, sizeof(FFCluster) * (size_t)t.total, err)) ||
      (rc = ensure(&d_mom_, &c_mom_, sizeof(double) * kFFMom * (size_t)t.total, err)) ||
      (rc = ensure(&d_tiles_, &c_tiles_, sizeof(double4) * (size_t)ntiles, err)) ||
      (rc = ensure(&d_near_, &c_near_, sizeof(int) * (size_t)ntiles * n0, err)) ||
      (rc = ensure(&d_far_, &c_far_, sizeof(int) * (size_t)ntiles * t.total, err)) ||
      (rc = ensure(&d_cnt_, &c_cnt_, sizeof(int2) * (size_t)ntiles, err)) ||
      (rc = ensure(&d_front_, &c_front_, sizeof(int) * 2 * (size_t)ntiles * n0, err)) ||
      (rc = ensure(&d_rec_, &c_rec_, sizeof(double) * kRec * (size_t)n0, err)) ||
      (rc = ensure(&d_part_, &c_part_, sizeof(double) * 3 * 2 * kGroups * (size_t)nt, err)))
    return rc;
  FFCluster* cl = (FFCluster*)d_cl_;
  double* mom = (double*)d_mom_;
  hipLaunchKernelGGL(k_ff_level0, dim3(n0), dim3(kMomThreads), 0, s, t, nodes, gamma, cl, mom,
                     (double*)d_rec_);
  for (int l = 1; l < t.levels; ++l) {
    const int nl = num * t.nbr[l] * t.nbc[l];
    hipLaunchKernelGGL(k_ff_parent, dim3(nl), dim3(kMomThreads), 0, s, t, l, cl, mom);
  }
  hipLaunchKernelGGL(k_ff_tiles, dim3(ntiles), dim3(64), 0, s, targets, nt, tile_targets,
                     ntiles, (double4*)d_tiles_);
  hipLaunchKernelGGL(k_ff_lists, dim3(ntiles), dim3(64), 0, s, t, (const FFCluster*)cl,
                     (const double4*)d_tiles_, ntiles, theta, (int*)d_near_, (int*)d_far_,
                     (int2*)d_cnt_, (int*)d_front_);
  hipLaunchKernelGGL(k_ff_near, dim3(ntiles, kGroups), dim3(64), 0, s, (const FFCluster*)cl,
                     (const double*)d_rec_, targets, nt, tile_targets, n0, (const int*)d_near_,
                     (const int2*)d_cnt_, (double*)d_part_);
  if (mlp_on_) {
    hipLaunchKernelGGL(k_ff_far<true>, dim3(ntiles, kGroups), dim3(64), 0, s, t,
                       (const FFCluster*)cl, (const double*)mom, (const double*)d_rec_, targets,
                       nt, tile_targets, (const int*)d_far_, (const int2*)d_cnt_,
                       mlp_dev_view(d_w_), ymax_, (unsigned long long*)d_mstat_,
                       (double*)d_part_);
  } else {
    hipLaunchKernelGGL(k_ff_far<false>, dim3(ntiles, kGroups), dim3(64), 0, s, t,
                       (const FFCluster*)cl, (const double*)mom, (const double*)d_rec_, targets,
                       nt, tile_targets, (const int*)d_far_, (const int2*)d_cnt_, MlpDev{},
                       0.0f, (unsigned long long*)nullptr, (double*)d_part_);
  }
  hipLaunchKernelGGL(k_ff_sum, dim3(cdiv(nt, 256)), dim3(256), 0, s, (const double*)d_part_, nt,
                     out);
  if (hipGetLastError() != hipSuccess) {
    *err = "far-field launch failed";
    return UVLM_ERR_HIP;
  }
  last_ntiles_ = ntiles;
  last_nt_ = nt;
  last_map_ = tile_targets;
  last_theta_ = theta;
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
    pick[q] = int4{tile, (int)(e - pre[tile]), (int)(next() % 64), 0};
  }
  int rc;
  const size_t sb = (size_t)max_samples;
  if ((rc = ensure(&d_samp_, &c_samp_, sb * (sizeof(int4) + sizeof(float) * (kMlpIn + 3) +
                                             sizeof(int)) + 256, err)))
    return rc;
  int4* dpick = (int4*)d_samp_;
  float* df = (float*)(dpick + sb);
  float* dr = df + sb * kMlpIn;
  int* dv = (int*)(dr + sb * 3);
  if (hipMemcpyAsync(dpick, pick.data(), sizeof(int4) * sb, hipMemcpyHostToDevice, s) !=
      hipSuccess) {
    *err = "sample upload failed";
    return UVLM_ERR_HIP;
  }
  hipLaunchKernelGGL(k_ff_samples, dim3(cdiv(max_samples, 64)), dim3(64), 0, s, tree_,
                     (const FFCluster*)d_cl_, (const double*)d_mom_, (const double*)d_rec_,
                     targets, nt, last_map_, (const int*)d_far_, (const int4*)dpick,
                     max_samples, df, dr, dv);
  std::vector<float> hf(sb * kMlpIn), hr(sb * 3);
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
      resid->insert(resid->end(), hr.begin() + q * 3, hr.begin() + (q + 1) * 3);
    }
  return UVLM_OK;
}

}  // namespace uvlm
