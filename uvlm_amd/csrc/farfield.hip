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

// Accumulate one segment a -> a + e of strength g into Q (relative to the
// centre: A = a - c).
__device__ __forceinline__ void add_segment_moments(double* q, int stride, double g,
                                                    const double A[3], const double E[3]) {
  for (int i = 0; i < 3; ++i) {
    const double ge = g * E[i];
    for (int j = 0; j < 3; ++j) {
      q[(i * 3 + j) * stride] += ge * (A[j] + 0.5 * E[j]);
      for (int k = 0; k < 3; ++k) {
        const double p2 = A[j] * A[k] + 0.5 * (A[j] * E[k] + E[j] * A[k]) + E[j] * E[k] / 3.0;
        q[(9 + (i * 3 + j) * 3 + k) * stride] += ge * p2;
        for (int l = 0; l < 3; ++l) {
          const double p3 = A[j] * A[k] * A[l] +
                            0.5 * (A[j] * A[k] * E[l] + A[j] * E[k] * A[l] + E[j] * A[k] * A[l]) +
                            (A[j] * E[k] * E[l] + E[j] * A[k] * E[l] + E[j] * E[k] * A[l]) / 3.0 +
                            0.25 * E[j] * E[k] * E[l];
          q[(36 + ((i * 3 + j) * 3 + k) * 3 + l) * stride] += ge * p3;
        }
      }
    }
  }
}

// One lane per level-0 block; moments accumulate in LDS (column per lane).
__global__ __launch_bounds__(64) void k_ff_level0(FFTree t, const double* __restrict__ nodes,
                                                  const double* __restrict__ gamma,
                                                  FFCluster* __restrict__ cl,
                                                  double* __restrict__ mom) {
  __shared__ double acc[kFFMom * 65];
  const int lane = threadIdx.x;
  const int q = blockIdx.x * 64 + lane;
  const int n0 = t.num * t.nbr[0] * t.nbc[0];
  for (int m = 0; m < kFFMom; ++m) acc[m * 65 + lane] = 0.0;
  if (q >= n0) return;  // (no barrier follows)
  const int bj = q % t.nbc[0], bi = (q / t.nbc[0]) % t.nbr[0], n = q / (t.nbc[0] * t.nbr[0]);
  const int C1 = t.cols + 1;
  const int r0 = bi * kFFRows, r1 = min(t.rows, r0 + kFFRows);
  const int c0 = bj * kFFCols, c1 = min(t.cols, c0 + kFFCols);
  const double* P = nodes + 3 * (long)n * (t.rows + 1) * C1;
  const double* G = gamma + (long)n * t.rows * t.cols;
  // centre and radius over the block's nodes
  double cx = 0.0, cy = 0.0, cz = 0.0;
  for (int i = r0; i <= r1; ++i)
    for (int j = c0; j <= c1; ++j) {
      const double* p = P + 3 * ((long)i * C1 + j);
      cx += p[0];
      cy += p[1];
      cz += p[2];
    }
  const double inv = 1.0 / ((r1 - r0 + 1) * (c1 - c0 + 1));
  cx *= inv;
  cy *= inv;
  cz *= inv;
  double rr = 0.0;
  for (int i = r0; i <= r1; ++i)
    for (int j = c0; j <= c1; ++j) {
      const double* p = P + 3 * ((long)i * C1 + j);
      const double dx = p[0] - cx, dy = p[1] - cy, dz = p[2] - cz;
      rr = fmax(rr, dx * dx + dy * dy + dz * dz);
    }
  double* a = acc + lane;
  auto seg = [&](int ia, int ja, int ib, int jb, double g) {
    if (g == 0.0) return;
    const double* pa = P + 3 * ((long)ia * C1 + ja);
    const double* pb = P + 3 * ((long)ib * C1 + jb);
    const double A[3] = {pa[0] - cx, pa[1] - cy, pa[2] - cz};
    const double E[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    add_segment_moments(a, 65, g, A, E);
  };
  // spanwise (i, j) -> (i, j+1): G(i, j) - G(i-1, j)
  for (int i = r0; i <= r1; ++i)
    for (int j = c0; j < c1; ++j)
      seg(i, j, i, j + 1, gin(G, t.cols, i, j, r0, r1, c0, c1) - gin(G, t.cols, i - 1, j, r0, r1, c0, c1));
  // chordwise (i, j) -> (i+1, j): G(i, j-1) - G(i, j)
  for (int j = c0; j <= c1; ++j)
    for (int i = r0; i < r1; ++i)
      seg(i, j, i + 1, j, gin(G, t.cols, i, j - 1, r0, r1, c0, c1) - gin(G, t.cols, i, j, r0, r1, c0, c1));
  FFCluster o;
  o.c[0] = cx;
  o.c[1] = cy;
  o.c[2] = cz;
  o.r = sqrt(rr);
  o.n = n;
  o.r0 = r0;
  o.r1 = r1;
  o.c0 = c0;
  o.c1 = c1;
  o.level = 0;
  o.pad0 = o.pad1 = 0;
  cl[q] = o;
  for (int m = 0; m < kFFMom; ++m) mom[(long)q * kFFMom + m] = a[m * 65];
}

// One thread per block of level l: children translated to the parent centre.
__global__ void k_ff_parent(FFTree t, int l, FFCluster* __restrict__ cl,
                            double* __restrict__ mom) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int nl = t.num * t.nbr[l] * t.nbc[l];
  if (q >= nl) return;
  const int bj = q % t.nbc[l], bi = (q / t.nbc[l]) % t.nbr[l], n = q / (t.nbc[l] * t.nbr[l]);
  const int r0 = bi * t.hr[l], r1 = min(t.rows, r0 + t.hr[l]);
  const int c0 = bj * t.hc[l], c1 = min(t.cols, c0 + t.hc[l]);
  const int lc = l - 1;
  const int ci0 = r0 / t.hr[lc], ci1 = min(t.nbr[lc], cdiv(r1, t.hr[lc]));
  const int cj0 = c0 / t.hc[lc], cj1 = min(t.nbc[lc], cdiv(c1, t.hc[lc]));
  double cx = 0.0, cy = 0.0, cz = 0.0, w = 0.0;
  for (int ci = ci0; ci < ci1; ++ci)
    for (int cj = cj0; cj < cj1; ++cj) {
      const FFCluster& ch = cl[ff_id(t, lc, n, ci, cj)];
      const double wc = (double)(ch.r1 - ch.r0) * (ch.c1 - ch.c0);
      cx += wc * ch.c[0];
      cy += wc * ch.c[1];
      cz += wc * ch.c[2];
      w += wc;
    }
  cx /= w;
  cy /= w;
  cz /= w;
  double rad = 0.0;
  double Q[kFFMom];
  for (int m = 0; m < kFFMom; ++m) Q[m] = 0.0;
  for (int ci = ci0; ci < ci1; ++ci)
    for (int cj = cj0; cj < cj1; ++cj) {
      const int id = ff_id(t, lc, n, ci, cj);
      const FFCluster& ch = cl[id];
      const double tv[3] = {ch.c[0] - cx, ch.c[1] - cy, ch.c[2] - cz};
      rad = fmax(rad, sqrt(tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2]) + ch.r);
      const double* q1 = mom + (long)id * kFFMom;
      const double* q2 = q1 + 9;
      const double* q3 = q1 + 36;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const double a1 = q1[i * 3 + j];
          Q[i * 3 + j] += a1;
          for (int k = 0; k < 3; ++k) {
            const double a2 = q2[(i * 3 + j) * 3 + k];
            Q[9 + (i * 3 + j) * 3 + k] += a2 + a1 * tv[k] + q1[i * 3 + k] * tv[j];
            for (int m = 0; m < 3; ++m) {
              Q[36 + ((i * 3 + j) * 3 + k) * 3 + m] +=
                  q3[((i * 3 + j) * 3 + k) * 3 + m] + q2[(i * 3 + j) * 3 + k] * tv[m] +
                  q2[(i * 3 + j) * 3 + m] * tv[k] + q2[(i * 3 + k) * 3 + m] * tv[j] +
                  a1 * tv[k] * tv[m] + q1[i * 3 + k] * tv[j] * tv[m] + q1[i * 3 + m] * tv[j] * tv[k];
            }
          }
        }
    }
  const int id = ff_id(t, l, n, bi, bj);
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
  for (int m = 0; m < kFFMom; ++m) mom[(long)id * kFFMom + m] = Q[m];
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

// One thread per tile: traversal of the block tree from the top level.
// near: level-0 blocks summed exactly; far: blocks used through their
// expansion.  Capacities: all level-0 blocks / all blocks.
__global__ void k_ff_lists(FFTree t, const FFCluster* __restrict__ cl,
                           const double4* __restrict__ tiles, int ntiles, double theta,
                           int* __restrict__ near, int* __restrict__ far, int2* __restrict__ cnt) {
  const int tile = blockIdx.x * blockDim.x + threadIdx.x;
  if (tile >= ntiles) return;
  const double4 tc = tiles[tile];
  const int n0 = t.num * t.nbr[0] * t.nbc[0];
  int* my_near = near + (long)tile * n0;
  int* my_far = far + (long)tile * t.total;
  int nn = 0, nf = 0;
  int stack[160];
  int sp = 0;
  const int L = t.levels - 1;
  for (int n = 0; n < t.num; ++n)
    for (int bi = 0; bi < t.nbr[L]; ++bi)
      for (int bj = 0; bj < t.nbc[L]; ++bj)
        if (sp < 160) stack[sp++] = ff_id(t, L, n, bi, bj);
  while (sp > 0) {
    const int id = stack[--sp];
    const FFCluster& c = cl[id];
    const double dx = c.c[0] - tc.x, dy = c.c[1] - tc.y, dz = c.c[2] - tc.z;
    const double dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (dist - tc.w > theta * c.r) {
      my_far[nf++] = id;
    } else if (c.level == 0) {
      my_near[nn++] = id - t.off[0];
    } else {
      const int lc = c.level - 1;
      const int ci0 = c.r0 / t.hr[lc], ci1 = min(t.nbr[lc], cdiv(c.r1, t.hr[lc]));
      const int cj0 = c.c0 / t.hc[lc], cj1 = min(t.nbc[lc], cdiv(c.c1, t.hc[lc]));
      for (int ci = ci0; ci < ci1; ++ci)
        for (int cj = cj0; cj < cj1; ++cj)
          if (sp < 160) stack[sp++] = ff_id(t, lc, c.n, ci, cj);
    }
  }
  cnt[tile] = int2{nn, nf};
}

// --- exact sum of near blocks --------------------------------------------------
// Group g of a tile sums the tile's near blocks g, g + kGroups, ...: each
// block's nodes and circulations are staged in LDS and walked as chains
// (node rows, then node columns) with the carried node of devmath.h.
__global__ __launch_bounds__(64) void k_ff_near(FFTree t, const FFCluster* __restrict__ cl,
                                                const double* __restrict__ nodes,
                                                const double* __restrict__ gamma,
                                                const double* __restrict__ x, int nt,
                                                const int* __restrict__ map,
                                                const int* __restrict__ near,
                                                const int2* __restrict__ cnt,
                                                double* __restrict__ part) {
  __shared__ double sp[(kFFRows + 1) * (kFFCols + 1) * 3];
  __shared__ double sg[kFFRows * kFFCols];
  const int tile = blockIdx.x, g = blockIdx.y, lane = threadIdx.x;
  const int tg = tile_target(map, tile, lane, nt);
  const double px = tg >= 0 ? x[3 * tg] : 0.0, py = tg >= 0 ? x[3 * tg + 1] : 0.0,
               pz = tg >= 0 ? x[3 * tg + 2] : 0.0;
  const int n0 = t.num * t.nbr[0] * t.nbc[0];
  const int count = cnt[tile].x;
  const int* list = near + (long)tile * n0;
  const int C1 = t.cols + 1;
  double ux = 0.0, uy = 0.0, uz = 0.0;
  for (int k = g; k < count; k += kGroups) {
    const FFCluster c = cl[list[k]];
    const int nr = c.r1 - c.r0, nc = c.c1 - c.c0;
    const double* P = nodes + 3 * ((long)c.n * (t.rows + 1) * C1);
    const double* G = gamma + (long)c.n * t.rows * t.cols;
    __syncthreads();
    for (int q = lane; q < (nr + 1) * (nc + 1); q += 64) {
      const int i = q / (nc + 1), j = q % (nc + 1);
      const double* p = P + 3 * ((long)(c.r0 + i) * C1 + c.c0 + j);
      sp[3 * q] = p[0];
      sp[3 * q + 1] = p[1];
      sp[3 * q + 2] = p[2];
    }
    for (int q = lane; q < nr * nc; q += 64)
      sg[q] = G[(long)(c.r0 + q / nc) * t.cols + c.c0 + q % nc] * kInv4Pi;
    __syncthreads();
    auto Gl = [&](int i, int j) -> double {
      return (i >= 0 && i < nr && j >= 0 && j < nc) ? sg[i * nc + j] : 0.0;
    };
    double e[8];
    // spanwise chains: node rows 0..nr
    for (int i = 0; i <= nr; ++i) {
      dev::Carry64 cr;
      dev::carry_reset(cr);
      for (int j = 0; j <= nc; ++j) {
        const double* p = sp + 3 * (i * (nc + 1) + j);
        e[0] = p[0];
        e[1] = p[1];
        e[2] = p[2];
        if (j > 0) {
          const double* pp = p - 3;
          e[3] = Gl(i, j - 1) - Gl(i - 1, j - 1);
          e[4] = p[0] - pp[0];
          e[5] = p[1] - pp[1];
          e[6] = p[2] - pp[2];
        } else {
          e[3] = e[4] = e[5] = e[6] = 0.0;
        }
        e[7] = 0.0;
        dev::chain_step64(e, px, py, pz, cr, ux, uy, uz);
      }
    }
    // chordwise chains: node columns 0..nc
    for (int j = 0; j <= nc; ++j) {
      dev::Carry64 cr;
      dev::carry_reset(cr);
      for (int i = 0; i <= nr; ++i) {
        const double* p = sp + 3 * (i * (nc + 1) + j);
        e[0] = p[0];
        e[1] = p[1];
        e[2] = p[2];
        if (i > 0) {
          const double* pp = p - 3 * (nc + 1);
          e[3] = Gl(i - 1, j - 1) - Gl(i - 1, j);
          e[4] = p[0] - pp[0];
          e[5] = p[1] - pp[1];
          e[6] = p[2] - pp[2];
        } else {
          e[3] = e[4] = e[5] = e[6] = 0.0;
        }
        e[7] = 0.0;
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

__global__ __launch_bounds__(64) void k_ff_far(const FFCluster* __restrict__ cl,
                                               const double* __restrict__ mom,
                                               const double* __restrict__ x, int nt,
                                               const int* __restrict__ map, int total,
                                               const int* __restrict__ far,
                                               const int2* __restrict__ cnt,
                                               double* __restrict__ part) {
  __shared__ double sq[kFFMom];
  __shared__ double sc[4];
  const int tile = blockIdx.x, g = blockIdx.y, lane = threadIdx.x;
  const int tg = tile_target(map, tile, lane, nt);
  const double px = tg >= 0 ? x[3 * tg] : 0.0, py = tg >= 0 ? x[3 * tg + 1] : 0.0,
               pz = tg >= 0 ? x[3 * tg + 2] : 0.0;
  const int count = cnt[tile].y;
  const int* list = far + (long)tile * total;
  double ux = 0.0, uy = 0.0, uz = 0.0;
  for (int k = g; k < count; k += kGroups) {
    const int id = list[k];
    __syncthreads();
    for (int m = lane; m < kFFMom; m += 64) sq[m] = mom[(long)id * kFFMom + m];
    if (lane < 3) sc[lane] = cl[id].c[lane];
    __syncthreads();
    double u[3];
    expand_moments(sq, px - sc[0], py - sc[1], pz - sc[2], u);
    ux += u[0] * kInv4Pi;
    uy += u[1] * kInv4Pi;
    uz += u[2] * kInv4Pi;
  }
  if (tg >= 0) {
    part[((long)(kGroups + g) * 3 + 0) * nt + tg] = ux;
    part[((long)(kGroups + g) * 3 + 1) * nt + tg] = uy;
    part[((long)(kGroups + g) * 3 + 2) * nt + tg] = uz;
  }
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
  void* bufs[] = {d_cl_, d_mom_, d_tiles_, d_near_, d_far_, d_cnt_, d_part_, d_map_,
                  d_w_,  d_samp_};
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
      (rc = ensure(&d_part_, &c_part_, sizeof(double) * 3 * 2 * kGroups * (size_t)nt, err)))
    return rc;
  FFCluster* cl = (FFCluster*)d_cl_;
  double* mom = (double*)d_mom_;
  hipLaunchKernelGGL(k_ff_level0, dim3(cdiv(n0, 64)), dim3(64), 0, s, t, nodes, gamma, cl, mom);
  for (int l = 1; l < t.levels; ++l) {
    const int nl = num * t.nbr[l] * t.nbc[l];
    hipLaunchKernelGGL(k_ff_parent, dim3(cdiv(nl, 64)), dim3(64), 0, s, t, l, cl, mom);
  }
  hipLaunchKernelGGL(k_ff_tiles, dim3(ntiles), dim3(64), 0, s, targets, nt, tile_targets,
                     ntiles, (double4*)d_tiles_);
  hipLaunchKernelGGL(k_ff_lists, dim3(cdiv(ntiles, 64)), dim3(64), 0, s, t, (const FFCluster*)cl,
                     (const double4*)d_tiles_, ntiles, theta, (int*)d_near_, (int*)d_far_,
                     (int2*)d_cnt_);
  hipLaunchKernelGGL(k_ff_near, dim3(ntiles, kGroups), dim3(64), 0, s, t, (const FFCluster*)cl,
                     nodes, gamma, targets, nt, tile_targets, (const int*)d_near_,
                     (const int2*)d_cnt_, (double*)d_part_);
  hipLaunchKernelGGL(k_ff_far, dim3(ntiles, kGroups), dim3(64), 0, s, (const FFCluster*)cl,
                     (const double*)mom, targets, nt, tile_targets, t.total,
                     (const int*)d_far_, (const int2*)d_cnt_, (double*)d_part_);
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

int FarField::set_mlp(hipStream_t, const float*, bool, std::string* err) {
  *err = "network correction not available";
  return UVLM_ERR_ARG;
}

int FarField::sample_residuals(hipStream_t, const double*, const double*, const double*, int,
                               int, unsigned, std::vector<float>*, std::vector<float>*,
                               std::string* err) {
  *err = "network correction not available";
  return UVLM_ERR_ARG;
}

}  // namespace uvlm
