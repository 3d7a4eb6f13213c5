// batch.hip -- many independent UVLM cases of uneven lattice sizes advanced
// together on gfx950.
//
// All cases step in lockstep, so the wake shape of every case follows from
// one step counter kept in device memory: at step s a case has s+1 wake node
// rows and s circulation rows.  Every kernel reads the counter, so the whole
// timestep is one fixed launch sequence that is captured once into a
// hipGraph and replayed every step (the host only launches the graph).
//
//   kinematics   k_b_morph, k_b_vmorph   plunge/flap/pitch on the device
//   shed         k_b_shed                ring-buffer row per case
//   geometry     k_b_geometry, k_b_pack_rings
//   matrix       k_b_matrix              static (case, target tile, ring tile) items
//   N-body       k_b_nbody<PHASE>        ONE persistent launch per phase over
//                                        (case, target tile) items fetched with an
//                                        atomic counter; sources are the case's
//                                        segment chains staged through LDS
//   solve        checked no-pivot panels (k_b_panel_np) + k_b_trsm + k_b_gemm
//                per panel over all cases (static item lists); cases whose
//                pivoting checks failed are restored and refactored with
//                partial pivoting (k_b_getf2/laswp); k_b_getrs per case
//   loads        k_b_loads               one workgroup per case
//   advection    Euler / RK2 / RK4 (reference formula) stages, k_b_update
//
// The N-body launch shape (workgroup size, LDS tile) is chosen by
// BatchEngine::autotune, which times every compiled variant on the current
// state and keeps the fastest.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/uvlm_gpu.h"
#include "batch.h"
#include "devmath.h"
#include "lu_dev.h"

namespace uvlm {
namespace {

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

struct d3 {
  double x, y, z;
};
__device__ __forceinline__ d3 ld3(const double* p) { return d3{p[0], p[1], p[2]}; }
__device__ __forceinline__ void st3(double* p, d3 a) {
  p[0] = a.x;
  p[1] = a.y;
  p[2] = a.z;
}
__device__ __forceinline__ d3 add(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 sub(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 scl(d3 a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ d3 dvs(d3 a, double s) { return d3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double dot(d3 a, d3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ double sqn(d3 a) { return (a.x * a.x + a.y * a.y) + a.z * a.z; }
__device__ __forceinline__ d3 crs(d3 a, d3 b) {
  return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ d3 normalized(d3 a) {
  const double z = sqn(a);
  if (z > 0) {
    const double n = sqrt(z);
    a.x /= n;
    a.y /= n;
    a.z /= n;
  }
  return a;
}

// morphing.cpp:23-43 with twist = bend = 0 (the batch's motion laws):
// x = z_hat plug(t) + (attack * flap) x_ref, mirrored for y < 0.
__device__ d3 perfome(const BCase& c, d3 x0, double t) {
  d3 xr = d3{x0.x - c.origin[0], x0.y - c.origin[1], x0.z - c.origin[2]};
  const bool neg = xr.y < 0;
  if (neg) xr.y *= -1;
  const double phi = c.flap[0] * cos(c.flap[1] * t + c.flap[2]);
  const double cph = cos(phi), sph = sin(phi);
  const double ca = cos(c.alpha), sa = sin(c.alpha);
  // attack * flap, rows: (ca, sa*sph, sa*cph), (0, cph, -sph), (-sa, ca*sph, ca*cph)
  const double T[3][3] = {{(ca * 1 + 0 * 0) + sa * 0, (ca * 0 + 0 * cph) + sa * sph,
                           (ca * 0 + 0 * -sph) + sa * cph},
                          {(0 * 1 + 1 * 0) + 0 * 0, (0 * 0 + 1 * cph) + 0 * sph,
                           (0 * 0 + 1 * -sph) + 0 * cph},
                          {(-sa * 1 + 0 * 0) + ca * 0, (-sa * 0 + 0 * cph) + ca * sph,
                           (-sa * 0 + 0 * -sph) + ca * cph}};
  const double gz = c.plug[0] * cos(c.plug[1] * t + c.plug[2]);
  d3 x = d3{0 * gz + ((T[0][0] * xr.x + T[0][1] * xr.y) + T[0][2] * xr.z),
            0 * gz + ((T[1][0] * xr.x + T[1][1] * xr.y) + T[1][2] * xr.z),
            1 * gz + ((T[2][0] * xr.x + T[2][1] * xr.y) + T[2][2] * xr.z)};
  if (xr.y == 0) x.y = 0;
  if (neg) x.y *= -1;
  return d3{x.x + c.origin[0], x.y + c.origin[1], x.z + c.origin[2]};
}

__device__ d3 morph_velocity(const BCase& c, d3 x0, double t) {
  const double h = 1e-6;
  const d3 a = perfome(c, x0, t - h), b = perfome(c, x0, t + h);
  return dvs(sub(b, a), 2 * h);
}

// ---- per-element kernels (index -> case through a case-index array) --------
__global__ void k_b_begin(BState* st) {
  if (threadIdx.x < 8) st->ctr[threadIdx.x] = 0u;
}
__global__ void k_b_end(BState* st) {
  if (threadIdx.x == 0) st->step += 1;
}

__global__ void k_b_morph(const BCase* __restrict__ cases, const BState* __restrict__ st,
                          const int* __restrict__ node_case, const double* __restrict__ init,
                          double* __restrict__ wing, long total) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const BCase& c = cases[node_case[k]];
  const double t = st->step * c.dt;
  st3(wing + 3 * k, perfome(c, ld3(init + 3 * k), t));
}

__global__ void k_b_vmorph(const BCase* __restrict__ cases, const BState* __restrict__ st,
                           const int* __restrict__ pcase, const double* __restrict__ p0,
                           double* __restrict__ vm, long total) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const BCase& c = cases[pcase[k]];
  st3(vm + 3 * k, morph_velocity(c, ld3(p0 + 3 * k), st->step * c.dt));
}

__global__ void k_b_copy(const double* __restrict__ a, double* __restrict__ b, long n) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) b[k] = a[k];
}

// Shed: TE node row -> wake row cap-(s+1); TE Gamma row -> Gamma row cap-s.
__global__ void k_b_shed(const BCase* __restrict__ cases, const BState* __restrict__ st,
                         const int* __restrict__ te_case, const int* __restrict__ te_j,
                         const double* __restrict__ wing, const double* __restrict__ gamma,
                         double* __restrict__ wake, double* __restrict__ wgam, int total) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const BCase& c = cases[te_case[k]];
  const int j = te_j[k], s = st->step, C1 = c.C + 1;
  const double* src = wing + 3 * (c.node_off + (long)c.R * C1 + j);
  double* dst = wake + 3 * (c.wake_off + b_node_row0(c, s) * C1 + j);
  dst[0] = src[0];
  dst[1] = src[1];
  dst[2] = src[2];
  if (s > 0 && j < c.C)
    wgam[c.wgam_off + b_gam_row0(c, s) * c.C + j] = gamma[c.ring_off + (long)(c.R - 1) * c.C + j];
}

__global__ void k_b_geometry(const BCase* __restrict__ cases, const int* __restrict__ ring_case,
                             const double* __restrict__ wing, double* __restrict__ cp,
                             double* __restrict__ nrm, double* __restrict__ area,
                             RingRec* __restrict__ rings, long total) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const BCase& c = cases[ring_case[k]];
  const int K = (int)(k - c.ring_off), i = K / c.C, j = K % c.C, C1 = c.C + 1;
  auto P = [&](int ii, int jj) { return ld3(wing + 3 * (c.node_off + (long)ii * C1 + jj)); };
  const d3 p00 = P(i, j), p10 = P(i + 1, j), p11 = P(i + 1, j + 1), p01 = P(i, j + 1);
  st3(cp + 3 * k, dvs(add(add(add(p00, p10), p11), p01), 4));
  st3(nrm + 3 * k, normalized(crs(sub(p11, p00), sub(p01, p10))));
  area[k] = sqrt(sqn(crs(sub(p10, p00), sub(p01, p00))));
  RingRec r;
  st3(r.p + 0, p00);
  st3(r.p + 3, p01);
  st3(r.p + 6, p11);
  st3(r.p + 9, p10);
  r.g = 1.0;
  r.pad[0] = r.pad[1] = r.pad[2] = 0.0;
  rings[k] = r;
}

// FP64 fast segment for the matrix (VORING order, Gamma = 1).
__device__ __forceinline__ d3 seg64(d3 a, d3 b, d3 x) {
  const double r0x = b.x - a.x, r0y = b.y - a.y, r0z = b.z - a.z;
  const double r1x = x.x - a.x, r1y = x.y - a.y, r1z = x.z - a.z;
  const double r2x = x.x - b.x, r2y = x.y - b.y, r2z = x.z - b.z;
  const double dx = __builtin_fma(r1y, r2z, -r1z * r2y);
  const double dy = __builtin_fma(r1z, r2x, -r1x * r2z);
  const double dz = __builtin_fma(r1x, r2y, -r1y * r2x);
  const double dsq = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
  const double r1s = __builtin_fma(r1z, r1z, __builtin_fma(r1y, r1y, r1x * r1x));
  const double r2s = __builtin_fma(r2z, r2z, __builtin_fma(r2y, r2y, r2x * r2x));
  const double o1 = __builtin_fma(r0z, r1z, __builtin_fma(r0y, r1y, r0x * r1x));
  const double o2 = __builtin_fma(r0z, r2z, __builtin_fma(r0y, r2y, r0x * r2x));
  const bool cut = (dsq < kCutoff) | (r1s < kCutoff * kCutoff) | (r2s < kCutoff * kCutoff);
  double K = kInv4Pi * __builtin_fma(o1, dev::rsq64(r1s), -o2 * dev::rsq64(r2s)) * dev::rcp64(dsq);
  K = cut ? 0.0 : K;
  return d3{dx * K, dy * K, dz * K};
}

constexpr int kMatRings = 32;
__global__ __launch_bounds__(256) void k_b_matrix(const BCase* __restrict__ cases,
                                                  const int4* __restrict__ items,
                                                  const RingRec* __restrict__ rings,
                                                  const double* __restrict__ cp,
                                                  const double* __restrict__ nrm,
                                                  double* __restrict__ A) {
  const int4 it = items[blockIdx.x];  // (case, target tile, ring tile, -)
  const BCase& c = cases[it.x];
  __shared__ double2 tile[kMatRings * 8];
  const int l0 = it.z * kMatRings;
  const int cnt = min(kMatRings, c.N - l0);
  const double2* src = reinterpret_cast<const double2*>(rings + c.ring_off + l0);
  for (int q = threadIdx.x; q < cnt * 8; q += 256) tile[q] = src[q];
  __syncthreads();
  const int K = it.y * 256 + threadIdx.x;
  if (K >= c.N) return;
  const d3 x = ld3(cp + 3 * (c.ring_off + K));
  const d3 nl = ld3(nrm + 3 * (c.ring_off + K));
  const double* tl = reinterpret_cast<const double*>(tile);
  for (int l = 0; l < cnt; ++l) {
    const double* r = tl + 16 * l;
    const d3 p[4] = {ld3(r), ld3(r + 3), ld3(r + 6), ld3(r + 9)};
    d3 u{0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) u = add(u, seg64(p[k], p[(k + 1) & 3], x));
    A[c.A_off + (long)(l0 + l) * c.N + K] = dot(u, nl);
  }
}

// Chains of a case lattice.  WAKE: nodes from the ring buffer (start row
// cap-(s+1)) or from a compact stage array (start row 0), Gamma rows s.
template <bool WAKE>
__global__ void k_b_pack_chains(const BCase* __restrict__ cases, const BState* __restrict__ st,
                                const long* __restrict__ chain_start, int n_cases,
                                const double* __restrict__ nodes, int compact,
                                const double* __restrict__ gam, ChainRec* __restrict__ out,
                                long total) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  // case of element q: last case with chain_start <= q
  int lo = 0, hi = n_cases - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (chain_start[mid] <= q)
      lo = mid;
    else
      hi = mid - 1;
  }
  const BCase& c = cases[lo];
  const int s = st->step;
  const int rows = WAKE ? s : c.R;
  const int C1 = c.C + 1;
  const long r = q - chain_start[lo];
  const long per = (long)(rows + 1) * C1;
  if (rows == 0 || r >= 2 * per) return;
  long row0, gstride_off;
  const double* base;
  if (WAKE) {
    row0 = compact ? 0 : b_node_row0(c, s);
    base = nodes + 3 * (compact ? c.wake_off : c.wake_off);
    gstride_off = c.wgam_off + b_gam_row0(c, s) * c.C;
  } else {
    row0 = 0;
    base = nodes + 3 * c.node_off;
    gstride_off = c.ring_off;
  }
  auto P = [&](int i, int j) { return base + 3 * ((row0 + i) * C1 + j); };
  auto G = [&](int i, int j) -> double {
    if (i < 0 || i >= rows || j < 0 || j >= c.C) return 0.0;
    return gam[gstride_off + (long)i * c.C + j];
  };
  int i, j, pi = -1, pj = -1;
  double g = 0.0;
  if (r < per) {
    i = (int)(r / C1);
    j = (int)(r % C1);
    if (j > 0) {
      pi = i;
      pj = j - 1;
      g = G(i, j - 1) - G(i - 1, j - 1);
    }
  } else {
    const long s2 = r - per;
    j = (int)(s2 / (rows + 1));
    i = (int)(s2 % (rows + 1));
    if (i > 0) {
      pi = i - 1;
      pj = j;
      g = G(i - 1, j - 1) - G(i - 1, j);
    }
  }
  ChainRec e;
  const double* p = P(i, j);
  e.p[0] = p[0];
  e.p[1] = p[1];
  e.p[2] = p[2];
  if (pi >= 0) {
    const double* a = P(pi, pj);
    e.r0[0] = p[0] - a[0];
    e.r0[1] = p[1] - a[1];
    e.r0[2] = p[2] - a[2];
  } else {
    e.r0[0] = e.r0[1] = e.r0[2] = 0.0;
  }
  e.r0sq = (e.r0[0] * e.r0[0] + e.r0[1] * e.r0[1]) + e.r0[2] * e.r0[2];
  e.g = g * kInv4Pi;
  out[q] = e;
}

// ---- persistent batched N-body ---------------------------------------------
// PHASE 0: RHS    targets cp (N),        sources wake chains
// PHASE 1: loads  targets edge mids (E), sources bound + wake chains
// PHASE 2: advect targets stage nodes ((s+1)(C+1)), bound + wake, TE row -ff
template <int PHASE, int BLOCK, int TILE>
__global__ __launch_bounds__(BLOCK) void k_b_nbody(
    const BCase* __restrict__ cases, BState* __restrict__ st, const int2* __restrict__ items,
    int n_items, int ctr_slot, const double* __restrict__ targets,
    const ChainRec* __restrict__ bchain, const ChainRec* __restrict__ wchain,
    double* __restrict__ out) {
  __shared__ double2 tile[TILE * 4];
  __shared__ int s_item;
  const int s = st->step;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(&st->ctr[ctr_slot], 1u);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= n_items) break;
    const int2 w = items[it];
    const BCase& c = cases[w.x];
    const int C1 = c.C + 1;
    long toff;
    int nt;
    if (PHASE == 0) {
      toff = c.ring_off;
      nt = c.N;
    } else if (PHASE == 1) {
      toff = c.edge_off;
      nt = c.E;
    } else {
      toff = c.wake_off;
      nt = (s + 1) * C1;
    }
    if (w.y >= nt) continue;  // tile beyond this step's targets (block-uniform)
    const int t = w.y + threadIdx.x;
    double x = 0.0, y = 0.0, z = 0.0;
    if (t < nt) {
      const double* p = targets + 3 * (toff + t);
      x = p[0];
      y = p[1];
      z = p[2];
    }
    double acc[2][3] = {{0, 0, 0}, {0, 0, 0}};
    const int ew = s > 0 ? 2 * (s + 1) * C1 : 0;
    const int eb = 2 * (c.R + 1) * C1;
    for (int part = (PHASE == 0 ? 1 : 0); part < 2; ++part) {
      const ChainRec* src = part == 0 ? bchain + c.chb_off : wchain + c.chw_off;
      const int ne = part == 0 ? eb : ew;
      dev::Carry64 cr;
      dev::carry_reset(cr);
      double ax = 0.0, ay = 0.0, az = 0.0;
      const double2* s2 = reinterpret_cast<const double2*>(src);
      for (int base = 0; base < ne; base += TILE) {
        const int cnt = min(TILE, ne - base);
        __syncthreads();
        int lng = 0;  // a segment of the tile too long for the short form
        for (int q = threadIdx.x; q < cnt * 4; q += BLOCK) {
          const double2 v = s2[(long)base * 4 + q];
          tile[q] = v;
          lng |= (q & 3) == 3 && !(v.y < dev::kShortSeg2);
        }
        const bool any_long = __syncthreads_or(lng);
        const double* tl = reinterpret_cast<const double*>(tile);
        if (any_long) {
#pragma unroll 1
          for (int k = 0; k < cnt; ++k) dev::chain_step64<true>(tl + 8 * k, x, y, z, cr, ax, ay, az);
        } else {
#pragma unroll 2
          for (int k = 0; k < cnt; ++k) dev::chain_step64<false>(tl + 8 * k, x, y, z, cr, ax, ay, az);
        }
      }
      acc[part][0] = ax;
      acc[part][1] = ay;
      acc[part][2] = az;
    }
    if (t < nt) {
      double* o = out + 3 * (toff + t);
      if (PHASE == 0) {
        o[0] = acc[1][0];
        o[1] = acc[1][1];
        o[2] = acc[1][2];
      } else if (PHASE == 2 && t < C1) {  // trailing-edge wake row
        o[0] = -c.ff[0];
        o[1] = -c.ff[1];
        o[2] = -c.ff[2];
      } else {
        o[0] = (-c.ff[0] + acc[0][0]) + acc[1][0];
        o[1] = (-c.ff[1] + acc[0][1]) + acc[1][1];
        o[2] = (-c.ff[2] + acc[0][2]) + acc[1][2];
      }
    }
  }
}

__global__ void k_b_rhs(const BCase* __restrict__ cases, const int* __restrict__ ring_case,
                        const double* __restrict__ uw, const double* __restrict__ vm,
                        const double* __restrict__ nrm, double* __restrict__ rhs, long total) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const BCase& c = cases[ring_case[k]];
  const d3 u = sub(ld3(uw + 3 * k), add(ld3(vm + 3 * k), d3{c.ff[0], c.ff[1], c.ff[2]}));
  rhs[k] = -dot(u, ld3(nrm + 3 * k));
}

// ---- batched LU (column-major A per case, lda = N) --------------------------
constexpr int kNB = 32;
typedef double v4d __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(1024) void k_b_getf2(const BCase* __restrict__ cases, int k0,
                                                  double* __restrict__ Ab, int* __restrict__ piv,
                                                  const int* __restrict__ only) {
  if (!only[blockIdx.x]) return;
  const BCase& c = cases[blockIdx.x];
  const int n = c.N;
  if (k0 >= n) return;
  const int kb = min(kNB, n - k0);
  double* A = Ab + c.A_off;
  int* ipiv = piv + c.ring_off;
  __shared__ double s_val[16];
  __shared__ int s_idx[16];
  __shared__ int s_piv;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int j = 0; j < kb; ++j) {
    const int col = k0 + j;
    double* Acol = A + (long)col * n;
    double best = -1.0;
    int bi = n;
    for (int r = col + tid; r < n; r += 1024) {
      const double v = fabs(Acol[r]);
      if (v > best) {
        best = v;
        bi = r;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_down(best, off);
      const int oi = __shfl_down(bi, off);
      if (ov > best || (ov == best && oi < bi)) {
        best = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      s_val[wid] = best;
      s_idx[wid] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      double b = s_val[0];
      int bj = s_idx[0];
      for (int w = 1; w < 16; ++w)
        if (s_val[w] > b || (s_val[w] == b && s_idx[w] < bj)) {
          b = s_val[w];
          bj = s_idx[w];
        }
      if (!(b > 0.0) || bj >= n) bj = col;
      s_piv = bj;
      ipiv[col] = bj;
    }
    __syncthreads();
    const int pv = s_piv;
    if (pv != col && tid < kb) {
      double* a = A + (long)(k0 + tid) * n;
      const double tmp = a[col];
      a[col] = a[pv];
      a[pv] = tmp;
    }
    __syncthreads();
    const double d = Acol[col];
    if (d != 0.0)
      for (int r = col + 1 + tid; r < n; r += 1024) Acol[r] = Acol[r] / d;
    __syncthreads();
    for (int cc = col + 1; cc < k0 + kb; ++cc) {
      double* Ac = A + (long)cc * n;
      const double u = Ac[col];
      for (int r = col + 1 + tid; r < n; r += 1024) Ac[r] = Ac[r] - Acol[r] * u;
    }
    __syncthreads();
  }
}

__global__ void k_b_laswp(const BCase* __restrict__ cases, int k0, double* __restrict__ Ab,
                          const int* __restrict__ piv, const int* __restrict__ only) {
  if (!only[blockIdx.y]) return;
  const BCase& c = cases[blockIdx.y];
  const int n = c.N;
  if (k0 >= n) return;
  const int kb = min(kNB, n - k0);
  const int cc = blockIdx.x * blockDim.x + threadIdx.x;
  if (cc >= n - kb) return;
  const int col = cc < k0 ? cc : cc + kb;
  double* a = Ab + c.A_off + (long)col * n;
  const int* ipiv = piv + c.ring_off;
  for (int j = 0; j < kb; ++j) {
    const int r = k0 + j, p = ipiv[r];
    if (p != r) {
      const double t = a[r];
      a[r] = a[p];
      a[p] = t;
    }
  }
}

// U12 solve per case.  dblk (nullable): the diagonal blocks k_b_panel_np
// left (one kNB x kNB block per case, column-major); L11 is read from there
// and workgroup 0 of each case stores its block in A.  Without dblk (the
// pivoted path, factored in place by k_b_getf2) L11 is read from A.
__global__ __launch_bounds__(64) void k_b_trsm(const BCase* __restrict__ cases, int k0,
                                               double* __restrict__ Ab,
                                               const int* __restrict__ only,
                                               const double* __restrict__ dblk) {
  if (only && !only[blockIdx.y]) return;
  const BCase& c = cases[blockIdx.y];
  const int n = c.N;
  if (k0 >= n) return;
  const int kb = min(kNB, n - k0);
  double* A = Ab + c.A_off;
  __shared__ double L[kNB][kNB + 1];
  const double* db = dblk ? dblk + (long)blockIdx.y * kNB * kNB : nullptr;
  for (int q = threadIdx.x; q < kNB * kNB; q += 64) {
    const int r = q % kNB, cc = q / kNB;
    const bool in = r < kb && cc < kb;
    const double v = !in ? 0.0 : db ? db[cc * kNB + r] : A[(long)(k0 + cc) * n + k0 + r];
    L[r][cc] = v;
    if (db && in && blockIdx.x == 0) A[(long)(k0 + cc) * n + k0 + r] = v;
  }
  const int t = threadIdx.x;
  const int col = k0 + kb + blockIdx.x * 64 + t;
  double* a = A + (long)col * n + k0;
  double x[kNB];  // this lane's column, in registers (zero beyond kb)
#pragma unroll
  for (int i = 0; i < kNB; ++i) x[i] = (col < n && i < kb) ? a[i] : 0.0;
  __syncthreads();
  if (col >= n) return;
  const double* Lb = &L[0][0];
#pragma unroll
  for (int j = 0; j < kNB; ++j) {
    int o = j;  // column offset through an empty asm (see k_b_panel_np)
    asm volatile("" : "+s"(o) : "v"(x[j > 0 ? j - 1 : 0]));
    const double xj = x[j];
#pragma unroll
    for (int i = j + 1; i < kNB; ++i) x[i] = x[i] - Lb[i * (kNB + 1) + o] * xj;
  }
#pragma unroll
  for (int i = 0; i < kNB; ++i)
    if (i < kb) a[i] = x[i];
}

// items: (case, row tile, col tile, -) of the trailing update after panel k0
__global__ __launch_bounds__(256) void k_b_gemm(const BCase* __restrict__ cases,
                                                const int4* __restrict__ items, int k0,
                                                double* __restrict__ Ab,
                                                const int* __restrict__ only) {
  const int4 it = items[blockIdx.x];
  if (only && !only[it.x]) return;
  const BCase& c = cases[it.x];
  const int n = c.N;
  const int kb = min(kNB, n - k0);
  double* A = Ab + c.A_off;
  __shared__ double sA[64][kNB + 1];
  __shared__ double sB[kNB][64 + 1];
  const int s0 = k0 + kb;
  const int m0 = s0 + it.y * 64;
  const int c0 = s0 + it.z * 64;
  for (int q = threadIdx.x; q < 64 * kNB; q += 256) {
    const int r = q % 64, k = q / 64;
    sA[r][k] = (m0 + r < n && k < kb) ? A[(long)(k0 + k) * n + m0 + r] : 0.0;
  }
  for (int q = threadIdx.x; q < kNB * 64; q += 256) {
    const int k = q % kNB, cc = q / kNB;
    sB[k][cc] = (c0 + cc < n && k < kb) ? A[(long)(c0 + cc) * n + k0 + k] : 0.0;
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wr = w >> 1, wc = w & 1;
  const int l16 = lane & 15, lq = lane >> 4;
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < kNB; k += 4) {
    double ua[2], lb[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) ua[j] = sB[k + lq][wc * 32 + j * 16 + l16];
#pragma unroll
    for (int i = 0; i < 2; ++i) lb[i] = sA[wr * 32 + i * 16 + l16][k + lq];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[j], lb[i], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wr * 32 + i * 16 + l16;
        const int col = c0 + wc * 32 + j * 16 + 4 * r + lq;  // f64 result rows interleave
        if (row < n && col < n) {
          double* p = A + (long)col * n + row;
          *p = *p - acc[i][j][r];
        }
      }
}

// Checked no-pivot panel (the single-case path's k_panel_np, per case):
// grid (row blocks of 64, cases); every workgroup factors the case's top
// block itself; workgroup 0 writes it to the case's dblk slot (k_b_trsm
// stores it in A: a workgroup of this launch dispatched after workgroup 0
// finished would otherwise load the factored block), and into A directly
// when it is the case's only workgroup; flag[case] <- 1 when a comparison
// of partial pivoting would have chosen an interchange.
__global__ __launch_bounds__(64) void k_b_panel_np(const BCase* __restrict__ cases, int k0,
                                                   double* __restrict__ Ab,
                                                   int* __restrict__ piv,
                                                   int* __restrict__ flag,
                                                   double* __restrict__ dblk) {
  const BCase& c = cases[blockIdx.y];
  const int n = c.N;
  if (k0 >= n) return;
  const int kb = min(kNB, n - k0);
  if ((int)blockIdx.x * 64 >= max(n - k0 - kb, 1)) return;
  double* A = Ab + c.A_off;
  __shared__ double D[kNB][kNB + 1];
  __shared__ int s_flag;
  const int t = threadIdx.x;
  if (t == 0) s_flag = 0;
  // rows in registers (one lane per row, columns unrolled), identity padding
  // beyond kb; the diagonal block's final rows are published through LDS
  const int row = k0 + kb + blockIdx.x * 64 + t;
  double x[kNB];
#pragma unroll
  for (int cc = 0; cc < kNB; ++cc)
    x[cc] = (row < n && cc < kb) ? A[(long)(k0 + cc) * n + row] : 0.0;
  double d[kNB];
#pragma unroll
  for (int cc = 0; cc < kNB; ++cc)
    d[cc] = (t < kb && cc < kb) ? A[(long)(k0 + cc) * n + k0 + t] : (cc == t ? 1.0 : 0.0);
  bool bad = false;
#pragma unroll
  for (int j = 0; j < kNB; ++j) {
    if (t == j) {
#pragma unroll
      for (int cc = 0; cc < kNB; ++cc) D[j][cc] = d[cc];
    }
    __syncthreads();
    const double pv = D[j][j];
    if (t > j && t < kb) {
      double l = d[j];
      bad |= fabs(l) > fabs(pv) || pv == 0.0;
      if (pv != 0.0) l = l / pv;
      d[j] = l;
#pragma unroll
      for (int cc = j + 1; cc < kNB; ++cc) d[cc] = d[cc] - l * D[j][cc];
    }
  }
  if (blockIdx.x == 0) {
    const bool alone = n - k0 - kb <= 64;  // the case's only active workgroup
    double* db = dblk + (long)blockIdx.y * kNB * kNB;
    for (int q = t; q < kb * kb; q += 64) {
      const int r = q % kb, cc = q / kb;
      db[cc * kNB + r] = D[r][cc];
      if (alone) A[(long)(k0 + cc) * n + k0 + r] = D[r][cc];
    }
    if (t < kb) piv[c.ring_off + k0 + t] = k0 + t;
  }
  const double* Db = &D[0][0];
#pragma unroll
  for (int j = 0; j < kNB; ++j) {
    // row offset through an empty asm that takes x[j - 1]: row j's reads are
    // issued once step j - 1 starts, not hoisted all to the top
    int o = j * (kNB + 1);
    asm volatile("" : "+s"(o) : "v"(x[j > 0 ? j - 1 : 0]));
    const double pv = Db[o + j];
    double xj = x[j];
    bad |= fabs(xj) > fabs(pv);
    if (pv != 0.0) xj = xj / pv;
    x[j] = xj;
#pragma unroll
    for (int cc = j + 1; cc < kNB; ++cc) x[cc] = x[cc] - xj * Db[o + cc];
  }
  if (row < n) {
#pragma unroll
    for (int cc = 0; cc < kNB; ++cc)
      if (cc < kb) A[(long)(k0 + cc) * n + row] = x[cc];
  }
  if (bad) s_flag = 1;
  __syncthreads();
  if (t == 0 && s_flag) atomicOr(&flag[blockIdx.y], 1);
}

__global__ void k_b_init_flags(int* flag, int n, int value) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = value;
}

// Flagged cases only: A <- saved copy (the pivoted factorisation restarts).
__global__ void k_b_restore(const BCase* __restrict__ cases, const int* __restrict__ flag,
                            const double* __restrict__ saved, double* __restrict__ Ab) {
  const BCase& c = cases[blockIdx.y];
  if (!flag[blockIdx.y]) return;
  const long nn = (long)c.N * c.N;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nn;
       q += (long)gridDim.x * blockDim.x)
    Ab[c.A_off + q] = saved[c.A_off + q];
}

// Substitution per case (one workgroup each): x = U^-1 L^-1 P rhs -> gamma,
// x resident in LDS (dynamic, maxN doubles) or, for large cases, in gamma.
__global__ __launch_bounds__(1024) void k_b_getrs(const BCase* __restrict__ cases,
                                                  const double* __restrict__ Ab,
                                                  const int* __restrict__ piv,
                                                  const double* __restrict__ rhs,
                                                  double* __restrict__ gamma, int lds_n) {
  extern __shared__ unsigned char s_raw[];
  __shared__ double D[64][65];
  const BCase& c = cases[blockIdx.x];
  const int n = c.N;
  double* x = n <= lds_n ? reinterpret_cast<double*>(s_raw) : gamma + c.ring_off;
  const int* ipiv = piv + c.ring_off;
  for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = rhs[c.ring_off + i];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int i = 0; i < n; ++i) {
      const int p = ipiv[i];
      if (p != i) {
        const double t = x[i];
        x[i] = x[p];
        x[p] = t;
      }
    }
  lu_substitute(Ab + c.A_off, n, x, D);
  if (n <= lds_n)
    for (int i = threadIdx.x; i < n; i += blockDim.x) gamma[c.ring_off + i] = x[i];
}

__global__ void k_b_edges(const BCase* __restrict__ cases, const int* __restrict__ edge_case,
                          const double* __restrict__ wing, double* __restrict__ mid, long total) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const BCase& c = cases[edge_case[k]];
  const int e = (int)(k - c.edge_off), C1 = c.C + 1;
  int ia, ja, ib, jb;
  if (e < c.R * c.C) {
    ia = ib = e / c.C;
    ja = e % c.C;
    jb = ja + 1;
  } else {
    const int q = e - c.R * c.C;
    ia = q / C1;
    ib = ia + 1;
    ja = jb = q % C1;
  }
  const d3 a = ld3(wing + 3 * (c.node_off + (long)ia * C1 + ja));
  const d3 b = ld3(wing + 3 * (c.node_off + (long)ib * C1 + jb));
  st3(mid + 3 * k, dvs(add(a, b), 2));
}

// Joukowski loads of case blockIdx.x into hist[step][case][2].
__global__ __launch_bounds__(256) void k_b_loads(
    const BCase* __restrict__ cases, const BState* __restrict__ st,
    const double* __restrict__ wing, const double* __restrict__ gamma,
    const double* __restrict__ gprev, const double* __restrict__ nrm,
    const double* __restrict__ area, const double* __restrict__ uflow,
    const double* __restrict__ vme, int n_cases, double* __restrict__ hist) {
  const BCase& c = cases[blockIdx.x];
  const int R = c.R, C = c.C, C1 = C + 1;
  auto P = [&](int i, int j) { return ld3(wing + 3 * (c.node_off + (long)i * C1 + j)); };
  auto U = [&](int e) {
    const long q = 3 * (c.edge_off + e);
    return sub(ld3(uflow + q), ld3(vme + q));
  };
  double f[6] = {0, 0, 0, 0, 0, 0};
  for (int K = threadIdx.x; K < R * C; K += blockDim.x) {
    const int i = K / C, j = K % C;
    const double g = gamma[c.ring_off + K];
    const d3 p0 = P(i, j), p1 = P(i, j + 1), p2 = P(i + 1, j + 1), p3 = P(i + 1, j);
    d3 sm = scl(crs(U(i * C + j), sub(p1, p0)), g);
    sm = add(sm, scl(crs(U(R * C + i * C1 + j + 1), sub(p2, p1)), g));
    if (i != R - 1) sm = add(sm, scl(crs(U((i + 1) * C + j), sub(p3, p2)), g));
    sm = add(sm, scl(crs(U(R * C + i * C1 + j), sub(p0, p3)), g));
    f[0] += sm.x;
    f[1] += sm.y;
    f[2] += sm.z;
    const long KK = c.ring_off + K;
    const double dg_dt = (g - gprev[KK]) / c.dt;
    const d3 df = scl(scl(ld3(nrm + 3 * KK), dg_dt), area[KK]);
    f[3] += df.x;
    f[4] += df.y;
    f[5] += df.z;
  }
  __shared__ double red[6][256];
  for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = f[q];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const d3 F = add(d3{red[0][0], red[1][0], red[2][0]}, d3{red[3][0], red[4][0], red[5][0]});
    const double den = 0.5 * c.Q * c.Q * c.ref_area;
    double* o = hist + b_hist_index(st->step, n_cases, blockIdx.x);
    o[0] = F.x / den;
    o[1] = F.z / den;
  }
}

// ring buffer rows [cap-(s+1), cap) -> compact stage array (same offsets)
__global__ void k_b_gather(const BCase* __restrict__ cases, const BState* __restrict__ st,
                           const int* __restrict__ wn_case, const double* __restrict__ wake,
                           double* __restrict__ pos, long total) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const BCase& c = cases[wn_case[k]];
  const int C1 = c.C + 1, s = st->step;
  const long local = k - c.wake_off;
  if (local >= (long)(s + 1) * C1) return;
  const double* src = wake + 3 * (c.wake_off + b_node_row0(c, s) * C1 + local);
  pos[3 * k + 0] = src[0];
  pos[3 * k + 1] = src[1];
  pos[3 * k + 2] = src[2];
}

__global__ void k_b_stage(const BCase* __restrict__ cases, const BState* __restrict__ st,
                          const int* __restrict__ wn_case, const double* __restrict__ pos,
                          const double* __restrict__ kv, double div, double* __restrict__ out,
                          long total) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const BCase& c = cases[wn_case[k]];
  if (k - c.wake_off >= (long)(st->step + 1) * (c.C + 1)) return;
  for (int d = 0; d < 3; ++d) {
    const double v = kv[3 * k + d] * c.dt;
    out[3 * k + d] = pos[3 * k + d] + (div == 1.0 ? v : v / div);
  }
}

__global__ void k_b_update(const BCase* __restrict__ cases, const BState* __restrict__ st,
                           const int* __restrict__ wn_case, int scheme, int ign,
                           const double* __restrict__ pos, const double* __restrict__ k1,
                           const double* __restrict__ k2, const double* __restrict__ k3,
                           double* __restrict__ wake, long total) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const BCase& c = cases[wn_case[k]];
  const int C1 = c.C + 1, s = st->step;
  const long local = k - c.wake_off;
  if (local >= (long)(s + 1) * C1) return;
  const int i = (int)(local / C1);
  const d3 w = ld3(pos + 3 * k), a = ld3(k1 + 3 * k);
  d3 r;
  if (scheme == UVLM_ADVECT_RK2 && i > ign) {
    r = add(w, dvs(scl(add(a, ld3(k2 + 3 * k)), c.dt), 2));
  } else if (scheme == UVLM_ADVECT_RK4 && i > ign) {
    const d3 b = ld3(k2 + 3 * k), cc = ld3(k3 + 3 * k);
    r = add(w, dvs(scl(add(add(add(a, scl(b, 2)), scl(cc, 2)), cc), c.dt), 6));
  } else {
    r = add(w, scl(a, c.dt));
  }
  st3(wake + 3 * (c.wake_off + b_node_row0(c, s) * C1 + local), r);
}

// ---- launch-shape variants of the persistent N-body (autotuned) -----------
// (workgroup size = targets per item, LDS source tile in chain elements):
// (256, 64), (256, 128), (128, 128), (512, 128)
constexpr int kVariants = 4;
constexpr int kVariantBlock[kVariants] = {256, 256, 128, 512};

template <int PHASE>
void launch_variant(int v, hipStream_t s, int grid, const BCase* cases, BState* st,
                    const int2* items, int n, int slot, const double* targets,
                    const ChainRec* bch, const ChainRec* wch, double* out) {
  switch (v) {
    case 0:
      hipLaunchKernelGGL((k_b_nbody<PHASE, 256, 64>), dim3(grid), dim3(256), 0, s, cases, st,
                         items, n, slot, targets, bch, wch, out);
      break;
    case 1:
      hipLaunchKernelGGL((k_b_nbody<PHASE, 256, 128>), dim3(grid), dim3(256), 0, s, cases, st,
                         items, n, slot, targets, bch, wch, out);
      break;
    case 2:
      hipLaunchKernelGGL((k_b_nbody<PHASE, 128, 128>), dim3(grid), dim3(128), 0, s, cases, st,
                         items, n, slot, targets, bch, wch, out);
      break;
    default:
      hipLaunchKernelGGL((k_b_nbody<PHASE, 512, 128>), dim3(grid), dim3(512), 0, s, cases, st,
                         items, n, slot, targets, bch, wch, out);
      break;
  }
}

template <typename T>
int dev_alloc(std::vector<void*>& allocs, T** p, size_t n, std::string* err) {
  void* q = nullptr;
  if (hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) {
    (void)hipGetLastError();
    *err = "device allocation failed";
    return UVLM_ERR_OOM;
  }
  allocs.push_back(q);
  *p = reinterpret_cast<T*>(q);
  return UVLM_OK;
}

}  // namespace

#define BTRY(x)                     \
  do {                              \
    int rc_ = (x);                  \
    if (rc_ != UVLM_OK) return rc_; \
  } while (0)
#define BHIP(x)                                            \
  do {                                                     \
    hipError_t e_ = (x);                                   \
    if (e_ != hipSuccess) {                                \
      *err = std::string(#x ": ") + hipGetErrorString(e_); \
      return UVLM_ERR_HIP;                                 \
    }                                                      \
  } while (0)

// ===========================================================================
BatchEngine::~BatchEngine() { release(); }

void BatchEngine::release() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graph_exec_ = nullptr;
  graph_ = nullptr;
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (void* p : allocs_) (void)hipFree(p);
  allocs_.clear();
}

int batch_plan(const BatchCaseSpec* specs, size_t n, size_t max_steps, BatchPlan* plan,
               std::string* err) {
  const int cap = (int)max_steps + 2;
  BatchPlan& P = *plan;
  P = BatchPlan{};
  P.cases.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const BatchCaseSpec& sp = specs[i];
    if (sp.rows == 0 || sp.cols == 0 || (sp.cols & 1)) {
      *err = "case " + std::to_string(i) + ": columns must be even and the lattice non-empty";
      return UVLM_ERR_ARG;
    }
    if (sp.rows * sp.cols > 8192) {
      *err = "case " + std::to_string(i) + ": lattice too large for the batched solve";
      return UVLM_ERR_ARG;
    }
    BCase& c = P.cases[i];
    std::memset(&c, 0, sizeof(c));
    c.R = (int)sp.rows;
    c.C = (int)sp.cols;
    c.N = c.R * c.C;
    c.E = c.R * (2 * c.C + 1);
    c.cap = cap;
    c.node_off = P.nodes;
    c.ring_off = P.rings;
    c.edge_off = P.edges;
    c.A_off = P.A;
    c.wake_off = P.wake;
    c.wgam_off = P.wgam;
    c.chb_off = P.cb;
    c.chw_off = P.cw;
    P.nodes += (long)(c.R + 1) * (c.C + 1);
    P.rings += c.N;
    P.edges += c.E;
    P.A += (long)c.N * c.N;
    P.wake += (long)cap * (c.C + 1);
    P.wgam += (long)cap * c.C;
    P.cb += b_chain_elems(c.R, c.C);
    P.cw += b_chain_elems(cap - 1, c.C);
    P.te += c.C + 1;
    c.dt = sp.dt;
    for (int d = 0; d < 3; ++d) {
      c.ff[d] = sp.ff[d];
      c.origin[d] = sp.origin[d];
      c.plug[d] = sp.plug[d];
      c.flap[d] = sp.flap[d];
    }
    c.alpha = sp.alpha;
    c.Q = std::sqrt((sp.ff[0] * sp.ff[0] + sp.ff[1] * sp.ff[1]) + sp.ff[2] * sp.ff[2]);
    c.ref_area = sp.chord * sp.span;
    P.maxN = std::max(P.maxN, c.N);
  }
  P.hist = 2 * (long)n * (long)(max_steps + 2);
  return UVLM_OK;
}

int batch_check_extents(const BatchPlan& P, size_t max_steps, std::string* err) {
  const int n = (int)P.cases.size();
  char buf[256];
  auto bad = [&](const char* what, int cs, int s) {
    snprintf(buf, sizeof(buf), "batch plan: %s out of bounds (case %d, step %d)", what, cs, s);
    *err = buf;
    return UVLM_ERR_CAPACITY;
  };
  // [lo, hi) must lie in [base, base + len) and in [0, total)
  auto inside = [](long lo, long hi, long base, long len, long total) {
    return lo >= base && hi <= base + len && lo >= 0 && hi <= total && lo <= hi;
  };
  for (int i = 0; i < n; ++i) {
    const BCase& c = P.cases[i];
    const int C1 = c.C + 1;
    const long wlen = (long)c.cap * C1, glen = (long)c.cap * c.C;
    for (int s = 0; s <= (int)max_steps; ++s) {
      const long r0 = b_node_row0(c, s), g0 = b_gam_row0(c, s);
      // k_b_shed: node row 0 (C + 1 nodes), circulation row 0 (s > 0)
      if (!inside(c.wake_off + r0 * C1, c.wake_off + r0 * C1 + C1, c.wake_off, wlen, P.wake))
        return bad("shed node row", i, s);
      if (s > 0 && !inside(c.wgam_off + g0 * c.C, c.wgam_off + g0 * c.C + c.C, c.wgam_off, glen,
                           P.wgam))
        return bad("shed circulation row", i, s);
      // k_b_gather / k_b_update / k_b_pack_chains<true> (ring buffer): node
      // rows 0..s; circulation rows 0..s-1
      if (!inside(c.wake_off + r0 * C1, c.wake_off + (r0 + s + 1) * C1, c.wake_off, wlen, P.wake))
        return bad("wake node rows", i, s);
      if (s > 0 && !inside(c.wgam_off + g0 * c.C, c.wgam_off + (g0 + s) * c.C, c.wgam_off, glen,
                           P.wgam))
        return bad("wake circulation rows", i, s);
      // compact stage arrays (same offsets, rows 0..s) and the advection
      // N-body's targets / outputs
      if (!inside(c.wake_off, c.wake_off + (long)(s + 1) * C1, c.wake_off, wlen, P.wake))
        return bad("stage rows", i, s);
      // wake chain elements written and read per step
      if (!inside(c.chw_off, c.chw_off + b_chain_elems(s, c.C), c.chw_off,
                  b_chain_elems(c.cap - 1, c.C), P.cw))
        return bad("wake chain elements", i, s);
      // k_b_loads: (CD, CL) of this step
      const long h = b_hist_index(s, n, i);
      if (!inside(h, h + 2, 0, P.hist, P.hist)) return bad("load history", i, s);
    }
    if (!inside(c.chb_off, c.chb_off + b_chain_elems(c.R, c.C), c.chb_off,
                b_chain_elems(c.R, c.C), P.cb))
      return bad("bound chain elements", i, 0);
    if (!inside(c.A_off, c.A_off + (long)c.N * c.N, c.A_off, (long)c.N * c.N, P.A))
      return bad("influence matrix", i, 0);
  }
  return UVLM_OK;
}

int BatchEngine::init(hipStream_t s, const BatchCaseSpec* specs, const double* wing_init,
                      size_t n, size_t max_steps, int scheme, int ignore_steps,
                      std::string* err) {
  release();
  stream_ = s;
  n_ = (int)n;
  max_steps_ = (int)max_steps;
  scheme_ = scheme;
  ign_ = ignore_steps;
  steps_done_ = 0;
  if (n == 0) {
    *err = "no cases";
    return UVLM_ERR_ARG;
  }
  if (scheme != UVLM_ADVECT_EULER && scheme != UVLM_ADVECT_RK2 && scheme != UVLM_ADVECT_RK4) {
    *err = "batched advection supports euler, RK2 and RK4";
    return UVLM_ERR_ARG;
  }
  BatchPlan plan;
  BTRY(batch_plan(specs, n, max_steps, &plan, err));
  BTRY(batch_check_extents(plan, max_steps, err));
  const int cap = (int)max_steps + 2;
  const std::vector<BCase>& hc = plan.cases;
  const long nodes = plan.nodes, rings = plan.rings, edges = plan.edges, Aoff = plan.A,
             woff = plan.wake, goff = plan.wgam, cb = plan.cb, cw = plan.cw;
  maxN_ = plan.maxN;
  tot_nodes_ = nodes;
  tot_rings_ = rings;
  tot_edges_ = edges;
  tot_wake_ = woff;
  tot_cb_ = cb;
  tot_cw_ = cw;
  tot_te_ = plan.te;
  cases_h_ = hc;

  std::vector<int> node_case(nodes), ring_case(rings), edge_case(edges), wn_case(woff), te_case,
      te_j;
  std::vector<double> cp0(3 * rings), mid0(3 * edges);
  std::vector<long> chb_start(n), chw_start(n);
  for (size_t i = 0; i < n; ++i) {
    const BCase& c = hc[i];
    const int C1 = c.C + 1;
    for (long k = 0; k < (long)(c.R + 1) * C1; ++k) node_case[c.node_off + k] = (int)i;
    for (long k = 0; k < c.N; ++k) ring_case[c.ring_off + k] = (int)i;
    for (long k = 0; k < c.E; ++k) edge_case[c.edge_off + k] = (int)i;
    for (long k = 0; k < (long)cap * C1; ++k) wn_case[c.wake_off + k] = (int)i;
    for (int j = 0; j < C1; ++j) {
      te_case.push_back((int)i);
      te_j.push_back(j);
    }
    chb_start[i] = c.chb_off;
    chw_start[i] = c.chw_off;
    auto P = [&](int a, int b, int d) {
      return wing_init[3 * (c.node_off + (long)a * C1 + b) + d];
    };
    // reference-frame collocation points and unique edge midpoints
    // (simple_simulator.cpp:19-34; edge order of uvlm_sim_loads)
    for (int a = 0; a < c.R; ++a)
      for (int b = 0; b < c.C; ++b)
        for (int d = 0; d < 3; ++d)
          cp0[3 * (c.ring_off + (long)a * c.C + b) + d] =
              (((P(a, b, d) + P(a + 1, b, d)) + P(a + 1, b + 1, d)) + P(a, b + 1, d)) / 4;
    long e = c.edge_off;
    for (int a = 0; a < c.R; ++a)
      for (int b = 0; b < c.C; ++b, ++e)
        for (int d = 0; d < 3; ++d) mid0[3 * e + d] = (P(a, b, d) + P(a, b + 1, d)) / 2;
    for (int a = 0; a < c.R; ++a)
      for (int b = 0; b <= c.C; ++b, ++e)
        for (int d = 0; d < 3; ++d) mid0[3 * e + d] = (P(a, b, d) + P(a + 1, b, d)) / 2;
  }
  // LPT order of the persistent work: most expensive cases first
  auto final_sources = [&](const BCase& c) {
    return 2.0 * (c.R + 1) * (c.C + 1) + 2.0 * (max_steps + 1) * (c.C + 1);
  };
  order_.resize(n);
  std::iota(order_.begin(), order_.end(), 0);
  std::stable_sort(order_.begin(), order_.end(), [&](int a, int b) {
    return final_sources(hc[a]) > final_sources(hc[b]);
  });
  // matrix items (case, target tile, ring tile); LU update items per panel
  std::vector<int4> it_mat;
  for (size_t i = 0; i < n; ++i)
    for (int tt = 0; tt < cdiv(hc[i].N, 256); ++tt)
      for (int rt = 0; rt < cdiv(hc[i].N, kMatRings); ++rt)
        it_mat.push_back(int4{(int)i, tt, rt, 0});
  n_mat_ = (int)it_mat.size();
  std::vector<int4> it_gemm;
  gemm_off_.clear();
  for (int k0 = 0; k0 < maxN_; k0 += kNB) {
    gemm_off_.push_back((int)it_gemm.size());
    for (size_t i = 0; i < n; ++i) {
      const int nn = hc[i].N;
      if (k0 >= nn) continue;
      const int rest = nn - k0 - std::min(kNB, nn - k0);
      for (int a = 0; a < cdiv(std::max(rest, 0), 64); ++a)
        for (int b = 0; b < cdiv(std::max(rest, 0), 64); ++b)
          it_gemm.push_back(int4{(int)i, a, b, 0});
    }
  }
  gemm_off_.push_back((int)it_gemm.size());

  BTRY(dev_alloc(allocs_, &d_cases_, n, err));
  BTRY(dev_alloc(allocs_, &d_state_, 1, err));
  BTRY(dev_alloc(allocs_, &d_node_case_, nodes, err));
  BTRY(dev_alloc(allocs_, &d_ring_case_, rings, err));
  BTRY(dev_alloc(allocs_, &d_edge_case_, edges, err));
  BTRY(dev_alloc(allocs_, &d_wn_case_, woff, err));
  BTRY(dev_alloc(allocs_, &d_te_case_, te_case.size(), err));
  BTRY(dev_alloc(allocs_, &d_te_j_, te_j.size(), err));
  BTRY(dev_alloc(allocs_, &d_chb_start_, n, err));
  BTRY(dev_alloc(allocs_, &d_chw_start_, n, err));
  BTRY(dev_alloc(allocs_, &d_init_, 3 * nodes, err));
  BTRY(dev_alloc(allocs_, &d_wing_, 3 * nodes, err));
  BTRY(dev_alloc(allocs_, &d_cp0_, 3 * rings, err));
  BTRY(dev_alloc(allocs_, &d_mid0_, 3 * edges, err));
  BTRY(dev_alloc(allocs_, &d_vm_, 3 * rings, err));
  BTRY(dev_alloc(allocs_, &d_vme_, 3 * edges, err));
  BTRY(dev_alloc(allocs_, &d_gamma_, rings, err));
  BTRY(dev_alloc(allocs_, &d_gprev_, rings, err));
  BTRY(dev_alloc(allocs_, &d_rhs_, rings, err));
  BTRY(dev_alloc(allocs_, &d_uw_, 3 * rings, err));
  BTRY(dev_alloc(allocs_, &d_cp_, 3 * rings, err));
  BTRY(dev_alloc(allocs_, &d_nrm_, 3 * rings, err));
  BTRY(dev_alloc(allocs_, &d_area_, rings, err));
  BTRY(dev_alloc(allocs_, &d_rings_, rings, err));
  BTRY(dev_alloc(allocs_, &d_mid_, 3 * edges, err));
  BTRY(dev_alloc(allocs_, &d_uflow_, 3 * edges, err));
  BTRY(dev_alloc(allocs_, &d_A_, Aoff, err));
  BTRY(dev_alloc(allocs_, &d_Asave_, Aoff, err));
  BTRY(dev_alloc(allocs_, &d_flag_, n, err));
  BTRY(dev_alloc(allocs_, &d_dblk_, (size_t)n * kNB * kNB, err));
  tot_A_ = Aoff;
  BTRY(dev_alloc(allocs_, &d_piv_, rings, err));
  BTRY(dev_alloc(allocs_, &d_wake_, 3 * woff, err));
  BTRY(dev_alloc(allocs_, &d_wgam_, goff, err));
  BTRY(dev_alloc(allocs_, &d_pos_, 3 * woff, err));
  BTRY(dev_alloc(allocs_, &d_pst_, 3 * woff, err));
  BTRY(dev_alloc(allocs_, &d_k1_, 3 * woff, err));
  BTRY(dev_alloc(allocs_, &d_k2_, 3 * woff, err));
  BTRY(dev_alloc(allocs_, &d_k3_, 3 * woff, err));
  BTRY(dev_alloc(allocs_, &d_bch_, cb, err));
  BTRY(dev_alloc(allocs_, &d_wch_, cw, err));
  BTRY(dev_alloc(allocs_, &d_hist_, 2 * (long)n * (max_steps + 2), err));
  BTRY(dev_alloc(allocs_, &d_it_mat_, it_mat.size(), err));
  BTRY(dev_alloc(allocs_, &d_it_gemm_, it_gemm.size(), err));

  auto up = [&](void* dst, const void* src, size_t bytes) -> int {
    if (bytes) BHIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_));
    return UVLM_OK;
  };
  BTRY(up(d_cases_, hc.data(), sizeof(BCase) * n));
  BState zero{};
  BTRY(up(d_state_, &zero, sizeof(BState)));
  BTRY(up(d_node_case_, node_case.data(), sizeof(int) * nodes));
  BTRY(up(d_ring_case_, ring_case.data(), sizeof(int) * rings));
  BTRY(up(d_edge_case_, edge_case.data(), sizeof(int) * edges));
  BTRY(up(d_wn_case_, wn_case.data(), sizeof(int) * woff));
  BTRY(up(d_te_case_, te_case.data(), sizeof(int) * te_case.size()));
  BTRY(up(d_te_j_, te_j.data(), sizeof(int) * te_j.size()));
  BTRY(up(d_chb_start_, chb_start.data(), sizeof(long) * n));
  BTRY(up(d_chw_start_, chw_start.data(), sizeof(long) * n));
  BTRY(up(d_init_, wing_init, sizeof(double) * 3 * nodes));
  BTRY(up(d_cp0_, cp0.data(), sizeof(double) * 3 * rings));
  BTRY(up(d_mid0_, mid0.data(), sizeof(double) * 3 * edges));
  BTRY(up(d_it_mat_, it_mat.data(), sizeof(int4) * it_mat.size()));
  BTRY(up(d_it_gemm_, it_gemm.data(), sizeof(int4) * it_gemm.size()));
  BHIP(hipMemsetAsync(d_gamma_, 0, sizeof(double) * rings, stream_));
  BHIP(hipMemsetAsync(d_gprev_, 0, sizeof(double) * rings, stream_));
  BHIP(hipMemsetAsync(d_hist_, 0, sizeof(double) * 2 * n * (max_steps + 2), stream_));
  BHIP(hipStreamSynchronize(stream_));
  int dev = 0, cus = 0;
  BHIP(hipGetDevice(&dev));
  BHIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  cus_ = std::max(cus, 1);
  variant_ = 1;
  BTRY(build_items(kVariantBlock[variant_], err));
  return UVLM_OK;
}

// N-body work items (case, first target) in units of `tile` targets.
int BatchEngine::build_items(int tile, std::string* err) {
  std::vector<int2> rhs, loads, adv;
  for (int i : order_) {
    const BCase& c = cases_h_[i];
    for (int t = 0; t < c.N; t += tile) rhs.push_back(int2{i, t});
    for (int t = 0; t < c.E; t += tile) loads.push_back(int2{i, t});
  }
  // advection targets grow with the step: tile-major order puts the live
  // tiles of every step first
  long max_tiles = 0;
  for (const auto& c : cases_h_) max_tiles = std::max<long>(max_tiles, cdiv((long)c.cap * (c.C + 1), tile));
  for (long t = 0; t < max_tiles; ++t)
    for (int i : order_)
      if (t * tile < (long)cases_h_[i].cap * (cases_h_[i].C + 1))
        adv.push_back(int2{i, (int)(t * tile)});
  auto put = [&](int2** d, size_t* cap, const std::vector<int2>& h) -> int {
    if (h.size() > *cap) {
      BTRY(dev_alloc(allocs_, d, h.size(), err));
      *cap = h.size();
    }
    if (!h.empty())
      BHIP(hipMemcpyAsync(*d, h.data(), sizeof(int2) * h.size(), hipMemcpyHostToDevice, stream_));
    return UVLM_OK;
  };
  BTRY(put(&d_it_rhs_, &cap_rhs_, rhs));
  BTRY(put(&d_it_loads_, &cap_loads_, loads));
  BTRY(put(&d_it_adv_, &cap_adv_, adv));
  BHIP(hipStreamSynchronize(stream_));
  n_rhs_ = (int)rhs.size();
  n_loads_ = (int)loads.size();
  n_adv_ = (int)adv.size();
  items_tile_ = tile;
  return UVLM_OK;
}

int BatchEngine::nbody(int phase, const double* targets, double* out, int slot) {
  const int2* items = phase == 0 ? d_it_rhs_ : phase == 1 ? d_it_loads_ : d_it_adv_;
  const int n = phase == 0 ? n_rhs_ : phase == 1 ? n_loads_ : n_adv_;
  if (n == 0) return UVLM_OK;
  const int per_cu = kVariantBlock[variant_] >= 512 ? 2 : kVariantBlock[variant_] >= 256 ? 4 : 8;
  const int grid = std::min(n, cus_ * per_cu);
  if (phase == 0)
    launch_variant<0>(variant_, stream_, grid, d_cases_, d_state_, items, n, slot, targets,
                      d_bch_, d_wch_, out);
  else if (phase == 1)
    launch_variant<1>(variant_, stream_, grid, d_cases_, d_state_, items, n, slot, targets,
                      d_bch_, d_wch_, out);
  else
    launch_variant<2>(variant_, stream_, grid, d_cases_, d_state_, items, n, slot, targets,
                      d_bch_, d_wch_, out);
  return UVLM_OK;
}

// One timestep of every case (simple_simulator.cpp:240-286 per case).
int BatchEngine::step_once(std::string* err) {
  hipStream_t s = stream_;
  const dim3 b256(256);
  auto g = [](long n) { return dim3(std::max(1, cdiv(n, 256))); };
  hipLaunchKernelGGL(k_b_begin, dim3(1), dim3(64), 0, s, d_state_);
  hipLaunchKernelGGL(k_b_copy, g(tot_rings_), b256, 0, s, d_gamma_, d_gprev_, tot_rings_);
  hipLaunchKernelGGL(k_b_morph, g(tot_nodes_), b256, 0, s, d_cases_, d_state_, d_node_case_,
                     d_init_, d_wing_, tot_nodes_);
  hipLaunchKernelGGL(k_b_vmorph, g(tot_rings_), b256, 0, s, d_cases_, d_state_, d_ring_case_,
                     d_cp0_, d_vm_, tot_rings_);
  hipLaunchKernelGGL(k_b_vmorph, g(tot_edges_), b256, 0, s, d_cases_, d_state_, d_edge_case_,
                     d_mid0_, d_vme_, tot_edges_);
  hipLaunchKernelGGL(k_b_shed, g(tot_te_), b256, 0, s, d_cases_, d_state_, d_te_case_, d_te_j_,
                     d_wing_, d_gamma_, d_wake_, d_wgam_, (int)tot_te_);
  hipLaunchKernelGGL(k_b_geometry, g(tot_rings_), b256, 0, s, d_cases_, d_ring_case_, d_wing_,
                     d_cp_, d_nrm_, d_area_, d_rings_, tot_rings_);
  hipLaunchKernelGGL(k_b_matrix, dim3(n_mat_), b256, 0, s, d_cases_, d_it_mat_, d_rings_, d_cp_,
                     d_nrm_, d_A_);
  hipLaunchKernelGGL((k_b_pack_chains<true>), g(tot_cw_), b256, 0, s, d_cases_, d_state_,
                     d_chw_start_, n_, d_wake_, 0, d_wgam_, d_wch_, tot_cw_);
  BTRY(nbody(0, d_cp_, d_uw_, 0));
  hipLaunchKernelGGL(k_b_rhs, g(tot_rings_), b256, 0, s, d_cases_, d_ring_case_, d_uw_, d_vm_,
                     d_nrm_, d_rhs_, tot_rings_);
  // batched LU: checked no-pivot panels for every case; cases whose checks
  // failed are restored and refactored with partial pivoting (those
  // launches return at once for the other cases)
  hipLaunchKernelGGL(k_b_init_flags, g(n_), b256, 0, s, d_flag_, n_, force_pivoted_ ? 1 : 0);
  hipLaunchKernelGGL(k_b_copy, g(tot_A_), b256, 0, s, d_A_, d_Asave_, tot_A_);
  for (size_t p = 0; p + 1 < gemm_off_.size(); ++p) {
    const int k0 = (int)p * kNB;
    hipLaunchKernelGGL(k_b_panel_np, dim3(std::max(1, cdiv(maxN_ - k0, 64)), n_), dim3(64), 0, s,
                       d_cases_, k0, d_A_, d_piv_, d_flag_, d_dblk_);
    hipLaunchKernelGGL(k_b_trsm, dim3(cdiv(maxN_, 64), n_), dim3(64), 0, s, d_cases_, k0, d_A_,
                       (const int*)nullptr, (const double*)d_dblk_);
    const int ng = gemm_off_[p + 1] - gemm_off_[p];
    if (ng > 0)
      hipLaunchKernelGGL(k_b_gemm, dim3(ng), b256, 0, s, d_cases_, d_it_gemm_ + gemm_off_[p], k0,
                         d_A_, (const int*)nullptr);
  }
  hipLaunchKernelGGL(k_b_restore, dim3(64, n_), b256, 0, s, d_cases_, d_flag_, d_Asave_, d_A_);
  for (size_t p = 0; p + 1 < gemm_off_.size(); ++p) {
    const int k0 = (int)p * kNB;
    hipLaunchKernelGGL(k_b_getf2, dim3(n_), dim3(1024), 0, s, d_cases_, k0, d_A_, d_piv_,
                       (const int*)d_flag_);
    hipLaunchKernelGGL(k_b_laswp, dim3(cdiv(maxN_, 256), n_), b256, 0, s, d_cases_, k0, d_A_,
                       d_piv_, (const int*)d_flag_);
    hipLaunchKernelGGL(k_b_trsm, dim3(cdiv(maxN_, 64), n_), dim3(64), 0, s, d_cases_, k0, d_A_,
                       (const int*)d_flag_, (const double*)nullptr);
    const int ng = gemm_off_[p + 1] - gemm_off_[p];
    if (ng > 0)
      hipLaunchKernelGGL(k_b_gemm, dim3(ng), b256, 0, s, d_cases_, d_it_gemm_ + gemm_off_[p], k0,
                         d_A_, (const int*)d_flag_);
  }
  const int lds_n = maxN_ <= 4096 ? maxN_ : 0;
  hipLaunchKernelGGL(k_b_getrs, dim3(n_), dim3(1024), sizeof(double) * lds_n, s, d_cases_, d_A_,
                     d_piv_, d_rhs_, d_gamma_, lds_n);
  hipLaunchKernelGGL((k_b_pack_chains<false>), g(tot_cb_), b256, 0, s, d_cases_, d_state_,
                     d_chb_start_, n_, d_wing_, 0, d_gamma_, d_bch_, tot_cb_);
  // loads
  hipLaunchKernelGGL(k_b_edges, g(tot_edges_), b256, 0, s, d_cases_, d_edge_case_, d_wing_, d_mid_,
                     tot_edges_);
  BTRY(nbody(1, d_mid_, d_uflow_, 1));
  hipLaunchKernelGGL(k_b_loads, dim3(n_), b256, 0, s, d_cases_, d_state_, d_wing_, d_gamma_,
                     d_gprev_, d_nrm_, d_area_, d_uflow_, d_vme_, n_, d_hist_);
  // advection
  hipLaunchKernelGGL(k_b_gather, g(tot_wake_), b256, 0, s, d_cases_, d_state_, d_wn_case_,
                     d_wake_, d_pos_, tot_wake_);
  BTRY(nbody(2, d_pos_, d_k1_, 2));
  if (scheme_ == UVLM_ADVECT_RK2 || scheme_ == UVLM_ADVECT_RK4) {
    const double div = scheme_ == UVLM_ADVECT_RK2 ? 1.0 : 2.0;
    hipLaunchKernelGGL(k_b_stage, g(tot_wake_), b256, 0, s, d_cases_, d_state_, d_wn_case_, d_pos_,
                       d_k1_, div, d_pst_, tot_wake_);
    hipLaunchKernelGGL((k_b_pack_chains<true>), g(tot_cw_), b256, 0, s, d_cases_, d_state_,
                       d_chw_start_, n_, d_pst_, 1, d_wgam_, d_wch_, tot_cw_);
    BTRY(nbody(2, d_pst_, d_k2_, 3));
    if (scheme_ == UVLM_ADVECT_RK4) {
      hipLaunchKernelGGL(k_b_stage, g(tot_wake_), b256, 0, s, d_cases_, d_state_, d_wn_case_,
                         d_pos_, d_k2_, 2.0, d_pst_, tot_wake_);
      hipLaunchKernelGGL((k_b_pack_chains<true>), g(tot_cw_), b256, 0, s, d_cases_, d_state_,
                         d_chw_start_, n_, d_pst_, 1, d_wgam_, d_wch_, tot_cw_);
      BTRY(nbody(2, d_pst_, d_k3_, 4));
    }
  }
  hipLaunchKernelGGL(k_b_update, g(tot_wake_), b256, 0, s, d_cases_, d_state_, d_wn_case_, scheme_,
                     ign_, d_pos_, d_k1_, d_k2_, d_k3_, d_wake_, tot_wake_);
  hipLaunchKernelGGL(k_b_end, dim3(1), dim3(64), 0, s, d_state_);
  BHIP(hipGetLastError());
  return UVLM_OK;
}

void BatchEngine::set_force_pivoted(bool on) {
  if (on != force_pivoted_ && graph_exec_) {
    (void)hipGraphExecDestroy(graph_exec_);
    (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
  }
  force_pivoted_ = on;
}

int BatchEngine::capture(std::string* err) {
  if (graph_exec_) return UVLM_OK;
  BHIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  std::string e2;
  const int rc = step_once(&e2);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(stream_, &g);
  if (rc != UVLM_OK) {
    if (g) (void)hipGraphDestroy(g);
    *err = e2;
    return rc;
  }
  BHIP(ec);
  graph_ = g;
  BHIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
  return UVLM_OK;
}

int BatchEngine::run(size_t steps, bool use_graph, std::string* err) {
  if ((size_t)steps_done_ + steps > (size_t)max_steps_ + 1) {
    *err = "batch capacity of " + std::to_string(max_steps_ + 1) + " steps exceeded";
    return UVLM_ERR_CAPACITY;
  }
  if (use_graph) BTRY(capture(err));
  // One step at a time: a step enqueues several hundred launches on the
  // eager path (the batched LU alone is 7 per panel), and an unbounded
  // backlog of them (4,382 packets were outstanding when a traced eager
  // run lost a dispatch packet) is not something to rely on.
  for (size_t k = 0; k < steps; ++k) {
    if (use_graph)
      BHIP(hipGraphLaunch(graph_exec_, stream_));
    else
      BTRY(step_once(err));
    BHIP(hipStreamSynchronize(stream_));
    ++steps_done_;
  }
  return UVLM_OK;
}

int BatchEngine::autotune(std::string* err, int* chosen, double* times_ms, int max_variants) {
  // Time the advection N-body of the current state for every launch shape;
  // the state is left unchanged (the launches only write the stage buffer,
  // which every step recomputes before use).
  BHIP(hipStreamSynchronize(stream_));
  hipEvent_t e0, e1;
  BHIP(hipEventCreate(&e0));
  BHIP(hipEventCreate(&e1));
  double best = 1e300;
  int best_v = variant_;
  for (int v = 0; v < kVariants; ++v) {
    variant_ = v;
    BTRY(build_items(kVariantBlock[v], err));
    float ms_total = 0.f;
    for (int rep = 0; rep < 3; ++rep) {
      hipLaunchKernelGGL(k_b_begin, dim3(1), dim3(64), 0, stream_, d_state_);
      BHIP(hipEventRecord(e0, stream_));
      BTRY(nbody(2, d_pos_, d_k2_, 2));
      BHIP(hipEventRecord(e1, stream_));
      BHIP(hipEventSynchronize(e1));
      float ms = 0.f;
      BHIP(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0) ms_total += ms;
    }
    const double t = ms_total / 2.0;
    if (v < max_variants && times_ms) times_ms[v] = t;
    if (t < best) {
      best = t;
      best_v = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  variant_ = best_v;
  BTRY(build_items(kVariantBlock[best_v], err));
  if (graph_exec_) {
    (void)hipGraphExecDestroy(graph_exec_);
    (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
  }
  if (chosen) *chosen = best_v;
  return UVLM_OK;
}

int BatchEngine::get_gamma(int c, double* out, std::string* err) {
  const BCase& cs = cases_h_[c];
  BHIP(hipStreamSynchronize(stream_));
  BHIP(hipMemcpy(out, d_gamma_ + cs.ring_off, sizeof(double) * cs.N, hipMemcpyDeviceToHost));
  return UVLM_OK;
}

int BatchEngine::get_wing(int c, double* out, std::string* err) {
  const BCase& cs = cases_h_[c];
  BHIP(hipStreamSynchronize(stream_));
  BHIP(hipMemcpy(out, d_wing_ + 3 * cs.node_off,
                 sizeof(double) * 3 * (cs.R + 1) * (cs.C + 1), hipMemcpyDeviceToHost));
  return UVLM_OK;
}

// nodes [s+1][C+1][3] and gamma [s][C] after steps_done_ = s+1 steps
int BatchEngine::get_wake(int c, double* nodes, double* gamma, std::string* err) {
  const BCase& cs = cases_h_[c];
  BHIP(hipStreamSynchronize(stream_));
  const int s = steps_done_ - 1;
  if (s < 0) return UVLM_OK;
  const int C1 = cs.C + 1;
  if (nodes)
    BHIP(hipMemcpy(nodes, d_wake_ + 3 * (cs.wake_off + b_node_row0(cs, s) * C1),
                   sizeof(double) * 3 * (s + 1) * C1, hipMemcpyDeviceToHost));
  if (gamma && s > 0)
    BHIP(hipMemcpy(gamma, d_wgam_ + cs.wgam_off + b_gam_row0(cs, s) * cs.C,
                   sizeof(double) * s * cs.C, hipMemcpyDeviceToHost));
  return UVLM_OK;
}

int BatchEngine::get_loads(int c, double* out, size_t* count, std::string* err) {
  BHIP(hipStreamSynchronize(stream_));
  const size_t n = std::min(*count, (size_t)steps_done_);
  std::vector<double> h(2 * (size_t)n_ * std::max<size_t>(n, 1));
  if (n)
    BHIP(hipMemcpy(h.data(), d_hist_, sizeof(double) * 2 * n_ * n, hipMemcpyDeviceToHost));
  for (size_t k = 0; k < n && out; ++k) {
    out[2 * k + 0] = h[b_hist_index((int)k, n_, c) + 0];
    out[2 * k + 1] = h[b_hist_index((int)k, n_, c) + 1];
  }
  *count = (size_t)steps_done_;
  return UVLM_OK;
}

// Segment evaluations of step s: RHS N x wake, loads E x (bound + wake),
// advection stages V x (bound + wake).
double BatchEngine::pairs_per_step_at(int s) const {
  const int stages = scheme_ == UVLM_ADVECT_RK2 ? 2 : scheme_ == UVLM_ADVECT_RK4 ? 3 : 1;
  double tot = 0.0;
  for (const auto& c : cases_h_) {
    const double sb = chain_segments(1, c.R, c.C), sw = chain_segments(1, s, c.C);
    const double V = (double)(s + 1) * (c.C + 1);
    tot += c.N * sw + c.E * (sb + sw) + stages * V * (sb + sw);
  }
  return tot;
}

}  // namespace uvlm

