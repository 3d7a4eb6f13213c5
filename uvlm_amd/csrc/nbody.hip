// nbody.hip -- Biot-Savart N-body sums over vortex rings on gfx950.
//
// One target point per lane; the block stages 64 source rings (128-byte
// records) at a time in LDS and every lane of every wavefront reads the same
// record (LDS broadcast), so HBM traffic per pair is ~0 and the kernel is
// bound by the FP64 VALU.  Sources are split into chunks along blockIdx.y
// when there are too few targets to fill the 256 CUs; each chunk's partial
// sum is written as SoA and the chunks are combined in chunk order.
//
// F64_EXACT evaluates BiotSavartLaw exactly as the reference does
// (vortex.cpp:21-39): no FMA contraction (the file is compiled with
// -ffp-contract=off), correctly rounded division and square root, and the
// reference's per-ring / per-target summation order (vortex.cpp:145-160,
// velocity.h:14-24).  With one chunk per range the sums are bitwise those of
// the reference's loop order.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace uvlm {
namespace {

constexpr int kBlock = 256;
constexpr int kTile = 64;

// ---------------------------------------------------------------------------
// vortex.cpp:21-39, bit-faithful.  td replaces `d_len * d_len < eps` by the
// equivalent `dsq < td` (see cutoff_dsq_threshold), saving one sqrt.
__device__ __forceinline__ void seg_exact(double ax, double ay, double az,
                                          double bx, double by, double bz,
                                          double x, double y, double z,
                                          const CutOff& k, double& ox, double& oy,
                                          double& oz) {
  const double r0x = bx - ax, r0y = by - ay, r0z = bz - az;
  const double r1x = x - ax, r1y = y - ay, r1z = z - az;
  const double r2x = x - bx, r2y = y - by, r2z = z - bz;
  const double dx = r1y * r2z - r1z * r2y;
  const double dy = r1z * r2x - r1x * r2z;
  const double dz = r1x * r2y - r1y * r2x;
  const double dsq = (dx * dx + dy * dy) + dz * dz;
  const double r1n = sqrt((r1x * r1x + r1y * r1y) + r1z * r1z);
  const double r2n = sqrt((r2x * r2x + r2y * r2y) + r2z * r2z);
  const bool cut = (dsq < k.td) | (r1n < k.eps) | (r2n < k.eps);
  const double wx = r1x / r1n - r2x / r2n;
  const double wy = r1y / r1n - r2y / r2n;
  const double wz = r1z / r1n - r2z / r2n;
  const double K = kInv4Pi / dsq * ((r0x * wx + r0y * wy) + r0z * wz);
  ox = cut ? 0.0 : dx * K;
  oy = cut ? 0.0 : dy * K;
  oz = cut ? 0.0 : dz * K;
}

// Refined hardware reciprocal square root / reciprocal (two Newton steps).
__device__ __forceinline__ double rsq64(double a) {
  double y = __builtin_amdgcn_rsq(a);
  double h = a * y;
  double e = __builtin_fma(-h, y, 1.0);
  y = __builtin_fma(0.5 * y, e, y);
  h = a * y;
  e = __builtin_fma(-h, y, 1.0);
  return __builtin_fma(0.5 * y, e, y);
}
__device__ __forceinline__ double rcp64(double a) {
  double r = __builtin_amdgcn_rcp(a);
  double e = __builtin_fma(-a, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-a, r, 1.0);
  return __builtin_fma(r, e, r);
}

// FP64 with FMA: u = d * [(r0.r1)/|r1| - (r0.r2)/|r2|] / (4 pi |d|^2) * g,
// d = r0 x r1 (= r1 x r2), same cut-off set up to rounding.
__device__ __forceinline__ void seg_fast(double ax, double ay, double az,
                                         double bx, double by, double bz,
                                         double x, double y, double z,
                                         double g, const CutOff& k, double& ux,
                                         double& uy, double& uz) {
  const double r0x = bx - ax, r0y = by - ay, r0z = bz - az;
  const double r1x = x - ax, r1y = y - ay, r1z = z - az;
  const double r2x = x - bx, r2y = y - by, r2z = z - bz;
  const double dx = __builtin_fma(r1y, r2z, -r1z * r2y);
  const double dy = __builtin_fma(r1z, r2x, -r1x * r2z);
  const double dz = __builtin_fma(r1x, r2y, -r1y * r2x);
  const double dsq = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
  const double r1s = __builtin_fma(r1z, r1z, __builtin_fma(r1y, r1y, r1x * r1x));
  const double r2s = __builtin_fma(r2z, r2z, __builtin_fma(r2y, r2y, r2x * r2x));
  const double o1 = __builtin_fma(r0z, r1z, __builtin_fma(r0y, r1y, r0x * r1x));
  const double o2 = __builtin_fma(r0z, r2z, __builtin_fma(r0y, r2y, r0x * r2x));
  const bool cut = (dsq < k.eps) | (r1s < k.eps2) | (r2s < k.eps2);
  const double t = __builtin_fma(o1, rsq64(r1s), -o2 * rsq64(r2s));
  double K = (kInv4Pi * g) * t * rcp64(dsq);
  K = cut ? 0.0 : K;
  ux = __builtin_fma(dx, K, ux);
  uy = __builtin_fma(dy, K, uy);
  uz = __builtin_fma(dz, K, uz);
}

// ---------------------------------------------------------------------------
// Ring sums in the reference's order.
template <int SEG>
__device__ __forceinline__ void ring_exact(const double* r, double x, double y,
                                           double z, const CutOff& cut, double& ax,
                                           double& ay, double& az) {
  const double g = r[12];
  double sx, sy, sz;
  double ux = 0.0, uy = 0.0, uz = 0.0;
  if (SEG == kSegVoring) {
    // VORING: u += VORTEX(x, p_k, p_k+1, g) for k = 0..3
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int a = 3 * k, b = 3 * ((k + 1) & 3);
      seg_exact(r[a], r[a + 1], r[a + 2], r[b], r[b + 1], r[b + 2], x, y, z, cut,
                sx, sy, sz);
      ux = ux + sx * g;
      uy = uy + sy * g;
      uz = uz + sz * g;
    }
  } else {
    // VortexRing::ForEachSegment: start node[(k+1)%4], end node[k]
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (SEG == kSegRingChord && (k & 1)) continue;
      const int a = 3 * ((k + 1) & 3), b = 3 * k;
      seg_exact(r[a], r[a + 1], r[a + 2], r[b], r[b + 1], r[b + 2], x, y, z, cut,
                sx, sy, sz);
      if (SEG == kSegRingPerSeg) {
        ux = ux + sx * g;
        uy = uy + sy * g;
        uz = uz + sz * g;
      } else {
        ux = ux + sx;
        uy = uy + sy;
        uz = uz + sz;
      }
    }
    if (SEG != kSegRingPerSeg) {
      ux = ux * g;
      uy = uy * g;
      uz = uz * g;
    }
  }
  ax = ax + ux;
  ay = ay + uy;
  az = az + uz;
}

template <int SEG>
__device__ __forceinline__ void ring_fast(const double* r, double x, double y,
                                          double z, const CutOff& cut, double& ax,
                                          double& ay, double& az) {
  const double g = r[12];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (SEG == kSegRingChord && (k & 1)) continue;
    int a, b;
    if (SEG == kSegVoring) {
      a = 3 * k;
      b = 3 * ((k + 1) & 3);
    } else {
      a = 3 * ((k + 1) & 3);
      b = 3 * k;
    }
    seg_fast(r[a], r[a + 1], r[a + 2], r[b], r[b + 1], r[b + 2], x, y, z, g, cut, ax,
             ay, az);
  }
}

__device__ __forceinline__ void chunk_range(const ChunkPlan& p, int c, int& b,
                                            int& e) {
  if (c < p.cb) {
    b = c * p.szb;
    e = min(b + p.szb, p.nb);
  } else {
    b = p.nb + (c - p.cb) * p.szw;
    e = min(b + p.szw, p.nb + p.nw);
  }
}

template <int PREC, int SEG>
__global__ __launch_bounds__(kBlock) void k_nbody(
    const double* __restrict__ targets, int nt, const RingRec* __restrict__ recs,
    ChunkPlan plan, double* __restrict__ partial, int ld, CutOff cut) {
  __shared__ double2 tile[kTile * 8];
  int rb, re;
  chunk_range(plan, blockIdx.y, rb, re);
  const int t = blockIdx.x * kBlock + threadIdx.x;
  double x = 0.0, y = 0.0, z = 0.0;
  if (t < nt) {
    x = targets[3 * t + 0];
    y = targets[3 * t + 1];
    z = targets[3 * t + 2];
  }
  double ax = 0.0, ay = 0.0, az = 0.0;
  const double2* src = reinterpret_cast<const double2*>(recs);
  for (int base = rb; base < re; base += kTile) {
    const int cnt = min(kTile, re - base);
    __syncthreads();
    for (int q = threadIdx.x; q < cnt * 8; q += kBlock) tile[q] = src[base * 8 + q];
    __syncthreads();
    const double* tl = reinterpret_cast<const double*>(tile);
    for (int k = 0; k < cnt; ++k) {
      if (PREC == kF64Exact)
        ring_exact<SEG>(tl + 16 * k, x, y, z, cut, ax, ay, az);
      else
        ring_fast<SEG>(tl + 16 * k, x, y, z, cut, ax, ay, az);
    }
  }
  if (t < nt) {
    const size_t c = blockIdx.y;
    partial[(c * 3 + 0) * (size_t)ld + t] = ax;
    partial[(c * 3 + 1) * (size_t)ld + t] = ay;
    partial[(c * 3 + 2) * (size_t)ld + t] = az;
  }
}

// ---------------------------------------------------------------------------
__global__ void k_pack_lattice(const double* __restrict__ nodes, int sheet_rows,
                               int row_off, const double* __restrict__ gamma,
                               int g_sheet_rows, int g_row_off, int num,
                               int rows, int cols, RingRec* __restrict__ out) {
  const int K = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = num * rows * cols;
  if (K >= total) return;
  const int j = K % cols;
  const int i = (K / cols) % rows;
  const int n = K / (cols * rows);
  const int C1 = cols + 1;
  const double* base = nodes + (size_t)3 * ((size_t)(n * sheet_rows + row_off + i) * C1);
  const double* nxt = base + 3 * C1;
  RingRec r;
  // p0=(i,j) p1=(i,j+1) p2=(i+1,j+1) p3=(i+1,j)
  for (int d = 0; d < 3; ++d) {
    r.p[0 + d] = base[3 * j + d];
    r.p[3 + d] = base[3 * (j + 1) + d];
    r.p[6 + d] = nxt[3 * (j + 1) + d];
    r.p[9 + d] = nxt[3 * j + d];
  }
  r.g = gamma ? gamma[(size_t)(n * g_sheet_rows + g_row_off + i) * cols + j] : 1.0;
  r.pad[0] = r.pad[1] = r.pad[2] = 0.0;
  out[K] = r;
}

__global__ void k_pack_rings(const double* __restrict__ rings,
                             const double* __restrict__ gamma, int n,
                             RingRec* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  RingRec r;
  for (int d = 0; d < 12; ++d) r.p[d] = rings[12 * (size_t)k + d];
  r.g = gamma[k];
  r.pad[0] = r.pad[1] = r.pad[2] = 0.0;
  out[k] = r;
}

__global__ void k_set_gamma(const double* __restrict__ gamma, int n,
                            RingRec* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k].g = gamma[k];
}

// Chunk partial sums [chunk][3][ld] -> per-target sums.  Few chunks: one
// lane per target, chunks in order.  Many chunks: 64 targets x 16 chunk
// groups per workgroup (coalesced along targets), the 16 group sums added in
// LDS in group order.
constexpr int kSumGroups = 16;
constexpr int kWideChunks = 24;

__device__ __forceinline__ double chunk_range_sum(const double* __restrict__ partial, int ld,
                                                  int t, int d, int c0, int c1, int step) {
  double s = 0.0;
  for (int c = c0; c < c1; c += step) s = s + partial[((size_t)c * 3 + d) * ld + t];
  return s;
}

__global__ void k_sum_chunks(const double* __restrict__ partial, int ld, int nt,
                             int chunks, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  for (int d = 0; d < 3; ++d) {
    double s = chunks > 0 ? partial[(size_t)d * ld + t] : 0.0;
    for (int c = 1; c < chunks; ++c) s = s + partial[((size_t)c * 3 + d) * ld + t];
    out[3 * (size_t)t + d] = s;
  }
}

// sums[t][3] of chunks [c0, c1) with the 64 x kSumGroups layout
__device__ void wide_chunk_sums(const double* __restrict__ partial, int ld, int nt, int c0,
                                int c1, double (*red)[kSumGroups][64], double out3[3]) {
  const int tl = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int t = blockIdx.x * 64 + tl;
  for (int d = 0; d < 3; ++d)
    red[d][g][tl] = t < nt ? chunk_range_sum(partial, ld, t, d, c0 + g, c1, kSumGroups) : 0.0;
  __syncthreads();
  if (g == 0)
    for (int d = 0; d < 3; ++d) {
      double s = red[d][0][tl];
      for (int q = 1; q < kSumGroups; ++q) s = s + red[d][q][tl];
      out3[d] = s;
    }
  __syncthreads();
}

__global__ __launch_bounds__(64 * kSumGroups) void k_sum_chunks_wide(
    const double* __restrict__ partial, int ld, int nt, int chunks, double* __restrict__ out) {
  __shared__ double red[3][kSumGroups][64];
  double s[3];
  wide_chunk_sums(partial, ld, nt, 0, chunks, red, s);
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (threadIdx.x < 64 && t < nt)
    for (int d = 0; d < 3; ++d) out[3 * (size_t)t + d] = s[d];
}

// advect.h:16-24: ((-ff + bound) + wake)
__global__ void k_flow_velocity(const double* __restrict__ partial, int ld,
                                int nt, int cb, int cw, double fx, double fy,
                                double fz, double* __restrict__ out,
                                int te_period, int te_width) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const double mff[3] = {-fx, -fy, -fz};
  if (te_period > 0 && (t % te_period) < te_width) {
    for (int d = 0; d < 3; ++d) out[3 * (size_t)t + d] = mff[d];
    return;
  }
  for (int d = 0; d < 3; ++d) {
    double b = 0.0, w = 0.0;
    if (cb > 0) {
      b = partial[(size_t)d * ld + t];
      for (int c = 1; c < cb; ++c) b = b + partial[((size_t)c * 3 + d) * ld + t];
    }
    if (cw > 0) {
      w = partial[((size_t)cb * 3 + d) * ld + t];
      for (int c = cb + 1; c < cb + cw; ++c)
        w = w + partial[((size_t)c * 3 + d) * ld + t];
    }
    out[3 * (size_t)t + d] = (mff[d] + b) + w;
  }
}

__global__ __launch_bounds__(64 * kSumGroups) void k_flow_velocity_wide(
    const double* __restrict__ partial, int ld, int nt, int cb, int cw, double fx, double fy,
    double fz, double* __restrict__ out, int te_period, int te_width) {
  __shared__ double red[3][kSumGroups][64];
  double b[3] = {0.0, 0.0, 0.0}, w[3] = {0.0, 0.0, 0.0};
  if (cb > 0) wide_chunk_sums(partial, ld, nt, 0, cb, red, b);
  if (cw > 0) wide_chunk_sums(partial, ld, nt, cb, cb + cw, red, w);
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (threadIdx.x >= 64 || t >= nt) return;
  const double mff[3] = {-fx, -fy, -fz};
  const bool te = te_period > 0 && (t % te_period) < te_width;
  for (int d = 0; d < 3; ++d) out[3 * (size_t)t + d] = te ? mff[d] : (mff[d] + b[d]) + w[d];
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

template <int PREC>
__global__ void k_segments(const double* __restrict__ a, const double* __restrict__ b,
                           const double* __restrict__ x, int n, CutOff cut,
                           double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double* A = a + 3 * (size_t)k;
  const double* B = b + 3 * (size_t)k;
  const double* X = x + 3 * (size_t)k;
  double ox = 0.0, oy = 0.0, oz = 0.0;
  if (PREC == kF64Exact)
    seg_exact(A[0], A[1], A[2], B[0], B[1], B[2], X[0], X[1], X[2], cut, ox, oy, oz);
  else
    seg_fast(A[0], A[1], A[2], B[0], B[1], B[2], X[0], X[1], X[2], 1.0, cut, ox, oy, oz);
  out[3 * (size_t)k + 0] = ox;
  out[3 * (size_t)k + 1] = oy;
  out[3 * (size_t)k + 2] = oz;
}

}  // namespace

// ---------------------------------------------------------------------------
double cutoff_dsq_threshold(double eps) {
  // f(x) = fl(fl(sqrt(x))^2) < eps is monotone in x >= 0; find the
  // smallest x with f(x) false by bisection on the IEEE bit patterns.
  auto f = [eps](double x) {
    volatile double s = std::sqrt(x);
    volatile double p = s * s;
    return p < eps;
  };
  uint64_t lo, hi;
  double dlo = 0.0, dhi = 2.0 * eps;
  std::memcpy(&lo, &dlo, 8);
  std::memcpy(&hi, &dhi, 8);
  while (hi - lo > 1) {
    uint64_t mid = lo + (hi - lo) / 2;
    double dm;
    std::memcpy(&dm, &mid, 8);
    if (f(dm))
      lo = mid;
    else
      hi = mid;
  }
  double T;
  std::memcpy(&T, &hi, 8);
  return T;
}

CutOff make_cutoff(double eps) {
  CutOff c;
  c.eps = eps;
  c.td = cutoff_dsq_threshold(eps);
  c.eps2 = eps * eps;
  c.short2 = 0.1 / eps;
  return c;
}

ChunkPlan plan_chunks(int nt, int nb, int nw, bool sequential, int slots) {
  ChunkPlan p{};
  p.nb = nb;
  p.nw = nw;
  if (sequential) {
    p.cb = nb > 0 ? 1 : 0;
    p.cw = nw > 0 ? 1 : 0;
  } else {
    const int tblocks = std::max(1, cdiv(nt, kBlock));
    int want;
    if (slots > 0)  // whole rounds of resident workgroups, at most 3
      want = std::max(1, 3 * slots / tblocks);
    else
      want = std::max(1, cdiv(4096, tblocks));
    const int total = nb + nw;
    want = std::min(want, std::max(1, cdiv(total, 2 * kTile)));
    want = std::min(want, 1024);
    if (nb > 0 && nw > 0) {
      int cb = (int)std::lround((double)want * nb / total);
      cb = std::max(1, std::min(cb, want - 1));
      p.cb = cb;
      p.cw = std::max(1, want - cb);
    } else {
      p.cb = nb > 0 ? want : 0;
      p.cw = nw > 0 ? want : 0;
    }
  }
  p.szb = p.cb ? cdiv(nb, p.cb) : 0;
  p.szw = p.cw ? cdiv(nw, p.cw) : 0;
  // drop empty trailing chunks created by rounding
  if (p.cb) p.cb = cdiv(nb, p.szb);
  if (p.cw) p.cw = cdiv(nw, p.szw);
  return p;
}

void launch_pack_lattice(hipStream_t s, const double* nodes, int sheet_rows,
                         int row_off, const double* gamma, int g_sheet_rows,
                         int g_row_off, int num, int rows, int cols,
                         RingRec* out) {
  const int total = num * rows * cols;
  if (total <= 0) return;
  hipLaunchKernelGGL(k_pack_lattice, dim3(cdiv(total, 256)), dim3(256), 0, s,
                     nodes, sheet_rows, row_off, gamma, g_sheet_rows, g_row_off,
                     num, rows, cols, out);
}

void launch_pack_rings(hipStream_t s, const double* rings, const double* gamma,
                       int n, RingRec* out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pack_rings, dim3(cdiv(n, 256)), dim3(256), 0, s, rings,
                     gamma, n, out);
}

void launch_set_gamma(hipStream_t s, const double* gamma, int n, RingRec* out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_set_gamma, dim3(cdiv(n, 256)), dim3(256), 0, s, gamma, n,
                     out);
}

void launch_nbody(hipStream_t s, int prec, int seg, const double* targets,
                  int nt, const RingRec* recs, const ChunkPlan& plan,
                  double* partial, int ld, const CutOff& cut) {
  if (nt <= 0 || plan.total() <= 0) return;
  dim3 grid(cdiv(nt, kBlock), plan.total());
#define UVLM_NB(P, S)                                                          \
  hipLaunchKernelGGL((k_nbody<P, S>), grid, dim3(kBlock), 0, s, targets, nt, \
                     recs, plan, partial, ld, cut)
  if (prec == kF64Exact) {
    switch (seg) {
      case kSegVoring: UVLM_NB(kF64Exact, kSegVoring); break;
      case kSegRingPerRing: UVLM_NB(kF64Exact, kSegRingPerRing); break;
      case kSegRingChord: UVLM_NB(kF64Exact, kSegRingChord); break;
      default: UVLM_NB(kF64Exact, kSegRingPerSeg); break;
    }
  } else {
    switch (seg) {
      case kSegVoring: UVLM_NB(kF64Fast, kSegVoring); break;
      case kSegRingPerRing: UVLM_NB(kF64Fast, kSegRingPerRing); break;
      case kSegRingChord: UVLM_NB(kF64Fast, kSegRingChord); break;
      default: UVLM_NB(kF64Fast, kSegRingPerSeg); break;
    }
  }
#undef UVLM_NB
}

void launch_segments(hipStream_t s, int prec, const double* a, const double* b,
                     const double* x, int n, const CutOff& cut, double* out) {
  if (n <= 0) return;
  if (prec == kF64Exact)
    hipLaunchKernelGGL((k_segments<kF64Exact>), dim3(cdiv(n, 256)), dim3(256), 0,
                       s, a, b, x, n, cut, out);
  else
    hipLaunchKernelGGL((k_segments<kF64Fast>), dim3(cdiv(n, 256)), dim3(256), 0,
                       s, a, b, x, n, cut, out);
}

void launch_sum_chunks(hipStream_t s, const double* partial, int ld, int nt,
                       int chunks, double* out) {
  if (nt <= 0) return;
  if (chunks > kWideChunks)
    hipLaunchKernelGGL(k_sum_chunks_wide, dim3(cdiv(nt, 64)), dim3(64 * kSumGroups), 0, s,
                       partial, ld, nt, chunks, out);
  else
    hipLaunchKernelGGL(k_sum_chunks, dim3(cdiv(nt, 256)), dim3(256), 0, s,
                       partial, ld, nt, chunks, out);
}

void launch_flow_velocity(hipStream_t s, const double* partial, int ld, int nt,
                          int cb, int cw, const double* ff, double* out,
                          int te_period, int te_width) {
  if (nt <= 0) return;
  if (cb + cw > kWideChunks)
    hipLaunchKernelGGL(k_flow_velocity_wide, dim3(cdiv(nt, 64)), dim3(64 * kSumGroups), 0, s,
                       partial, ld, nt, cb, cw, ff[0], ff[1], ff[2], out, te_period, te_width);
  else
    hipLaunchKernelGGL(k_flow_velocity, dim3(cdiv(nt, 256)), dim3(256), 0, s,
                       partial, ld, nt, cb, cw, ff[0], ff[1], ff[2], out,
                       te_period, te_width);
}

}  // namespace uvlm
