// chain.hip -- node-shared Biot-Savart N-body over segment chains (gfx950).
//
// The fast modes evaluate every vortex lattice as chains of merged segments
// (ChainRec): an interior lattice segment is shared by two rings, so it is
// evaluated once with the difference of their circulations, and consecutive
// segments of a chain share a node, so 1/|x - p| of each node is computed
// once per target and carried to the next segment.  Per target and lattice
// node this is one reciprocal square root, against two per segment and four
// segments per ring in the reference's ring-by-ring order.
//
//   F64  : FP64 with FMA, refined hardware rsqrt / reciprocal.
//   F32C : x - p in FP64 (so the FP32 difference is exact to FP32 rounding
//          relative to |x - p|), then FP32 arithmetic with the hardware
//          rsqrt / reciprocal; contributions are summed in FP32 over runs
//          of 16 elements and carried in FP64 per target.  Pairs within a
//          band around the reference's cut-off (|r| < 1e-9, |d|^2 < 1e-9 plus
//          the FP32 rounding bound of d) are re-evaluated in FP64 exactly as
//          the reference does.
#include <algorithm>
#include <cmath>

#include "devmath.h"
#include "internal.h"

namespace uvlm {
namespace {

constexpr int kBlock = 256;
constexpr int kTile = 128;  // chain elements per LDS tile (8 KB)

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

__device__ __forceinline__ void chunk_range(const ChunkPlan& p, int c, int& b, int& e) {
  if (c < p.cb) {
    b = c * p.szb;
    e = min(b + p.szb, p.nb);
  } else {
    b = p.nb + (c - p.cb) * p.szw;
    e = min(b + p.szw, p.nb + p.nw);
  }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_chain_f64(
    const double* __restrict__ targets, int nt, const ChainRec* __restrict__ recs,
    ChunkPlan plan, double* __restrict__ partial, int ld, CutOff cut) {
  __shared__ double2 tile[kTile * 4];
  int rb, re;
  chunk_range(plan, blockIdx.y, rb, re);
  const int t = blockIdx.x * kBlock + threadIdx.x;
  double x = 0.0, y = 0.0, z = 0.0;
  if (t < nt) {
    x = targets[3 * t + 0];
    y = targets[3 * t + 1];
    z = targets[3 * t + 2];
  }
  // carried previous node
  dev::Carry64 cr;
  dev::carry_reset(cr);
  if (rb > 0) {
    const double* p = recs[rb - 1].p;
    cr.rx = x - p[0];
    cr.ry = y - p[1];
    cr.rz = z - p[2];
    cr.rs = __builtin_fma(cr.rz, cr.rz, __builtin_fma(cr.ry, cr.ry, cr.rx * cr.rx));
    cr.inv = dev::rsq64_1(cr.rs);
  }
  double ux = 0.0, uy = 0.0, uz = 0.0;
  const double2* src = reinterpret_cast<const double2*>(recs);
  for (int base = rb; base < re; base += kTile) {
    const int cnt = min(kTile, re - base);
    __syncthreads();
    int lng = 0;  // a segment of the tile too long for the short form
    for (int q = threadIdx.x; q < cnt * 4; q += kBlock) {
      const double2 v = src[(size_t)base * 4 + q];
      tile[q] = v;
      lng |= (q & 3) == 3 && !(v.y < cut.short2);
    }
    const bool any_long = __syncthreads_or(lng);
    const double* tl = reinterpret_cast<const double*>(tile);
    if (any_long) {
#pragma unroll 1
      for (int k = 0; k < cnt; ++k)
        dev::chain_step64<true>(tl + 8 * k, x, y, z, cr, ux, uy, uz, cut.eps, cut.eps2);
    } else {
#pragma unroll 2
      for (int k = 0; k < cnt; ++k)
        dev::chain_step64<false>(tl + 8 * k, x, y, z, cr, ux, uy, uz, cut.eps, cut.eps2);
    }
  }
  if (t < nt) {
    const size_t c = blockIdx.y;
    partial[(c * 3 + 0) * (size_t)ld + t] = ux;
    partial[(c * 3 + 1) * (size_t)ld + t] = uy;
    partial[(c * 3 + 2) * (size_t)ld + t] = uz;
  }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_chain_f32c(
    const double* __restrict__ targets, int nt, const ChainRec* __restrict__ recs,
    ChunkPlan plan, double* __restrict__ partial, int ld) {
  __shared__ double2 tile[kTile * 4];
  int rb, re;
  chunk_range(plan, blockIdx.y, rb, re);
  const int t = blockIdx.x * kBlock + threadIdx.x;
  double x = 0.0, y = 0.0, z = 0.0;
  if (t < nt) {
    x = targets[3 * t + 0];
    y = targets[3 * t + 1];
    z = targets[3 * t + 2];
  }
  double pax = 0.0, pay = 0.0, paz = 0.0;  // previous node (FP64, near-core path)
  dev::End32 A = {0.f, 0.f, 0.f, 0.f, 0.f}, B;  // element 0 starts a chain (g = 0)
  if (rb > 0) {
    const double* p = recs[rb - 1].p;
    pax = p[0];
    pay = p[1];
    paz = p[2];
    dev::end32(p, x, y, z, A);
  }
  double ux = 0.0, uy = 0.0, uz = 0.0;
  const double2* src = reinterpret_cast<const double2*>(recs);
  for (int base = rb; base < re; base += kTile) {
    const int cnt = min(kTile, re - base);
    __syncthreads();
    for (int q = threadIdx.x; q < cnt * 4; q += kBlock) tile[q] = src[(size_t)base * 4 + q];
    __syncthreads();
    const double* tl = reinterpret_cast<const double*>(tile);
    for (int k0 = 0; k0 < cnt; k0 += 16) {
      float fx = 0.f, fy = 0.f, fz = 0.f;
      const int kend = min(cnt, k0 + 16);
      // run start: previous node for the FP64 re-evaluation pass below
      const double qx = pax, qy = pay, qz = paz;
      // lanes with an element of the run to re-evaluate in FP64 (wave mask)
      unsigned long long redo = 0ull;
      const dev::End32 A0 = A;  // the run's first end, for that pass
      // Two elements per trip, the end point alternating between A and B;
      // chain starts (g = 0, r0 = 0) fall in the band and are not re-evaluated.
      int k = k0;
      for (; k + 2 <= kend; k += 2) {
        const double* e0 = tl + 8 * k;  // ChainRec32
        const double* e1 = e0 + 8;
        const float* f0 = reinterpret_cast<const float*>(e0 + 4);
        const float* f1 = reinterpret_cast<const float*>(e1 + 4);
        dev::end32(e0, x, y, z, B);
        redo |= __ballot(dev::elem32(A, B, f0, f0[4], fx, fy, fz) && e0[3] != 0.0);
        dev::end32(e1, x, y, z, A);
        redo |= __ballot(dev::elem32(B, A, f1, f1[4], fx, fy, fz) && e1[3] != 0.0);
      }
      if (k < kend) {
        const double* e0 = tl + 8 * k;
        const float* f0 = reinterpret_cast<const float*>(e0 + 4);
        dev::end32(e0, x, y, z, B);
        redo |= __ballot(dev::elem32(A, B, f0, f0[4], fx, fy, fz) && e0[3] != 0.0);
        A = B;
      }
      {
        const double* el = tl + 8 * (kend - 1);
        pax = el[0];
        pay = el[1];
        paz = el[2];
      }
      if (redo) {  // rare: pairs near the cut-off again, in FP64
        const bool mine = (redo >> __lane_id()) & 1ull;
        double ax = qx, ay = qy, az = qz;
        dev::End32 P = A0, Q;
        for (int k = k0; k < kend; ++k) {
          const double* e = tl + 8 * k;
          dev::end32(e, x, y, z, Q);
          if (mine && e[3] != 0.0 && dev::in32(P, Q, reinterpret_cast<const float*>(e + 4)))
            dev::seg_ref(ax, ay, az, e[0], e[1], e[2], x, y, z, e[3], ux, uy, uz);
          P = Q;
          ax = e[0];
          ay = e[1];
          az = e[2];
        }
      }
      ux += (double)fx;
      uy += (double)fy;
      uz += (double)fz;
    }
  }
  if (t < nt) {
    const size_t c = blockIdx.y;
    partial[(c * 3 + 0) * (size_t)ld + t] = ux;
    partial[(c * 3 + 1) * (size_t)ld + t] = uy;
    partial[(c * 3 + 2) * (size_t)ld + t] = uz;
  }
}

// ---------------------------------------------------------------------------
// Lattice -> chains.  Per sheet: spanwise chains (node row i = 0..rows, nodes
// j = 0..cols; segment (i,j-1)->(i,j) carries G(i,j-1) - G(i-1,j-1)), then
// chordwise chains (column j = 0..cols, nodes i = 0..rows; segment
// (i-1,j)->(i,j) carries G(i-1,j-1) - G(i-1,j)), G = 0 outside the lattice.
// These are the VORING segments p0->p1 (+G) and p2->p3 (-G of the ring
// below) and p1->p2 (+G) and p3->p0 (-G of the ring to the right).
__global__ void k_pack_chains(const double* __restrict__ nodes, int sheet_rows,
                              int row_off, const double* __restrict__ gamma,
                              int g_sheet_rows, int g_row_off, int num, int rows,
                              int cols, ChainRec* __restrict__ out, bool f32) {
  const int per = (rows + 1) * (cols + 1);
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 2 * num * per) return;
  const int n = q / (2 * per);
  const int r = q % (2 * per);
  const int C1 = cols + 1;
  auto P = [&](int i, int j) {
    return nodes + 3 * ((size_t)(n * sheet_rows + row_off + i) * C1 + j);
  };
  auto G = [&](int i, int j) -> double {
    if (i < 0 || i >= rows || j < 0 || j >= cols) return 0.0;
    return gamma[(size_t)(n * g_sheet_rows + g_row_off + i) * cols + j];
  };
  int i, j, pi = -1, pj = -1;
  double g = 0.0;
  if (r < per) {  // spanwise
    i = r / C1;
    j = r % C1;
    if (j > 0) {
      pi = i;
      pj = j - 1;
      g = G(i, j - 1) - G(i - 1, j - 1);
    }
  } else {  // chordwise
    const int s = r - per;
    j = s / (rows + 1);
    i = s % (rows + 1);
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
  if (!f32) {
    out[q] = e;
    return;
  }
  ChainRec32 f;
  f.p[0] = e.p[0];
  f.p[1] = e.p[1];
  f.p[2] = e.p[2];
  f.g = e.g;
  f.r0[0] = (float)e.r0[0];
  f.r0[1] = (float)e.r0[1];
  f.r0[2] = (float)e.r0[2];
  f.band = dev::kBand32 * (float)e.r0sq;
  f.gf = (float)e.g;
  f.pad[0] = f.pad[1] = f.pad[2] = 0.f;
  reinterpret_cast<ChainRec32*>(out)[q] = f;
}

}  // namespace

void launch_pack_chains(hipStream_t s, const double* nodes, int sheet_rows, int row_off,
                        const double* gamma, int g_sheet_rows, int g_row_off, int num,
                        int rows, int cols, ChainRec* out, bool f32) {
  const int total = chain_elems(num, rows, cols);
  if (total <= 0) return;
  hipLaunchKernelGGL(k_pack_chains, dim3(cdiv(total, 256)), dim3(256), 0, s, nodes,
                     sheet_rows, row_off, gamma, g_sheet_rows, g_row_off, num, rows, cols,
                     out, f32);
}

int chain_slots(int prec) {
  static int cached[2][64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  const int k = prec == kF32C ? 1 : 0;
  if (cached[k][dev] == 0) {
    int per_cu = 0, cus = 0;
    const hipError_t e1 =
        k ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_chain_f32c, kBlock, 0)
          : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_chain_f64, kBlock, 0);
    const hipError_t e2 = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e1 != hipSuccess || e2 != hipSuccess || per_cu <= 0 || cus <= 0) {
      (void)hipGetLastError();
      cached[k][dev] = -1;
    } else {
      cached[k][dev] = per_cu * cus;
    }
  }
  return cached[k][dev] > 0 ? cached[k][dev] : 0;
}

void launch_nbody_chain(hipStream_t s, int prec, const double* targets, int nt,
                        const ChainRec* recs, const ChunkPlan& plan, double* partial,
                        int ld, const CutOff& cut) {
  if (nt <= 0 || plan.total() <= 0) return;
  dim3 grid(cdiv(nt, kBlock), plan.total());
  if (prec == kF32C)
    hipLaunchKernelGGL(k_chain_f32c, grid, dim3(kBlock), 0, s, targets, nt, recs, plan,
                       partial, ld);
  else
    hipLaunchKernelGGL(k_chain_f64, grid, dim3(kBlock), 0, s, targets, nt, recs, plan,
                       partial, ld, cut);
}

}  // namespace uvlm
