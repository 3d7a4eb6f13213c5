// step.hip -- per-timestep kernels of the SimpleSimulator loop on gfx950.
//
// Geometry, influence matrix, right-hand side, shedding into the wake ring
// buffer, the advection updates of every AdvectFactory scheme and the
// Joukowski loads.  Arithmetic follows the reference's operation order
// (file compiled with -ffp-contract=off); the reductions of the loads use a
// tree order instead of the reference's OpenMP reduction.
#include <cmath>

#include "devmath.h"
#include "internal.h"
#include "step.h"

namespace uvlm {
namespace {

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

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// simple_simulator.cpp:19-49 and joukowski.cpp:55-67
__global__ void k_geometry(const double* __restrict__ nodes, int num, int R,
                           int C, double* __restrict__ cp,
                           double* __restrict__ nrm, double* __restrict__ area) {
  const int K = blockIdx.x * blockDim.x + threadIdx.x;
  if (K >= num * R * C) return;
  const int j = K % C, i = (K / C) % R, n = K / (R * C);
  const int C1 = C + 1;
  auto P = [&](int ii, int jj) {
    return ld3(nodes + 3 * ((size_t)(n * (R + 1) + ii) * C1 + jj));
  };
  const d3 p00 = P(i, j), p10 = P(i + 1, j), p11 = P(i + 1, j + 1), p01 = P(i, j + 1);
  st3(cp + 3 * (size_t)K, dvs(add(add(add(p00, p10), p11), p01), 4));
  st3(nrm + 3 * (size_t)K, normalized(crs(sub(p11, p00), sub(p01, p10))));
  area[K] = sqrt(sqn(crs(sub(p10, p00), sub(p01, p00))));
}

// BiotSavartLaw (vortex.cpp:21-39), same arithmetic as the N-body kernel.
__device__ __forceinline__ d3 seg_exact(d3 a, d3 b, d3 x, const CutOff& k) {
  const d3 r0 = sub(b, a), r1 = sub(x, a), r2 = sub(x, b);
  const d3 d = crs(r1, r2);
  const double dsq = sqn(d);
  const double r1n = sqrt(sqn(r1));
  const double r2n = sqrt(sqn(r2));
  const bool cut = (dsq < k.td) | (r1n < k.eps) | (r2n < k.eps);
  const d3 w = sub(dvs(r1, r1n), dvs(r2, r2n));
  const double K = kInv4Pi / dsq * dot(r0, w);
  return cut ? d3{0.0, 0.0, 0.0} : scl(d, K);
}

// A[L*N + K] = VORING_L(cp_K; Gamma = 1) . n_K  (simple_simulator.cpp:76-101)
// LEGACY: ring segments node[(k+1)%4] -> node[k] summed, times Gamma = 1
// (VortexRing::BiotSavartLaw, linear.inl:14-32).
// Block: 256 targets x kMatRings source rings staged in LDS.
constexpr int kMatRings = 32;
template <bool LEGACY>
__global__ __launch_bounds__(256) void k_matrix(const RingRec* __restrict__ recs,
                                                const double* __restrict__ cp,
                                                const double* __restrict__ nrm,
                                                int N, CutOff cut,
                                                double* __restrict__ A) {
  __shared__ double2 tile[kMatRings * 8];
  const int l0 = blockIdx.y * kMatRings;
  const int cnt = min(kMatRings, N - l0);
  const double2* src = reinterpret_cast<const double2*>(recs + l0);
  for (int q = threadIdx.x; q < cnt * 8; q += 256) tile[q] = src[q];
  __syncthreads();
  const int K = blockIdx.x * 256 + threadIdx.x;
  if (K >= N) return;
  const d3 x = ld3(cp + 3 * (size_t)K);
  const d3 nl = ld3(nrm + 3 * (size_t)K);
  const double* tl = reinterpret_cast<const double*>(tile);
  for (int l = 0; l < cnt; ++l) {
    const double* r = tl + 16 * l;
    const d3 p[4] = {ld3(r), ld3(r + 3), ld3(r + 6), ld3(r + 9)};
    d3 u{0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (LEGACY)
        u = add(u, seg_exact(p[(k + 1) & 3], p[k], x, cut));
      else
        u = add(u, seg_exact(p[k], p[(k + 1) & 3], x, cut));
    }
    A[(size_t)(l0 + l) * N + K] = dot(u, nl);
  }
}

// Rsqrt form of the same matrix for the FP64-chain and F32C modes: per
// (collocation point, ring) one refined hardware rsqrt per corner (shared by
// the two segments that meet there) and one refined reciprocal per segment,
// FMA-contracted, each segment's term projected on the normal before the
// sum; A = (sum of d.n r0.(r1/|r1| - r2/|r2|) / |d|^2) / 4pi.  The cut-off
// tests are the segment chains' (|d|^2 < eps, |r1|, |r2| < eps on the
// squares).  Agrees with the reference order to ~1e-14 of max |A| (the
// rsqrt and reciprocal carry one Newton step each, as in the chains).
__global__ __launch_bounds__(256) void k_matrix_fast(const RingRec* __restrict__ recs,
                                                     const double* __restrict__ cp,
                                                     const double* __restrict__ nrm, int N,
                                                     CutOff cut, double* __restrict__ A) {
  __shared__ double2 tile[kMatRings * 8];
  const int l0 = blockIdx.y * kMatRings;
  const int cnt = min(kMatRings, N - l0);
  const double2* src = reinterpret_cast<const double2*>(recs + l0);
  for (int q = threadIdx.x; q < cnt * 8; q += 256) tile[q] = src[q];
  __syncthreads();
  const int K = blockIdx.x * 256 + threadIdx.x;
  if (K >= N) return;
  const d3 x = ld3(cp + 3 * (size_t)K);
  const d3 nl = ld3(nrm + 3 * (size_t)K);
  const double* tl = reinterpret_cast<const double*>(tile);
  for (int l = 0; l < cnt; ++l) {
    const double* r = tl + 16 * l;
    double rx[4], ry[4], rz[4], rs[4], inv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      rx[k] = x.x - r[3 * k];
      ry[k] = x.y - r[3 * k + 1];
      rz[k] = x.z - r[3 * k + 2];
      rs[k] = __builtin_fma(rz[k], rz[k], __builtin_fma(ry[k], ry[k], rx[k] * rx[k]));
      inv[k] = dev::rsq64_1(rs[k]);
    }
    double a = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int b = (k + 1) & 3;
      const double r0x = r[3 * b] - r[3 * k], r0y = r[3 * b + 1] - r[3 * k + 1],
                   r0z = r[3 * b + 2] - r[3 * k + 2];
      const double dx = __builtin_fma(r0y, rz[k], -r0z * ry[k]);
      const double dy = __builtin_fma(r0z, rx[k], -r0x * rz[k]);
      const double dz = __builtin_fma(r0x, ry[k], -r0y * rx[k]);
      const double dsq = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
      const double o1 = __builtin_fma(r0z, rz[k], __builtin_fma(r0y, ry[k], r0x * rx[k]));
      const double o2 = __builtin_fma(r0z, rz[b], __builtin_fma(r0y, ry[b], r0x * rx[b]));
      const double tt = __builtin_fma(o1, inv[k], -o2 * inv[b]);
      const double dn = __builtin_fma(dz, nl.z, __builtin_fma(dy, nl.y, dx * nl.x));
      const bool cutk = (dsq < cut.eps) | (rs[k] < cut.eps2) | (rs[b] < cut.eps2);
      const double t = dn * tt * dev::rcp64_1(dsq);
      a += cutk ? 0.0 : t;
    }
    A[(size_t)(l0 + l) * N + K] = a * kInv4Pi;
  }
}

// rhs_K = -(Uw - (Uls + ff)) . n_K  (simple_simulator.cpp:103-124)
__global__ void k_rhs(const double* __restrict__ uw, const double* __restrict__ vm,
                      const double* __restrict__ nrm, int N, double fx,
                      double fy, double fz, double* __restrict__ rhs) {
  const int K = blockIdx.x * blockDim.x + threadIdx.x;
  if (K >= N) return;
  const d3 Uw = ld3(uw + 3 * (size_t)K);
  const d3 Uls = ld3(vm + 3 * (size_t)K);
  const d3 u = sub(Uw, add(Uls, d3{fx, fy, fz}));
  rhs[K] = -dot(u, ld3(nrm + 3 * (size_t)K));
}

// SimpleSimulator::Shed: TE node row -> wake physical row `row`; TE Gamma
// row -> wake Gamma physical row `grow` (grow < 0: no Gamma row).
__global__ void k_shed(const double* __restrict__ wing, const double* __restrict__ gamma,
                       int num, int R, int C, double* __restrict__ wake, int cap,
                       int row, double* __restrict__ wgam, int capg, int grow) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int C1 = C + 1;
  if (k < num * C1) {
    const int n = k / C1, j = k % C1;
    const double* src = wing + 3 * ((size_t)(n * (R + 1) + R) * C1 + j);
    double* dst = wake + 3 * ((size_t)(n * cap + row) * C1 + j);
    dst[0] = src[0];
    dst[1] = src[1];
    dst[2] = src[2];
  }
  if (grow >= 0 && k < num * C) {
    const int n = k / C, j = k % C;
    wgam[(size_t)(n * capg + grow) * C + j] = gamma[(size_t)(n * R + R - 1) * C + j];
  }
}

// ring buffer rows [row0, row0+rows) of each sheet -> compact [num][rows][C+1]
__global__ void k_gather_rows(const double* __restrict__ rb, int cap, int row0,
                              int num, int rows, int C1, double* __restrict__ out) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)num * rows * C1;
  if (k >= total) return;
  const size_t j = k % C1, i = (k / C1) % rows, n = k / ((size_t)C1 * rows);
  const double* s = rb + 3 * ((n * cap + row0 + i) * C1 + j);
  out[3 * k + 0] = s[0];
  out[3 * k + 1] = s[1];
  out[3 * k + 2] = s[2];
}

// p = wake + k * (dt / div)  evaluated as wake + ((k*dt)/div) (advect.cpp)
__global__ void k_stage_pos(const double* __restrict__ wake,
                            const double* __restrict__ k1, double dt, double div,
                            size_t total, double* __restrict__ out) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  for (int d = 0; d < 3; ++d) {
    const double v = k1[3 * k + d] * dt;
    out[3 * k + d] = wake[3 * k + d] + (div == 1.0 ? v : v / div);
  }
}

// Final updates (advect.cpp:11-216); pos = compact current wake, written
// into the ring buffer rows [row0, row0+rows).
__global__ void k_update(int scheme, const double* __restrict__ pos,
                         const double* __restrict__ k1,
                         const double* __restrict__ k2,
                         const double* __restrict__ k3,
                         const double* __restrict__ k4,
                         const double* __restrict__ vprev, int prev_rows,
                         int ign, double dt, int num, int rows, int C1,
                         double* __restrict__ rb, int cap, int row0) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)num * rows * C1;
  if (k >= total) return;
  const int j = (int)(k % C1), i = (int)((k / C1) % rows), n = (int)(k / ((size_t)C1 * rows));
  const d3 w = ld3(pos + 3 * k);
  const d3 a = ld3(k1 + 3 * k);
  d3 r;
  if (scheme == 0) {  // Euler: wake + vel*dt
    r = add(w, scl(a, dt));
  } else if (scheme == 1) {  // RK2
    if (i > ign)
      r = add(w, dvs(scl(add(a, ld3(k2 + 3 * k)), dt), 2));
    else
      r = add(w, scl(a, dt));
  } else if (scheme == 2 || scheme == 4) {  // RK4 (reference) / classic
    if (i > ign) {
      const d3 b = ld3(k2 + 3 * k), c = ld3(k3 + 3 * k);
      const d3 last = scheme == 4 ? ld3(k4 + 3 * k) : c;
      const d3 s = add(add(add(a, scl(b, 2)), scl(c, 2)), last);
      r = add(w, dvs(scl(s, dt), 6));
    } else {
      r = add(w, scl(a, dt));
    }
  } else {  // AB2
    if (vprev != nullptr && i >= 1 && i - 1 < prev_rows) {
      const d3 vp = ld3(vprev + 3 * ((size_t)(n * prev_rows + i - 1) * C1 + j));
      r = add(w, dvs(scl(sub(scl(a, 3), vp), dt), 2));
    } else {
      r = add(w, scl(a, dt));
    }
  }
  st3(rb + 3 * (((size_t)n * cap + row0 + i) * C1 + j), r);
}

// Unique non-TE edges of each wing: spanwise (i,j)-(i,j+1), i<R, j<C, then
// chordwise (i,j)-(i+1,j), i<R, j<=C.  Midpoint = (p0 + p1) / 2.
__device__ __forceinline__ void edge_ends(int e, int R, int C, int& ia, int& ja,
                                          int& ib, int& jb) {
  if (e < R * C) {
    ia = ib = e / C;
    ja = e % C;
    jb = ja + 1;
  } else {
    const int q = e - R * C;
    ia = q / (C + 1);
    ib = ia + 1;
    ja = jb = q % (C + 1);
  }
}

__global__ void k_edge_midpoints(const double* __restrict__ nodes, int num, int R,
                                 int C, double* __restrict__ mid) {
  const int E = R * (2 * C + 1);
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= num * E) return;
  const int n = k / E, e = k % E;
  int ia, ja, ib, jb;
  edge_ends(e, R, C, ia, ja, ib, jb);
  const int C1 = C + 1;
  const d3 a = ld3(nodes + 3 * ((size_t)(n * (R + 1) + ia) * C1 + ja));
  const d3 b = ld3(nodes + 3 * ((size_t)(n * (R + 1) + ib) * C1 + jb));
  st3(mid + 3 * (size_t)k, dvs(add(a, b), 2));
}

// Joukowski loads of wing n = blockIdx.x (joukowski.cpp:12-92,
// joukowski.h:133-149, simple_simulator.cpp:195-233).  uflow = flow
// velocity ((-ff+bound)+wake) at the unique edge midpoints.
__global__ __launch_bounds__(256) void k_loads(
    const double* __restrict__ nodes, const double* __restrict__ gamma,
    const double* __restrict__ gamma_prev, const double* __restrict__ nrm,
    const double* __restrict__ area, const double* __restrict__ uflow,
    const double* __restrict__ vm, int R, int C, double dt, double Q,
    const double* __restrict__ ref_area, double* __restrict__ out) {
  const int n = blockIdx.x;
  const int E = R * (2 * C + 1);
  const int C1 = C + 1;
  auto P = [&](int i, int j) {
    return ld3(nodes + 3 * ((size_t)(n * (R + 1) + i) * C1 + j));
  };
  auto U = [&](int e) {
    const size_t q = 3 * ((size_t)n * E + e);
    return sub(ld3(uflow + q), ld3(vm + q));
  };
  double f[6] = {0, 0, 0, 0, 0, 0};
  for (int K = threadIdx.x; K < R * C; K += blockDim.x) {
    const int i = K / C, j = K % C;
    const double g = gamma[(size_t)n * R * C + K];
    const d3 p0 = P(i, j), p1 = P(i, j + 1), p2 = P(i + 1, j + 1), p3 = P(i + 1, j);
    // k=0: span(i,j) p0->p1; k=1: chord(i,j+1) p1->p2;
    // k=2: span(i+1,j) p2->p3 (skipped on the TE row); k=3: chord(i,j) p3->p0
    d3 s = scl(crs(U(i * C + j), sub(p1, p0)), g);
    s = add(s, scl(crs(U(R * C + i * C1 + j + 1), sub(p2, p1)), g));
    if (i != R - 1) s = add(s, scl(crs(U((i + 1) * C + j), sub(p3, p2)), g));
    s = add(s, scl(crs(U(R * C + i * C1 + j), sub(p0, p3)), g));
    f[0] += s.x;
    f[1] += s.y;
    f[2] += s.z;
    const size_t KK = (size_t)n * R * C + K;
    const double dg_dt = (g - gamma_prev[KK]) / dt;
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
    const double den = 0.5 * Q * Q * ref_area[n];
    out[2 * n + 0] = F.x / den;
    out[2 * n + 1] = F.z / den;
  }
}

}  // namespace

void launch_geometry(hipStream_t s, const double* nodes, int num, int R, int C,
                     double* cp, double* nrm, double* area) {
  const int N = num * R * C;
  hipLaunchKernelGGL(k_geometry, dim3(cdiv(N, 256)), dim3(256), 0, s, nodes,
                     num, R, C, cp, nrm, area);
}

void launch_matrix(hipStream_t s, const RingRec* recs, const double* cp,
                   const double* nrm, int N, const CutOff& cut, double* A) {
  dim3 g(cdiv(N, 256), cdiv(N, kMatRings));
  hipLaunchKernelGGL(k_matrix<false>, g, dim3(256), 0, s, recs, cp, nrm, N, cut, A);
}

void launch_matrix_fast(hipStream_t s, const RingRec* recs, const double* cp,
                        const double* nrm, int N, const CutOff& cut, double* A) {
  dim3 g(cdiv(N, 256), cdiv(N, kMatRings));
  hipLaunchKernelGGL(k_matrix_fast, g, dim3(256), 0, s, recs, cp, nrm, N, cut, A);
}

void launch_ring_matrix(hipStream_t s, const RingRec* recs, const double* cp,
                        const double* nrm, int N, const CutOff& cut, double* A) {
  dim3 g(cdiv(N, 256), cdiv(N, kMatRings));
  hipLaunchKernelGGL(k_matrix<true>, g, dim3(256), 0, s, recs, cp, nrm, N, cut, A);
}

void launch_rhs(hipStream_t s, const double* uw, const double* vm,
                const double* nrm, int N, const double* ff, double* rhs) {
  hipLaunchKernelGGL(k_rhs, dim3(cdiv(N, 256)), dim3(256), 0, s, uw, vm, nrm, N,
                     ff[0], ff[1], ff[2], rhs);
}

void launch_shed(hipStream_t s, const double* wing, const double* gamma, int num,
                 int R, int C, double* wake, int cap, int row, double* wgam,
                 int capg, int grow) {
  const int n = num * (C + 1);
  hipLaunchKernelGGL(k_shed, dim3(cdiv(n, 256)), dim3(256), 0, s, wing, gamma,
                     num, R, C, wake, cap, row, wgam, capg, grow);
}

void launch_gather_rows(hipStream_t s, const double* rb, int cap, int row0,
                        int num, int rows, int C1, double* out) {
  const size_t total = (size_t)num * rows * C1;
  if (!total) return;
  hipLaunchKernelGGL(k_gather_rows, dim3(cdiv(total, 256)), dim3(256), 0, s, rb,
                     cap, row0, num, rows, C1, out);
}

void launch_stage_pos(hipStream_t s, const double* wake, const double* k,
                      double dt, double div, size_t total, double* out) {
  if (!total) return;
  hipLaunchKernelGGL(k_stage_pos, dim3(cdiv(total, 256)), dim3(256), 0, s, wake,
                     k, dt, div, total, out);
}

void launch_update(hipStream_t s, int scheme, const double* pos,
                   const double* k1, const double* k2, const double* k3,
                   const double* k4, const double* vprev, int prev_rows, int ign,
                   double dt, int num, int rows, int C1, double* rb, int cap,
                   int row0) {
  const size_t total = (size_t)num * rows * C1;
  if (!total) return;
  hipLaunchKernelGGL(k_update, dim3(cdiv(total, 256)), dim3(256), 0, s, scheme,
                     pos, k1, k2, k3, k4, vprev, prev_rows, ign, dt, num, rows,
                     C1, rb, cap, row0);
}

void launch_edge_midpoints(hipStream_t s, const double* nodes, int num, int R,
                           int C, double* mid) {
  const int n = num * R * (2 * C + 1);
  hipLaunchKernelGGL(k_edge_midpoints, dim3(cdiv(n, 256)), dim3(256), 0, s,
                     nodes, num, R, C, mid);
}

void launch_loads(hipStream_t s, const double* nodes, const double* gamma,
                  const double* gamma_prev, const double* nrm,
                  const double* area, const double* uflow, const double* vm,
                  int num, int R, int C, double dt, double Q,
                  const double* ref_area, double* out) {
  hipLaunchKernelGGL(k_loads, dim3(num), dim3(256), 0, s, nodes, gamma,
                     gamma_prev, nrm, area, uflow, vm, R, C, dt, Q, ref_area, out);
}

}  // namespace uvlm
