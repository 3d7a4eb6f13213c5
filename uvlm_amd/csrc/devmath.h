// devmath.h -- device arithmetic shared by the chain N-body kernels.
#pragma once

#include "internal.h"

namespace uvlm {
namespace dev {

// Refined hardware reciprocal square root / reciprocal (two Newton steps
// each: full FP64 accuracy from the hardware seed).
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

// One Newton step each from the hardware seed: relative error below 2^-47
// (rsqrt, ~20 ulp) and 2^-49 (reciprocal), for the fast chains whose pair
// sums carry rounding of that order anyway.
__device__ __forceinline__ double rsq64_1(double a) {
  const double y = __builtin_amdgcn_rsq(a);
  const double e = __builtin_fma(-(a * y), y, 1.0);
  return __builtin_fma(0.5 * y, e, y);
}
__device__ __forceinline__ double rcp64_1(double a) {
  const double r = __builtin_amdgcn_rcp(a);
  return __builtin_fma(r, __builtin_fma(-a, r, 1.0), r);
}

// Carried state of a chain walk for one target: x - p of the previous node,
// its squared norm and 1/|x - p|.
struct Carry64 {
  double rx, ry, rz, rs, inv;
};

__device__ __forceinline__ void carry_reset(Carry64& c) {
  c.rx = c.ry = c.rz = c.rs = c.inv = 0.0;
}

// Squared segment length below which the end-point cut-offs are implied by
// the |d|^2 cut-off: |r1| < 1e-10 gives |d| <= |r1| |r2| < 1e-10 (|r0| +
// 1e-10), far below sqrt(1e-10) while |r0| < 3e4 (likewise for r2).
constexpr double kShortSeg2 = 1e9;

// One chain element e (ChainRec as 8 doubles: p[3], g, r0[3], r0sq) at
// target (x, y, z): adds g * BiotSavartLaw(prev -> p) to u and advances the
// carry to p.  d = r1 x r2 = r0 x (x - prev), r0 . r2 = r0 . r1 - |r0|^2;
// the cut-off set is the reference's (|d|^2, |r1|, |r2| against 1e-10) up
// to rounding.  Ends = false drops the end-point tests, exact when the
// segment is shorter than sqrt(kShortSeg2) (a NaN from a target on a node is
// then removed by the |d|^2 test, since d = 0 there).
template <bool Ends = true>
__device__ __forceinline__ void chain_step64(const double* e, double x, double y, double z,
                                             Carry64& c, double& ux, double& uy,
                                             double& uz, double eps = kCutoff,
                                             double eps2 = kCutoff * kCutoff) {
  const double rbx = x - e[0], rby = y - e[1], rbz = z - e[2];
  const double rbs = __builtin_fma(rbz, rbz, __builtin_fma(rby, rby, rbx * rbx));
  const double ib = rsq64_1(rbs);
  const double r0x = e[4], r0y = e[5], r0z = e[6];
  const double dx = __builtin_fma(r0y, c.rz, -r0z * c.ry);
  const double dy = __builtin_fma(r0z, c.rx, -r0x * c.rz);
  const double dz = __builtin_fma(r0x, c.ry, -r0y * c.rx);
  const double dsq = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
  const double oa = __builtin_fma(r0z, c.rz, __builtin_fma(r0y, c.ry, r0x * c.rx));
  const double ob = oa - e[7];
  const double tt = __builtin_fma(oa, c.inv, -ob * ib);
  bool cut = dsq < eps;
  if (Ends) cut = cut | (c.rs < eps2) | (rbs < eps2);
  double K = tt * e[3] * rcp64_1(dsq);
  K = cut ? 0.0 : K;
  ux = __builtin_fma(dx, K, ux);
  uy = __builtin_fma(dy, K, uy);
  uz = __builtin_fma(dz, K, uz);
  c.rx = rbx;
  c.ry = rby;
  c.rz = rbz;
  c.rs = rbs;
  c.inv = ib;
}

// BiotSavartLaw of segment a->b at x, times g (vortex.cpp:21-39 formula,
// FP64, used for the near-core pairs of F32C).
__device__ __forceinline__ void seg_ref(double ax, double ay, double az, double bx,
                                        double by, double bz, double x, double y,
                                        double z, double g, double& ux, double& uy,
                                        double& uz) {
  const double r0x = bx - ax, r0y = by - ay, r0z = bz - az;
  const double r1x = x - ax, r1y = y - ay, r1z = z - az;
  const double r2x = x - bx, r2y = y - by, r2z = z - bz;
  const double dx = r1y * r2z - r1z * r2y;
  const double dy = r1z * r2x - r1x * r2z;
  const double dz = r1x * r2y - r1y * r2x;
  const double dsq = (dx * dx + dy * dy) + dz * dz;
  const double r1n = sqrt((r1x * r1x + r1y * r1y) + r1z * r1z);
  const double r2n = sqrt((r2x * r2x + r2y * r2y) + r2z * r2z);
  const bool cut = (sqrt(dsq) * sqrt(dsq) < kCutoff) | (r1n < kCutoff) | (r2n < kCutoff);
  const double K = ((r0x * (r1x / r1n - r2x / r2n) + r0y * (r1y / r1n - r2y / r2n)) +
                    r0z * (r1z / r1n - r2z / r2n)) /
                   dsq;
  if (!cut) {
    ux += dx * K * g;
    uy += dy * K * g;
    uz += dz * K * g;
  }
}

// --- the FP32 segment law with compensation ---------------------------------
// 64 (2^-24)^2: the FP32 law's bound on the rounding of |d|^2 is
// kBand32 |b - a|^2 |x - a|^2; records hold kBand32 |b - a|^2 (an exact
// scaling).
constexpr float kBand32 = 0x1p-42f;

// End point of a segment seen from the target, in FP32: offsets (x - p
// formed in FP64, so exact to FP32 rounding relative to |x - p|), squared
// distance and its reciprocal root.
struct End32 {
  float rx, ry, rz, rs, inv;
};

__device__ __forceinline__ void end32(const double* p, double x, double y, double z, End32& e) {
  e.rx = (float)(x - p[0]);
  e.ry = (float)(y - p[1]);
  e.rz = (float)(z - p[2]);
  e.rs = __builtin_fmaf(e.rz, e.rz, __builtin_fmaf(e.ry, e.ry, e.rx * e.rx));
  e.inv = __builtin_amdgcn_rsqf(e.rs);
}

// d = r0 x (x - a) of a segment a -> b (f = (float)(b - a), ...) and |d|^2.
__device__ __forceinline__ void cross32(const End32& a, const float* f, float& dx, float& dy,
                                        float& dz, float& dsq) {
  const float r0x = f[0], r0y = f[1], r0z = f[2];
  dx = __builtin_fmaf(r0y, a.rz, -r0z * a.ry);
  dy = __builtin_fmaf(r0z, a.rx, -r0x * a.rz);
  dz = __builtin_fmaf(r0x, a.ry, -r0y * a.rx);
  dsq = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// The cut-off band of the FP32 law: |d|^2 below 2e-9 plus the FP32 rounding
// bound of d, or |r| < 1e-9 at either end.  Ends = false drops the end
// tests, which the first implies for segments shorter than sqrt(kShortSeg2):
// an end within 1e-9 of the target gives |d| <= |r0| 1e-9 plus the rounding
// of x - a, which the band's bound covers.
template <bool Ends>
__device__ __forceinline__ bool band32(const End32& a, const End32& b, const float* f,
                                       float dsq) {
  const bool core = dsq < __builtin_fmaf(f[3], a.rs, 2e-9f);
  return Ends ? core | (fminf(a.rs, b.rs) < 1e-18f) : core;
}

// One segment a -> b of strength g (with 1/4pi): f = (float)(b - a) and
// kBand32 |b - a|^2 (so |b - a|^2 = f[3] 2^42 exactly).  Pairs within the
// band add nothing here and are reported, for the caller's FP64
// re-evaluation; in32 tells again, for the same ends, whether elem32
// reported the pair (the same operations, so the same decision).
template <bool Ends = true>
__device__ __forceinline__ bool elem32(const End32& a, const End32& b, const float* f, float g,
                                       float& fx, float& fy, float& fz) {
  float dx, dy, dz, dsq;
  cross32(a, f, dx, dy, dz, dsq);
  const float r0x = f[0], r0y = f[1], r0z = f[2];
  const float oa = __builtin_fmaf(r0z, a.rz, __builtin_fmaf(r0y, a.ry, r0x * a.rx));
  const float ob = __builtin_fmaf(-f[3], 0x1p42f, oa);  // r0.(x - b) = r0.(x - a) - |r0|^2
  const float tt = __builtin_fmaf(oa, a.inv, -ob * b.inv);
  const bool nearcore = band32<Ends>(a, b, f, dsq);
  float K = tt * g * __builtin_amdgcn_rcpf(dsq);
  K = nearcore ? 0.f : K;
  fx = __builtin_fmaf(dx, K, fx);
  fy = __builtin_fmaf(dy, K, fy);
  fz = __builtin_fmaf(dz, K, fz);
  return nearcore;
}
template <bool Ends = true>
__device__ __forceinline__ bool in32(const End32& a, const End32& b, const float* f) {
  float dx, dy, dz, dsq;
  cross32(a, f, dx, dy, dz, dsq);
  return band32<Ends>(a, b, f, dsq);
}

}  // namespace dev
}  // namespace uvlm
