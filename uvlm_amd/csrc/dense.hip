// dense.hip -- dense circulation solve on gfx950.
//
// Replaces Eigen::FullPivLU (simple_simulator.cpp:263-268).  The default
// solve factors with row pivoting chosen by a tournament (the panel's pivot
// rows are elected tile by tile in LDS); the solution agrees with full
// pivoting to ~cond(A) * eps.  UVLM_SOLVE_F64_FULLPIV reproduces FullPivLU's
// elimination order exactly.
//
// Right-looking blocked LU, column-major.  Fast path (no row interchanges,
// every choice of partial pivoting checked), 64-column steps:
//   k_diag    the 64 x 64 diagonal block factored in registers, and the
//             inverses of its unit-lower and upper factors
//   L21 = A21 U11^-1 and U12 = L11^-1 A12 (one GEMM launch, both in place)
//   A22 -= L21 U12
// on the matrix cores (gemm.hip).  Fallback (a check failed): tournament
// pivoting with 32-column panels,
//   k_tslu    panel factorisation by tournament pivoting (LDS tiles + tree)
//   k_tslu_l21 L21 = A21 U11^-1 below the panel's pivot block
//   k_apply_moves  the panel's row permutation applied to the other columns
//   k_trsm    U12 = L11^-1 A12 (one lane per column in registers, L11 in LDS)
//   A22 -= L21 U12 (gemm.hip),
// then the inverses of the factors' 64 x 64 diagonal blocks.  The solve is
// the permutation and blocked substitution through those inverses
// (k_fwd, k_bwd: one launch per 64-row block).  The mixed-precision modes
// factor a copy in FP32 (the trailing updates in FP32, or with BF16
// operands and FP32 accumulation) and refine the FP64 solution with FP64
// residuals r = b - A x.
#include <algorithm>
#include <climits>
#include <cmath>

#include "gemm.h"
#include "internal.h"

namespace uvlm {
namespace {

constexpr int kNB = 32;

__host__ __device__ inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }


// ---- panel factorisation ----------------------------------------------------
// Fallback panel (k_tslu): the panel rows are cut into tiles of kTR rows,
// each tile factored with partial pivoting in LDS nominates its kb pivot
// rows, and the nominations are merged kFan tiles at a time up a tree inside
// the same launch (the last workgroup to finish a group merges it, so no
// workgroup waits).  This is the pivoting of communication-avoiding LU
// (Grigori, Demmel, Xiang 2011).
constexpr int kTR = 128;  // rows per tile (= workgroup size)
constexpr int kFan = 4;   // tree fan-in: kFan * kNB nominations = one tile

// ---- 64-column steps of the fast path --------------------------------------
// The fast path factors A without row interchanges and checks every
// comparison partial pivoting would make: the pivot |a_jj| against each
// |a_ij| below it, on the updated values, i.e. |l_ij| <= 1 for every entry of
// L.  When no check fails, partial pivoting would have chosen no interchange
// and the factors are GEPP's (up to rounding in exact ties).  The influence
// matrix of a vortex lattice takes this path (a ring's self-influence
// dominates its column).  A failed check sets flag; the caller refactors a
// saved copy with tournament pivoting.
//
// An FP64 operation here has a latency of ~40 cycles and a division ~100,
// so the kernels keep their serial chains in registers of one wavefront:
// no barrier inside a chain, loads staged into registers before the
// arithmetic, pivot rows broadcast with readlane.
constexpr int kDB = 64;

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// 16 x 16 x 4 matrix-core step D += X Y: lane l supplies X[l%16][l/16] and
// Y[l/16][l%16]; result element r of lane l is D[rowof(r, l/16)][l%16].
template <typename T>
struct Mf;
template <>
struct Mf<double> {
  typedef v4d acc;
  static __device__ __forceinline__ acc mma(double x, double y, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int rowof(int r, int lq) { return 4 * r + lq; }
};
template <>
struct Mf<float> {
  typedef v4f acc;
  static __device__ __forceinline__ acc mma(float x, float y, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int rowof(int r, int lq) { return 4 * lq + r; }
};

constexpr int kSub = 16;

// v from lane `l` (a constant after unrolling) to every lane
__device__ __forceinline__ double bcast(double v, int l) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ float bcast(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// 16 x 16 block product on the matrix cores: acc += op(X) op(Y) with X, Y in
// LDS (row stride ldx, ldy), X negated when neg.
template <typename T, int LDX, int LDY>
__device__ __forceinline__ typename Mf<T>::acc mm16(const T* X, const T* Y, bool neg,
                                                   typename Mf<T>::acc acc, int l16, int lq) {
#pragma unroll
  for (int k = 0; k < kSub; k += 4) {
    const T x = X[l16 * LDX + k + lq];
    acc = Mf<T>::mma(neg ? -x : x, Y[(k + lq) * LDY + l16], acc);
  }
  return acc;
}

// 1 / d by the hardware reciprocal and one Newton step (a few ulp; far
// shorter than the division sequence, whose dependent steps cost ~40 cycles
// each on these chains)
__device__ __forceinline__ double fast_rcp(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
}
__device__ __forceinline__ float fast_rcp(float d) {
  const float r = __builtin_amdgcn_rcpf(d);
  return __builtin_fmaf(r, __builtin_fmaf(-d, r, 1.0f), r);
}

// LU without interchanges of the 64 x 64 block M (LDS, identity beyond kb)
// by 256 lanes, in four 16-column stages: the 16 x 16 diagonal block
// factored in registers by 16 lanes of wavefront 0 (pivot row by readlane)
// and its inverses Li[b] = L_bb^-1, Ui[b] = U_bb^-1 formed (lane = column);
// then the blocks below and right of it, M[r][b] U_bb^-1 and L_bb^-1
// M[b][c], and the trailing blocks M[r][c] -= M[r][b] M[b][c] on the matrix
// cores.  *bad: an |l| > 1 (thread-local); *zero: first zero pivot (LDS).
template <typename T>
__device__ void diag_factor(T (*M)[kDB + 1], T (*Li)[kSub][kSub + 1], T (*Ui)[kSub][kSub + 1],
                            int kb, bool* bad, int* zero) {
  typedef typename Mf<T>::acc Acc;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63, l16 = lane & 15, lq = lane >> 4;
  for (int b = 0; b < 4; ++b) {
    const int ob = kSub * b;
    if (w == 0) {  // 16 x 16 LU, lane i = row ob + i
      const int i = l16;
      T x[kSub];
#pragma unroll
      for (int c = 0; c < kSub; ++c) x[c] = M[ob + i][ob + c];
#pragma unroll
      for (int j = 0; j < kSub; ++j) {
        const T piv = bcast(x[j], j);
        if (lane == 0 && piv == T(0) && ob + j < kb) atomicMin(zero, ob + j);
        const bool below = i > j && lane < kSub;
        T l = x[j];
        *bad |= below && fabs(l) > fabs(piv);
        if (piv != T(0)) l = l * fast_rcp(piv);
        if (below) x[j] = l;
        const T lm = below ? l : T(0);
#pragma unroll
        for (int c = j + 1; c < kSub; ++c) x[c] = __builtin_fma(-lm, bcast(x[c], j), x[c]);
      }
      if (lane < kSub) {
#pragma unroll
        for (int c = 0; c < kSub; ++c) M[ob + i][ob + c] = x[c];
      }
    }
    __syncthreads();
    if (w == 0 && lane < 2 * kSub) {  // lanes 0..15: column of L_bb^-1, 16..31: of U_bb^-1
      const int c = l16;
      T y[kSub];
#pragma unroll
      for (int r = 0; r < kSub; ++r) y[r] = r == c ? T(1) : T(0);
      if (lane < kSub) {
#pragma unroll
        for (int k = 0; k < kSub; ++k) {
          const T yk = y[k];
#pragma unroll
          for (int r = k + 1; r < kSub; ++r) y[r] = __builtin_fma(-M[ob + r][ob + k], yk, y[r]);
        }
#pragma unroll
        for (int r = 0; r < kSub; ++r) Li[b][r][c] = y[r];
      } else {
        T rd[kSub];
#pragma unroll
        for (int k = 0; k < kSub; ++k) rd[k] = fast_rcp(M[ob + k][ob + k]);
#pragma unroll
        for (int k = kSub - 1; k >= 0; --k) {
          const T yk = y[k] * rd[k];
          y[k] = yk;
#pragma unroll
          for (int r = 0; r < k; ++r) y[r] = __builtin_fma(-M[ob + r][ob + k], yk, y[r]);
        }
#pragma unroll
        for (int r = 0; r < kSub; ++r) Ui[b][r][c] = y[r];
      }
    }
    __syncthreads();
    const int rem = 3 - b;  // blocks below (right of) the diagonal block
    for (int task = w; task < 2 * rem; task += 4) {
      Acc acc = {T(0), T(0), T(0), T(0)};
      if (task < rem) {  // M[r][b] <- M[r][b] U_bb^-1
        const int orow = ob + kSub * (1 + task);
        acc = mm16<T, kDB + 1, kSub + 1>(&M[orow][ob], &Ui[b][0][0], false, acc, l16, lq);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int rr = Mf<T>::rowof(r, lq);
          M[orow + rr][ob + l16] = acc[r];
          *bad |= fabs(acc[r]) > T(1);
        }
      } else {  // M[b][c] <- L_bb^-1 M[b][c]
        const int ocol = ob + kSub * (1 + task - rem);
        acc = mm16<T, kSub + 1, kDB + 1>(&Li[b][0][0], &M[ob][ocol], false, acc, l16, lq);
#pragma unroll
        for (int r = 0; r < 4; ++r) M[ob + Mf<T>::rowof(r, lq)][ocol + l16] = acc[r];
      }
    }
    __syncthreads();
    for (int task = w; task < rem * rem; task += 4) {  // M[r][c] -= M[r][b] M[b][c]
      const int orow = ob + kSub * (1 + task % rem), ocol = ob + kSub * (1 + task / rem);
      Acc acc;
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = M[orow + Mf<T>::rowof(r, lq)][ocol + l16];
      acc = mm16<T, kDB + 1, kDB + 1>(&M[orow][ob], &M[ob][ocol], true, acc, l16, lq);
#pragma unroll
      for (int r = 0; r < 4; ++r) M[orow + Mf<T>::rowof(r, lq)][ocol + l16] = acc[r];
    }
    __syncthreads();
  }
}

// One 64-column step of the fast path.  Every workgroup factors the
// diagonal block at k0 (read from A, identity beyond kb) with diag_factor;
// workgroup 0 stores the factored block in dblk (not in A: a workgroup of
// the same launch dispatched after it finished would read the factored
// block as the original), reports |l| > 1 (flag) and the first zero pivot
// (info).  Then L21 = A21 U11^-1 (workgroups [0, nrb): 64 rows each) and
// U12 = L11^-1 A12 (workgroups from nrb: 64 columns each) in four 16-column
// (16-row) stages on the matrix cores: stage b forms
//   L21[:, b] = (A21[:, b] - L21[:, <b] U11[<b, b]) U_bb^-1
//   U12[b, :] = L_bb^-1 (A12[b, :] - L11[b, <b] U12[<b, :])
// Wavefront w owns the 16 rows (columns) w of the tile, so the stages need
// no exchange between wavefronts.  The diagonal blocks reach A through
// k_diag_inv after the factorisation.
template <typename T>
__global__ __launch_bounds__(256) void k_panel64(T* A, int lda, int n, int k0, int nrb,
                                                 T* __restrict__ dblk, int* flag, int* info) {
  typedef typename Mf<T>::acc Acc;
  __shared__ T D[kDB][kDB + 1];    // diagonal block, factored in place
  __shared__ T W[kDB][kDB + 1];    // the tile: [row][column]
  __shared__ T Tm[4][kSub][kSub + 1];
  __shared__ T Li[4][kSub][kSub + 1], Ui[4][kSub][kSub + 1];
  __shared__ int s_zero;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63, l16 = lane & 15, lq = lane >> 4;
  const int kb = min(kDB, n - k0), s0 = k0 + kb;
  const bool lower = (int)blockIdx.x < nrb;
  const bool tile = (int)blockIdx.x < 2 * nrb;
  // rows (lower) or columns (upper) of the tile, in the matrix
  const int base = s0 + kDB * (lower ? blockIdx.x : blockIdx.x - nrb);
  if (t == 0) s_zero = kDB;
  {
    const int r = t & 63, cq = t >> 6;
    T dv[16], wv[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int c = 16 * cq + e;
      dv[e] = A[(size_t)(k0 + min(c, kb - 1)) * lda + k0 + min(r, kb - 1)];
      wv[e] = !tile ? T(0)
              : lower ? A[(size_t)(k0 + min(c, kb - 1)) * lda + min(base + r, n - 1)]
                      : A[(size_t)min(base + c, n - 1) * lda + k0 + min(r, kb - 1)];
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int c = 16 * cq + e;
      D[r][c] = (r < kb && c < kb) ? dv[e] : (r == c ? T(1) : T(0));
      W[r][c] = (lower ? c < kb : r < kb) ? wv[e] : T(0);
    }
  }
  __syncthreads();
  bool dbad = false;
  diag_factor<T>(D, Li, Ui, kb, &dbad, &s_zero);
  if (blockIdx.x == 0) {
    const int r = t & 63, cq = t >> 6;
#pragma unroll
    for (int e = 0; e < 16; ++e) dblk[(16 * cq + e) * kDB + r] = D[r][16 * cq + e];
    if (__any(dbad) && lane == 0) atomicOr(flag, 1);
    if (t == 0 && s_zero < kDB) atomicCAS(info, 0, k0 + s_zero + 1);
  }
  if (!tile) return;
  bool bad = false;
  for (int b = 0; b < 4; ++b) {
    const int ob = kSub * b;
    Acc acc;
    if (lower) {  // rows 16w.. of the tile, columns ob..: A21 block minus L21[:, <b] U11[<b, b]
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = W[16 * w + Mf<T>::rowof(r, lq)][ob + l16];
      for (int k = 0; k < ob; k += 4)
        acc = Mf<T>::mma(-W[16 * w + l16][k + lq], D[k + lq][ob + l16], acc);
    } else {  // rows ob.. of U12, columns 16w..: A12 block minus L11[b, <b] U12[<b, :]
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = W[ob + Mf<T>::rowof(r, lq)][16 * w + l16];
      for (int k = 0; k < ob; k += 4)
        acc = Mf<T>::mma(-D[ob + l16][k + lq], W[k + lq][16 * w + l16], acc);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) Tm[w][Mf<T>::rowof(r, lq)][l16] = acc[r];
    __syncthreads();
    Acc out = {T(0), T(0), T(0), T(0)};
    if (lower) {  // (..) U_bb^-1
      out = mm16<T, kSub + 1, kSub + 1>(&Tm[w][0][0], &Ui[b][0][0], false, out, l16, lq);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        W[16 * w + Mf<T>::rowof(r, lq)][ob + l16] = out[r];
        bad |= fabs(out[r]) > T(1) && base + 16 * w + Mf<T>::rowof(r, lq) < n && ob + l16 < kb;
      }
    } else {  // L_bb^-1 (..)
      out = mm16<T, kSub + 1, kSub + 1>(&Li[b][0][0], &Tm[w][0][0], false, out, l16, lq);
#pragma unroll
      for (int r = 0; r < 4; ++r) W[ob + Mf<T>::rowof(r, lq)][16 * w + l16] = out[r];
    }
    __syncthreads();
  }
  {
    const int r = t & 63, cq = t >> 6;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int c = 16 * cq + e;
      if (lower) {
        if (base + r < n && c < kb) A[(size_t)(k0 + c) * lda + base + r] = W[r][c];
      } else {
        if (base + c < n && r < kb) A[(size_t)(base + c) * lda + k0 + r] = W[r][c];
      }
    }
  }
  if (__any(bad) && lane == 0) atomicOr(flag, 1);
}

// Inverses of the factored 64 x 64 diagonal blocks (one workgroup per block,
// identity beyond the matrix), column-major with ld 64 at invL, invU + 4096
// b: wavefront 0 forms L^-1, wavefront 1 U^-1, one lane per column, by
// forward and backward elimination on the identity in registers.  dblk
// (nullable): the blocks as the fast path left them (64 x 64 column-major
// each), stored into A here; without it the blocks are read from A.
template <typename T>
__global__ __launch_bounds__(128) void k_diag_inv(T* A, int lda, int n, const T* dblk, T* invL,
                                                  T* invU) {
  __shared__ T M[kDB][kDB + 1];
  __shared__ T rcp[kDB];
  const int k0 = kDB * blockIdx.x, kb = min(kDB, n - k0);
  const int t = threadIdx.x, lane = t & 63;
  invL += (size_t)kDB * kDB * blockIdx.x;
  invU += (size_t)kDB * kDB * blockIdx.x;
  if (t < 64) {
    T v[kDB];
    const T* db = dblk + (size_t)kDB * kDB * blockIdx.x;
#pragma unroll
    for (int c = 0; c < kDB; ++c)
      v[c] = dblk ? db[c * kDB + t]
                  : A[(size_t)(k0 + min(c, kb - 1)) * lda + k0 + min(t, kb - 1)];
#pragma unroll
    for (int c = 0; c < kDB; ++c) M[t][c] = (t < kb && c < kb) ? v[c] : (t == c ? T(1) : T(0));
    if (dblk && t < kb) {  // the fast path's factored diagonal block into A
#pragma unroll
      for (int c = 0; c < kDB; ++c)
        if (c < kb) A[(size_t)(k0 + c) * lda + k0 + t] = v[c];
    }
  }
  __syncthreads();
  if (t < 64) rcp[t] = T(1) / M[t][t];
  __syncthreads();
  T y[kDB];
#pragma unroll
  for (int r = 0; r < kDB; ++r) y[r] = r == lane ? T(1) : T(0);
  if (t < 64) {  // L y = e_lane, unit lower
#pragma unroll
    for (int k = 0; k < kDB; ++k) {
      const T yk = y[k];
#pragma unroll
      for (int r = k + 1; r < kDB; ++r) y[r] = __builtin_fma(-M[r][k], yk, y[r]);
    }
  } else {  // U y = e_lane
#pragma unroll
    for (int k = kDB - 1; k >= 0; --k) {
      const T yk = y[k] * rcp[k];
      y[k] = yk;
#pragma unroll
      for (int r = 0; r < k; ++r) y[r] = __builtin_fma(-M[r][k], yk, y[r]);
    }
  }
  T* out = (t < 64 ? invL : invU) + lane * kDB;
#pragma unroll
  for (int r = 0; r < kDB; ++r) out[r] = y[r];
}

// Partial-pivoting LU of tile rows [0, rows) x columns [0, kb) in LDS; the
// global row numbers in idx[] follow the interchanges.  Returns (in every
// thread) the first column with a zero pivot, or -1.
template <typename T>
__device__ int tile_gepp(T (*tile)[kNB + 1], int* idx, int rows, int kb, T* s_v, int* s_i) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int steps = min(kb, rows);
  int zero = -1;
  for (int j = 0; j < steps; ++j) {
    T v = (t >= j && t < rows) ? fabs(tile[t][j]) : T(-1);
    int bi = t;
    for (int off = 32; off > 0; off >>= 1) {
      const T ov = __shfl_down(v, off);
      const int oi = __shfl_down(bi, off);
      if (ov > v || (ov == v && oi < bi)) {
        v = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      s_v[w] = v;
      s_i[w] = bi;
    }
    __syncthreads();
    T best = s_v[0];
    int p = s_i[0];
    if (s_v[1] > best || (s_v[1] == best && s_i[1] < p)) {
      best = s_v[1];
      p = s_i[1];
    }
    if (!(best > T(0))) {
      p = j;
      if (zero < 0) zero = j;
    }
    if (p != j) {
      if (t < kb) {
        const T a = tile[j][t];
        tile[j][t] = tile[p][t];
        tile[p][t] = a;
      }
      if (t == 0) {
        const int a = idx[j];
        idx[j] = idx[p];
        idx[p] = a;
      }
    }
    __syncthreads();
    const T d = tile[j][j];
    if (t > j && t < rows) {
      T l = tile[t][j];
      if (d != T(0)) {
        l = l / d;
        tile[t][j] = l;
      }
#pragma unroll 8
      for (int c = j + 1; c < kb; ++c) tile[t][c] = tile[t][c] - l * tile[j][c];
    }
    __syncthreads();
  }
  return zero;
}

// Pivoted panel, one launch, grid = number of tiles.  cand: nominations per
// tree node; ctr: arrival counters per tree node (zero on entry, reset by
// the workgroup that consumes them).  The root writes the panel's row moves
// (mv: dst, src pairs and their count), moves the displaced rows of the
// panel columns and stores the factored kb x kb block.
template <typename T>
__global__ __launch_bounds__(kTR) void k_tslu(T* A, int lda, int n, int k0, int kb,
                                              int* __restrict__ cand, int* __restrict__ ctr,
                                              int* __restrict__ mv, int* __restrict__ info) {
  __shared__ T tile[kTR][kNB + 1];
  __shared__ int idx[kTR];
  __shared__ T s_v[2];
  __shared__ int s_i[2], s_cnt[2], s_last, s_nm, s_dst[2 * kNB], s_src[2 * kNB];
  const int t = threadIdx.x, lane = t & 63;
  const int m = n - k0;
  int count = cdiv(m, kTR);  // nodes of the current level
  int node = blockIdx.x, level_off = 0, ctr_off = 0;
  int rows;
  {
    const int r0 = k0 + node * kTR;
    rows = min(kTR, n - r0);
    if (t < rows) {
      idx[t] = r0 + t;
      for (int c = 0; c < kb; ++c) tile[t][c] = A[(size_t)(k0 + c) * lda + r0 + t];
    }
    __syncthreads();
  }
  int zero;
  for (;;) {
    zero = tile_gepp(tile, idx, rows, kb, s_v, s_i);
    if (count == 1) break;
    const int nc = min(kb, rows);
    if (t < kNB) cand[(size_t)(level_off + node) * kNB + t] = t < nc ? idx[t] : -1;
    __threadfence();
    __syncthreads();
    const int parent = node / kFan;
    const int kids = min(kFan, count - parent * kFan);
    if (t == 0) {
      const int old = atomicAdd(&ctr[ctr_off + parent], 1);
      s_last = old == kids - 1;
      if (s_last) ctr[ctr_off + parent] = 0;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    // gather the group's nominations (original rows of A), compacted
    int r = -1;
    if (t < kids * kNB)
      r = __atomic_load_n(&cand[(size_t)(level_off + parent * kFan) * kNB + t], __ATOMIC_RELAXED);
    const unsigned long long bal = __ballot(r >= 0);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) s_cnt[t >> 6] = __popcll(bal);
    __syncthreads();
    const int pos = before + ((t >> 6) ? s_cnt[0] : 0);
    rows = s_cnt[0] + s_cnt[1];
    if (r >= 0) {
      idx[pos] = r;
      for (int c = 0; c < kb; ++c) tile[pos][c] = A[(size_t)(k0 + c) * lda + r];
    }
    __syncthreads();
    level_off += count;
    ctr_off += cdiv(count, kFan);
    count = cdiv(count, kFan);
    node = parent;
  }
  // Root.  The panel's row permutation as a list of moves dst <- src
  // (positions in the matrix before this panel): target k0+j receives winner
  // idx[j], and every row displaced from a target by the interchanges of
  // LAPACK's sequential form lands at a vacated winner position.  Wave 0
  // replays the interchanges with the displaced rows as a lane-distributed
  // (row, position) list.
  if (t < 64) {
    int my_id = -1, my_pos = -1, nlist = 0;
    for (int j = 0; j < kb; ++j) {
      const int r = idx[j], target = k0 + j;
      const unsigned long long mr = __ballot(my_id == r);
      const int a = mr ? __ffsll((long long)mr) - 1 : -1;
      const int p = a >= 0 ? __shfl(my_pos, a) : r;
      if (p != target) {
        const unsigned long long mt = __ballot(my_id >= 0 && my_pos == target);
        const int b = mt ? __ffsll((long long)mt) - 1 : -1;
        const int rt = b >= 0 ? __shfl(my_id, b) : target;
        if (b >= 0) {
          if (t == b) my_pos = p;
        } else {
          if (t == nlist) {
            my_id = rt;
            my_pos = p;
          }
          ++nlist;
        }
        if (a >= 0 && t == a) my_pos = target;
      }
    }
    const bool out = my_id >= 0 && (my_pos < k0 || my_pos >= k0 + kb);
    const unsigned long long mo = __ballot(out);
    const int slot = __popcll(mo & ((1ull << t) - 1ull));
    if (out) {
      s_dst[kb + slot] = my_pos;
      s_src[kb + slot] = my_id;
    }
    if (t < kb) {
      s_dst[t] = k0 + t;
      s_src[t] = idx[t];
    }
    if (t == 0) s_nm = kb + __popcll(mo);
  }
  __syncthreads();
  const int nm = s_nm, nd = nm - kb;
  // panel columns: displaced rows move (read all, then write), the targets
  // receive the factored block
  constexpr int kHeld = (kNB * kNB) / kTR;
  T held[kHeld];
#pragma unroll
  for (int u = 0; u < kHeld; ++u) {
    const int q = t + u * kTR;
    if (q < nd * kb) held[u] = A[(size_t)(k0 + q / nd) * lda + s_src[kb + q % nd]];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kHeld; ++u) {
    const int q = t + u * kTR;
    if (q < nd * kb) A[(size_t)(k0 + q / nd) * lda + s_dst[kb + q % nd]] = held[u];
  }
  for (int q = t; q < kb * kb; q += kTR) {
    const int j = q % kb, c = q / kb;
    A[(size_t)(k0 + c) * lda + k0 + j] = tile[j][c];
  }
  if (t < nm) {
    mv[2 * t] = s_dst[t];
    mv[2 * t + 1] = s_src[t];
  }
  if (t == 0) {
    mv[4 * kNB] = nm;
    if (zero >= 0) atomicCAS(info, 0, k0 + zero + 1);
  }
}

// L21 = A21 U11^-1 for the rows [k0+kb, n) of a pivoted panel: one lane per
// row, column sweep in LDS.
template <typename T>
__global__ __launch_bounds__(64) void k_tslu_l21(T* A, int lda, int n, int k0, int kb) {
  __shared__ T U[kNB][kNB + 1];
  __shared__ T X[kNB][64 + 1];
  const int t = threadIdx.x;
  for (int q = t; q < kNB * kNB; q += 64) {
    const int r = q % kNB, c = q / kNB;
    U[r][c] = (r < kb && c < kb) ? A[(size_t)(k0 + c) * lda + k0 + r] : T(0);
  }
  const int row = k0 + kb + blockIdx.x * 64 + t;
  if (row < n)
    for (int c = 0; c < kb; ++c) X[c][t] = A[(size_t)(k0 + c) * lda + row];
  __syncthreads();
  if (row >= n) return;
  for (int j = 0; j < kb; ++j) {
    const T d = U[j][j];
    T xj = X[j][t];
    if (d != T(0)) xj = xj / d;
    X[j][t] = xj;
#pragma unroll 8
    for (int c = j + 1; c < kb; ++c) X[c][t] = X[c][t] - xj * U[j][c];
  }
  for (int c = 0; c < kb; ++c) A[(size_t)(k0 + c) * lda + row] = X[c][t];
}

// Zero the pivot report, the fallback flag and the tree counters; perm =
// identity.
__global__ void k_lu_init(int* info, int* ctr, int nctr, int* perm, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    info[0] = 0;
    info[1] = 0;
  }
  if (i < nctr) ctr[i] = 0;
  if (i < n) perm[i] = i;
}

// The panel's row moves applied to the columns outside it and to the row
// permutation perm (one extra column): 64 lanes per column, one move each;
// every source is read before any destination is written.
template <typename T>
__global__ __launch_bounds__(256) void k_apply_moves(T* A, int lda, int n, int k0, int kb,
                                                     const int* __restrict__ mv,
                                                     int* __restrict__ perm) {
  const int e = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nm = mv[4 * kNB];
  const bool act = c <= n - kb && e < nm;
  const int dst = act ? mv[2 * e] : 0, src = act ? mv[2 * e + 1] : 0;
  if (c == n - kb) {
    const int v = act ? perm[src] : 0;
    __syncthreads();
    if (act) perm[dst] = v;
    return;
  }
  T* a = A + (size_t)(c < k0 ? c : c + kb) * lda;
  const T v = act ? a[src] : T(0);
  __syncthreads();
  if (act) a[dst] = v;
}

// U12 = L11^-1 A12 for the columns [k0+kb, n): one lane per column held in
// registers, L11 in LDS (zero beyond kb).
template <typename T>
__global__ __launch_bounds__(64) void k_trsm(T* A, int lda, int n, int k0, int kb) {
  __shared__ T L[kNB][kNB + 1];
  const int t = threadIdx.x;
  for (int q = t; q < kNB * kNB; q += 64) {
    const int r = q % kNB, c = q / kNB;
    L[r][c] = (r < kb && c < kb) ? A[(size_t)(k0 + c) * lda + k0 + r] : T(0);
  }
  const int col = k0 + kb + blockIdx.x * 64 + t;
  T* a = A + (size_t)col * lda + k0;
  T x[kNB];
#pragma unroll
  for (int i = 0; i < kNB; ++i) x[i] = (col < n && i < kb) ? a[i] : T(0);
  __syncthreads();
  if (col >= n) return;
  const T* Lb = &L[0][0];
#pragma unroll
  for (int j = 0; j < kNB; ++j) {
    // column offset through an empty asm that also takes x[j - 1]: the reads
    // of column j are issued once step j - 1 starts, not all at the top
    int o = j;
    asm volatile("" : "+s"(o) : "v"(x[j > 0 ? j - 1 : 0]));
    const T xj = x[j];
#pragma unroll
    for (int i = j + 1; i < kNB; ++i) x[i] = x[i] - Lb[i * (kNB + 1) + o] * xj;
  }
#pragma unroll
  for (int i = 0; i < kNB; ++i)
    if (i < kb) a[i] = x[i];
}

// x = U^-1 L^-1 (P b) through the inverses of the 64 x 64 diagonal blocks,
// one launch per group of 256 rows.  Forward launch for the group [j0, j0 +
// 256): every workgroup solves the group itself in LDS (x_s = L_ss^-1 w_s
// for its four blocks in turn, each applied to the group's later rows),
// workgroup 0 stores the group's solution in y, and each workgroup subtracts
// L_(rows, group) x_group from w at its 64 rows below the group (four
// partial sums per row over 64 columns each, combined in a fixed order).
// Backward launches go up from the last group with U.  A launch updates
// only rows outside the group it reads and writes the group's solution to
// another vector, so no workgroup reads what another writes in the same
// launch (forward: w = tmp, y = out; backward: w = out, y = tmp, then copied
// to out).  Loads use clamped in-bounds addresses and zero weights instead
// of branches, so they are all in flight together.
constexpr int kGB = 256;

template <typename T>
__global__ void k_permute(const T* __restrict__ b, const int* __restrict__ perm, int n,
                          T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = b[perm[i]];
}

template <typename T>
__global__ void k_copy_t(const T* __restrict__ a, int n, T* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// LU column `col` (clamped to n - 1) at row `row` (clamped)
template <typename T>
__device__ __forceinline__ T lu_at(const T* __restrict__ LU, int n, int row, int col) {
  return LU[(size_t)min(col, n - 1) * n + min(row, n - 1)];
}

// sum_{c < N} LU(row, col0 + c) wl[c]: loads staged 32 at a time, four
// partial sums combined in a fixed order.
template <typename T, int N>
__device__ __forceinline__ T dot_cols(const T* __restrict__ LU, int n, int row, int col0,
                                      const T* wl) {
  T acc[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int h = 0; h < N; h += 32) {
    T v[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) v[c] = lu_at(LU, n, row, col0 + h + c);
#pragma unroll
    for (int c = 0; c < 32; ++c) acc[c & 3] = __builtin_fma(v[c], wl[h + c], acc[c & 3]);
  }
  return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

// One block of the group: wl[s0 .. s0 + 64) <- Dinv wl[s0 .. s0 + 64) (the
// block's inverse, column-major, ld 64, zero-padded like wl).
template <typename T>
__device__ __forceinline__ void group_block(const T* __restrict__ Dinv, T* wl, int s0,
                                            T (*part)[64]) {
  const int t = threadIdx.x, r = t & 63, q = t >> 6;
  T v[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) v[c] = Dinv[(16 * q + c) * kDB + r];
  T acc[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int c = 0; c < 16; ++c) acc[c & 3] = __builtin_fma(v[c], wl[s0 + 16 * q + c], acc[c & 3]);
  part[q][r] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  if (t < 64) wl[s0 + t] = ((part[0][t] + part[1][t]) + part[2][t]) + part[3][t];
  __syncthreads();
}

// w[row0 + r] -= sum over the group's gb columns of LU(row0 + r, j0 + c) wl[c]
// for r < 64 and row0 + r < lim.
template <typename T>
__device__ __forceinline__ void group_rows(const T* __restrict__ LU, int n, int j0, int gb,
                                           int row0, int lim, const T* wl, T* w,
                                           T (*part)[64]) {
  const int t = threadIdx.x, r = t & 63, q = t >> 6;
  T s = T(0);
  if (64 * q < gb) s = dot_cols<T, 64>(LU, n, row0 + r, j0 + 64 * q, wl + 64 * q);
  part[q][r] = s;
  __syncthreads();
  if (t < 64 && row0 + t < lim)
    w[row0 + t] = w[row0 + t] - (((part[0][t] + part[1][t]) + part[2][t]) + part[3][t]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_fwd(const T* __restrict__ LU, int n,
                                             const T* __restrict__ invL, T* w, T* y, int j0) {
  __shared__ T wl[kGB];
  __shared__ T part[4][64];
  const int t = threadIdx.x;
  const int gb = min(kGB, n - j0);
  wl[t] = t < gb ? w[j0 + t] : T(0);
  __syncthreads();
  for (int s0 = 0; s0 < gb; s0 += kDB) {
    group_block(invL + (size_t)(j0 + s0) * kDB, wl, s0, part);
    // the group's later rows
    const int row = s0 + kDB + t;  // row within the group
    T acc = T(0);
    if (s0 + kDB < gb) acc = dot_cols<T, kDB>(LU, n, j0 + row, j0 + s0, wl + s0);
    __syncthreads();
    if (row < gb) wl[row] = wl[row] - acc;
    __syncthreads();
  }
  if (blockIdx.x == 0 && t < gb) y[j0 + t] = wl[t];
  group_rows(LU, n, j0, gb, j0 + gb + 64 * blockIdx.x, n, wl, w, part);
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd(const T* __restrict__ LU, int n,
                                             const T* __restrict__ invU, T* w, T* y, int j0) {
  __shared__ T wl[kGB];
  __shared__ T part[4][64];
  const int t = threadIdx.x;
  const int gb = min(kGB, n - j0);
  wl[t] = t < gb ? w[j0 + t] : T(0);
  __syncthreads();
  for (int s0 = (gb - 1) / kDB * kDB; s0 >= 0; s0 -= kDB) {
    group_block(invU + (size_t)(j0 + s0) * kDB, wl, s0, part);
    // the group's earlier rows
    T acc = T(0);
    if (s0 > 0) acc = dot_cols<T, kDB>(LU, n, j0 + t, j0 + s0, wl + s0);
    __syncthreads();
    if (t < s0) wl[t] = wl[t] - acc;
    __syncthreads();
  }
  if (blockIdx.x == 0 && t < gb) y[j0 + t] = wl[t];
  group_rows(LU, n, j0, gb, 64 * blockIdx.x, j0, wl, w, part);
}

// inv: the diagonal blocks' inverses (L^-1 blocks, then U^-1 blocks, each
// cdiv(n, 64) blocks of 64 x 64).  tmp: n values of scratch (distinct from b
// and out).
template <typename T>
void launch_getrs(hipStream_t s, const T* LU, int n, const int* perm, const T* inv,
                  const T* b, T* out, T* tmp) {
  if (n <= 0) return;
  const T* invL = inv;
  const T* invU = inv + (size_t)cdiv(n, kDB) * kDB * kDB;
  hipLaunchKernelGGL((k_permute<T>), dim3(cdiv(n, 256)), dim3(256), 0, s, b, perm, n, tmp);
  for (int j0 = 0; j0 < n; j0 += kGB) {
    const int rest = n - j0 - std::min(kGB, n - j0);
    hipLaunchKernelGGL((k_fwd<T>), dim3(std::max(1, cdiv(rest, 64))), dim3(256), 0, s, LU, n,
                       invL, tmp, out, j0);
  }
  for (int j0 = (n - 1) / kGB * kGB; j0 >= 0; j0 -= kGB)
    hipLaunchKernelGGL((k_bwd<T>), dim3(std::max(1, cdiv(j0, 64))), dim3(256), 0, s, LU, n,
                       invU, out, tmp, j0);
  hipLaunchKernelGGL((k_copy_t<T>), dim3(cdiv(n, 256)), dim3(256), 0, s, (const T*)tmp, n, out);
}

// Workspace of the fallback factorisation (ints): move list of the current
// panel, tree nominations and arrival counters of the panel kernel.
size_t lu_work_ints(int n) {
  const int tiles = cdiv(n, kTR);
  return 4 * kNB + 64 + (size_t)2 * tiles * kNB + 2 * (size_t)tiles + 64;
}

// Elements of the diagonal blocks' inverses: the L^-1 blocks, then the U^-1
// blocks, 64 x 64 each, then the fast path's factored diagonal blocks.
size_t lu_inv_elems(int n) { return 3 * (size_t)cdiv(n, kDB) * kDB * kDB; }

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// A22 -= L21 U12 after the panel [k0, k0 + kb).
template <typename T>
void lu_trailing(hipStream_t s, int prec, T* A, int n, int k0, int kb) {
  const int rest = n - k0 - kb;
  const GemmArgs u{rest, rest, kb, A + (size_t)k0 * n + k0 + kb, n,
                   A + (size_t)(k0 + kb) * n + k0, n, A + (size_t)(k0 + kb) * n + k0 + kb, n,
                   -1.0, 1.0, nullptr};
  launch_gemm(s, prec, &u, 1);
}

// Blocked right-looking LU.  prec: GEMM_F64 (T = double), or GEMM_F32 /
// GEMM_BF16 (T = float; the trailing updates with FP32 or BF16 operands,
// the panels in FP32).  pivoted = false: the checked
// no-pivot path (info[1] <- 1 when a check fails: the factors are then not
// GEPP's and the caller refactors with pivoted = true).  perm = the row
// permutation; inv <- the diagonal blocks' inverses (lu_inv_elems).
template <typename T>
void lu_factor_t(hipStream_t s, int prec, T* A, int n, int* perm, int* info, int* work, T* inv,
                 bool pivoted) {
  const int tiles = cdiv(n, kTR);
  int* mv = work;
  int* cand = work + 4 * kNB + 64;
  int* ctr = cand + (size_t)2 * tiles * kNB;
  const int nctr = 2 * tiles;
  T* invL = inv;
  T* invU = inv + (size_t)cdiv(n, kDB) * kDB * kDB;
  T* dblk = invU + (size_t)cdiv(n, kDB) * kDB * kDB;  // the fast path's diagonal blocks
  hipLaunchKernelGGL(k_lu_init, dim3(cdiv(std::max(n, nctr), 256)), dim3(256), 0, s, info,
                     ctr, nctr, perm, n);
  if (!pivoted) {
    for (int k0 = 0; k0 < n; k0 += kDB) {
      const int kb = std::min(kDB, n - k0), rest = n - k0 - kb;
      const int nb = cdiv(rest, kDB);
      hipLaunchKernelGGL((k_panel64<T>), dim3(std::max(1, 2 * nb)), dim3(256), 0, s, A, n, n, k0,
                         nb, dblk + (size_t)k0 * kDB, info + 1, info);
      if (rest > 0) lu_trailing(s, prec, A, n, k0, kb);
    }
  } else {
    for (int k0 = 0; k0 < n; k0 += kNB) {
      const int kb = std::min(kNB, n - k0);
      const int rest = n - k0 - kb;
      hipLaunchKernelGGL((k_tslu<T>), dim3(cdiv(n - k0, kTR)), dim3(kTR), 0, s, A, n, n, k0, kb,
                         cand, ctr, mv, info);
      if (rest > 0)
        hipLaunchKernelGGL((k_tslu_l21<T>), dim3(cdiv(rest, 64)), dim3(64), 0, s, A, n, n, k0,
                           kb);
      hipLaunchKernelGGL((k_apply_moves<T>), dim3(cdiv(n - kb + 1, 4)), dim3(256), 0, s, A, n, n,
                         k0, kb, (const int*)mv, perm);
      if (rest > 0) {
        hipLaunchKernelGGL((k_trsm<T>), dim3(cdiv(rest, 64)), dim3(64), 0, s, A, n, n, k0, kb);
        lu_trailing(s, prec, A, n, k0, kb);
      }
    }
  }
  hipLaunchKernelGGL((k_diag_inv<T>), dim3(cdiv(n, kDB)), dim3(128), 0, s, A, n, n,
                     pivoted ? (const T*)nullptr : (const T*)dblk, invL, invU);
}

// Factor with the checked no-pivot path; if a check fails, restore A from
// `saved` and factor with tournament pivoting.  Synchronises the stream once
// (to read the check).  Returns 1 when pivoting was needed, -1 on a HIP
// error.
template <typename T>
int lu_factor_checked(hipStream_t s, int prec, T* A, int n, int* perm, int* info, int* work,
                      T* inv, T* saved) {
  const size_t bytes = sizeof(T) * (size_t)n * n;
  if (hipMemcpyAsync(saved, A, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess) return -1;
  lu_factor_t<T>(s, prec, A, n, perm, info, work, inv, false);
  int flag = 0;
  if (hipMemcpyAsync(&flag, info + 1, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return -1;
  if (!flag) return 0;
  if (hipMemcpyAsync(A, saved, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess) return -1;
  lu_factor_t<T>(s, prec, A, n, perm, info, work, inv, true);
  return 1;
}

// --- mixed precision helpers ------------------------------------------------
__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b,
                         size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// r = b - A x in FP64 (column-major A): workgroup (i, s) sums columns
// [s cols, (s + 1) cols) for 256 rows into part[s][row]; k_resid_sum adds
// the slices in increasing s.
__global__ __launch_bounds__(256) void k_resid_part(const double* __restrict__ A, int n,
                                                    const double* __restrict__ x, int cols,
                                                    double* __restrict__ part) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int j0 = blockIdx.y * cols, j1 = min(n, j0 + cols);
  const double* p = A + (size_t)j0 * n + i;
  double acc = 0.0;
  int j = j0;
  for (; j + 4 <= j1; j += 4) {
    const double a0 = p[0], a1 = p[n], a2 = p[2L * n], a3 = p[3L * n];
    acc = __builtin_fma(a0, x[j], acc);
    acc = __builtin_fma(a1, x[j + 1], acc);
    acc = __builtin_fma(a2, x[j + 2], acc);
    acc = __builtin_fma(a3, x[j + 3], acc);
    p += 4L * n;
  }
  for (; j < j1; ++j, p += n) acc = __builtin_fma(*p, x[j], acc);
  part[(size_t)blockIdx.y * n + i] = acc;
}

__global__ __launch_bounds__(256) void k_resid_sum(const double* __restrict__ part, int S,
                                                   int n, const double* __restrict__ b,
                                                   double* __restrict__ r,
                                                   float* __restrict__ rf) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int k = 0; k < S; ++k) acc += part[(size_t)k * n + i];
  const double v = b[i] - acc;
  r[i] = v;
  rf[i] = (float)v;
}

// Column slices of the residual: about 1024 workgroups, at most 64 slices.
inline void resid_split(int n, int* S, int* cols) {
  const int rb = cdiv(n, 256);
  int want = std::min(64, std::max(1, cdiv(1024, rb)));
  *cols = cdiv(n, want);
  *S = cdiv(n, *cols);
}

// x += dx; norms |dx|_inf and |x|_inf for the convergence test.
__global__ __launch_bounds__(1024) void k_update(double* __restrict__ x,
                                                 const float* __restrict__ dx,
                                                 int n, double* __restrict__ norms) {
  __shared__ double s1[16], s2[16];
  double m1 = 0.0, m2 = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) {
    const double v = x[i] + (double)dx[i];
    x[i] = v;
    m1 = fmax(m1, fabs((double)dx[i]));
    m2 = fmax(m2, fabs(v));
  }
  for (int off = 32; off > 0; off >>= 1) {
    m1 = fmax(m1, __shfl_down(m1, off));
    m2 = fmax(m2, __shfl_down(m2, off));
  }
  if ((threadIdx.x & 63) == 0) {
    s1[threadIdx.x >> 6] = m1;
    s2[threadIdx.x >> 6] = m2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 16; ++w) {
      m1 = fmax(m1, s1[w]);
      m2 = fmax(m2, s2[w]);
    }
    norms[0] = s1[0] > m1 ? s1[0] : m1;
    norms[1] = s2[0] > m2 ? s2[0] : m2;
  }
}

__global__ void k_f32_to_f64(const float* __restrict__ a, double* __restrict__ b,
                             int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

// --- full pivoting in the elimination order of Eigen::FullPivLU ------------
// (simple_simulator.cpp:265-268 calls FullPivLU<MatrixXd>(A).solve(rhs)).
// Every entry of the factors is produced by the same single IEEE operation
// as in a sequential elimination (the rank-1 update is elementwise), so with
// identical pivots the factors and the solution are bitwise reproducible.
// The pivot of step k is the first maximum of |m| over the trailing block in
// column-major scan order.  Per step: k_fp_argmax (grid partials),
// k_fp_pivot (reduce, row and column interchange, column scale),
// k_fp_update (rank-1 update); then k_fp_solve in one workgroup.
struct FpState {
  double maxpivot;
  int nonzero, done;
};
constexpr int kFpParts = 512;

__global__ void k_fp_init(FpState* st, int n) {
  st->maxpivot = 0.0;
  st->nonzero = n;
  st->done = 0;
}

__device__ __forceinline__ void fp_better(double& v, long& q, double ov, long oq) {
  if (ov > v || (ov == v && oq < q)) {
    v = ov;
    q = oq;
  }
}

__global__ __launch_bounds__(256) void k_fp_argmax(const double* __restrict__ m, int n, int k,
                                                   const FpState* __restrict__ st,
                                                   double* __restrict__ pv,
                                                   long* __restrict__ pq) {
  if (st->done) return;
  const long w = n - k, total = w * w;
  double best = -1.0;
  long bq = LONG_MAX;
  for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < total; q += (long)gridDim.x * 256) {
    const double a = fabs(m[(size_t)(k + q / w) * n + k + q % w]);
    if (a > best) {
      best = a;
      bq = q;
    }
  }
  for (int off = 32; off > 0; off >>= 1)
    fp_better(best, bq, __shfl_down(best, off), __shfl_down(bq, off));
  __shared__ double sv[4];
  __shared__ long sq[4];
  if ((threadIdx.x & 63) == 0) {
    sv[threadIdx.x >> 6] = best;
    sq[threadIdx.x >> 6] = bq;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w2 = 1; w2 < 4; ++w2) fp_better(best, bq, sv[w2], sq[w2]);
    pv[blockIdx.x] = best;
    pq[blockIdx.x] = bq;
  }
}

__global__ __launch_bounds__(1024) void k_fp_pivot(double* __restrict__ m, int n, int k,
                                                   int parts, FpState* __restrict__ st,
                                                   const double* __restrict__ pv,
                                                   const long* __restrict__ pq,
                                                   int* __restrict__ rt, int* __restrict__ ct) {
  if (st->done) return;
  const int tid = threadIdx.x;
  double best = -1.0;
  long bq = LONG_MAX;
  for (int p = tid; p < parts; p += 1024) fp_better(best, bq, pv[p], pq[p]);
  for (int off = 32; off > 0; off >>= 1)
    fp_better(best, bq, __shfl_down(best, off), __shfl_down(bq, off));
  __shared__ double sv[16];
  __shared__ long sq[16];
  __shared__ int s_br, s_bc, s_stop;
  if ((tid & 63) == 0) {
    sv[tid >> 6] = best;
    sq[tid >> 6] = bq;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w2 = 1; w2 < 16; ++w2) fp_better(best, bq, sv[w2], sq[w2]);
    const long w = n - k;
    s_stop = best == 0.0;
    if (best == 0.0) {
      st->nonzero = k;
      st->done = 1;
    } else {
      if (best > st->maxpivot) st->maxpivot = best;
      if (bq == LONG_MAX) bq = 0;  // no comparable entry (NaN block): keep (k, k)
      s_br = k + (int)(bq % w);
      s_bc = k + (int)(bq / w);
      rt[k] = s_br;
      ct[k] = s_bc;
    }
  }
  __syncthreads();
  if (s_stop) {
    for (int i = k + tid; i < n; i += 1024) rt[i] = ct[i] = i;
    return;
  }
  const int br = s_br, bc = s_bc;
  if (br != k)
    for (int cc = tid; cc < n; cc += 1024) {
      const double t = m[(size_t)cc * n + k];
      m[(size_t)cc * n + k] = m[(size_t)cc * n + br];
      m[(size_t)cc * n + br] = t;
    }
  __syncthreads();
  if (bc != k)
    for (int r = tid; r < n; r += 1024) {
      const double t = m[(size_t)k * n + r];
      m[(size_t)k * n + r] = m[(size_t)bc * n + r];
      m[(size_t)bc * n + r] = t;
    }
  __syncthreads();
  const double piv = m[(size_t)k * n + k];
  for (int r = k + 1 + tid; r < n; r += 1024) m[(size_t)k * n + r] /= piv;
}

__global__ __launch_bounds__(256) void k_fp_update(double* __restrict__ m, int n, int k,
                                                   const FpState* __restrict__ st) {
  if (st->done) return;
  const long w = n - k - 1, total = w * w;
  for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < total; q += (long)gridDim.x * 256) {
    const int cc = k + 1 + (int)(q / w), r = k + 1 + (int)(q % w);
    const double u = m[(size_t)cc * n + k];
    m[(size_t)cc * n + r] -= m[(size_t)k * n + r] * u;
  }
}

// Rank with FullPivLU's default threshold, then P, unit-lower, upper on the
// rank x rank block, Q -- in the order of a sequential solve.
__global__ __launch_bounds__(1024) void k_fp_solve(const double* __restrict__ m, int n,
                                                   const FpState* __restrict__ st,
                                                   const int* __restrict__ rt,
                                                   const int* __restrict__ ct,
                                                   const double* __restrict__ b,
                                                   double* __restrict__ c, int* rank_out) {
  const int tid = threadIdx.x;
  const double thr = st->maxpivot * (2.220446049250313e-16 * (double)n);
  int cnt = 0;
  for (int k = tid; k < st->nonzero; k += 1024) cnt += fabs(m[(size_t)k * n + k]) > thr;
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off);
  __shared__ int s_cnt[16];
  __shared__ int s_rank;
  if ((tid & 63) == 0) s_cnt[tid >> 6] = cnt;
  for (int i = tid; i < n; i += 1024) c[i] = b[i];
  __syncthreads();
  if (tid == 0) {
    int r = 0;
    for (int w2 = 0; w2 < 16; ++w2) r += s_cnt[w2];
    s_rank = r;
    if (rank_out) *rank_out = r;
    for (int k = 0; k < n; ++k) {
      const double t = c[k];
      c[k] = c[rt[k]];
      c[rt[k]] = t;
    }
  }
  __syncthreads();
  const int rank = s_rank;
  for (int k = 0; k < n; ++k) {
    const double ck = c[k];
    for (int r = k + 1 + tid; r < n; r += 1024) c[r] -= m[(size_t)k * n + r] * ck;
    __syncthreads();
  }
  for (int kk = rank - 1; kk >= 0; --kk) {
    if (tid == 0) c[kk] /= m[(size_t)kk * n + kk];
    __syncthreads();
    const double ck = c[kk];
    for (int r = tid; r < kk; r += 1024) c[r] -= m[(size_t)kk * n + r] * ck;
    __syncthreads();
  }
  for (int k = rank + tid; k < n; k += 1024) c[k] = 0.0;
  __syncthreads();
  if (tid == 0)
    for (int k = n - 1; k >= 0; --k) {
      const double t = c[k];
      c[k] = c[ct[k]];
      c[ct[k]] = t;
    }
}

}  // namespace

// LU workspace: ints [info, flag | pad 64][perm n][moves, tree], then an
// n x n copy of A for the pivoted retry, then the diagonal blocks' inverses.
static size_t lu_ints(int n) { return 64 + (size_t)n + lu_work_ints(n); }
size_t dense_lu_work_bytes(int n) {
  return align256(sizeof(int) * lu_ints(n)) + align256(sizeof(double) * (size_t)n * n) +
         sizeof(double) * lu_inv_elems(n);
}
int* dense_lu_info(void* work) { return (int*)work; }

static double* lu_saved(void* work, int n) {
  return (double*)((char*)work + align256(sizeof(int) * lu_ints(n)));
}
static double* lu_inv(void* work, int n) {
  return (double*)((char*)lu_saved(work, n) + align256(sizeof(double) * (size_t)n * n));
}

int dense_lu_factor(hipStream_t s, double* A, int n, void* work) {
  int* w = (int*)work;
  return lu_factor_checked<double>(s, GEMM_F64, A, n, w + 64, w, w + 64 + (size_t)n,
                                   lu_inv(work, n), lu_saved(work, n));
}

void dense_lu_solve(hipStream_t s, const double* LU, int n, void* work, const double* b,
                    double* x) {
  const int* w = (const int*)work;
  // the copy of A kept for a pivoted retry is free once A is factored
  launch_getrs<double>(s, LU, n, w + 64, lu_inv(work, n), b, x, lu_saved(work, n));
}

size_t dense_mixed_work_bytes(int n) {
  const size_t nn = (size_t)n * n;
  return 2 * align256(nn * sizeof(float)) + align256(sizeof(int) * lu_ints(n)) +
         (size_t)n * sizeof(double) + 64 * sizeof(double) + 2 * (size_t)n * sizeof(float) +
         256 + align256(64 * (size_t)n * sizeof(double)) + sizeof(float) * lu_inv_elems(n);
}

// Mixed precision: LU of fl32(A) (FP32 or BF16-operand MFMA updates), then
// x_{k+1} = x_k + LU^-1 (b - A x_k) with the residual in FP64.
int dense_mixed_solve(hipStream_t s, const double* A, int n, const double* b,
                      double* x, int mode, double tol, int max_iter, void* work,
                      size_t work_bytes, int* iters) {
  if (work_bytes < dense_mixed_work_bytes(n)) return -1;
  char* w = (char*)work;
  const size_t nn = (size_t)n * n;
  float* Af = (float*)w;
  w += align256(nn * sizeof(float));
  float* Asaved = (float*)w;
  w += align256(nn * sizeof(float));
  int* lw = (int*)w;
  w += align256(sizeof(int) * lu_ints(n));
  double* r = (double*)w;
  w += (size_t)n * sizeof(double);
  double* norms = (double*)w;
  w += 64 * sizeof(double);
  float* rf = (float*)w;
  w += (size_t)n * sizeof(float);
  float* dxf = (float*)w;
  w += (size_t)n * sizeof(float);
  w = (char*)(((uintptr_t)w + 255) & ~(uintptr_t)255);
  double* part = (double*)w;  // residual slices [S][n], 256-byte aligned
  w += align256(64 * (size_t)n * sizeof(double));
  float* inv = (float*)w;
  int S = 1, cols = n;
  resid_split(n, &S, &cols);
  int* perm = lw + 64;
  int* tree = lw + 64 + (size_t)n;
  hipLaunchKernelGGL(k_to_f32, dim3(cdiv(nn, 256)), dim3(256), 0, s, A, Af, nn);
  const int rc = lu_factor_checked<float>(s, mode == 2 ? GEMM_BF16 : GEMM_F32, Af, n, perm, lw,
                                          tree, inv, Asaved);
  if (rc < 0) return -2;
  (void)hipMemsetAsync(x, 0, sizeof(double) * n, s);
  int it = 0;
  double h_norms[2] = {0, 0};
  for (it = 0; it < max_iter; ++it) {
    hipLaunchKernelGGL(k_resid_part, dim3(cdiv(n, 256), S), dim3(256), 0, s, A, n,
                       (const double*)x, cols, part);
    hipLaunchKernelGGL(k_resid_sum, dim3(cdiv(n, 256)), dim3(256), 0, s, (const double*)part,
                       S, n, b, r, rf);
    launch_getrs<float>(s, Af, n, perm, inv, rf, dxf, Asaved);  // Asaved is free after the LU
    hipLaunchKernelGGL(k_update, dim3(1), dim3(1024), 0, s, x, (const float*)dxf,
                       n, norms);
    if (hipMemcpyAsync(h_norms, norms, sizeof(h_norms), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return -2;
    if (!(h_norms[0] <= tol * h_norms[1])) continue;
    ++it;
    break;
  }
  if (iters) *iters = it;
  if (!(h_norms[0] <= tol * h_norms[1])) return 1;
  return 0;
}

size_t dense_fullpiv_work_bytes(int n) {
  return 256 + kFpParts * (sizeof(double) + sizeof(long)) + 2 * (size_t)n * sizeof(int) + 64;
}

// A (column-major, destroyed) x = b with the elimination order of
// Eigen::FullPivLU; rank_dev (nullable) receives the numerical rank.
int dense_fullpiv_solve(hipStream_t s, double* A, int n, const double* b, double* x,
                        void* work, size_t work_bytes, int* rank_dev) {
  if (work_bytes < dense_fullpiv_work_bytes(n)) return -1;
  char* w = (char*)work;
  FpState* st = (FpState*)w;
  w += 256;
  double* pv = (double*)w;
  w += kFpParts * sizeof(double);
  long* pq = (long*)w;
  w += kFpParts * sizeof(long);
  int* rt = (int*)w;
  w += (size_t)n * sizeof(int);
  int* ct = (int*)w;
  hipLaunchKernelGGL(k_fp_init, dim3(1), dim3(1), 0, s, st, n);
  for (int k = 0; k < n; ++k) {
    const long wk = n - k;
    const int parts = (int)std::min<long>(kFpParts, cdiv(wk * wk, 256));
    hipLaunchKernelGGL(k_fp_argmax, dim3(parts), dim3(256), 0, s, (const double*)A, n, k,
                       (const FpState*)st, pv, pq);
    hipLaunchKernelGGL(k_fp_pivot, dim3(1), dim3(1024), 0, s, A, n, k, parts, st,
                       (const double*)pv, (const long*)pq, rt, ct);
    const long wu = n - k - 1;
    if (wu > 0)
      hipLaunchKernelGGL(k_fp_update, dim3((int)std::min<long>(2048, cdiv(wu * wu, 256))),
                         dim3(256), 0, s, A, n, k, (const FpState*)st);
  }
  hipLaunchKernelGGL(k_fp_solve, dim3(1), dim3(1024), 0, s, (const double*)A, n,
                     (const FpState*)st, (const int*)rt, (const int*)ct, b, x, rank_dev);
  return 0;
}

}  // namespace uvlm
