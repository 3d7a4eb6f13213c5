// dense.hip -- dense circulation solve on gfx950.
//
// Replaces Eigen::FullPivLU (simple_simulator.cpp:263-268).  The influence
// matrix of a vortex lattice is diagonally dominant, so the GPU factors with
// partial (row) pivoting; the solution agrees with full pivoting to
// ~cond(A) * eps.
//
// Right-looking blocked LU, panel width kNB = 32, column-major:
//   k_getf2   panel factorisation by one 1024-thread workgroup
//   k_laswp   row interchanges applied to the columns outside the panel
//   k_trsm    U12 = L11^-1 A12 (one lane per column, L11 and X in LDS)
//   k_gemm    A22 -= L21 U12 on the FP64 matrix cores
//             (v_mfma_f64_16x16x4f64, 64x64 tile per workgroup)
// followed by k_getrs (pivots, unit-lower and upper substitution) in one
// workgroup.  The mixed-precision modes factor a copy in FP32 (or with
// BF16-rounded MFMA inputs) and refine the FP64 solution with FP64
// residuals r = b - A x.
#include <algorithm>
#include <climits>
#include <cmath>

#include "internal.h"

namespace uvlm {
namespace {

constexpr int kNB = 32;
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

__global__ void k_zero_int(int* p) { *p = 0; }

// Panel factorisation of columns [k0, k0+kb), rows [k0, n).
template <typename T>
__global__ __launch_bounds__(1024) void k_getf2(T* A, int lda, int n, int k0,
                                               int kb, int* ipiv, int* info) {
  __shared__ T s_val[16];
  __shared__ int s_idx[16];
  __shared__ int s_piv;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  for (int j = 0; j < kb; ++j) {
    const int col = k0 + j;
    T* Acol = A + (size_t)col * lda;
    T best = T(-1);
    int bi = n;
    for (int r = col + tid; r < n; r += 1024) {
      const T v = fabs(Acol[r]);
      if (v > best) {
        best = v;
        bi = r;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const T ov = __shfl_down(best, off);
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
      T b = s_val[0];
      int bj = s_idx[0];
      for (int w = 1; w < 16; ++w)
        if (s_val[w] > b || (s_val[w] == b && s_idx[w] < bj)) {
          b = s_val[w];
          bj = s_idx[w];
        }
      if (!(b > T(0)) || bj >= n) {
        if (*info == 0) *info = col + 1;
        bj = col;
      }
      s_piv = bj;
      ipiv[col] = bj;
    }
    __syncthreads();
    const int piv = s_piv;
    if (piv != col && tid < kb) {
      T* a = A + (size_t)(k0 + tid) * lda;
      const T tmp = a[col];
      a[col] = a[piv];
      a[piv] = tmp;
    }
    __syncthreads();
    const T d = Acol[col];
    if (d != T(0))
      for (int r = col + 1 + tid; r < n; r += 1024) Acol[r] = Acol[r] / d;
    __syncthreads();
    const int ncols = k0 + kb - col - 1;
    const int nrows = n - col - 1;
    if (ncols > 0 && nrows > 0) {
      for (int c = 0; c < ncols; ++c) {
        T* Ac = A + (size_t)(col + 1 + c) * lda;
        const T u = Ac[col];
        for (int r = col + 1 + tid; r < n; r += 1024) Ac[r] = Ac[r] - Acol[r] * u;
      }
    }
    __syncthreads();
  }
}

// Row interchanges of the panel applied to the columns outside it.
template <typename T>
__global__ void k_laswp(T* A, int lda, int n, int k0, int kb,
                        const int* ipiv) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n - kb) return;
  const int col = c < k0 ? c : c + kb;
  T* a = A + (size_t)col * lda;
  for (int j = 0; j < kb; ++j) {
    const int r = k0 + j, p = ipiv[r];
    if (p != r) {
      const T t = a[r];
      a[r] = a[p];
      a[p] = t;
    }
  }
}

// U12 = L11^-1 A12 for the columns [k0+kb, n): one wavefront per 64
// columns, one lane per column; L11 and the 64 partial solutions in LDS
// (a register-resident solution vector would need a fully unrolled 32x32
// triangle, which the compiler spills).
template <typename T>
__global__ __launch_bounds__(64) void k_trsm(T* A, int lda, int n, int k0,
                                             int kb) {
  __shared__ T L[kNB][kNB + 1];
  __shared__ T X[kNB][64 + 1];
  for (int q = threadIdx.x; q < kNB * kNB; q += 64) {
    const int r = q % kNB, c = q / kNB;
    L[r][c] = (r < kb && c < kb) ? A[(size_t)(k0 + c) * lda + k0 + r] : T(0);
  }
  const int t = threadIdx.x;
  const int col = k0 + kb + blockIdx.x * 64 + t;
  T* a = A + (size_t)col * lda + k0;
  if (col < n)
    for (int i = 0; i < kb; ++i) X[i][t] = a[i];
  __syncthreads();
  if (col >= n) return;
  for (int j = 0; j < kb; ++j) {
    const T xj = X[j][t];
    for (int i = j + 1; i < kb; ++i) X[i][t] = X[i][t] - L[i][j] * xj;
  }
  for (int i = 0; i < kb; ++i) a[i] = X[i][t];
}

// A22 -= L21 * U12 with the FP64 matrix cores.  The product is formed
// transposed (D^T = U12^T L21^T) so that each group of 16 lanes holds 16
// consecutive rows of one column: stores are contiguous in column-major A.
// v_mfma_f64_16x16x4f64 operand layout: a[lane] = X[lane%16][lane/16],
// b[lane] = Y[lane/16][lane%16], d[r] = D[4*r+lane/16][lane%16] (the
// result rows interleave across the four lane groups; the f32 variant
// below keeps them blocked: d[r] = D[4*(lane/16)+r][lane%16]).
__global__ __launch_bounds__(256) void k_gemm_f64(double* A, int lda, int n,
                                                  int k0, int kb) {
  __shared__ double sA[64][kNB + 1];  // L21 tile: rows x k
  __shared__ double sB[kNB][64 + 1];  // U12 tile: k x cols
  const int s0 = k0 + kb;
  const int m0 = s0 + blockIdx.y * 64;
  const int c0 = s0 + blockIdx.x * 64;
  for (int q = threadIdx.x; q < 64 * kNB; q += 256) {
    const int r = q % 64, k = q / 64;
    sA[r][k] = (m0 + r < n && k < kb) ? A[(size_t)(k0 + k) * lda + m0 + r] : 0.0;
  }
  for (int q = threadIdx.x; q < kNB * 64; q += 256) {
    const int k = q % kNB, c = q / kNB;
    sB[k][c] = (c0 + c < n && k < kb) ? A[(size_t)(c0 + c) * lda + k0 + k] : 0.0;
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
        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[j], lb[i], acc[i][j],
                                                          0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wr * 32 + i * 16 + l16;
        const int col = c0 + wc * 32 + j * 16 + 4 * r + lq;
        if (row < n && col < n) {
          double* p = A + (size_t)col * lda + row;
          *p = *p - acc[i][j][r];
        }
      }
}

// FP32 trailing update with the FP32 matrix cores (v_mfma_f32_16x16x4f32,
// same operand layout as the f64 variant for one 16x16 block), or with
// BF16-rounded operands when BF16 is set (inputs rounded to bfloat16,
// products accumulated in FP32: the precision of v_mfma_f32_*_bf16).
__device__ __forceinline__ float round_bf16(float x) {
  unsigned u = __float_as_uint(x);
  const unsigned lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  u &= 0xffff0000u;
  return __uint_as_float(u);
}

template <bool BF16>
__global__ __launch_bounds__(256) void k_gemm_f32(float* A, int lda, int n,
                                                  int k0, int kb) {
  __shared__ float sA[64][kNB + 1];
  __shared__ float sB[kNB][64 + 1];
  const int s0 = k0 + kb;
  const int m0 = s0 + blockIdx.y * 64;
  const int c0 = s0 + blockIdx.x * 64;
  for (int q = threadIdx.x; q < 64 * kNB; q += 256) {
    const int r = q % 64, k = q / 64;
    float v = (m0 + r < n && k < kb) ? A[(size_t)(k0 + k) * lda + m0 + r] : 0.f;
    sA[r][k] = BF16 ? round_bf16(v) : v;
  }
  for (int q = threadIdx.x; q < kNB * 64; q += 256) {
    const int k = q % kNB, c = q / kNB;
    float v = (c0 + c < n && k < kb) ? A[(size_t)(c0 + c) * lda + k0 + k] : 0.f;
    sB[k][c] = BF16 ? round_bf16(v) : v;
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wr = w >> 1, wc = w & 1;
  const int l16 = lane & 15, lq = lane >> 4;
  v4f acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < kNB; k += 4) {
    float ua[2], lb[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) ua[j] = sB[k + lq][wc * 32 + j * 16 + l16];
#pragma unroll
    for (int i = 0; i < 2; ++i) lb[i] = sA[wr * 32 + i * 16 + l16][k + lq];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ua[j], lb[i], acc[i][j],
                                                          0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wr * 32 + i * 16 + l16;
        const int col = c0 + wc * 32 + j * 16 + 4 * lq + r;
        if (row < n && col < n) {
          float* p = A + (size_t)col * lda + row;
          *p = *p - acc[i][j][r];
        }
      }
}

// Forward (unit lower) and backward (upper) substitution after applying the
// row interchanges, one 1024-thread workgroup, 64-row blocks.
template <typename T>
__global__ __launch_bounds__(1024) void k_getrs(const T* LU, int n,
                                               const int* ipiv, T* b) {
  __shared__ T xs[64];
  const int tid = threadIdx.x;
  if (tid == 0)
    for (int i = 0; i < n; ++i) {
      const int p = ipiv[i];
      if (p != i) {
        const T t = b[i];
        b[i] = b[p];
        b[p] = t;
      }
    }
  __syncthreads();
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int ib = min(64, n - i0);
    if (tid < ib) xs[tid] = b[i0 + tid];
    __syncthreads();
    for (int i = 0; i < ib; ++i) {
      if (tid > i && tid < ib)
        xs[tid] = xs[tid] - LU[(size_t)(i0 + i) * n + i0 + tid] * xs[i];
      __syncthreads();
    }
    if (tid < ib) b[i0 + tid] = xs[tid];
    for (int r = i0 + ib + tid; r < n; r += 1024) {
      T s = b[r];
      for (int i = 0; i < ib; ++i) s = s - LU[(size_t)(i0 + i) * n + r] * xs[i];
      b[r] = s;
    }
    __syncthreads();
  }
  for (int iend = n; iend > 0; iend -= 64) {
    const int i0 = max(0, iend - 64);
    const int ib = iend - i0;
    if (tid < ib) xs[tid] = b[i0 + tid];
    __syncthreads();
    for (int i = ib - 1; i >= 0; --i) {
      if (tid == i) xs[i] = xs[i] / LU[(size_t)(i0 + i) * n + i0 + i];
      __syncthreads();
      if (tid < i) xs[tid] = xs[tid] - LU[(size_t)(i0 + i) * n + i0 + tid] * xs[i];
      __syncthreads();
    }
    if (tid < ib) b[i0 + tid] = xs[tid];
    for (int r = tid; r < i0; r += 1024) {
      T s = b[r];
      for (int i = 0; i < ib; ++i) s = s - LU[(size_t)(i0 + i) * n + r] * xs[i];
      b[r] = s;
    }
    __syncthreads();
  }
}

template <typename T, int GEMM>
int lu_factor_t(hipStream_t s, T* A, int n, int* ipiv, int* info) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, info);
  for (int k0 = 0; k0 < n; k0 += kNB) {
    const int kb = std::min(kNB, n - k0);
    hipLaunchKernelGGL((k_getf2<T>), dim3(1), dim3(1024), 0, s, A, n, n, k0, kb,
                       ipiv, info);
    if (n - kb > 0)
      hipLaunchKernelGGL((k_laswp<T>), dim3(cdiv(n - kb, 256)), dim3(256), 0, s,
                         A, n, n, k0, kb, (const int*)ipiv);
    const int rest = n - k0 - kb;
    if (rest > 0) {
      hipLaunchKernelGGL((k_trsm<T>), dim3(cdiv(rest, 64)), dim3(64), 0, s, A,
                         n, n, k0, kb);
      dim3 g(cdiv(rest, 64), cdiv(rest, 64));
      if constexpr (GEMM == 0) {
        hipLaunchKernelGGL(k_gemm_f64, g, dim3(256), 0, s, (double*)A, n, n, k0, kb);
      } else if constexpr (GEMM == 1) {
        hipLaunchKernelGGL((k_gemm_f32<false>), g, dim3(256), 0, s, (float*)A, n,
                           n, k0, kb);
      } else {
        hipLaunchKernelGGL((k_gemm_f32<true>), g, dim3(256), 0, s, (float*)A, n,
                           n, k0, kb);
      }
    }
  }
  return 0;
}

// --- mixed precision helpers ------------------------------------------------
__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b,
                         size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// r = b - A x  (FP64, column-major A, one row per lane; columns streamed).
__global__ __launch_bounds__(256) void k_residual(const double* __restrict__ A,
                                                  int n,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ r,
                                                  float* __restrict__ rf) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < n; ++j) s = __builtin_fma(A[(size_t)j * n + i], x[j], s);
  const double v = b[i] - s;
  r[i] = v;
  rf[i] = (float)v;
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

int dense_lu_factor(hipStream_t s, double* A, int n, int* ipiv, int* info_dev) {
  return lu_factor_t<double, 0>(s, A, n, ipiv, info_dev);
}

void dense_lu_solve(hipStream_t s, const double* LU, int n, const int* ipiv,
                    double* b) {
  hipLaunchKernelGGL((k_getrs<double>), dim3(1), dim3(1024), 0, s, LU, n, ipiv,
                     b);
}

size_t dense_mixed_work_bytes(int n) {
  const size_t nn = (size_t)n * n;
  return nn * sizeof(float) + (size_t)n * (sizeof(int) + 2 * sizeof(double) +
                                           2 * sizeof(float)) +
         64 * sizeof(double);
}

// Mixed precision: LU of fl32(A) (FP32 or BF16-input MFMA updates), then
// x_{k+1} = x_k + LU^-1 (b - A x_k) with the residual in FP64.
int dense_mixed_solve(hipStream_t s, const double* A, int n, const double* b,
                      double* x, int mode, double tol, int max_iter, void* work,
                      size_t work_bytes, int* iters) {
  if (work_bytes < dense_mixed_work_bytes(n)) return -1;
  char* w = (char*)work;
  const size_t nn = (size_t)n * n;
  float* Af = (float*)w;
  w += nn * sizeof(float);
  int* ipiv = (int*)w;
  w += (size_t)n * sizeof(int);
  double* r = (double*)w;
  w += (size_t)n * sizeof(double);
  double* norms = (double*)w;
  w += 64 * sizeof(double);
  float* rf = (float*)w;
  w += (size_t)n * sizeof(float);
  int* info = (int*)(w);
  hipLaunchKernelGGL(k_to_f32, dim3(cdiv(nn, 256)), dim3(256), 0, s, A, Af, nn);
  if (mode == 2)
    lu_factor_t<float, 2>(s, Af, n, ipiv, info);
  else
    lu_factor_t<float, 1>(s, Af, n, ipiv, info);
  // x0 = LU^-1 b
  (void)hipMemsetAsync(x, 0, sizeof(double) * n, s);
  int it = 0;
  double h_norms[2] = {0, 0};
  for (it = 0; it < max_iter; ++it) {
    hipLaunchKernelGGL(k_residual, dim3(cdiv(n, 256)), dim3(256), 0, s, A, n, x,
                       b, r, rf);
    hipLaunchKernelGGL((k_getrs<float>), dim3(1), dim3(1024), 0, s,
                       (const float*)Af, n, (const int*)ipiv, rf);
    hipLaunchKernelGGL(k_update, dim3(1), dim3(1024), 0, s, x, (const float*)rf,
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
