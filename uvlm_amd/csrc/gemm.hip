// gemm.hip -- matrix-core GEMM of the dense solves on gfx950.
//
// C = alpha A B + beta C, column-major.  A workgroup of 256 lanes (2 x 2
// wavefronts) computes a BM x BN tile of C; each wavefront a (BM/2) x (BN/2)
// block of 16 x 16 MFMA tiles held in accumulator registers.  Operand tiles
// of BK columns of A and BK rows of B are staged through LDS, double
// buffered: the global loads of step t + 1 are issued before the matrix
// instructions of step t, so one barrier per step suffices.
//
// The product is formed transposed (D = B^T A^T) so that the 16 lanes of a
// lane group hold 16 consecutive rows of one column of C: the epilogue's
// loads and stores of C are contiguous in column-major memory.
//
// v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32: a[lane] = X[lane%16]
// [lane/16], b[lane] = Y[lane/16][lane%16]; the result element r of a lane is
// D[4r + lane/16][lane%16] (f64) or D[4(lane/16) + r][lane%16] (f32).
// v_mfma_f32_16x16x32_bf16: lane l holds X[l%16][8(l/16) .. 8(l/16)+7] and
// Y[8(l/16) .. 8(l/16)+7][l%16]; result as the f32 form.
#include "gemm.h"

#include <algorithm>
#include <cstdint>
#include <type_traits>

namespace uvlm {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

inline int gdiv(long a, long b) { return (int)((a + b - 1) / b); }

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef v4d acc_t;
  static __device__ __forceinline__ acc_t mma(double x, double y, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
  }
  // column of C (within the 16-column block) of result element r
  static __device__ __forceinline__ int col(int r, int lq) { return 4 * r + lq; }
};
template <>
struct Mfma<float> {
  typedef v4f acc_t;
  static __device__ __forceinline__ acc_t mma(float x, float y, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int col(int r, int lq) { return 4 * lq + r; }
};

struct GemmPair {
  GemmArgs p0, p1;
  int tiles0;  // tiles of p0; blocks beyond belong to p1
};

// Tile of this workgroup: problem fields and the tile's first row / column.
template <int BM, int BN>
__device__ __forceinline__ GemmArgs pick(const GemmPair& pp, int* r0, int* c0) {
  int t = blockIdx.x;
  const bool second = t >= pp.tiles0;
  if (second) t -= pp.tiles0;
  const GemmArgs p = second ? pp.p1 : pp.p0;
  const int tm = (p.m + BM - 1) / BM;
  *r0 = (t % tm) * BM;
  *c0 = (t / tm) * BN;
  return p;
}

// Epilogue of one 16-row block of a wavefront: the JN x 4 elements of C
// this lane updates are read together (clamped in-bounds addresses, so the
// reads do not wait on one another), then alpha acc + beta C is stored (C
// not read when beta == 0).  Returns true when a result has |r| > 1.
template <typename T, int JN, typename Acc, typename ColOf>
__device__ __forceinline__ bool epilogue_rows(const GemmArgs& p, T* C, int row, int cbase,
                                              const Acc* acc, ColOf colof) {
  const T alpha = (T)p.alpha, beta = (T)p.beta;
  const int rc = min(row, p.m - 1);
  T cv[JN][4];
  if (beta != T(0)) {
#pragma unroll
    for (int j = 0; j < JN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = min(cbase + 16 * j + colof(r), p.n - 1);
        cv[j][r] = C[(size_t)col * p.ldc + rc];
      }
  }
  bool big = false;
#pragma unroll
  for (int j = 0; j < JN; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int col = cbase + 16 * j + colof(r);
      if (row < p.m && col < p.n) {
        T v = alpha * acc[j][r];
        if (beta != T(0)) v = v + beta * cv[j][r];
        C[(size_t)col * p.ldc + row] = v;
        big |= fabs(v) > T(1);
      }
    }
  return big;
}

template <typename T, int BM, int BN>
__global__ __launch_bounds__(256) void k_gemm(GemmPair pp) {
  constexpr int BK = 16;
  constexpr int PAD = 16 / (int)sizeof(T);  // 16 bytes per LDS row
  constexpr int WM = BM / 2, WN = BN / 2, IM = WM / 16, JN = WN / 16;
  constexpr int LA = BM * BK / 256, LB = BK * BN / 256;
  __shared__ T sA[2][BK][BM + PAD];
  __shared__ T sB[2][BK][BN + PAD];
  int r0, c0;
  const GemmArgs p = pick<BM, BN>(pp, &r0, &c0);
  const T* __restrict__ A = (const T*)p.A;
  const T* __restrict__ B = (const T*)p.B;
  T* C = (T*)p.C;
  const int t = threadIdx.x;
  // global -> register staging: A element e at row ar, column ak + e * (256 / BM);
  // B element e at row bk, column bc + e * (256 / BK)
  const int ar = t % BM, ak = t / BM;
  const int bk = t % BK, bc = t / BK;
  T ra[LA], rb[LB];
  auto load = [&](int k0) {
#pragma unroll
    for (int e = 0; e < LA; ++e) {  // rows beyond m only reach unstored results
      const int k = k0 + ak + e * (256 / BM);
      const auto v = A[(size_t)min(k, p.k - 1) * p.lda + min(r0 + ar, p.m - 1)];
      ra[e] = k < p.k ? v : T(0);
    }
#pragma unroll
    for (int e = 0; e < LB; ++e) {
      const int k = k0 + bk;
      const auto v = B[(size_t)min(c0 + bc + e * (256 / BK), p.n - 1) * p.ldb + min(k, p.k - 1)];
      rb[e] = k < p.k ? v : T(0);
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int e = 0; e < LA; ++e) sA[buf][ak + e * (256 / BM)][ar] = ra[e];
#pragma unroll
    for (int e = 0; e < LB; ++e) sB[buf][bk][bc + e * (256 / BK)] = rb[e];
  };
  const int w = t >> 6, lane = t & 63, l16 = lane & 15, lq = lane >> 4;
  const int wr = w >> 1, wc = w & 1;
  typename Mfma<T>::acc_t acc[IM][JN];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) acc[i][j] = typename Mfma<T>::acc_t{0, 0, 0, 0};
  const int nk = (p.k + BK - 1) / BK;
  if (nk > 0) {
    load(0);
    stage(0);
  }
  __syncthreads();
  for (int s = 0; s < nk; ++s) {
    const int cur = s & 1;
    if (s + 1 < nk) load((s + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T fa[IM], fb[JN];
#pragma unroll
      for (int i = 0; i < IM; ++i) fa[i] = sA[cur][kk + lq][wr * WM + 16 * i + l16];
#pragma unroll
      for (int j = 0; j < JN; ++j) fb[j] = sB[cur][kk + lq][wc * WN + 16 * j + l16];
#pragma unroll
      for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < JN; ++j) acc[i][j] = Mfma<T>::mma(fb[j], fa[i], acc[i][j]);
    }
    if (s + 1 < nk) stage(cur ^ 1);
    __syncthreads();
  }
  bool big = false;
#pragma unroll
  for (int i = 0; i < IM; ++i)
    big |= epilogue_rows<T, JN>(p, C, r0 + wr * WM + 16 * i + l16, c0 + wc * WN, acc[i],
                                [&](int r) { return Mfma<T>::col(r, lq); });
  if (big && p.bad) atomicOr(p.bad, 1);
}

// FP32 data, BF16 operands: BK = 32 (one v_mfma_f32_16x16x32_bf16 deep).
// LDS rows hold the k index contiguously (8 BF16 = 16 bytes per lane read).
__device__ __forceinline__ unsigned short to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)(u >> 16);
}

// IN16: the operands are already BF16 in memory (their rounding is to_bf16's,
// so the products equal the FP32-data form's bit for bit).
template <int BM, int BN, bool IN16>
__global__ __launch_bounds__(256) void k_gemm_bf16(GemmPair pp) {
  constexpr int BK = 32, KP = BK + 8;  // row pitch: 80 bytes
  constexpr int WM = BM / 2, WN = BN / 2, IM = WM / 16, JN = WN / 16;
  constexpr int LA = BM * BK / 256, LB = BK * BN / 256;
  __shared__ __align__(16) unsigned short sA[2][BM][KP];
  __shared__ __align__(16) unsigned short sB[2][BN][KP];
  int r0, c0;
  const GemmArgs p = pick<BM, BN>(pp, &r0, &c0);
  typedef typename std::conditional<IN16, unsigned short, float>::type E;
  const E* __restrict__ A = (const E*)p.A;
  const E* __restrict__ B = (const E*)p.B;
  float* C = (float*)p.C;
  const int t = threadIdx.x;
  // Each thread stages LA (LB) consecutive k of one row of A (column of B),
  // so its BF16 values go to LDS as 16-byte writes (LA / 8 of them) instead
  // of one 2-byte write each.  A (column-major, rows contiguous): lanes take
  // consecutive rows, so every load instruction is coalesced.  B (k
  // contiguous in a column): a thread's LB values are contiguous, loaded 16
  // bytes at a time when the run lies inside the column and is aligned.
  constexpr int GA = 256 / BM, GB = 256 / BN;  // k groups per row / column
  const int ar = t % BM, ak = t / BM;          // row of A, k group
  const int bc = t / GB, bk = t % GB;          // column of B, k group
  E ra[LA], rb[LB];
  constexpr int VEC = 16 / (int)sizeof(E);  // elements per 16-byte load of B
  const bool bvec = (p.ldb % VEC) == 0 && ((uintptr_t)B & 15) == 0;
  auto load = [&](int k0) {
#pragma unroll
    for (int e = 0; e < LA; ++e) {  // rows beyond m only reach unstored results
      const int k = k0 + ak * LA + e;
      const E v = A[(size_t)min(k, p.k - 1) * p.lda + min(r0 + ar, p.m - 1)];
      ra[e] = k < p.k ? v : E(0);
    }
    const int kb = k0 + bk * LB;
    const E* bcol = B + (size_t)min(c0 + bc, p.n - 1) * p.ldb;
    if (bvec && kb + LB <= p.k) {
#pragma unroll
      for (int e = 0; e < LB; e += VEC) {
        const uint4 v = *reinterpret_cast<const uint4*>(bcol + kb + e);
        if constexpr (IN16) {
          rb[e] = (E)(v.x & 0xffffu);
          rb[e + 1] = (E)(v.x >> 16);
          rb[e + 2] = (E)(v.y & 0xffffu);
          rb[e + 3] = (E)(v.y >> 16);
          rb[e + 4] = (E)(v.z & 0xffffu);
          rb[e + 5] = (E)(v.z >> 16);
          rb[e + 6] = (E)(v.w & 0xffffu);
          rb[e + 7] = (E)(v.w >> 16);
        } else {
          rb[e] = __uint_as_float(v.x);
          rb[e + 1] = __uint_as_float(v.y);
          rb[e + 2] = __uint_as_float(v.z);
          rb[e + 3] = __uint_as_float(v.w);
        }
      }
    } else {
#pragma unroll
      for (int e = 0; e < LB; ++e) {
        const int k = kb + e;
        const E v = bcol[min(k, p.k - 1)];
        rb[e] = k < p.k ? v : E(0);
      }
    }
  };
  auto pack = [](const E* r) {  // 8 operands as BF16, 16 bytes
    uint4 q;
    if constexpr (IN16) {
      q.x = r[0] | ((unsigned)r[1] << 16);
      q.y = r[2] | ((unsigned)r[3] << 16);
      q.z = r[4] | ((unsigned)r[5] << 16);
      q.w = r[6] | ((unsigned)r[7] << 16);
    } else {
      q.x = to_bf16(r[0]) | ((unsigned)to_bf16(r[1]) << 16);
      q.y = to_bf16(r[2]) | ((unsigned)to_bf16(r[3]) << 16);
      q.z = to_bf16(r[4]) | ((unsigned)to_bf16(r[5]) << 16);
      q.w = to_bf16(r[6]) | ((unsigned)to_bf16(r[7]) << 16);
    }
    return q;
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int e = 0; e < LA; e += 8) *reinterpret_cast<uint4*>(&sA[buf][ar][ak * LA + e]) = pack(ra + e);
#pragma unroll
    for (int e = 0; e < LB; e += 8) *reinterpret_cast<uint4*>(&sB[buf][bc][bk * LB + e]) = pack(rb + e);
  };
  static_assert(LA % 8 == 0 && LB % 8 == 0 && GA * LA == BK && GB * LB == BK,
                "each thread stages whole 16-byte runs of k");
  const int w = t >> 6, lane = t & 63, l16 = lane & 15, lq = lane >> 4;
  const int wr = w >> 1, wc = w & 1;
  v4f acc[IM][JN];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  const int nk = (p.k + BK - 1) / BK;
  if (nk > 0) {
    load(0);
    stage(0);
  }
  __syncthreads();
  for (int s = 0; s < nk; ++s) {
    const int cur = s & 1;
    if (s + 1 < nk) load((s + 1) * BK);
    v8bf fa[IM], fb[JN];
#pragma unroll
    for (int i = 0; i < IM; ++i)
      fa[i] = *reinterpret_cast<const v8bf*>(&sA[cur][wr * WM + 16 * i + l16][8 * lq]);
#pragma unroll
    for (int j = 0; j < JN; ++j)
      fb[j] = *reinterpret_cast<const v8bf*>(&sB[cur][wc * WN + 16 * j + l16][8 * lq]);
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
      for (int j = 0; j < JN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[j], fa[i], acc[i][j], 0, 0, 0);
    if (s + 1 < nk) stage(cur ^ 1);
    __syncthreads();
  }
  bool big = false;
#pragma unroll
  for (int i = 0; i < IM; ++i)
    big |= epilogue_rows<float, JN>(p, C, r0 + wr * WM + 16 * i + l16, c0 + wc * WN, acc[i],
                                    [&](int r) { return 4 * lq + r; });
  if (big && p.bad) atomicOr(p.bad, 1);
}

template <int BM, int BN>
int tiles_of(const GemmArgs& p) {
  if (p.m <= 0 || p.n <= 0) return 0;
  return gdiv(p.m, BM) * gdiv(p.n, BN);
}

template <int BM, int BN>
void launch_tile(hipStream_t s, int prec, GemmPair& pp) {
  pp.tiles0 = tiles_of<BM, BN>(pp.p0);
  const int total = pp.tiles0 + tiles_of<BM, BN>(pp.p1);
  if (total == 0) return;
  if (prec == GEMM_F64)
    hipLaunchKernelGGL((k_gemm<double, BM, BN>), dim3(total), dim3(256), 0, s, pp);
  else if (prec == GEMM_F32)
    hipLaunchKernelGGL((k_gemm<float, BM, BN>), dim3(total), dim3(256), 0, s, pp);
  else if (prec == GEMM_BF16)
    hipLaunchKernelGGL((k_gemm_bf16<BM, BN, false>), dim3(total), dim3(256), 0, s, pp);
  else
    hipLaunchKernelGGL((k_gemm_bf16<BM, BN, true>), dim3(total), dim3(256), 0, s, pp);
}

__global__ void k_round_bf16(const float* __restrict__ x, long n, unsigned short* __restrict__ y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = to_bf16(x[i]);
}

}  // namespace

void launch_round_bf16(hipStream_t s, const float* x, long n, unsigned short* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_round_bf16, dim3((unsigned)std::min<long>(2048, gdiv(n, 256))), dim3(256), 0,
                     s, x, n, y);
}

void launch_gemm(hipStream_t s, int prec, const GemmArgs* p, int count) {
  GemmPair pp;
  pp.p0 = p[0];
  if (count > 1) {
    pp.p1 = p[1];
  } else {
    pp.p1 = p[0];
    pp.p1.m = 0;
  }
  // 128 x 128 tiles for deep products (k >= 256) that fill the chip (256
  // CUs) when the problems leave in-place use valid; 64 x 64 tiles
  // otherwise (measured on the LU's rank-64 updates: 30 against 22 TFLOP/s
  // at 4032 x 4032, four wavefronts per SIMD hiding the loads of C).
  const bool in_place = (pp.p0.C == pp.p0.A || pp.p0.C == pp.p0.B) ||
                        (pp.p1.m > 0 && (pp.p1.C == pp.p1.A || pp.p1.C == pp.p1.B));
  const int kmin = pp.p1.m > 0 ? std::min(pp.p0.k, pp.p1.k) : pp.p0.k;
  if (!in_place && kmin >= 256 &&
      tiles_of<128, 128>(pp.p0) + tiles_of<128, 128>(pp.p1) >= 256)
    launch_tile<128, 128>(s, prec, pp);
  else
    launch_tile<64, 64>(s, prec, pp);
}

}  // namespace uvlm
