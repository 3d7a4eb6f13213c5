// lu_dev.h -- device substitution shared by the single-case and batched
// dense solves.
#pragma once

#include <hip/hip_runtime.h>

namespace uvlm {

// x <- U^-1 L^-1 x for a column-major LU (lda = n, unit lower), one
// 1024-thread workgroup; x (already permuted) in LDS or global memory.
// Blocks of 64 rows: the 64x64 diagonal block is staged in LDS (D) and
// solved by wavefront 0 with lane shuffles, then all lanes apply the block
// to the remaining rows.
template <typename T>
__device__ void lu_substitute(const T* __restrict__ LU, int n, T* x, T (*D)[65]) {
  const int tid = threadIdx.x, lane = tid & 63;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int ib = min(64, n - i0);
    __syncthreads();
    for (int q = tid; q < 64 * 64; q += blockDim.x) {
      const int r = q % 64, c = q / 64;
      D[r][c] = (r < ib && c < ib) ? LU[(size_t)(i0 + c) * n + i0 + r] : T(0);
    }
    __syncthreads();
    if (tid < 64) {
      T v = lane < ib ? x[i0 + lane] : T(0);
      for (int i = 0; i < ib; ++i) {
        const T xi = __shfl(v, i);
        if (lane > i && lane < ib) v = v - D[lane][i] * xi;
      }
      if (lane < ib) x[i0 + lane] = v;
    }
    __syncthreads();
    for (int r = i0 + ib + tid; r < n; r += blockDim.x) {
      T s = x[r];
      const T* row = LU + (size_t)i0 * n + r;
#pragma unroll 16
      for (int i = 0; i < ib; ++i) s = s - row[(size_t)i * n] * x[i0 + i];
      x[r] = s;
    }
  }
  for (int iend = n; iend > 0; iend -= 64) {
    const int i0 = max(0, iend - 64), ib = iend - i0;
    __syncthreads();
    for (int q = tid; q < 64 * 64; q += blockDim.x) {
      const int r = q % 64, c = q / 64;
      D[r][c] = (r < ib && c < ib) ? LU[(size_t)(i0 + c) * n + i0 + r] : T(0);
    }
    __syncthreads();
    if (tid < 64) {
      T v = lane < ib ? x[i0 + lane] : T(0);
      const T dl = lane < ib ? D[lane][lane] : T(1);
      for (int i = ib - 1; i >= 0; --i) {
        if (lane == i) v = v / dl;
        const T xi = __shfl(v, i);
        if (lane < i) v = v - D[lane][i] * xi;
      }
      if (lane < ib) x[i0 + lane] = v;
    }
    __syncthreads();
    for (int r = tid; r < i0; r += blockDim.x) {
      T s = x[r];
      const T* row = LU + (size_t)i0 * n + r;
#pragma unroll 16
      for (int i = 0; i < ib; ++i) s = s - row[(size_t)i * n] * x[i0 + i];
      x[r] = s;
    }
  }
  __syncthreads();
}

}  // namespace uvlm
