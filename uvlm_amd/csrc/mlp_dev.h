// mlp_dev.h -- BF16 MFMA inference of the far-field correction network for
// one wavefront of 64 samples (one per lane): used by the tile-correction
// kernel of the far field and by the stand-alone inference entry point.
//
// Network: h1 = silu(W1 f + b1), h2 = silu(W2 h1 + b2), h3 = silu(W3 h2 + b3),
// y = W4 h3 + b4 (32 -> 128 -> 128 -> 128 -> 30).  Weights and activations
// are rounded to BF16 (round to nearest even), products accumulate in FP32
// on gfx950's v_mfma_f32_16x16x32_bf16, biases and activations are FP32.
//
// MFMA operand layout (16 x 16 x 32, one wavefront): lane l holds
// A[l % 16][8 (l / 16) .. 8 (l / 16) + 7] and B[8 (l / 16) .. + 7][l % 16]
// (8 BF16 each); the result D[4 (l / 16) + r][l % 16] is element r of the
// lane's accumulator.  Here A = activations (row = sample), B[k][j] =
// W[j][k], so a lane holds 4 samples' values of one output.
#pragma once

#include <hip/hip_runtime.h>

#include "farfield.h"

namespace uvlm {

// Inference copy of the network: weights as BF16 bits, row-major [out][in]
// (W4 padded with zero rows to 32 outputs), biases FP32 (b4 padded to 32).
struct MlpDev {
  const unsigned short* w1;  // [128][32]
  const unsigned short* w2;  // [128][128]
  const unsigned short* w3;  // [128][128]
  const unsigned short* w4;  // [32][128]
  const float* b1;
  const float* b2;
  const float* b3;
  const float* b4;  // [32]
};
constexpr int kMlpW4Rows = 32;
constexpr size_t kMlpDevHalfs =
    (size_t)kMlpHidden * kMlpIn + 2 * (size_t)kMlpHidden * kMlpHidden + kMlpW4Rows * kMlpHidden;
constexpr size_t kMlpDevFloats = 3 * kMlpHidden + kMlpW4Rows;
// LDS row stride (BF16 elements) of the activation buffers: 128 + 8 keeps
// rows 16-byte aligned and the 16-byte fragment reads of a tile's 16 rows
// on distinct banks.
constexpr int kMlpLdsStride = kMlpHidden + 8;
constexpr int kMlpLdsHalfs = 64 * kMlpLdsStride;
// Row stride (floats) of the output buffer: 32 outputs + 1 against bank
// conflicts of the per-lane reads.
constexpr int kMlpYStride = kMlpW4Rows + 1;

typedef __bf16 mlp_v8bf __attribute__((ext_vector_type(8)));
typedef float mlp_v4f __attribute__((ext_vector_type(4)));

__host__ __device__ inline unsigned short f32_to_bf16(float f) {
  unsigned u;
#ifdef __HIP_DEVICE_COMPILE__
  u = __float_as_uint(f);
#else
  __builtin_memcpy(&u, &f, 4);
#endif
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ inline float mlp_silu(float z) { return z / (1.0f + __expf(-z)); }

// out[64 x N] = act(in[64 x K] W^T + b) for one wavefront.  in/out: LDS,
// BF16, row stride kMlpLdsStride; without SILU the FP32 results go to yout
// (row stride kMlpYStride).  NT = N / 16 output tiles, K a multiple of 32.
// The output tiles are formed kMlpChunk at a time (4 x kMlpChunk
// accumulators of 4 floats), so the registers stay within one wavefront's
// budget without spilling.
constexpr int kMlpChunk = 2;
template <int K, int NT, bool SILU>
__device__ inline void mlp_layer(const unsigned short* __restrict__ W, const float* __restrict__ b,
                                 const unsigned short* in, unsigned short* out, float* yout) {
  const int lane = threadIdx.x & 63;
  const int l16 = lane & 15, lq = lane >> 4;
  constexpr int NC = NT < kMlpChunk ? NT : kMlpChunk;
  static_assert(NT % NC == 0, "output tiles in whole chunks");
  for (int n0 = 0; n0 < NT; n0 += NC) {
    mlp_v4f acc[4][NC];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[mt][c] = mlp_v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int k0 = 0; k0 < K; k0 += 32) {
      const int k = k0 + 8 * lq;
      mlp_v8bf bf[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c)
        bf[c] = *reinterpret_cast<const mlp_v8bf*>(W + (size_t)(16 * (n0 + c) + l16) * K + k);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        const mlp_v8bf af =
            *reinterpret_cast<const mlp_v8bf*>(in + (16 * mt + l16) * kMlpLdsStride + k);
#pragma unroll
        for (int c = 0; c < NC; ++c)
          acc[mt][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf[c], acc[mt][c], 0, 0, 0);
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int n = 16 * (n0 + c) + l16;
      const float bn = b[n];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * mt + 4 * lq + r;
          const float z = acc[mt][c][r] + bn;
          if (SILU)
            out[row * kMlpLdsStride + n] = f32_to_bf16(mlp_silu(z));
          else
            yout[row * kMlpYStride + n] = z;
        }
    }
  }
}

// Forward pass of the 64 lanes' features f[32]: outputs of lane l in
// ylds[l * kMlpYStride + o], o < kMlpOut.  lds: 2 x kMlpLdsHalfs BF16
// buffers; ylds: 64 x kMlpYStride floats, which may be the first of the
// two buffers (the last layer reads only the second).  For one-wavefront
// workgroups; every lane calls it together (it contains workgroup barriers).
static_assert(64 * kMlpYStride * 4 <= kMlpLdsHalfs * 2, "outputs fit one activation buffer");
__device__ inline void mlp_wave_bf16(const MlpDev& net, const float* f, unsigned short* lds,
                                     float* ylds) {
  const int lane = threadIdx.x & 63;
  unsigned short* A = lds;
  unsigned short* B = lds + kMlpLdsHalfs;
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) A[lane * kMlpLdsStride + k] = f32_to_bf16(f[k]);
  __syncthreads();
  mlp_layer<kMlpIn, kMlpHidden / 16, true>(net.w1, net.b1, A, B, nullptr);
  __syncthreads();
  mlp_layer<kMlpHidden, kMlpHidden / 16, true>(net.w2, net.b2, B, A, nullptr);
  __syncthreads();
  mlp_layer<kMlpHidden, kMlpHidden / 16, true>(net.w3, net.b3, A, B, nullptr);
  __syncthreads();
  mlp_layer<kMlpHidden, kMlpW4Rows / 16, false>(net.w4, net.b4, B, nullptr, ylds);
  __syncthreads();
}

}  // namespace uvlm
