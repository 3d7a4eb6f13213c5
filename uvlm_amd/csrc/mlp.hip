// mlp.hip -- training (FP32) and BF16 inference of the far-field correction
// network on gfx950.
//
// Training keeps a batch's activations in HBM and runs each layer as a
// plain kernel with one thread per output element: forward Z = H W^T + b
// (weights transposed once per iteration so the lanes read them
// coalesced), backward dH = dZ W, dZ_prev = dH silu'(Z_prev), weight
// gradients as per-chunk partial sums over the batch rows (bias = column
// K) reduced in chunk order, then Adam.  Every sum runs in a fixed order, so
// training is deterministic.  Inference (the far-field kernel and
// launch_mlp_infer) runs the network on MFMA in BF16 (mlp_dev.h).
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/uvlm_gpu.h"

namespace uvlm {
namespace {

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

constexpr int H = kMlpHidden, IN = kMlpIn;
constexpr int kChunkRows = 128;  // batch rows per weight-gradient partial

// Inference copy offsets (halfs, then floats after an aligned boundary).
constexpr size_t kOffW1 = 0, kOffW2 = kOffW1 + (size_t)H * IN, kOffW3 = kOffW2 + (size_t)H * H,
                 kOffW4 = kOffW3 + (size_t)H * H;
constexpr size_t kHalfBytes = ((kMlpDevHalfs * 2) + 255) / 256 * 256;

__device__ inline float silu_f(float z) { return z / (1.0f + expf(-z)); }
__device__ inline float dsilu_f(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// FP32 parameters -> BF16 weights (W4 padded to 16 rows) + FP32 biases.
__global__ void k_mlp_to_dev(const float* __restrict__ w, unsigned short* __restrict__ hw,
                             float* __restrict__ fb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  using S = MlpShape;
  if (t < H * IN) hw[kOffW1 + t] = f32_to_bf16(w[S::off_w1 + t]);
  if (t < H * H) {
    hw[kOffW2 + t] = f32_to_bf16(w[S::off_w2 + t]);
    hw[kOffW3 + t] = f32_to_bf16(w[S::off_w3 + t]);
  }
  if (t < kMlpW4Rows * H) hw[kOffW4 + t] = t < kMlpOut * H ? f32_to_bf16(w[S::off_w4 + t]) : 0;
  if (t < H) {
    fb[t] = w[S::off_b1 + t];
    fb[H + t] = w[S::off_b2 + t];
    fb[2 * H + t] = w[S::off_b3 + t];
  }
  if (t < kMlpW4Rows) fb[3 * H + t] = t < kMlpOut ? w[S::off_b4 + t] : 0.0f;
}

// One wavefront per 64 samples; rows beyond n run on zero features and are
// not written.
__global__ __launch_bounds__(64) void k_mlp_infer(MlpDev net, const float* __restrict__ X, int n,
                                                  float* __restrict__ Y) {
  __shared__ unsigned short lds[2 * kMlpLdsHalfs];
  __shared__ float ylds[64 * kMlpYStride];
  const int row = blockIdx.x * 64 + threadIdx.x;
  float f[kMlpIn];
  for (int k = 0; k < kMlpIn; ++k) f[k] = row < n ? X[(long)row * kMlpIn + k] : 0.0f;
  mlp_wave_bf16(net, f, lds, ylds);
  if (row < n)
    for (int o = 0; o < kMlpOut; ++o)
      Y[(long)row * kMlpOut + o] = ylds[threadIdx.x * kMlpYStride + o];
}

// WT[k][n] = W[n][k] for the four layers (k-major copies for the forward).
__global__ void k_transpose_all(const float* __restrict__ w, float* __restrict__ wt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  using S = MlpShape;
  if (t < H * IN) {  // W1 [H][IN] -> WT1 [IN][H]
    const int n = t / IN, k = t % IN;
    wt[k * H + n] = w[S::off_w1 + t];
  }
  if (t < H * H) {
    const int n = t / H, k = t % H;
    wt[H * IN + k * H + n] = w[S::off_w2 + t];
    wt[H * IN + H * H + k * H + n] = w[S::off_w3 + t];
  }
  if (t < kMlpOut * H) {  // W4 [OUT][H] -> WT4 [H][OUT]
    const int n = t / H, k = t % H;
    wt[H * IN + 2 * H * H + k * kMlpOut + n] = w[S::off_w4 + t];
  }
}

// Z[b][n] = b[n] + sum_k in[b][k] WT[k][n]; Hout = silu(Z) (when given).
__global__ void k_fwd(const float* __restrict__ in, int K, const float* __restrict__ wt,
                      const float* __restrict__ bias, int N, int B, float* __restrict__ Z,
                      float* __restrict__ Hout) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)B * N) return;
  const int b = (int)(t / N), n = (int)(t % N);
  const float* a = in + (long)b * K;
  float z = 0.0f;
  for (int k = 0; k < K; ++k) z = fmaf(a[k], wt[(long)k * N + n], z);
  z += bias[n];
  Z[t] = z;
  if (Hout) Hout[t] = silu_f(z);
}

// dY = (Y - V) / B; lrow[b] = 1/2 |Y - V|^2.
__global__ void k_out_grad(const float* __restrict__ Y, const float* __restrict__ V, int B,
                           float* __restrict__ dY, float* __restrict__ lrow) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float l = 0.0f;
  for (int o = 0; o < kMlpOut; ++o) {
    const long q = (long)kMlpOut * b + o;
    const float e = Y[q] - V[q];
    dY[q] = e / (float)B;
    l = fmaf(e, e, l);
  }
  lrow[b] = 0.5f * l;
}

// dZp[b][k] = (sum_n dZ[b][n] W[n][k]) silu'(Zp[b][k]).
__global__ void k_bwd_data(const float* __restrict__ dZ, int N, const float* __restrict__ W,
                           int K, const float* __restrict__ Zp, int B, float* __restrict__ dZp) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)B * K) return;
  const int b = (int)(t / K), k = (int)(t % K);
  const float* d = dZ + (long)b * N;
  float a = 0.0f;
  for (int n = 0; n < N; ++n) a = fmaf(d[n], W[(long)n * K + k], a);
  dZp[t] = a * dsilu_f(Zp[t]);
}

// P[c][n][k] = sum over rows b of chunk c of dZ[b][n] Hin[b][k] (k = K: 1).
__global__ void k_wgrad(const float* __restrict__ dZ, int N, const float* __restrict__ Hin, int K,
                        int B, float* __restrict__ P) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int K1 = K + 1;
  if (t >= N * K1) return;
  const int n = t / K1, k = t % K1, c = blockIdx.y;
  const int b0 = c * kChunkRows, b1 = min(B, b0 + kChunkRows);
  float a = 0.0f;
  for (int b = b0; b < b1; ++b)
    a = fmaf(dZ[(long)b * N + n], k < K ? Hin[(long)b * K + k] : 1.0f, a);
  P[((long)c * N + n) * K1 + k] = a;
}

// g[offw + n K + k] (k < K) and g[offb + n] (k = K) = sum over chunks.
__global__ void k_wgrad_reduce(const float* __restrict__ P, int C, int N, int K,
                               float* __restrict__ g, int offw, int offb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int K1 = K + 1;
  if (t >= N * K1) return;
  const int n = t / K1, k = t % K1;
  float a = 0.0f;
  for (int c = 0; c < C; ++c) a += P[((long)c * N + n) * K1 + k];
  if (k < K)
    g[offw + n * K + k] = a;
  else
    g[offb + n] = a;
}

// loss = sum_b lrow[b] / B (fixed tree order, one workgroup).
__global__ __launch_bounds__(256) void k_loss_sum(const float* __restrict__ lrow, int B,
                                                  float* __restrict__ loss) {
  __shared__ float s[256];
  float a = 0.0f;
  for (int b = threadIdx.x; b < B; b += 256) a += lrow[b];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = s[0] / (float)B;
}

__global__ void k_gather_rows(const float* __restrict__ X, const float* __restrict__ V,
                              const int* __restrict__ idx, int B, float* __restrict__ Xb,
                              float* __restrict__ Vb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * (IN + kMlpOut)) return;
  const int b = t / (IN + kMlpOut), k = t % (IN + kMlpOut);
  const long r = idx[b];
  if (k < IN)
    Xb[(long)b * IN + k] = X[r * IN + k];
  else
    Vb[(long)b * kMlpOut + k - IN] = V[r * kMlpOut + k - IN];
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = 0.9f * m[t] + 0.1f * gt;
  const float vt = 0.999f * v[t] + 0.001f * gt * gt;
  m[t] = mt;
  v[t] = vt;
  w[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + 1e-8f);
}

__global__ void k_fill_f(float* p, int n, float v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

}  // namespace

size_t mlp_dev_bytes() { return kHalfBytes + sizeof(float) * kMlpDevFloats; }

MlpDev mlp_dev_view(const void* base) {
  const unsigned short* h = (const unsigned short*)base;
  const float* f = (const float*)((const char*)base + kHalfBytes);
  return MlpDev{h + kOffW1, h + kOffW2, h + kOffW3, h + kOffW4, f, f + H, f + 2 * H, f + 3 * H};
}

void launch_mlp_to_dev(hipStream_t s, const float* w, void* dev) {
  hipLaunchKernelGGL(k_mlp_to_dev, dim3(cdiv(H * H, 256)), dim3(256), 0, s, w,
                     (unsigned short*)dev, (float*)((char*)dev + kHalfBytes));
}

void launch_mlp_infer(hipStream_t s, MlpDev net, const float* X, int n, float* Y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mlp_infer, dim3(cdiv(n, 64)), dim3(64), 0, s, net, X, n, Y);
}

// ---------------------------------------------------------------------------
MlpTrainer::~MlpTrainer() {
  float* bufs[] = {Z1_, Z2_, Z3_, H1_, H2_, H3_, Y_, dY_, D3_, D2_, D1_, WT_, P_, lrow_,
                   Xb_, Vb_, g_, m_, v_, loss_};
  for (float* p : bufs)
    if (p) (void)hipFree(p);
  if (idx_) (void)hipFree(idx_);
}

int MlpTrainer::ensure(int rows, std::string* err) {
  if (rows <= cap_ && g_) return UVLM_OK;
  float** bufs[] = {&Z1_, &Z2_, &Z3_, &H1_, &H2_, &H3_, &Y_, &dY_, &D3_, &D2_, &D1_, &WT_, &P_,
                    &lrow_, &Xb_, &Vb_, &g_, &m_, &v_, &loss_};
  for (float** p : bufs) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  const size_t R = (size_t)rows, C = (size_t)cdiv(rows, kChunkRows);
  const size_t O = kMlpOut;
  const size_t sizes[] = {R * H, R * H, R * H, R * H, R * H, R * H, R * O, R * O, R * H, R * H,
                          R * H, (size_t)MlpShape::total, C * H * (H + 1), R, R * IN, R * O,
                          (size_t)MlpShape::total, (size_t)MlpShape::total,
                          (size_t)MlpShape::total, 4};
  for (int i = 0; i < 20; ++i)
    if (hipMalloc((void**)bufs[i], sizeof(float) * sizes[i]) != hipSuccess) {
      (void)hipGetLastError();
      *err = "network buffer allocation failed";
      cap_ = 0;
      return UVLM_ERR_OOM;
    }
  cap_ = rows;
  return UVLM_OK;
}

void MlpTrainer::fwd(hipStream_t s, const float* w, const float* X, int B) {
  using S = MlpShape;
  hipLaunchKernelGGL(k_transpose_all, dim3(cdiv(H * H, 256)), dim3(256), 0, s, w, WT_);
  const float* wt1 = WT_;
  const float* wt2 = wt1 + H * IN;
  const float* wt3 = wt2 + H * H;
  const float* wt4 = wt3 + H * H;
  const int g = cdiv((long)B * H, 256);
  hipLaunchKernelGGL(k_fwd, dim3(g), dim3(256), 0, s, X, IN, wt1, w + S::off_b1, H, B, Z1_, H1_);
  hipLaunchKernelGGL(k_fwd, dim3(g), dim3(256), 0, s, (const float*)H1_, H, wt2, w + S::off_b2, H,
                     B, Z2_, H2_);
  hipLaunchKernelGGL(k_fwd, dim3(g), dim3(256), 0, s, (const float*)H2_, H, wt3, w + S::off_b3, H,
                     B, Z3_, H3_);
  hipLaunchKernelGGL(k_fwd, dim3(cdiv((long)B * kMlpOut, 256)), dim3(256), 0, s,
                     (const float*)H3_, H, wt4, w + S::off_b4, kMlpOut, B, Y_, (float*)nullptr);
}

void MlpTrainer::bwd(hipStream_t s, const float* w, const float* X, const float* V, int B,
                     float* g, float* loss) {
  using S = MlpShape;
  const int C = cdiv(B, kChunkRows);
  hipLaunchKernelGGL(k_out_grad, dim3(cdiv(B, 256)), dim3(256), 0, s, (const float*)Y_, V, B,
                     dY_, lrow_);
  hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(256), 0, s, (const float*)lrow_, B, loss);
  const int gh = cdiv((long)B * H, 256);
  // layer 4
  hipLaunchKernelGGL(k_wgrad, dim3(cdiv(kMlpOut * (H + 1), 256), C), dim3(256), 0, s,
                     (const float*)dY_, kMlpOut, (const float*)H3_, H, B, P_);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(cdiv(kMlpOut * (H + 1), 256)), dim3(256), 0, s,
                     (const float*)P_, C, kMlpOut, H, g, S::off_w4, S::off_b4);
  hipLaunchKernelGGL(k_bwd_data, dim3(gh), dim3(256), 0, s, (const float*)dY_, kMlpOut,
                     w + S::off_w4, H, (const float*)Z3_, B, D3_);
  // layer 3
  hipLaunchKernelGGL(k_wgrad, dim3(cdiv(H * (H + 1), 256), C), dim3(256), 0, s,
                     (const float*)D3_, H, (const float*)H2_, H, B, P_);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(cdiv(H * (H + 1), 256)), dim3(256), 0, s,
                     (const float*)P_, C, H, H, g, S::off_w3, S::off_b3);
  hipLaunchKernelGGL(k_bwd_data, dim3(gh), dim3(256), 0, s, (const float*)D3_, H, w + S::off_w3,
                     H, (const float*)Z2_, B, D2_);
  // layer 2
  hipLaunchKernelGGL(k_wgrad, dim3(cdiv(H * (H + 1), 256), C), dim3(256), 0, s,
                     (const float*)D2_, H, (const float*)H1_, H, B, P_);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(cdiv(H * (H + 1), 256)), dim3(256), 0, s,
                     (const float*)P_, C, H, H, g, S::off_w2, S::off_b2);
  hipLaunchKernelGGL(k_bwd_data, dim3(gh), dim3(256), 0, s, (const float*)D2_, H, w + S::off_w2,
                     H, (const float*)Z1_, B, D1_);
  // layer 1
  hipLaunchKernelGGL(k_wgrad, dim3(cdiv(H * (IN + 1), 256), C), dim3(256), 0, s,
                     (const float*)D1_, H, X, IN, B, P_);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(cdiv(H * (IN + 1), 256)), dim3(256), 0, s,
                     (const float*)P_, C, H, IN, g, S::off_w1, S::off_b1);
  hipLaunchKernelGGL(k_fill_f, dim3(1), dim3(64), 0, s, g + S::off_b4 + kMlpOut, 2, 0.0f);  // pad
}

int MlpTrainer::forward(hipStream_t s, const float* w, const float* X, int n, float* Y,
                        std::string* err) {
  if (n <= 0) return UVLM_OK;
  int rc = ensure(n, err);
  if (rc != UVLM_OK) return rc;
  fwd(s, w, X, n);
  if (hipMemcpyAsync(Y, Y_, sizeof(float) * kMlpOut * (size_t)n, hipMemcpyDeviceToDevice, s) !=
      hipSuccess) {
    *err = "network output copy failed";
    return UVLM_ERR_HIP;
  }
  return UVLM_OK;
}

int MlpTrainer::grad(hipStream_t s, const float* w, const float* X, const float* V, int n,
                     float* g, float* loss, std::string* err) {
  if (n <= 0) {
    *err = "no samples";
    return UVLM_ERR_ARG;
  }
  int rc = ensure(n, err);
  if (rc != UVLM_OK) return rc;
  fwd(s, w, X, n);
  bwd(s, w, X, V, n, g, loss);
  return UVLM_OK;
}

int MlpTrainer::train(hipStream_t s, float* w, const float* X, const float* V, int n, int iters,
                      int batch, float lr, unsigned seed, std::vector<float>* losses,
                      std::string* err) {
  if (n <= 0 || iters <= 0 || batch <= 0) {
    *err = "training needs samples, iterations and a batch size";
    return UVLM_ERR_ARG;
  }
  batch = std::min(batch, n);
  int rc = ensure(batch, err);
  if (rc != UVLM_OK) return rc;
  // mini-batch rows: a seeded permutation per epoch (xorshift32 Fisher-Yates)
  std::vector<int> rows((size_t)iters * batch);
  std::vector<int> perm(n);
  unsigned st = seed ? seed : 0x9e3779b9u;
  size_t pos = (size_t)n;
  for (size_t q = 0; q < rows.size(); ++q) {
    if (pos == (size_t)n) {
      for (int i = 0; i < n; ++i) perm[i] = i;
      for (int i = n - 1; i > 0; --i) {
        st ^= st << 13;
        st ^= st >> 17;
        st ^= st << 5;
        std::swap(perm[i], perm[st % (unsigned)(i + 1)]);
      }
      pos = 0;
    }
    rows[q] = perm[pos++];
  }
  if (rows.size() > idx_cap_) {
    if (idx_) (void)hipFree(idx_);
    idx_ = nullptr;
    idx_cap_ = 0;
    if (hipMalloc((void**)&idx_, sizeof(int) * rows.size()) != hipSuccess) {
      (void)hipGetLastError();
      *err = "index buffer allocation failed";
      return UVLM_ERR_OOM;
    }
    idx_cap_ = rows.size();
  }
  if (hipMemcpy(idx_, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice) !=
      hipSuccess) {
    *err = "index upload failed";
    return UVLM_ERR_HIP;
  }
  const int P = MlpShape::total;
  hipLaunchKernelGGL(k_fill_f, dim3(cdiv(P, 256)), dim3(256), 0, s, m_, P, 0.0f);
  hipLaunchKernelGGL(k_fill_f, dim3(cdiv(P, 256)), dim3(256), 0, s, v_, P, 0.0f);
  if (losses) losses->assign(iters, 0.0f);
  std::vector<float> lh;
  float* dloss = nullptr;
  if (losses && hipMalloc((void**)&dloss, sizeof(float) * iters) != hipSuccess) {
    (void)hipGetLastError();
    *err = "loss buffer allocation failed";
    return UVLM_ERR_OOM;
  }
  for (int it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(k_gather_rows, dim3(cdiv((long)batch * (IN + kMlpOut), 256)), dim3(256), 0, s,
                       X, V, (const int*)(idx_ + (size_t)it * batch), batch, Xb_, Vb_);
    fwd(s, w, Xb_, batch);
    bwd(s, w, Xb_, Vb_, batch, g_, dloss ? dloss + it : loss_);
    const double t = it + 1;
    hipLaunchKernelGGL(k_adam, dim3(cdiv(P, 256)), dim3(256), 0, s, w, (const float*)g_, m_, v_,
                       P, lr, (float)(1.0 - std::pow(0.9, t)), (float)(1.0 - std::pow(0.999, t)));
  }
  if (dloss) {
    const hipError_t e1 =
        hipMemcpyAsync(losses->data(), dloss, sizeof(float) * iters, hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(dloss);
    if (e1 != hipSuccess || e2 != hipSuccess) {
      *err = "loss read-back failed";
      return UVLM_ERR_HIP;
    }
  }
  if (hipGetLastError() != hipSuccess) {
    *err = "training launch failed";
    return UVLM_ERR_HIP;
  }
  return UVLM_OK;
}

}  // namespace uvlm
