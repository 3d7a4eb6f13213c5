// inv_solve.hip -- dense circulation solve by iterative refinement on a kept
// approximate inverse (UVLM_SOLVE_F64_INV).
//
// The influence matrix of a time-stepping run changes little from one step
// to the next.  The solver keeps X ~ A^-1 of a recent step and refines
//     x_{k+1} = x_k + X (b - A x_k)
// with the residual in FP64: the error contracts by rho(I - X A) per sweep
// and the result solves the current A x = b to FP64 accuracy.  X only sets
// the contraction, so it is kept in FP32: its rounding (~6e-8 cond(A))
// bounds rho far below the refresh threshold, halves the bytes of its
// matrix-vector product and lets its Newton-Schulz steps run on the FP32
// matrix cores.  A matrix whose FP32 inverse cannot reach that accuracy
// (the build stalls above 1e-3) is handled with X in FP64.
// A sweep is two matrix-vector products (split-column partial sums, reduced
// in a fixed order, so results are deterministic) instead of a
// latency-bound LU.  The start is extrapolated from the last two solutions
// of a run.  X is improved in place by Newton-Schulz steps
// X <- X (2I - A X) (two matrix-core GEMMs, gemm.hip) when the contraction
// slows, and built initially from X0 = A^T / (||A||_1 ||A||_inf), for which
// the iteration converges for every non-singular A.
// A batch of sweeps is queued at once; a one-thread check after each sweep
// raises a device flag on convergence (or divergence) and the later sweeps
// of the batch return at once, so the host usually reads back once per solve.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/uvlm_gpu.h"
#include "gemm.h"
#include "internal.h"

namespace uvlm {
namespace {

constexpr int kMaxBatch = 16;           // sweeps queued per host read
constexpr int kNsSlot = 2 * kMaxBatch;  // Newton-Schulz deviation word
constexpr int kFlagSlot = kNsSlot + 1;  // 0 running, 1 + k converged at k, -1 - k diverged
constexpr int kCtrSlot = kFlagSlot + 1;  // workgroups of k_update_sum done (reset by the last)
constexpr int kStatWords = kCtrSlot + 1;
constexpr double kRefreshRho = 0.05;    // contraction above which X is refreshed
// A warm solve starts with a BF16 refresh when the previous solve contracted
// more slowly than this (the matrix of a time-stepping run drifts by about
// as much per step).
constexpr double kPreRefreshRho = 0.02;
// ... for matrices of at most this order: the refresh's two GEMMs grow as
// N^3 against the sweeps' N^2 (measured: at N = 2048 the queued refreshes
// take a run from 0.83 to 0.5 ms of solve per step, at N = 4096 they cost
// 1.4 ms per step more than the sweeps they save).
constexpr int kPreRefreshMaxN = 3072;
constexpr int kRedGroups = 16;          // partial-sum groups per row in the reductions

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

__device__ inline void atomic_max_nonneg(unsigned long long* p, double v) {
  atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// Max over a 256-thread workgroup of a non-negative value, then one atomic.
__device__ inline void block_max_to(double v, unsigned long long* out, double* lds) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
    atomic_max_nonneg(out, v);
  }
}

// T <- 2I - T and out <- max |I - T| (grid-stride, one atomic per workgroup)
template <typename T_>
__global__ __launch_bounds__(256) void k_two_minus_dev(T_* T, int n, unsigned long long* out) {
  __shared__ double lds[4];
  const long nn = (long)n * n;
  double e = 0.0;
  for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < nn; k += (long)gridDim.x * 256) {
    const long i = k % n, j = k / n;
    const T_ id = i == j ? (T_)1 : (T_)0;
    const T_ t = T[k];
    e = fmax(e, (double)fabs(id - t));
    T[k] = (T_)2 * id - t;
  }
  block_max_to(e, out, lds);
}

// X <- alpha A^T
template <typename T_>
__global__ void k_scaled_transpose(const double* A, int n, double alpha, T_* X) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)n * n) return;
  const long i = k % n, j = k / n;
  X[k] = (T_)(alpha * A[i * n + j]);
}

// Af <- A rounded to FP32 (the operand of the FP32 Newton-Schulz GEMM)
__global__ void k_round_f32(const double* __restrict__ A, long nn, float* __restrict__ Af) {
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < nn;
       k += (long)gridDim.x * blockDim.x)
    Af[k] = (float)A[k];
}

// out[0] = max_j sum_i |A_ij| (1-norm), out[1] = max_i sum_j |A_ij| (inf-norm),
// one workgroup (run once per inverse build).
// |A|_1 and |A|_inf of a column-major A: out[0], out[1] hold the bit
// patterns of the maxima (non-negative doubles order like their bits; an
// atomic max is order-independent, so the result is deterministic).
__global__ void k_norms_zero(unsigned long long* out) {
  out[0] = 0ull;
  out[1] = 0ull;
}

// One workgroup per column: the column's absolute sum (fixed tree order).
__global__ __launch_bounds__(256) void k_colsum_max(const double* __restrict__ A, int n,
                                                    unsigned long long* __restrict__ out) {
  __shared__ double sw[4];
  const int c = blockIdx.x;
  double acc = 0.0;
  for (int r = threadIdx.x; r < n; r += 256) acc += fabs(A[(long)c * n + r]);
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicMax(out, (unsigned long long)__double_as_longlong((sw[0] + sw[1]) + (sw[2] + sw[3])));
}

// part[s][i] = sum_{j in column slice s} |A[j n + i]|: the row sums of |A|
// over column slices (one row per thread, coalesced across the workgroup).
__global__ __launch_bounds__(256) void k_absrow_part(const double* __restrict__ A, int n, int cols,
                                                     double* __restrict__ part) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int j0 = blockIdx.y * cols;
  const int j1 = min(n, j0 + cols);
  const double* p = A + (long)j0 * n + i;
  double acc = 0.0;
  for (int j = j0; j < j1; ++j, p += n) acc += fabs(*p);
  part[(long)blockIdx.y * n + i] = acc;
}
// part[s][i] = sum_{j in column slice s} M[j n + i] v[j] (column-major M,
// FP64 or FP32 elements, FP64 products and sums), s = blockIdx.y; one row
// per thread, columns in increasing order.
template <typename T_>
__global__ __launch_bounds__(256) void k_gemv_part(const T_* __restrict__ M, int n,
                                                   const double* __restrict__ v, int cols,
                                                   double* __restrict__ part,
                                                   const long long* flag) {
  if (*flag != 0) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int j0 = blockIdx.y * cols;
  const int j1 = min(n, j0 + cols);
  const T_* p = M + (long)j0 * n + i;
  double acc = 0.0;
  int j = j0;
  for (; j + 4 <= j1; j += 4) {
    const double a0 = p[0], a1 = p[n], a2 = p[2L * n], a3 = p[3L * n];
    acc = fma(a0, v[j], acc);
    acc = fma(a1, v[j + 1], acc);
    acc = fma(a2, v[j + 2], acc);
    acc = fma(a3, v[j + 3], acc);
    p += 4L * n;
  }
  for (; j < j1; ++j, p += n) acc = fma((double)*p, v[j], acc);
  part[(long)blockIdx.y * n + i] = acc;
}

// Sum over s of part[s][i] in increasing s: 16 rows per workgroup, 16 groups
// of consecutive slices per row, combined in group order.
__device__ inline double sum_parts(const double* __restrict__ part, int S, int n, int i,
                                   double (*lds)[17]) {
  const int r = threadIdx.x & 15, g = threadIdx.x >> 4;
  const int per = (S + kRedGroups - 1) / kRedGroups;
  const int s0 = g * per, s1 = min(S, s0 + per);
  double acc = 0.0;
  if (i < n)
#pragma unroll 4
    for (int s = s0; s < s1; ++s) acc += part[(long)s * n + i];
  lds[g][r] = acc;
  __syncthreads();
  double t = 0.0;
  if (g == 0)
    for (int q = 0; q < kRedGroups; ++q) t += lds[q][r];
  return t;  // valid in group 0
}

// r = b - sum_s part[s]
__global__ __launch_bounds__(256) void k_residual_sum(const double* __restrict__ part, int S,
                                                      int n, const double* __restrict__ b,
                                                      double* __restrict__ r,
                                                      const long long* flag) {
  __shared__ double lds[kRedGroups][17];
  if (*flag != 0) return;
  const int i = blockIdx.x * 16 + (threadIdx.x & 15);
  const double t = sum_parts(part, S, n, i, lds);
  if (threadIdx.x < 16 && i < n) r[i] = b[i] - t;
}

__device__ void check_sweep(unsigned long long* norms, int k, double prev, double tol,
                            long long* flag);

// x += sum_s part[s]; norms[2k], norms[2k + 1] <- max |dx|, max |x| (after
// the update); the last workgroup to finish then checks the sweep
// (check_sweep), so a sweep is four launches.
__global__ __launch_bounds__(256) void k_update_sum(const double* __restrict__ part, int S,
                                                    int n, double* __restrict__ x,
                                                    unsigned long long* norms, int k,
                                                    double prev, double tol, long long* flag,
                                                    unsigned* ctr) {
  __shared__ double lds[kRedGroups][17];
  if (*flag != 0) return;
  unsigned long long* out = norms + 2 * k;
  const int i = blockIdx.x * 16 + (threadIdx.x & 15);
  const double dx = sum_parts(part, S, n, i, lds);
  if (threadIdx.x < 64) {  // wave 0; its lanes 16..63 contribute zeros
    double a = 0.0, m = 0.0;
    if (threadIdx.x < 16 && i < n) {
      const double v = x[i] + dx;
      x[i] = v;
      a = fabs(dx);
      m = fabs(v);
    }
    for (int off = 32; off > 0; off >>= 1) {
      a = fmax(a, __shfl_xor(a, off));
      m = fmax(m, __shfl_xor(m, off));
    }
    if (threadIdx.x == 0) {
      atomic_max_nonneg(out, a);
      atomic_max_nonneg(out + 1, m);
      __threadfence();
      if (atomicAdd(ctr, 1u) == gridDim.x - 1) {  // every workgroup's maxima are in
        *ctr = 0u;
        __threadfence();
        check_sweep(norms, k, prev, tol, flag);
      }
    }
  }
}

// out[1] <- max_i sum_s part[s][i] (the inf-norm from k_absrow_part's slices)
__global__ __launch_bounds__(256) void k_rowmax_sum(const double* __restrict__ part, int S, int n,
                                                    unsigned long long* out) {
  __shared__ double lds[kRedGroups][17];
  const int i = blockIdx.x * 16 + (threadIdx.x & 15);
  double t = sum_parts(part, S, n, i, lds);
  if (threadIdx.x < 64) {
    if (!(threadIdx.x < 16 && i < n)) t = 0.0;
    for (int off = 32; off > 0; off >>= 1) t = fmax(t, __shfl_xor(t, off));
    if (threadIdx.x == 0) atomic_max_nonneg(out + 1, t);
  }
}

// After sweep k: raise the flag on convergence (max|dx| <= tol max|x|, or a
// stall at the rounding level: contraction above 1/2, growth included, with
// max|dx| <= 1e3 tol max|x|) or divergence (a non-finite correction, or one
// that grows above the rounding level).  Run by the last workgroup of the
// sweep's update (the norms are read as atomics: the other workgroups'
// maxima reach it through the device-scope atomics and fences).
__device__ void check_sweep(unsigned long long* norms, int k, double prev, double tol,
                            long long* flag) {
  auto ld = [](unsigned long long* q) {
    return __longlong_as_double(
        (long long)__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  };
  const double d = ld(norms + 2 * k);
  const double m = ld(norms + 2 * k + 1);
  const double p = k > 0 ? ld(norms + 2 * k - 2) : prev;
  if (!(d <= 1e300) || !(m <= 1e300)) {
    *flag = -1 - k;
    return;
  }
  const double rho = p > 0.0 ? d / p : -1.0;
  if (d <= tol * m || (rho > 0.5 && d <= 1e3 * tol * m)) {
    *flag = 1 + k;
    return;
  }
  if (p > 0.0 && d > p) *flag = -1 - k;
}

// x <- 2 x - xold, xold <- x (linear extrapolation of the start)
__global__ void k_extrapolate(double* x, double* xold, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = x[i];
  x[i] = 2.0 * v - xold[i];
  xold[i] = v;
}

__global__ void k_copy_vec(const double* a, double* b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_fill_vec(double* a, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Reads n bytes from the device after the stream's work; false with the
// reason in *err on failure.
bool read_back(hipStream_t s, void* host, const void* dev, size_t n, const char* what,
               std::string* err) {
  hipError_t e = hipMemcpyAsync(host, dev, n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) return true;
  *err = std::string("inverse solver: ") + what + ": " + hipGetErrorString(e);
  return false;
}

double bits_to_double(unsigned long long u) {
  double d;
  std::memcpy(&d, &u, sizeof(d));
  return d;
}

}  // namespace

InvSolver::~InvSolver() { release(); }

void InvSolver::release() {
  void* bufs[] = {X_, T_, X2_, Af_, Ab_, Xb_, Tb_, r_, xold_, part_, stat_};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  X_ = T_ = X2_ = nullptr;
  Af_ = nullptr;
  Ab_ = Xb_ = Tb_ = nullptr;
  r_ = xold_ = part_ = nullptr;
  stat_ = nullptr;
  n_ = 0;
  wide_ = false;
  valid_ = false;
  streak_ = 0;
}

int InvSolver::prepare(hipStream_t s, int n, bool wide, std::string* err) {
  if (n == n_ && wide == wide_) return UVLM_OK;
  release();
  // split-column shape of the matrix-vector products: ~1024 workgroups
  const int rb = cdiv(n, 256);
  const int want = std::max(1, std::min(n, 1024 / rb));
  cols_ = cdiv(n, want);
  S_ = cdiv(n, cols_);
  const size_t nn = (size_t)n * n;
  const size_t el = wide ? sizeof(double) : sizeof(float);
  if (hipMalloc(&X_, el * nn) != hipSuccess || hipMalloc(&T_, el * nn) != hipSuccess ||
      hipMalloc(&X2_, el * nn) != hipSuccess ||
      (!wide && hipMalloc((void**)&Af_, sizeof(float) * nn) != hipSuccess) ||
      (!wide && hipMalloc((void**)&Ab_, sizeof(unsigned short) * nn) != hipSuccess) ||
      (!wide && hipMalloc((void**)&Xb_, sizeof(unsigned short) * nn) != hipSuccess) ||
      (!wide && hipMalloc((void**)&Tb_, sizeof(unsigned short) * nn) != hipSuccess) ||
      hipMalloc((void**)&r_, sizeof(double) * n) != hipSuccess ||
      hipMalloc((void**)&xold_, sizeof(double) * n) != hipSuccess ||
      hipMalloc((void**)&part_, sizeof(double) * (size_t)S_ * n) != hipSuccess ||
      hipMalloc((void**)&stat_, sizeof(unsigned long long) * kStatWords) != hipSuccess) {
    (void)hipGetLastError();
    release();
    *err = "inverse solver allocation failed";
    return UVLM_ERR_OOM;
  }
  n_ = n;
  wide_ = wide;
  // the sweep counter starts at zero (its last user resets it)
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, stat_, kStatWords);
  return UVLM_OK;
}

// One Newton-Schulz step from the current X; returns max |I - A X_old| (or
// < 0 on error, reason in last_err_).  FP32: A is rounded to FP32 and both
// products run on the FP32 matrix cores, or with low on the BF16 matrix
// cores (FP32 data rounded to BF16 operands, FP32 accumulation: the step
// then squares the deviation down to a floor of ~1e-2, measured 0.007 in
// spectral radius on the tandem wings' matrix, enough to contract each
// sweep ~100x).  The step synchronises the stream.
double InvSolver::newton_schulz(hipStream_t s, const double* A, bool low, bool wait) {
  last_err_.clear();
  const int n = n_;
  const long nn = (long)n * n;
  const int prec = wide_ ? GEMM_F64 : low ? GEMM_BF16 : GEMM_F32;
  const void* Aop = A;
  if (!wide_) {
    hipLaunchKernelGGL(k_round_f32, dim3(std::min(2048, cdiv(nn, 256))), dim3(256), 0, s, A, nn,
                       Af_);
    Aop = Af_;
  }
  // BF16 products read operands rounded once to BF16 (half the bytes of
  // the FP32 data per tile; the same roundings as rounding while staging)
  const bool b16 = prec == GEMM_BF16;
  if (b16) {
    launch_round_bf16(s, Af_, nn, Ab_);
    launch_round_bf16(s, (const float*)X_, nn, Xb_);
  }
  const GemmArgs ax{n, n, n, b16 ? (const void*)Ab_ : Aop, n, b16 ? (const void*)Xb_ : X_, n,
                    T_, n, 1.0, 0.0, nullptr};  // T = A X
  launch_gemm(s, b16 ? GEMM_BF16_IN : prec, &ax, 1);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, stat_ + kNsSlot, 1);
  const dim3 g2(std::min(1024, cdiv(nn, 256)));
  if (wide_)
    hipLaunchKernelGGL(k_two_minus_dev<double>, g2, dim3(256), 0, s, (double*)T_, n,
                       stat_ + kNsSlot);
  else
    hipLaunchKernelGGL(k_two_minus_dev<float>, g2, dim3(256), 0, s, (float*)T_, n,
                       stat_ + kNsSlot);
  if (b16) launch_round_bf16(s, (const float*)T_, nn, Tb_);
  const GemmArgs xt{n, n, n, b16 ? (const void*)Xb_ : X_, n, b16 ? (const void*)Tb_ : T_, n,
                    X2_, n, 1.0, 0.0, nullptr};  // X2 = X (2I - A X)
  launch_gemm(s, b16 ? GEMM_BF16_IN : prec, &xt, 1);
  std::swap(X_, X2_);
  ++refreshes_;
  if (!wait) return 0.0;
  unsigned long long u = 0;
  if (!read_back(s, &u, stat_ + kNsSlot, sizeof(u), "Newton-Schulz step", &last_err_))
    return -1.0;
  return bits_to_double(u);
}

// X from X0 = A^T / (||A||_1 ||A||_inf) by Newton-Schulz steps until
// max |I - A X| < 1e-3.  In FP32 a deviation that stops falling above that
// (the FP32 rounding of an ill-conditioned A) switches to FP64 and starts over.
int InvSolver::build(hipStream_t s, const double* A, std::string* err) {
  const int n = n_;
  double* nrm = reinterpret_cast<double*>(stat_);  // words 0 and 1
  unsigned long long* nb = reinterpret_cast<unsigned long long*>(stat_);
  hipLaunchKernelGGL(k_norms_zero, dim3(1), dim3(1), 0, s, nb);
  hipLaunchKernelGGL(k_colsum_max, dim3(n), dim3(256), 0, s, A, n, nb);
  hipLaunchKernelGGL(k_absrow_part, dim3(cdiv(n, 256), S_), dim3(256), 0, s, A, n, cols_, part_);
  hipLaunchKernelGGL(k_rowmax_sum, dim3(cdiv(n, 16)), dim3(256), 0, s, (const double*)part_, S_,
                     n, nb);
  double h[2] = {0, 0};
  if (!read_back(s, h, nrm, sizeof(h), "matrix norms", err)) return UVLM_ERR_HIP;
  if (!(h[0] > 0.0) || !(h[1] > 0.0)) {
    *err = "Matrix is not invertible (zero matrix)";
    return UVLM_ERR_SINGULAR;
  }
  const long nn = (long)n * n;
  if (wide_)
    hipLaunchKernelGGL(k_scaled_transpose<double>, dim3(cdiv(nn, 256)), dim3(256), 0, s, A, n,
                       1.0 / (h[0] * h[1]), (double*)X_);
  else
    hipLaunchKernelGGL(k_scaled_transpose<float>, dim3(cdiv(nn, 256)), dim3(256), 0, s, A, n,
                       1.0 / (h[0] * h[1]), (float*)X_);
  double best = 1e300;
  int since = 0;  // steps since the deviation last fell below 0.7 of its best
  for (int it = 0; it < 200; ++it) {
    const double e = newton_schulz(s, A, false, true);
    if (e < 0.0) {
      *err = last_err_.empty() ? "inverse solver: Newton-Schulz GEMM failed" : last_err_;
      return UVLM_ERR_HIP;
    }
    if (e < 1e-3) {  // the step just taken squared a deviation below 1e-3
      valid_ = true;
      return UVLM_OK;
    }
    if (!(e < 1e300)) break;
    // below 1/4 the deviation falls at least by half per step (its spectral
    // part squares); four steps without a 30 % gain there, or 60 steps in
    // all (1 - e doubles per step from 1 - 2^-60), mean the FP32 floor
    if (e < 0.7 * best) {
      best = e;
      since = 0;
    } else {
      ++since;
    }
    if (!wide_ && ((best < 0.25 && since >= 4) || it >= 60)) {
      int rc = prepare(s, n, true, err);
      if (rc != UVLM_OK) return rc;
      return build(s, A, err);
    }
  }
  *err = "inverse solver: Newton-Schulz did not converge (singular matrix?)";
  return UVLM_ERR_NOCONV;
}

// Enqueues sweep k of a batch: norms of its correction to stat_[2k], [2k+1],
// then the convergence check (prev: last correction of the previous batch).
void InvSolver::enqueue_sweep(hipStream_t s, const double* A, const double* b, double* x,
                              int k, double prev, double tol) {
  const int n = n_;
  const long long* flag = reinterpret_cast<const long long*>(stat_ + kFlagSlot);
  const dim3 g(cdiv(n, 256), S_);
  const int rb = cdiv(n, 16);
  hipLaunchKernelGGL(k_gemv_part<double>, g, dim3(256), 0, s, A, n, (const double*)x, cols_,
                     part_, flag);
  hipLaunchKernelGGL(k_residual_sum, dim3(rb), dim3(256), 0, s, (const double*)part_, S_, n, b,
                     r_, flag);
  if (wide_)
    hipLaunchKernelGGL(k_gemv_part<double>, g, dim3(256), 0, s, (const double*)X_, n,
                       (const double*)r_, cols_, part_, flag);
  else
    hipLaunchKernelGGL(k_gemv_part<float>, g, dim3(256), 0, s, (const float*)X_, n,
                       (const double*)r_, cols_, part_, flag);
  hipLaunchKernelGGL(k_update_sum, dim3(rb), dim3(256), 0, s, (const double*)part_, S_, n, x,
                     stat_, k, prev, tol, reinterpret_cast<long long*>(stat_ + kFlagSlot),
                     reinterpret_cast<unsigned*>(stat_ + kCtrSlot));
}

int InvSolver::solve(hipStream_t s, const double* A, int n, const double* b, double* x,
                     bool warm, double tol, int max_sweeps, int* sweeps, std::string* err) {
  int rc = prepare(s, n, n == n_ && wide_, err);
  if (rc != UVLM_OK) return rc;
  if (sweeps) *sweeps = 0;
  const int nb = cdiv(n, 256);
  if (!warm) {
    streak_ = 0;
    hipLaunchKernelGGL(k_fill_vec, dim3(nb), dim3(256), 0, s, x, n, 0.0);
  } else if (streak_ >= 2) {
    hipLaunchKernelGGL(k_extrapolate, dim3(nb), dim3(256), 0, s, x, xold_, n);
  } else if (streak_ == 1) {
    hipLaunchKernelGGL(k_copy_vec, dim3(nb), dim3(256), 0, s, (const double*)x, xold_, n);
  }
  if (!valid_ && (rc = build(s, A, err)) != UVLM_OK) return rc;
  else if (warm && end_rho_ > kPreRefreshRho && n <= kPreRefreshMaxN)
    (void)newton_schulz(s, A, true, false);  // queued ahead of the sweeps, no read-back
  double prev = -1.0;
  int refreshed = 0, done = 0;
  bool rebuilt = false;
  // first batch: the last solve's count and two more (sweeps after the
  // converged one are skipped on the device)
  int batch = std::min(kMaxBatch, std::max(4, last_sweeps_ + 2));
  while (done < max_sweeps) {
    batch = std::max(1, std::min(batch, max_sweeps - done));
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, stat_, 2 * kMaxBatch);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, stat_ + kFlagSlot, 1);
    for (int k = 0; k < batch; ++k) enqueue_sweep(s, A, b, x, k, prev, tol);
    unsigned long long u[kStatWords];
    if (!read_back(s, u, stat_, sizeof(u), "correction norms", err)) return UVLM_ERR_HIP;
    long long flag;
    std::memcpy(&flag, &u[kFlagSlot], sizeof(flag));
    const int ran = flag > 0 ? (int)flag : flag < 0 ? (int)(-flag) : batch;
    done += ran;
    sweeps_total_ += ran;
    if (sweeps) *sweeps = done;
    if (flag > 0) {
      last_sweeps_ = done;
      if (warm) ++streak_;
      // contraction of this solve: geometric mean over the sweeps before
      // the converged one (that one sits at the rounding level)
      const int kc = (int)flag - 1;
      const double d0 = bits_to_double(u[0]);
      if (kc >= 2 && d0 > 0.0) {
        const double dl = bits_to_double(u[2 * (kc - 1)]);
        end_rho_ = std::pow(std::max(dl, 1e-300) / d0, 1.0 / (kc - 1));
      } else {
        end_rho_ = kc >= 2 ? 1.0 : 0.0;
      }
      if (refreshed > 0 || prev > 0.0) end_rho_ = std::max(end_rho_, kPreRefreshRho * 2);
      return UVLM_OK;
    }
    if (flag < 0) {
      // growing corrections: X no longer inverts A closely enough
      if (rebuilt) break;
      rebuilt = true;
      ++rebuilds_;
      streak_ = 0;
      hipLaunchKernelGGL(k_fill_vec, dim3(nb), dim3(256), 0, s, x, n, 0.0);
      valid_ = false;
      if ((rc = build(s, A, err)) != UVLM_OK) return rc;
      prev = -1.0;
      batch = 4;
      continue;
    }
    const double d = bits_to_double(u[2 * (batch - 1)]);
    const double m = bits_to_double(u[2 * (batch - 1) + 1]);
    const double dp = batch > 1 ? bits_to_double(u[2 * (batch - 2)]) : prev;
    const double rho = dp > 0.0 ? d / dp : -1.0;
    prev = d;
    // slow contraction: improve X (each step squares I - A X); the first
    // refresh of a solve on the BF16 matrix cores, later ones in FP32 (a
    // contraction the BF16 floor does not bring under the threshold)
    if (rho > kRefreshRho && refreshed < 3) {
      if (newton_schulz(s, A, refreshed == 0, true) < 0.0) {
        *err = last_err_.empty() ? "inverse solver: Newton-Schulz GEMM failed" : last_err_;
        return UVLM_ERR_HIP;
      }
      ++refreshed;
      prev = -1.0;  // the contraction changes with X
      batch = 4;
      continue;
    }
    // sweeps still needed at the observed contraction, and two more
    if (rho > 0.0) {
      const double need = std::log(std::max(tol * m, 1e-300) / d) / std::log(rho);
      batch = (int)std::ceil(std::max(1.0, need)) + 2;
    } else {
      batch = 4;
    }
    batch = std::min(batch, kMaxBatch);
  }
  streak_ = 0;
  last_sweeps_ = 0;
  *err = "inverse solver: refinement did not converge";
  return UVLM_ERR_NOCONV;
}

}  // namespace uvlm
