// inv_solve.hip -- dense circulation solve by iterative refinement on a kept
// approximate inverse (UVLM_SOLVE_F64_INV).
//
// The influence matrix of a time-stepping run changes little from one step
// to the next.  The solver keeps X ~ A^-1 of a recent step and refines
//     x_{k+1} = x_k + X (b - A x_k)
// from the previous step's circulations, with the residual in FP64: the
// error contracts by ||I - X A|| per sweep and the result solves the current
// A x = b to FP64 accuracy.  Each sweep is two matrix-vector products over
// all of HBM bandwidth instead of a latency-bound LU.  X is improved in
// place by Newton-Schulz steps X <- X (2I - A X) (two library GEMMs) when
// the contraction slows, and built initially from X0 = A^T / (||A||_1
// ||A||_inf), for which the iteration converges for every non-singular A.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/uvlm_gpu.h"
#include "internal.h"

namespace uvlm {
namespace {

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// T <- 2I - T
__global__ void k_two_minus(double* T, int n) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)n * n) return;
  const long i = k % n, j = k / n;
  T[k] = (i == j ? 2.0 : 0.0) - T[k];
}

// X <- alpha A^T
__global__ void k_scaled_transpose(const double* A, int n, double alpha, double* X) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)n * n) return;
  const long i = k % n, j = k / n;
  X[k] = alpha * A[i * n + j];
}

// out[0] = max_j sum_i |A_ij| (1-norm), out[1] = max_i sum_j |A_ij| (inf-norm),
// one workgroup.
__global__ __launch_bounds__(1024) void k_norms1inf(const double* A, int n, double* out) {
  __shared__ double s1[16], s2[16];
  double m1 = 0.0, m2 = 0.0;
  for (int c = threadIdx.x; c < n; c += 1024) {
    double cs = 0.0, rs = 0.0;
    for (int r = 0; r < n; ++r) {
      cs += fabs(A[(long)c * n + r]);
      rs += fabs(A[(long)r * n + c]);
    }
    m1 = fmax(m1, cs);
    m2 = fmax(m2, rs);
  }
  for (int off = 32; off > 0; off >>= 1) {
    m1 = fmax(m1, __shfl_xor(m1, off));
    m2 = fmax(m2, __shfl_xor(m2, off));
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
    out[0] = m1;
    out[1] = m2;
  }
}

// out[0] <- bits of max |I - T|; atomic max on non-negative doubles
__global__ void k_max_dev_identity(const double* T, int n, unsigned long long* out) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (k < (long)n * n) {
    const long i = k % n, j = k / n;
    e = fabs((i == j ? 1.0 : 0.0) - T[k]);
  }
  for (int off = 32; off > 0; off >>= 1) e = fmax(e, __shfl_xor(e, off));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(e));
}

// x += dx; out[0], out[1] <- bits of max |dx|, max |x| (after the update)
__global__ void k_add_norms(double* x, const double* dx, int n, unsigned long long* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0.0, b = 0.0;
  if (k < n) {
    const double v = x[k] + dx[k];
    x[k] = v;
    a = fabs(dx[k]);
    b = fabs(v);
  }
  for (int off = 32; off > 0; off >>= 1) {
    a = fmax(a, __shfl_xor(a, off));
    b = fmax(b, __shfl_xor(b, off));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(out, (unsigned long long)__double_as_longlong(a));
    atomicMax(out + 1, (unsigned long long)__double_as_longlong(b));
  }
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
  if (handle_) (void)rocblas_destroy_handle((rocblas_handle)handle_);
  handle_ = nullptr;
  void* bufs[] = {X_, T_, X2_, r_, dx_, stat_};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  X_ = T_ = X2_ = r_ = dx_ = nullptr;
  stat_ = nullptr;
  n_ = 0;
  valid_ = false;
}

int InvSolver::prepare(hipStream_t s, int n, std::string* err) {
  if (!handle_) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) {
      *err = "rocblas_create_handle failed";
      return UVLM_ERR_HIP;
    }
    handle_ = h;
  }
  if (rocblas_set_stream((rocblas_handle)handle_, s) != rocblas_status_success) {
    *err = "rocblas_set_stream failed";
    return UVLM_ERR_HIP;
  }
  if (n == n_) return UVLM_OK;
  void* h = handle_;
  handle_ = nullptr;
  release();
  handle_ = h;
  const size_t nn = (size_t)n * n;
  if (hipMalloc(&X_, sizeof(double) * nn) != hipSuccess ||
      hipMalloc(&T_, sizeof(double) * nn) != hipSuccess ||
      hipMalloc(&X2_, sizeof(double) * nn) != hipSuccess ||
      hipMalloc(&r_, sizeof(double) * n) != hipSuccess ||
      hipMalloc(&dx_, sizeof(double) * n) != hipSuccess ||
      hipMalloc((void**)&stat_, sizeof(unsigned long long) * 4) != hipSuccess) {
    (void)hipGetLastError();
    *err = "inverse solver allocation failed";
    return UVLM_ERR_OOM;
  }
  n_ = n;
  valid_ = false;
  return UVLM_OK;
}

// One Newton-Schulz step from the current X; returns max |I - A X_old| (or
// < 0 on error).  The step synchronises the stream.
double InvSolver::newton_schulz(hipStream_t s, const double* A) {
  last_err_.clear();
  rocblas_handle h = (rocblas_handle)handle_;
  const int n = n_;
  const double one = 1.0, zero = 0.0;
  if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, A, n, X_,
                    n, &zero, T_, n) != rocblas_status_success)
    return -1.0;
  const long nn = (long)n * n;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, stat_, 4);
  hipLaunchKernelGGL(k_max_dev_identity, dim3(cdiv(nn, 256)), dim3(256), 0, s, T_, n, stat_);
  hipLaunchKernelGGL(k_two_minus, dim3(cdiv(nn, 256)), dim3(256), 0, s, T_, n);
  if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, X_, n, T_,
                    n, &zero, X2_, n) != rocblas_status_success)
    return -1.0;
  std::swap(X_, X2_);
  unsigned long long u = 0;
  if (!read_back(s, &u, stat_, sizeof(u), "Newton-Schulz step", &last_err_)) return -1.0;
  ++refreshes_;
  return bits_to_double(u);
}

int InvSolver::build(hipStream_t s, const double* A, std::string* err) {
  const int n = n_;
  double* nrm = reinterpret_cast<double*>(stat_);  // 2 of its 4 words
  hipLaunchKernelGGL(k_norms1inf, dim3(1), dim3(1024), 0, s, A, n, nrm);
  double h[2] = {0, 0};
  if (!read_back(s, h, nrm, sizeof(h), "matrix norms", err)) return UVLM_ERR_HIP;
  if (!(h[0] > 0.0) || !(h[1] > 0.0)) {
    *err = "Matrix is not invertible (zero matrix)";
    return UVLM_ERR_SINGULAR;
  }
  const long nn = (long)n * n;
  hipLaunchKernelGGL(k_scaled_transpose, dim3(cdiv(nn, 256)), dim3(256), 0, s, A, n,
                     1.0 / (h[0] * h[1]), X_);
  for (int it = 0; it < 200; ++it) {
    const double e = newton_schulz(s, A);
    if (e < 0.0) {
      *err = last_err_.empty() ? "inverse solver: Newton-Schulz GEMM failed" : last_err_;
      return UVLM_ERR_HIP;
    }
    if (e < 1e-3) {  // X was already good; the step just taken squares that
      valid_ = true;
      return UVLM_OK;
    }
    if (!(e < 1e300)) break;
  }
  *err = "inverse solver: Newton-Schulz did not converge (singular matrix?)";
  return UVLM_ERR_NOCONV;
}

int InvSolver::sweep(hipStream_t s, const double* A, const double* b, double* x, double* d,
                     double* m, std::string* err) {
  rocblas_handle h = (rocblas_handle)handle_;
  const int n = n_;
  const double one = 1.0, mone = -1.0, zero = 0.0;
  if (hipMemcpyAsync(r_, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
      rocblas_dgemv(h, rocblas_operation_none, n, n, &mone, A, n, x, 1, &one, r_, 1) !=
          rocblas_status_success ||
      rocblas_dgemv(h, rocblas_operation_none, n, n, &one, X_, n, r_, 1, &zero, dx_, 1) !=
          rocblas_status_success) {
    *err = "inverse solver: matrix-vector product failed";
    return UVLM_ERR_HIP;
  }
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, stat_, 2);
  hipLaunchKernelGGL(k_add_norms, dim3(cdiv(n, 256)), dim3(256), 0, s, x, (const double*)dx_,
                     n, stat_);
  unsigned long long u[2] = {0, 0};
  if (!read_back(s, u, stat_, sizeof(u), "correction norms", err)) return UVLM_ERR_HIP;
  *d = bits_to_double(u[0]);
  *m = bits_to_double(u[1]);
  ++sweeps_total_;
  return UVLM_OK;
}

int InvSolver::solve(hipStream_t s, const double* A, int n, const double* b, double* x,
                     bool warm, double tol, int max_sweeps, int* sweeps, std::string* err) {
  int rc = prepare(s, n, err);
  if (rc != UVLM_OK) return rc;
  if (sweeps) *sweeps = 0;
  if (!warm && hipMemsetAsync(x, 0, sizeof(double) * n, s) != hipSuccess) {
    *err = "hipMemsetAsync failed";
    return UVLM_ERR_HIP;
  }
  if (!valid_ && (rc = build(s, A, err)) != UVLM_OK) return rc;
  double prev = -1.0;
  int refreshed = 0;
  bool rebuilt = false;
  for (int k = 0; k < max_sweeps; ++k) {
    double d = 0.0, m = 0.0;
    if ((rc = sweep(s, A, b, x, &d, &m, err)) != UVLM_OK) return rc;
    if (sweeps) *sweeps = k + 1;
    if (!(d <= 1e300) || (prev > 0.0 && d > prev)) {
      // growing corrections: X no longer inverts A closely enough
      if (rebuilt) break;
      rebuilt = true;
      ++rebuilds_;
      if (hipMemsetAsync(x, 0, sizeof(double) * n, s) != hipSuccess) {
        *err = "hipMemsetAsync failed";
        return UVLM_ERR_HIP;
      }
      valid_ = false;
      if ((rc = build(s, A, err)) != UVLM_OK) return rc;
      prev = -1.0;
      continue;
    }
    if (d <= tol * m) return UVLM_OK;
    // stalled at the rounding level of the residual
    if (prev > 0.0 && d > 0.5 * prev && d <= 1e3 * tol * m) return UVLM_OK;
    // slow contraction: improve X (each step squares I - A X)
    if (prev > 0.0 && d > 0.05 * prev && refreshed < 3) {
      if (newton_schulz(s, A) < 0.0) {
        *err = last_err_.empty() ? "inverse solver: Newton-Schulz GEMM failed" : last_err_;
        return UVLM_ERR_HIP;
      }
      ++refreshed;
    }
    prev = d;
  }
  *err = "inverse solver: refinement did not converge";
  return UVLM_ERR_NOCONV;
}

}  // namespace uvlm
