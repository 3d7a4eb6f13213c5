// gemm.h -- matrix-core GEMM of the dense solves (gemm.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace uvlm {

// C = alpha A B + beta C, column-major, A m x k, B k x n.  beta == 0 does
// not read C.  bad (nullable): set to 1 when a result r has |r| > 1 (the
// LU's pivoting check on L21; NaN does not set it).
//
// In place: C may be the memory of A when one tile spans all n columns
// (n <= 64), or the memory of B when one tile spans all m rows (m <= 64):
// a workgroup reads every element of its rows (columns) before it stores.
struct GemmArgs {
  int m, n, k;
  const void* A;
  long lda;
  const void* B;
  long ldb;
  void* C;
  long ldc;
  double alpha, beta;
  int* bad;
};

enum { GEMM_F64 = 0, GEMM_F32 = 1, GEMM_BF16 = 2, GEMM_BF16_IN = 3 };

// One launch for one or two problems (count 1 or 2), in FP64 on
// v_mfma_f64_16x16x4f64, FP32 on v_mfma_f32_16x16x4f32, or FP32 data with
// BF16 operands on v_mfma_f32_16x16x32_bf16 (FP32 accumulation).
// GEMM_BF16_IN: A and B already BF16 (2-byte elements, as launch_round_bf16
// writes them), C FP32; the same products as GEMM_BF16 of the FP32 data.
void launch_gemm(hipStream_t s, int prec, const GemmArgs* p, int count);

// y[i] = x[i] rounded to BF16 (nearest even), the GEMM's operand rounding.
void launch_round_bf16(hipStream_t s, const float* x, long n, unsigned short* y);

}  // namespace uvlm
