// Dense-layer building blocks of the learned step: a BF16 MFMA GEMM with
// fused epilogues, and the small element-wise / graph kernels around it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tit_amd {

typedef __bf16 bf16;

enum NnAct { kActNone = 0, kActHswish = 1 };

// The network's activation, hard-swish x min(max(x / 6 + 1/2, 0), 1): three
// full-rate VALU operations (an FMA, a clamp, a product), used by every
// layer (training GEMM epilogues, the derivative, the fused inference kernel).
__device__ __forceinline__ float nn_gate(float x) {
  return __builtin_amdgcn_fmed3f(__builtin_fmaf(x, 1.0f / 6.0f, 0.5f), 0.0f, 1.0f);
}
__device__ __forceinline__ float nn_act(float x) { return x * nn_gate(x); }
// 0 below -3, x / 3 + 1/2 on (-3, 3), 1 above 3
__device__ __forceinline__ float nn_act_grad(float x) {
  const float g = nn_gate(x);
  return (x > -3.0f && x < 3.0f) ? __builtin_fmaf(x, 1.0f / 6.0f, g) : g;
}

// C[M x N] = act(op(A)[M x K] . op(B)[K x N] + bias + G1[r1[m]] + G2[r2[m]])
// with BF16 operands and FP32 accumulation.  op(A) is A (row-major, lda) or,
// with transA, the transpose of a row-major K x M matrix; likewise for B.
// Outputs (any may be null): pre (FP32 before the activation), out32 (FP32
// after), out16 (BF16 after).  With accumulate, out32 += result through
// atomics (split-K over k_splits slices of K; out32 must hold the running
// sum).
struct GemmArgs {
  int M = 0, N = 0, K = 0;
  const bf16* A = nullptr;
  int lda = 0;
  bool transA = false;
  const bf16* B = nullptr;
  int ldb = 0;
  bool transB = false;
  const float* bias = nullptr;
  const float* G1 = nullptr;  // row-gathered addends, ld = N
  const uint32_t* r1 = nullptr;
  const float* G2 = nullptr;
  const uint32_t* r2 = nullptr;
  int act = kActNone;
  float* pre = nullptr;
  int ldpre = 0;
  float* out32 = nullptr;
  int ldo32 = 0;
  bf16* out16 = nullptr;
  int ldo16 = 0;
  bool accumulate = false;
  int k_splits = 1;
};

void gemm_bf16(hipStream_t s, const GemmArgs& g);

// dx = dy * act'(pre) for an M x N block; dy row m is read from row
// m / group (leading dimension ld_dy), so group > 1 broadcasts one upstream
// row to `group` rows.  Writes FP32 (dx32) and/or BF16 (dx16) copies, both
// M x N.  In place (dx32 == dy) is allowed when group == 1 and ld_dy == N.
void act_backward(hipStream_t s, int M, int N, const float* dy, int ld_dy,
                   int group, const float* pre, float* dx32, bf16* dx16);
// out[r][n] = sum_{k < group} in[r * group + k][n]; optional BF16 copy with
// its own leading dimension and column offset.
void group_sum(hipStream_t s, int rows, int N, int group, const float* in,
               float* out, bf16* out16, int ld16, int col16);
// out[idx[m]][n] += in[m][n] (atomics).
void scatter_add_rows(hipStream_t s, int M, int N, const float* in,
                      const uint32_t* idx, float* out);
// col_sum[n] += sum_m in[m][n].
void column_sum(hipStream_t s, int M, int N, const float* in, float* col_sum);
// BF16 copy of an FP32 block into a wider matrix: dst[m][col0 + n] = src[m][n].
void to_bf16_block(hipStream_t s, int M, int N, const float* src, int lds,
                   bf16* dst, int ldd, int col0);
// dst[m][n] = src[m][col0 + n] (FP32 column block copy).
void copy_cols(hipStream_t s, int M, int N, const float* src, int lds, int col0,
               float* dst, int ldd);
// dst[m][n] += src[m][col0 + n].
void add_cols(hipStream_t s, int M, int N, const float* src, int lds, int col0,
              float* dst, int ldd);
// Adam on FP32 master weights, refreshing the BF16 copy.
void adam_update(hipStream_t s, int n, float* w, bf16* w16, const float* g,
                 float* m, float* v, float lr, float b1, float b2, float eps,
                 float bc1, float bc2);
void fill_f32(hipStream_t s, size_t n, float* p, float value);

}  // namespace tit_amd
