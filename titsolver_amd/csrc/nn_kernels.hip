// BF16 MFMA GEMM and the element-wise / graph kernels of the learned step.
//
// GEMM tile: 64 x 64 outputs per 256-thread workgroup, four waves of 32 x 32
// each, one v_mfma_f32_32x32x16_bf16 per wave and K-step of 16.  Operand
// fragments (gfx950, 64 lanes): A[m = l % 32][k = 8 (l / 32) + i] and
// B[k = 8 (l / 32) + i][n = l % 32], i < 8; the accumulator holds
// C[m = 8 j + 4 (l / 32) + i][n = l % 32] in element 4 j + i.  A is staged
// row-major and B column-major in LDS so both fragments are 16-byte reads.
#include "nn_kernels.hpp"

namespace tit_amd {

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;

constexpr int kBM = 64, kBN = 64, kBK = 64;
constexpr int kLdsRow = kBK + 8;  // padded LDS row (144 B): conflict-free 16 B reads

// Eight consecutive BF16 of src starting at element `at`, `valid` of them
// in range (the rest zero): one 16-byte load when all are valid and aligned.
__device__ __forceinline__ uint4 load8(const bf16* src, size_t at, int valid) {
  if (valid >= 8 && (reinterpret_cast<uintptr_t>(src + at) & 15u) == 0) {
    return *reinterpret_cast<const uint4*>(src + at);
  }
  union {
    uint4 u;
    bf16 h[8];
  } r;
  r.u = make_uint4(0u, 0u, 0u, 0u);
  for (int t = 0; t < 8 && t < valid; ++t) r.h[t] = src[at + t];
  return r.u;
}

// Stages a 64 x 64 operand tile as tile[row][k] (row = m for A, n for B).
// `major_k`: the source rows run along k (A row-major, or B transposed);
// otherwise they run along m / n and the tile is transposed into LDS.
__device__ __forceinline__ void stage_tile(bf16 (*tile)[kLdsRow], const bf16* src, int ld,
                                           bool major_k, int r0, int rows, int k0, int kend,
                                           int tid) {
  for (int u = tid; u < kBM * kBK / 8; u += 256) {
    if (major_k) {
      const int row = u >> 3, kq = (u & 7) * 8;
      const int r = r0 + row, k = k0 + kq;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (r < rows && k < kend) v = load8(src, static_cast<size_t>(r) * ld + k, kend - k);
      *reinterpret_cast<uint4*>(&tile[row][kq]) = v;
    } else {
      const int krow = u >> 3, rq = (u & 7) * 8;
      const int k = k0 + krow, r = r0 + rq;
      union {
        uint4 u;
        bf16 h[8];
      } v;
      v.u = make_uint4(0u, 0u, 0u, 0u);
      if (k < kend && r < rows) v.u = load8(src, static_cast<size_t>(k) * ld + r, rows - r);
#pragma unroll
      for (int t = 0; t < 8; ++t) tile[rq + t][krow] = v.h[t];
    }
  }
}

__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 sA[kBM][kLdsRow];
  __shared__ __attribute__((aligned(16))) bf16 sB[kBN][kLdsRow];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.x * kBM, n0 = blockIdx.y * kBN;
  int kbeg = 0, kend = g.K;
  if (g.k_splits > 1) {
    const int chunk = ((g.K + g.k_splits - 1) / g.k_splits + kBK - 1) / kBK * kBK;
    kbeg = blockIdx.z * chunk;
    kend = min(g.K, kbeg + chunk);
  }
  f32x16 acc = {};
  for (int k0 = kbeg; k0 < kend; k0 += kBK) {
    stage_tile(sA, g.A, g.lda, !g.transA, m0, g.M, k0, kend, tid);
    stage_tile(sB, g.B, g.ldb, g.transB, n0, g.N, k0, kend, tid);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 16) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[wm * 32 + (lane & 31)][kk + 8 * (lane >> 5)]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sB[wn * 32 + (lane & 31)][kk + 8 * (lane >> 5)]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  const int n = n0 + wn * 32 + (lane & 31);
  if (n >= g.N) return;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + wm * 32 + 8 * j + 4 * (lane >> 5) + i;
      if (m >= g.M) continue;
      float x = acc[4 * j + i];
      if (g.accumulate) {
        atomicAdd(g.out32 + static_cast<size_t>(m) * g.ldo32 + n, x);
        continue;
      }
      if (g.bias) x += g.bias[n];
      if (g.G1) x += g.G1[static_cast<size_t>(g.r1[m]) * g.N + n];
      if (g.G2) x += g.G2[static_cast<size_t>(g.r2[m]) * g.N + n];
      if (g.pre) g.pre[static_cast<size_t>(m) * g.ldpre + n] = x;
      const float y = g.act == kActHswish ? nn_act(x) : x;
      if (g.out32) g.out32[static_cast<size_t>(m) * g.ldo32 + n] = y;
      if (g.out16) g.out16[static_cast<size_t>(m) * g.ldo16 + n] = static_cast<bf16>(y);
    }
  }
}

__global__ void k_act_backward(int M, int N, const float* dy, int ld_dy,
                                int group, const float* pre, float* dx32,
                                bf16* dx16) {
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<size_t>(M) * N) return;
  const size_t m = t / N, n = t % N;
  const float g = dy[(m / group) * ld_dy + n] * nn_act_grad(pre[t]);
  if (dx32) dx32[t] = g;
  if (dx16) dx16[t] = static_cast<bf16>(g);
}

__global__ void k_group_sum(int rows, int N, int group, const float* in,
                            float* out, bf16* out16, int ld16, int col16) {
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<size_t>(rows) * N) return;
  const size_t r = t / N, n = t % N;
  float s = 0.0f;
  for (int k = 0; k < group; ++k) s += in[(r * group + k) * N + n];
  if (out) out[t] = s;
  if (out16) out16[r * ld16 + col16 + n] = static_cast<bf16>(s);
}

__global__ void k_scatter_add_rows(int M, int N, const float* in,
                                   const uint32_t* idx, float* out) {
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<size_t>(M) * N) return;
  const size_t m = t / N, n = t % N;
  atomicAdd(out + static_cast<size_t>(idx[m]) * N + n, in[t]);
}

__global__ void k_column_sum(int M, int N, const float* in, float* col_sum) {
  // one block per 64 columns x 256 row slices
  const int n = blockIdx.x * 64 + (threadIdx.x & 63);
  const int slice = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int per = (M + 4 * gridDim.y - 1) / (4 * gridDim.y);
  if (n >= N) return;
  float s = 0.0f;
  for (int m = slice * per; m < min(M, (slice + 1) * per); ++m) s += in[static_cast<size_t>(m) * N + n];
  atomicAdd(col_sum + n, s);
}

__global__ void k_to_bf16_block(int M, int N, const float* src, int lds, bf16* dst,
                                int ldd, int col0) {
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<size_t>(M) * N) return;
  const size_t m = t / N, n = t % N;
  dst[m * ldd + col0 + n] = static_cast<bf16>(src[m * lds + n]);
}

__global__ void k_copy_cols(int M, int N, const float* src, int lds, int col0,
                            float* dst, int ldd, int add) {
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<size_t>(M) * N) return;
  const size_t m = t / N, n = t % N;
  const float v = src[m * lds + col0 + n];
  if (add) {
    dst[m * ldd + n] += v;
  } else {
    dst[m * ldd + n] = v;
  }
}

__global__ void k_adam(int n, float* w, bf16* w16, const float* g, float* m,
                       float* v, float lr, float b1, float b2, float eps,
                       float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  w[i] = wi;
  if (w16) w16[i] = static_cast<bf16>(wi);
}

__global__ void k_fill(size_t n, float* p, float value) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

inline unsigned blocks(size_t n, int b = 256) {
  return static_cast<unsigned>((n + b - 1) / b);
}

}  // namespace

void gemm_bf16(hipStream_t s, const GemmArgs& g) {
  if (g.M <= 0 || g.N <= 0) return;
  const dim3 grid((g.M + kBM - 1) / kBM, (g.N + kBN - 1) / kBN,
                  g.k_splits > 1 ? g.k_splits : 1);
  k_gemm<<<grid, 256, 0, s>>>(g);
}

void act_backward(hipStream_t s, int M, int N, const float* dy, int ld_dy,
                   int group, const float* pre, float* dx32, bf16* dx16) {
  const size_t n = static_cast<size_t>(M) * N;
  if (n) {
    k_act_backward<<<blocks(n), 256, 0, s>>>(M, N, dy, ld_dy, group, pre, dx32,
                                              dx16);
  }
}

void group_sum(hipStream_t s, int rows, int N, int group, const float* in,
               float* out, bf16* out16, int ld16, int col16) {
  const size_t n = static_cast<size_t>(rows) * N;
  if (n) k_group_sum<<<blocks(n), 256, 0, s>>>(rows, N, group, in, out, out16, ld16, col16);
}

void scatter_add_rows(hipStream_t s, int M, int N, const float* in,
                      const uint32_t* idx, float* out) {
  const size_t n = static_cast<size_t>(M) * N;
  if (n) k_scatter_add_rows<<<blocks(n), 256, 0, s>>>(M, N, in, idx, out);
}

void column_sum(hipStream_t s, int M, int N, const float* in, float* col_sum) {
  if (M <= 0 || N <= 0) return;
  const int slices = std::min(1024, std::max(1, M / 256));
  k_column_sum<<<dim3((N + 63) / 64, slices), 256, 0, s>>>(M, N, in, col_sum);
}

void to_bf16_block(hipStream_t s, int M, int N, const float* src, int lds,
                   bf16* dst, int ldd, int col0) {
  const size_t n = static_cast<size_t>(M) * N;
  if (n) k_to_bf16_block<<<blocks(n), 256, 0, s>>>(M, N, src, lds, dst, ldd, col0);
}

void copy_cols(hipStream_t s, int M, int N, const float* src, int lds, int col0,
               float* dst, int ldd) {
  const size_t n = static_cast<size_t>(M) * N;
  if (n) k_copy_cols<<<blocks(n), 256, 0, s>>>(M, N, src, lds, col0, dst, ldd, 0);
}

void add_cols(hipStream_t s, int M, int N, const float* src, int lds, int col0,
              float* dst, int ldd) {
  const size_t n = static_cast<size_t>(M) * N;
  if (n) k_copy_cols<<<blocks(n), 256, 0, s>>>(M, N, src, lds, col0, dst, ldd, 1);
}

void adam_update(hipStream_t s, int n, float* w, bf16* w16, const float* g,
                 float* m, float* v, float lr, float b1, float b2, float eps,
                 float bc1, float bc2) {
  if (n) k_adam<<<blocks(n), 256, 0, s>>>(n, w, w16, g, m, v, lr, b1, b2, eps, bc1, bc2);
}

void fill_f32(hipStream_t s, size_t n, float* p, float value) {
  if (n) k_fill<<<blocks(n), 256, 0, s>>>(n, p, value);
}

}  // namespace tit_amd
