// Fused inference of one message-passing round (network in surrogate.hpp).
//
// A persistent workgroup of 8 waves keeps the round's weights in LDS as
// ready-made MFMA operand fragments and walks tiles of 64 consecutive
// (cell-sorted) particles = 512 edges:
//
//   edge phase (each wave: 2 blocks of 32 edges)
//     layer 1, transposed:  C1[f][e] = W1^T X^T     A = W1^T fragment (L1/L2),
//                                                   B = x_e gathered from HBM
//     layer 2, direct:      C2[e][f] = h1 W2        A = h1 straight from C1's
//                                                   registers, B = W2 (LDS)
//     aggregation: a block's edge rows are assigned so that the 8 edges of a
//     particle land in one lane's accumulators -> in-register sums
//   node phase (wave w: 32 features x 32 particles), through LDS
//     y = [H | M] -> layer 3 -> layer 4 -> layer 5, all transposed
//
// MFMA v_mfma_f32_32x32x16_bf16 fragments (64 lanes, h = lane / 32):
// A[i = lane % 32][k = 8h + t], B[k = 8h + t][j = lane % 32] (t < 8) and
// C[8q + 4h + s][lane % 32] in accumulator element 4q + s.  Layer 1's output
// rows f = 32 ob + 8q + 4h + s are exactly the K-slots of layer 2's A operand
// when layer 2's K-step (ob, jj) takes accumulator elements 8 jj .. 8 jj + 7;
// W2's B fragments are packed in that K order.
#include "surrogate.hpp"

namespace tit_amd {

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;

constexpr int kFrag = 512;            // BF16 elements of one fragment
constexpr int kNF1 = 4 * 5;           // W1^T: 4 output blocks x 5 K-steps
constexpr int kNF2 = 4 * 8;           // W2:   4 output blocks x 8 K-steps
constexpr int kNF3 = 4 * 10;          // W3^T
constexpr int kNF4 = 4 * 8;           // W4^T
constexpr int kNF5 = 8;               // W5^T (16 outputs padded to 32)
constexpr int kOffF1 = 0;
constexpr int kOffF2 = kOffF1 + kNF1 * kFrag;
constexpr int kOffF3 = kOffF2 + kNF2 * kFrag;
constexpr int kOffF4 = kOffF3 + kNF3 * kFrag;
constexpr int kOffF5 = kOffF4 + kNF4 * kFrag;
constexpr int kRoundFrag = kOffF5 + kNF5 * kFrag;  // per round
constexpr int kLdsFrag = kRoundFrag - kOffF2;      // W2..W5 staged in LDS

constexpr int kTileP = 64;
constexpr int kThreads = 512;
constexpr int kYld = 168;  // padded LDS row of y = [H | M]
constexpr int kHld = 136;  // padded LDS row of hidden node activations


// ---- fragment packing -------------------------------------------------------------

// Writes the fragments of both rounds from the BF16 parameter vector.
__global__ void k_pack(const bf16* __restrict__ w16, bf16* __restrict__ frag) {
  const int f = blockIdx.x;  // fragment index within a round
  const int r = blockIdx.y;  // round
  const int l = threadIdx.x;  // lane 0..63
  const bf16* W = w16 + static_cast<size_t>(r) * RoundLayout::size;
  bf16* out = frag + static_cast<size_t>(r) * kRoundFrag;
  const int h = l >> 5, col = l & 31;
  bf16 v[8];
  const bf16 zero = static_cast<bf16>(0.0f);
  int off;
  if (f < kNF1) {  // A = W1^T: row = out feature, k = input feature
    const int ob = f / 5, ks = f % 5;
    off = kOffF1 + f * kFrag;
    for (int t = 0; t < 8; ++t) {
      const int k = 16 * ks + 8 * h + t;
      v[t] = W[RoundLayout::W1 + k * kHidden + 32 * ob + col];
    }
  } else if (f < kNF1 + kNF2) {  // B = W2 in the permuted K order
    const int g = f - kNF1, ob = g / 8, ks = g % 8;
    const int ob_in = ks >> 1, jj = ks & 1;
    off = kOffF2 + g * kFrag;
    for (int t = 0; t < 8; ++t) {
      const int k = 32 * ob_in + 8 * (2 * jj + (t >> 2)) + 4 * h + (t & 3);
      v[t] = W[RoundLayout::W2 + k * kHidden + 32 * ob + col];
    }
  } else if (f < kNF1 + kNF2 + kNF3) {  // A = W3^T
    const int g = f - kNF1 - kNF2, ob = g / 10, ks = g % 10;
    off = kOffF3 + g * kFrag;
    for (int t = 0; t < 8; ++t) {
      const int k = 16 * ks + 8 * h + t;
      v[t] = W[RoundLayout::W3 + k * kHidden + 32 * ob + col];
    }
  } else if (f < kNF1 + kNF2 + kNF3 + kNF4) {  // A = W4^T
    const int g = f - kNF1 - kNF2 - kNF3, ob = g / 8, ks = g % 8;
    off = kOffF4 + g * kFrag;
    for (int t = 0; t < 8; ++t) {
      const int k = 16 * ks + 8 * h + t;
      v[t] = W[RoundLayout::W4 + k * kHidden + 32 * ob + col];
    }
  } else {  // A = W5^T, rows >= 16 zero
    const int ks = f - kNF1 - kNF2 - kNF3 - kNF4;
    off = kOffF5 + ks * kFrag;
    for (int t = 0; t < 8; ++t) {
      const int k = 16 * ks + 8 * h + t;
      v[t] = col < kOut ? W[RoundLayout::W5 + k * kOut + col] : zero;
    }
  }
  for (int t = 0; t < 8; ++t) out[off + 8 * l + t] = v[t];
}

// ---- fused round --------------------------------------------------------------------

struct FusedArgs {
  uint32_t n;
  uint32_t n_tiles;
  const bf16* frag;      // this round's fragments
  const float* w32;      // this round's FP32 parameters (biases)
  const uint32_t* nbr;
  const bf16* ef;
  const bf16* H;         // n x kNodeIn
  bf16* Hn;              // next round's node input (null in the last round)
  float* u;              // n x kOut (null except in the last round)
};

__device__ __forceinline__ bf16x8 ld8(const bf16* p) {
  return *reinterpret_cast<const bf16x8*>(p);
}

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Accumulators that start from the bias: of the output rows
// f = 32 ob + 8q + 4h + s (transposed layers), or of column f (direct layers).
__device__ __forceinline__ f32x16 bias_rows(const float* b, int ob, int h) {
  f32x16 c;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int s = 0; s < 4; ++s) c[4 * q + s] = b[32 * ob + 8 * q + 4 * h + s];
  }
  return c;
}
__device__ __forceinline__ f32x16 bias_all(float b) {
  f32x16 c;
#pragma unroll
  for (int k = 0; k < 16; ++k) c[k] = b;
  return c;
}

// Tiles in XCD-major bands: workgroup b runs on XCD b % 8 and takes tiles of
// that XCD's eighth of the particle range, so each L2 sees one region.
__global__ __launch_bounds__(kThreads) void k_round_fused(FusedArgs A) {
  __shared__ __attribute__((aligned(16))) bf16 sF[kLdsFrag];
  __shared__ float sBias[4 * kHidden + kOut];
  __shared__ __attribute__((aligned(16))) bf16 sY[kTileP * kYld];
  __shared__ __attribute__((aligned(16))) bf16 sH3[kTileP * kHld];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int h = lane >> 5, col = lane & 31;

  // stage W2..W5 fragments and the biases
  {
    const uint4* src = reinterpret_cast<const uint4*>(A.frag + kOffF2);
    uint4* dst = reinterpret_cast<uint4*>(sF);
    for (int k = tid; k < kLdsFrag / 8; k += kThreads) dst[k] = src[k];
    for (int k = tid; k < 4 * kHidden + kOut; k += kThreads) {
      float b;
      if (k < kHidden) {
        b = A.w32[RoundLayout::b1 + k];
      } else if (k < 2 * kHidden) {
        b = A.w32[RoundLayout::b2 + k - kHidden];
      } else if (k < 3 * kHidden) {
        b = A.w32[RoundLayout::b3 + k - 2 * kHidden];
      } else if (k < 4 * kHidden) {
        b = A.w32[RoundLayout::b4 + k - 3 * kHidden];
      } else {
        b = A.w32[RoundLayout::b5 + k - 4 * kHidden];
      }
      sBias[k] = b;
    }
  }
  __syncthreads();
  const float* b1 = sBias;
  const float* b2 = sBias + kHidden;
  const float* b3 = sBias + 2 * kHidden;
  const float* b4 = sBias + 3 * kHidden;
  const float* b5 = sBias + 4 * kHidden;

  // XCD band of tiles
  const uint32_t xcd = blockIdx.x & 7u, per_xcd = (gridDim.x + 7u - xcd) >> 3;
  const uint32_t k_in = blockIdx.x >> 3;
  const uint32_t band = (A.n_tiles + 7u) >> 3;
  const uint32_t t_beg = min(A.n_tiles, xcd * band), t_end = min(A.n_tiles, t_beg + band);

  // edge slot of this lane within a 32-edge block: particle (of 4), edge (of 8)
  const int slot = col;
  const int p_slot = 2 * ((slot >> 2) & 1) + (slot >> 4);
  const int k_slot = 4 * ((slot >> 3) & 1) + (slot & 3);

  // edge inputs x_e of this wave's 2 blocks of 32 edges of a tile
  auto gather_x = [&](uint32_t tile, bf16x8 (&X)[2][5]) {
    const uint32_t p0 = tile * kTileP;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int blk = 2 * wave + c;                      // 0..15
      uint32_t p = p0 + 4 * blk + p_slot;
      p = min(p, A.n - 1);                               // tail: valid memory
      const size_t e = static_cast<size_t>(p) * kKnn + k_slot;
      const uint32_t b = A.nbr[e];
      X[c][0] = ld8(A.ef + e * kEdgeFeat + 8 * h);
      X[c][1] = ld8(A.H + static_cast<size_t>(p) * kNodeIn + 8 * h);
      X[c][2] = ld8(A.H + static_cast<size_t>(p) * kNodeIn + 16 + 8 * h);
      X[c][3] = ld8(A.H + static_cast<size_t>(b) * kNodeIn + 8 * h);
      X[c][4] = ld8(A.H + static_cast<size_t>(b) * kNodeIn + 16 + 8 * h);
    }
  };
  bf16x8 X[2][5];
  if (t_beg + k_in < t_end) gather_x(t_beg + k_in, X);
  for (uint32_t tile = t_beg + k_in; tile < t_end; tile += per_xcd) {
    const uint32_t p0 = tile * kTileP;
    // ---------------- edge phase: wave handles edge blocks 2w, 2w + 1 --------------
    // layer 1 (transposed) -> layer-2 A operands, BF16
    bf16x8 H1[2][8];  // [block][K-step of layer 2 = 2 ob + jj]
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
      const f32x16 b0 = bias_rows(b1, ob, h);
      f32x16 acc[2] = {b0, b0};
#pragma unroll
      for (int ks = 0; ks < 5; ++ks) {
        const bf16x8 a = ld8(A.frag + kOffF1 + (ob * 5 + ks) * kFrag + 8 * lane);
        acc[0] = mfma(a, X[0][ks], acc[0]);
        acc[1] = mfma(a, X[1][ks], acc[1]);
      }
#pragma unroll
      for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          bf16x8 v;
#pragma unroll
          for (int t = 0; t < 8; ++t) v[t] = static_cast<bf16>(nn_act(acc[c][8 * jj + t]));
          H1[c][2 * ob + jj] = v;
        }
      }
    }
    // layer 2 (direct) + aggregation -> M into LDS
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
      const int f = 32 * ob + col;
      const f32x16 b0 = bias_all(b2[f]);
      f32x16 acc[2] = {b0, b0};
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
        const bf16x8 bw = ld8(sF + (kOffF2 - kOffF2) + (ob * 8 + ks) * kFrag + 8 * lane);
        acc[0] = mfma(H1[0][ks], bw, acc[0]);
        acc[1] = mfma(H1[1][ks], bw, acc[1]);
      }
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int blk = 2 * wave + c;
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          float s = 0.0f;
#pragma unroll
          for (int q = 2 * jj; q < 2 * jj + 2; ++q) {
#pragma unroll
            for (int e = 0; e < 4; ++e) s += nn_act(acc[c][4 * q + e]);
          }
          const int pl = 4 * blk + 2 * h + jj;  // tile-local particle
          sY[pl * kYld + kNodeIn + f] = static_cast<bf16>(s);
        }
      }
    }
    // the next tile's edge inputs load while this tile finishes
    if (tile + per_xcd < t_end) gather_x(tile + per_xcd, X);
    // node inputs H into y[:, 0:32]
    if (tid < kTileP * (kNodeIn / 8)) {
      const int pl = tid >> 2, g = tid & 3;
      const uint32_t p = min(p0 + pl, A.n - 1);
      *reinterpret_cast<uint4*>(sY + pl * kYld + 8 * g) =
          *reinterpret_cast<const uint4*>(A.H + static_cast<size_t>(p) * kNodeIn + 8 * g);
    }
    __syncthreads();
    // ---------------- node phase: wave -> (output block, particle block) ----------
    const int ob = wave & 3, pb = wave >> 2;
    const int prow = 32 * pb + col;  // particle of this lane's B column
    {  // layer 3: y (160) -> 128
      f32x16 acc = bias_rows(b3, ob, h);
#pragma unroll
      for (int ks = 0; ks < 10; ++ks) {
        const bf16x8 a = ld8(sF + (kOffF3 - kOffF2) + (ob * 10 + ks) * kFrag + 8 * lane);
        const bf16x8 b = ld8(sY + prow * kYld + 16 * ks + 8 * h);
        acc = mfma(a, b, acc);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int f = 32 * ob + 8 * q + 4 * h + s;
          sH3[prow * kHld + f] = static_cast<bf16>(nn_act(acc[4 * q + s]));
        }
      }
    }
    __syncthreads();
    {  // layer 4: 128 -> 128, output into the y rows (y is no longer read)
      f32x16 acc = bias_rows(b4, ob, h);
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
        const bf16x8 a = ld8(sF + (kOffF4 - kOffF2) + (ob * 8 + ks) * kFrag + 8 * lane);
        const bf16x8 b = ld8(sH3 + prow * kHld + 16 * ks + 8 * h);
        acc = mfma(a, b, acc);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int f = 32 * ob + 8 * q + 4 * h + s;
          sY[prow * kHld + f] = static_cast<bf16>(nn_act(acc[4 * q + s]));
        }
      }
    }
    __syncthreads();
    if (wave < 2) {  // layer 5: 128 -> 16, waves 0 and 1 (particle blocks)
      const int pr = 32 * wave + col;
      f32x16 acc = {};
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
        const bf16x8 a = ld8(sF + (kOffF5 - kOffF2) + ks * kFrag + 8 * lane);
        const bf16x8 b = ld8(sY + pr * kHld + 16 * ks + 8 * h);
        acc = mfma(a, b, acc);
      }
      const uint32_t p = p0 + pr;
      if (p < A.n) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int o = 8 * q + 4 * h + s;
            const float val = acc[4 * q + s] + b5[o];
            if (A.u) A.u[static_cast<size_t>(p) * kOut + o] = val;
            if (A.Hn) A.Hn[static_cast<size_t>(p) * kNodeIn + kNodeFeat + o] = static_cast<bf16>(val);
          }
        }
        if (A.Hn && h == 0) {  // node features carry over
          const uint4* src = reinterpret_cast<const uint4*>(A.H + static_cast<size_t>(p) * kNodeIn);
          uint4* dst = reinterpret_cast<uint4*>(A.Hn + static_cast<size_t>(p) * kNodeIn);
          dst[0] = src[0];
          dst[1] = src[1];
        }
      }
    }
    __syncthreads();  // y / h3 rows are rewritten by the next tile
  }
}

}  // namespace

size_t fused_fragment_elems() { return static_cast<size_t>(kRounds) * kRoundFrag; }

void pack_fragments(hipStream_t s, const bf16* w16, bf16* frag) {
  k_pack<<<dim3(kNF1 + kNF2 + kNF3 + kNF4 + kNF5, kRounds), 64, 0, s>>>(w16, frag);
}

void launch_round_fused(hipStream_t s, int n_cu, uint32_t n, int round, const bf16* frag,
                        const float* w32, const uint32_t* nbr, const bf16* ef,
                        const bf16* H, bf16* Hn, float* u) {
  if (n == 0) return;
  FusedArgs a{};
  a.n = n;
  a.n_tiles = (n + kTileP - 1) / kTileP;
  a.frag = frag + static_cast<size_t>(round) * kRoundFrag;
  a.w32 = w32 + static_cast<size_t>(round) * RoundLayout::size;
  a.nbr = nbr;
  a.ef = ef;
  a.H = H;
  a.Hn = Hn;
  a.u = u;
  // one workgroup per CU (LDS-bound), a multiple of the 8 XCDs
  const uint32_t grid = std::max<uint32_t>(8u, static_cast<uint32_t>(n_cu) / 8u * 8u);
  k_round_fused<<<grid, kThreads, 0, s>>>(a);
}

}  // namespace tit_amd
