// Per-target trimming of the 5^d cell stencil of the stage kernels.
#pragma once

#include "internal.hpp"

namespace tit_amd {

// Per-target stencil trimming.  A particle of cell (ix, iy, z) lies within
// skin/2 of that cell's box (it is sorted into it at the last rebuild and
// has moved less than skin/2 since; only border cells hold particles outside
// the grid, clamped in).  So for a target whose whole 5^d window is interior
// (no border cell), a window cell can hold a neighbour only if its box grown
// by skin/2 lies within R of the target: gx^2 (+ gy^2) + gz^2 <= R^2 with the
// per-axis gaps from the target to the grown box.  Along a column those
// cells are contiguous (the distance to consecutive boxes falls, then
// rises).  Trimming removes only candidates the exact FP64 predicate
// rejects, so the accepted set and its order are unchanged; it therefore
// only has to be conservative, and runs in FP32 with widened bounds: the
// gaps (FP64, then rounded) shrunk by 1e-6 relative, R^2 grown by 1e-5
// relative.
//
// The row windows of the NCOL stencil columns are packed 6 bits per column,
// ten columns per 64-bit word (3-bit offsets of the first and last row from
// zc - 2; 7 = empty column), held in named words so that no array is indexed
// at run time.
template <int DIM, int NCOL>
struct Windows {
  uint64_t w0 = 0, w1 = 0, w2 = 0;
  static_assert(NCOL <= 30, "three words of ten columns");
  __device__ __forceinline__ void set(int k, uint64_t v) {
    const uint64_t x = v << (6 * (k % 10));
    if (k < 10) {
      w0 |= x;
    } else if (k < 20) {
      w1 |= x;
    } else {
      w2 |= x;
    }
  }
  __device__ __forceinline__ uint32_t get(int k) const {
    const uint64_t w = k < 10 ? w0 : (k < 20 ? w1 : w2);
    return static_cast<uint32_t>(w >> (6 * (k % 10))) & 63u;
  }
};

template <int DIM, int NCOL>
__device__ __forceinline__ Windows<DIM, NCOL> stencil_windows(const DevParams& P,
                                                              const DevTank& T,
                                                              const double* xa, int cx, int cy,
                                                              int zc) {
  constexpr int L = DIM - 1;
  Windows<DIM, NCOL> W;
  bool inner = cx - kStencil > 0 && cx + kStencil < T.nc[0] - 1 && zc - kStencil > 0 &&
               zc + kStencil < T.nc[L] - 1;
  if constexpr (DIM == 3) inner = inner && cy - kStencil > 0 && cy + kStencil < T.nc[1] - 1;
  constexpr uint64_t kFull = 0u | (4u << 3);  // rows zc - 2 .. zc + 2
  if (!inner) {
#pragma unroll
    for (int k = 0; k < NCOL; ++k) W.set(k, kFull);
    return W;
  }
  // gaps from the target to the grown boxes of the five stencil columns of
  // each column axis and of the five rows
  constexpr int NC = 2 * kStencil + 1;
  float gx[NC], gz2[NC];
  [[maybe_unused]] float gy[NC];
  double cw[DIM];
#pragma unroll
  for (int ax = 0; ax < DIM; ++ax) cw[ax] = 1.0 / T.inv_c[ax];
  auto gap = [&](int ax, int ic) {
    const double a0 = T.lo[ax] + ic * cw[ax] - P.skin_half;
    const double a1 = T.lo[ax] + (ic + 1) * cw[ax] + P.skin_half;
    return static_cast<float>(fmax(0.0, fmax(a0 - xa[ax], xa[ax] - a1))) * (1.0f - 1e-6f);
  };
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    gx[q] = gap(0, cx - kStencil + q);
    if constexpr (DIM == 3) gy[q] = gap(1, cy - kStencil + q);
    const float g = gap(L, zc - kStencil + q);
    gz2[q] = g * g;
  }
  const float R2 = static_cast<float>(P.R2) * (1.0f + 1e-5f);
#pragma unroll
  for (int k = 0; k < NCOL; ++k) {
    float g2;
    if constexpr (DIM == 2) {
      g2 = gx[k] * gx[k];
    } else {
      g2 = gx[k / NC] * gx[k / NC] + gy[k % NC] * gy[k % NC];
    }
    uint32_t rows = 0;  // bit q: row zc - 2 + q can hold a neighbour
#pragma unroll
    for (int q = 0; q < NC; ++q) rows |= (g2 + gz2[q] <= R2 ? 1u : 0u) << q;
    uint64_t v = 7u | (7u << 3);
    if (rows) {
      v = static_cast<uint64_t>(__builtin_ctz(rows)) |
          (static_cast<uint64_t>(31 - __builtin_clz(rows)) << 3);
    }
    W.set(k, v);
  }
  return W;
}

}  // namespace tit_amd
