// Per-target trimming of the 5^d cell stencil, shared by the stage kernels
// and the boundary extrapolation.
#pragma once

#include "internal.hpp"

namespace tit_amd {

// Per-target stencil trimming.  A particle of cell (ix, iy, z) lies within
// skin/2 of that cell's box (it is sorted into it at the last rebuild and
// has moved less than skin/2 since; only border cells hold particles outside
// the grid, clamped in).  So for a target whose whole 5^d window is interior
// (no border cell), a column whose grown box is farther than R from the
// target in the column plane holds no neighbour, and only the rows within
// sqrt(R^2 - gap^2) + skin/2 of the target can.  Trimming removes only
// candidates the exact FP64 predicate rejects, so the accepted set and its
// order are unchanged; it therefore only has to be conservative, and runs in
// FP32 with widened bounds: the gaps (FP64, then rounded) shrunk by 1e-6
// relative, R^2 grown by 1e-5 relative, the row bounds widened by 1e-4 cell.
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
  // each column axis
  constexpr int NC = 2 * kStencil + 1;
  float gx[NC];
  [[maybe_unused]] float gy[NC];
  auto gap = [&](int ax, int ic) {
    const double c = 1.0 / T.inv_c[ax];
    const double a0 = T.lo[ax] + ic * c - P.skin_half;
    const double a1 = T.lo[ax] + (ic + 1) * c + P.skin_half;
    return static_cast<float>(fmax(0.0, fmax(a0 - xa[ax], xa[ax] - a1))) * (1.0f - 1e-6f);
  };
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    gx[q] = gap(0, cx - kStencil + q);
    if constexpr (DIM == 3) gy[q] = gap(1, cy - kStencil + q);
  }
  const float R2 = static_cast<float>(P.R2) * (1.0f + 1e-5f);
  const float inv_cz = static_cast<float>(T.inv_c[L]);
  const float zr = static_cast<float>((xa[L] - T.lo[L]) * T.inv_c[L]);  // cell units
  const float sz = static_cast<float>(P.skin_half * T.inv_c[L]) * (1.0f + 1e-5f) + 1e-4f;
#pragma unroll
  for (int k = 0; k < NCOL; ++k) {
    float g2;
    if constexpr (DIM == 2) {
      g2 = gx[k] * gx[k];
    } else {
      g2 = gx[k / NC] * gx[k / NC] + gy[k % NC] * gy[k % NC];
    }
    uint64_t v = 7u | (7u << 3);
    if (!(g2 > R2)) {
      const float half = __builtin_sqrtf(R2 - g2) * inv_cz * (1.0f + 1e-5f) + sz;
      const int w0 = max(zc - kStencil, static_cast<int>(floorf(zr - half)));
      const int w1 = min(zc + kStencil, static_cast<int>(floorf(zr + half)));
      if (w0 <= w1) {
        v = static_cast<uint64_t>(w0 - (zc - kStencil)) |
            (static_cast<uint64_t>(w1 - (zc - kStencil)) << 3);
      }
    }
    W.set(k, v);
  }
  return W;
}

}  // namespace tit_amd
