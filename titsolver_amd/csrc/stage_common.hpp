// Device code shared by the stage kernels (stage_kernel.hip: 2D and the
// column-chunk global path; stage3d.hip: 3D block chunks): candidate
// accessors, the per-target stencil walk with acceptance masks, the pair
// loops, the per-target SSPRK update and the per-wave reductions.  Every
// path visits a target's candidates in the same order and rounds
// identically, so all schedules produce the same bits.
#pragma once

#include <algorithm>
#include <type_traits>

#include "pair_math.hpp"
#include "sph_kernels.hpp"
#include "stencil.hpp"

namespace tit_amd {

namespace {

constexpr int kWords = 5;      // acceptance masks held per lane (LDS path)
constexpr int kWordsG = 24;    // 32-candidate masks held per lane (global path)
constexpr uint32_t kWordSpan = 1u << 22;  // candidate global-index span of an LDS chunk
static_assert(kStageLdsCap <= 1024, "LDS index packs into 10 bits");

// Positions from LDS (2D).
struct LdsPos {
  const double2* xy;
  const uint32_t (*coff)[kRowsMax + 1];  // [column][row - zlo]
  const uint32_t* gb;                    // global index of each column's row zlo
  const uint32_t* lb;                    // LDS index of each column's row zlo
  int zlo, nz, ix, ncx;
  uint32_t gbase;   // global index of the chunk's first candidate

  // Candidates of column k in rows [w0, w1] (inside the target's 5-row window).
  __device__ __forceinline__ bool range(int k, int w0, int w1, uint32_t& b,
                                        uint32_t& e) const {
    const int ixk = ix - kStencil + k;
    if (ixk < 0 || ixk >= ncx) return false;
    const int z0 = max(w0, 0), z1 = min(w1, nz - 1);
    if (z0 > z1) return false;
    b = coff[k][z0 - zlo];
    e = coff[k][z1 + 1 - zlo];
    return true;
  }
};

// Positions from global memory (any dimension).
template <int DIM>
struct GlobalPos {
  const DevTank& T;
  const DevState& S;
  const uint32_t* coff;
  int cx, cy;    // target column
  uint32_t* wm;  // this wave's [kWordsG][64] 32-candidate acceptance masks
  uint32_t* wj;  // global index of each mask's bit 0

  __device__ __forceinline__ bool range(int k, int w0, int w1, uint32_t& b,
                                        uint32_t& e) const {
    constexpr int L = DIM - 1;
    int ix, iy = 0;
    if constexpr (DIM == 2) {
      ix = cx - kStencil + k;
    } else {
      ix = cx - kStencil + k / kCols2;
      iy = cy - kStencil + k % kCols2;
      if (iy < 0 || iy >= T.nc[1]) return false;
    }
    if (ix < 0 || ix >= T.nc[0]) return false;
    const int nz = T.nc[L];
    const int z0 = max(w0, 0), z1 = min(w1, nz - 1);
    if (z0 > z1) return false;
    uint32_t row;
    if constexpr (DIM == 2) {
      row = T.cell_off + static_cast<uint32_t>(ix) * nz;
    } else {
      row = T.cell_off +
            (static_cast<uint32_t>(ix) * T.nc[1] + static_cast<uint32_t>(iy)) * nz;
    }
    b = coff[row + z0];
    e = coff[row + z1 + 1];
    return true;
  }
};

template <int DIM>
struct PairIn {
  double x[DIM];
  float4 pv;
  float2 aux;
  float vol;  // 3D with per-particle masses: m / rho
};

template <int DIM, int KERNEL, int FLAVOUR>
__device__ __forceinline__ void term_of(const DevParams& P, const Target<DIM>& a,
                                        const PairIn<DIM>& p, float& c, float* f) {
  double d[DIM];
#pragma unroll
  for (int q = 0; q < DIM; ++q) d[q] = p.x[q] - a.x[q];
  pair_term<DIM, KERNEL, FLAVOUR>(P, a, d, p.pv, p.aux.x, p.aux.y, p.vol, c, f);
}

template <int DIM, int KERNEL, int FLAVOUR>
__device__ __forceinline__ void sum_pair(const DevParams& P, const Target<DIM>& a,
                                         const PairIn<DIM>& p, Sums<DIM>& acc) {
  float c, f[DIM];
  term_of<DIM, KERNEL, FLAVOUR>(P, a, p, c, f);
  acc.add(c, f);
}

template <int DIM>
__device__ __forceinline__ bool within(const DevParams& P, const double* xb,
                                       const double* xa) {
  const double d0 = xb[0] - xa[0];
  double s = d0 * d0;
#pragma unroll
  for (int q = 1; q < DIM; ++q) {
    const double dq = xb[q] - xa[q];
    s = s + dq * dq;
  }
  return s <= P.R2;
}

__device__ __forceinline__ uint32_t pop_lowest(uint64_t& mask) {
  const uint32_t t = static_cast<uint32_t>(__builtin_ctzll(mask));
  mask &= mask - 1;
  return t;
}

// Candidate order, the same on every path so that all schedules produce the
// same bits: stencil columns in order, within a column ascending index; the
// accepted candidates' terms enter the compensated sums two at a time.
//
// 2D with LDS positions.  The acceptance masks of up to kWords candidate
// batches (normally one per stencil column) are queued per lane in registers
// and their accepted candidates are summed in one loop, so a wave's pair loop
// runs for the largest accepted count among its lanes, not for the sum over
// columns of each column's largest count.
struct WordQueue {
  uint64_t m[kWords];   // acceptance masks, oldest first (never empty)
  uint32_t jl[kWords];  // their bit 0: global index - gbase (bits 10..31), LDS index

  // Appends at slot n (a run-time index: one select per slot).
  __device__ __forceinline__ void put(int n, uint64_t mask, uint32_t v) {
#pragma unroll
    for (int q = 0; q < kWords; ++q) {
      if (q == n) {
        m[q] = mask;
        jl[q] = v;
      }
    }
  }
  // Removes the oldest word.
  __device__ __forceinline__ void pop(uint64_t& mask, uint32_t& v) {
    mask = m[0];
    v = jl[0];
#pragma unroll
    for (int q = 0; q + 1 < kWords; ++q) {
      m[q] = m[q + 1];
      jl[q] = jl[q + 1];
    }
  }
};

template <int KERNEL, int FLAVOUR>
__device__ __forceinline__ void sum_words(const DevParams& P, const LdsPos& ps,
                                          const DevState& S, const Target<2>& a,
                                          uint32_t n_acc, Sums<2>& acc, WordQueue& wq) {
  uint64_t m = 0;
  uint32_t l0 = 0, j0 = 0;
  // The next of the n_acc accepted candidates.
  auto take = [&](PairIn<2>& p) {
    if (m == 0) {
      uint32_t jl;
      wq.pop(m, jl);
      j0 = ps.gbase + (jl >> 10);
      l0 = jl & 1023u;
    }
    const uint32_t t = pop_lowest(m);
    const double2 x = ps.xy[l0 + t];
    p.x[0] = x.x;
    p.x[1] = x.y;
    p.pv = S.pv[j0 + t];
    p.aux = S.aux[j0 + t];
    p.vol = 0.0f;
  };
  uint32_t rem = n_acc;
  PairIn<2> A, B;
  if (rem == 0) return;
  if (acc.pend) {  // an odd term of an earlier batch of words waits for its partner
    take(A);
    sum_pair<2, KERNEL, FLAVOUR>(P, a, A, acc);
    if (--rem == 0) return;
  }
  // two register sets alternate (the next candidate's loads in flight while
  // the current one is summed); their terms enter the compensated sums
  // together
  take(A);
  --rem;
  while (true) {
    const bool hb = rem > 0;
    if (hb) {
      take(B);
      --rem;
    }
    float ca, fa[2];
    term_of<2, KERNEL, FLAVOUR>(P, a, A, ca, fa);
    if (!hb) {
      acc.add(ca, fa);
      break;
    }
    const bool ha = rem > 0;
    if (ha) {
      take(A);
      --rem;
    }
    float cb, fb[2];
    term_of<2, KERNEL, FLAVOUR>(P, a, B, cb, fb);
    acc.add2(ca, fa, cb, fb);
    if (!ha) break;
  }
}

template <int KERNEL, int FLAVOUR, int NCOL>
__device__ __forceinline__ void gather_words(const DevParams& P, const LdsPos& ps,
                                             const DevState& S, const Target<2>& a,
                                             uint32_t i, int zc, const Windows<2, NCOL>& W,
                                             Sums<2>& acc, uint32_t& count) {
  WordQueue wq;
  int k = 0, kc = 0;
  uint32_t c0 = 0, e = 0;
  bool more = true;
  while (more) {
    int nw = 0;
    uint32_t n_acc = 0;
    while (nw < kWords) {
      if (c0 >= e) {  // next non-empty column range
        bool found = false;
        for (; k < NCOL && !found; ++k) {
          const uint32_t v = W.get(k);
          if ((v & 7u) == 7u) continue;
          const int w0 = zc - kStencil + static_cast<int>(v & 7u);
          const int w1 = zc - kStencil + static_cast<int>(v >> 3);
          uint32_t b, eb;
          if (ps.range(k, w0, w1, b, eb) && b < eb) {
            found = true;
            kc = k;
            c0 = b;
            e = eb;
          }
        }
        if (!found) {
          more = false;
          break;
        }
      }
      const uint32_t len = min(e - c0, 64u);
      const uint32_t lc = ps.lb[kc] + (c0 - ps.gb[kc]);
      uint64_t mask = 0;
      uint32_t t = 0;
      for (; t + 4 <= len; t += 4) {
        const double2 p0 = ps.xy[lc + t], p1 = ps.xy[lc + t + 1];
        const double2 p2 = ps.xy[lc + t + 2], p3 = ps.xy[lc + t + 3];
        const double x0[2] = {p0.x, p0.y}, x1[2] = {p1.x, p1.y};
        const double x2[2] = {p2.x, p2.y}, x3[2] = {p3.x, p3.y};
        const uint64_t mm = (within<2>(P, x0, a.x) ? 1ull : 0ull) |
                            (within<2>(P, x1, a.x) ? 2ull : 0ull) |
                            (within<2>(P, x2, a.x) ? 4ull : 0ull) |
                            (within<2>(P, x3, a.x) ? 8ull : 0ull);
        mask |= mm << t;
      }
      for (; t < len; ++t) {
        const double2 p0 = ps.xy[lc + t];
        const double x0[2] = {p0.x, p0.y};
        if (within<2>(P, x0, a.x)) mask |= 1ull << t;
      }
      if (i >= c0 && i < c0 + len) mask &= ~(1ull << (i - c0));  // not self
      const uint32_t na = static_cast<uint32_t>(__builtin_popcountll(mask));
      count += na;
      n_acc += na;
      if (mask) {
        wq.put(nw, mask, ((c0 - ps.gbase) << 10) | lc);
        ++nw;
      }
      c0 += len;
    }
    sum_words<KERNEL, FLAVOUR>(P, ps, S, a, n_acc, acc, wq);
  }
}

// Global-memory path, the same scheme with 32-candidate masks (the 3D stencil
// has 25 short column ranges): all columns' accepted candidates are summed
// in one loop per lane, in the common candidate order.
template <int DIM, int KERNEL, int FLAVOUR>
__device__ __forceinline__ void sum_words_g(const DevParams& P, const GlobalPos<DIM>& ps,
                                            const DevState& S, const Target<DIM>& a,
                                            uint32_t n_acc, Sums<DIM>& acc) {
  const int lane = threadIdx.x & 63;
  int w = 0;
  uint32_t m = 0, m_next = ps.wm[lane], j0 = 0, j_next = ps.wj[lane];
  // the next of the n_acc accepted candidates (stored masks are never empty)
  auto take = [&](PairIn<DIM>& p) {
    if (m == 0) {
      m = m_next;
      j0 = j_next;
      w = min(w + 1, kWordsG - 1);
      m_next = ps.wm[w * 64 + lane];
      j_next = ps.wj[w * 64 + lane];
    }
    const uint32_t j = j0 + static_cast<uint32_t>(__builtin_ctz(m));
    m &= m - 1u;
#pragma unroll
    for (int k = 0; k < DIM; ++k) p.x[k] = S.x[k][j];
    p.pv = S.pv[j];
    p.aux = S.aux[j];
    p.vol = (DIM == 3 && mass_flavour(FLAVOUR)) ? S.vol[j] : 0.0f;
  };
  uint32_t rem = n_acc;
  PairIn<DIM> A, B;
  if (rem == 0) return;
  if (acc.pend) {  // an odd term of an earlier batch of masks waits for its partner
    take(A);
    sum_pair<DIM, KERNEL, FLAVOUR>(P, a, A, acc);
    if (--rem == 0) return;
  }
  take(A);
  --rem;
  while (true) {
    const bool hb = rem > 0;
    if (hb) {
      take(B);
      --rem;
    }
    float ca, fa[DIM];
    term_of<DIM, KERNEL, FLAVOUR>(P, a, A, ca, fa);
    if (!hb) {
      acc.add(ca, fa);
      break;
    }
    const bool ha = rem > 0;
    if (ha) {
      take(A);
      --rem;
    }
    float cb, fb[DIM];
    term_of<DIM, KERNEL, FLAVOUR>(P, a, B, cb, fb);
    acc.add2(ca, fa, cb, fb);
    if (!ha) break;
  }
}

template <int DIM, int KERNEL, int FLAVOUR, int NCOL>
__device__ __forceinline__ void gather_words_g(const DevParams& P, const GlobalPos<DIM>& ps,
                                               const DevState& S, const Target<DIM>& a,
                                               uint32_t i, int zc, const Windows<DIM, NCOL>& W,
                                               Sums<DIM>& acc, uint32_t& count) {
  const int lane = threadIdx.x & 63;
  int k = 0;
  uint32_t c0 = 0, e = 0;
  bool more = true;
  while (more) {
    int nw = 0;
    uint32_t n_acc = 0;
    while (nw < kWordsG) {
      if (c0 >= e) {  // next non-empty column range
        bool found = false;
        for (; k < NCOL && !found; ++k) {
          const uint32_t v = W.get(k);
          if ((v & 7u) == 7u) continue;
          const int w0 = zc - kStencil + static_cast<int>(v & 7u);
          const int w1 = zc - kStencil + static_cast<int>(v >> 3);
          uint32_t b, eb;
          if (ps.range(k, w0, w1, b, eb) && b < eb) {
            found = true;
            c0 = b;
            e = eb;
          }
        }
        if (!found) {
          more = false;
          break;
        }
      }
      const uint32_t len = min(e - c0, 32u);
      uint32_t mask = 0;
      uint32_t t = 0;
      for (; t + 4 <= len; t += 4) {
        double x0[DIM], x1[DIM], x2[DIM], x3[DIM];
#pragma unroll
        for (int q = 0; q < DIM; ++q) {
          x0[q] = S.x[q][c0 + t];
          x1[q] = S.x[q][c0 + t + 1];
          x2[q] = S.x[q][c0 + t + 2];
          x3[q] = S.x[q][c0 + t + 3];
        }
        const uint32_t mm = (within<DIM>(P, x0, a.x) ? 1u : 0u) |
                            (within<DIM>(P, x1, a.x) ? 2u : 0u) |
                            (within<DIM>(P, x2, a.x) ? 4u : 0u) |
                            (within<DIM>(P, x3, a.x) ? 8u : 0u);
        mask |= mm << t;
      }
      for (; t < len; ++t) {
        double x0[DIM];
#pragma unroll
        for (int q = 0; q < DIM; ++q) x0[q] = S.x[q][c0 + t];
        if (within<DIM>(P, x0, a.x)) mask |= 1u << t;
      }
      if (i >= c0 && i < c0 + len) mask &= ~(1u << (i - c0));  // not self
      const uint32_t na = static_cast<uint32_t>(__builtin_popcount(mask));
      count += na;
      n_acc += na;
      if (mask) {
        ps.wm[nw * 64 + lane] = mask;
        ps.wj[nw * 64 + lane] = c0;
        ++nw;
      }
      c0 += len;
    }
    sum_words_g<DIM, KERNEL, FLAVOUR>(P, ps, S, a, n_acc, acc);
  }
}

// Positions from LDS, 3D block chunks.  A chunk's targets are the particles
// of a block of kBlk3 x kBlk3 cell columns in a range of rows; its window is
// the block grown by the stencil, kWin3 x kWin3 columns over rows
// [zlo, zlo + nrows).  Each window column's candidates sit contiguously in
// LDS from lb[kw] on (FP64 x, y, z in three arrays), and delta[kw] turns an
// LDS index of column kw into the particle's global index.
struct LdsPos3 {
  const double* sx;
  const double* sy;
  const double* sz;
  const uint32_t (*coff)[kRows3 + 1];  // [window column][row - zlo]: global index
  const uint32_t* lb;
  const uint32_t* delta;
  int zlo, nz, wx0, wy0, ncx, ncy;

  // LDS range of stencil column k of target column (cx, cy), rows [w0, w1]
  // (inside the target's 5-row window, so inside the chunk's window).
  __device__ __forceinline__ bool range(int cx, int cy, int k, int w0, int w1, uint32_t& b,
                                        uint32_t& e, int& kw) const {
    const int ix = cx - kStencil + k / kCols2, iy = cy - kStencil + k % kCols2;
    if (ix < 0 || ix >= ncx || iy < 0 || iy >= ncy) return false;
    const int z0 = max(w0, 0), z1 = min(w1, nz - 1);
    if (z0 > z1) return false;
    kw = (ix - wx0) * kWin3 + (iy - wy0);
    const uint32_t base = coff[kw][0];
    b = lb[kw] + (coff[kw][z0 - zlo] - base);
    e = lb[kw] + (coff[kw][z1 + 1 - zlo] - base);
    return true;
  }
};

// gather_words3's words, queued per lane in registers: 32-bit acceptance mask
// | LDS index of its bit 0 << 32 | window column << 48 (never an empty mask).
struct WordQueue3 {
  uint64_t w[kWords3];  // oldest first

  // Appends at slot n (a run-time index: one select per slot).
  __device__ __forceinline__ void put(int n, uint64_t v) {
#pragma unroll
    for (int q = 0; q < kWords3; ++q) {
      if (q == n) w[q] = v;
    }
  }
  // Removes the oldest word.
  __device__ __forceinline__ uint64_t pop() {
    const uint64_t v = w[0];
#pragma unroll
    for (int q = 0; q + 1 < kWords3; ++q) w[q] = w[q + 1];
    return v;
  }
};

// The pair loop of gather_words3's queued words (as sum_words_g).
template <int KERNEL, int FLAVOUR>
__device__ __forceinline__ void sum_words3(const DevParams& P, const LdsPos3& ps,
                                           const DevState& S, const Target<3>& a,
                                           uint32_t n_acc, Sums<3>& acc, WordQueue3& wq) {
  uint32_t m = 0, l0 = 0, d0 = 0;
  // the next of the n_acc accepted candidates
  auto take = [&](PairIn<3>& p) {
    if (m == 0) {
      const uint64_t next = wq.pop();
      m = static_cast<uint32_t>(next);
      l0 = static_cast<uint32_t>(next >> 32) & 0xFFFFu;
      d0 = ps.delta[static_cast<uint32_t>(next >> 48)];
    }
    const uint32_t l = l0 + static_cast<uint32_t>(__builtin_ctz(m));
    m &= m - 1u;
    p.x[0] = ps.sx[l];
    p.x[1] = ps.sy[l];
    p.x[2] = ps.sz[l];
    const uint32_t j = l + d0;
    p.pv = S.pv[j];
    p.aux = S.aux[j];
    p.vol = mass_flavour(FLAVOUR) ? S.vol[j] : 0.0f;
  };
  uint32_t rem = n_acc;
  PairIn<3> A, B;
  if (rem == 0) return;
  if (acc.pend) {  // an odd term of an earlier batch of words waits for its partner
    take(A);
    sum_pair<3, KERNEL, FLAVOUR>(P, a, A, acc);
    if (--rem == 0) return;
  }
  take(A);
  --rem;
  while (true) {
    const bool hb = rem > 0;
    if (hb) {
      take(B);
      --rem;
    }
    float ca, fa[3];
    term_of<3, KERNEL, FLAVOUR>(P, a, A, ca, fa);
    if (!hb) {
      acc.add(ca, fa);
      break;
    }
    const bool ha = rem > 0;
    if (ha) {
      take(A);
      --rem;
    }
    float cb, fb[3];
    term_of<3, KERNEL, FLAVOUR>(P, a, B, cb, fb);
    acc.add2(ca, fa, cb, fb);
    if (!ha) break;
  }
}

// 3D with LDS positions: the global path's walk (stencil columns in order,
// 32-candidate masks), the predicate reading the staged positions, up to
// kWords3 masks per pair loop.
template <int KERNEL, int FLAVOUR>
__device__ __forceinline__ void gather_words3(const DevParams& P, const LdsPos3& ps,
                                              const DevState& S, const Target<3>& a,
                                              uint32_t i, int cx, int cy, int zc,
                                              const Windows<3, kCols2 * kCols2>& W,
                                              Sums<3>& acc, uint32_t& count) {
  constexpr int NCOL = kCols2 * kCols2;
  WordQueue3 wq;
  int k = 0, kw = 0;
  uint32_t c0 = 0, e = 0;
  bool more = true;
  while (more) {
    int nw = 0;
    uint32_t n_acc = 0;
    while (nw < kWords3) {
      if (c0 >= e) {  // next non-empty column range
        bool found = false;
        for (; k < NCOL && !found; ++k) {
          const uint32_t v = W.get(k);
          if ((v & 7u) == 7u) continue;
          const int w0 = zc - kStencil + static_cast<int>(v & 7u);
          const int w1 = zc - kStencil + static_cast<int>(v >> 3);
          uint32_t b, eb;
          int kk;
          if (ps.range(cx, cy, k, w0, w1, b, eb, kk) && b < eb) {
            found = true;
            c0 = b;
            e = eb;
            kw = kk;
          }
        }
        if (!found) {
          more = false;
          break;
        }
      }
      const uint32_t len = min(e - c0, 32u);
      uint32_t mask = 0;
      uint32_t t = 0;
      for (; t + 4 <= len; t += 4) {
        const double x0[3] = {ps.sx[c0 + t], ps.sy[c0 + t], ps.sz[c0 + t]};
        const double x1[3] = {ps.sx[c0 + t + 1], ps.sy[c0 + t + 1], ps.sz[c0 + t + 1]};
        const double x2[3] = {ps.sx[c0 + t + 2], ps.sy[c0 + t + 2], ps.sz[c0 + t + 2]};
        const double x3[3] = {ps.sx[c0 + t + 3], ps.sy[c0 + t + 3], ps.sz[c0 + t + 3]};
        const uint32_t mm = (within<3>(P, x0, a.x) ? 1u : 0u) |
                            (within<3>(P, x1, a.x) ? 2u : 0u) |
                            (within<3>(P, x2, a.x) ? 4u : 0u) |
                            (within<3>(P, x3, a.x) ? 8u : 0u);
        mask |= mm << t;
      }
      for (; t < len; ++t) {
        const double x0[3] = {ps.sx[c0 + t], ps.sy[c0 + t], ps.sz[c0 + t]};
        if (within<3>(P, x0, a.x)) mask |= 1u << t;
      }
      const uint32_t g0 = c0 + ps.delta[kw];  // global index of bit 0
      if (i >= g0 && i < g0 + len) mask &= ~(1u << (i - g0));  // not self
      const uint32_t na = static_cast<uint32_t>(__builtin_popcount(mask));
      count += na;
      n_acc += na;
      if (mask) {
        wq.put(nw, static_cast<uint64_t>(mask) | (static_cast<uint64_t>(c0) << 32) |
                       (static_cast<uint64_t>(kw) << 48));
        ++nw;
      }
      c0 += len;
    }
    sum_words3<KERNEL, FLAVOUR>(P, ps, S, a, n_acc, acc, wq);
  }
}

__device__ __forceinline__ float wave_max_f32(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  return v;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}

// Per-wave running reductions of a persistent loop: the chunks a wave takes
// in turn mostly belong to one tank, so the tank's slots are updated once per
// change of tank (and at the end) instead of once per chunk.
struct WaveReduce {
  uint32_t t = ~0u;
  float vmax = 0.0f;
  unsigned long long dtmin = ~0ull;
  __device__ __forceinline__ void flush(const StageArgs& A) {
    if (t == ~0u) return;
    if ((threadIdx.x & 63) == 0) {
      if (vmax > 0.0f) atomicMax(red_slot(A.vmax, t), __float_as_uint(vmax));
      if (dtmin != ~0ull) atomicMin(red_slot(A.dt_min, t), dtmin);
    }
    vmax = 0.0f;
    dtmin = ~0ull;
  }
  __device__ __forceinline__ void add(const StageArgs& A, uint32_t tank, float vm,
                                      unsigned long long dm) {
    if (tank != t) {
      flush(A);
      t = tank;
    }
    vmax = fmaxf(vmax, vm);
    dtmin = dm < dtmin ? dm : dtmin;
  }
};

struct ChunkInfo {
  uint32_t t, col_row, t0, t1;
  int cx, cy, nz;
  DevTank T;
};

template <int DIM>
__device__ __forceinline__ ChunkInfo chunk_info(const StageArgs& A, uint32_t q) {
  ChunkInfo c;
  const uint4 ch = A.chunks[q];
  const uint4 col = A.cols[ch.x];
  c.t = col.x;
  c.T = A.tanks[c.t];
  c.cx = static_cast<int>(col.y);
  c.cy = static_cast<int>(col.z);
  c.col_row = col.w;
  c.nz = c.T.nc[DIM - 1];
  c.t0 = ch.y;
  c.t1 = ch.z;
  return c;
}

// One target (one lane), particle i of cell column (cx, cy) of tank `tank`
// (col_row: the global index of the column's first cell): pair sums, then
// the SSPRK substep / combination of time_integrator.hpp:188-223 in FP64
// and, on the last stage, the next step's CFL candidate.  The lane's speed
// and CFL candidate come back in vmax and dtmin for the per-wave reductions.
template <int DIM, int KERNEL, int FLAVOUR, int KIND, bool FINAL, class Pos>
__device__ __forceinline__ void target_step(const StageArgs& A, const DevTank& T,
                                            uint32_t tank, int cx, int cy, uint32_t col_row,
                                            uint32_t i, bool in_range, const Pos& ps,
                                            float& vmax, unsigned long long& dtmin) {
  constexpr int NCOL = DIM == 2 ? kCols2 : kCols2 * kCols2;
  const DevParams& P = A.P;
  const uint32_t tag = in_range ? A.tag[i] : 1u;
  const bool fluid = in_range && !(tag & 1u);
  vmax = 0.0f;
  dtmin = ~0ull;
  if (in_range && !fluid) {
    if constexpr (KIND == kRhsOnly) {
      A.rhs_drho[i] = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) A.rhs_dv[k][i] = 0.0;
      if (A.rhs_count) A.rhs_count[i] = 0u;
    } else {
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        A.out.x[k][i] = A.in.x[k][i];
        A.out.v[k][i] = 0.0;
      }
      A.out.rho[i] = A.in.rho[i];
      A.out.aux[i] = A.in.aux[i];
      A.out.pv[i] = A.in.pv[i];
      if constexpr (DIM == 3 && mass_flavour(FLAVOUR)) A.out.vol[i] = A.in.vol[i];
    }
  }
  if (fluid) {
    double xa[DIM], va[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      xa[k] = A.in.x[k][i];
      va[k] = A.in.v[k][i];
    }
    const double rho_a = A.in.rho[i];
    Target<DIM> a;
    make_target<DIM, FLAVOUR>(P, xa, va, rho_a, A.in.aux[i], a);
    const int zc = static_cast<int>(A.key[i] - col_row);
    Sums<DIM> acc;
    acc.init();
    const Windows<DIM, NCOL> W =
        stencil_windows<DIM, NCOL>(P, T, xa, cx, cy, zc);
    uint32_t count = 0;
    if constexpr (std::is_same_v<Pos, LdsPos>) {
      gather_words<KERNEL, FLAVOUR, NCOL>(P, ps, A.in, a, i, zc, W, acc, count);
    } else if constexpr (std::is_same_v<Pos, LdsPos3>) {
      gather_words3<KERNEL, FLAVOUR>(P, ps, A.in, a, i, cx, cy, zc, W, acc, count);
    } else {
      gather_words_g<DIM, KERNEL, FLAVOUR, NCOL>(P, ps, A.in, a, i, zc, W, acc, count);
    }
    acc.finish();
    if constexpr (KIND == kRhsOnly) {
      if (A.rhs_count) A.rhs_count[i] = count;
    }
    double drho = static_cast<double>(acc.rho.value());
    double dv[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) dv[k] = static_cast<double>(acc.v[k].value());
    if constexpr (FLAVOUR == kFlavourWall) {
      // face terms (already divided by gamma_a) + pair sums / gamma_a; none
      // without walls (per-particle masses alone)
      if (A.wall) {
        const double* wt = A.wall + 4 * static_cast<size_t>(i);
        const double ig = 1.0 / wt[3];
        drho = wt[0] + drho * ig;
#pragma unroll
        for (int k = 0; k < DIM; ++k) dv[k] = wt[1 + k] + dv[k] * ig;
      }
    }
    dv[1] = dv[1] - P.g;
    if constexpr (KIND == kRhsOnly) {
      A.rhs_drho[i] = drho;
#pragma unroll
      for (int k = 0; k < DIM; ++k) A.rhs_dv[k][i] = dv[k];
    } else {
      double v2 = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) v2 = v2 + va[k] * va[k];
      vmax = static_cast<float>(sqrt(v2));
      const double h = A.dt[tank];
      double rn[DIM], vn[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        rn[k] = xa[k] + h * va[k];
        vn[k] = va[k] + h * dv[k];
      }
      double rhon = rho_a + h * drho;
      if constexpr (KIND == kCombine) {
        const double wn = A.w, wo = 1.0 - A.w;
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
          rn[k] = wo * A.old.x[k][i] + wn * rn[k];
          vn[k] = wo * A.old.v[k][i] + wn * vn[k];
        }
        rhon = wo * A.old.rho[i] + wn * rhon;
      }
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        A.out.x[k][i] = rn[k];
        A.out.v[k][i] = vn[k];
      }
      A.out.rho[i] = rhon;
      write_pair_data<DIM>(P, A.out, i, vn, rhon, P.mass ? P.mass[A.id[i]] : P.m);
      if constexpr (FINAL) {
        dtmin = dbl_bits(particle_time_step<DIM>(P, rhon, vn, dv));
        if (A.dv_last[0]) {  // kept for the time step after post-integration
#pragma unroll
          for (int k = 0; k < DIM; ++k) A.dv_last[k][i] = dv[k];
        }
      }
    }
  }
}

// The per-wave reductions of one chunk's targets (all of one tank).
template <int KIND, bool FINAL>
__device__ __forceinline__ void reduce_chunk(const StageArgs& A, uint32_t tank, float vmax,
                                             unsigned long long dtmin, WaveReduce& red) {
  if constexpr (KIND != kRhsOnly) {
    const float vm = wave_max_f32(vmax);
    unsigned long long dm = ~0ull;
    if constexpr (FINAL) dm = wave_min_u64(dtmin);
    red.add(A, tank, vm, dm);
  }
}

// Targets of one column chunk, one per lane.
template <int DIM, int KERNEL, int FLAVOUR, int KIND, bool FINAL, class Pos>
__device__ __forceinline__ void chunk_targets(const StageArgs& A, const ChunkInfo& c,
                                              const Pos& ps, WaveReduce& red) {
  const uint32_t i = c.t0 + threadIdx.x;
  float vmax;
  unsigned long long dtmin;
  target_step<DIM, KERNEL, FLAVOUR, KIND, FINAL>(A, c.T, c.t, c.cx, c.cy, c.col_row, i,
                                                 i < c.t1, ps, vmax, dtmin);
  reduce_chunk<KIND, FINAL>(A, c.t, vmax, dtmin, red);
}

struct Band {
  uint32_t lo, hi, step;
};
__device__ __forceinline__ Band xcd_band(uint32_t n) {
  const uint32_t x = blockIdx.x & 7u, slot = blockIdx.x >> 3;
  Band b;
  b.lo = static_cast<uint32_t>((static_cast<uint64_t>(n) * x) / 8) + slot;
  b.hi = static_cast<uint32_t>((static_cast<uint64_t>(n) * (x + 1)) / 8);
  b.step = gridDim.x >> 3;
  return b;
}

}  // namespace

}  // namespace tit_amd
