// Fused SSPRK stage kernel: continuity + momentum pair sums, EOS, update.
//
// Work unit: a chunk of at most chunk_targets consecutive (cell-sorted)
// particles of one cell column of one tank, one particle per lane.  After
// every cell rebuild the chunks are recut on the device so that each column's
// particles split into equal chunks; a persistent grid then walks the chunk
// list in XCD-major bands (launch index b runs on XCD b % 8 and takes chunks
// from that XCD's eighth of the list, so each XCD's L2 holds a contiguous
// band of columns).
//
// 2D: the FP64 positions of the chunk's candidates (the 5 neighbouring
// columns, the chunk's rows +- 2) are staged in LDS.  Each lane walks its own
// 5 x 5 cell stencil one column range at a time: the exact FP64 predicate
// (four LDS loads per batch) builds a 64-bit acceptance mask per batch, kept
// in LDS; then the FP32 pair terms run over the accepted candidates of all
// batches in one loop, reading their pair records (velocity, m/rho,
// rho - rho0 and p/rho^2, cs) from L2 one candidate ahead of the arithmetic.
// Chunks whose candidates exceed the LDS capacity, and 3D, take positions
// from global memory too.  Every path visits candidates in the same order and
// rounds identically, so all schedules produce the same bits.
#include <algorithm>
#include <type_traits>

#include "pair_math.hpp"
#include "sph_kernels.hpp"

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
  uint64_t* wmask;  // this wave's [kWords][64] acceptance masks
  uint32_t* wjl;    // their bit 0: global index - gbase (bits 10..31), LDS index

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
  __device__ __forceinline__ void pos(int k, uint32_t j, double* x) const {
    const double2 p = xy[lb[k] + (j - gb[k])];
    x[0] = p.x;
    x[1] = p.y;
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
  __device__ __forceinline__ void pos(int, uint32_t j, double* x) const {
#pragma unroll
    for (int k = 0; k < DIM; ++k) x[k] = S.x[k][j];
  }
};

template <int DIM>
struct PairIn {
  double x[DIM];
  float4 pv;
  float2 aux;
};

template <int DIM, class Pos>
__device__ __forceinline__ void load_pair(const Pos& ps, const DevState& S, int k,
                                          uint32_t j, PairIn<DIM>& p) {
  ps.pos(k, j, p.x);
  p.pv = S.pv[j];
  p.aux = S.aux[j];
}

template <int DIM, int KERNEL, int FLAVOUR>
__device__ __forceinline__ void term_of(const DevParams& P, const Target<DIM>& a,
                                        const PairIn<DIM>& p, float& c, float* f) {
  double d[DIM];
#pragma unroll
  for (int q = 0; q < DIM; ++q) d[q] = p.x[q] - a.x[q];
  pair_term<DIM, KERNEL, FLAVOUR>(P, a, d, p.pv, p.aux.x, p.aux.y, c, f);
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

// Per-target stencil trimming.  A particle of cell (ix, iy, z) lies within
// skin/2 of that cell's box (it is sorted into it at the last rebuild and
// has moved less than skin/2 since; only border cells hold particles outside
// the grid, clamped in).  So for a target whose whole 5^d window is interior
// (no border cell), a column whose grown box is farther than R from the
// target in the column plane holds no neighbour, and only the rows within
// sqrt(R^2 - gap^2) + skin/2 of the target can.  The windows are widened by
// a relative 1e-6 margin; trimming removes only candidates the exact FP64
// predicate rejects, so the accepted set and its order are unchanged.
struct Trim {
  bool on;
  double x[3];
  double lo[3], c[3], inv_c[3];
  double R2, s;
};

template <int DIM>
__device__ __forceinline__ Trim make_trim(const DevParams& P, const DevTank& T,
                                          const double* xa, int cx, int cy, int zc) {
  constexpr int L = DIM - 1;
  Trim t;
  bool inner = cx - kStencil > 0 && cx + kStencil < T.nc[0] - 1 && zc - kStencil > 0 &&
               zc + kStencil < T.nc[L] - 1;
  if constexpr (DIM == 3) inner = inner && cy - kStencil > 0 && cy + kStencil < T.nc[1] - 1;
  t.on = inner;
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    t.x[k] = xa[k];
    t.lo[k] = T.lo[k];
    t.inv_c[k] = T.inv_c[k];
    t.c[k] = 1.0 / T.inv_c[k];
  }
  t.R2 = P.R2 * (1.0 + 1e-6);
  t.s = P.skin_half * (1.0 + 1e-6);
  return t;
}

// Row window [w0, w1] of stencil column k; false: no neighbour in column k.
template <int DIM>
__device__ __forceinline__ bool trim_window(const Trim& t, int cx, int cy, int zc, int k,
                                            int& w0, int& w1) {
  constexpr int L = DIM - 1;
  w0 = zc - kStencil;
  w1 = zc + kStencil;
  if (!t.on) return true;
  int ixk, iyk = 0;
  if constexpr (DIM == 2) {
    ixk = cx - kStencil + k;
  } else {
    ixk = cx - kStencil + k / kCols2;
    iyk = cy - kStencil + k % kCols2;
  }
  auto gap = [&](int ax, int ic) {
    const double a0 = t.lo[ax] + ic * t.c[ax] - t.s;
    const double a1 = t.lo[ax] + (ic + 1) * t.c[ax] + t.s;
    return fmax(0.0, fmax(a0 - t.x[ax], t.x[ax] - a1));
  };
  const double gx = gap(0, ixk);
  double g2 = gx * gx;
  if constexpr (DIM == 3) {
    const double gy = gap(1, iyk);
    g2 = g2 + gy * gy;
  }
  if (g2 > t.R2) return false;
  const double half = sqrt(t.R2 - g2) + t.s;
  const double z = t.x[L] - t.lo[L];
  w0 = max(w0, static_cast<int>(floor((z - half) * t.inv_c[L] - 1e-6)));
  w1 = min(w1, static_cast<int>(floor((z + half) * t.inv_c[L] + 1e-6)));
  return w0 <= w1;
}

// Pair sums of one target over its stencil, column range by column range.
// The row windows of the NCOL stencil columns, packed 6 bits per column
// (3-bit offsets of the first and last row from zc - 2; 7 = empty column), so
// no FP64 trimming state stays live through the pair loop.
template <int DIM, int NCOL>
struct Windows {
  uint32_t w[(NCOL * 6 + 31) / 32];
  __device__ __forceinline__ void set(int k, uint32_t v) {
    const int bit = 6 * k;
    w[bit >> 5] |= v << (bit & 31);
    if ((bit & 31) > 26) w[(bit >> 5) + 1] |= v >> (32 - (bit & 31));
  }
  __device__ __forceinline__ uint32_t get(int k) const {
    const int bit = 6 * k;
    uint32_t v = w[bit >> 5] >> (bit & 31);
    if ((bit & 31) > 26) v |= w[(bit >> 5) + 1] << (32 - (bit & 31));
    return v & 63u;
  }
};

template <int DIM, int NCOL>
__device__ __forceinline__ Windows<DIM, NCOL> stencil_windows(const Trim& tr, int cx, int cy,
                                                              int zc) {
  Windows<DIM, NCOL> W;
#pragma unroll
  for (int q = 0; q < (NCOL * 6 + 31) / 32; ++q) W.w[q] = 0u;
  for (int k = 0; k < NCOL; ++k) {
    int w0, w1;
    uint32_t v = 7u | (7u << 3);
    if (trim_window<DIM>(tr, cx, cy, zc, k, w0, w1)) {
      v = static_cast<uint32_t>(w0 - (zc - kStencil)) |
          (static_cast<uint32_t>(w1 - (zc - kStencil)) << 3);
    }
    W.set(k, v);
  }
  return W;
}

template <int DIM, int KERNEL, int FLAVOUR, int NCOL, class Pos>
__device__ __forceinline__ void gather(const DevParams& P, const Pos& ps,
                                       const DevState& S, const Target<DIM>& a,
                                       uint32_t i, int zc, const Windows<DIM, NCOL>& W,
                                       Sums<DIM>& acc, uint32_t& count) {
  for (int k = 0; k < NCOL; ++k) {
    uint32_t b, e;
    const uint32_t v = W.get(k);
    if ((v & 7u) == 7u) continue;
    const int w0 = zc - kStencil + static_cast<int>(v & 7u);
    const int w1 = zc - kStencil + static_cast<int>(v >> 3);
    if (!ps.range(k, w0, w1, b, e)) continue;
    for (uint32_t c0 = b; c0 < e; c0 += 64) {
      const uint32_t len = min(e - c0, 64u);
      uint64_t mask = 0;
      uint32_t t = 0;
      for (; t + 4 <= len; t += 4) {
        double x0[DIM], x1[DIM], x2[DIM], x3[DIM];
        ps.pos(k, c0 + t, x0);
        ps.pos(k, c0 + t + 1, x1);
        ps.pos(k, c0 + t + 2, x2);
        ps.pos(k, c0 + t + 3, x3);
        const uint64_t m = (within<DIM>(P, x0, a.x) ? 1ull : 0ull) |
                           (within<DIM>(P, x1, a.x) ? 2ull : 0ull) |
                           (within<DIM>(P, x2, a.x) ? 4ull : 0ull) |
                           (within<DIM>(P, x3, a.x) ? 8ull : 0ull);
        mask |= m << t;
      }
      for (; t < len; ++t) {
        double x0[DIM];
        ps.pos(k, c0 + t, x0);
        if (within<DIM>(P, x0, a.x)) mask |= 1ull << t;
      }
      if (i >= c0 && i < c0 + len) mask &= ~(1ull << (i - c0));  // not self
      count += static_cast<uint32_t>(__builtin_popcountll(mask));
      if (!mask) continue;
      PairIn<DIM> A, B;
      if (acc.pend) {  // the odd term of an earlier batch waits for its partner
        load_pair<DIM>(ps, S, k, c0 + pop_lowest(mask), A);
        sum_pair<DIM, KERNEL, FLAVOUR>(P, a, A, acc);
        if (!mask) continue;
      }
      // Two register sets alternate: the next accepted candidate's loads are
      // in flight while the current one is summed; two terms enter the
      // compensated sums together.
      load_pair<DIM>(ps, S, k, c0 + pop_lowest(mask), A);
      while (true) {
        const bool hb = mask != 0;
        if (hb) load_pair<DIM>(ps, S, k, c0 + pop_lowest(mask), B);
        float ca, fa[DIM];
        term_of<DIM, KERNEL, FLAVOUR>(P, a, A, ca, fa);
        if (!hb) {
          acc.add(ca, fa);
          break;
        }
        const bool ha = mask != 0;
        if (ha) load_pair<DIM>(ps, S, k, c0 + pop_lowest(mask), A);
        float cb, fb[DIM];
        term_of<DIM, KERNEL, FLAVOUR>(P, a, B, cb, fb);
        acc.add2(ca, fa, cb, fb);
        if (!ha) break;
      }
    }
  }
}

// 2D with LDS positions.  The acceptance masks of up to kWords candidate
// batches (normally one per stencil column) are kept per lane in LDS and
// their accepted candidates are summed in one loop, so a wave's pair loop
// runs for the largest accepted count among its lanes, not for the sum over
// columns of each column's largest count.  Candidates are visited in the
// same order as in gather(), so the sums are identical.
template <int KERNEL, int FLAVOUR>
__device__ __forceinline__ void sum_words(const DevParams& P, const LdsPos& ps,
                                          const DevState& S, const Target<2>& a,
                                          uint32_t n_acc, Sums<2>& acc) {
  const int lane = threadIdx.x & 63;
  int w = 0;
  uint64_t m = 0, m_next = ps.wmask[lane];  // the next word is read one word ahead
  uint32_t l0 = 0, j0 = 0, jl_next = ps.wjl[lane];
  // The next of the n_acc accepted candidates (stored words are never
  // empty; slots past the last stored word are read but never used).
  auto take = [&](PairIn<2>& p) {
    if (m == 0) {
      m = m_next;
      j0 = ps.gbase + (jl_next >> 10);
      l0 = jl_next & 1023u;
      w = min(w + 1, kWords - 1);
      m_next = ps.wmask[w * 64 + lane];
      jl_next = ps.wjl[w * 64 + lane];
    }
    const uint32_t t = pop_lowest(m);
    const double2 x = ps.xy[l0 + t];
    p.x[0] = x.x;
    p.x[1] = x.y;
    p.pv = S.pv[j0 + t];
    p.aux = S.aux[j0 + t];
  };
  uint32_t rem = n_acc;
  PairIn<2> A, B;
  if (rem == 0) return;
  if (acc.pend) {  // an odd term of an earlier batch of words waits for its partner
    take(A);
    sum_pair<2, KERNEL, FLAVOUR>(P, a, A, acc);
    if (--rem == 0) return;
  }
  // two register sets alternate, as in gather(); their terms enter the
  // compensated sums together
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
  const int lane = threadIdx.x & 63;
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
        ps.wmask[nw * 64 + lane] = mask;
        ps.wjl[nw * 64 + lane] = ((c0 - ps.gbase) << 10) | lc;
        ++nw;
      }
      c0 += len;
    }
    sum_words<KERNEL, FLAVOUR>(P, ps, S, a, n_acc, acc);
  }
}

// Global-memory path, the same scheme with 32-candidate masks (the 3D stencil
// has 25 short column ranges): all columns' accepted candidates are summed
// in one loop per lane, in the same order as gather().
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

// Targets of one chunk (one per lane): pair sums, then the SSPRK substep /
// combination of time_integrator.hpp:188-223 in FP64 and, on the last stage,
// the next step's CFL candidate; per-wave reductions at the end.
template <int DIM, int KERNEL, int FLAVOUR, int KIND, bool FINAL, class Pos>
__device__ __forceinline__ void chunk_targets(const StageArgs& A, const ChunkInfo& c,
                                              const Pos& ps, WaveReduce& red) {
  constexpr int NCOL = DIM == 2 ? kCols2 : kCols2 * kCols2;
  const DevParams& P = A.P;
  const uint32_t i = c.t0 + threadIdx.x;
  const bool in_range = i < c.t1;
  const uint32_t tag = in_range ? A.tag[i] : 1u;
  const bool fluid = in_range && !(tag & 1u);
  float vmax = 0.0f;
  unsigned long long dtmin = ~0ull;
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
    const int zc = static_cast<int>(A.key[i] - c.col_row);
    Sums<DIM> acc;
    acc.init();
    const Windows<DIM, NCOL> W =
        stencil_windows<DIM, NCOL>(make_trim<DIM>(P, c.T, xa, c.cx, c.cy, zc), c.cx, c.cy, zc);
    uint32_t count = 0;
    if constexpr (std::is_same_v<Pos, LdsPos>) {
      gather_words<KERNEL, FLAVOUR, NCOL>(P, ps, A.in, a, i, zc, W, acc, count);
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
      // face terms (already divided by gamma_a) + pair sums / gamma_a
      const double* wt = A.wall + 4 * static_cast<size_t>(i);
      const double ig = 1.0 / wt[3];
      drho = wt[0] + drho * ig;
#pragma unroll
      for (int k = 0; k < DIM; ++k) dv[k] = wt[1 + k] + dv[k] * ig;
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
      const double h = A.dt[c.t];
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
      write_pair_data<DIM>(P, A.out, i, vn, rhon);
      if constexpr (FINAL) {
        dtmin = dbl_bits(particle_time_step<DIM>(P, rhon, vn, dv));
        if (A.dv_last[0]) {  // kept for the time step after post-integration
#pragma unroll
          for (int k = 0; k < DIM; ++k) A.dv_last[k][i] = dv[k];
        }
      }
    }
  }
  if constexpr (KIND != kRhsOnly) {
    // all targets of a chunk belong to one tank
    const float vm = wave_max_f32(vmax);
    unsigned long long dm = ~0ull;
    if constexpr (FINAL) dm = wave_min_u64(dtmin);
    red.add(A, c.t, vm, dm);
  }
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

// 2D chunks with LDS-staged positions; chunks that do not fit go to the
// overflow list.
template <int KERNEL, int FLAVOUR, int KIND, bool FINAL, int NT, int CAP>
__global__ __launch_bounds__(NT) void k_stage_lds(StageArgs A) {
  __shared__ double2 s_xy[CAP];
  __shared__ uint32_t s_coff[kCols2][kRowsMax + 1];
  __shared__ uint32_t s_gb[kCols2], s_lb[kCols2 + 1];
  __shared__ int s_fits;
  __shared__ uint64_t s_wmask[NT / 64][kWords * 64];
  __shared__ uint32_t s_wjl[NT / 64][kWords * 64];
  __shared__ uint32_t s_gbase;
  const int tid = threadIdx.x;
  const Band band = xcd_band(A.counts[0]);
  WaveReduce red;
  for (uint32_t q = band.lo; q < band.hi; q += band.step) {
    const ChunkInfo c = chunk_info<2>(A, q);
    const int zf = static_cast<int>(A.key[c.t0] - c.col_row);
    const int zl = static_cast<int>(A.key[c.t1 - 1] - c.col_row);
    const int zlo = max(zf - kStencil, 0), zhi = min(zl + kStencil, c.nz - 1);
    const int nrows = zhi - zlo + 1;
    __syncthreads();  // the previous chunk's LDS reads are done
    const bool rows_fit = nrows <= kRowsMax;
    if (rows_fit) {
      for (int r = tid; r < kCols2 * (nrows + 1); r += NT) {
        const int k = r / (nrows + 1), row = r % (nrows + 1);
        const int ixk = c.cx - kStencil + k;
        s_coff[k][row] =
            (ixk >= 0 && ixk < c.T.nc[0])
                ? A.cell_off[c.T.cell_off + static_cast<uint32_t>(ixk) * c.nz + zlo + row]
                : 0u;
      }
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t lb = 0, g0 = ~0u, g1 = 0;
      if (rows_fit) {
        for (int k = 0; k < kCols2; ++k) {
          s_gb[k] = s_coff[k][0];
          s_lb[k] = lb;
          lb += s_coff[k][nrows] - s_coff[k][0];
          const int ixk = c.cx - kStencil + k;
          if (ixk >= 0 && ixk < c.T.nc[0]) {
            g0 = min(g0, s_coff[k][0]);
            g1 = max(g1, s_coff[k][nrows]);
          }
        }
        s_lb[kCols2] = lb;
        s_gbase = g0;
      }
      s_fits = rows_fit && lb <= A.lds_cap && g1 - g0 < kWordSpan;
      if (!s_fits) A.overflow[atomicAdd(A.counts + 1, 1u)] = q;
    }
    __syncthreads();
    if (!s_fits) continue;
    const uint32_t total = s_lb[kCols2];
    for (uint32_t l = tid; l < total; l += NT) {
      int k = 0;
      while (l >= s_lb[k + 1]) ++k;
      const uint32_t j = s_gb[k] + (l - s_lb[k]);
      s_xy[l] = make_double2(A.in.x[0][j], A.in.x[1][j]);
    }
    __syncthreads();
    const LdsPos ps{s_xy,   s_coff,   s_gb,      s_lb,           zlo,           c.nz,
                    c.cx,   c.T.nc[0], s_gbase,  s_wmask[tid >> 6], s_wjl[tid >> 6]};
    chunk_targets<2, KERNEL, FLAVOUR, KIND, FINAL>(A, c, ps, red);
  }
  red.flush(A);
}

// Chunks with positions from global memory: 3D, and 2D overflow chunks.
template <int DIM, int KERNEL, int FLAVOUR, int KIND, bool FINAL, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_stage_global(
    StageArgs A, const uint32_t* list, const uint32_t* count) {
  __shared__ uint32_t s_wm[NT / 64][kWordsG * 64];
  __shared__ uint32_t s_wj[NT / 64][kWordsG * 64];
  const Band band = xcd_band(*count);
  WaveReduce red;
  for (uint32_t q = band.lo; q < band.hi; q += band.step) {
    const ChunkInfo c = chunk_info<DIM>(A, list ? list[q] : q);
    const GlobalPos<DIM> ps{c.T,  A.in, A.cell_off, c.cx, c.cy, s_wm[threadIdx.x >> 6],
                            s_wj[threadIdx.x >> 6]};
    chunk_targets<DIM, KERNEL, FLAVOUR, KIND, FINAL>(A, c, ps, red);
  }
  red.flush(A);
}

// ---- chunk list (after every cell rebuild) ----------------------------------

constexpr int kScanBlock = 1024;

// The chunks of column col_i, in order: the column's particles in
// ceil(n / chunk_targets) equal parts, each cut further into pieces whose
// targets span at most kChunkRows cells (a piece starting at particle p ends
// before the first particle with key >= key[p] + kChunkRows).  Calls
// emit(k, p0, p1) for the k-th chunk and returns the count.  A column gains
// at most (nz - 1) / kChunkRows cuts: each cut starts a piece kChunkRows
// cells or more above the last, and the parts' spans do not overlap.
template <class F>
__device__ __forceinline__ uint32_t column_chunks(const StageArgs& A, uint32_t col_i, F&& emit) {
  const uint4 col = A.cols[col_i];
  const DevTank& T = A.tanks[col.x];
  const int nz = T.nc[A.P.dim - 1];
  const uint32_t beg = A.cell_off[col.w];
  const uint32_t n = A.cell_off[col.w + nz] - beg;
  const uint32_t nch = (n + A.chunk_targets - 1) / A.chunk_targets;
  uint32_t count = 0;
  for (uint32_t k = 0; k < nch; ++k) {
    const uint32_t p0 = beg + static_cast<uint32_t>((static_cast<uint64_t>(n) * k) / nch);
    const uint32_t p1 = beg + static_cast<uint32_t>((static_cast<uint64_t>(n) * (k + 1)) / nch);
    uint32_t p = p0;
    while (p < p1) {
      const uint32_t lim = A.key[p] + static_cast<uint32_t>(kChunkRows);
      uint32_t q = p1;
      if (A.key[p1 - 1] >= lim) {
        // the first particle in (p, p1) with key >= lim (keys ascend)
        uint32_t lo = p + 1, hi = p1 - 1;
        while (lo < hi) {
          const uint32_t mid = lo + (hi - lo) / 2;
          if (A.key[mid] >= lim) {
            hi = mid;
          } else {
            lo = mid + 1;
          }
        }
        q = lo;
      }
      emit(count, p, q);
      ++count;
      p = q;
    }
  }
  return count;
}

// Exclusive scan over a 1024-thread block: wave shuffles, then the 16 wave
// totals.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum,
                                                    uint32_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) {
      const uint32_t t = wsum[w];
      wsum[w] = s;
      s += t;
    }
    wsum[kScanBlock / 64] = s;
  }
  __syncthreads();
  const uint32_t r = wsum[wave] + x - v;
  *total = wsum[kScanBlock / 64];
  __syncthreads();
  return r;
}

// Chunk list in column order, in three passes: chunks per column and their
// per-workgroup totals; an exclusive scan of the totals (one workgroup, also
// writes the chunk count); each column's chunks at its scanned offset.
__global__ __launch_bounds__(kScanBlock) void k_chunk_count(StageArgs A, uint32_t* block_tot,
                                                           const uint32_t* ctl) {
  if (!ctl[0]) return;
  __shared__ uint32_t wsum[kScanBlock / 64 + 1];
  const uint32_t c = blockIdx.x * kScanBlock + threadIdx.x;
  uint32_t nch = 0;
  if (c < A.n_cols) nch = column_chunks(A, c, [](uint32_t, uint32_t, uint32_t) {});
  uint32_t tot;
  (void)block_excl_scan(nch, wsum, &tot);
  if (threadIdx.x == 0) block_tot[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanBlock) void k_chunk_base(StageArgs A, uint32_t* block_tot,
                                                          uint32_t n_blocks, const uint32_t* ctl) {
  if (!ctl[0]) return;
  __shared__ uint32_t wsum[kScanBlock / 64 + 1];
  uint32_t base_off = 0;
  for (uint32_t base = 0; base < n_blocks; base += kScanBlock) {
    const uint32_t b = base + threadIdx.x;
    const uint32_t v = b < n_blocks ? block_tot[b] : 0u;
    uint32_t tot;
    const uint32_t off = base_off + block_excl_scan(v, wsum, &tot);
    if (b < n_blocks) block_tot[b] = off;
    base_off += tot;
  }
  if (threadIdx.x == 0) {
    A.counts[0] = base_off;  // chunks
    A.counts[1] = 0;         // overflow chunks (appended by the stage kernel)
  }
}

__global__ __launch_bounds__(kScanBlock) void k_chunk_emit(StageArgs A,
                                                          const uint32_t* block_tot,
                                                          const uint32_t* ctl) {
  if (!ctl[0]) return;
  __shared__ uint32_t wsum[kScanBlock / 64 + 1];
  const uint32_t c = blockIdx.x * kScanBlock + threadIdx.x;
  // the count again (the scan needs every column's before the emission)
  uint32_t nch = 0;
  if (c < A.n_cols) nch = column_chunks(A, c, [](uint32_t, uint32_t, uint32_t) {});
  uint32_t tot;
  const uint32_t off = block_tot[blockIdx.x] + block_excl_scan(nch, wsum, &tot);
  if (c < A.n_cols) {
    (void)column_chunks(A, c, [&](uint32_t k, uint32_t p0, uint32_t p1) {
      A.chunks[off + k] = make_uint4(c, p0, p1, 0u);
    });
  }
}

__global__ void k_reset_overflow(uint32_t* counts) { counts[1] = 0u; }

// ---- dispatch -------------------------------------------------------------------

constexpr int kThreads = kChunkMax;

template <int DIM, int KERNEL, int FLAVOUR, int KIND, bool FINAL>
void launch_kind(hipStream_t s, const StageArgs& a) {
  if constexpr (DIM == 2) {
    k_reset_overflow<<<1, 1, 0, s>>>(a.counts);
    k_stage_lds<KERNEL, FLAVOUR, KIND, FINAL, kThreads, kStageLdsCap>
        <<<a.grid_lds, kThreads, 0, s>>>(a);
    k_stage_global<2, KERNEL, FLAVOUR, KIND, FINAL, kThreads>
        <<<a.grid_global, kThreads, 0, s>>>(a, a.overflow, a.counts + 1);
  } else {
    k_stage_global<3, KERNEL, FLAVOUR, KIND, FINAL, kThreads>
        <<<a.grid_global, kThreads, 0, s>>>(a, nullptr, a.counts);
  }
}

template <int DIM, int KERNEL, int FLAVOUR>
void dispatch_kind(hipStream_t s, const StageArgs& a, int kind, bool final) {
  if (kind == kRhsOnly) {
    launch_kind<DIM, KERNEL, FLAVOUR, kRhsOnly, false>(s, a);
  } else if (kind == kEuler) {
    if (final) {
      launch_kind<DIM, KERNEL, FLAVOUR, kEuler, true>(s, a);
    } else {
      launch_kind<DIM, KERNEL, FLAVOUR, kEuler, false>(s, a);
    }
  } else {
    if (final) {
      launch_kind<DIM, KERNEL, FLAVOUR, kCombine, true>(s, a);
    } else {
      launch_kind<DIM, KERNEL, FLAVOUR, kCombine, false>(s, a);
    }
  }
}

template <int DIM, int KERNEL>
void dispatch_flavour(hipStream_t s, const StageArgs& a, int kind, bool final) {
  if (a.P.flavour == kFlavourCfg) {
    dispatch_kind<DIM, KERNEL, kFlavourCfg>(s, a, kind, final);
  } else if (a.P.flavour == kFlavourWall) {
    if constexpr (DIM == 2) dispatch_kind<DIM, KERNEL, kFlavourWall>(s, a, kind, final);
  } else {
    dispatch_kind<DIM, KERNEL, kFlavourRef>(s, a, kind, final);
  }
}

template <int DIM>
void dispatch_kernel(hipStream_t s, const StageArgs& a, int kind, bool final) {
  if (a.P.kernel == kKernelC4) {
    dispatch_flavour<DIM, kKernelC4>(s, a, kind, final);
  } else {
    dispatch_flavour<DIM, kKernelC2>(s, a, kind, final);
  }
}

template <class F>
int occupancy(F* f) {
  int n = 0;
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(f),
                                                     kThreads, 0);
  return std::max(n, 1);
}

}  // namespace

uint32_t stage_lds_capacity(int dim) { return dim == 2 ? kStageLdsCap : 0u; }

void stage_grids(const DevParams& P, int n_cu, uint32_t* grid_lds,
                 uint32_t* grid_global) {
  int bl, bg;
  if (P.dim == 2) {
    bl = occupancy(k_stage_lds<kKernelC2, kFlavourCfg, kCombine, false, kThreads, kStageLdsCap>);
    bg = occupancy(k_stage_global<2, kKernelC2, kFlavourCfg, kCombine, false, kThreads>);
  } else {
    bl = 1;
    bg = occupancy(k_stage_global<3, kKernelC2, kFlavourCfg, kCombine, false, kThreads>);
  }
  // multiples of 8: every XCD owns the same number of slots
  *grid_lds = 8u * static_cast<uint32_t>((bl * n_cu + 7) / 8);
  *grid_global = 8u * static_cast<uint32_t>((bg * n_cu + 7) / 8);
}

void launch_build_chunks(hipStream_t s, const StageArgs& a, uint32_t* block_tot,
                         const uint32_t* ctl) {
  const uint32_t nb = (a.n_cols + kScanBlock - 1) / kScanBlock;
  if (nb == 0) return;
  k_chunk_count<<<nb, kScanBlock, 0, s>>>(a, block_tot, ctl);
  k_chunk_base<<<1, kScanBlock, 0, s>>>(a, block_tot, nb, ctl);
  k_chunk_emit<<<nb, kScanBlock, 0, s>>>(a, block_tot, ctl);
}

void launch_stage(hipStream_t s, const StageArgs& a, int kind, bool final) {
  if (a.n == 0 || a.n_cols == 0) return;
  if (a.P.dim == 2) {
    dispatch_kernel<2>(s, a, kind, final);
  } else {
    dispatch_kernel<3>(s, a, kind, final);
  }
}

}  // namespace tit_amd
