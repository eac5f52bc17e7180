// WCSPH step kernels for gfx950.
//
// Particles live in cell order (sorted by global cell key).  A fluid target
// gathers over the 5^d-cell stencil around the cell it was sorted into; with
// the last axis fastest in the cell key (source/tit/geom/grid.hpp:120-128),
// every stencil row along that axis is one contiguous particle range.  The
// inclusive FP64 predicate decides membership, so the neighbour set is the
// reference's exactly, independent of how the cells were built.
#include <algorithm>

#include "pair_math.hpp"
#include "sph_kernels.hpp"

namespace tit_amd {

namespace {

constexpr int kBlock = 256;

inline uint32_t grid_for(uint32_t n, int block = kBlock) {
  return (n + block - 1) / block;
}

// ---- stencil traversal ------------------------------------------------------

// Calls f(j, d, s) for every particle j != i with s = norm2(d) <= (2h)^2,
// d = r_j - r_i, in stencil order.
template <int DIM, class F>
__device__ __forceinline__ void for_each_neighbor(
    const DevParams& P, const DevTank& T, const DevState& S,
    const uint32_t* __restrict__ coff, uint32_t key, uint32_t i,
    const double* xa, F&& f) {
  const uint32_t local = key - T.cell_off;
  int c[3];
  if constexpr (DIM == 2) {
    c[0] = static_cast<int>(local / T.nc[1]);
    c[1] = static_cast<int>(local % T.nc[1]);
  } else {
    const uint32_t t = local / T.nc[2];
    c[2] = static_cast<int>(local % T.nc[2]);
    c[1] = static_cast<int>(t % T.nc[1]);
    c[0] = static_cast<int>(t / T.nc[1]);
  }
  constexpr int L = DIM - 1;
  const int zlo = max(c[L] - kStencil, 0);
  const int zhi = min(c[L] + kStencil, T.nc[L] - 1);
  const int x0 = max(c[0] - kStencil, 0), x1 = min(c[0] + kStencil, T.nc[0] - 1);
  int y0 = 0, y1 = 0;
  if constexpr (DIM == 3) {
    y0 = max(c[1] - kStencil, 0);
    y1 = min(c[1] + kStencil, T.nc[1] - 1);
  }
  for (int ix = x0; ix <= x1; ++ix) {
    for (int iy = y0; iy <= y1; ++iy) {
      uint32_t row;
      if constexpr (DIM == 2) {
        row = T.cell_off + static_cast<uint32_t>(ix) * T.nc[1];
      } else {
        row = T.cell_off +
              (static_cast<uint32_t>(ix) * T.nc[1] + static_cast<uint32_t>(iy)) *
                  T.nc[2];
      }
      const uint32_t b = coff[row + zlo], e = coff[row + zhi + 1];
      for (uint32_t j = b; j < e; ++j) {
        double d[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) d[k] = S.x[k][j] - xa[k];
        double s = d[0] * d[0];
#pragma unroll
        for (int k = 1; k < DIM; ++k) s = s + d[k] * d[k];
        if (s <= P.R2 && j != i) f(j, d, s);
      }
    }
  }
}

// ---- wave-level reductions to per-tank slots ----------------------------------

__device__ __forceinline__ void wave_atomic_min_u64(unsigned long long* base,
                                                    uint32_t t,
                                                    unsigned long long v,
                                                    bool active) {
  const uint64_t act = __ballot(active);
  if (act == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll(static_cast<long long>(act)) - 1;
  const uint32_t t0 = __shfl(t, leader);
  const bool same = __ballot(active && t == t0) == act;
  if (same) {
    unsigned long long m = active ? v : ~0ull;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(m, off);
      m = o < m ? o : m;
    }
    if (lane == leader) atomicMin(red_slot(base, t0), m);
  } else if (active) {
    atomicMin(red_slot(base, t), v);
  }
}

// ---- import / keys / permutation --------------------------------------------

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_import(
    DevParams P, uint32_t n, const double* __restrict__ r, const double* __restrict__ v,
    const double* __restrict__ rho, DevState st, uint32_t* __restrict__ id) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double vi[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    st.x[k][i] = r[static_cast<size_t>(i) * DIM + k];
    vi[k] = v ? v[static_cast<size_t>(i) * DIM + k] : 0.0;
    st.v[k][i] = vi[k];
  }
  st.rho[i] = rho[i];
  write_pair_data<DIM>(P, st, i, vi, rho[i], particle_mass(P, i));
  id[i] = i;
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_cell_keys(
    DevParams P, const DevTank* __restrict__ tanks, uint32_t n, DevState st,
    const uint32_t* __restrict__ tag, uint32_t* __restrict__ key,
    uint32_t* __restrict__ flags, const uint32_t* __restrict__ ctl,
    const uint32_t* __restrict__ need) {
  if (!ctl[0]) return;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i == 0) {
    flags[-1] = 0u;  // fixed-particle counter, filled by k_permute
    flags[1] = 0u;   // listed fixed particles with fluid near, filled by k_fixed_active
  }
  if (i >= n) return;
  const uint32_t t = tag[i] >> 1;
  if (!need[t]) return;  // the tank keeps its keys, hence its order
  const DevTank& T = tanks[t];
  uint32_t flat = 0;
  bool bad = false;
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    const double x = st.x[k][i];
    const double f = (x - T.lo[k]) * T.inv_c[k];
    int c;
    if (!(f == f)) {
      bad = true;
      c = 0;
    } else if (f <= 0.0) {
      c = 0;
    } else if (f >= static_cast<double>(T.nc[k] - 1)) {
      c = T.nc[k] - 1;
    } else {
      c = static_cast<int>(f);
    }
    flat = flat * static_cast<uint32_t>(T.nc[k]) + static_cast<uint32_t>(c);
  }
  key[i] = T.cell_off + flat;
  if (bad) atomicOr(flags, 2u);
}

// Gathers the state into cell order and lists the fixed particles (the list
// order is irrelevant: each fixed particle is extrapolated on its own).
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_permute(
    DevParams P, uint32_t n, const uint32_t* __restrict__ perm, DevState src, DevState dst,
    const uint32_t* __restrict__ id_src, uint32_t* __restrict__ id_dst,
    const uint32_t* __restrict__ tag_src, uint32_t* __restrict__ tag_dst,
    uint32_t* __restrict__ fixed_list, uint32_t* __restrict__ fixed_count,
    const uint32_t* __restrict__ ctl) {
  if (!ctl[0]) return;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool in = i < n;
  const uint32_t tg = in ? tag_src[perm[i]] : 0u;
  const uint64_t m = __ballot(in && (tg & 1u));
  if (m) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll(static_cast<long long>(m)) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(fixed_count, static_cast<uint32_t>(__popcll(m)));
    base = __shfl(base, leader);
    if (in && (tg & 1u)) fixed_list[base + __popcll(m & ((1ull << lane) - 1ull))] = i;
  }
  if (!in) return;
  const uint32_t s = perm[i];
  if (s == i) return;  // stays in its slot (its tank was not re-sorted, or it kept its place)
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    dst.x[k][i] = src.x[k][s];
    dst.v[k][i] = src.v[k][s];
  }
  const double rho = src.rho[s];
  dst.rho[i] = rho;
  double v[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) v[k] = dst.v[k][i];
  const uint32_t pid = id_src[s];
  write_pair_data<DIM>(P, dst, i, v, rho, particle_mass(P, pid));
  id_dst[i] = pid;
  tag_dst[i] = tg;
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_restore_order(
    uint32_t n, const uint32_t* __restrict__ perm, DevState src, DevState dst,
    const uint32_t* __restrict__ id_src, uint32_t* __restrict__ id_dst,
    const uint32_t* __restrict__ tag_src, uint32_t* __restrict__ tag_dst,
    const uint32_t* __restrict__ ctl) {
  if (!ctl[0]) return;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || perm[i] == i) return;  // unmoved: dst already holds it
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    dst.x[k][i] = src.x[k][i];
    dst.v[k][i] = src.v[k][i];
  }
  dst.rho[i] = src.rho[i];
  dst.aux[i] = src.aux[i];
  dst.pv[i] = src.pv[i];
  if (dst.vol) dst.vol[i] = src.vol[i];
  id_dst[i] = id_src[i];
  tag_dst[i] = tag_src[i];
}

// CSR offsets: off[c] = first sorted position with key >= c.  Thread i
// (0..n) owns the cells c in (key[i-1], key[i]] (key[-1] = -1, key[n] =
// n_cells), so every cell is written exactly once: off[c] = i, and the cells
// before key[i] are marked as holding no fluid.  A run of up to kGapInline
// cells is filled by the thread; a longer run (empty space above or beside
// the water) is listed and filled by a whole wave in k_cell_gaps.  The first
// particle of cell key[i] marks whether the cell holds a fluid particle.  A
// fixed particle with no marked cell in its stencil has no fluid neighbour
// (fluid particles stay within skin/2 of their cell).
constexpr int64_t kGapInline = 32;

__global__ void k_cell_gaps_clear(uint32_t* n_gaps, const uint32_t* __restrict__ ctl) {
  if (ctl[0]) n_gaps[0] = 0u;
}

__global__ __launch_bounds__(kBlock) void k_cell_offsets(
    uint32_t n_cells, uint32_t n, const uint32_t* __restrict__ key,
    const uint32_t* __restrict__ tag, uint32_t* __restrict__ off,
    uint8_t* __restrict__ fluid, uint4* __restrict__ gaps, uint32_t* __restrict__ n_gaps,
    const uint32_t* __restrict__ ctl) {
  if (!ctl[0]) return;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i > n) return;
  const int64_t prev = i > 0 ? static_cast<int64_t>(key[i - 1]) : -1;
  const int64_t cur = i < n ? static_cast<int64_t>(key[i]) : static_cast<int64_t>(n_cells);
  if (cur - prev <= kGapInline) {
    for (int64_t c = prev + 1; c <= cur; ++c) {
      off[c] = i;
      if (c < cur) fluid[c] = 0;
    }
  } else {
    off[cur] = i;
    const uint32_t g = atomicAdd(n_gaps, 1u);
    gaps[g] = make_uint4(static_cast<uint32_t>(prev + 1), static_cast<uint32_t>(cur), i, 0u);
  }
  if (i < n && prev < cur) {
    uint8_t any = 0;
    for (uint32_t j = i; j < n && key[j] == static_cast<uint32_t>(cur); ++j) {
      any |= static_cast<uint8_t>(!(tag[j] & 1u));
    }
    fluid[cur] = any;
  }
}

// The listed runs [first, end) of empty cells: off = the owning thread's
// index, no fluid.  Persistent waves over the list.
__global__ __launch_bounds__(kBlock) void k_cell_gaps(uint32_t* __restrict__ off,
                                                      uint8_t* __restrict__ fluid,
                                                      const uint4* __restrict__ gaps,
                                                      const uint32_t* __restrict__ n_gaps,
                                                      const uint32_t* __restrict__ ctl) {
  if (!ctl[0]) return;
  const uint32_t ng = n_gaps[0];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t waves = gridDim.x * (kBlock / 64);
  for (uint32_t g = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); g < ng; g += waves) {
    const uint4 r = gaps[g];
    for (uint32_t c = r.x + lane; c < r.y; c += 64) {
      off[c] = r.z;
      fluid[c] = 0;
    }
  }
}

// ---- boundary extrapolation ---------------------------------------------------
// FluidEquations::setup_boundary (source/tit/sph/fluid_equations.hpp:122-164):
// Shepard average of H(rho_b) + g (r_be . n_e) n_e[1] over fluid neighbours,
// solved for rho_e; v_e = 0.  FP64 throughout.  A group of kBoundaryLanes
// lanes per fixed particle: the lanes split the stencil's candidates and the
// two sums are reduced with shuffles inside the group.

constexpr int kBoundaryLanes = 16;
// candidates whose loads a lane of k_boundary has in flight at once
constexpr int kBoundaryUnroll = 2;
// fixed particles per workgroup of k_fixed_active (its LDS list)
constexpr uint32_t kFixedPerBlock = 1024;

// Cell coordinates (per axis) of particle i's cell in its tank.
template <int DIM>
__device__ __forceinline__ void cell_coords(const BoundaryArgs& A, const DevTank& T, uint32_t i,
                                            int* c) {
  const uint32_t local = A.key[i] - T.cell_off;
  if constexpr (DIM == 2) {
    c[0] = static_cast<int>(local / T.nc[1]);
    c[1] = static_cast<int>(local % T.nc[1]);
  } else {
    const uint32_t q = local / T.nc[2];
    c[2] = static_cast<int>(local % T.nc[2]);
    c[1] = static_cast<int>(q % T.nc[1]);
    c[0] = static_cast<int>(q / T.nc[1]);
  }
}

// Whether a cell of the 5^d stencil around cell c held fluid at the last
// rebuild; the G lanes of a group split the cells.
template <int DIM, int G>
__device__ __forceinline__ bool stencil_has_fluid(const BoundaryArgs& A, const DevTank& T,
                                                  const int* c, int lane) {
  constexpr int NCOL = DIM == 2 ? 2 * kStencil + 1 : (2 * kStencil + 1) * (2 * kStencil + 1);
  constexpr int L = DIM - 1;
  const int zlo = max(c[L] - kStencil, 0), zhi = min(c[L] + kStencil, T.nc[L] - 1);
  const int nrow = zhi - zlo + 1;
  int any = 0;
  for (int t = lane; t < NCOL * nrow; t += G) {
    const int q = t / nrow, z = zlo + t % nrow;
    int ix, iy = 0;
    bool ok;
    if constexpr (DIM == 2) {
      ix = c[0] - kStencil + q;
      ok = ix >= 0 && ix < T.nc[0];
    } else {
      ix = c[0] - kStencil + q / (2 * kStencil + 1);
      iy = c[1] - kStencil + q % (2 * kStencil + 1);
      ok = ix >= 0 && ix < T.nc[0] && iy >= 0 && iy < T.nc[1];
    }
    if (ok) {
      const uint32_t row = DIM == 2
                               ? T.cell_off + static_cast<uint32_t>(ix) * T.nc[1]
                               : T.cell_off + (static_cast<uint32_t>(ix) * T.nc[1] +
                                               static_cast<uint32_t>(iy)) * T.nc[2];
      any |= A.cell_fluid[row + static_cast<uint32_t>(z)];
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) any |= __shfl_xor(any, off, G);
  return any != 0;
}

// The fixed particle's extrapolated state from the Shepard sums S and H.
template <int DIM>
__device__ __forceinline__ void write_fixed(const BoundaryArgs& A, uint32_t i, double S,
                                            double H) {
  const DevParams& P = A.P;
  const double tiny = 6.0554544523933395e-06;  // cbrt(DBL_EPSILON)
  const double Hs = fabs(S) <= tiny ? 0.0 : H / S;
  const double rho_e = eos_density_from_potential(P, Hs);
  A.s.rho[i] = rho_e;
  const double zero[DIM] = {};
  const double m_e = particle_mass(P, A.id[i]);
  write_pair_data<DIM>(P, A.s, i, zero, rho_e, m_e);
  if (A.wall_rho) {
    const uint32_t k = A.id[i] - A.wall_id0;
    A.wall_rho[k] = rho_e;
    A.s.pv[i].z = static_cast<float>(A.wall_mu[k] * m_e / rho_e);  // V_e = m_e / rho_e
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) A.s.v[d][i] = 0.0;
}

// After a cell rebuild: lists the fixed particles with fluid in their
// stencil (k_boundary extrapolates those, every stage), and writes the
// others' state once: with no fluid neighbour S = H = 0, the value a full
// scan gives, and it holds until the next rebuild because the stencil covers
// every neighbour until then.  The list's order is irrelevant (each fixed
// particle is extrapolated on its own).  A workgroup compacts its share in
// LDS and takes its slots with one atomic.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_fixed_active(BoundaryArgs A, const uint32_t* ctl,
                                                         uint32_t per_block) {
  if (!ctl[0]) return;
  constexpr int G = kBoundaryLanes;
  constexpr int kGroups = kBlock / G;
  __shared__ uint32_t s_list[kFixedPerBlock];
  __shared__ uint32_t s_n, s_base;
  const int lane = threadIdx.x % G;
  const uint32_t grp = threadIdx.x / G;
  const uint32_t beg = blockIdx.x * per_block;
  const uint32_t end = min(beg + per_block, A.n_fixed);
  if (threadIdx.x == 0) s_n = 0u;
  __syncthreads();
  for (uint32_t base = beg; base < end; base += kGroups) {
    const uint32_t k = base + grp;
    if (k < end) {  // uniform in the group
      const uint32_t i = A.fixed_list[k];
      const DevTank T = A.tanks[A.tag[i] >> 1];
      int c[3];
      cell_coords<DIM>(A, T, i, c);
      const bool any = stencil_has_fluid<DIM, G>(A, T, c, lane);
      if (lane == 0) {
        if (any) {
          s_list[atomicAdd(&s_n, 1u)] = i;
        } else {
          write_fixed<DIM>(A, i, 0.0, 0.0);
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) s_base = s_n ? atomicAdd(A.n_active, s_n) : 0u;
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < s_n; t += kBlock) A.active[s_base + t] = s_list[t];
}

// Extrapolation of the listed fixed particles, one group of G lanes each.
// XI_INT: the Tait exponent is a small integer (P.xi_int), so H(rho) takes
// integer powers; the general FP64 pow otherwise (it alone needs about 55
// more VGPRs, which would cost a wave per SIMD).
template <int DIM, int KERNEL, bool XI_INT>
__global__ __launch_bounds__(kBlock) void k_boundary(BoundaryArgs A) {
  constexpr int NCOL = DIM == 2 ? 2 * kStencil + 1 : (2 * kStencil + 1) * (2 * kStencil + 1);
  constexpr int G = kBoundaryLanes;
  const uint32_t k = blockIdx.x * (kBlock / G) + (threadIdx.x / G);
  const int lane = threadIdx.x % G;
  if (k >= *A.n_active) return;  // uniform in the group
  const DevParams& P = A.P;
  const uint32_t i = A.active[k];
  const DevTank T = A.tanks[A.tag[i] >> 1];
  double xe[DIM], ne[DIM];
  const double* nrm = A.normal + static_cast<size_t>(A.id[i]) * DIM;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    xe[d] = A.s.x[d][i];
    ne[d] = nrm[d];
  }
  double n2 = ne[0] * ne[0];
#pragma unroll
  for (int d = 1; d < DIM; ++d) n2 = n2 + ne[d] * ne[d];
  const double tiny = 6.0554544523933395e-06;  // cbrt(DBL_EPSILON)
  if (n2 >= tiny * tiny) {
    const double s = sqrt(n2);
#pragma unroll
    for (int d = 0; d < DIM; ++d) ne[d] = ne[d] / s;
  } else {
#pragma unroll
    for (int d = 0; d < DIM; ++d) ne[d] = 0.0;
  }
  const double omega = KERNEL == kKernelC2
                           ? (DIM == 2 ? 7.0 / (4.0 * M_PI) : 21.0 / (16.0 * M_PI))
                           : (DIM == 2 ? 9.0 / (4.0 * M_PI) : 495.0 / (256.0 * M_PI));
  const double h_inv = 1.0 / P.h;
  double w0 = omega;
#pragma unroll
  for (int d = 0; d < DIM; ++d) w0 *= h_inv;

  // Stencil column ranges (wave-uniform).
  int c[3];
  cell_coords<DIM>(A, T, i, c);
  constexpr int L = DIM - 1;
  const int zlo = max(c[L] - kStencil, 0), zhi = min(c[L] + kStencil, T.nc[L] - 1);
  // Stencil trimming (interior windows only): a fluid particle lies within
  // skin/2 of its cell's box, so one in a column whose box (grown by skin/2)
  // is g away in the column plane has |dz| <= sqrt((2h)^2 - g^2); the rows
  // beyond cannot hold a neighbour (FP64, bounds widened by 1e-9).
  bool inner = c[0] - kStencil > 0 && c[0] + kStencil < T.nc[0] - 1 && c[L] - kStencil > 0 &&
               c[L] + kStencil < T.nc[L] - 1;
  if constexpr (DIM == 3) inner = inner && c[1] - kStencil > 0 && c[1] + kStencil < T.nc[1] - 1;
  auto gap = [&](int ax, int ic) {
    const double cw = 1.0 / T.inv_c[ax];
    const double a0 = T.lo[ax] + ic * cw - P.skin_half;
    const double a1 = T.lo[ax] + (ic + 1) * cw + P.skin_half;
    return fmax(0.0, fmax(a0 - xe[ax], xe[ax] - a1)) * (1.0 - 1e-9);
  };
  // rows [w0, w1] of a column at squared in-plane gap g2; false: none
  auto rows = [&](double g2, int& w0, int& w1) {
    w0 = zlo;
    w1 = zhi;
    if (!inner) return true;
    if (g2 > P.R2 * (1.0 + 1e-9)) return false;
    const double hz = (sqrt(fmax(P.R2 * (1.0 + 1e-9) - g2, 0.0)) + P.skin_half) *
                          T.inv_c[L] * (1.0 + 1e-9) + 1e-9;
    const double zr = (xe[L] - T.lo[L]) * T.inv_c[L];
    w0 = max(w0, static_cast<int>(floor(zr - hz)));
    w1 = min(w1, static_cast<int>(floor(zr + hz)));
    return w0 <= w1;
  };
  // 2D: the five column ranges and their prefix
  uint32_t rb[DIM == 2 ? NCOL : 1], rl[DIM == 2 ? NCOL + 1 : 1];
  rl[0] = 0;
  if constexpr (DIM == 2) {
#pragma unroll
    for (int q = 0; q < NCOL; ++q) {
      const int ix = c[0] - kStencil + q;
      uint32_t b = 0, e = 0;
      int w0, w1;
      if (ix >= 0 && ix < T.nc[0]) {
        const double gx = gap(0, ix);
        if (rows(gx * gx, w0, w1)) {
          const uint32_t row = T.cell_off + static_cast<uint32_t>(ix) * T.nc[1];
          b = A.cell_off[row + w0];
          e = A.cell_off[row + w1 + 1];
        }
      }
      rb[q] = b;
      rl[q + 1] = rl[q] + (e - b);
    }
  }
  double S = 0.0, H = 0.0;
  const double xi_1 = P.xi - 1.0, inv_xi_1 = 1.0 / xi_1, cs2 = P.cs0 * P.cs0;
  // U candidates per lane at a time.  3D: three rounds of loads that each
  // have all U candidates in flight (the tags; the positions of the fluid
  // ones; ρ and the id, for per-particle masses, of those within the radius);
  // 2D: one round of all of them.  The terms are then added in candidate
  // order.
  constexpr int U = kBoundaryUnroll;
  auto batch = [&](const uint32_t (&j)[U], const bool (&live)[U]) {
    bool ok[U];
    double dd[U][DIM], s2[U], rho[U], mb[U];
    if constexpr (DIM == 2) {
      // 2D (about 100 candidates): every field of a live candidate in one
      // round (measured faster here; 3D is bound by the load traffic)
      uint32_t tg[U], id[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!live[u]) continue;
        tg[u] = A.tag[j[u]];
        id[u] = P.mass ? A.id[j[u]] : 0u;
#pragma unroll
        for (int d = 0; d < DIM; ++d) dd[u][d] = A.s.x[d][j[u]] - xe[d];
        rho[u] = A.s.rho[j[u]];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        ok[u] = live[u] && !(tg[u] & 1u);  // fluid only
        if (!ok[u]) continue;
        s2[u] = dd[u][0] * dd[u][0];
#pragma unroll
        for (int d = 1; d < DIM; ++d) s2[u] = s2[u] + dd[u][d] * dd[u][d];
        ok[u] = s2[u] <= P.R2;
        if (ok[u]) mb[u] = P.mass ? P.mass[id[u]] : P.m;
      }
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) ok[u] = live[u] && !(A.tag[j[u]] & 1u);  // fluid only
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!ok[u]) continue;
#pragma unroll
        for (int d = 0; d < DIM; ++d) dd[u][d] = A.s.x[d][j[u]] - xe[d];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!ok[u]) continue;
        s2[u] = dd[u][0] * dd[u][0];
#pragma unroll
        for (int d = 1; d < DIM; ++d) s2[u] = s2[u] + dd[u][d] * dd[u][d];
        ok[u] = s2[u] <= P.R2;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!ok[u]) continue;
        rho[u] = A.s.rho[j[u]];
        mb[u] = P.mass ? P.mass[A.id[j[u]]] : P.m;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) continue;
      const double rho_b = rho[u];
      const double V_b = mb[u] / rho_b;
      const double W = w0 * kernel_value_poly_f64<KERNEL>(h_inv * sqrt(s2[u]));
      double rn = dd[u][0] * ne[0];
#pragma unroll
      for (int d = 1; d < DIM; ++d) rn = rn + dd[u][d] * ne[d];
      // H(rho_b) = cs0^2 ((rho_b/rho0)^(xi-1) - 1) / (xi - 1)
      const double ratio = rho_b * P.inv_rho0;
      const double rp = XI_INT ? ipow_f64(ratio, P.xi_int - 1) : pow_general(ratio, xi_1);
      const double H_b = cs2 * (rp - 1.0) * inv_xi_1;
      S += V_b * W;
      H += V_b * (H_b + P.g * rn * ne[1]) * W;
    }
  };
  if constexpr (DIM == 2) {
    // the lanes take every G-th candidate of the five column ranges in turn
    const uint32_t n_cand = rl[NCOL];
    auto cand = [&](uint32_t t) {
      int q = 0;
      while (t >= rl[q + 1]) ++q;
      return rb[q] + (t - rl[q]);
    };
    for (uint32_t t = lane; t < n_cand; t += U * G) {
      uint32_t j[U];
      bool live[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        live[u] = t + u * G < n_cand;
        j[u] = live[u] ? cand(t + u * G) : 0u;
      }
      batch(j, live);
    }
  } else {
    // The 25 trimmed column ranges, laid end to end: each lane loads two
    // columns' bounds, the group scans their lengths, and the lanes then
    // take every G-th candidate of the whole list (no round trip per
    // column).
    __shared__ uint32_t s_rb[kBlock / G][NCOL];
    __shared__ uint32_t s_rl[kBlock / G][NCOL + 1];
    const int grp = threadIdx.x / G;
    uint32_t beg[2] = {0u, 0u}, len[2] = {0u, 0u};
#pragma unroll
    for (int hq = 0; hq < 2; ++hq) {
      const int q = lane + hq * G;
      if (q >= NCOL) continue;
      const int ix = c[0] - kStencil + q / (2 * kStencil + 1);
      const int iy = c[1] - kStencil + q % (2 * kStencil + 1);
      if (ix < 0 || ix >= T.nc[0] || iy < 0 || iy >= T.nc[1]) continue;
      const double gx = gap(0, ix), gy = gap(1, iy);
      int w0, w1;
      if (!rows(gx * gx + gy * gy, w0, w1)) continue;
      const uint32_t row =
          T.cell_off + (static_cast<uint32_t>(ix) * T.nc[1] + static_cast<uint32_t>(iy)) *
                           T.nc[2];
      beg[hq] = A.cell_off[row + w0];
      len[hq] = A.cell_off[row + w1 + 1] - beg[hq];
    }
    // columns 0..15 on the lanes' first slot, 16..24 on the second
    uint32_t x0 = len[0], x1 = len[1];
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
      const uint32_t y0 = __shfl_up(x0, off, G), y1 = __shfl_up(x1, off, G);
      if (lane >= off) {
        x0 += y0;
        x1 += y1;
      }
    }
    const uint32_t tot0 = __shfl(x0, G - 1, G), tot1 = __shfl(x1, G - 1, G);
    s_rb[grp][lane] = beg[0];
    s_rl[grp][lane] = x0 - len[0];
    if (lane + G < NCOL) {
      s_rb[grp][lane + G] = beg[1];
      s_rl[grp][lane + G] = tot0 + x1 - len[1];
    }
    if (lane == 0) s_rl[grp][NCOL] = tot0 + tot1;
    // the group's lanes read what the others wrote: wavefront-scope ordering
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t n_cand = tot0 + tot1;
    int q = 0;
    auto cand = [&](uint32_t t) {  // candidates come in increasing t
      while (t >= s_rl[grp][q + 1]) ++q;
      return s_rb[grp][q] + (t - s_rl[grp][q]);
    };
    for (uint32_t t = lane; t < n_cand; t += U * G) {
      uint32_t j[U];
      bool live[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        live[u] = t + u * G < n_cand;
        j[u] = live[u] ? cand(t + u * G) : 0u;
      }
      batch(j, live);
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    S += __shfl_xor(S, off);
    H += __shfl_xor(H, off);
  }
  if (lane == 0) write_fixed<DIM>(A, i, S, H);
}

// Time step of the current state with a given dv/dt (reference order) or
// with dv/dt = 0.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_time_step(
    DevParams P, uint32_t n, DevState st, const uint32_t* __restrict__ tag,
    const uint32_t* __restrict__ id, const double* __restrict__ dv_ref,
    unsigned long long* __restrict__ dt_min) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool in_range = i < n;
  const uint32_t tg = in_range ? tag[i] : 1u;
  const bool fluid = in_range && !(tg & 1u);
  double dt = 0.0;
  if (fluid) {
    double v[DIM], dv[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      v[k] = st.v[k][i];
      dv[k] = dv_ref ? dv_ref[static_cast<size_t>(id[i]) * DIM + k] : 0.0;
    }
    dt = particle_time_step<DIM>(P, st.rho[i], v, dv);
  }
  wave_atomic_min_u64(dt_min, tg >> 1, dbl_bits(dt), fluid);
}

__global__ void k_rebuild_clear(uint32_t* ctl) { ctl[0] = 0u; }

// One wave per tank (the reduction slots' copies folded lane by lane).
__global__ __launch_bounds__(64) void k_step_begin(unsigned long long* dt_min, double* dt,
                                                   uint32_t* vmax, const double* disp,
                                                   double skin_half, double span, uint32_t* ctl,
                                                   uint32_t* need) {
  const uint32_t t = blockIdx.x;
  const unsigned long long b = red_fold<false>(dt_min, t, true);
  const uint32_t vb = red_fold<true>(vmax, t, true);
  if (threadIdx.x != 0) return;
  // A tank without fluid keeps dt = 0 (nothing moves).
  const double h = b == ~0ull ? 0.0 : __longlong_as_double(static_cast<long long>(b));
  dt[t] = h;
  const double vprev = static_cast<double>(__uint_as_float(vb));
  const bool nd = ctl[1] != 0u || !(disp[t] + 2.0 * span * h * vprev <= 0.9 * skin_half);
  need[t] = nd ? 1u : 0u;
  if (nd) atomicOr(ctl, 1u);
}

__global__ void k_force_rebuild(int n_tanks, uint32_t* ctl, uint32_t* need) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) ctl[0] = 1u;
  if (t < n_tanks) need[t] = 1u;
}

// Rebuild only if a forced one is pending (the current index is otherwise
// valid: every particle is within half the skin of its cell).
__global__ void k_pending_rebuild(int n_tanks, uint32_t* ctl, uint32_t* need) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t f = ctl[1];
  if (t < n_tanks) need[t] = f;
  if (t == 0) ctl[0] = f;
}

// One wave per tank.
__global__ __launch_bounds__(64) void k_step_end(const double* dt, uint32_t* vmax, double* time,
                                                 double* disp, double skin_half, uint32_t* flags,
                                                 double span) {
  const uint32_t t = blockIdx.x;
  const uint32_t vb = red_fold<true>(vmax, t, false);
  if (threadIdx.x != 0) return;
  const double T = span * dt[t];
  time[t] += T;
  const double d = disp[t] + T * static_cast<double>(__uint_as_float(vb));
  disp[t] = d;
  if (!(d <= skin_half)) atomicOr(flags, 1u);
}

__global__ void k_reset_disp(int n_tanks, double* disp, uint32_t* ctl,
                             const uint32_t* need, unsigned long long* count) {
  if (!ctl[0]) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    ctl[1] = 0u;  // no forced rebuild pending
    ctl[2] = 1u;  // the cell index is valid
    *count += 1ull;
  }
  if (t < n_tanks && need[t]) disp[t] = 0.0;
}

// ---- export ---------------------------------------------------------------------

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_export(
    DevParams P, uint32_t n, DevState st, const uint32_t* __restrict__ id,
    double* __restrict__ r, double* __restrict__ v, double* __restrict__ rho,
    double* __restrict__ p) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const size_t o = id[i];
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    r[o * DIM + k] = st.x[k][i];
    v[o * DIM + k] = st.v[k][i];
  }
  rho[o] = st.rho[i];
  p[o] = eos_pressure(P, st.rho[i]);
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_export_rhs(
    uint32_t n, const uint32_t* __restrict__ id, const double* __restrict__ drho,
    const double* __restrict__ dv0, const double* __restrict__ dv1,
    const double* __restrict__ dv2, double* __restrict__ drho_ref,
    double* __restrict__ dv_ref) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const size_t o = id[i];
  drho_ref[o] = drho[i];
  dv_ref[o * DIM + 0] = dv0[i];
  dv_ref[o * DIM + 1] = dv1[i];
  if constexpr (DIM == 3) dv_ref[o * DIM + 2] = dv2[i];
}

// ---- neighbour lists (test hook) ----------------------------------------------

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_neighbor_count(
    DevParams P, const DevTank* __restrict__ tanks, uint32_t n, DevState st,
    const uint32_t* __restrict__ key, const uint32_t* __restrict__ tag,
    const uint32_t* __restrict__ id, const uint32_t* __restrict__ coff,
    uint64_t* __restrict__ count_ref) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const DevTank T = tanks[tag[i] >> 1];
  double xa[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) xa[k] = st.x[k][i];
  uint64_t c = 1;  // self
  for_each_neighbor<DIM>(P, T, st, coff, key[i], i, xa,
                         [&](uint32_t, const double*, double) { ++c; });
  count_ref[id[i]] = c;
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_neighbor_fill(
    DevParams P, const DevTank* __restrict__ tanks, uint32_t n, DevState st,
    const uint32_t* __restrict__ key, const uint32_t* __restrict__ tag,
    const uint32_t* __restrict__ id, const uint32_t* __restrict__ coff,
    const uint64_t* __restrict__ off_ref, uint32_t* __restrict__ idx_ref) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const DevTank T = tanks[tag[i] >> 1];
  double xa[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) xa[k] = st.x[k][i];
  const uint32_t self = id[i];
  uint32_t* out = idx_ref + off_ref[self];
  uint32_t c = 0;
  out[c++] = self;
  for_each_neighbor<DIM>(P, T, st, coff, key[i], i, xa,
                         [&](uint32_t j, const double*, double) { out[c++] = id[j]; });
  // ascending order (particle_mesh.hpp:146)
  for (uint32_t a = 1; a < c; ++a) {
    const uint32_t v = out[a];
    uint32_t b = a;
    while (b > 0 && out[b - 1] > v) {
      out[b] = out[b - 1];
      --b;
    }
    out[b] = v;
  }
}

// ---- dispatch helpers --------------------------------------------------------------

}  // namespace

void launch_import(hipStream_t s, const DevParams& P, uint32_t n, const double* r,
                   const double* v, const double* rho, DevState st,
                   uint32_t* id) {
  if (n == 0) return;
  if (P.dim == 2) {
    k_import<2><<<grid_for(n), kBlock, 0, s>>>(P, n, r, v, rho, st, id);
  } else {
    k_import<3><<<grid_for(n), kBlock, 0, s>>>(P, n, r, v, rho, st, id);
  }
}

void launch_cell_keys(hipStream_t s, const DevParams& P, const DevTank* tanks,
                      uint32_t n, DevState st, const uint32_t* tag,
                      uint32_t* key, uint32_t* flags, RebuildCtl rc) {
  if (n == 0) return;
  if (P.dim == 2) {
    k_cell_keys<2><<<grid_for(n), kBlock, 0, s>>>(P, tanks, n, st, tag, key, flags, rc.ctl,
                                                  rc.need);
  } else {
    k_cell_keys<3><<<grid_for(n), kBlock, 0, s>>>(P, tanks, n, st, tag, key, flags, rc.ctl,
                                                  rc.need);
  }
}

void launch_permute(hipStream_t s, const DevParams& P, uint32_t n, const uint32_t* perm,
                    DevState src, DevState dst, const uint32_t* id_src,
                    uint32_t* id_dst, const uint32_t* tag_src,
                    uint32_t* tag_dst, uint32_t* fixed_list, uint32_t* fixed_count,
                    const uint32_t* ctl) {
  if (n == 0) return;
  if (P.dim == 2) {
    k_permute<2><<<grid_for(n), kBlock, 0, s>>>(P, n, perm, src, dst, id_src,
                                                id_dst, tag_src, tag_dst, fixed_list,
                                                fixed_count, ctl);
  } else {
    k_permute<3><<<grid_for(n), kBlock, 0, s>>>(P, n, perm, src, dst, id_src,
                                                id_dst, tag_src, tag_dst, fixed_list,
                                                fixed_count, ctl);
  }
}

void launch_restore_order(hipStream_t s, int dim, uint32_t n, const uint32_t* perm,
                          DevState src, DevState dst, const uint32_t* id_src,
                          uint32_t* id_dst, const uint32_t* tag_src, uint32_t* tag_dst,
                          const uint32_t* ctl) {
  if (n == 0) return;
  if (dim == 2) {
    k_restore_order<2><<<grid_for(n), kBlock, 0, s>>>(n, perm, src, dst, id_src, id_dst,
                                                      tag_src, tag_dst, ctl);
  } else {
    k_restore_order<3><<<grid_for(n), kBlock, 0, s>>>(n, perm, src, dst, id_src, id_dst,
                                                      tag_src, tag_dst, ctl);
  }
}

void launch_cell_offsets(hipStream_t s, uint32_t n_cells, uint32_t n,
                         const uint32_t* key, const uint32_t* tag, uint32_t* cell_off,
                         uint8_t* fluid, uint4* gaps, uint32_t* n_gaps, const uint32_t* ctl) {
  k_cell_gaps_clear<<<1, 1, 0, s>>>(n_gaps, ctl);
  k_cell_offsets<<<grid_for(n + 1), kBlock, 0, s>>>(n_cells, n, key, tag, cell_off, fluid,
                                                     gaps, n_gaps, ctl);
  k_cell_gaps<<<1024, kBlock, 0, s>>>(cell_off, fluid, gaps, n_gaps, ctl);
}

void launch_fixed_active(hipStream_t s, const BoundaryArgs& a, const uint32_t* ctl) {
  if (a.n_fixed == 0) return;
  // at most kFixedPerBlock per workgroup; at least 16 (one round) when there
  // are few, so that the list is built by many workgroups
  const uint32_t grid = std::max((a.n_fixed + kFixedPerBlock - 1) / kFixedPerBlock,
                                 std::min<uint32_t>(1024u, (a.n_fixed + 15) / 16));
  const uint32_t per_block = (a.n_fixed + grid - 1) / grid;
  if (a.P.dim == 2) {
    k_fixed_active<2><<<grid, kBlock, 0, s>>>(a, ctl, per_block);
  } else {
    k_fixed_active<3><<<grid, kBlock, 0, s>>>(a, ctl, per_block);
  }
}

void launch_boundary(hipStream_t s, const BoundaryArgs& a) {
  if (a.n_fixed == 0) return;
  const uint32_t g = grid_for(a.n_fixed, kBlock / kBoundaryLanes);
  const bool c4 = a.P.kernel == kKernelC4, xi_int = a.P.xi_int != 0;
  if (a.P.dim == 2) {
    if (c4) {
      if (xi_int) k_boundary<2, kKernelC4, true><<<g, kBlock, 0, s>>>(a);
      else k_boundary<2, kKernelC4, false><<<g, kBlock, 0, s>>>(a);
    } else {
      if (xi_int) k_boundary<2, kKernelC2, true><<<g, kBlock, 0, s>>>(a);
      else k_boundary<2, kKernelC2, false><<<g, kBlock, 0, s>>>(a);
    }
  } else {
    if (c4) {
      if (xi_int) k_boundary<3, kKernelC4, true><<<g, kBlock, 0, s>>>(a);
      else k_boundary<3, kKernelC4, false><<<g, kBlock, 0, s>>>(a);
    } else {
      if (xi_int) k_boundary<3, kKernelC2, true><<<g, kBlock, 0, s>>>(a);
      else k_boundary<3, kKernelC2, false><<<g, kBlock, 0, s>>>(a);
    }
  }
}

void launch_time_step(hipStream_t s, const DevParams& P, uint32_t n,
                      DevState st, const uint32_t* tag, const uint32_t* id,
                      const double* dv_ref, unsigned long long* dt_min) {
  if (n == 0) return;
  if (P.dim == 2) {
    k_time_step<2><<<grid_for(n), kBlock, 0, s>>>(P, n, st, tag, id, dv_ref, dt_min);
  } else {
    k_time_step<3><<<grid_for(n), kBlock, 0, s>>>(P, n, st, tag, id, dv_ref, dt_min);
  }
}

void launch_step_begin(hipStream_t s, int n_tanks, unsigned long long* dt_min,
                       double* dt, uint32_t* vmax, const double* disp,
                       double skin_half, RebuildCtl rc, double span) {
  k_rebuild_clear<<<1, 1, 0, s>>>(rc.ctl);
  k_step_begin<<<n_tanks, 64, 0, s>>>(dt_min, dt, vmax, disp, skin_half, span, rc.ctl,
                                      rc.need);
}

void launch_force_rebuild(hipStream_t s, int n_tanks, RebuildCtl rc) {
  k_force_rebuild<<<grid_for(n_tanks, 64), 64, 0, s>>>(n_tanks, rc.ctl, rc.need);
}

__global__ void k_mark_pending(uint32_t* ctl) { ctl[1] = 1u; }

void launch_mark_pending(hipStream_t s, RebuildCtl rc) { k_mark_pending<<<1, 1, 0, s>>>(rc.ctl); }

void launch_pending_rebuild(hipStream_t s, int n_tanks, RebuildCtl rc) {
  k_pending_rebuild<<<grid_for(n_tanks, 64), 64, 0, s>>>(n_tanks, rc.ctl, rc.need);
}

void launch_step_end(hipStream_t s, int n_tanks, const double* dt,
                     uint32_t* vmax, double* time, double* disp,
                     double skin_half, uint32_t* flags, double span) {
  k_step_end<<<n_tanks, 64, 0, s>>>(dt, vmax, time, disp, skin_half, flags, span);
}

void launch_reset_disp(hipStream_t s, int n_tanks, double* disp, RebuildCtl rc) {
  k_reset_disp<<<grid_for(n_tanks, 64), 64, 0, s>>>(n_tanks, disp, rc.ctl, rc.need,
                                                     rc.count);
}

void launch_export(hipStream_t s, const DevParams& P, uint32_t n, DevState st,
                   const uint32_t* id, double* r, double* v, double* rho,
                   double* p) {
  if (n == 0) return;
  if (P.dim == 2) {
    k_export<2><<<grid_for(n), kBlock, 0, s>>>(P, n, st, id, r, v, rho, p);
  } else {
    k_export<3><<<grid_for(n), kBlock, 0, s>>>(P, n, st, id, r, v, rho, p);
  }
}

__global__ void k_export_u32(uint32_t n, const uint32_t* __restrict__ id,
                             const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[id[i]] = src[i];
}

void launch_export_u32(hipStream_t s, uint32_t n, const uint32_t* id,
                       const uint32_t* src, uint32_t* dst) {
  if (n) k_export_u32<<<grid_for(n), kBlock, 0, s>>>(n, id, src, dst);
}

void launch_export_rhs(hipStream_t s, int dim, uint32_t n, const uint32_t* id,
                       const double* drho, double* const* dv, double* drho_ref,
                       double* dv_ref) {
  if (n == 0) return;
  if (dim == 2) {
    k_export_rhs<2><<<grid_for(n), kBlock, 0, s>>>(n, id, drho, dv[0], dv[1],
                                                   nullptr, drho_ref, dv_ref);
  } else {
    k_export_rhs<3><<<grid_for(n), kBlock, 0, s>>>(n, id, drho, dv[0], dv[1],
                                                   dv[2], drho_ref, dv_ref);
  }
}

void launch_neighbor_count(hipStream_t s, const DevParams& P,
                           const DevTank* tanks, uint32_t n, DevState st,
                           const uint32_t* key, const uint32_t* tag,
                           const uint32_t* id, const uint32_t* cell_off,
                           uint64_t* count_ref) {
  if (n == 0) return;
  if (P.dim == 2) {
    k_neighbor_count<2><<<grid_for(n), kBlock, 0, s>>>(P, tanks, n, st, key, tag,
                                                       id, cell_off, count_ref);
  } else {
    k_neighbor_count<3><<<grid_for(n), kBlock, 0, s>>>(P, tanks, n, st, key, tag,
                                                       id, cell_off, count_ref);
  }
}

void launch_neighbor_fill(hipStream_t s, const DevParams& P,
                          const DevTank* tanks, uint32_t n, DevState st,
                          const uint32_t* key, const uint32_t* tag,
                          const uint32_t* id, const uint32_t* cell_off,
                          const uint64_t* off_ref, uint32_t* idx_ref) {
  if (n == 0) return;
  if (P.dim == 2) {
    k_neighbor_fill<2><<<grid_for(n), kBlock, 0, s>>>(P, tanks, n, st, key, tag,
                                                      id, cell_off, off_ref, idx_ref);
  } else {
    k_neighbor_fill<3><<<grid_for(n), kBlock, 0, s>>>(P, tanks, n, st, key, tag,
                                                      id, cell_off, off_ref, idx_ref);
  }
}

}  // namespace tit_amd
