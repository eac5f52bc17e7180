// Post-integration step of FluidEquations (fluid_equations.hpp:313-512):
// particle shifting with renormalised gradients (apply_shifts, :327-471) and
// the free-surface density correction (apply_free_surface_correction,
// :480-512).  Without walls gamma = 1 and there are no face terms; with the
// semi-analytic wall the sums start from the face terms (computed by the wall
// kernel, 16 lanes per particle), the pair terms take per-particle masses
// over gamma_a, the velocity correction applies where gamma = 1, and the
// free-surface correction compares the Shepard sum with gamma.
//
// Everything runs in FP64, as the reference does, one particle per thread;
// the neighbour sets are those of the post-step positions (the exact
// predicate over the 5^d cell stencil, the index covering them by the skin
// bound of the step).  Passes:
//   volume   V = m / rho per particle; in 2D also packed double2 copies of
//            (x, y), (v_x, v_y) and (rho, V), which the walks below read
//   sums     N = sum V_b grad W_ab, L = sum V_b r_ba (x) grad W_ab, grad v,
//            grad rho (every particle); renormalise by (L^T)^-1; N normalised
//   phi      free-surface indicator: a fluid particle is "far" (1) when a
//            neighbour lies within 45 degrees of -N behind it, splashes
//            (<= 8 / 26 neighbours) are on the surface (DBL_MIN)
//   shift    far particles next to a surface particle get phi = |N_b . r_ab|
//            / 2h; particles still far are shifted by -0.08 h^2 N_raw with
//            v and rho corrected along the shift (into B)
//   fsc      surface and near-surface particles: Shepard-blended density on
//            the same neighbour sets at the shifted positions (into C)
//   commit   A = (B.r, B.v, C.rho), pair records, the next step size from
//            the last stage's acceleration, and the shift into the
//            displacement bound of the cell index.
#include <algorithm>
#include <cfloat>

#include "pair_math.hpp"
#include "sph_kernels.hpp"

namespace tit_amd {

namespace {

constexpr int kPiBlock = 256;
constexpr double kTiny = 6.0554544523933395e-06;  // cbrt(DBL_EPSILON)
constexpr double kPhiMax = 1.0;

__device__ __forceinline__ bool same_bits(double a, double b) {
  return __double_as_longlong(a) == __double_as_longlong(b);
}

// Per-tank atomic min / max of a wave's values (every lane of the wave must
// call it; `part` marks the lanes with a value).  A tank's particles are
// contiguous, so the participating lanes of a wave nearly always share one
// tank: then the wave reduces first and one lane updates the tank's slot.
template <bool MAX, class T>
__device__ __forceinline__ void tank_atomic(bool part, uint32_t t, T v, T* slot) {
  const unsigned long long m = __ballot(part);
  if (m == 0) return;
  const int first = __ffsll(static_cast<long long>(m)) - 1;
  const uint32_t t0 = __shfl(t, first);
  if (__all(!part || t == t0)) {
    T r = part ? v : (MAX ? T(0) : ~T(0));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const T o = __shfl_xor(r, off);
      r = MAX ? (o > r ? o : r) : (o < r ? o : r);
    }
    if ((threadIdx.x & 63) == static_cast<unsigned>(first)) {
      if constexpr (MAX) {
        atomicMax(red_slot(slot, t0), r);
      } else {
        atomicMin(red_slot(slot, t0), r);
      }
    }
  } else if (part) {
    if constexpr (MAX) {
      atomicMax(red_slot(slot, t), v);
    } else {
      atomicMin(red_slot(slot, t), v);
    }
  }
}

// The 5^d cell window around particle i's cell, clamped to its tank: cells
// (ix, iy, z) for ix in [xlo, xhi], iy in [ylo, yhi] (3D; 0 in 2D), z in
// [zlo, zhi] along the last axis; row(ix, iy) is the global index of (ix, iy, 0).
struct Window {
  uint32_t cell_off, nc1, nc2;
  int xlo, xhi, ylo, yhi, zlo, zhi;
};

template <int DIM>
__device__ __forceinline__ Window window_of(const PiArgs& A, uint32_t i) {
  const DevTank T = A.tanks[A.tag[i] >> 1];
  const uint32_t local = A.key[i] - T.cell_off;
  int c[3];
  if constexpr (DIM == 2) {
    c[0] = static_cast<int>(local / T.nc[1]);
    c[1] = static_cast<int>(local % T.nc[1]);
  } else {
    const uint32_t q = local / T.nc[2];
    c[2] = static_cast<int>(local % T.nc[2]);
    c[1] = static_cast<int>(q % T.nc[1]);
    c[0] = static_cast<int>(q / T.nc[1]);
  }
  constexpr int L = DIM - 1;
  Window w;
  w.cell_off = T.cell_off;
  w.nc1 = static_cast<uint32_t>(T.nc[1]);
  w.nc2 = DIM == 3 ? static_cast<uint32_t>(T.nc[2]) : 1u;
  w.xlo = max(c[0] - kStencil, 0);
  w.xhi = min(c[0] + kStencil, T.nc[0] - 1);
  w.ylo = DIM == 3 ? max(c[1] - kStencil, 0) : 0;
  w.yhi = DIM == 3 ? min(c[1] + kStencil, T.nc[1] - 1) : 0;
  w.zlo = max(c[L] - kStencil, 0);
  w.zhi = min(c[L] + kStencil, T.nc[L] - 1);
  return w;
}

template <int DIM>
__device__ __forceinline__ uint32_t window_row(const Window& w, int ix, int iy) {
  return DIM == 2 ? w.cell_off + static_cast<uint32_t>(ix) * w.nc1
                  : w.cell_off + (static_cast<uint32_t>(ix) * w.nc1 + static_cast<uint32_t>(iy)) * w.nc2;
}

// Position of particle j in state S (= A.a): one 16-byte load of the packed
// copy in 2D.
template <int DIM>
__device__ __forceinline__ void pos_of(const PiArgs& A, const DevState& S, uint32_t j,
                                       double* x) {
  if constexpr (DIM == 2) {
    const double2 q = A.pk_x[j];
    x[0] = q.x;
    x[1] = q.y;
  } else {
#pragma unroll
    for (int k = 0; k < DIM; ++k) x[k] = S.x[k][j];
  }
}

// Calls f(j, d, s2) for every particle j (self included) with
// s2 = |r_j - r_a|^2 <= (2h)^2, d = r_j - r_a, positions from S, in column
// order and ascending j.  With ONLY_MARKED, only the particles of marked
// cells are visited (still in ascending j).
template <int DIM, bool ONLY_MARKED = false, class F>
__device__ __forceinline__ void for_neighbours(const PiArgs& A, uint32_t i, const double* xa,
                                               const DevState& S, F&& f) {
  const Window w = window_of<DIM>(A, i);
  auto visit = [&](uint32_t b, uint32_t e) {
    for (uint32_t j = b; j < e; ++j) {
      double d[DIM];
      pos_of<DIM>(A, S, j, d);
#pragma unroll
      for (int k = 0; k < DIM; ++k) d[k] = d[k] - xa[k];
      double s2 = d[0] * d[0];
#pragma unroll
      for (int k = 1; k < DIM; ++k) s2 = s2 + d[k] * d[k];
      if (s2 <= A.P.R2) f(j, d, s2);
    }
  };
  for (int ix = w.xlo; ix <= w.xhi; ++ix) {
    for (int iy = w.ylo; iy <= w.yhi; ++iy) {
      const uint32_t row = window_row<DIM>(w, ix, iy);
      if constexpr (ONLY_MARKED) {
        for (int z = w.zlo; z <= w.zhi; ++z) {
          if (A.cell_mark[row + z] != 0u) visit(A.cell_off[row + z], A.cell_off[row + z + 1]);
        }
      } else {
        visit(A.cell_off[row + w.zlo], A.cell_off[row + w.zhi + 1]);
      }
    }
  }
}

// Accepted candidates buffered per lane: kPiWords windows of 32 consecutive
// candidates (a mask and the window's first index each), in LDS.
constexpr int kPiWords = 12;

// A neighbour's fields, loaded ahead of its turn.
template <int DIM>
struct NbFields {
  double x[DIM], v[DIM], rho, vol;
};

// The same neighbours in the same order as for_neighbours, for a callback
// too heavy to run under the acceptance branch: a wave's lanes walk
// different candidates, so under the branch the wave pays the callback on
// nearly every candidate iteration.  Here the predicate only fills masks
// (positions loaded a batch at a time), and the callback runs in one flat
// loop over the buffered masks, a wave paying it about max-over-lanes
// (neighbour count) times, with the next neighbour's fields (S's x, v, rho;
// vol; in 2D from the packed copies of A, which hold S = A.a) loaded while
// the current one is summed.  f(j, d, s2, fields).
// s_m, s_j: this workgroup's kPiWords x NT buffers (NT threads).
template <int DIM, int NT, class F>
__device__ __forceinline__ void for_neighbours_packed(const PiArgs& A, uint32_t i,
                                                      const double* xa, const DevState& S,
                                                      uint32_t* s_m, uint32_t* s_j, F&& f) {
  const uint32_t lane = threadIdx.x;
  int nw = 0;
  auto drain = [&]() {
    int w = 0;
    uint32_t m = nw > 0 ? s_m[lane] : 0u;
    // the next buffered neighbour, if any
    auto next = [&](uint32_t& j) -> bool {
      while (m == 0u && ++w < nw) m = s_m[w * NT + lane];
      if (m == 0u) return false;
      const uint32_t t = static_cast<uint32_t>(__builtin_ctz(m));
      m &= m - 1u;
      j = s_j[w * NT + lane] + t;
      return true;
    };
    auto load = [&](uint32_t j, NbFields<DIM>& nb) {
      if constexpr (DIM == 2) {
        const double2 x = A.pk_x[j], v = A.pk_v[j], rv = A.pk_rv[j];
        nb.x[0] = x.x;
        nb.x[1] = x.y;
        nb.v[0] = v.x;
        nb.v[1] = v.y;
        nb.rho = rv.x;
        nb.vol = rv.y;
      } else {
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
          nb.x[k] = S.x[k][j];
          nb.v[k] = S.v[k][j];
        }
        nb.rho = S.rho[j];
        nb.vol = A.vol[j];
      }
    };
    uint32_t j = 0;
    bool have = next(j);
    NbFields<DIM> cur;
    if (have) load(j, cur);
    while (have) {
      uint32_t j2 = 0;
      const bool have2 = next(j2);
      NbFields<DIM> nxt;
      if (have2) load(j2, nxt);
      double d[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) d[k] = cur.x[k] - xa[k];
      double s2 = d[0] * d[0];
#pragma unroll
      for (int k = 1; k < DIM; ++k) s2 = s2 + d[k] * d[k];
      f(j, d, s2, cur);
      j = j2;
      cur = nxt;
      have = have2;
    }
    nw = 0;
  };
  const Window win = window_of<DIM>(A, i);
  for (int ix = win.xlo; ix <= win.xhi; ++ix) {
    for (int iy = win.ylo; iy <= win.yhi; ++iy) {
      const uint32_t row = window_row<DIM>(win, ix, iy);
      const uint32_t b = A.cell_off[row + win.zlo], e = A.cell_off[row + win.zhi + 1];
      for (uint32_t c0 = b; c0 < e; c0 += 32u) {
        const uint32_t len = min(32u, e - c0);
        uint32_t mask = 0u;
        // kBatch candidates' positions loaded together, then tested (the
        // loads of one batch overlap; indices past the window's end repeat
        // its last candidate and are masked out)
        constexpr uint32_t kBatch = DIM == 2 ? 8u : 4u;
        for (uint32_t t0 = 0; t0 < len; t0 += kBatch) {
          double px[kBatch][DIM];
#pragma unroll
          for (uint32_t u = 0; u < kBatch; ++u) pos_of<DIM>(A, S, c0 + min(t0 + u, len - 1u), px[u]);
#pragma unroll
          for (uint32_t u = 0; u < kBatch; ++u) {
            double d[DIM];
#pragma unroll
            for (int k = 0; k < DIM; ++k) d[k] = px[u][k] - xa[k];
            double s2 = d[0] * d[0];
#pragma unroll
            for (int k = 1; k < DIM; ++k) s2 = s2 + d[k] * d[k];
            if (t0 + u < len && s2 <= A.P.R2) mask |= 1u << (t0 + u);
          }
        }
        if (mask != 0u) {
          s_m[nw * NT + lane] = mask;
          s_j[nw * NT + lane] = c0;
          if (++nw == kPiWords) drain();
        }
      }
    }
  }
  drain();
}

// f'(q) / q of the unit kernel (f' = -5 q t^3 for Wendland C2,
// -14/3 q (1 + 5q/2) t^5 for C4, t = 1 - q/2; sph/kernel.gen.cpp:539-546):
// grad W = w f'(q) / h * r_ab / |r_ab| = (w / h^2) (f'(q) / q) r_ab, with no
// division by |r_ab|.
template <int KERNEL>
__device__ __forceinline__ double unit_deriv_over_q(double q) {
  if (!(q < 2.0)) return 0.0;
  const double t = 1.0 - q / 2.0;
  if constexpr (KERNEL == kKernelC2) {
    return -5.0 * (t * t * t);
  } else {
    const double t5 = t * t * t * t * t;
    return -14.0 / 3.0 * (1.0 + 2.5 * q) * t5;
  }
}

template <int DIM, int KERNEL>
__device__ __forceinline__ double kernel_factor(const DevParams& P) {
  const double omega = KERNEL == kKernelC2
                           ? (DIM == 2 ? 7.0 / (4.0 * M_PI) : 21.0 / (16.0 * M_PI))
                           : (DIM == 2 ? 9.0 / (4.0 * M_PI) : 495.0 / (256.0 * M_PI));
  const double h_inv = 1.0 / P.h;
  double hd = h_inv;  // omega / h^d (sph/kernel.hpp:142-151)
#pragma unroll
  for (int d = 1; d < DIM; ++d) hd *= h_inv;
  return omega * hd;
}

template <int DIM>
__device__ __forceinline__ void normalize_f64(const double* a, double* out) {
  double n2 = a[0] * a[0];
#pragma unroll
  for (int k = 1; k < DIM; ++k) n2 = n2 + a[k] * a[k];
  if (n2 >= kTiny * kTiny) {
    const double s = sqrt(n2);
#pragma unroll
    for (int k = 0; k < DIM; ++k) out[k] = a[k] / s;
  } else {
#pragma unroll
    for (int k = 0; k < DIM; ++k) out[k] = 0.0;
  }
}

// lu + FactLU::inverse (core/_mat/fact.hpp:60-107): Doolittle factors, near
// singular when a pivot is tiny, then L U X = I column by column.
template <int DIM>
__device__ __forceinline__ bool lu_inverse(const double* M, double* inv) {
  double LU[DIM * DIM];
#pragma unroll
  for (int i = 0; i < DIM; ++i) {
#pragma unroll
    for (int j = 0; j < i; ++j) {
      double x = M[i * DIM + j];
#pragma unroll
      for (int k = 0; k < j; ++k) x -= LU[i * DIM + k] * LU[k * DIM + j];
      LU[i * DIM + j] = x / LU[j * DIM + j];
    }
#pragma unroll
    for (int j = i; j < DIM; ++j) {
      double x = M[i * DIM + j];
#pragma unroll
      for (int k = 0; k < i; ++k) x -= LU[i * DIM + k] * LU[k * DIM + j];
      LU[i * DIM + j] = x;
    }
    if (fabs(LU[i * DIM + i]) <= kTiny) return false;
  }
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    double y[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      double x = i == c ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < i; ++k) x -= LU[i * DIM + k] * y[k];
      y[i] = x;
    }
#pragma unroll
    for (int i = DIM - 1; i >= 0; --i) {
      double x = y[i];
#pragma unroll
      for (int k = i + 1; k < DIM; ++k) x -= LU[i * DIM + k] * y[k];
      y[i] = x / LU[i * DIM + i];
    }
#pragma unroll
    for (int i = 0; i < DIM; ++i) inv[i * DIM + c] = y[i];
  }
  return true;
}

// Mass of particle j: its uploaded mass (or the uniform m), times gamma for a
// fixed particle with walls (FluidEquations::initialize,
// fluid_equations.hpp:87-88).
__device__ __forceinline__ double mass_of(const PiArgs& A, uint32_t j) {
  if (!A.P.mass && !A.wall_mu) return A.P.m;
  const uint32_t id = A.id[j];
  const double m = particle_mass(A.P, id);
  if (A.wall_mu && (A.tag[j] & 1u)) return m * A.wall_mu[id - A.wall_id0];
  return m;
}

// gamma of fluid particle i (1 without walls; fixed particles' sums are not used).
__device__ __forceinline__ double gamma_of(const PiArgs& A, uint32_t i) {
  return A.wall && !(A.tag[i] & 1u) ? A.wall[4 * static_cast<size_t>(i) + 3] : 1.0;
}

__global__ __launch_bounds__(kPiBlock) void k_pi_volume(PiArgs A) {
  const uint32_t i = blockIdx.x * kPiBlock + threadIdx.x;
  if (i >= A.n) return;
  const double rho = A.a.rho[i], vol = mass_of(A, i) / rho;
  A.vol[i] = vol;
  if (A.pk_x) {  // 2D
    A.pk_x[i] = make_double2(A.a.x[0][i], A.a.x[1][i]);
    A.pk_v[i] = make_double2(A.a.v[0][i], A.a.v[1][i]);
    A.pk_rv[i] = make_double2(rho, vol);
  }
}

// apply_shifts, sums and renormalisation (:335-365) of particle i at xa:
// walk(f) calls f(j, d, s2, fields) for each neighbour j (self included).
template <int DIM, int KERNEL, class Walk>
__device__ __forceinline__ void pi_sums_one(const PiArgs& A, uint32_t i, const double* xa,
                                            Walk&& walk) {
  const DevParams& P = A.P;
  double va[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) va[k] = A.a.v[k][i];
  const double rho_a = A.a.rho[i];
  const double h_inv = 1.0 / P.h, w2 = kernel_factor<DIM, KERNEL>(P) * (h_inv * h_inv);
  double N[DIM] = {}, Lm[DIM * DIM] = {}, G[DIM * DIM] = {}, g[DIM] = {};
  if constexpr (DIM == 2) {
    // the face terms of a particle near a wall (fluid_equations.hpp:342-349)
    const uint32_t q = A.near_idx ? A.near_idx[i] : ~0u;
    if (q != ~0u) {
      const double* t = A.pi_face + 12 * static_cast<size_t>(q);
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        N[p] = t[p];
        g[p] = t[10 + p];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        Lm[k] = t[2 + k];
        G[k] = t[6 + k];
      }
    }
  }
  const double ga = gamma_of(A, i);
  uint32_t cnt = 0;
  walk([&](uint32_t j, const double* d, double s2, const NbFields<DIM>& nb) {
    ++cnt;
    if (j == i) return;
    const double rho_b = nb.rho;
    // V_b / gamma_a (x / 1 is x: no division away from walls)
    const double V_b = ga == 1.0 ? nb.vol : nb.vol / ga;
    // grad W_ab, r_ab = -d (zero below the tiny separation, where the
    // reference's normalize gives a zero direction)
    const double s = s2 >= kTiny * kTiny ? w2 * unit_deriv_over_q<KERNEL>(h_inv * sqrt(s2)) : 0.0;
    double gW[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) gW[k] = s * -d[k];
    // V_b grad W once: the sums below are V_b (x grad W) re-associated as
    // x (V_b grad W) (rounding differs from the reference's order by an ulp)
    double Vg[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) Vg[k] = V_b * gW[k];
    const double drho = rho_b - rho_a;
#pragma unroll
    for (int p = 0; p < DIM; ++p) {
      N[p] += Vg[p];
      const double v_ba = nb.v[p] - va[p];
#pragma unroll
      for (int q = 0; q < DIM; ++q) {
        Lm[p * DIM + q] += d[p] * Vg[q];
        G[p * DIM + q] += v_ba * Vg[q];
      }
      g[p] += drho * Vg[p];
    }
  });
  // renormalisation (:354-365)
  double LT[DIM * DIM], inv[DIM * DIM];
#pragma unroll
  for (int p = 0; p < DIM; ++p) {
    A.dr0[p][i] = N[p];
#pragma unroll
    for (int q = 0; q < DIM; ++q) LT[p * DIM + q] = Lm[q * DIM + p];
  }
  if (lu_inverse<DIM>(LT, inv)) {
    double t[DIM * DIM];
#pragma unroll
    for (int p = 0; p < DIM; ++p) {
      double x = 0.0;
#pragma unroll
      for (int q = 0; q < DIM; ++q) x += inv[p * DIM + q] * N[q];
      t[p] = x;
    }
#pragma unroll
    for (int p = 0; p < DIM; ++p) N[p] = t[p];
#pragma unroll
    for (int p = 0; p < DIM; ++p) {
#pragma unroll
      for (int q = 0; q < DIM; ++q) {
        double x = 0.0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) x += G[p * DIM + k] * inv[q * DIM + k];
        t[p * DIM + q] = x;
      }
    }
#pragma unroll
    for (int k = 0; k < DIM * DIM; ++k) G[k] = t[k];
#pragma unroll
    for (int p = 0; p < DIM; ++p) {
      double x = 0.0;
#pragma unroll
      for (int q = 0; q < DIM; ++q) x += inv[p * DIM + q] * g[q];
      t[p] = x;
    }
#pragma unroll
    for (int p = 0; p < DIM; ++p) g[p] = t[p];
  }
  double nn[DIM];
  normalize_f64<DIM>(N, nn);
#pragma unroll
  for (int p = 0; p < DIM; ++p) {
    A.N[p][i] = nn[p];
    A.grad_rho[p][i] = g[p];
#pragma unroll
    for (int q = 0; q < DIM; ++q) A.grad_v[p * DIM + q][i] = G[p * DIM + q];
  }
  A.count[i] = cnt;
}


// One thread per particle, neighbours from global memory.
template <int DIM, int KERNEL>
__global__ __launch_bounds__(kPiBlock) void k_pi_sums(PiArgs A) {
  __shared__ uint32_t s_m[kPiWords * kPiBlock], s_j[kPiWords * kPiBlock];
  const uint32_t i = blockIdx.x * kPiBlock + threadIdx.x;
  if (i >= A.n) return;
  double xa[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) xa[k] = A.a.x[k][i];
  pi_sums_one<DIM, KERNEL>(A, i, xa, [&](auto&& f) {
    for_neighbours_packed<DIM, kPiBlock>(A, i, xa, A.a, s_m, s_j, f);
  });
}

// Whether a neighbour j != i lies inside the 45-degree cone behind particle
// i along its normal N (:392-417).  Any one decides, so the order of the
// walk does not change the answer: each axis of the window is walked from
// the side the cone opens to (descending where N_k < 0, since the cone
// opens along -N), so that an interior particle meets one within its
// first candidates instead of after most of its window.
template <int DIM>
__device__ __forceinline__ bool any_in_cone(const PiArgs& A, uint32_t i, const double* xa,
                                            const double* Na) {
  const double cos_fov = 0.7071067811865476;  // cos(pi / 4) in FP64
  const Window w = window_of<DIM>(A, i);
  const bool rx = Na[0] < 0.0, ry = DIM == 3 && Na[1] < 0.0, rz = Na[DIM - 1] < 0.0;
  const int nx = w.xhi - w.xlo + 1, ny = w.yhi - w.ylo + 1;
  for (int a = 0; a < nx; ++a) {
    const int ix = rx ? w.xhi - a : w.xlo + a;
    for (int c = 0; c < ny; ++c) {
      const int iy = ry ? w.yhi - c : w.ylo + c;
      const uint32_t row = window_row<DIM>(w, ix, iy);
      const uint32_t b = A.cell_off[row + w.zlo], e = A.cell_off[row + w.zhi + 1];
      // kBatch candidates loaded together, then tested (indices past the
      // column's end repeat its last candidate: a repeat decides nothing new)
      constexpr uint32_t kBatch = 4u;
      const uint32_t len = e - b;
      for (uint32_t k0 = 0; k0 < len; k0 += kBatch) {
        uint32_t jj[kBatch];
        double px[kBatch][DIM];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
          const uint32_t k = min(k0 + u, len - 1u);
          jj[u] = rz ? e - 1u - k : b + k;
          pos_of<DIM>(A, A.a, jj[u], px[u]);
        }
        bool hit = false;
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
          double d[DIM];
#pragma unroll
          for (int q = 0; q < DIM; ++q) d[q] = px[u][q] - xa[q];
          double s2 = d[0] * d[0];
#pragma unroll
          for (int q = 1; q < DIM; ++q) s2 = s2 + d[q] * d[q];
          const double fov = (cos_fov * cos_fov) * s2;
          double n_a = Na[0] * -d[0];
#pragma unroll
          for (int q = 1; q < DIM; ++q) n_a = n_a + Na[q] * -d[q];
          hit = hit || (jj[u] != i && s2 <= A.P.R2 && n_a > 0.0 && n_a * n_a >= fov);
        }
        if (hit) return true;
      }
    }
  }
  return false;
}

// Free-surface indicator (:375-426): far when a neighbour lies inside the
// 45-degree cone behind the particle along N; splashes are on the surface.
template <int DIM>
__global__ __launch_bounds__(kPiBlock) void k_pi_phi(PiArgs A) {
  const uint32_t i = blockIdx.x * kPiBlock + threadIdx.x;
  if (i >= A.n) return;
  if (A.tag[i] & 1u) {
    A.phi[i] = kPhiMax;
    return;
  }
  double phi = DBL_MIN;
  const uint32_t cutoff = DIM == 2 ? 8u : 26u;
  if (A.count[i] > cutoff) {
    double xa[DIM], Na[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      xa[k] = A.a.x[k][i];
      Na[k] = A.N[k][i];
    }
    if (any_in_cone<DIM>(A, i, xa, Na)) phi = kPhiMax;
  }
  A.phi[i] = phi;
  if (same_bits(phi, DBL_MIN)) {
    // a surface particle's cell, and every cell whose window holds it (the
    // window relation is symmetric: the cells within kStencil of its cell)
    A.cell_mark[A.key[i]] = 1u;
    const Window w = window_of<DIM>(A, i);
    for (int ix = w.xlo; ix <= w.xhi; ++ix) {
      for (int iy = w.ylo; iy <= w.yhi; ++iy) {
        const uint32_t row = window_row<DIM>(w, ix, iy);
        for (int z = w.zlo; z <= w.zhi; ++z) A.cell_near[row + z] = 1u;
      }
    }
  }
}

__global__ __launch_bounds__(kPiBlock) void k_pi_clear(uint32_t n, uint32_t* mark) {
  const uint32_t i = blockIdx.x * kPiBlock + threadIdx.x;
  if (i < n) mark[i] = 0u;
}

// A lane's running per-tank value: the lanes of a persistent workgroup walk
// a contiguous range of particles, so a lane's tank changes only at tank
// boundaries (then the old tank's value goes out with one atomic); at the
// end each wave combines its lanes' values (one atomic when they share a
// tank).
template <bool MAX, class T>
struct TankRun {
  uint32_t t = ~0u;
  T v = MAX ? T(0) : ~T(0);
  __device__ __forceinline__ void add(uint32_t tank, T x, T* slot) {
    if (tank != t) {
      if (t != ~0u) {
        if constexpr (MAX) {
          atomicMax(red_slot(slot, t), v);
        } else {
          atomicMin(red_slot(slot, t), v);
        }
      }
      t = tank;
      v = x;
    } else {
      v = MAX ? (x > v ? x : v) : (x < v ? x : v);
    }
  }
  __device__ __forceinline__ void flush(T* slot) { tank_atomic<MAX>(t != ~0u, t, v, slot); }
};

// This workgroup's contiguous share of the n particles, walked in steps of
// the workgroup size (a uniform trip count).
struct Share {
  uint32_t beg, end;
};
__device__ __forceinline__ Share share(uint32_t n) {
  const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
  const uint32_t beg = min(n, blockIdx.x * per);
  return Share{beg, min(n, beg + per)};
}

// Near-surface classification and the shifts (:439-470), into B.
template <int DIM>
__device__ __forceinline__ void shift_one(const PiArgs& A, uint32_t i, bool& shifted,
                                          unsigned long long& shift_bits) {
  const DevParams& P = A.P;
  double xa[DIM], va[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    xa[k] = A.a.x[k][i];
    va[k] = A.a.v[k][i];
  }
  double rho = A.a.rho[i];
  double phi = A.phi[i];
  const bool fluid = !(A.tag[i] & 1u);
  // only a particle whose window holds a surface particle can have one as a
  // neighbour
  if (fluid && same_bits(phi, kPhiMax) && A.cell_near[A.key[i]] != 0u) {
    // the nearest surface neighbour; the reference's lists are in index
    // order, so equal distances go to the smaller index
    uint32_t best = ~0u, best_id = ~0u;
    double best_d2 = 0.0, bx[DIM] = {};
    // only the cells holding a surface particle can hold one
    for_neighbours<DIM, true>(A, i, xa, A.a, [&](uint32_t j, const double* d, double s2) {
      if (!same_bits(A.phi[j], DBL_MIN)) return;
      const uint32_t jid = A.id[j];
      if (best == ~0u || s2 < best_d2 || (s2 == best_d2 && jid < best_id)) {
        best = j;
        best_id = jid;
        best_d2 = s2;
#pragma unroll
        for (int k = 0; k < DIM; ++k) bx[k] = -d[k];
      }
    });
    if (best != ~0u) {
      double nb = A.N[0][best] * bx[0];
#pragma unroll
      for (int k = 1; k < DIM; ++k) nb = nb + A.N[k][best] * bx[k];
      phi *= fabs(nb) / (2.0 * P.h);
    }
  }
  A.phi_near[i] = phi;
  if (fluid && same_bits(phi, kPhiMax)) {
    const double coef = -0.4 * 0.2 * (P.h * P.h);  // -CFL_ C_shift_ h^2
    double dr[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) dr[k] = A.dr0[k][i] * coef;
    double dv[DIM];
#pragma unroll
    for (int p = 0; p < DIM; ++p) {
      double x = 0.0;
#pragma unroll
      for (int q = 0; q < DIM; ++q) x += A.grad_v[p * DIM + q][i] * dr[q];
      dv[p] = x;
    }
    double drho = A.grad_rho[0][i] * dr[0];
#pragma unroll
    for (int k = 1; k < DIM; ++k) drho = drho + A.grad_rho[k][i] * dr[k];
    // the velocity correction only where gamma = 1 (:465-468)
    const bool correct_v = fabs(gamma_of(A, i) - 1.0) <= kTiny;
    double d2 = 0.0;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      xa[k] += dr[k];
      if (correct_v) va[k] += dv[k];
      d2 += dr[k] * dr[k];
    }
    rho += drho;
    shifted = true;
    shift_bits = dbl_bits(sqrt(d2));
  }
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    A.b.x[k][i] = xa[k];
    A.b.v[k][i] = va[k];
  }
  A.b.rho[i] = rho;
}

// One particle per thread; the largest shift goes to the tank's slots with
// one atomic per wave.
template <int DIM>
__global__ __launch_bounds__(kPiBlock) void k_pi_shift(PiArgs A) {
  const uint32_t i = blockIdx.x * kPiBlock + threadIdx.x;
  bool shifted = false;
  unsigned long long bits = 0ull;
  if (i < A.n) shift_one<DIM>(A, i, shifted, bits);
  tank_atomic<true>(shifted, shifted ? A.tag[i] >> 1 : 0u, bits, A.shift_max);
}

// apply_free_surface_correction (:483-512): on the neighbour sets of the
// unshifted positions, with W at the shifted ones; densities into C.
template <int DIM, int KERNEL>
__global__ __launch_bounds__(kPiBlock) void k_pi_fsc(PiArgs A) {
  const uint32_t i = blockIdx.x * kPiBlock + threadIdx.x;
  if (i >= A.n) return;
  const DevParams& P = A.P;
  double rho = A.b.rho[i];
  if (!(A.tag[i] & 1u) && !same_bits(A.phi_near[i], kPhiMax)) {
    double xa[DIM], xs[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      xa[k] = A.a.x[k][i];
      xs[k] = A.b.x[k][i];
    }
    const double w = kernel_factor<DIM, KERNEL>(P), h_inv = 1.0 / P.h;
    double alpha = 0.0, rho_t = 0.0;
    for_neighbours<DIM>(A, i, xa, A.a, [&](uint32_t j, const double*, double) {
      double x2 = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        const double x = xs[k] - A.b.x[k][j];
        x2 += x * x;
      }
      const double W = w * kernel_value_poly_f64<KERNEL>(h_inv * sqrt(x2));
      const double m_b = mass_of(A, j);
      alpha += m_b / A.b.rho[j] * W;
      rho_t += m_b * W;
    });
    const double ga = gamma_of(A, i);
    const double ratio = fmin(1.0, alpha / ga);
    if (!(ratio > 0.99)) {
      const double K = -log(0.05) / (0.01 * 0.01);
      const double beta = exp(-K * ((ratio - 1.0) * (ratio - 1.0)));
      const double corr = beta * ga + (1.0 - beta) * alpha;
      if (fabs(corr) > kTiny) rho = rho_t / corr;
    }
  }
  A.c.rho[i] = rho;
}

// Shift displacement into the index's bound; the next step size is taken
// again after the shifts.  One wave per tank.  The passes above read the
// neighbour sets of the unshifted positions, which the index still covers; a
// shift that takes a tank past skin/2 marks a rebuild of every tank pending
// (ctl[1]), which the next step or hook runs before it searches again.
__global__ __launch_bounds__(64) void k_pi_bound(unsigned long long* shift_max, double* disp,
                                                 double skin_half, uint32_t* ctl,
                                                 unsigned long long* dt_min) {
  const uint32_t t = blockIdx.x;
  const unsigned long long sm = red_fold<true>(shift_max, t, true);
  dt_min[static_cast<size_t>(t) * kRedStripes + threadIdx.x] = ~0ull;
  if (threadIdx.x != 0) return;
  const double d = disp[t] + __longlong_as_double(static_cast<long long>(sm));
  disp[t] = d;
  if (!(d <= skin_half)) ctl[1] = 1u;
}

// Persistent: a lane walks a contiguous run of particles and keeps running
// per-tank values, so the grid's flushes stay few (a one-particle-per-thread
// form, with two atomics per wave over 31k waves, measured 2.6x slower).
template <int DIM>
__global__ __launch_bounds__(kPiBlock) void k_pi_commit(PiArgs A) {
  const Share sh = share(A.n);
  TankRun<false, unsigned long long> dt_run;
  TankRun<true, uint32_t> v_run;
  for (uint32_t base = sh.beg; base < sh.end; base += kPiBlock) {
    const uint32_t i = base + threadIdx.x;
    if (i >= sh.end) continue;
    double v[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      v[k] = A.b.v[k][i];
      A.a.x[k][i] = A.b.x[k][i];
      A.a.v[k][i] = v[k];
    }
    const double rho = A.c.rho[i];
    A.a.rho[i] = rho;
    write_pair_data<DIM>(A.P, A.a, i, v, rho, mass_of(A, i));
    const uint32_t tg = A.tag[i];
    if (!(tg & 1u)) {
      double dv[DIM], v2 = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        dv[k] = A.dv_last[k][i];
        v2 += v[k] * v[k];
      }
      dt_run.add(tg >> 1, dbl_bits(particle_time_step<DIM>(A.P, rho, v, dv)), A.dt_min);
      v_run.add(tg >> 1, __float_as_uint(static_cast<float>(sqrt(v2))), A.vmax);
    }
  }
  dt_run.flush(A.dt_min);
  v_run.flush(A.vmax);
}

__global__ void k_export_f64(uint32_t n, const uint32_t* __restrict__ id,
                             const double* __restrict__ src, double* __restrict__ dst) {
  const uint32_t i = blockIdx.x * kPiBlock + threadIdx.x;
  if (i < n) dst[id[i]] = src[i];
}

// Workgroups of K resident per CU (registers and LDS), asked once: the
// persistent passes launch exactly that many per CU, so that no part of
// their grid waits for a second round.
template <auto K>
uint32_t resident_per_cu() {
  static const uint32_t n = [] {
    int b = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void*>(K),
                                                       kPiBlock, 0);
    return static_cast<uint32_t>(std::max(b, 1));
  }();
  return n;
}

}  // namespace

void launch_export_f64(hipStream_t s, uint32_t n, const uint32_t* id, const double* src,
                       double* dst) {
  if (n) k_export_f64<<<(n + kPiBlock - 1) / kPiBlock, kPiBlock, 0, s>>>(n, id, src, dst);
}

void launch_post_integrate(hipStream_t s, const PiArgs& a, int n_tanks, double* disp,
                           double skin_half, uint32_t* ctl, int n_cu) {
  if (a.n == 0) return;
  const uint32_t nb = (a.n + kPiBlock - 1) / kPiBlock;
  const int dim = a.P.dim, kernel = a.P.kernel;
  // the commit pass runs persistent, as many workgroups as stay resident
  const uint32_t cus = static_cast<uint32_t>(std::max(n_cu, 1));
  const uint32_t np_commit =
      std::min(nb, cus * (dim == 2 ? resident_per_cu<k_pi_commit<2>>() : resident_per_cu<k_pi_commit<3>>()));
  // cell_mark and cell_near lie back to back, n_cells + 1 entries each
  const uint32_t n_marks = 2u * (a.n_cells + 1u);
  k_pi_volume<<<nb, kPiBlock, 0, s>>>(a);
  if (dim == 2) {
    if (kernel == kKernelC2) {
      k_pi_sums<2, kKernelC2><<<nb, kPiBlock, 0, s>>>(a);
    } else {
      k_pi_sums<2, kKernelC4><<<nb, kPiBlock, 0, s>>>(a);
    }
    k_pi_clear<<<(n_marks + kPiBlock - 1) / kPiBlock, kPiBlock, 0, s>>>(n_marks, a.cell_mark);
    k_pi_phi<2><<<nb, kPiBlock, 0, s>>>(a);
    k_pi_shift<2><<<nb, kPiBlock, 0, s>>>(a);
    if (kernel == kKernelC2) {
      k_pi_fsc<2, kKernelC2><<<nb, kPiBlock, 0, s>>>(a);
    } else {
      k_pi_fsc<2, kKernelC4><<<nb, kPiBlock, 0, s>>>(a);
    }
  } else {
    if (kernel == kKernelC2) {
      k_pi_sums<3, kKernelC2><<<nb, kPiBlock, 0, s>>>(a);
    } else {
      k_pi_sums<3, kKernelC4><<<nb, kPiBlock, 0, s>>>(a);
    }
    k_pi_clear<<<(n_marks + kPiBlock - 1) / kPiBlock, kPiBlock, 0, s>>>(n_marks, a.cell_mark);
    k_pi_phi<3><<<nb, kPiBlock, 0, s>>>(a);
    k_pi_shift<3><<<nb, kPiBlock, 0, s>>>(a);
    if (kernel == kKernelC2) {
      k_pi_fsc<3, kKernelC2><<<nb, kPiBlock, 0, s>>>(a);
    } else {
      k_pi_fsc<3, kKernelC4><<<nb, kPiBlock, 0, s>>>(a);
    }
  }
  k_pi_bound<<<n_tanks, 64, 0, s>>>(a.shift_max, disp, skin_half, ctl, a.dt_min);
  if (dim == 2) {
    k_pi_commit<2><<<np_commit, kPiBlock, 0, s>>>(a);
  } else {
    k_pi_commit<3><<<np_commit, kPiBlock, 0, s>>>(a);
  }
}

}  // namespace tit_amd
