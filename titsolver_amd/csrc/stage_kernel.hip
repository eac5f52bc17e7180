// Fused SSPRK stage kernel: continuity + momentum pair sums, EOS, update.
//
// Work is split into strips: one workgroup takes S consecutive cells along
// the fast (last) axis of one cell column of one tank, and every particle in
// those cells is a target.  In 2D the candidates of the whole strip (5
// neighbouring columns, S + 4 rows each) are staged in LDS once: FP64
// positions for the exact predicate, FP32 velocity, FP64 density and the
// per-particle EOS terms p/rho^2 and m/rho.  Each target lane then walks its
// own 5 x 5 cell stencil in LDS in two passes per column range: the FP64
// predicate builds a 64-bit mask of accepted candidates, and the FP32 pair
// terms run over the set bits only, so rejected candidates never occupy the
// pair arithmetic.  A strip whose candidates exceed the LDS capacity, and 3D,
// read the same candidates from global memory in the same order, so every
// path produces bitwise identical sums.
//
// Workgroups are assigned XCD-major: launch index b runs on XCD b % 8, and
// takes strip (b % 8) * ceil(n/8) + b / 8, so each XCD's L2 sees a contiguous
// band of columns and the halo columns it stages are mostly its own.
#include "pair_math.hpp"
#include "sph_kernels.hpp"

namespace tit_amd {

namespace {

constexpr int kCols2 = 2 * kStencil + 1;

template <int CAP>
struct StripLds {
  double2 xy[CAP];
  float2 v[CAP];
  double rho[CAP];
  float P[CAP];
  float V[CAP];
};

// Candidate source reading global memory (any dimension).
template <int DIM>
struct GlobalSrc {
  const DevParams& P;
  const DevTank& T;
  const DevState& S;
  const uint32_t* coff;
  int c[3];  // target column (and the fast-axis cell is passed per call)

  __device__ __forceinline__ bool range(int k, int zc, uint32_t& b,
                                        uint32_t& e) const {
    constexpr int L = DIM - 1;
    int ix, iy = 0;
    if constexpr (DIM == 2) {
      ix = c[0] - kStencil + k;
    } else {
      ix = c[0] - kStencil + k / kCols2;
      iy = c[1] - kStencil + k % kCols2;
      if (iy < 0 || iy >= T.nc[1]) return false;
    }
    if (ix < 0 || ix >= T.nc[0]) return false;
    const int nz = T.nc[L];
    const int z0 = max(zc - kStencil, 0), z1 = min(zc + kStencil, nz - 1);
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
  __device__ __forceinline__ void data(int, uint32_t j, float* v, double& rho,
                                       float& Pb, float& Vb) const {
#pragma unroll
    for (int k = 0; k < DIM; ++k) v[k] = static_cast<float>(S.v[k][j]);
    rho = S.rho[j];
    Pb = pressure_over_rho2(P, rho);
    Vb = P.m_f / static_cast<float>(rho);
  }
};

// Candidate source reading a strip staged in LDS (2D).
template <int CAP, int ROWS>
struct LdsSrc {
  const StripLds<CAP>& L;
  const uint32_t (*coff)[ROWS];  // [column][row - zlo]
  const uint32_t* gb;            // global index of each column's first row
  const uint32_t* lb;            // LDS index of each column's first row
  int zlo, nz;
  bool valid[kCols2];

  __device__ __forceinline__ bool range(int k, int zc, uint32_t& b,
                                        uint32_t& e) const {
    if (!valid[k]) return false;
    const int z0 = max(zc - kStencil, 0), z1 = min(zc + kStencil, nz - 1);
    b = coff[k][z0 - zlo];
    e = coff[k][z1 + 1 - zlo];
    return true;
  }
  __device__ __forceinline__ uint32_t lds(int k, uint32_t j) const {
    return lb[k] + (j - gb[k]);
  }
  __device__ __forceinline__ void pos(int k, uint32_t j, double* x) const {
    const double2 p = L.xy[lds(k, j)];
    x[0] = p.x;
    x[1] = p.y;
  }
  __device__ __forceinline__ void data(int k, uint32_t j, float* v, double& rho,
                                       float& Pb, float& Vb) const {
    const uint32_t l = lds(k, j);
    const float2 vv = L.v[l];
    v[0] = vv.x;
    v[1] = vv.y;
    rho = L.rho[l];
    Pb = L.P[l];
    Vb = L.V[l];
  }
};

// Pair sums of one target over its stencil, column range by column range:
// an FP64 predicate pass builds the acceptance mask of up to 64 candidates,
// then the pair terms run over the accepted ones.
template <int DIM, int KERNEL, int FLAVOUR, int NCOL, class Src>
__device__ __forceinline__ void gather(const DevParams& P, const Src& src,
                                       const Target<DIM>& a, uint32_t i, int zc,
                                       Sums<DIM>& acc) {
  for (int k = 0; k < NCOL; ++k) {
    uint32_t b, e;
    if (!src.range(k, zc, b, e)) continue;
    for (uint32_t c0 = b; c0 < e; c0 += 64) {
      const uint32_t len = min(e - c0, 64u);
      uint64_t mask = 0;
      for (uint32_t t = 0; t < len; ++t) {
        const uint32_t j = c0 + t;
        double xb[DIM];
        src.pos(k, j, xb);
        double d0 = xb[0] - a.x[0];
        double s = d0 * d0;
#pragma unroll
        for (int q = 1; q < DIM; ++q) {
          const double dq = xb[q] - a.x[q];
          s = s + dq * dq;
        }
        if (s <= P.R2 && j != i) mask |= 1ull << t;
      }
      while (mask) {
        const uint32_t t = static_cast<uint32_t>(__builtin_ctzll(mask));
        mask &= mask - 1;
        const uint32_t j = c0 + t;
        double xb[DIM], d[DIM];
        src.pos(k, j, xb);
#pragma unroll
        for (int q = 0; q < DIM; ++q) d[q] = xb[q] - a.x[q];
        float vb[DIM], Pb, Vb;
        double rho_b;
        src.data(k, j, vb, rho_b, Pb, Vb);
        pair_term<DIM, KERNEL, FLAVOUR>(P, a, d, vb, rho_b, Pb, Vb, acc);
      }
    }
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

template <int DIM>
__device__ __forceinline__ double particle_dt(const DevParams& P, double rho,
                                             const double* v, const double* dv) {
  double v2 = v[0] * v[0];
  double a2 = dv[0] * dv[0];
#pragma unroll
  for (int k = 1; k < DIM; ++k) {
    v2 = v2 + v[k] * v[k];
    a2 = a2 + dv[k] * dv[k];
  }
  double dt = P.cfl * P.h / (eos_sound_speed(P, rho) + sqrt(v2));
  if (P.mu > 0.0) dt = fmin(dt, P.c_visc * (P.h * P.h) * rho / P.mu);
  const double an = fmax(sqrt(a2), P.g);
  return fmin(dt, P.c_force * sqrt(P.h / an));
}

constexpr int kRowsMax = kStripMax + 2 * kStencil + 1;

template <int DIM, int KERNEL, int FLAVOUR, int KIND, bool FINAL, int WAVES,
          int CAP>
__global__ __launch_bounds__(64 * WAVES) void k_strip(StageArgs A) {
  constexpr int NT = 64 * WAVES;
  constexpr bool kLds = DIM == 2 && CAP > 0;
  constexpr int NCOL = DIM == 2 ? kCols2 : kCols2 * kCols2;
  __shared__ StripLds<kLds ? CAP : 1> lds;
  __shared__ uint32_t s_coff[kLds ? kCols2 : 1][kLds ? kRowsMax : 1];
  __shared__ uint32_t s_gb[kCols2], s_lb[kCols2 + 1];
  __shared__ int s_use_lds;

  const DevParams& P = A.P;
  const uint32_t per = (A.n_work + 7) / 8;
  const uint32_t w = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  if (w >= A.n_work) return;
  const uint4 item = A.work[w];  // tank, column x, column y, first cell z
  const uint32_t t = item.x;
  const DevTank T = A.tanks[t];
  constexpr int L = DIM - 1;
  const int nz = T.nc[L];
  const int z0 = static_cast<int>(item.w);
  const int z1 = min(z0 + static_cast<int>(A.strip_len), nz);
  uint32_t col_row;
  if constexpr (DIM == 2) {
    col_row = T.cell_off + item.y * static_cast<uint32_t>(nz);
  } else {
    col_row = T.cell_off + (item.y * T.nc[1] + item.z) * static_cast<uint32_t>(nz);
  }
  const uint32_t t0 = A.cell_off[col_row + z0], t1 = A.cell_off[col_row + z1];
  if (t0 == t1) return;
  const int tid = threadIdx.x;
  const int ix = static_cast<int>(item.y);

  bool use_lds = false;
  const int zlo = max(z0 - kStencil, 0), zhi = min(z1 - 1 + kStencil, nz - 1);
  if constexpr (kLds) {
    const int nrows = zhi - zlo + 1;
    for (int q = tid; q < kCols2 * (nrows + 1); q += NT) {
      const int k = q / (nrows + 1), r = q % (nrows + 1);
      const int ixk = ix - kStencil + k;
      s_coff[k][r] = (ixk >= 0 && ixk < T.nc[0])
                         ? A.cell_off[T.cell_off + static_cast<uint32_t>(ixk) * nz + zlo + r]
                         : 0u;
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t lb = 0;
      for (int k = 0; k < kCols2; ++k) {
        s_gb[k] = s_coff[k][0];
        s_lb[k] = lb;
        lb += s_coff[k][nrows] - s_coff[k][0];
      }
      s_lb[kCols2] = lb;
      s_use_lds = lb <= A.lds_cap;
    }
    __syncthreads();
    use_lds = s_use_lds;
    if (use_lds) {
      const uint32_t total = s_lb[kCols2];
      for (uint32_t l = tid; l < total; l += NT) {
        int k = 0;
        while (l >= s_lb[k + 1]) ++k;
        const uint32_t j = s_gb[k] + (l - s_lb[k]);
        const double rho = A.in.rho[j];
        lds.xy[l] = make_double2(A.in.x[0][j], A.in.x[1][j]);
        lds.v[l] = make_float2(static_cast<float>(A.in.v[0][j]),
                               static_cast<float>(A.in.v[1][j]));
        lds.rho[l] = rho;
        lds.P[l] = pressure_over_rho2(P, rho);
        lds.V[l] = P.m_f / static_cast<float>(rho);
      }
      __syncthreads();
    }
  }

  float vmax = 0.0f;
  unsigned long long dtmin = ~0ull;
  for (uint32_t base = t0; base < t1; base += NT) {
    const uint32_t i = base + tid;
    const bool in_range = i < t1;
    const uint32_t tag = in_range ? A.tag[i] : 1u;
    const bool fluid = in_range && !(tag & 1u);
    if (in_range && !fluid) {
      if constexpr (KIND == kRhsOnly) {
        A.rhs_drho[i] = 0.0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) A.rhs_dv[k][i] = 0.0;
      } else {
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
          A.out.x[k][i] = A.in.x[k][i];
          A.out.v[k][i] = 0.0;
        }
        A.out.rho[i] = A.in.rho[i];
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
      make_target<DIM>(P, xa, va, rho_a, a);
      const int zc = static_cast<int>(A.key[i] - col_row);
      Sums<DIM> acc;
      acc.init();
      if constexpr (kLds) {
        if (use_lds) {
          LdsSrc<CAP, kRowsMax> src{lds, s_coff, s_gb, s_lb, zlo, nz, {}};
#pragma unroll
          for (int k = 0; k < kCols2; ++k) {
            const int ixk = ix - kStencil + k;
            src.valid[k] = ixk >= 0 && ixk < T.nc[0];
          }
          gather<DIM, KERNEL, FLAVOUR, NCOL>(P, src, a, i, zc, acc);
        }
      }
      if (!use_lds) {
        GlobalSrc<DIM> src{P, T, A.in, A.cell_off, {ix, static_cast<int>(item.z), 0}};
        gather<DIM, KERNEL, FLAVOUR, NCOL>(P, src, a, i, zc, acc);
      }
      const double drho = static_cast<double>(acc.rho.value());
      double dv[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) dv[k] = static_cast<double>(acc.v[k].value());
      dv[1] = dv[1] - P.g;
      if constexpr (KIND == kRhsOnly) {
        A.rhs_drho[i] = drho;
#pragma unroll
        for (int k = 0; k < DIM; ++k) A.rhs_dv[k][i] = dv[k];
      } else {
        double v2 = 0.0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) v2 = v2 + va[k] * va[k];
        vmax = fmaxf(vmax, static_cast<float>(sqrt(v2)));
        const double h = A.dt[t];
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
        if constexpr (FINAL) {
          const unsigned long long b = dbl_bits(particle_dt<DIM>(P, rhon, vn, dv));
          dtmin = b < dtmin ? b : dtmin;
        }
      }
    }
  }
  if constexpr (KIND != kRhsOnly) {
    // every target of the strip belongs to tank t: one atomic per wave
    vmax = wave_max_f32(vmax);
    if ((tid & 63) == 0 && vmax > 0.0f) atomicMax(A.vmax + t, __float_as_uint(vmax));
    if constexpr (FINAL) {
      dtmin = wave_min_u64(dtmin);
      if ((tid & 63) == 0 && dtmin != ~0ull) atomicMin(A.dt_min + t, dtmin);
    }
  }
}

template <int DIM, int KERNEL, int FLAVOUR, int WAVES, int CAP>
void dispatch_kind(hipStream_t s, const StageArgs& a, int kind, bool final) {
  const uint32_t per = (a.n_work + 7) / 8;
  const dim3 grid(8 * per), block(64 * WAVES);
  if (kind == kRhsOnly) {
    k_strip<DIM, KERNEL, FLAVOUR, kRhsOnly, false, WAVES, CAP><<<grid, block, 0, s>>>(a);
  } else if (kind == kEuler) {
    if (final) {
      k_strip<DIM, KERNEL, FLAVOUR, kEuler, true, WAVES, CAP><<<grid, block, 0, s>>>(a);
    } else {
      k_strip<DIM, KERNEL, FLAVOUR, kEuler, false, WAVES, CAP><<<grid, block, 0, s>>>(a);
    }
  } else {
    if (final) {
      k_strip<DIM, KERNEL, FLAVOUR, kCombine, true, WAVES, CAP><<<grid, block, 0, s>>>(a);
    } else {
      k_strip<DIM, KERNEL, FLAVOUR, kCombine, false, WAVES, CAP><<<grid, block, 0, s>>>(a);
    }
  }
}

template <int DIM, int KERNEL, int WAVES, int CAP>
void dispatch_flavour(hipStream_t s, const StageArgs& a, int kind, bool final) {
  if (a.P.flavour == kFlavourCfg) {
    dispatch_kind<DIM, KERNEL, kFlavourCfg, WAVES, CAP>(s, a, kind, final);
  } else {
    dispatch_kind<DIM, KERNEL, kFlavourRef, WAVES, CAP>(s, a, kind, final);
  }
}

template <int DIM, int WAVES, int CAP>
void dispatch_kernel(hipStream_t s, const StageArgs& a, int kind, bool final) {
  if (a.P.kernel == kKernelC4) {
    dispatch_flavour<DIM, kKernelC4, WAVES, CAP>(s, a, kind, final);
  } else {
    dispatch_flavour<DIM, kKernelC2, WAVES, CAP>(s, a, kind, final);
  }
}

}  // namespace

int strip_threads(int dim) { return dim == 2 ? 128 : 128; }
uint32_t strip_lds_capacity(int dim) { return dim == 2 ? kStripLdsCap : 0u; }

void launch_stage(hipStream_t s, const StageArgs& a, int kind, bool final) {
  if (a.n == 0 || a.n_work == 0) return;
  if (a.P.dim == 2) {
    dispatch_kernel<2, 2, kStripLdsCap>(s, a, kind, final);
  } else {
    dispatch_kernel<3, 2, 0>(s, a, kind, final);
  }
}

}  // namespace tit_amd
