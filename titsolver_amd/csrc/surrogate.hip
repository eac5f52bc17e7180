// Learned step (network layout in surrogate.hpp): k-nearest graph, features,
// training forward/backward on the BF16 MFMA GEMM, Adam, the 4 dt update, the
// judge/fallback controller and their C ABI.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "context.hpp"
#include "pair_math.hpp"
#include "surrogate.hpp"

using namespace tit_amd;

namespace tit_amd {

namespace {

constexpr int kBlk = 256;

inline unsigned nblk(size_t n) { return static_cast<unsigned>((n + kBlk - 1) / kBlk); }

// ---- graph ----------------------------------------------------------------------

// The kKnn nearest other particles within h, by (distance^2, id).  The cell
// size is at least h, so a 3^d-cell stencil holds every candidate right after
// a rebuild.
template <int DIM>
__global__ __launch_bounds__(kBlk) void k_knn(DevParams P, const DevTank* __restrict__ tanks,
                                              uint32_t n, DevState st,
                                              const uint32_t* __restrict__ key,
                                              const uint32_t* __restrict__ tag,
                                              const uint32_t* __restrict__ id,
                                              const uint32_t* __restrict__ coff,
                                              uint32_t* __restrict__ nbr) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  const DevTank T = tanks[tag[i] >> 1];
  double xa[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) xa[k] = st.x[k][i];
  const double h2 = P.h * P.h;
  double bd[kKnn];
  uint32_t bid[kKnn], bj[kKnn];
#pragma unroll
  for (int q = 0; q < kKnn; ++q) {
    bd[q] = INFINITY;
    bid[q] = 0xffffffffu;
    bj[q] = i;
  }
  const uint32_t local = key[i] - T.cell_off;
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
  const int zlo = max(c[L] - 1, 0), zhi = min(c[L] + 1, T.nc[L] - 1);
  const int x0 = max(c[0] - 1, 0), x1 = min(c[0] + 1, T.nc[0] - 1);
  int y0 = 0, y1 = 0;
  if constexpr (DIM == 3) {
    y0 = max(c[1] - 1, 0);
    y1 = min(c[1] + 1, T.nc[1] - 1);
  }
  for (int ix = x0; ix <= x1; ++ix) {
    for (int iy = y0; iy <= y1; ++iy) {
      const uint32_t row =
          DIM == 2 ? T.cell_off + static_cast<uint32_t>(ix) * T.nc[1]
                   : T.cell_off + (static_cast<uint32_t>(ix) * T.nc[1] +
                                   static_cast<uint32_t>(iy)) * T.nc[2];
      const uint32_t b = coff[row + zlo], e = coff[row + zhi + 1];
      for (uint32_t j = b; j < e; ++j) {
        if (j == i) continue;
        double d[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) d[k] = st.x[k][j] - xa[k];
        double s = d[0] * d[0];
#pragma unroll
        for (int k = 1; k < DIM; ++k) s = s + d[k] * d[k];
        if (!(s <= h2)) continue;
        const uint32_t jd = id[j];
        if (!(s < bd[kKnn - 1] || (s == bd[kKnn - 1] && jd < bid[kKnn - 1]))) continue;
        // insert, keeping (bd, bid) ascending
        double cs = s;
        uint32_t cid = jd, cj = j;
#pragma unroll
        for (int q = 0; q < kKnn; ++q) {
          const bool before = cs < bd[q] || (cs == bd[q] && cid < bid[q]);
          if (before) {
            const double td = bd[q];
            const uint32_t ti = bid[q], tj = bj[q];
            bd[q] = cs;
            bid[q] = cid;
            bj[q] = cj;
            cs = td;
            cid = ti;
            cj = tj;
          }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kKnn; ++q) nbr[static_cast<size_t>(i) * kKnn + q] = bj[q];
}

// ---- features -------------------------------------------------------------------

// Node input rows [nf | 0] (BF16, kNodeIn wide).
template <int DIM>
__global__ __launch_bounds__(kBlk) void k_node_feat(DevParams P, LearnedScales sc,
                                                    uint32_t n, DevState st,
                                                    const uint32_t* __restrict__ tag,
                                                    const uint32_t* __restrict__ id,
                                                    const double* __restrict__ normal,
                                                    bf16* __restrict__ H) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  float f[kNodeIn];
#pragma unroll
  for (int k = 0; k < kNodeIn; ++k) f[k] = 0.0f;
  const float inv_vs = 1.0f / sc.v_s;
#pragma unroll
  for (int k = 0; k < DIM; ++k) f[k] = static_cast<float>(st.v[k][i]) * inv_vs;
  f[3] = static_cast<float>((st.rho[i] * P.inv_rho0 - 1.0) / sc.rho_s);
  const bool fixed = tag[i] & 1u;
  f[4] = fixed ? 1.0f : 0.0f;
  if (fixed) {
    const double* nr = normal + static_cast<size_t>(id[i]) * DIM;
    double n2 = 0.0;
#pragma unroll
    for (int k = 0; k < DIM; ++k) n2 += nr[k] * nr[k];
    const double inv = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
#pragma unroll
    for (int k = 0; k < DIM; ++k) f[5 + k] = static_cast<float>(nr[k] * inv);
  }
  bf16* row = H + static_cast<size_t>(i) * kNodeIn;
#pragma unroll
  for (int k = 0; k < kNodeIn; ++k) row[k] = static_cast<bf16>(f[k]);
}

template <int DIM, int KERNEL>
__global__ __launch_bounds__(kBlk) void k_edge_feat(DevParams P, LearnedScales sc,
                                                    uint32_t n, DevState st,
                                                    const uint32_t* __restrict__ nbr,
                                                    bf16* __restrict__ ef) {
  const size_t e = static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x;
  if (e >= static_cast<size_t>(n) * kKnn) return;
  const uint32_t a = static_cast<uint32_t>(e / kKnn), b = nbr[e];
  float f[kEdgeFeat];
#pragma unroll
  for (int k = 0; k < kEdgeFeat; ++k) f[k] = 0.0f;
  if (b != a) {
    const float inv_h = static_cast<float>(1.0 / P.h), inv_vs = 1.0f / sc.v_s;
    float r2 = 0.0f;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const float d = static_cast<float>(st.x[k][b] - st.x[k][a]) * inv_h;
      f[k] = d;
      r2 += d * d;
      f[5 + k] = static_cast<float>(st.v[k][b] - st.v[k][a]) * inv_vs;
    }
    const float q = sqrtf(r2);
    f[3] = q;
    f[4] = kernel_value_poly<KERNEL>(q) / kernel_value_poly<KERNEL>(0.0f);
    f[8] = static_cast<float>((st.rho[b] - st.rho[a]) * P.inv_rho0 / sc.rho_s);
    f[9] = 1.0f;
  }
  bf16* row = ef + e * kEdgeFeat;
#pragma unroll
  for (int k = 0; k < kEdgeFeat; ++k) row[k] = static_cast<bf16>(f[k]);
}

// x_e = [ef_e | H_a | H_b].
__global__ __launch_bounds__(kBlk) void k_gather_edges(uint32_t n, const bf16* __restrict__ ef,
                                                       const bf16* __restrict__ H,
                                                       const uint32_t* __restrict__ nbr,
                                                       bf16* __restrict__ X) {
  // one thread per (edge, 8-element group); kEdgeIn / 8 = 10 groups
  constexpr int G = kEdgeIn / 8;
  const size_t t = static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x;
  if (t >= static_cast<size_t>(n) * kKnn * G) return;
  const size_t e = t / G;
  const int g = static_cast<int>(t % G);
  const uint4* src;
  if (g < kEdgeFeat / 8) {
    src = reinterpret_cast<const uint4*>(ef + e * kEdgeFeat) + g;
  } else if (g < (kEdgeFeat + kNodeIn) / 8) {
    src = reinterpret_cast<const uint4*>(H + (e / kKnn) * kNodeIn) + (g - kEdgeFeat / 8);
  } else {
    src = reinterpret_cast<const uint4*>(H + static_cast<size_t>(nbr[e]) * kNodeIn) +
          (g - (kEdgeFeat + kNodeIn) / 8);
  }
  reinterpret_cast<uint4*>(X + e * kEdgeIn)[g] = *src;
}

// dst[m][col0 + k] = src[m][k] for k < width (BF16, width a multiple of 8).
__global__ __launch_bounds__(kBlk) void k_copy_bf16_cols(uint32_t M, int width,
                                                         const bf16* __restrict__ src,
                                                         int lds, bf16* __restrict__ dst,
                                                         int ldd, int col0) {
  const int G = width / 8;
  const size_t t = static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x;
  if (t >= static_cast<size_t>(M) * G) return;
  const size_t m = t / G;
  const int g = static_cast<int>(t % G);
  *reinterpret_cast<uint4*>(dst + m * ldd + col0 + 8 * g) =
      *reinterpret_cast<const uint4*>(src + m * lds + 8 * g);
}

// ---- loss -------------------------------------------------------------------------

// L = sum over fluid rows and used channels of (u - y / s)^2 / count; du the
// gradient (zero elsewhere).  Channels: acceleration 0..dim-1 and 3.
__global__ __launch_bounds__(kBlk) void k_loss(uint32_t n, int dim, const float* __restrict__ u,
                                               const float* __restrict__ target,
                                               const uint8_t* __restrict__ fluid,
                                               LearnedScales sc, float inv_count,
                                               float* __restrict__ du,
                                               double* __restrict__ loss) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  float l = 0.0f;
  if (i < n) {
    float* g = du + static_cast<size_t>(i) * kOut;
#pragma unroll
    for (int c = 0; c < kOut; ++c) g[c] = 0.0f;
    if (fluid[i]) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c >= dim && c != 3) continue;
        const float r = u[static_cast<size_t>(i) * kOut + c] -
                        target[static_cast<size_t>(i) * 4 + c] / sc.out_scale[c];
        l += r * r;
        g[c] = 2.0f * r * inv_count;
      }
    }
  }
  // block reduction, one FP64 atomic per block
  __shared__ float red[kBlk / 64];
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.0f;
    for (int w = 0; w < kBlk / 64; ++w) s += red[w];
    atomicAdd(loss, static_cast<double>(s) * inv_count);
  }
}

// du_prev[a][c] = dY[a][16 + c] + sum_k dX[a kKnn + k][32 + c].
__global__ __launch_bounds__(kBlk) void k_latent_grad_gather(uint32_t n, const float* __restrict__ dY,
                                                             const float* __restrict__ dX,
                                                             float* __restrict__ du) {
  const size_t t = static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x;
  if (t >= static_cast<size_t>(n) * kLatent) return;
  const size_t a = t / kLatent;
  const int c = static_cast<int>(t % kLatent);
  float s = dY[a * kNodeCat + kNodeFeat + c];
  for (int k = 0; k < kKnn; ++k) s += dX[(a * kKnn + k) * kEdgeIn + kEdgeFeat + kNodeFeat + c];
  du[t] = s;
}

// du_prev[nbr[e]][c] += dX[e][64 + c].
__global__ __launch_bounds__(kBlk) void k_latent_grad_scatter(uint32_t n, const float* __restrict__ dX,
                                                              const uint32_t* __restrict__ nbr,
                                                              float* __restrict__ du) {
  const size_t t = static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x;
  if (t >= static_cast<size_t>(n) * kKnn * kLatent) return;
  const size_t e = t / kLatent;
  const int c = static_cast<int>(t % kLatent);
  atomicAdd(du + static_cast<size_t>(nbr[e]) * kLatent + c,
            dX[e * kEdgeIn + kEdgeFeat + kNodeIn + kNodeFeat + c]);
}

// ---- learned update ---------------------------------------------------------------

// Fluid: v' = v + T a, r' = r + T (v + v') / 2, rho' = rho + T drho/dt with
// T = factor dt of the particle's tank (FP64); then the time-step candidate of
// the new state (compute_time_step, fluid_equations.hpp:199-222, with the
// predicted acceleration) into the tank's pending minimum.
template <int DIM>
__global__ __launch_bounds__(kBlk) void k_learned_update(DevParams P, LearnedScales sc,
                                                         uint32_t n, DevState st,
                                                         const uint32_t* __restrict__ tag,
                                                         const double* __restrict__ dt,
                                                         double factor,
                                                         const float* __restrict__ u,
                                                         unsigned long long* __restrict__ dt_min,
                                                         uint32_t* __restrict__ vmax) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  bool fluid = false;
  uint32_t t = 0;
  unsigned long long cand = ~0ull;
  float vm = 0.0f;  // largest speed over the step (displacement bound)
  if (i < n) {
    const uint32_t tg = tag[i];
    t = tg >> 1;
    fluid = !(tg & 1u);
    if (fluid) {
      const double T = factor * dt[t];
      const float* ui = u + static_cast<size_t>(i) * kOut;
      double a[DIM], v1[DIM];
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        a[k] = static_cast<double>(ui[k]) * sc.out_scale[k];
        const double v0 = st.v[k][i];
        v1[k] = v0 + T * a[k];
        st.x[k][i] = st.x[k][i] + T * (0.5 * (v0 + v1[k]));
        st.v[k][i] = v1[k];
        s0 += v0 * v0;
        s1 += v1[k] * v1[k];
      }
      // the position moves by T (v0 + v1) / 2: |.| <= T max(|v0|, |v1|)
      vm = static_cast<float>(sqrt(fmax(s0, s1))) * (1.0f + 1e-6f);
      const double rho = st.rho[i] + T * (static_cast<double>(ui[3]) * sc.out_scale[3]);
      st.rho[i] = rho;
      write_pair_data<DIM>(P, st, i, v1, rho, P.m);  // uniform masses only
      cand = dbl_bits(particle_time_step<DIM>(P, rho, v1, a));
    }
  }
  // wave minimum when the wave's fluid lanes share a tank, then one atomic
  // per workgroup and tank (on the workgroup's copy of the tank's slots)
  __shared__ unsigned long long s_min[kBlk / 64];
  __shared__ float s_max[kBlk / 64];
  __shared__ uint32_t s_t[kBlk / 64];
  const int w = threadIdx.x >> 6;
  const uint64_t act = __ballot(fluid);
  const int leader = act ? __ffsll(static_cast<long long>(act)) - 1 : 0;
  const uint32_t t0 = __shfl(t, leader);
  const bool uniform = act != 0 && __ballot(fluid && t != t0) == 0;
  if (act != 0 && !uniform && fluid) {
    atomicMin(red_slot(dt_min, t), cand);
    atomicMax(red_slot(vmax, t), __float_as_uint(vm));
  }
  unsigned long long m = uniform ? cand : ~0ull;
  float mx = uniform ? vm : 0.0f;
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long q = __shfl_xor(m, o);
    m = q < m ? q : m;
    mx = fmaxf(mx, __shfl_xor(mx, o));
  }
  if ((threadIdx.x & 63) == 0) {
    s_min[w] = m;
    s_max[w] = mx;
    s_t[w] = uniform ? t0 : ~0u;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int a = 0; a < kBlk / 64; ++a) {
      if (s_t[a] == ~0u) continue;
      unsigned long long mm = s_min[a];
      float xx = s_max[a];
      for (int b = a + 1; b < kBlk / 64; ++b) {
        if (s_t[b] == s_t[a]) {
          mm = s_min[b] < mm ? s_min[b] : mm;
          xx = fmaxf(xx, s_max[b]);
          s_t[b] = ~0u;
        }
      }
      // the workgroup's copy of the tank's slots
      const size_t slot = static_cast<size_t>(s_t[a]) * kRedStripes + (blockIdx.x & (kRedStripes - 1));
      atomicMin(dt_min + slot, mm);
      atomicMax(vmax + slot, __float_as_uint(xx));
    }
  }
}


// ---- judge ------------------------------------------------------------------------

// Keeps the state's density, x and velocity by particle id.
template <int DIM>
__global__ __launch_bounds__(kBlk) void k_keep_by_id(uint32_t n, DevState st,
                                                     const uint32_t* __restrict__ id,
                                                     double* __restrict__ rho_by_id,
                                                     double* __restrict__ x_by_id,
                                                     double* __restrict__ v_by_id) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = id[i];
  if (rho_by_id) rho_by_id[j] = st.rho[i];
  if (x_by_id) x_by_id[j] = st.x[0][i];
#pragma unroll
  for (int k = 0; k < DIM; ++k) v_by_id[static_cast<size_t>(j) * DIM + k] = st.v[k][i];
}

__device__ __forceinline__ unsigned long long ordered_bits(double x) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Judge of one learned step L against the plain reference P (the current
// state), both from the start state S, over fluid particles, each result in
// kRedStripes copies (one atomic per wave on the wave's copy; the host folds):
//   err_bits[s]: max |rho_L - rho_P| / rho_P (FP32 bits)
//   front[s], front[kRedStripes + s]: largest x of P and of L (ordered bits)
//   dv[s] += |v_L - v_P|^2, dv[kRedStripes + s] += |v_P - v_S|^2
template <int DIM>
__global__ __launch_bounds__(kBlk) void k_judge(uint32_t n, DevState st,
                                                const uint32_t* __restrict__ id,
                                                const uint32_t* __restrict__ tag,
                                                const double* __restrict__ rho_l,
                                                const double* __restrict__ x_l,
                                                const double* __restrict__ v_l,
                                                const double* __restrict__ v_s,
                                                uint32_t* __restrict__ err_bits,
                                                unsigned long long* __restrict__ front,
                                                double* __restrict__ dv) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  double num = 0.0, den = 0.0;
  uint32_t eb = 0u;                   // below every non-negative error
  unsigned long long f0 = 0, f1 = 0;  // below every ordered double
  if (i < n && !(tag[i] & 1u)) {
    const uint32_t j = id[i];
    const double rp = st.rho[i];
    eb = __float_as_uint(static_cast<float>(fabs(rho_l[j] - rp) / rp));
    f0 = ordered_bits(st.x[0][i]);
    f1 = ordered_bits(x_l[j]);
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const double vp = st.v[k][i];
      const double a = v_l[static_cast<size_t>(j) * DIM + k] - vp;
      const double b = vp - v_s[static_cast<size_t>(j) * DIM + k];
      num += a * a;
      den += b * b;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    num += __shfl_xor(num, o);
    den += __shfl_xor(den, o);
    const uint32_t e2 = __shfl_xor(eb, o);
    const unsigned long long g0 = __shfl_xor(f0, o), g1 = __shfl_xor(f1, o);
    eb = e2 > eb ? e2 : eb;
    f0 = g0 > f0 ? g0 : f0;
    f1 = g1 > f1 ? g1 : f1;
  }
  if ((threadIdx.x & 63) == 0) {
    const uint32_t s = red_stripe();
    atomicMax(err_bits + s, eb);
    atomicMax(front + s, f0);
    atomicMax(front + kRedStripes + s, f1);
    atomicAdd(dv + s, num);
    atomicAdd(dv + kRedStripes + s, den);
  }
}

// ---- recording --------------------------------------------------------------------

template <int DIM>
__global__ __launch_bounds__(kBlk) void k_targets(uint32_t n, const uint32_t* __restrict__ id0,
                                                  const uint32_t* __restrict__ tag0,
                                                  DevState s0, const double* __restrict__ v1,
                                                  const double* __restrict__ rho1,
                                                  const double* __restrict__ t0,
                                                  const double* __restrict__ t1,
                                                  float* __restrict__ target,
                                                  uint8_t* __restrict__ fluid) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  const uint32_t tg = tag0[i], j = id0[i];
  float* y = target + static_cast<size_t>(i) * 4;
  y[0] = y[1] = y[2] = y[3] = 0.0f;
  fluid[i] = (tg & 1u) ? 0 : 1;
  if (tg & 1u) return;
  const double T = t1[tg >> 1] - t0[tg >> 1];
  const double inv = T > 0.0 ? 1.0 / T : 0.0;
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    y[k] = static_cast<float>((v1[static_cast<size_t>(j) * DIM + k] - s0.v[k][i]) * inv);
  }
  y[3] = static_cast<float>((rho1[j] - s0.rho[i]) * inv);
}

// Export helpers (sorted order -> id order).
__global__ __launch_bounds__(kBlk) void k_export_graph(uint32_t n, const uint32_t* __restrict__ id,
                                                       const uint32_t* __restrict__ nbr,
                                                       const bf16* __restrict__ H,
                                                       const bf16* __restrict__ ef,
                                                       uint32_t* __restrict__ nbr_ref,
                                                       float* __restrict__ nf_ref,
                                                       float* __restrict__ ef_ref) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = id[i];
  for (int k = 0; k < kKnn; ++k) {
    const size_t e = static_cast<size_t>(i) * kKnn + k;
    if (nbr_ref) nbr_ref[static_cast<size_t>(j) * kKnn + k] = id[nbr[e]];
    if (ef_ref) {
      for (int c = 0; c < kEdgeFeat; ++c) {
        ef_ref[(static_cast<size_t>(j) * kKnn + k) * kEdgeFeat + c] =
            static_cast<float>(ef[e * kEdgeFeat + c]);
      }
    }
  }
  if (nf_ref) {
    for (int c = 0; c < kNodeFeat; ++c) {
      nf_ref[static_cast<size_t>(j) * kNodeFeat + c] =
          static_cast<float>(H[static_cast<size_t>(i) * kNodeIn + c]);
    }
  }
}

__global__ __launch_bounds__(kBlk) void k_export_out(uint32_t n, const uint32_t* __restrict__ id,
                                                     const float* __restrict__ u,
                                                     float* __restrict__ out) {
  const size_t t = static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x;
  if (t >= static_cast<size_t>(n) * kOut) return;
  const size_t i = t / kOut;
  out[static_cast<size_t>(id[i]) * kOut + t % kOut] = u[t];
}

__global__ void k_to_bf16(size_t n, const float* __restrict__ src, bf16* __restrict__ dst) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = static_cast<bf16>(src[i]);
}

}  // namespace

// ---- learned-step state ---------------------------------------------------------

// Activations of one forward pass over n nodes (E = kKnn n edges).  The
// training form keeps every round's pre-activations for the backward pass.
struct Activations {
  uint32_t cap = 0;  // nodes
  bool training = false;
  bf16* H[kRounds + 1] = {};     // node inputs per round, n x kNodeIn
  bf16* X[kRounds] = {};         // edge inputs, E x kEdgeIn
  float* pre1[kRounds] = {};     // E x kHidden (training)
  bf16* a1[kRounds] = {};        // E x kHidden
  float* pre2[kRounds] = {};     // E x kHidden (training)
  float* msg = nullptr;          // E x kHidden (one round at a time)
  bf16* Y[kRounds] = {};         // n x kNodeCat
  float* pre3[kRounds] = {};     // n x kHidden (training)
  bf16* a3[kRounds] = {};
  float* pre4[kRounds] = {};     // (training)
  bf16* a4[kRounds] = {};
  float* u[kRounds] = {};        // n x kOut
  // backward scratch
  float* dE = nullptr;           // E x kHidden
  bf16* dE16 = nullptr;
  float* dX = nullptr;           // E x kEdgeIn
  float* dN = nullptr;           // n x kHidden
  bf16* dN16 = nullptr;
  float* dY = nullptr;           // n x kNodeCat
  float* du = nullptr;           // n x kOut
  bf16* du16 = nullptr;
  std::vector<void*> mem;
};

struct Sample {
  uint32_t n = 0;
  uint32_t* nbr = nullptr;  // n x kKnn (sample index space)
  bf16* H0 = nullptr;       // n x kNodeIn, [nf | 0]
  bf16* ef = nullptr;       // n kKnn x kEdgeFeat
  float* target = nullptr;  // n x 4 (raw: acceleration, density rate)
  uint8_t* fluid = nullptr;
  uint32_t n_fluid = 0;
  std::vector<void*> mem;
};

// Everything a plain step reads besides the parameters: the state in cell
// order with its cell index (keys, offsets, the cells holding fluid, chunk
// lists, the fixed particles and those with fluid near), the clocks and the
// rebuild bookkeeping.  Restoring it makes the following
// plain steps repeat bit for bit what they would have done.
struct Snapshot {
  DevState s{};
  uint32_t* id = nullptr;
  uint32_t* tag = nullptr;
  uint32_t* key = nullptr;
  uint32_t* cell_off = nullptr;
  uint4* chunks = nullptr;
  uint32_t* chunk_counts = nullptr;
  uint32_t* fixed_list = nullptr;
  uint32_t* fixed_active = nullptr;
  uint8_t* cell_fluid = nullptr;
  uint4* bchunks = nullptr;  // 3D block chunks and their per-block counts
  uint32_t* bcount = nullptr;
  uint32_t* counter = nullptr;
  uint32_t* ctl = nullptr;
  uint32_t* need = nullptr;
  uint32_t* vmax = nullptr;
  double* disp = nullptr;
  double* time = nullptr;
  double* dt = nullptr;
  unsigned long long* dt_min = nullptr;  // pending step size of each tank
  size_t chunk_cap = 0, bchunk_cap = 0;
  uint32_t n_blocks = 0;
  bool alloc = false;
};

struct Learned {
  float* w32 = nullptr;
  bf16* w16 = nullptr;
  bf16* frag = nullptr;      // MFMA fragments of w16 for the fused rounds
  bool fused = true;         // inference through the fused round kernel
  float* grad = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  long adam_t = 0;
  LearnedScales sc{};
  bool scales_set = false;
  // graph and features of the context's current state
  uint32_t cap = 0;
  uint32_t* nbr = nullptr;
  bf16* ef = nullptr;
  Activations infer;   // forward only
  Activations train;   // training form
  std::vector<Sample> samples;
  double* loss = nullptr;  // device accumulators
  // judge
  Snapshot accepted, before;
  double* rho_l = nullptr;
  double* x_l = nullptr;
  double* v_l = nullptr;     // n x dim, by id
  double* v_s = nullptr;     // n x dim, by id (judged step's start)
  double* dv = nullptr;      // 2 sums
  uint32_t* err_bits = nullptr;
  unsigned long long* front = nullptr;
  // recording
  double* v_ref = nullptr;
  double* rho_ref = nullptr;
  double* r_ref = nullptr;
  double* p_ref = nullptr;
  double* t0 = nullptr;
  std::vector<void*> mem;    // parameters, optimiser state, scalars
  std::vector<void*> mem_n;  // buffers sized by the context's particle count
  uint32_t sized_n = 0;
  int sized_tanks = 0;
};

namespace {

template <class T>
hipError_t lalloc(std::vector<void*>& mem, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e == hipSuccess) mem.push_back(q);
  *p = static_cast<T*>(q);
  return e;
}

void free_mem(std::vector<void*>& mem) {
  for (void* p : mem) (void)hipFree(p);
  mem.clear();
}

#define TIT_TRY_E(call)             \
  do {                              \
    hipError_t e_ = (call);         \
    if (e_ != hipSuccess) return e_; \
  } while (0)

hipError_t alloc_acts(Activations& A, uint32_t n, bool training) {
  free_mem(A.mem);
  A = Activations{};
  A.cap = n;
  A.training = training;
  const size_t E = static_cast<size_t>(n) * kKnn;
  auto& m = A.mem;
  for (int r = 0; r <= kRounds; ++r) TIT_TRY_E(lalloc(m, &A.H[r], static_cast<size_t>(n) * kNodeIn));
  const int rounds_kept = training ? kRounds : 1;
  for (int r = 0; r < kRounds; ++r) {
    if (r < rounds_kept) {
      TIT_TRY_E(lalloc(m, &A.X[r], E * kEdgeIn));
      TIT_TRY_E(lalloc(m, &A.a1[r], E * kHidden));
      TIT_TRY_E(lalloc(m, &A.Y[r], static_cast<size_t>(n) * kNodeCat));
      TIT_TRY_E(lalloc(m, &A.a3[r], static_cast<size_t>(n) * kHidden));
      TIT_TRY_E(lalloc(m, &A.a4[r], static_cast<size_t>(n) * kHidden));
      if (training) {
        TIT_TRY_E(lalloc(m, &A.pre1[r], E * kHidden));
        TIT_TRY_E(lalloc(m, &A.pre2[r], E * kHidden));
        TIT_TRY_E(lalloc(m, &A.pre3[r], static_cast<size_t>(n) * kHidden));
        TIT_TRY_E(lalloc(m, &A.pre4[r], static_cast<size_t>(n) * kHidden));
      }
    } else {
      A.X[r] = A.X[0];
      A.a1[r] = A.a1[0];
      A.Y[r] = A.Y[0];
      A.a3[r] = A.a3[0];
      A.a4[r] = A.a4[0];
    }
    TIT_TRY_E(lalloc(m, &A.u[r], static_cast<size_t>(n) * kOut));
  }
  TIT_TRY_E(lalloc(m, &A.msg, E * kHidden));
  if (training) {
    TIT_TRY_E(lalloc(m, &A.dE, E * kHidden));
    TIT_TRY_E(lalloc(m, &A.dE16, E * kHidden));
    TIT_TRY_E(lalloc(m, &A.dX, E * kEdgeIn));
    TIT_TRY_E(lalloc(m, &A.dN, static_cast<size_t>(n) * kHidden));
    TIT_TRY_E(lalloc(m, &A.dN16, static_cast<size_t>(n) * kHidden));
    TIT_TRY_E(lalloc(m, &A.dY, static_cast<size_t>(n) * kNodeCat));
    TIT_TRY_E(lalloc(m, &A.du, static_cast<size_t>(n) * kOut));
    TIT_TRY_E(lalloc(m, &A.du16, static_cast<size_t>(n) * kOut));
  }
  return hipSuccess;
}

// Split-K count for a weight-gradient GEMM over K samples.
int wgrad_splits(size_t K) {
  return static_cast<int>(std::clamp<size_t>(K / 4096, 1, 512));
}

struct RoundW {
  const float* w32;
  const bf16* w16;
  float* g;
};

RoundW round_w(Learned* L, int r) {
  const size_t o = static_cast<size_t>(r) * RoundLayout::size;
  return RoundW{L->w32 + o, L->w16 + o, L->grad ? L->grad + o : nullptr};
}

// Forward of round r: reads A.H[r], writes A.u[r] and, for r + 1 < kRounds,
// A.H[r + 1] = [nf | bf16(u)].
void forward_round(hipStream_t s, Learned* L, Activations& A, int r, uint32_t n,
                   const bf16* ef, const uint32_t* nbr) {
  const size_t E = static_cast<size_t>(n) * kKnn;
  const RoundW W = round_w(L, r);
  constexpr int G = kEdgeIn / 8;
  k_gather_edges<<<nblk(E * G), kBlk, 0, s>>>(n, ef, A.H[r], nbr, A.X[r]);
  GemmArgs g{};
  // edge layer 1
  g.M = static_cast<int>(E);
  g.N = kHidden;
  g.K = kEdgeIn;
  g.A = A.X[r];
  g.lda = kEdgeIn;
  g.B = W.w16 + RoundLayout::W1;
  g.ldb = kHidden;
  g.bias = W.w32 + RoundLayout::b1;
  g.act = kActHswish;
  g.pre = A.training ? A.pre1[r] : nullptr;
  g.ldpre = kHidden;
  g.out16 = A.a1[r];
  g.ldo16 = kHidden;
  gemm_bf16(s, g);
  // edge layer 2 -> messages
  g = GemmArgs{};
  g.M = static_cast<int>(E);
  g.N = kHidden;
  g.K = kHidden;
  g.A = A.a1[r];
  g.lda = kHidden;
  g.B = W.w16 + RoundLayout::W2;
  g.ldb = kHidden;
  g.bias = W.w32 + RoundLayout::b2;
  g.act = kActHswish;
  g.pre = A.training ? A.pre2[r] : nullptr;
  g.ldpre = kHidden;
  g.out32 = A.msg;
  g.ldo32 = kHidden;
  gemm_bf16(s, g);
  // y = [H | sum of messages]
  k_copy_bf16_cols<<<nblk(static_cast<size_t>(n) * (kNodeIn / 8)), kBlk, 0, s>>>(
      n, kNodeIn, A.H[r], kNodeIn, A.Y[r], kNodeCat, 0);
  group_sum(s, static_cast<int>(n), kHidden, kKnn, A.msg, nullptr, A.Y[r], kNodeCat, kNodeIn);
  // node layers
  g = GemmArgs{};
  g.M = static_cast<int>(n);
  g.N = kHidden;
  g.K = kNodeCat;
  g.A = A.Y[r];
  g.lda = kNodeCat;
  g.B = W.w16 + RoundLayout::W3;
  g.ldb = kHidden;
  g.bias = W.w32 + RoundLayout::b3;
  g.act = kActHswish;
  g.pre = A.training ? A.pre3[r] : nullptr;
  g.ldpre = kHidden;
  g.out16 = A.a3[r];
  g.ldo16 = kHidden;
  gemm_bf16(s, g);
  g.K = kHidden;
  g.A = A.a3[r];
  g.lda = kHidden;
  g.B = W.w16 + RoundLayout::W4;
  g.bias = W.w32 + RoundLayout::b4;
  g.pre = A.training ? A.pre4[r] : nullptr;
  g.out16 = A.a4[r];
  gemm_bf16(s, g);
  g = GemmArgs{};
  g.M = static_cast<int>(n);
  g.N = kOut;
  g.K = kHidden;
  g.A = A.a4[r];
  g.lda = kHidden;
  g.B = W.w16 + RoundLayout::W5;
  g.ldb = kOut;
  g.bias = W.w32 + RoundLayout::b5;
  g.act = kActNone;
  g.out32 = A.u[r];
  g.ldo32 = kOut;
  if (r + 1 < kRounds) {
    k_copy_bf16_cols<<<nblk(static_cast<size_t>(n) * (kNodeFeat / 8)), kBlk, 0, s>>>(
        n, kNodeFeat, A.H[r], kNodeIn, A.H[r + 1], kNodeIn, 0);
    g.out16 = A.H[r + 1] + kNodeFeat;
    g.ldo16 = kNodeIn;
  }
  gemm_bf16(s, g);
}

// dW += X^T dY over K rows (X: K x M BF16 row-major, dY: K x N BF16).
void wgrad(hipStream_t s, int M, int N, size_t K, const bf16* X, const bf16* dY,
           float* gW) {
  GemmArgs g{};
  g.M = M;
  g.N = N;
  g.K = static_cast<int>(K);
  g.A = X;
  g.lda = M;
  g.transA = true;
  g.B = dY;
  g.ldb = N;
  g.accumulate = true;
  g.out32 = gW;
  g.ldo32 = N;
  g.k_splits = wgrad_splits(K);
  gemm_bf16(s, g);
}

// dX = dY W^T (dY: M x N BF16, W: K x N BF16 row-major) -> M x K FP32.
void dgrad(hipStream_t s, size_t M, int N, int K, const bf16* dY, const bf16* W,
           float* dX) {
  GemmArgs g{};
  g.M = static_cast<int>(M);
  g.N = K;
  g.K = N;
  g.A = dY;
  g.lda = N;
  g.B = W;
  g.ldb = N;
  g.transB = true;
  g.out32 = dX;
  g.ldo32 = K;
  gemm_bf16(s, g);
}

// Backward of round r from A.du (n x kOut FP32).  For r > 0 leaves the
// gradient of the previous round's output in du_prev.
void backward_round(hipStream_t s, Learned* L, Activations& A, int r, uint32_t n,
                    const uint32_t* nbr, float* du_prev) {
  const size_t E = static_cast<size_t>(n) * kKnn;
  const RoundW W = round_w(L, r);
  float* g = W.g;
  // output layer
  k_to_bf16<<<nblk(static_cast<size_t>(n) * kOut), kBlk, 0, s>>>(
      static_cast<size_t>(n) * kOut, A.du, A.du16);
  column_sum(s, static_cast<int>(n), kOut, A.du, g + RoundLayout::b5);
  wgrad(s, kHidden, kOut, n, A.a4[r], A.du16, g + RoundLayout::W5);
  dgrad(s, n, kOut, kHidden, A.du16, W.w16 + RoundLayout::W5, A.dN);
  // node layer 2
  act_backward(s, static_cast<int>(n), kHidden, A.dN, kHidden, 1, A.pre4[r], A.dN, A.dN16);
  column_sum(s, static_cast<int>(n), kHidden, A.dN, g + RoundLayout::b4);
  wgrad(s, kHidden, kHidden, n, A.a3[r], A.dN16, g + RoundLayout::W4);
  dgrad(s, n, kHidden, kHidden, A.dN16, W.w16 + RoundLayout::W4, A.dN);
  // node layer 1
  act_backward(s, static_cast<int>(n), kHidden, A.dN, kHidden, 1, A.pre3[r], A.dN, A.dN16);
  column_sum(s, static_cast<int>(n), kHidden, A.dN, g + RoundLayout::b3);
  wgrad(s, kNodeCat, kHidden, n, A.Y[r], A.dN16, g + RoundLayout::W3);
  dgrad(s, n, kHidden, kNodeCat, A.dN16, W.w16 + RoundLayout::W3, A.dY);
  // edge layer 2 (message gradient = dY[:, kNodeIn:] broadcast to the edges)
  act_backward(s, static_cast<int>(E), kHidden, A.dY + kNodeIn, kNodeCat, kKnn,
                A.pre2[r], A.dE, A.dE16);
  column_sum(s, static_cast<int>(E), kHidden, A.dE, g + RoundLayout::b2);
  wgrad(s, kHidden, kHidden, E, A.a1[r], A.dE16, g + RoundLayout::W2);
  dgrad(s, E, kHidden, kHidden, A.dE16, W.w16 + RoundLayout::W2, A.dE);
  // edge layer 1
  act_backward(s, static_cast<int>(E), kHidden, A.dE, kHidden, 1, A.pre1[r], A.dE, A.dE16);
  column_sum(s, static_cast<int>(E), kHidden, A.dE, g + RoundLayout::b1);
  wgrad(s, kEdgeIn, kHidden, E, A.X[r], A.dE16, g + RoundLayout::W1);
  if (r > 0) {
    dgrad(s, E, kHidden, kEdgeIn, A.dE16, W.w16 + RoundLayout::W1, A.dX);
    k_latent_grad_gather<<<nblk(static_cast<size_t>(n) * kLatent), kBlk, 0, s>>>(
        n, A.dY, A.dX, du_prev);
    k_latent_grad_scatter<<<nblk(E * kLatent), kBlk, 0, s>>>(n, A.dX, nbr, du_prev);
  }
}

// Loss and parameter gradient of one sample (gradient accumulated into
// L->grad, which the caller clears).
hipError_t loss_and_grad(hipStream_t s, Learned* L, int dim, const Sample& S,
                         double* loss) {
  Activations& A = L->train;
  TIT_TRY_E(hipMemcpyAsync(A.H[0], S.H0, static_cast<size_t>(S.n) * kNodeIn * sizeof(bf16),
                           hipMemcpyDeviceToDevice, s));
  for (int r = 0; r < kRounds; ++r) forward_round(s, L, A, r, S.n, S.ef, S.nbr);
  const float inv_count =
      1.0f / static_cast<float>(std::max<uint32_t>(S.n_fluid, 1) * (dim + 1));
  k_loss<<<nblk(S.n), kBlk, 0, s>>>(S.n, dim, A.u[kRounds - 1], S.target, S.fluid, L->sc,
                                    inv_count, A.du, loss);
  for (int r = kRounds - 1; r >= 0; --r) {
    float* du_prev = r > 0 ? A.u[r - 1] : nullptr;  // u[r-1] is no longer needed
    backward_round(s, L, A, r, S.n, S.nbr, du_prev);
    if (r > 0) {
      TIT_TRY_E(hipMemcpyAsync(A.du, du_prev, static_cast<size_t>(S.n) * kOut * sizeof(float),
                               hipMemcpyDeviceToDevice, s));
    }
  }
  return hipSuccess;
}

// splitmix64 -> uniform [0, 1)
struct Rng {
  uint64_t s;
  double next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z = z ^ (z >> 31);
    return static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
  }
};

void init_weights(std::vector<float>& w, uint64_t seed) {
  w.assign(kParams, 0.0f);
  Rng rng{seed};
  auto fill = [&](size_t off, int in, int out, double gain) {
    const double a = gain * std::sqrt(6.0 / (in + out));
    for (size_t k = 0; k < static_cast<size_t>(in) * out; ++k) {
      w[off + k] = static_cast<float>((2.0 * rng.next() - 1.0) * a);
    }
  };
  for (int r = 0; r < kRounds; ++r) {
    const size_t o = static_cast<size_t>(r) * RoundLayout::size;
    fill(o + RoundLayout::W1, kEdgeIn, kHidden, 1.0);
    fill(o + RoundLayout::W2, kHidden, kHidden, 1.0);
    fill(o + RoundLayout::W3, kNodeCat, kHidden, 1.0);
    fill(o + RoundLayout::W4, kHidden, kHidden, 1.0);
    fill(o + RoundLayout::W5, kHidden, kOut, 0.1);
  }
}

}  // namespace

void destroy_learned(tit_hip_ctx* c) {
  if (!c->learned) return;
  Learned* L = c->learned;
  for (Sample& S : L->samples) free_mem(S.mem);
  free_mem(L->infer.mem);
  free_mem(L->train.mem);
  free_mem(L->mem_n);
  free_mem(L->mem);
  delete L;
  c->learned = nullptr;
}

namespace {

// Checks that the learned step exists and that its particle-count-sized
// buffers match the uploaded particle set (freed on a change of size; the
// weights, scales and recorded samples are kept).
tit_status need_learned(tit_hip_ctx* c) {
  if (!c) return TIT_E_ARG;
  if (!c->learned) return fail(c, TIT_E_STATE, "learned step not initialised");
  if (!c->uploaded) return fail(c, TIT_E_STATE, "no particles uploaded");
  if (c->mass) return fail(c, TIT_E_ARG, "the learned step needs uniform masses");
  Learned* L = c->learned;
  if (L->sized_n != c->n || L->sized_tanks != c->n_tanks) {
    TIT_HIP_TRY(c, hipSetDevice(c->device));
    TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
    free_mem(L->mem_n);
    free_mem(L->infer.mem);
    L->infer = Activations{};
    L->cap = 0;
    L->nbr = nullptr;
    L->ef = nullptr;
    L->accepted = Snapshot{};
    L->before = Snapshot{};
    L->v_ref = L->rho_ref = L->r_ref = L->p_ref = L->t0 = nullptr;
    TIT_HIP_TRY(c, lalloc(L->mem_n, &L->rho_l, c->n));
    TIT_HIP_TRY(c, lalloc(L->mem_n, &L->x_l, c->n));
    TIT_HIP_TRY(c, lalloc(L->mem_n, &L->v_l, static_cast<size_t>(c->n) * c->prm.dim));
    TIT_HIP_TRY(c, lalloc(L->mem_n, &L->v_s, static_cast<size_t>(c->n) * c->prm.dim));
    L->sized_n = c->n;
    L->sized_tanks = c->n_tanks;
  }
  return TIT_OK;
}

LearnedScales default_scales(const tit_hip_ctx* c) {
  LearnedScales s{};
  s.v_s = static_cast<float>(c->prm.cs0 / 10.0);
  s.rho_s = 0.01f;
  for (int k = 0; k < 4; ++k) s.out_scale[k] = 1.0f;
  return s;
}

// Graph + features of the context's current (freshly rebuilt) state.
tit_status build_graph(tit_hip_ctx* c) {
  Learned* L = c->learned;
  hipStream_t s = c->stream;
  const uint32_t n = c->n;
  if (L->cap < n) {
    TIT_HIP_TRY(c, lalloc(L->mem_n, &L->nbr, static_cast<size_t>(n) * kKnn));
    TIT_HIP_TRY(c, lalloc(L->mem_n, &L->ef, static_cast<size_t>(n) * kKnn * kEdgeFeat));
    L->cap = n;
  }
  if (L->infer.cap < n) TIT_HIP_TRY(c, alloc_acts(L->infer, n, false));
  const DevParams& P = c->dp;
  const size_t E = static_cast<size_t>(n) * kKnn;
  ProfScope ps(c->prof, s, "learned_graph");
  if (P.dim == 2) {
    k_knn<2><<<nblk(n), kBlk, 0, s>>>(P, c->tanks, n, c->A, c->key, c->tag, c->id,
                                      c->cell_off, L->nbr);
    k_node_feat<2><<<nblk(n), kBlk, 0, s>>>(P, L->sc, n, c->A, c->tag, c->id, c->normal,
                                            L->infer.H[0]);
    if (P.kernel == kKernelC2) {
      k_edge_feat<2, kKernelC2><<<nblk(E), kBlk, 0, s>>>(P, L->sc, n, c->A, L->nbr, L->ef);
    } else {
      k_edge_feat<2, kKernelC4><<<nblk(E), kBlk, 0, s>>>(P, L->sc, n, c->A, L->nbr, L->ef);
    }
  } else {
    k_knn<3><<<nblk(n), kBlk, 0, s>>>(P, c->tanks, n, c->A, c->key, c->tag, c->id,
                                      c->cell_off, L->nbr);
    k_node_feat<3><<<nblk(n), kBlk, 0, s>>>(P, L->sc, n, c->A, c->tag, c->id, c->normal,
                                            L->infer.H[0]);
    if (P.kernel == kKernelC2) {
      k_edge_feat<3, kKernelC2><<<nblk(E), kBlk, 0, s>>>(P, L->sc, n, c->A, L->nbr, L->ef);
    } else {
      k_edge_feat<3, kKernelC4><<<nblk(E), kBlk, 0, s>>>(P, L->sc, n, c->A, L->nbr, L->ef);
    }
  }
  return check_launch(c, "graph");
}

// Rebuild, boundary values, graph, features and the network's output on the
// current state (in L->infer.u[kRounds - 1]).
tit_status infer_current(tit_hip_ctx* c, RebuildMode mode) {
  enqueue_rebuild(c, mode);
  enqueue_boundary(c, c->A);
  tit_status st = build_graph(c);
  if (st != TIT_OK) return st;
  Learned* L = c->learned;
  Activations& A = L->infer;
  ProfScope ps(c->prof, c->stream, "learned_network");
  if (L->fused) {
    for (int r = 0; r < kRounds; ++r) {
      const bool last = r + 1 == kRounds;
      launch_round_fused(c->stream, c->n_cu, c->n, r, L->frag, L->w32, L->nbr, L->ef, A.H[r],
                         last ? nullptr : A.H[r + 1], last ? A.u[r] : nullptr);
    }
  } else {
    for (int r = 0; r < kRounds; ++r) forward_round(c->stream, L, A, r, c->n, L->ef, L->nbr);
  }
  return check_launch(c, "learned forward");
}

// One learned step of 4 dt, dt being the tank's pending step size (taken as a
// plain step takes it); leaves the next pending step size.
tit_status learned_step(tit_hip_ctx* c) {
  constexpr double kFactor = 4.0;
  launch_step_begin(c->stream, c->n_tanks, c->dt_min, c->dt, c->vmax, c->disp,
                    c->dp.skin_half, rebuild_ctl(c), kFactor);
  tit_status st = infer_current(c, kRebuildMarked);
  if (st != TIT_OK) return st;
  Learned* L = c->learned;
  ProfScope ps(c->prof, c->stream, "learned_update");
  if (c->dp.dim == 2) {
    k_learned_update<2><<<nblk(c->n), kBlk, 0, c->stream>>>(
        c->dp, L->sc, c->n, c->A, c->tag, c->dt, kFactor, L->infer.u[kRounds - 1],
        c->dt_min, c->vmax);
  } else {
    k_learned_update<3><<<nblk(c->n), kBlk, 0, c->stream>>>(
        c->dp, L->sc, c->n, c->A, c->tag, c->dt, kFactor, L->infer.u[kRounds - 1],
        c->dt_min, c->vmax);
  }
  // clock and displacement bound of the step (the next step rebuilds the
  // cell index if the bound requires)
  launch_step_end(c->stream, c->n_tanks, c->dt, c->vmax, c->time, c->disp, c->dp.skin_half,
                  c->counter + 1, kFactor);
  return check_launch(c, "learned update");
}

hipError_t alloc_snapshot(tit_hip_ctx* c, Snapshot& S) {
  if (S.alloc && S.chunk_cap == c->chunk_cap && S.bchunk_cap == c->bchunk_cap &&
      S.n_blocks == c->n_blocks) {
    return hipSuccess;
  }
  Learned* L = c->learned;
  auto& m = L->mem_n;
  if (!S.alloc) {
    for (int k = 0; k < 3; ++k) S.s.x[k] = S.s.v[k] = nullptr;
    for (int k = 0; k < c->prm.dim; ++k) {
      TIT_TRY_E(lalloc(m, &S.s.x[k], c->n));
      TIT_TRY_E(lalloc(m, &S.s.v[k], c->n));
    }
    TIT_TRY_E(lalloc(m, &S.s.rho, c->n));
    TIT_TRY_E(lalloc(m, &S.s.aux, c->n));
    TIT_TRY_E(lalloc(m, &S.s.pv, c->n));
    S.s.vol = nullptr;
    if (c->A.vol) TIT_TRY_E(lalloc(m, &S.s.vol, c->n));
    TIT_TRY_E(lalloc(m, &S.id, c->n));
    TIT_TRY_E(lalloc(m, &S.tag, c->n));
    TIT_TRY_E(lalloc(m, &S.key, c->n));
    TIT_TRY_E(lalloc(m, &S.cell_off, static_cast<size_t>(c->n_cells) + 1));
    TIT_TRY_E(lalloc(m, &S.chunk_counts, 4));
    TIT_TRY_E(lalloc(m, &S.fixed_list, std::max<uint32_t>(c->n_fixed, 1)));
    TIT_TRY_E(lalloc(m, &S.fixed_active, std::max<uint32_t>(c->n_fixed, 1)));
    TIT_TRY_E(lalloc(m, &S.cell_fluid, std::max<uint32_t>(c->n_cells, 1)));
    TIT_TRY_E(lalloc(m, &S.counter, 4));
    TIT_TRY_E(lalloc(m, &S.ctl, 4));
    TIT_TRY_E(lalloc(m, &S.need, c->n_tanks));
    TIT_TRY_E(lalloc(m, &S.vmax, static_cast<size_t>(c->n_tanks) * kRedStripes));
    TIT_TRY_E(lalloc(m, &S.disp, c->n_tanks));
    TIT_TRY_E(lalloc(m, &S.time, c->n_tanks));
    TIT_TRY_E(lalloc(m, &S.dt, c->n_tanks));
    TIT_TRY_E(lalloc(m, &S.dt_min, static_cast<size_t>(c->n_tanks) * kRedStripes));
  }
  // the chunk buffers follow the context's schedule
  TIT_TRY_E(lalloc(m, &S.chunks, c->chunk_cap));
  S.chunk_cap = c->chunk_cap;
  if (c->bchunk_cap) {
    TIT_TRY_E(lalloc(m, &S.bchunks, c->bchunk_cap));
    TIT_TRY_E(lalloc(m, &S.bcount, static_cast<size_t>(c->n_blocks) + 1));
  }
  S.bchunk_cap = c->bchunk_cap;
  S.n_blocks = c->n_blocks;
  S.alloc = true;
  return hipSuccess;
}

template <class T>
hipError_t dcopy(T* to, const T* from, size_t count, hipStream_t s) {
  return hipMemcpyAsync(to, from, count * sizeof(T), hipMemcpyDeviceToDevice, s);
}

// to_snapshot: context -> S, else S -> context.
hipError_t copy_state(tit_hip_ctx* c, Snapshot& S, bool to_snapshot) {
  hipStream_t s = c->stream;
  const size_t n = c->n, nt = c->n_tanks;
  auto cp = [&](auto* ctx_p, auto* snap_p, size_t count) {
    return to_snapshot ? dcopy(snap_p, ctx_p, count, s) : dcopy(ctx_p, snap_p, count, s);
  };
  for (int k = 0; k < c->prm.dim; ++k) {
    TIT_TRY_E(cp(c->A.x[k], S.s.x[k], n));
    TIT_TRY_E(cp(c->A.v[k], S.s.v[k], n));
  }
  TIT_TRY_E(cp(c->A.rho, S.s.rho, n));
  TIT_TRY_E(cp(c->A.aux, S.s.aux, n));
  TIT_TRY_E(cp(c->A.pv, S.s.pv, n));
  if (c->A.vol) TIT_TRY_E(cp(c->A.vol, S.s.vol, n));
  TIT_TRY_E(cp(c->id, S.id, n));
  TIT_TRY_E(cp(c->tag, S.tag, n));
  TIT_TRY_E(cp(c->key, S.key, n));
  TIT_TRY_E(cp(c->cell_off, S.cell_off, static_cast<size_t>(c->n_cells) + 1));
  TIT_TRY_E(cp(c->chunks, S.chunks, c->chunk_cap));
  TIT_TRY_E(cp(c->chunk_counts, S.chunk_counts, 4));
  TIT_TRY_E(cp(c->fixed_list, S.fixed_list, std::max<uint32_t>(c->n_fixed, 1)));
  TIT_TRY_E(cp(c->fixed_active, S.fixed_active, std::max<uint32_t>(c->n_fixed, 1)));
  TIT_TRY_E(cp(c->cell_fluid, S.cell_fluid, std::max<uint32_t>(c->n_cells, 1)));
  if (c->bchunk_cap) {
    TIT_TRY_E(cp(c->bchunks, S.bchunks, c->bchunk_cap));
    TIT_TRY_E(cp(c->bcount, S.bcount, static_cast<size_t>(c->n_blocks) + 1));
  }
  TIT_TRY_E(cp(c->counter, S.counter, 4));
  TIT_TRY_E(cp(c->rebuild_ctl, S.ctl, 4));
  TIT_TRY_E(cp(c->tank_need, S.need, nt));
  TIT_TRY_E(cp(c->vmax, S.vmax, nt * kRedStripes));
  TIT_TRY_E(cp(c->disp, S.disp, nt));
  TIT_TRY_E(cp(c->time, S.time, nt));
  TIT_TRY_E(cp(c->dt, S.dt, nt));
  TIT_TRY_E(cp(c->dt_min, S.dt_min, nt * kRedStripes));
  return hipSuccess;
}

hipError_t save(tit_hip_ctx* c, Snapshot& S) {
  TIT_TRY_E(alloc_snapshot(c, S));
  return copy_state(c, S, true);
}

hipError_t restore(tit_hip_ctx* c, Snapshot& S) { return copy_state(c, S, false); }

tit_status upload_weights(tit_hip_ctx* c, const float* w) {
  Learned* L = c->learned;
  TIT_HIP_TRY(c, hipMemcpy(L->w32, w, kParams * sizeof(float), hipMemcpyHostToDevice));
  k_to_bf16<<<nblk(kParams), kBlk, 0, c->stream>>>(kParams, L->w32, L->w16);
  pack_fragments(c->stream, L->w16, L->frag);
  TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "weights");
}

}  // namespace

}  // namespace tit_amd

extern "C" {

tit_status tit_hip_learned_init(tit_hip_ctx* c, uint64_t seed) {
  if (!c) return TIT_E_ARG;
  if (!c->uploaded) return fail(c, TIT_E_STATE, "learned init before upload");
  if (c->walls) return fail(c, TIT_E_ARG, "the learned step runs without walls");
  if (c->mass) return fail(c, TIT_E_ARG, "the learned step needs uniform masses");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  destroy_learned(c);
  auto* L = new (std::nothrow) Learned();
  if (!L) return TIT_E_OOM;
  c->learned = L;
  TIT_HIP_TRY(c, lalloc(L->mem, &L->w32, kParams));
  TIT_HIP_TRY(c, lalloc(L->mem, &L->w16, kParams));
  TIT_HIP_TRY(c, lalloc(L->mem, &L->frag, fused_fragment_elems()));
  TIT_HIP_TRY(c, lalloc(L->mem, &L->grad, kParams));
  TIT_HIP_TRY(c, lalloc(L->mem, &L->adam_m, kParams));
  TIT_HIP_TRY(c, lalloc(L->mem, &L->adam_v, kParams));
  TIT_HIP_TRY(c, lalloc(L->mem, &L->loss, 1));
  TIT_HIP_TRY(c, lalloc(L->mem, &L->err_bits, kRedStripes));
  TIT_HIP_TRY(c, lalloc(L->mem, &L->front, 2 * kRedStripes));
  TIT_HIP_TRY(c, lalloc(L->mem, &L->dv, 2 * kRedStripes));
  L->sc = default_scales(c);
  std::vector<float> w;
  init_weights(w, seed);
  fill_f32(c->stream, kParams, L->adam_m, 0.0f);
  fill_f32(c->stream, kParams, L->adam_v, 0.0f);
  L->adam_t = 0;
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  return upload_weights(c, w.data());
}

tit_status tit_hip_learned_params(tit_hip_ctx* c, float* w, size_t cap, size_t* n) {
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  if (n) *n = kParams;
  if (!w) return TIT_OK;
  if (cap < static_cast<size_t>(kParams)) return fail(c, TIT_E_ARG, "parameter buffer too small");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
  TIT_HIP_TRY(c, hipMemcpy(w, c->learned->w32, kParams * sizeof(float), hipMemcpyDeviceToHost));
  return TIT_OK;
}

tit_status tit_hip_learned_set_params(tit_hip_ctx* c, const float* w, size_t n) {
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  if (!w || n != static_cast<size_t>(kParams)) return fail(c, TIT_E_ARG, "parameter count mismatch");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  c->learned->adam_t = 0;
  fill_f32(c->stream, kParams, c->learned->adam_m, 0.0f);
  fill_f32(c->stream, kParams, c->learned->adam_v, 0.0f);
  return upload_weights(c, w);
}

tit_status tit_hip_learned_set_fused(tit_hip_ctx* c, int enable) {
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  c->learned->fused = enable != 0;
  return TIT_OK;
}

tit_status tit_hip_learned_scales(tit_hip_ctx* c, double* out) {
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  if (!out) return fail(c, TIT_E_ARG, "null output");
  const LearnedScales& s = c->learned->sc;
  out[0] = s.v_s;
  out[1] = s.rho_s;
  for (int k = 0; k < 4; ++k) out[2 + k] = s.out_scale[k];
  return TIT_OK;
}

tit_status tit_hip_learned_set_scales(tit_hip_ctx* c, const double* in) {
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  if (!in) return fail(c, TIT_E_ARG, "null scales");
  for (int k = 0; k < 6; ++k) {
    if (!(in[k] > 0.0) || !std::isfinite(in[k])) return fail(c, TIT_E_ARG, "scales must be positive");
  }
  LearnedScales& s = c->learned->sc;
  s.v_s = static_cast<float>(in[0]);
  s.rho_s = static_cast<float>(in[1]);
  for (int k = 0; k < 4; ++k) s.out_scale[k] = static_cast<float>(in[2 + k]);
  c->learned->scales_set = true;
  return TIT_OK;
}

tit_status tit_hip_learned_graph(tit_hip_ctx* c, uint32_t* nbr_ref, float* node_feat,
                                 float* edge_feat) {
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  Learned* L = c->learned;
  const uint32_t n = c->n;
  uint32_t* d_nbr = nullptr;
  float *d_nf = nullptr, *d_ef = nullptr;
  std::vector<void*> tmp;
  auto cleanup = [&] { free_mem(tmp); };
  enqueue_rebuild(c, kRebuildPending);
  enqueue_boundary(c, c->A);
  st = build_graph(c);
  if (st != TIT_OK) return st;
  if (lalloc(tmp, &d_nbr, static_cast<size_t>(n) * kKnn) != hipSuccess ||
      lalloc(tmp, &d_nf, static_cast<size_t>(n) * kNodeFeat) != hipSuccess ||
      lalloc(tmp, &d_ef, static_cast<size_t>(n) * kKnn * kEdgeFeat) != hipSuccess) {
    cleanup();
    return fail(c, TIT_E_OOM, "graph export buffers");
  }
  k_export_graph<<<nblk(n), kBlk, 0, c->stream>>>(n, c->id, L->nbr, L->infer.H[0], L->ef, d_nbr,
                                                  d_nf, d_ef);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && nbr_ref) {
    e = hipMemcpy(nbr_ref, d_nbr, static_cast<size_t>(n) * kKnn * 4, hipMemcpyDeviceToHost);
  }
  if (e == hipSuccess && node_feat) {
    e = hipMemcpy(node_feat, d_nf, static_cast<size_t>(n) * kNodeFeat * 4, hipMemcpyDeviceToHost);
  }
  if (e == hipSuccess && edge_feat) {
    e = hipMemcpy(edge_feat, d_ef, static_cast<size_t>(n) * kKnn * kEdgeFeat * 4,
                  hipMemcpyDeviceToHost);
  }
  cleanup();
  if (e != hipSuccess) return fail(c, TIT_E_HIP, std::string("graph export: ") + hipGetErrorString(e));
  return check_launch(c, "graph export");
}

tit_status tit_hip_learned_forward(tit_hip_ctx* c, float* out) {
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  if (!out) return fail(c, TIT_E_ARG, "null output");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  st = infer_current(c, kRebuildPending);
  if (st != TIT_OK) return st;
  float* d_out = nullptr;
  std::vector<void*> tmp;
  if (lalloc(tmp, &d_out, static_cast<size_t>(c->n) * kOut) != hipSuccess) {
    return fail(c, TIT_E_OOM, "forward export buffer");
  }
  k_export_out<<<nblk(static_cast<size_t>(c->n) * kOut), kBlk, 0, c->stream>>>(
      c->n, c->id, c->learned->infer.u[kRounds - 1], d_out);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) {
    e = hipMemcpy(out, d_out, static_cast<size_t>(c->n) * kOut * 4, hipMemcpyDeviceToHost);
  }
  free_mem(tmp);
  if (e != hipSuccess) return fail(c, TIT_E_HIP, std::string("forward export: ") + hipGetErrorString(e));
  return check_launch(c, "forward export");
}

tit_status tit_hip_learned_record(tit_hip_ctx* c, int plain_steps, int* sample_index) {
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  if (plain_steps < 1) return fail(c, TIT_E_ARG, "plain_steps must be positive");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  Learned* L = c->learned;
  const uint32_t n = c->n;
  const int dim = c->prm.dim;
  hipStream_t s = c->stream;
  enqueue_rebuild(c, kRebuildPending);
  enqueue_boundary(c, c->A);
  st = build_graph(c);
  if (st != TIT_OK) return st;
  Sample S;
  S.n = n;
  auto& m = S.mem;
  if (lalloc(m, &S.nbr, static_cast<size_t>(n) * kKnn) != hipSuccess ||
      lalloc(m, &S.H0, static_cast<size_t>(n) * kNodeIn) != hipSuccess ||
      lalloc(m, &S.ef, static_cast<size_t>(n) * kKnn * kEdgeFeat) != hipSuccess ||
      lalloc(m, &S.target, static_cast<size_t>(n) * 4) != hipSuccess ||
      lalloc(m, &S.fluid, n) != hipSuccess) {
    free_mem(m);
    return fail(c, TIT_E_OOM, "sample buffers");
  }
  if (!L->v_ref) {
    TIT_HIP_TRY(c, lalloc(L->mem_n, &L->v_ref, static_cast<size_t>(n) * dim));
    TIT_HIP_TRY(c, lalloc(L->mem_n, &L->r_ref, static_cast<size_t>(n) * dim));
    TIT_HIP_TRY(c, lalloc(L->mem_n, &L->rho_ref, n));
    TIT_HIP_TRY(c, lalloc(L->mem_n, &L->p_ref, n));
    TIT_HIP_TRY(c, lalloc(L->mem_n, &L->t0, c->n_tanks));
  }
  TIT_HIP_TRY(c, hipMemcpyAsync(S.nbr, L->nbr, static_cast<size_t>(n) * kKnn * 4,
                                hipMemcpyDeviceToDevice, s));
  TIT_HIP_TRY(c, hipMemcpyAsync(S.H0, L->infer.H[0], static_cast<size_t>(n) * kNodeIn * 2,
                                hipMemcpyDeviceToDevice, s));
  TIT_HIP_TRY(c, hipMemcpyAsync(S.ef, L->ef, static_cast<size_t>(n) * kKnn * kEdgeFeat * 2,
                                hipMemcpyDeviceToDevice, s));
  TIT_HIP_TRY(c, save(c, L->before));
  st = run_plain(c, plain_steps);
  if (st != TIT_OK) {
    free_mem(m);
    return st;
  }
  launch_export(s, c->dp, n, c->A, c->id, L->r_ref, L->v_ref, L->rho_ref, L->p_ref);
  const Snapshot& B = L->before;
  if (dim == 2) {
    k_targets<2><<<nblk(n), kBlk, 0, s>>>(n, B.id, B.tag, B.s, L->v_ref, L->rho_ref, B.time,
                                          c->time, S.target, S.fluid);
  } else {
    k_targets<3><<<nblk(n), kBlk, 0, s>>>(n, B.id, B.tag, B.s, L->v_ref, L->rho_ref, B.time,
                                          c->time, S.target, S.fluid);
  }
  TIT_HIP_TRY(c, hipStreamSynchronize(s));
  st = check_launch(c, "record");
  if (st != TIT_OK) {
    free_mem(m);
    return st;
  }
  S.n_fluid = n - c->n_fixed;
  if (!L->scales_set) {
    // output scales: RMS of the recorded accelerations / density rates
    std::vector<float> y(static_cast<size_t>(n) * 4);
    std::vector<uint8_t> fl(n);
    TIT_HIP_TRY(c, hipMemcpy(y.data(), S.target, y.size() * 4, hipMemcpyDeviceToHost));
    TIT_HIP_TRY(c, hipMemcpy(fl.data(), S.fluid, n, hipMemcpyDeviceToHost));
    double sa = 0.0, sr = 0.0;
    size_t cnt = 0;
    for (uint32_t i = 0; i < n; ++i) {
      if (!fl[i]) continue;
      for (int k = 0; k < dim; ++k) sa += static_cast<double>(y[4 * i + k]) * y[4 * i + k];
      sr += static_cast<double>(y[4 * i + 3]) * y[4 * i + 3];
      ++cnt;
    }
    const double ra = cnt ? std::sqrt(sa / (cnt * dim)) : 1.0;
    const double rr = cnt ? std::sqrt(sr / cnt) : 1.0;
    for (int k = 0; k < 3; ++k) L->sc.out_scale[k] = static_cast<float>(ra > 0 ? ra : 1.0);
    L->sc.out_scale[3] = static_cast<float>(rr > 0 ? rr : 1.0);
    L->scales_set = true;
  }
  L->samples.push_back(std::move(S));
  if (sample_index) *sample_index = static_cast<int>(L->samples.size()) - 1;
  return TIT_OK;
}

tit_status tit_hip_learned_clear_samples(tit_hip_ctx* c) {
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (Sample& S : c->learned->samples) free_mem(S.mem);
  c->learned->samples.clear();
  return TIT_OK;
}

tit_status tit_hip_learned_sample(tit_hip_ctx* c, int index, size_t* n, uint32_t* nbr,
                                  float* node_feat, float* edge_feat, float* target,
                                  uint8_t* fluid) {
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  Learned* L = c->learned;
  if (index < 0 || index >= static_cast<int>(L->samples.size())) {
    return fail(c, TIT_E_ARG, "no such sample");
  }
  const Sample& S = L->samples[index];
  if (n) *n = S.n;
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
  const size_t N = S.n, E = N * kKnn;
  if (nbr) TIT_HIP_TRY(c, hipMemcpy(nbr, S.nbr, E * 4, hipMemcpyDeviceToHost));
  if (target) TIT_HIP_TRY(c, hipMemcpy(target, S.target, N * 16, hipMemcpyDeviceToHost));
  if (fluid) TIT_HIP_TRY(c, hipMemcpy(fluid, S.fluid, N, hipMemcpyDeviceToHost));
  if (node_feat) {
    std::vector<bf16> h(N * kNodeIn);
    TIT_HIP_TRY(c, hipMemcpy(h.data(), S.H0, h.size() * 2, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < N; ++i) {
      for (int k = 0; k < kNodeFeat; ++k) node_feat[i * kNodeFeat + k] = static_cast<float>(h[i * kNodeIn + k]);
    }
  }
  if (edge_feat) {
    std::vector<bf16> e(E * kEdgeFeat);
    TIT_HIP_TRY(c, hipMemcpy(e.data(), S.ef, e.size() * 2, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < e.size(); ++k) edge_feat[k] = static_cast<float>(e[k]);
  }
  return TIT_OK;
}

tit_status tit_hip_learned_loss_grad(tit_hip_ctx* c, int index, double* loss, float* grad) {
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  Learned* L = c->learned;
  if (index < 0 || index >= static_cast<int>(L->samples.size())) {
    return fail(c, TIT_E_ARG, "no such sample");
  }
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  const Sample& S = L->samples[index];
  if (L->train.cap < S.n) TIT_HIP_TRY(c, alloc_acts(L->train, S.n, true));
  hipStream_t s = c->stream;
  fill_f32(s, kParams, L->grad, 0.0f);
  TIT_HIP_TRY(c, hipMemsetAsync(L->loss, 0, sizeof(double), s));
  TIT_HIP_TRY(c, loss_and_grad(s, L, c->prm.dim, S, L->loss));
  TIT_HIP_TRY(c, hipStreamSynchronize(s));
  st = check_launch(c, "loss/grad");
  if (st != TIT_OK) return st;
  if (loss) TIT_HIP_TRY(c, hipMemcpy(loss, L->loss, sizeof(double), hipMemcpyDeviceToHost));
  if (grad) TIT_HIP_TRY(c, hipMemcpy(grad, L->grad, kParams * 4, hipMemcpyDeviceToHost));
  return TIT_OK;
}

tit_status tit_hip_learned_train(tit_hip_ctx* c, int iters, double lr, double* losses) {
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  Learned* L = c->learned;
  if (L->samples.empty()) return fail(c, TIT_E_STATE, "no recorded samples");
  if (iters < 0 || !(lr > 0.0)) return fail(c, TIT_E_ARG, "bad training arguments");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  uint32_t nmax = 0;
  for (const Sample& S : L->samples) nmax = std::max(nmax, S.n);
  if (L->train.cap < nmax) TIT_HIP_TRY(c, alloc_acts(L->train, nmax, true));
  hipStream_t s = c->stream;
  double* d_losses = nullptr;
  std::vector<void*> tmp;
  if (lalloc(tmp, &d_losses, std::max(iters, 1)) != hipSuccess) return fail(c, TIT_E_OOM, "losses");
  TIT_HIP_TRY(c, hipMemsetAsync(d_losses, 0, std::max(iters, 1) * sizeof(double), s));
  constexpr float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int it = 0; it < iters; ++it) {
    const Sample& S = L->samples[it % L->samples.size()];
    fill_f32(s, kParams, L->grad, 0.0f);
    hipError_t e = loss_and_grad(s, L, c->prm.dim, S, d_losses + it);
    if (e != hipSuccess) {
      free_mem(tmp);
      return fail(c, TIT_E_HIP, std::string("train: ") + hipGetErrorString(e));
    }
    L->adam_t += 1;
    const float bc1 = 1.0f - std::pow(b1, static_cast<float>(L->adam_t));
    const float bc2 = 1.0f - std::pow(b2, static_cast<float>(L->adam_t));
    adam_update(s, kParams, L->w32, L->w16, L->grad, L->adam_m, L->adam_v,
                static_cast<float>(lr), b1, b2, eps, bc1, bc2);
  }
  pack_fragments(s, L->w16, L->frag);
  hipError_t e = hipStreamSynchronize(s);
  if (e == hipSuccess && losses && iters > 0) {
    e = hipMemcpy(losses, d_losses, iters * sizeof(double), hipMemcpyDeviceToHost);
  }
  free_mem(tmp);
  if (e != hipSuccess) return fail(c, TIT_E_HIP, std::string("train: ") + hipGetErrorString(e));
  return check_launch(c, "train");
}

tit_status tit_hip_learned_run(tit_hip_ctx* c, int n_steps, int judge_every, double tol,
                               double dv_tol, tit_hip_learned_report* rep) {
  tit_status st = need_learned(c);
  if (st != TIT_OK) return st;
  if (n_steps < 0 || judge_every < 1 || !(tol >= 0.0) || !(dv_tol >= 0.0)) {
    return fail(c, TIT_E_ARG, "bad learned-run arguments");
  }
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  Learned* L = c->learned;
  hipStream_t s = c->stream;
  tit_hip_learned_report R{};
  constexpr int kPlainPerLearned = 4;
  TIT_HIP_TRY(c, save(c, L->accepted));
  int done = 0;
  while (done < n_steps) {
    const int group = std::min(judge_every, n_steps - done);
    for (int k = 0; k + 1 < group; ++k) {
      st = learned_step(c);
      if (st != TIT_OK) return st;
    }
    // judged step: learned and plain from the same state
    TIT_HIP_TRY(c, save(c, L->before));
    if (c->dp.dim == 2) {
      k_keep_by_id<2><<<nblk(c->n), kBlk, 0, s>>>(c->n, c->A, c->id, nullptr, nullptr, L->v_s);
    } else {
      k_keep_by_id<3><<<nblk(c->n), kBlk, 0, s>>>(c->n, c->A, c->id, nullptr, nullptr, L->v_s);
    }
    st = learned_step(c);
    if (st != TIT_OK) return st;
    if (c->dp.dim == 2) {
      k_keep_by_id<2><<<nblk(c->n), kBlk, 0, s>>>(c->n, c->A, c->id, L->rho_l, L->x_l, L->v_l);
    } else {
      k_keep_by_id<3><<<nblk(c->n), kBlk, 0, s>>>(c->n, c->A, c->id, L->rho_l, L->x_l, L->v_l);
    }
    TIT_HIP_TRY(c, restore(c, L->before));
    st = run_plain(c, kPlainPerLearned);
    if (st != TIT_OK) return st;
    TIT_HIP_TRY(c, hipMemsetAsync(L->err_bits, 0, kRedStripes * sizeof(uint32_t), s));
    TIT_HIP_TRY(c, hipMemsetAsync(L->front, 0, 2 * kRedStripes * sizeof(unsigned long long), s));
    TIT_HIP_TRY(c, hipMemsetAsync(L->dv, 0, 2 * kRedStripes * sizeof(double), s));
    if (c->dp.dim == 2) {
      k_judge<2><<<nblk(c->n), kBlk, 0, s>>>(c->n, c->A, c->id, c->tag, L->rho_l, L->x_l,
                                             L->v_l, L->v_s, L->err_bits, L->front, L->dv);
    } else {
      k_judge<3><<<nblk(c->n), kBlk, 0, s>>>(c->n, c->A, c->id, c->tag, L->rho_l, L->x_l,
                                             L->v_l, L->v_s, L->err_bits, L->front, L->dv);
    }
    uint32_t ebs[kRedStripes];
    unsigned long long frs[2 * kRedStripes];
    double dvss[2 * kRedStripes];
    TIT_HIP_TRY(c, hipMemcpyAsync(ebs, L->err_bits, sizeof(ebs), hipMemcpyDeviceToHost, s));
    TIT_HIP_TRY(c, hipMemcpyAsync(frs, L->front, sizeof(frs), hipMemcpyDeviceToHost, s));
    TIT_HIP_TRY(c, hipMemcpyAsync(dvss, L->dv, sizeof(dvss), hipMemcpyDeviceToHost, s));
    TIT_HIP_TRY(c, hipStreamSynchronize(s));
    st = check_launch(c, "judge");
    if (st != TIT_OK) return st;
    // fold the copies
    uint32_t eb = 0;
    unsigned long long fr[2] = {0, 0};
    double dvs[2] = {0.0, 0.0};
    for (int k = 0; k < kRedStripes; ++k) {
      eb = std::max(eb, ebs[k]);
      fr[0] = std::max(fr[0], frs[k]);
      fr[1] = std::max(fr[1], frs[kRedStripes + k]);
      dvs[0] += dvss[k];
      dvs[1] += dvss[kRedStripes + k];
    }
    float err;
    std::memcpy(&err, &eb, 4);
    auto unorder = [](unsigned long long b) {
      const unsigned long long u = (b >> 63) ? (b & 0x7fffffffffffffffull) : ~b;
      double x;
      std::memcpy(&x, &u, 8);
      return x;
    };
    const double fp = unorder(fr[0]), fl = unorder(fr[1]);
    const double ferr = fp != 0.0 ? std::fabs(fl - fp) / std::fabs(fp) : 0.0;
    R.learned_steps += group;
    R.judged += 1;
    R.max_rho_err = std::max(R.max_rho_err, static_cast<double>(err));
    R.max_front_err = std::max(R.max_front_err, ferr);
    const double dverr = dvs[1] > 0.0 ? std::sqrt(dvs[0] / dvs[1]) : 0.0;
    R.max_dv_err = std::max(R.max_dv_err, dverr);
    R.last_dv_err = dverr;
    R.last_rho_err = err;
    R.plain_steps += kPlainPerLearned;
    if (!(static_cast<double>(err) <= tol) || !(dverr <= dv_tol)) {
      // fallback: back to the last accepted state, plain steps over the group
      R.rejected += 1;
      TIT_HIP_TRY(c, restore(c, L->accepted));
      st = run_plain(c, kPlainPerLearned * group);
      if (st != TIT_OK) return st;
      R.plain_steps += kPlainPerLearned * group;
    }
    TIT_HIP_TRY(c, save(c, L->accepted));
    done += group;
  }
  TIT_HIP_TRY(c, hipStreamSynchronize(s));
  if (c->prof.enabled) c->prof.collect();
  st = read_flags(c);
  if (rep) *rep = R;
  return st;
}

}  // extern "C"

extern "C" {

// BF16 GEMM on the current device (test hook): host FP32 operands are
// rounded to BF16, C = act(op(A) op(B) + bias) in FP32, or C += op(A) op(B)
// through split-K atomics when k_splits > 1.
tit_status tit_hip_nn_gemm(int M, int N, int K, const float* A, int transA,
                           const float* B, int transB, const float* bias, int act,
                           int k_splits, float* C) {
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C || k_splits < 1 ||
      (act != kActNone && act != kActHswish) || (k_splits > 1 && (bias || act))) {
    return TIT_E_ARG;
  }
  std::vector<void*> mem;
  float *dA = nullptr, *dB = nullptr, *dC = nullptr, *dbias = nullptr;
  bf16 *hA = nullptr, *hB = nullptr;
  const size_t na = static_cast<size_t>(M) * K, nb = static_cast<size_t>(K) * N,
               nc = static_cast<size_t>(M) * N;
  hipError_t e = hipSuccess;
  auto step = [&](hipError_t x) {
    if (e == hipSuccess) e = x;
  };
  step(lalloc(mem, &dA, na));
  step(lalloc(mem, &dB, nb));
  step(lalloc(mem, &dC, nc));
  step(lalloc(mem, &hA, na));
  step(lalloc(mem, &hB, nb));
  if (bias) step(lalloc(mem, &dbias, N));
  if (e == hipSuccess) {
    step(hipMemcpy(dA, A, na * 4, hipMemcpyHostToDevice));
    step(hipMemcpy(dB, B, nb * 4, hipMemcpyHostToDevice));
    step(hipMemcpy(dC, C, nc * 4, hipMemcpyHostToDevice));
    if (bias) step(hipMemcpy(dbias, bias, N * 4, hipMemcpyHostToDevice));
  }
  if (e == hipSuccess) {
    k_to_bf16<<<nblk(na), kBlk>>>(na, dA, hA);
    k_to_bf16<<<nblk(nb), kBlk>>>(nb, dB, hB);
    GemmArgs g{};
    g.M = M;
    g.N = N;
    g.K = K;
    g.A = hA;
    g.transA = transA != 0;
    g.lda = transA ? M : K;
    g.B = hB;
    g.transB = transB != 0;
    g.ldb = transB ? K : N;
    g.bias = dbias;
    g.act = act;
    g.out32 = dC;
    g.ldo32 = N;
    g.accumulate = k_splits > 1;
    g.k_splits = k_splits;
    gemm_bf16(nullptr, g);
    step(hipGetLastError());
    step(hipDeviceSynchronize());
    if (e == hipSuccess) step(hipMemcpy(C, dC, nc * 4, hipMemcpyDeviceToHost));
  }
  free_mem(mem);
  return e == hipSuccess ? TIT_OK : (e == hipErrorOutOfMemory ? TIT_E_OOM : TIT_E_HIP);
}

}  // extern "C"
