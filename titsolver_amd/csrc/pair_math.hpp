// Per-pair arithmetic of the continuity and momentum right-hand sides, shared
// by every traversal (LDS-staged strips and the global-memory path) so that
// all paths round identically.
//
// With r_ab = r_a - r_b and grad W_ab = G(q) r_ab (device_common.hpp):
//   continuity  m v_ab . grad W_ab                  (fluid_equations.hpp:257)
//     + CFG  delta h c0 V_b 2 (rho_a - rho_b) G     (Molteni & Colagrossi 2009;
//            psi_ab . grad W = 2 rho_ab G, the |r_ab|^2 cancels)
//     + REF  m max(cs_a, cs_b) rho_ab / rho_b G |r_ab|   (Ferrari, :253-257)
//   momentum    m (Pi_ab - p_a/rho_a^2 - p_b/rho_b^2) G r_ab   (:297-303)
//     CFG  Pi_ab = alpha h cbar (v_ab . r_ab) / (rhobar (|r_ab|^2 + 0.01 h^2))
//          for approaching pairs (Monaghan 1992), else 0
//     REF  Pi_ab = 2 mu (v_ab . r_ab) / (rho_a rho_b |r_ab|^2)   (:299-300)
#pragma once

#include "internal.hpp"

namespace tit_amd {

// Sound speed of an FP32 density: cs0 (rho/rho0)^((xi-1)/2).
__device__ __forceinline__ float sound_speed_f32(const DevParams& P, float rho) {
  const float ratio = rho * P.inv_rho0_f;
  if (P.xi_int) {
    const int k = P.xi_int - 1;
    float r = 1.0f, b = ratio;
    for (int e = k >> 1; e > 0; e >>= 1) {
      if (e & 1) r *= b;
      b *= b;
    }
    return P.cs0_f * ((k & 1) ? r * __builtin_sqrtf(ratio) : r);
  }
  return P.cs0_f * __builtin_powf(ratio, 0.5f * (static_cast<float>(P.xi) - 1.0f));
}

template <int DIM>
struct Target {
  double x[DIM];
  float v[DIM];
  double rho;
  float drho_f, rho_f, Pa, csa;  // rho - rho0, rho (FP32)
};

template <int DIM>
struct Sums {
  Kahan rho;
  Kahan v[DIM];
  __device__ __forceinline__ void init() {
    rho.init();
#pragma unroll
    for (int k = 0; k < DIM; ++k) v[k].init();
  }
};

// p/rho^2 of a density, the only per-particle EOS quantity of the pair terms.
__device__ __forceinline__ float pressure_over_rho2(const DevParams& P,
                                                    double rho) {
  const double p = eos_pressure(P, rho);
  return static_cast<float>(p / (rho * rho));
}

// Per-particle pair inputs kept beside rho: (p / rho^2, cs).  Every kernel
// that writes a density writes this too.
__device__ __forceinline__ float2 make_aux(const DevParams& P, double rho) {
  return make_float2(pressure_over_rho2(P, rho),
                     sound_speed_f32(P, static_cast<float>(rho)));
}

// (v, rho - rho0) in FP32 for the neighbours' side of the pair terms.
template <int DIM>
__device__ __forceinline__ float4 make_pv(const DevParams& P, const double* v, double rho) {
  return make_float4(static_cast<float>(v[0]), static_cast<float>(v[1]),
                     DIM == 3 ? static_cast<float>(v[DIM - 1]) : 0.0f,
                     static_cast<float>(rho - P.rho0));
}

// Writes both per-particle pair records of a particle.
template <int DIM>
__device__ __forceinline__ void write_pair_data(const DevParams& P, const DevState& st,
                                                uint32_t i, const double* v, double rho) {
  st.aux[i] = make_aux(P, rho);
  st.pv[i] = make_pv<DIM>(P, v, rho);
}

template <int DIM>
__device__ __forceinline__ void make_target(const DevParams& P, const double* x,
                                            const double* v, double rho,
                                            float2 aux, Target<DIM>& a) {
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    a.x[k] = x[k];
    a.v[k] = static_cast<float>(v[k]);
  }
  a.rho = rho;
  a.drho_f = static_cast<float>(rho - P.rho0);
  a.rho_f = P.rho0_f + a.drho_f;
  a.Pa = aux.x;
  a.csa = aux.y;
}

// d = r_b - r_a in FP64 (as in the neighbour predicate); pvb = (v_b,
// rho_b - rho0), Pb = p_b/rho_b^2, csb = cs(rho_b) from the neighbour's
// pair records.
template <int DIM, int KERNEL, int FLAVOUR>
__device__ __forceinline__ void pair_term(const DevParams& P, const Target<DIM>& a,
                                          const double* d, const float4& pvb, float Pb,
                                          float csb, Sums<DIM>& s) {
  const float vb[3] = {pvb.x, pvb.y, pvb.z};
  float x[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) x[k] = static_cast<float>(-d[k]);
  float r2 = x[0] * x[0];
#pragma unroll
  for (int k = 1; k < DIM; ++k) r2 = __builtin_fmaf(x[k], x[k], r2);
  // hardware square root and reciprocal (about 1 ulp; the pair terms are
  // FP32 and compared with the FP64 oracle at a stated tolerance)
  const float r = __builtin_amdgcn_sqrtf(r2);
  const float q = r * P.inv_h;
  const float G = P.gsc * kernel_grad_poly<KERNEL>(q);
  float vr = 0.0f;
#pragma unroll
  for (int k = 0; k < DIM; ++k) vr = __builtin_fmaf(a.v[k] - vb[k], x[k], vr);
  const float drab = a.drho_f - pvb.w;
  const float rbf = P.rho0_f + pvb.w;
  const float Vb = P.m_f * __builtin_amdgcn_rcpf(rbf);
  float cont, Pi;
  if constexpr (FLAVOUR == kFlavourCfg) {
    cont = G * __builtin_fmaf(P.m_f, vr, P.dhc2 * Vb * drab);
    const float cbar = 0.5f * (a.csa + csb);
    const float rbar = 0.5f * (a.rho_f + rbf);
    const float visc = P.ahc * cbar * vr * __builtin_amdgcn_rcpf(rbar * (r2 + P.eta2));
    Pi = vr < 0.0f ? visc : 0.0f;
  } else {
    const float csab = fmaxf(a.csa, csb);
    cont = G * __builtin_fmaf(P.m_f, vr, csab * drab * Vb * r);
    Pi = P.two_mu * vr * __builtin_amdgcn_rcpf(a.rho_f * rbf * r2);
  }
  const float coef = P.m_f * (Pi - (a.Pa + Pb)) * G;
  s.rho.add(cont);
#pragma unroll
  for (int k = 0; k < DIM; ++k) s.v[k].add(coef * x[k]);
}

}  // namespace tit_amd
