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
// The uniform factors m gsc, kd, kv are folded in once (FP64, then rounded),
// so a pair costs the arithmetic that varies per pair only.  With walls
// (kFlavourWall) and with per-particle masses (kFlavourWall without walls,
// kFlavourCfgM) the masses differ per particle (fixed particles carry
// m gamma, fluid_equations.hpp:87-88): m_b = V_b rho_b comes from the pair
// record (V_b = m_b / rho_b in its third slot in 2D, in the volume array in
// 3D) and multiplies the whole term.
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
  return P.cs0_f * powf_general(ratio, 0.5f * (static_cast<float>(P.xi) - 1.0f));
}

// Per-target values of the pair terms, with the uniform factors folded in
// (DevParams::mg, kd, kv).
template <int DIM>
struct Target {
  double x[DIM];
  float v[DIM];
  float kd;  // kd (rho_a - rho0)
  float rb;  // CFG: (rho_a + rho0) / 2; REF: rho_a
  float cv;  // CFG: kv cs_a; REF: cs_a
  float pa;  // mg p_a / rho_a^2 (p_a / rho_a^2 with per-particle masses)
};

// Compensated per-particle sums.  Pair terms enter in twos: the FP32 sum of
// two consecutive terms is added with Kahan compensation, so every path that
// visits the same candidates in the same order produces the same bits.
template <int DIM>
struct Sums {
  Kahan rho;
  Kahan v[DIM];
  float p_rho, p_v[DIM];  // a pending single term
  bool pend;
  __device__ __forceinline__ void init() {
    rho.init();
#pragma unroll
    for (int k = 0; k < DIM; ++k) v[k].init();
    pend = false;
  }
  // two terms at once (no term may be pending)
  __device__ __forceinline__ void add2(float c0, const float* f0, float c1, const float* f1) {
    rho.add(c0 + c1);
#pragma unroll
    for (int k = 0; k < DIM; ++k) v[k].add(f0[k] + f1[k]);
  }
  __device__ __forceinline__ void add(float c, const float* f) {
    if (pend) {
      rho.add(p_rho + c);
#pragma unroll
      for (int k = 0; k < DIM; ++k) v[k].add(p_v[k] + f[k]);
    } else {
      p_rho = c;
#pragma unroll
      for (int k = 0; k < DIM; ++k) p_v[k] = f[k];
    }
    pend = !pend;
  }
  // a last single term enters as its sum with zero, as in add2
  __device__ __forceinline__ void finish() {
    if (pend) {
      const float z[DIM] = {};
      add2(p_rho, p_v, 0.0f, z);
      pend = false;
    }
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

// Mass of the particle with id `id` (its index in the uploaded order):
// ParticleArray's m (fluid_equations.hpp:41-48) when per-particle masses were
// uploaded, else the uniform m.
__device__ __forceinline__ double particle_mass(const DevParams& P, uint32_t id) {
  return P.mass ? P.mass[id] : P.m;
}

// (v, rho - rho0) in FP32 for the neighbours' side of the pair terms; in 2D
// the third slot holds the particle volume m / rho instead of v_z.
template <int DIM>
__device__ __forceinline__ float4 make_pv(const DevParams& P, const double* v, double rho,
                                          double m) {
  return make_float4(static_cast<float>(v[0]), static_cast<float>(v[1]),
                     DIM == 3 ? static_cast<float>(v[DIM - 1])
                              : static_cast<float>(m / rho),
                     static_cast<float>(rho - P.rho0));
}

// Writes the per-particle pair records of a particle of mass m (in 3D with
// per-particle masses also its volume m / rho).
template <int DIM>
__device__ __forceinline__ void write_pair_data(const DevParams& P, const DevState& st,
                                                uint32_t i, const double* v, double rho,
                                                double m) {
  st.aux[i] = make_aux(P, rho);
  st.pv[i] = make_pv<DIM>(P, v, rho, m);
  if constexpr (DIM == 3) {
    if (P.mass) st.vol[i] = static_cast<float>(m / rho);
  }
}

template <int DIM, int FLAVOUR>
__device__ __forceinline__ void make_target(const DevParams& P, const double* x,
                                            const double* v, double rho,
                                            float2 aux, Target<DIM>& a) {
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    a.x[k] = x[k];
    a.v[k] = static_cast<float>(v[k]);
  }
  const float drho = static_cast<float>(rho - P.rho0);
  a.kd = P.kd * drho;
  if constexpr (FLAVOUR == kFlavourCfg || FLAVOUR == kFlavourCfgM) {
    a.rb = 0.5f * (P.rho0_f + drho) + 0.5f * P.rho0_f;
    a.cv = P.kv * aux.y;
  } else {
    a.rb = P.rho0_f + drho;
    a.cv = aux.y;
  }
  a.pa = mass_flavour(FLAVOUR) ? aux.x : P.mg * aux.x;
}

// The kernel-gradient polynomial P(q) of grad W_ab = gsc P(q) r_ab as a
// function of r = |r_ab| (device_common.hpp: kernel_grad_poly).
template <int KERNEL>
__device__ __forceinline__ float grad_poly_r(const DevParams& P, float r) {
  if constexpr (KERNEL == kKernelC2) {
    const float t = fmaxf(__builtin_fmaf(r, -P.half_inv_h, 1.0f), 0.0f);
    return t * t * t;
  } else {
    const float q = r * P.inv_h;
    const float t = fmaxf(__builtin_fmaf(q, -0.5f, 1.0f), 0.0f);
    const float t2 = t * t;
    return __builtin_fmaf(2.5f, q, 1.0f) * (t2 * t2 * t);
  }
}

// One pair's continuity and momentum contributions.  d = r_b - r_a in FP64
// (as in the neighbour predicate); pvb = (v_b, rho_b - rho0) with m / rho_b
// in the third slot in 2D; Pb = p_b / rho_b^2, csb = cs(rho_b); volb = m_b /
// rho_b in 3D with per-particle masses (unused otherwise).
template <int DIM, int KERNEL, int FLAVOUR>
__device__ __forceinline__ void pair_term(const DevParams& P, const Target<DIM>& a,
                                          const double* d, const float4& pvb, float Pb,
                                          float csb, float volb, float& cont, float* mom) {
  const float vb[3] = {pvb.x, pvb.y, DIM == 3 ? pvb.z : 0.0f};
  float x[DIM];  // r_a - r_b
#pragma unroll
  for (int k = 0; k < DIM; ++k) x[k] = static_cast<float>(-d[k]);
  float r2 = x[0] * x[0];
#pragma unroll
  for (int k = 1; k < DIM; ++k) r2 = __builtin_fmaf(x[k], x[k], r2);
  // hardware square root and reciprocal (about 1 ulp; the pair terms are
  // FP32 and compared with the FP64 oracle at a stated tolerance)
  const float r = __builtin_amdgcn_sqrtf(r2);
  const float G = grad_poly_r<KERNEL>(P, r);
  float vr = 0.0f;
#pragma unroll
  for (int k = 0; k < DIM; ++k) vr = __builtin_fmaf(a.v[k] - vb[k], x[k], vr);
  const float Vb = DIM == 2 ? pvb.z
                  : (mass_flavour(FLAVOUR) ? volb
                                           : P.m_f * __builtin_amdgcn_rcpf(P.rho0_f + pvb.w));
  const float drk = __builtin_fmaf(pvb.w, -P.kd, a.kd);  // kd (rho_a - rho_b)
  float c, Pi;
  if constexpr (FLAVOUR == kFlavourCfg) {
    c = __builtin_fmaf(P.mg, vr, drk * Vb);
    const float rbar = __builtin_fmaf(pvb.w, 0.5f, a.rb);
    const float num = __builtin_fmaf(csb, P.kv, a.cv) * fminf(vr, 0.0f);
    Pi = num * __builtin_amdgcn_rcpf(rbar * (r2 + P.eta2));
  } else if constexpr (FLAVOUR == kFlavourRef) {
    c = __builtin_fmaf(P.mg, vr, fmaxf(a.cv, csb) * drk * Vb * r);
    const float rb = P.rho0_f + pvb.w;
    Pi = (P.kv * vr) * __builtin_amdgcn_rcpf(a.rb * rb * r2);
  } else if constexpr (FLAVOUR == kFlavourCfgM) {  // m_b gsc = V_b rho_b gsc
    const float mb = Vb * (P.rho0_f + pvb.w) * P.gsc;
    c = __builtin_fmaf(mb, vr, drk * Vb);
    const float rbar = __builtin_fmaf(pvb.w, 0.5f, a.rb);
    const float num = __builtin_fmaf(csb, P.kv, a.cv) * fminf(vr, 0.0f);
    Pi = num * __builtin_amdgcn_rcpf(rbar * (r2 + P.eta2));
    cont = c * G;
    const float coef = ((Pi - a.pa) - Pb) * (mb * G);
#pragma unroll
    for (int k = 0; k < DIM; ++k) mom[k] = coef * x[k];
    return;
  } else {  // kFlavourWall: m_b gsc = V_b rho_b gsc (kd = gsc here)
    const float rb = P.rho0_f + pvb.w;
    const float mb = Vb * rb * P.kd;
    c = __builtin_fmaf(mb, vr, fmaxf(a.cv, csb) * drk * Vb * r);
    Pi = (P.two_mu * vr) * __builtin_amdgcn_rcpf(a.rb * rb * r2);
    cont = c * G;
    const float coef = ((Pi - a.pa) - Pb) * (mb * G);
#pragma unroll
    for (int k = 0; k < DIM; ++k) mom[k] = coef * x[k];
    return;
  }
  cont = c * G;
  const float coef = __builtin_fmaf(Pb, -P.mg, Pi - a.pa) * G;
#pragma unroll
  for (int k = 0; k < DIM; ++k) mom[k] = coef * x[k];
}

}  // namespace tit_amd
