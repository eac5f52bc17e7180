// Device-side parameters and per-pair arithmetic of the WCSPH step.
//
// Precision split: positions, velocities, densities and every integrator
// update are FP64; the Tait EOS of each particle is evaluated in FP64; the
// pair terms run in FP32 on relative coordinates (r_a - r_b formed in FP64,
// then rounded) and are accumulated per particle with Kahan compensation.
// The neighbour predicate is the reference's FP64 test
// norm2(r_b - r_a) <= (2h)^2 (source/tit/geom/bsphere.hpp:52-54) with each
// product and sum rounded separately: this library is built with
// -ffp-contract=off, FMAs are written out explicitly where wanted.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tit_amd {

constexpr int kKernelC2 = 0;
constexpr int kKernelC4 = 1;
constexpr int kFlavourRef = 0;
constexpr int kFlavourCfg = 1;
// REF terms with the semi-analytic wall (2D): per-particle masses (fixed
// particles carry m gamma), pair sums divided by gamma_a, and the face terms
// added per particle.  Without walls (gamma = 1, no face terms; 2D or 3D)
// this is the REF flavour with per-particle masses.
constexpr int kFlavourWall = 2;
// CFG terms with per-particle masses: m_b = V_b rho_b (V_b from the pair
// record in 2D, the volume array in 3D) multiplies the whole term, as in
// kFlavourWall.
constexpr int kFlavourCfgM = 3;
// Flavours whose pair terms take each neighbour's own mass.
constexpr bool mass_flavour(int flavour) {
  return flavour == kFlavourWall || flavour == kFlavourCfgM;
}

// Uniform parameters, passed to every kernel by value.
struct DevParams {
  int dim, kernel, flavour, order;
  int xi_int;       // Tait index when it is an integer in [2, 15], else 0
  int n_tanks;
  double h, m, rho0, cs0, xi, g, mu, alpha, delta, cfl, c_force, c_visc;
  double R2;        // (2h)^2, the inclusive neighbour predicate bound
  double eos_B;     // rho0 cs0^2 / xi
  double inv_rho0;
  double skin_half; // allowed displacement since the last cell rebuild
  const double* mass; // per-particle masses by particle id, or null (uniform m)
  float inv_h;      // 1 / h
  float gsc;        // grad W_ab = gsc * P(q) * (r_a - r_b)
  float wsc;        // W_ab = wsc * w(q)
  float m_f;        // particle mass
  float dhc2;       // 2 delta h cs0   (delta-SPH)
  float ahc;        // alpha h         (Monaghan viscosity)
  float eta2;       // 0.01 h^2
  float two_mu;     // 2 mu            (physical viscosity)
  float inv_rho0_f; // 1 / rho0
  float cs0_f;      // cs0
  float rho0_f;     // rho0
  // factors folded into the pair terms
  float half_inv_h; // 1 / (2h)
  float mg;         // m gsc
  float kd;         // CFG: 2 delta h cs0 gsc; REF: gsc
  float kv;         // CFG: alpha h m gsc / 2; CFG with masses: alpha h / 2;
                    // REF: 2 mu m gsc
};

// Per-tank cell grid.  Cell size is at least (support radius + skin) / 2, so
// a 5^d-cell stencil covers every pair within 2h while particles have moved
// less than skin/2 since the cells were built.
struct DevTank {
  double lo[3];
  double inv_c[3];
  int nc[3];
  uint32_t cell_off;  // first global cell index of this tank
  uint32_t n_cells;
};

// Stencil half-width in cells.
constexpr int kStencil = 2;

// ---- smoothing kernel ------------------------------------------------------
// Wendland C2: w = (1+2q)(1-q/2)^4, w' = -5 q (1-q/2)^3.
// Wendland C4: w = (1+3q+35/12 q^2)(1-q/2)^6, w' = -14/3 q (1+5q/2)(1-q/2)^5
// (source/tit/sph/kernel.gen.cpp:539-546).  grad W = omega/h^(d+1) w'(q) x/|x|
// = [omega w'(q)/(q h^(d+2))] x, so only P(q) = w'(q)/(c q) is evaluated.
template <int KERNEL>
__device__ __forceinline__ float kernel_grad_poly(float q) {
  float t = fmaxf(1.0f - 0.5f * q, 0.0f);
  if constexpr (KERNEL == kKernelC2) {
    return t * t * t;
  } else {
    const float t2 = t * t;
    return __builtin_fmaf(2.5f, q, 1.0f) * (t2 * t2 * t);
  }
}

template <int KERNEL>
__device__ __forceinline__ float kernel_value_poly(float q) {
  float t = fmaxf(1.0f - 0.5f * q, 0.0f);
  if constexpr (KERNEL == kKernelC2) {
    const float t2 = t * t;
    return __builtin_fmaf(2.0f, q, 1.0f) * (t2 * t2);
  } else {
    const float t3 = t * t * t;
    return __builtin_fmaf(q, __builtin_fmaf(35.0f / 12.0f, q, 3.0f), 1.0f) *
           (t3 * t3);
  }
}

// FP64 kernel value, for the boundary extrapolation (weights of a Shepard
// sum, evaluated once per fixed-fluid pair).
template <int KERNEL>
__device__ __forceinline__ double kernel_value_poly_f64(double q) {
  if (!(q < 2.0)) return 0.0;
  const double t = 1.0 - q / 2.0;
  if constexpr (KERNEL == kKernelC2) {
    return (1.0 + 2.0 * q) * (t * t) * (t * t);
  } else {
    const double t3 = t * t * t;
    return (1.0 + 3.0 * q + 35.0 / 12.0 * q * q) * (t3 * t3);
  }
}

// ---- Tait equation of state (source/tit/sph/equation_of_state.hpp:19-72) ---

// General powers, kept out of line.  Inlined, the FP64 pow alone takes about
// 55 VGPRs, which every kernel evaluating the equation of state would pay for
// in waves per SIMD even where the Tait exponent is a small integer and the
// branch never runs.
[[maybe_unused]] static __device__ __noinline__ double pow_general(double x, double y) {
  return pow(x, y);
}
[[maybe_unused]] static __device__ __noinline__ float powf_general(float x, float y) {
  return __builtin_powf(x, y);
}

__device__ __forceinline__ double ipow_f64(double x, int n) {
  double r = 1.0, b = x;
  while (n > 0) {
    if (n & 1) r *= b;
    b *= b;
    n >>= 1;
  }
  return r;
}

// (rho/rho0)^xi
__device__ __forceinline__ double eos_ratio_pow_xi(const DevParams& P,
                                                   double ratio) {
  return P.xi_int ? ipow_f64(ratio, P.xi_int) : pow_general(ratio, P.xi);
}

// p(rho) = B ((rho/rho0)^xi - 1)
__device__ __forceinline__ double eos_pressure(const DevParams& P, double rho) {
  return P.eos_B * (eos_ratio_pow_xi(P, rho / P.rho0) - 1.0);
}

// cs(rho) = cs0 (rho/rho0)^((xi-1)/2)
__device__ __forceinline__ double eos_sound_speed(const DevParams& P,
                                                  double rho) {
  const double ratio = rho / P.rho0;
  if (P.xi_int) {
    const int k = P.xi_int - 1;
    const double rk = ipow_f64(ratio, k >> 1);
    return P.cs0 * ((k & 1) ? rk * sqrt(ratio) : rk);
  }
  return P.cs0 * pow_general(ratio, (P.xi - 1.0) / 2.0);
}

// H(rho) = cs0^2 ((rho/rho0)^(xi-1) - 1) / (xi - 1)
__device__ __forceinline__ double eos_potential(const DevParams& P,
                                                double rho) {
  const double xi_1 = P.xi - 1.0;
  const double ratio = rho / P.rho0;
  const double rp = P.xi_int ? ipow_f64(ratio, P.xi_int - 1) : pow_general(ratio, xi_1);
  return P.cs0 * P.cs0 * (rp - 1.0) / xi_1;
}

// rho(H) = rho0 (1 + (xi-1) H / cs0^2)^(1/(xi-1))
__device__ __forceinline__ double eos_density_from_potential(const DevParams& P,
                                                             double H) {
  const double xi_1 = P.xi - 1.0;
  return P.rho0 * pow_general(1.0 + xi_1 * H / (P.cs0 * P.cs0), 1.0 / xi_1);
}

// FluidEquations::compute_time_step for one fluid particle
// (source/tit/sph/fluid_equations.hpp:199-222).
template <int DIM>
__device__ __forceinline__ double particle_time_step(const DevParams& P,
                                                     double rho,
                                                     const double* v,
                                                     const double* dv) {
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
  dt = fmin(dt, P.c_force * sqrt(P.h / an));
  return dt;
}

// ---- compensated accumulation ----------------------------------------------
struct Kahan {
  float s, c;
  __device__ __forceinline__ void init() {
    s = 0.0f;
    c = 0.0f;
  }
  __device__ __forceinline__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ float value() const { return s - c; }
};

// FP64 bits of a positive double order like unsigned integers.
__device__ __forceinline__ unsigned long long dbl_bits(double x) {
  return static_cast<unsigned long long>(__double_as_longlong(x));
}

// ---- per-tank reduction slots ------------------------------------------------
// A per-tank minimum or maximum reduced over a whole grid (the next step
// size, the largest speed, the largest shift) lives in kRedStripes copies
// per tank, slot t * kRedStripes + s.  A wave updates the copy picked by its
// global wave index, so the atomics of a large grid spread over many
// addresses: atomics on one address serialise at about 11.5 ns each on
// MI355X (8192 wave flushes onto one slot take ~95 us, onto 64 slots ~2 us).
// The readers fold the copies with one wave per tank, lane s taking copy s.
constexpr int kRedStripes = 64;

__device__ __forceinline__ uint32_t red_stripe() {
  return (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kRedStripes - 1);
}

template <class T>
__device__ __forceinline__ T* red_slot(T* base, uint32_t t) {
  return base + static_cast<size_t>(t) * kRedStripes + red_stripe();
}

// Folds tank t's copies (called by all 64 lanes of a wave; lane s reads copy
// s) and, with `reset`, puts the identity back.  Every lane gets the result.
template <bool MAX, class T>
__device__ __forceinline__ T red_fold(T* base, uint32_t t, bool reset) {
  T* p = base + static_cast<size_t>(t) * kRedStripes + (threadIdx.x & 63);
  T v = *p;
  if (reset) *p = MAX ? T(0) : ~T(0);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const T o = __shfl_xor(v, off);
    v = MAX ? (o > v ? o : v) : (o < v ? o : v);
  }
  return v;
}

}  // namespace tit_amd
