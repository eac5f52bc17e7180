// Semi-analytic wall boundary, 2D: the kernel's flux and antigradient flux
// over wall segments (sph/kernel.hpp:193-314), the segment-sphere predicate of
// the face search (geom/segment.hpp:98-112) and the winding-number containment
// test (geom/segment.hpp:77-82).  Host and device share these functions: the
// host evaluates them once for the fixed particles at upload, the device per
// stage for the fluid particles near a wall.  All FP64.
//
// Unit kernels as monomials, f(q) = sum_k a_k q^k on [0, 2], and their tail
// moments M(q) = int_q^2 s f(s) ds = M0 + sum_j b_j q^(j+2).  For a face at
// distance eta from the point spanning z in [z0, z1] along it (unit
// coordinates, rho = sqrt(z^2 + eta^2)):
//   unit_flux          = int f(rho) dz           = sum_k a_k J_k
//   unit_antigrad_flux = eta int M(rho)/rho^2 dz = M0 atan2(z, eta) + eta sum_j b_j J_j
// with J_n = int rho^n dz: J_0 = z, J_-1 = L = asinh(z / eta),
// J_n = (z rho^n + n eta^2 J_(n-2)) / (n + 1), between the ends clipped to the
// support as unit_segment_integral does (kernel.hpp:282-312).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "device_common.hpp"

namespace tit_amd {

constexpr double kWallTiny = 6.0554544523933395e-06;  // cbrt(DBL_EPSILON)

// One wall segment: vertices a, b, unit normal n = normalize(cross(b - a)),
// cross(v) = (v1, -v0), pointing into the fluid, unit direction e of b - a,
// and the vertex indices (= fixed particles n_fluid + s0, n_fluid + s1).
struct WallFace {
  double ax, ay, bx, by;
  double nx, ny, ex, ey;
  uint32_t s0, s1;
  uint32_t pad[2];
};

template <int KERNEL>
__host__ __device__ constexpr int wall_na() { return KERNEL == kKernelC2 ? 6 : 9; }

template <int KERNEL>
__host__ __device__ constexpr double wall_a(int k) {
  if constexpr (KERNEL == kKernelC2) {
    constexpr double a[6] = {1.0, 0.0, -2.5, 2.5, -15.0 / 16.0, 1.0 / 8.0};
    return a[k];
  } else {
    constexpr double a[9] = {1.0,         0.0,          -7.0 / 3.0, 0.0,         35.0 / 8.0,
                             -14.0 / 3.0, 35.0 / 16.0, -0.5,        35.0 / 768.0};
    return a[k];
  }
}

template <int KERNEL>
__host__ __device__ constexpr double wall_m0() { return KERNEL == kKernelC2 ? 2.0 / 7.0 : 2.0 / 9.0; }

template <int KERNEL>
__host__ __device__ constexpr double wall_b(int k) {
  if constexpr (KERNEL == kKernelC2) {
    constexpr double b[6] = {-0.5, 0.0, 5.0 / 8.0, -0.5, 5.0 / 32.0, -1.0 / 56.0};
    return b[k];
  } else {
    constexpr double b[9] = {-0.5,        0.0,           7.0 / 12.0, 0.0,         -35.0 / 48.0,
                             2.0 / 3.0,   -35.0 / 128.0, 1.0 / 18.0, -7.0 / 1536.0};
    return b[k];
  }
}

// 2D kernel weight 1 / (2 pi M0).
template <int KERNEL>
__host__ __device__ constexpr double wall_weight() {
  return KERNEL == kKernelC2 ? 7.0 / (4.0 * 3.14159265358979323846)
                             : 9.0 / (4.0 * 3.14159265358979323846);
}

// Primitive at z of the flux (ANTI = false) or of the antigradient flux.
template <int KERNEL, bool ANTI>
__host__ __device__ inline double wall_primitive(double eta, double z) {
  const double rho = sqrt(z * z + eta * eta);
  const double L = fabs(eta) <= kWallTiny
                       ? 0.0
                       : copysign(log1p((fabs(z) + z * z / (rho + eta)) / eta), z);
  const double e2 = eta * eta;
  double J[9];
  double rp = rho;
  J[0] = z;
  J[1] = (z * rho + e2 * L) / 2.0;
#pragma unroll
  for (int k = 2; k < 9; ++k) {
    rp = rp * rho;
    J[k] = (z * rp + static_cast<double>(k) * e2 * J[k - 2]) / static_cast<double>(k + 1);
  }
  double r = 0.0;
#pragma unroll
  for (int k = 0; k < wall_na<KERNEL>(); ++k) r += (ANTI ? wall_b<KERNEL>(k) : wall_a<KERNEL>(k)) * J[k];
  if constexpr (ANTI) return wall_m0<KERNEL>() * atan2(z, eta) + eta * r;
  return r;
}

template <int KERNEL, bool ANTI>
__host__ __device__ inline double wall_segment_integral(double eta, double z_min, double z_max) {
  constexpr double cutoff = 2.0;
  if (eta >= cutoff) return 0.0;
  const double z_clip = sqrt(cutoff * cutoff - eta * eta);
  const double z_lo = z_min > -z_clip ? z_min : -z_clip;
  const double z_hi = z_max < z_clip ? z_max : z_clip;
  if (z_lo >= z_hi) return 0.0;
  return wall_primitive<KERNEL, ANTI>(eta, z_hi) - wall_primitive<KERNEL, ANTI>(eta, z_lo);
}

// Segment::intersects(BSphere{c, R}): boxes overlap (inclusive) and the
// segment point closest to c (vertex a when |b - a|^2 is tiny) lies within R.
__host__ __device__ inline bool wall_face_intersects(const WallFace& f, double cx, double cy,
                                                     double R) {
  const double lx = f.ax < f.bx ? f.ax : f.bx, hx = f.ax < f.bx ? f.bx : f.ax;
  const double ly = f.ay < f.by ? f.ay : f.by, hy = f.ay < f.by ? f.by : f.ay;
  if (!(cx - R <= hx && lx <= cx + R)) return false;
  if (!(cy - R <= hy && ly <= cy + R)) return false;
  const double bax = f.bx - f.ax, bay = f.by - f.ay;
  const double len2 = bax * bax + bay * bay;
  double px = f.ax, py = f.ay;
  if (!(fabs(len2) <= kWallTiny)) {
    const double t = ((cx - f.ax) * bax + (cy - f.ay) * bay) / len2;
    if (t > 1.0) {
      px = f.bx;
      py = f.by;
    } else if (!(t < 0.0)) {
      px = f.ax + t * bax;
      py = f.ay + t * bay;
    }
  }
  const double dx = px - cx, dy = py - cy;
  return dx * dx + dy * dy <= R * R;
}

// Distance (in units of h) and projected ends of the face seen from (x, y).
struct WallLocal {
  double d, z0, z1;
};
__host__ __device__ inline WallLocal wall_local(const WallFace& f, double x, double y,
                                                double h_inv) {
  WallLocal l;
  l.d = ((x - f.ax) * f.nx + (y - f.ay) * f.ny) * h_inv;
  l.z0 = ((f.ax - x) * f.ex + (f.ay - y) * f.ey) * h_inv;
  l.z1 = ((f.bx - x) * f.ex + (f.by - y) * f.ey) * h_inv;
  return l;
}

// Kernel::flux: n (w / h) unit_flux(|d|, z0, z1).
template <int KERNEL>
__host__ __device__ inline void wall_face_flux(const WallFace& f, double x, double y, double h,
                                               double* out) {
  const double h_inv = 1.0 / h;
  const double w = wall_weight<KERNEL>() * h_inv;
  const WallLocal l = wall_local(f, x, y, h_inv);
  const double u = wall_segment_integral<KERNEL, false>(fabs(l.d), l.z0, l.z1);
  out[0] = f.nx * w * u;
  out[1] = f.ny * w * u;
}

// Kernel::antigrad_flux: copysign(w, d) unit_antigrad_flux(|d|, z0, z1).
template <int KERNEL>
__host__ __device__ inline double wall_face_antigrad_flux(const WallFace& f, double x, double y,
                                                          double h) {
  const double h_inv = 1.0 / h;
  const WallLocal l = wall_local(f, x, y, h_inv);
  return copysign(wall_weight<KERNEL>(), l.d) *
         wall_segment_integral<KERNEL, true>(fabs(l.d), l.z0, l.z1);
}

// Exact generalised winding number of a polygon given as edges
// (ax, ay, bx, by) around (x, y).
__host__ __device__ inline double wall_winding(const double* edges, uint32_t n_edges, double x,
                                               double y) {
  const double two_pi = 6.283185307179586476925286766559;
  double r = 0.0;
  for (uint32_t k = 0; k < n_edges; ++k) {
    const double* e = edges + 4 * k;
    const double apx = e[0] - x, apy = e[1] - y, bpx = e[2] - x, bpy = e[3] - y;
    r += atan2(apx * bpy - apy * bpx, apx * bpx + apy * bpy) / two_pi;
  }
  return r;
}

// Host copy of the wall: faces, containment edges and the static face grid
// (CSR of ascending face indices per cell).
struct WallHost {
  std::vector<WallFace> faces;
  std::vector<double> cedges;  // 4 per containment edge
  std::vector<uint32_t> cell_off, cell_faces;
  // per cell without candidate faces: 1 when the whole cell is contained (no
  // face crosses it, so the winding number is that of its centre)
  std::vector<uint32_t> cell_inside;
  double lo[2] = {0.0, 0.0}, inv_cw[2] = {0.0, 0.0};
  int nc[2] = {0, 0};
};

// Faces, containment edges and the face grid over the box [lo, hi].
bool wall_build(const double* verts, uint32_t n_verts, const uint32_t* faces, uint32_t n_faces,
                const double* cverts, uint32_t n_cverts, const uint32_t* cfaces,
                uint32_t n_cfaces, const double* lo, const double* hi, double h, WallHost* W,
                std::string* err);

// compute_gamma for n points r (the fixed particles, once at upload).
void wall_fixed_gamma(int kernel, const WallHost& W, double h, const double* r, uint32_t n,
                      double* gamma, double* grad);

}  // namespace tit_amd
