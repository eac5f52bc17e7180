// Semi-analytic wall boundary, 2D (fluid_equations.hpp:168-193, 243-247,
// 277-291).  Per stage, after the boundary extrapolation, for every fluid
// particle:
//   grad gamma_a = sum over faces within 2h (ascending index) of Kernel::flux;
//   gamma_a      = [winding number > 1/2] - sum of Kernel::antigrad_flux at
//                  r_a + (2 gamma - 1) n_a h^2, when |grad gamma_a| is not tiny;
//   face terms   drho: - rho_s v_as . grad gamma_as / gamma_a
//                dv:   (P_as grad gamma_as - Pi_as |grad gamma_as|) / gamma_a
// with s the face's two vertex (fixed) particles, their fields averaged.  The
// stage kernel adds these to its pair sums divided by gamma_a.  The faces
// within 2h come from a static grid of cells of width 2h listing every face
// whose box reaches the cell grown by 2h; the exact segment-sphere predicate
// then selects them in ascending order, as the reference's sorted face search
// does (particle_mesh.hpp:150-162).  Cells that no face reaches are wholly
// inside or outside the containment polygon; the host classifies them once,
// so only particles near a wall evaluate the winding number.  A first pass (one
// thread per particle) settles the particles that no face reaches and lists
// the others; the second gives each listed particle 16 lanes, one per
// candidate face, so that a particle's face sums are a few evaluations deep
// instead of one chain of dependent FP64 transcendentals per face.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"
#include "sph_kernels.hpp"
#include "wall.hpp"

namespace tit_amd {

namespace {

constexpr int kWallBlock = 256;

__device__ __forceinline__ void normalize2(double x, double y, double* out) {
  const double n2 = x * x + y * y;
  if (n2 >= kWallTiny * kWallTiny) {
    const double s = sqrt(n2);
    out[0] = x / s;
    out[1] = y / s;
  } else {
    out[0] = 0.0;
    out[1] = 0.0;
  }
}

// Grid cell of a point (escaped points clamp to the border).
__device__ __forceinline__ uint32_t wall_cell(const WallArgs& A, double x, double y) {
  int c[2];
  const double xy[2] = {x, y};
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double f = (xy[k] - A.lo[k]) * A.inv_cw[k];
    c[k] = !(f > 0.0) ? 0 : (f >= static_cast<double>(A.nc[k] - 1) ? A.nc[k] - 1 : static_cast<int>(f));
  }
  return static_cast<uint32_t>(c[0]) * static_cast<uint32_t>(A.nc[1]) + static_cast<uint32_t>(c[1]);
}

__global__ void k_wall_clear(uint32_t* near_count) { near_count[0] = 0u; }

// Pass 1, one thread per particle.  A fluid particle whose cell no face
// reaches has gamma = 1 (or 0 outside the polygon) and no face terms; the
// others are listed for pass 2 (one atomic per wave).
__global__ __launch_bounds__(kWallBlock) void k_wall_scan(WallArgs A) {
  const uint32_t i = blockIdx.x * kWallBlock + threadIdx.x;
  bool near = false;
  if (i < A.n && !(A.tag[i] & 1u)) {
    const uint32_t cell = wall_cell(A, A.s.x[0][i], A.s.x[1][i]);
    near = A.cell_faces_off[cell] != A.cell_faces_off[cell + 1];
    if (!near) {
      double* o = A.out + 4 * static_cast<size_t>(i);
      o[0] = 0.0;
      o[1] = 0.0;
      o[2] = 0.0;
      o[3] = A.cell_inside[cell] ? 1.0 : 0.0;
    }
  }
  const uint64_t m = __ballot(near);
  if (m == 0) {
    if (A.near_idx && i < A.n) A.near_idx[i] = ~0u;
    return;
  }
  const int lane = static_cast<int>(threadIdx.x % 64);
  const int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(A.near_count, static_cast<uint32_t>(__popcll(m)));
  base = __shfl(base, leader);
  const uint64_t below = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
  const uint32_t g = base + static_cast<uint32_t>(__popcll(below));
  if (near) A.near[g] = i;
  if (A.near_idx && i < A.n) A.near_idx[i] = near ? g : ~0u;
}

constexpr int kWallLanes = 16;

__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = kWallLanes / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWallLanes);
  return v;
}

// Pass 2, kWallLanes lanes per listed particle: each lane takes every
// kWallLanes-th candidate face, the group sums in a fixed tree.
template <int KERNEL>
__global__ __launch_bounds__(kWallBlock) void k_wall_near(WallArgs A) {
  const uint32_t n_near = A.near_count[0];
  const int lane = static_cast<int>(threadIdx.x % kWallLanes);
  const uint32_t groups = gridDim.x * (kWallBlock / kWallLanes);
  const DevParams& P = A.P;
  const double h = P.h, R = 2.0 * h;
  for (uint32_t g = blockIdx.x * (kWallBlock / kWallLanes) + threadIdx.x / kWallLanes; g < n_near;
       g += groups) {
    const uint32_t i = A.near[g];
    const double x = A.s.x[0][i], y = A.s.x[1][i];
    const uint32_t cell = wall_cell(A, x, y);
    const uint32_t fb = A.cell_faces_off[cell], fe = A.cell_faces_off[cell + 1];

    double gx = 0.0, gy = 0.0, n_acc = 0.0;
    for (uint32_t j = fb + lane; j < fe; j += kWallLanes) {
      const WallFace f = A.faces[A.cell_faces[j]];
      if (!wall_face_intersects(f, x, y, R)) continue;
      double fl[2];
      wall_face_flux<KERNEL>(f, x, y, h, fl);
      gx += fl[0];
      gy += fl[1];
      n_acc += 1.0;
    }
    gx = group_sum(gx);
    gy = group_sum(gy);
    n_acc = group_sum(n_acc);
    double wind = 0.0;
    for (uint32_t e = lane; e < A.n_cedges; e += kWallLanes) {
      wind += wall_winding(A.cedges + 4 * e, 1, x, y);
    }
    wind = group_sum(wind);
    double gamma = wind > 0.5 ? 1.0 : 0.0;
    const double ng = sqrt(gx * gx + gy * gy);
    double d[3] = {0.0, 0.0, 0.0};
    if (n_acc > 0.0 && !(fabs(ng) <= kWallTiny)) {
      const double nax = gx / ng, nay = gy / ng;
      const double sgn = 2.0 * gamma - 1.0;
      const double xs = x + sgn * nax * (h * h), ys = y + sgn * nay * (h * h);
      double anti = 0.0;
      for (uint32_t j = fb + lane; j < fe; j += kWallLanes) {
        const WallFace f = A.faces[A.cell_faces[j]];
        if (!wall_face_intersects(f, x, y, R)) continue;
        anti += wall_face_antigrad_flux<KERNEL>(f, xs, ys, h);
      }
      gamma -= group_sum(anti);
    }
    if (n_acc > 0.0) {
      const double rho_a = A.s.rho[i];
      const double va[2] = {A.s.v[0][i], A.s.v[1][i]};
      const double pa = eos_pressure(P, rho_a) / (rho_a * rho_a);
      for (uint32_t j = fb + lane; j < fe; j += kWallLanes) {
        const WallFace f = A.faces[A.cell_faces[j]];
        if (!wall_face_intersects(f, x, y, R)) continue;
        double gf[2];
        wall_face_flux<KERNEL>(f, x, y, h, gf);
        const double r0 = A.vert_rho[f.s0], r1 = A.vert_rho[f.s1];
        const double rho_s = (r0 + r1) / 2.0;
        const double p_s = (eos_pressure(P, r0) + eos_pressure(P, r1)) / 2.0;
        // the wall particles do not move: v_as = v_a, r_as = r_a - face midpoint
        const double vg = va[0] * gf[0] + va[1] * gf[1];
        d[0] -= rho_s * vg / gamma;
        const double P_as = rho_s * (pa + p_s / (rho_s * rho_s));
        double n_s[2], t_as[2];
        normalize2(gf[0], gf[1], n_s);
        const double vn = va[0] * n_s[0] + va[1] * n_s[1];
        normalize2(va[0] - vn * n_s[0], va[1] - vn * n_s[1], t_as);
        const double rax = x - (f.ax + f.bx) / 2.0, ray = y - (f.ay + f.by) / 2.0;
        const double rn = rax * n_s[0] + ray * n_s[1];
        const double dr_as = h / 2.0 > rn ? h / 2.0 : rn;
        const double pc = 2.0 * P.mu / (rho_a * dr_as) * (va[0] * t_as[0] + va[1] * t_as[1]);
        const double gn = sqrt(gf[0] * gf[0] + gf[1] * gf[1]);
        d[1] += (P_as * gf[0] - pc * t_as[0] * gn) / gamma;
        d[2] += (P_as * gf[1] - pc * t_as[1] * gn) / gamma;
      }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = group_sum(d[k]);
    if (lane == 0) {
      double* o = A.out + 4 * static_cast<size_t>(i);
      o[0] = d[0];
      o[1] = d[1];
      o[2] = d[2];
      o[3] = gamma;
    }
    if (A.pi_face) {
      // apply_shifts face terms (fluid_equations.hpp:342-349), the wall
      // particles at rest: r_sa = face midpoint - r_a, v_sa = -v_a,
      // rho_sa = mean vertex density - rho_a
      double t[12] = {};
      const double rho_a = A.s.rho[i];
      const double va[2] = {A.s.v[0][i], A.s.v[1][i]};
      for (uint32_t j = fb + lane; j < fe; j += kWallLanes) {
        const WallFace f = A.faces[A.cell_faces[j]];
        if (!wall_face_intersects(f, x, y, R)) continue;
        double gf[2];
        wall_face_flux<KERNEL>(f, x, y, h, gf);
        const double r_sa[2] = {(f.ax + f.bx) / 2.0 - x, (f.ay + f.by) / 2.0 - y};
        const double v_sa[2] = {0.0 - va[0], 0.0 - va[1]};
        const double rho_sa = (A.vert_rho[f.s0] + A.vert_rho[f.s1]) / 2.0 - rho_a;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          t[p] -= gf[p] / gamma;
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            t[2 + 2 * p + q] -= r_sa[p] * gf[q] / gamma;
            t[6 + 2 * p + q] -= v_sa[p] * gf[q] / gamma;
          }
          t[10 + p] -= rho_sa * gf[p] / gamma;
        }
      }
#pragma unroll
      for (int k = 0; k < 12; ++k) t[k] = group_sum(t[k]);
      if (lane == 0) {
        double* o = A.pi_face + 12 * static_cast<size_t>(g);
#pragma unroll
        for (int k = 0; k < 12; ++k) o[k] = t[k];
      }
    }
  }
}

template <int KERNEL>
void fixed_gamma(const WallHost& W, double h, const double* r, uint32_t n, double* gamma,
                 double* grad) {
  const double R = 2.0 * h;
  for (uint32_t e = 0; e < n; ++e) {
    const double x = r[2 * e], y = r[2 * e + 1];
    int c[2];
    const double xy[2] = {x, y};
    for (int k = 0; k < 2; ++k) {
      const double f = (xy[k] - W.lo[k]) * W.inv_cw[k];
      c[k] = !(f > 0.0) ? 0 : (f >= static_cast<double>(W.nc[k] - 1) ? W.nc[k] - 1 : static_cast<int>(f));
    }
    const uint32_t cell = static_cast<uint32_t>(c[0]) * static_cast<uint32_t>(W.nc[1]) +
                          static_cast<uint32_t>(c[1]);
    double gg[2] = {0.0, 0.0};
    for (uint32_t j = W.cell_off[cell]; j < W.cell_off[cell + 1]; ++j) {
      const WallFace& f = W.faces[W.cell_faces[j]];
      if (!wall_face_intersects(f, x, y, R)) continue;
      double fl[2];
      wall_face_flux<KERNEL>(f, x, y, h, fl);
      gg[0] += fl[0];
      gg[1] += fl[1];
    }
    double g = wall_winding(W.cedges.data(), static_cast<uint32_t>(W.cedges.size() / 4), x, y) > 0.5
                   ? 1.0
                   : 0.0;
    const double ng = std::sqrt(gg[0] * gg[0] + gg[1] * gg[1]);
    if (!(std::fabs(ng) <= kWallTiny)) {
      const double nax = gg[0] / ng, nay = gg[1] / ng;
      const double sgn = 2.0 * g - 1.0;
      const double xs = x + sgn * nax * (h * h), ys = y + sgn * nay * (h * h);
      for (uint32_t j = W.cell_off[cell]; j < W.cell_off[cell + 1]; ++j) {
        const WallFace& f = W.faces[W.cell_faces[j]];
        if (!wall_face_intersects(f, x, y, R)) continue;
        g -= wall_face_antigrad_flux<KERNEL>(f, xs, ys, h);
      }
    }
    gamma[e] = g;
    grad[2 * e] = gg[0];
    grad[2 * e + 1] = gg[1];
  }
}

}  // namespace

bool wall_build(const double* verts, uint32_t n_verts, const uint32_t* faces, uint32_t n_faces,
                const double* cverts, uint32_t n_cverts, const uint32_t* cfaces,
                uint32_t n_cfaces, const double* lo, const double* hi, double h, WallHost* W,
                std::string* err) {
  W->faces.clear();
  W->cedges.clear();
  for (uint32_t f = 0; f < n_faces; ++f) {
    const uint32_t ia = faces[2 * f], ib = faces[2 * f + 1];
    if (ia >= n_verts || ib >= n_verts) {
      *err = "wall face names a vertex out of range";
      return false;
    }
    WallFace w{};
    w.ax = verts[2 * ia];
    w.ay = verts[2 * ia + 1];
    w.bx = verts[2 * ib];
    w.by = verts[2 * ib + 1];
    // normal = normalize(cross(b - a)), cross(v) = (v1, -v0); e = normalize(b - a)
    const double bax = w.bx - w.ax, bay = w.by - w.ay;
    const double wnx = bay, wny = -bax;
    const double n2 = wnx * wnx + wny * wny;
    if (n2 >= kWallTiny * kWallTiny) {
      const double s = std::sqrt(n2);
      w.nx = wnx / s;
      w.ny = wny / s;
    }
    const double e2 = bax * bax + bay * bay;
    if (e2 >= kWallTiny * kWallTiny) {
      const double s = std::sqrt(e2);
      w.ex = bax / s;
      w.ey = bay / s;
    }
    w.s0 = ia;
    w.s1 = ib;
    W->faces.push_back(w);
  }
  for (uint32_t f = 0; f < n_cfaces; ++f) {
    const uint32_t ia = cfaces[2 * f], ib = cfaces[2 * f + 1];
    if (ia >= n_cverts || ib >= n_cverts) {
      *err = "containment face names a vertex out of range";
      return false;
    }
    W->cedges.insert(W->cedges.end(),
                     {cverts[2 * ia], cverts[2 * ia + 1], cverts[2 * ib], cverts[2 * ib + 1]});
  }
  // grid of cells of width >= 2h over the tank box, at most 4096 per axis
  const double R = 2.0 * h;
  for (int k = 0; k < 2; ++k) {
    const double span = hi[k] - lo[k];
    if (!(span > 0.0)) {
      *err = "empty tank box";
      return false;
    }
    const double cw = std::max(R, span / 4096.0);
    W->nc[k] = std::max(1, static_cast<int>(std::ceil(span / cw)));
    W->lo[k] = lo[k];
    W->inv_cw[k] = 1.0 / cw;
  }
  const size_t n_cells = static_cast<size_t>(W->nc[0]) * static_cast<size_t>(W->nc[1]);
  std::vector<std::vector<uint32_t>> lists(n_cells);
  const double margin = R * (1.0 + 1e-6);
  for (uint32_t f = 0; f < n_faces; ++f) {
    const WallFace& w = W->faces[f];
    int c0[2], c1[2];
    const double fa[2] = {w.ax, w.ay}, fb2[2] = {w.bx, w.by};
    for (int k = 0; k < 2; ++k) {
      const double l = std::min(fa[k], fb2[k]) - margin, u = std::max(fa[k], fb2[k]) + margin;
      const double fl = (l - W->lo[k]) * W->inv_cw[k], fu = (u - W->lo[k]) * W->inv_cw[k];
      c0[k] = fl <= 0.0 ? 0 : std::min(W->nc[k] - 1, static_cast<int>(fl));
      c1[k] = fu <= 0.0 ? 0 : std::min(W->nc[k] - 1, static_cast<int>(fu));
    }
    for (int ix = c0[0]; ix <= c1[0]; ++ix) {
      for (int iy = c0[1]; iy <= c1[1]; ++iy) {
        lists[static_cast<size_t>(ix) * W->nc[1] + iy].push_back(f);
      }
    }
  }
  W->cell_off.assign(n_cells + 1, 0u);
  W->cell_faces.clear();
  W->cell_inside.assign(n_cells, 0u);
  const uint32_t n_ce = static_cast<uint32_t>(W->cedges.size() / 4);
  for (size_t c = 0; c < n_cells; ++c) {
    W->cell_faces.insert(W->cell_faces.end(), lists[c].begin(), lists[c].end());
    W->cell_off[c + 1] = static_cast<uint32_t>(W->cell_faces.size());
    if (lists[c].empty()) {
      const double cx = W->lo[0] + (static_cast<double>(c / W->nc[1]) + 0.5) / W->inv_cw[0];
      const double cy = W->lo[1] + (static_cast<double>(c % W->nc[1]) + 0.5) / W->inv_cw[1];
      W->cell_inside[c] = wall_winding(W->cedges.data(), n_ce, cx, cy) > 0.5 ? 1u : 0u;
    }
  }
  return true;
}

void wall_fixed_gamma(int kernel, const WallHost& W, double h, const double* r, uint32_t n,
                      double* gamma, double* grad) {
  if (kernel == kKernelC4) {
    fixed_gamma<kKernelC4>(W, h, r, n, gamma, grad);
  } else {
    fixed_gamma<kKernelC2>(W, h, r, n, gamma, grad);
  }
}

void launch_wall(hipStream_t s, const WallArgs& a, int n_cu) {
  if (a.n == 0) return;
  k_wall_clear<<<1, 1, 0, s>>>(a.near_count);
  k_wall_scan<<<(a.n + kWallBlock - 1) / kWallBlock, kWallBlock, 0, s>>>(a);
  // persistent groups: every wave leaves its loop once the list is done
  const uint32_t g = static_cast<uint32_t>(std::max(n_cu, 1)) * 8u;
  if (a.P.kernel == kKernelC4) {
    k_wall_near<kKernelC4><<<g, kWallBlock, 0, s>>>(a);
  } else {
    k_wall_near<kKernelC2><<<g, kWallBlock, 0, s>>>(a);
  }
}

}  // namespace tit_amd
