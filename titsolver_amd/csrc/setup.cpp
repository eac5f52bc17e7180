// Host-side scene set-up: the dam break of titwcsph (wcsph.cpp:35-142) on a
// lattice with layers of fixed wall particles.  No device is involved.
#include <cmath>
#include <cstddef>
#include <vector>

#include "../../include/tit_hip.h"

namespace {

struct Scene {
  int dim;
  double dr;
  long nx_tank, ny_tank, nz_tank;  // tank size in lattice steps
  long m, nc, kc;                  // fluid column in lattice steps
  int layers;
};

bool make_scene(int dim, double dr, const double* tank, const double* column,
                int layers, Scene* s) {
  if ((dim != 2 && dim != 3) || !(dr > 0.0) || !tank || !column || layers < 0) {
    return false;
  }
  s->dim = dim;
  s->dr = dr;
  s->layers = layers;
  s->nx_tank = std::lround(tank[0] / dr);
  s->ny_tank = std::lround(tank[1] / dr);
  s->nz_tank = dim == 3 ? std::lround(tank[2] / dr) : 0;
  s->m = std::lround(column[0] / dr);  // WATER_M, wcsph.cpp:44
  s->nc = std::lround(column[1] / dr); // WATER_N, wcsph.cpp:45
  s->kc = dim == 3 ? std::lround(column[2] / dr) : 1;
  if (s->m < 1 || s->nc < 1 || s->kc < 1 || s->nx_tank < s->m + 1 ||
      s->ny_tank < 1 || (dim == 3 && s->nz_tank < s->kc + 1)) {
    return false;
  }
  return true;
}

// Wall normal of a fixed particle: the sum of the unit normals of the walls it
// lies behind, normalised (corners and edges get the diagonal).
void wall_normal(const Scene& s, const long* ij, double* n) {
  for (int k = 0; k < s.dim; ++k) n[k] = 0.0;
  if (ij[1] <= 0) n[1] += 1.0;
  if (ij[0] <= 0) n[0] += 1.0;
  if (ij[0] >= s.nx_tank) n[0] -= 1.0;
  if (s.dim == 3) {
    if (ij[2] <= 0) n[2] += 1.0;
    if (ij[2] >= s.nz_tank) n[2] -= 1.0;
  }
  double l2 = 0.0;
  for (int k = 0; k < s.dim; ++k) l2 += n[k] * n[k];
  const double l = std::sqrt(l2);
  for (int k = 0; k < s.dim; ++k) n[k] /= l;
}

// Enumerates the fixed particles (lattice indices) in a fixed order: the
// bottom layers, then the walls normal to x, then (3D) the walls normal to z.
template <class F>
void for_each_fixed(const Scene& s, F&& f) {
  const long L = s.layers;
  long ij[3] = {0, 0, 0};
  const long zlo = s.dim == 3 ? -(L - 1) : 0;
  const long zhi = s.dim == 3 ? s.nz_tank + (L - 1) : 0;
  for (long l = 0; l < L; ++l) {  // bottom, y = -l dr
    for (long i = -(L - 1); i <= s.nx_tank + (L - 1); ++i) {
      for (long k = zlo; k <= zhi; ++k) {
        ij[0] = i;
        ij[1] = -l;
        ij[2] = k;
        f(ij);
      }
    }
  }
  for (long l = 0; l < L; ++l) {  // x walls, x = -l dr and x = W + l dr
    for (int side = 0; side < 2; ++side) {
      for (long j = 1; j <= s.ny_tank; ++j) {
        for (long k = zlo; k <= zhi; ++k) {
          ij[0] = side == 0 ? -l : s.nx_tank + l;
          ij[1] = j;
          ij[2] = k;
          f(ij);
        }
      }
    }
  }
  if (s.dim == 3) {  // z walls between the x walls
    for (long l = 0; l < L; ++l) {
      for (int side = 0; side < 2; ++side) {
        for (long j = 1; j <= s.ny_tank; ++j) {
          for (long i = 1; i <= s.nx_tank - 1; ++i) {
            ij[0] = i;
            ij[1] = j;
            ij[2] = side == 0 ? -l : s.nz_tank + l;
            f(ij);
          }
        }
      }
    }
  }
}

// Hydrostatic pressure of the column from the Poisson solution of
// titwcsph (wcsph.cpp:129-138): p = rho0 g (H - y)
//   - sum_{k odd < 100} 8 rho0 g H/(k pi)^2 exp(k pi (x - L)/(2H)) cos(k pi y/(2H)).
double column_pressure(double x, double y, double L, double H, double rho0,
                       double g) {
  double p = rho0 * g * (H - y);
  const double pi = 3.14159265358979323846;
  for (int k = 1; k < 100; k += 2) {
    const double kp = static_cast<double>(k) * pi;
    p -= 8 * rho0 * g * H / (kp * kp) *
         (std::exp(kp * (x - L) / (2 * H)) * std::cos(kp * y / (2 * H)));
  }
  return p;
}

}  // namespace

extern "C" {

tit_status tit_dam_break_count(int dim, double dr, const double* tank,
                               const double* column, int layers,
                               size_t* n_fluid, size_t* n_fixed) {
  Scene s;
  if (!make_scene(dim, dr, tank, column, layers, &s)) return TIT_E_ARG;
  if (n_fluid) *n_fluid = static_cast<size_t>(s.m * s.nc * s.kc);
  size_t nf = 0;
  for_each_fixed(s, [&](const long*) { ++nf; });
  if (n_fixed) *n_fixed = nf;
  return TIT_OK;
}

tit_status tit_dam_break_fill(int dim, double dr, const double* tank,
                              const double* column, int layers, double rho0,
                              double g, double cs0, double* r, double* rho,
                              double* normal) {
  Scene s;
  if (!make_scene(dim, dr, tank, column, layers, &s) || !r || !rho ||
      !(cs0 > 0.0)) {
    return TIT_E_ARG;
  }
  const double L = column[0], H = column[1];
  size_t a = 0;
  // Fluid particles, i outer, j inner (wcsph.cpp:104-109).
  for (long i = 0; i < s.m; ++i) {
    for (long j = 0; j < s.nc; ++j) {
      for (long k = 0; k < s.kc; ++k) {
        const double x = dr * (static_cast<double>(i) + 1.0);
        const double y = dr * (static_cast<double>(j) + 1.0);
        r[a * dim + 0] = x;
        r[a * dim + 1] = y;
        if (dim == 3) r[a * dim + 2] = dr * (static_cast<double>(k) + 1.0);
        const double p = column_pressure(x, y, L, H, rho0, g);
        rho[a] = rho0 + p / (cs0 * cs0);  // wcsph.cpp:141
        ++a;
      }
    }
  }
  size_t e = 0;
  for_each_fixed(s, [&](const long* ij) {
    for (int k = 0; k < dim; ++k) r[a * dim + k] = dr * static_cast<double>(ij[k]);
    rho[a] = rho0;  // wcsph.cpp:124-127
    if (normal) wall_normal(s, ij, normal + e * dim);
    ++a;
    ++e;
  });
  return TIT_OK;
}

}  // extern "C"
