/*
 * FP64 CPU oracle for the WCSPH step -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm (TitSolver, paths relative
 * to its source tree).  Built with -ffp-contract=off: every product and sum is
 * rounded on its own, left to right, which fixes the floating-point convention
 * that the reference leaves to -ffast-math/-ffp-contract=fast and Highway's
 * horizontal sums (cmake/clang.cmake:111-114, core/_vec/vec.hpp:628-640).
 *
 * What pins each part (tests under tests/ check these):
 *   - neighbour search: O(N^2) brute force with the production predicate and
 *     the reference's own naive-search fixture (geom/search.test.cpp:41-163);
 *     integer output, compared bit-exactly.
 *   - kernel: integral of W over the support equals 1 (kernel.test.cpp:33-91),
 *     grad W equals a central difference of W (kernel.test.cpp:95-122),
 *     weights equal the closed forms of the theory guide's table
 *     (manual/.../02-kernel-functions.md:145-152).
 *   - EOS: p(rho0) = 0, rho(p(rho)) = rho, rho(H(rho)) = rho, cs^2 = dp/drho.
 *   - right-hand sides: pairwise momentum conservation, Galilean invariance,
 *     rotation covariance, rest state (drho/dt = 0, dv/dt = -g), linear
 *     velocity field divergence and linear pressure gradient on a lattice,
 *     planted-mistake checks (mutation switch below).
 *   - SSPRK: one step on u' = lambda u equals the cubic Taylor polynomial,
 *     third-order convergence.
 *   - time step: closed form on single-particle states.
 *   - post-integration: the renormalised gradients reproduce linear velocity
 *     and density fields exactly on a perturbed lattice, the free-surface
 *     indicator classifies a lattice block's outer layer, interior and the
 *     layer within 2h of the surface (phi = depth / 2h there), the shift moves
 *     a displaced particle back toward its lattice site, the free-surface
 *     correction keeps rho0 on a uniform block's surface, planted mistakes
 *     5-8 each fail one of these.
 */
#include "tit_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Mutation switch used only by the oracle's own pinning tests: a non-zero
 * value plants a known mistake so the tests can show that their invariants
 * catch it.  0 = faithful restatement. */
static int g_mutation = 0;
void orc_set_mutation(int m) { g_mutation = m; }

/* ~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~ */
/* core/math.hpp:85-163 and core/_vec/vec.hpp:627-679 conventions. */

/* tiny_v<double> = eps^(1/3) (core/math.hpp:140-146). */
static double tiny_v(void) { return cbrt(DBL_EPSILON); }
static int is_tiny(double a) { return fabs(a) <= tiny_v(); }
static double pow2(double a) { return a * a; }

static double dot(int dim, const double* a, const double* b) {
  double r = a[0] * b[0];
  for (int k = 1; k < dim; ++k) r += a[k] * b[k];
  return r;
}
static double norm2(int dim, const double* a) { return dot(dim, a, a); }
static double norm(int dim, const double* a) { return sqrt(norm2(dim, a)); }

/* normalize returns zero for |a|^2 < tiny^2 (core/_vec/vec.hpp:663-679). */
static void normalize(int dim, const double* a, double* out) {
  const double n2 = norm2(dim, a);
  const double eps2 = pow2(tiny_v());
  if (n2 >= eps2) {
    const double s = sqrt(n2);
    for (int k = 0; k < dim; ++k) out[k] = a[k] / s;
  } else {
    for (int k = 0; k < dim; ++k) out[k] = 0.0;
  }
}

/* ~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~ */
/* Smoothing kernels.  Unit polynomials from sph/kernel.gen.cpp:539-546, both
 * with support q < 2 (unit radius 2, kernel.gen.cpp:501-507). */

double orc_kernel_weight(int kernel, int dim) {
  /* 1 / (zero moment * unit sphere area) (sph/kernel.hpp:126-131); closed
   * forms as in manual/.../02-kernel-functions.md:145-152. */
  const double pi = 3.14159265358979323846;
  if (kernel == ORC_KERNEL_WENDLAND_C2) {
    if (dim == 2) return 7.0 / (4.0 * pi);
    if (dim == 3) return 21.0 / (16.0 * pi);
  } else if (kernel == ORC_KERNEL_WENDLAND_C4) {
    if (dim == 2) return 9.0 / (4.0 * pi);
    if (dim == 3) return 495.0 / (256.0 * pi);
  }
  return NAN;
}

double orc_kernel_unit_value(int kernel, double q) {
  if (!(q < 2.0)) return 0.0;
  const double t = 1.0 - q / 2.0;
  if (kernel == ORC_KERNEL_WENDLAND_C2) {
    return (1.0 + 2.0 * q) * (t * t) * (t * t);
  }
  if (kernel == ORC_KERNEL_WENDLAND_C4) {
    const double t3 = t * t * t;
    return (1.0 + 3.0 * q + 35.0 / 12.0 * q * q) * (t3 * t3);
  }
  return NAN;
}

double orc_kernel_unit_deriv(int kernel, double q) {
  if (!(q < 2.0)) return 0.0;
  const double t = 1.0 - q / 2.0;
  if (kernel == ORC_KERNEL_WENDLAND_C2) {
    /* d/dq (1+2q)(1-q/2)^4 = -5 q (1-q/2)^3 */
    return -5.0 * q * (t * t * t);
  }
  if (kernel == ORC_KERNEL_WENDLAND_C4) {
    /* d/dq (1+3q+35/12 q^2)(1-q/2)^6 = -14/3 q (1 + 5q/2) (1-q/2)^5 */
    const double t5 = t * t * t * t * t;
    return -14.0 / 3.0 * q * (1.0 + 2.5 * q) * t5;
  }
  return NAN;
}

double orc_kernel_radius(int kernel, double h) {
  (void)kernel;
  return 2.0 * h; /* sph/kernel.hpp:134-139 with unit radius 2 */
}

static double ipow(double a, int d) {
  double r = a;
  for (int k = 1; k < d; ++k) r *= a;
  return r;
}

/* sph/kernel.hpp:142-151 */
double orc_kernel_value(int kernel, int dim, const double* x, double h) {
  const double h_inv = 1.0 / h;
  const double w = orc_kernel_weight(kernel, dim) * ipow(h_inv, dim);
  const double q = h_inv * norm(dim, x);
  return w * orc_kernel_unit_value(kernel, q);
}

/* sph/kernel.hpp:154-163 */
void orc_kernel_grad(int kernel, int dim, const double* x, double h,
                     double* out) {
  const double h_inv = 1.0 / h;
  const double w = orc_kernel_weight(kernel, dim) * ipow(h_inv, dim);
  const double q = h_inv * norm(dim, x);
  double nx[3] = {0.0, 0.0, 0.0};
  normalize(dim, x, nx);
  const double s = w * orc_kernel_unit_deriv(kernel, q);
  for (int k = 0; k < dim; ++k) out[k] = s * (nx[k] * h_inv);
}

/* ~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~ */
/* Tait equation of state (sph/equation_of_state.hpp:19-72). */

static double eos_B(const orc_params* prm) {
  return prm->rho0 * pow2(prm->cs0) / prm->xi;
}
double orc_eos_pressure(const orc_params* prm, double rho) {
  return eos_B(prm) * (pow(rho / prm->rho0, prm->xi) - 1.0); /* :37-40 */
}
double orc_eos_density_from_pressure(const orc_params* prm, double p) {
  return prm->rho0 * pow(1.0 + p / eos_B(prm), 1.0 / prm->xi); /* :43-46 */
}
double orc_eos_sound_speed(const orc_params* prm, double rho) {
  return prm->cs0 * pow(rho / prm->rho0, (prm->xi - 1.0) / 2.0); /* :49-52 */
}
double orc_eos_potential(const orc_params* prm, double rho) {
  const double xi_1 = prm->xi - 1.0; /* :55-58 */
  return pow2(prm->cs0) * (pow(rho / prm->rho0, xi_1) - 1.0) / xi_1;
}
double orc_eos_density_from_potential(const orc_params* prm, double H) {
  const double xi_1 = prm->xi - 1.0; /* :61-64 */
  return prm->rho0 * pow(1.0 + xi_1 * H / pow2(prm->cs0), 1.0 / xi_1);
}

void orc_eos_pressure_array(const orc_params* prm, size_t n, const double* rho,
                            double* p) {
  for (size_t i = 0; i < n; ++i) p[i] = orc_eos_pressure(prm, rho[i]);
}

/* ~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~ */
/* Uniform grid index (geom/search/grid_search.hpp:73-130, geom/grid.hpp:86-169,
 * geom/bbox.hpp:76-119, geom/point_range.hpp:82-87, geom/bsphere.hpp:47-54). */

typedef struct grid_index {
  int dim;
  double lo[3], hi[3];
  size_t nc[3];
  double cext[3], inv[3];
  size_t ncells;
  size_t* cell_off; /* ncells + 1 */
  uint32_t* cell_pts;
} grid_index;

static void grid_cell_index(const grid_index* g, const double* p, size_t* out) {
  for (int k = 0; k < g->dim; ++k) {
    const double f = (p[k] - g->lo[k]) * g->inv[k]; /* grid.hpp:105-112 */
    out[k] = (size_t)f;
  }
}
static size_t grid_flatten(const grid_index* g, const size_t* idx) {
  size_t flat = idx[0]; /* grid.hpp:120-128: last axis fastest */
  for (int k = 1; k < g->dim; ++k) flat = g->nc[k] * flat + idx[k];
  return flat;
}

static int grid_build(grid_index* g, int dim, size_t n, const double* pts,
                      double hint) {
  memset(g, 0, sizeof(*g));
  g->dim = dim;
  if (n == 0) return 0;
  for (int k = 0; k < dim; ++k) g->lo[k] = g->hi[k] = pts[k];
  for (size_t i = 1; i < n; ++i) {
    for (int k = 0; k < dim; ++k) {
      const double x = pts[i * dim + k];
      if (x < g->lo[k]) g->lo[k] = x;
      if (x > g->hi[k]) g->hi[k] = x;
    }
  }
  g->ncells = 1;
  for (int k = 0; k < dim; ++k) {
    g->lo[k] -= hint / 2.0; /* grow(size_hint / 2), grid_search.hpp:55 */
    g->hi[k] += hint / 2.0;
    const double ext = g->hi[k] - g->lo[k];
    double ncf = ceil(ext / hint); /* set_cell_extents, grid.hpp:90-98 */
    if (ncf < 1.0) ncf = 1.0;
    g->nc[k] = (size_t)ncf;
    g->cext[k] = ext / ncf;
    g->inv[k] = 1.0 / g->cext[k];
    g->ncells *= g->nc[k];
  }
  g->cell_off = (size_t*)calloc(g->ncells + 1, sizeof(size_t));
  g->cell_pts = (uint32_t*)malloc(n * sizeof(uint32_t));
  size_t* cell_of = (size_t*)malloc(n * sizeof(size_t));
  if (!g->cell_off || !g->cell_pts || !cell_of) {
    free(cell_of);
    return -1;
  }
  size_t ci[3];
  for (size_t i = 0; i < n; ++i) { /* count, grid_search.hpp:61-66 */
    grid_cell_index(g, pts + i * dim, ci);
    cell_of[i] = grid_flatten(g, ci);
    g->cell_off[cell_of[i] + 1] += 1;
  }
  for (size_t c = 0; c < g->ncells; ++c) g->cell_off[c + 1] += g->cell_off[c];
  size_t* fill = (size_t*)malloc(g->ncells * sizeof(size_t));
  if (!fill) {
    free(cell_of);
    return -1;
  }
  memcpy(fill, g->cell_off, g->ncells * sizeof(size_t));
  for (size_t i = 0; i < n; ++i) g->cell_pts[fill[cell_of[i]]++] = (uint32_t)i;
  free(fill);
  free(cell_of);
  return 0;
}

static void grid_free(grid_index* g) {
  free(g->cell_off);
  free(g->cell_pts);
  g->cell_off = NULL;
  g->cell_pts = NULL;
}

/* Inclusive containment predicate of BSphere::contains (bsphere.hpp:52-54):
 * d = point - center, s = d0*d0 + d1*d1 (+ d2*d2), s <= r^2. */
static int sphere_contains(int dim, const double* center, double r2,
                           const double* point) {
  double d[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < dim; ++k) d[k] = point[k] - center[k];
  return norm2(dim, d) <= r2;
}

typedef struct ubuf {
  uint32_t* data;
  size_t size, cap;
} ubuf;

static int ubuf_push(ubuf* b, uint32_t v) {
  if (b->size == b->cap) {
    size_t nc = b->cap ? 2 * b->cap : 1024;
    uint32_t* nd = (uint32_t*)realloc(b->data, nc * sizeof(uint32_t));
    if (!nd) return -1;
    b->data = nd;
    b->cap = nc;
  }
  b->data[b->size++] = v;
  return 0;
}

static void sort_u32(uint32_t* a, size_t n) { /* insertion: lists are short */
  for (size_t i = 1; i < n; ++i) {
    const uint32_t x = a[i];
    size_t j = i;
    while (j > 0 && a[j - 1] > x) {
      a[j] = a[j - 1];
      --j;
    }
    a[j] = x;
  }
}

/* GridIndex::search (grid_search.hpp:117-130) with cells_intersecting
 * (grid.hpp:160-169), followed by the per-particle sort of
 * ParticleMesh::search_ (particle_mesh.hpp:137-147). */
static int grid_query(const grid_index* g, const double* pts,
                      const double* center, double radius, ubuf* out) {
  const int dim = g->dim;
  const double r2 = pow2(radius);
  double slo[3], shi[3];
  size_t clo[3], chi[3];
  for (int k = 0; k < dim; ++k) {
    const double half = g->cext[k] / 2.0;
    double lo = center[k] - radius; /* BSphere::box: BBox{c}.grow(r) */
    double hi = center[k] + radius;
    lo -= half; /* grow(half) */
    hi += half;
    if (g->lo[k] > lo) lo = g->lo[k]; /* intersect(box) */
    if (g->hi[k] < hi) hi = g->hi[k];
    lo += half; /* shrink(half) */
    hi -= half;
    slo[k] = lo;
    shi[k] = hi;
  }
  grid_cell_index(g, slo, clo);
  grid_cell_index(g, shi, chi);
  const size_t start = out->size;
  size_t c[3] = {0, 0, 0};
  for (c[0] = clo[0]; c[0] <= chi[0]; ++c[0]) {
    for (c[1] = clo[1]; c[1] <= chi[1]; ++c[1]) {
      size_t z_lo = dim == 3 ? clo[2] : 0, z_hi = dim == 3 ? chi[2] : 0;
      for (size_t z = z_lo; z <= z_hi; ++z) {
        c[2] = z;
        const size_t f = grid_flatten(g, c);
        for (size_t j = g->cell_off[f]; j < g->cell_off[f + 1]; ++j) {
          const uint32_t b = g->cell_pts[j];
          if (sphere_contains(dim, center, r2, pts + (size_t)b * dim)) {
            if (ubuf_push(out, b)) return -1;
          }
        }
      }
    }
  }
  sort_u32(out->data + start, out->size - start);
  return 0;
}

/* Builds the full adjacency (CSR) of n points with a grid of the given hint. */
static int build_adjacency(int dim, size_t n, const double* pts, double radius,
                           double hint, uint64_t* off, ubuf* idx) {
  grid_index g;
  if (grid_build(&g, dim, n, pts, hint)) return -1;
  idx->size = 0;
  off[0] = 0;
  for (size_t a = 0; a < n; ++a) {
    if (grid_query(&g, pts, pts + a * dim, radius, idx)) {
      grid_free(&g);
      return -1;
    }
    off[a + 1] = idx->size;
  }
  grid_free(&g);
  return 0;
}

static long long copy_out(const ubuf* b, uint32_t* idx, size_t cap) {
  if (b->size > cap) return -(long long)b->size;
  if (b->size) memcpy(idx, b->data, b->size * sizeof(uint32_t));
  return (long long)b->size;
}

long long orc_grid_search(int dim, size_t n, const double* pts, double radius,
                          double size_hint, uint64_t* off, uint32_t* idx,
                          size_t cap) {
  if ((dim != 2 && dim != 3) || !(radius > 0.0) || !(size_hint > 0.0)) {
    return -1;
  }
  off[0] = 0;
  if (n == 0) return 0;
  ubuf b = {0, 0, 0};
  if (build_adjacency(dim, n, pts, radius, size_hint, off, &b)) {
    free(b.data);
    return -1;
  }
  long long r = copy_out(&b, idx, cap);
  free(b.data);
  return r;
}

long long orc_naive_search(int dim, size_t n, const double* pts,
                           double radius, int strict, uint64_t* off,
                           uint32_t* idx, size_t cap) {
  if (dim != 2 && dim != 3) return -1;
  const double r2 = pow2(radius);
  ubuf b = {0, 0, 0};
  off[0] = 0;
  for (size_t a = 0; a < n; ++a) {
    for (size_t j = 0; j < n; ++j) {
      double d[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < dim; ++k) d[k] = pts[j * dim + k] - pts[a * dim + k];
      const double s = norm2(dim, d);
      const int in = strict ? (s < r2 || j == a) : (s <= r2);
      if (in && ubuf_push(&b, (uint32_t)j)) {
        free(b.data);
        return -1;
      }
    }
    off[a + 1] = b.size;
  }
  long long r = copy_out(&b, idx, cap);
  free(b.data);
  return r;
}

/* ~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~ */
/* Semi-analytic wall, 2D (sph/kernel.hpp:193-314, geom/segment.hpp:66-112,
 * fluid_equations.hpp:168-193).
 *
 * The unit kernels as monomials, f(q) = sum_k a_k q^k on [0, 2], and their
 * tail moments M(q) = int_q^2 s f(s) ds = M0 + sum_j b_j q^(j+2) (no linear
 * term: M'(q) = -q f(q)).  The kernel's face integrals in unit coordinates,
 * for a face at distance eta from the point, spanning z in [z0, z1] along it
 * (rho = sqrt(z^2 + eta^2)):
 *   unit_flux          = int f(rho) dz              = sum_k a_k J_k
 *   unit_antigrad_flux = eta int M(rho)/rho^2 dz    = M0 atan2(z, eta) + eta sum_j b_j J_j
 * with J_n(z) = int rho^n dz: J_0 = z, J_-1 = asinh(z/eta) = L,
 * J_n = (z rho^n + n eta^2 J_(n-2)) / (n + 1), each evaluated between the
 * clipped ends as unit_segment_integral does (kernel.hpp:282-312).  The
 * tests check both against numerical quadrature of the defining integrals
 * and the resulting gamma against a direct area integral of W. */
typedef struct unit_poly {
  int na;
  double a[9];
  double M0;
  int nb;
  double b[9];
} unit_poly;
static const unit_poly kPolyC2 = {
    6, {1.0, 0.0, -2.5, 2.5, -15.0 / 16.0, 1.0 / 8.0}, 2.0 / 7.0,
    6, {-0.5, 0.0, 5.0 / 8.0, -0.5, 5.0 / 32.0, -1.0 / 56.0}};
static const unit_poly kPolyC4 = {
    9, {1.0, 0.0, -7.0 / 3.0, 0.0, 35.0 / 8.0, -14.0 / 3.0, 35.0 / 16.0, -0.5, 35.0 / 768.0},
    2.0 / 9.0,
    9, {-0.5, 0.0, 7.0 / 12.0, 0.0, -35.0 / 48.0, 2.0 / 3.0, -35.0 / 128.0, 1.0 / 18.0,
        -7.0 / 1536.0}};
static const unit_poly* unit_poly_of(int kernel) {
  return kernel == ORC_KERNEL_WENDLAND_C4 ? &kPolyC4 : &kPolyC2;
}

/* Primitive of the flux (which = 0) or of the antigradient flux (which = 1)
 * at z. */
static double segment_primitive(const unit_poly* up, int which, double eta, double z) {
  const double rho = sqrt(pow2(z) + pow2(eta));
  const double L = is_tiny(eta) ? 0.0 : copysign(log1p((fabs(z) + pow2(z) / (rho + eta)) / eta), z);
  const double e2 = pow2(eta);
  double J[9];
  double rp = rho; /* rho^n */
  J[0] = z;
  J[1] = (z * rho + e2 * L) / 2.0;
  for (int k = 2; k < 9; ++k) {
    rp = rp * rho;
    J[k] = (z * rp + (double)k * e2 * J[k - 2]) / (double)(k + 1);
  }
  double r = 0.0;
  if (which == 0) {
    for (int k = 0; k < up->na; ++k) r += up->a[k] * J[k];
    return r;
  }
  for (int k = 0; k < up->nb; ++k) r += up->b[k] * J[k];
  return up->M0 * atan2(z, eta) + eta * r;
}

static double unit_segment_integral(int kernel, int which, double eta, double z_min,
                                    double z_max) {
  const double cutoff = 2.0;
  if (eta >= cutoff) return 0.0;
  const double z_clip = sqrt(pow2(cutoff) - pow2(eta));
  const double z_lo = z_min > -z_clip ? z_min : -z_clip;
  const double z_hi = z_max < z_clip ? z_max : z_clip;
  if (z_lo >= z_hi) return 0.0;
  const unit_poly* up = unit_poly_of(kernel);
  return segment_primitive(up, which, eta, z_hi) - segment_primitive(up, which, eta, z_lo);
}

double orc_kernel_unit_flux(int kernel, double eta, double z0, double z1) {
  return unit_segment_integral(kernel, 0, eta, z0, z1);
}
double orc_kernel_unit_antigrad_flux(int kernel, double eta, double z0, double z1) {
  return unit_segment_integral(kernel, 1, eta, z0, z1);
}

/* Segment::normal = normalize(cross(b - a)), cross(v) = (v1, -v0), and the
 * unit direction of b - a (geom/segment.hpp:54-75, core/_vec/vec.hpp:692). */
static void face_frame(const double* a, const double* b, double* n, double* e) {
  const double ba[2] = {b[0] - a[0], b[1] - a[1]};
  const double wn[2] = {ba[1], -ba[0]};
  normalize(2, wn, n);
  normalize(2, ba, e);
}

/* Kernel::flux (kernel.hpp:193-205): n w/h unit_flux(|d|, z_a, z_b),
 * d = (x - a).n / h, z = (vertex - x).e / h (Segment::project). */
void orc_face_flux(int kernel, const double* a, const double* b, const double* x,
                   double h, double* out) {
  double n[2], e[2];
  face_frame(a, b, n, e);
  const double hi = 1.0 / h;
  const double w = orc_kernel_weight(kernel, 2) * hi;
  const double xa[2] = {x[0] - a[0], x[1] - a[1]};
  const double ax[2] = {a[0] - x[0], a[1] - x[1]};
  const double bx[2] = {b[0] - x[0], b[1] - x[1]};
  const double d = dot(2, xa, n) * hi;
  double f = orc_kernel_unit_flux(kernel, fabs(d), dot(2, ax, e) * hi, dot(2, bx, e) * hi);
  if (g_mutation == 9) f = -f; /* planted: face normal reversed */
  for (int k = 0; k < 2; ++k) out[k] = n[k] * w * f;
}

/* Kernel::antigrad_flux (kernel.hpp:207-220): copysign(w, d) unit_antigrad_flux. */
double orc_face_antigrad_flux(int kernel, const double* a, const double* b, const double* x,
                              double h) {
  double n[2], e[2];
  face_frame(a, b, n, e);
  const double hi = 1.0 / h;
  const double w = orc_kernel_weight(kernel, 2);
  const double xa[2] = {x[0] - a[0], x[1] - a[1]};
  const double ax[2] = {a[0] - x[0], a[1] - x[1]};
  const double bx[2] = {b[0] - x[0], b[1] - x[1]};
  const double d = dot(2, xa, n) * hi;
  return copysign(w, d) *
         orc_kernel_unit_antigrad_flux(kernel, fabs(d), dot(2, ax, e) * hi, dot(2, bx, e) * hi);
}

/* Segment::intersects(BSphere) (geom/segment.hpp:98-112, bbox.hpp:62-65,
 * bsphere.hpp:47-54): the boxes overlap (inclusive) and the point of the
 * segment closest to the centre lies within the radius (inclusive).  clamp
 * returns vertex a when |b - a|^2 is tiny, as the reference does. */
int orc_face_intersects(const double* a, const double* b, const double* c, double R) {
  for (int k = 0; k < 2; ++k) {
    const double lo = a[k] < b[k] ? a[k] : b[k];
    const double hi = a[k] < b[k] ? b[k] : a[k];
    if (!(c[k] - R <= hi && lo <= c[k] + R)) return 0;
  }
  const double ba[2] = {b[0] - a[0], b[1] - a[1]};
  const double len2 = norm2(2, ba);
  double p[2] = {a[0], a[1]};
  if (!is_tiny(len2)) {
    const double ca[2] = {c[0] - a[0], c[1] - a[1]};
    const double t = dot(2, ca, ba) / len2;
    if (t > 1.0) {
      p[0] = b[0];
      p[1] = b[1];
    } else if (!(t < 0.0)) {
      p[0] = a[0] + t * ba[0];
      p[1] = a[1] + t * ba[1];
    }
  }
  const double pc[2] = {p[0] - c[0], p[1] - c[1]};
  return norm2(2, pc) <= pow2(R);
}

/* Generalised winding number of the containment polygon around x, the exact
 * sum of Segment::winding_number (geom/segment.hpp:77-82); the reference's
 * fast estimator falls back to this sum near the threshold
 * (winding/fast_winding.hpp:80-92).  contains: winding > 1/2. */
double orc_winding(const orc_walls* w, const double* x) {
  const double two_pi = 6.283185307179586476925286766559;
  double r = 0.0;
  for (size_t f = 0; f < w->n_cfaces; ++f) {
    const double* a = w->cverts + 2 * w->cfaces[2 * f];
    const double* b = w->cverts + 2 * w->cfaces[2 * f + 1];
    const double ap[2] = {a[0] - x[0], a[1] - x[1]};
    const double bp[2] = {b[0] - x[0], b[1] - x[1]};
    r += atan2(ap[0] * bp[1] - ap[1] * bp[0], dot(2, ap, bp)) / two_pi;
  }
  return r;
}

static const double* wall_vert(const orc_walls* w, size_t f, int k) {
  return w->verts + 2 * w->faces[2 * f + k];
}

/* Candidate faces of a point: a uniform grid of cells of width >= R over the
 * faces' box, each cell listing (ascending) the faces whose box grown by R
 * reaches it, as GridFaceIndex does (geom/face_search/grid_face_search.hpp:
 * 43-111).  The exact predicate then selects the faces, so the sets and their
 * order are those of a scan over every face.  Rebuilt when the wall or R
 * changes. */
typedef struct face_grid {
  const orc_walls* w;
  const double* verts;
  const uint32_t* faces;
  size_t n_faces;
  double R;
  double lo[2], inv[2];
  int nc[2];
  uint32_t* off;
  uint32_t* idx;
} face_grid;
static face_grid g_fg;

static int face_cell(const face_grid* g, int k, double x) {
  const double f = (x - g->lo[k]) * g->inv[k];
  if (!(f > 0.0)) return 0;
  if (f >= (double)(g->nc[k] - 1)) return g->nc[k] - 1;
  return (int)f;
}

static const face_grid* faces_near(const orc_walls* w, double R) {
  face_grid* g = &g_fg;
  if (g->w == w && g->verts == w->verts && g->faces == w->faces && g->n_faces == w->n_faces &&
      g->R == R && g->off) {
    return g;
  }
  free(g->off);
  free(g->idx);
  memset(g, 0, sizeof *g);
  double lo[2] = {DBL_MAX, DBL_MAX}, hi[2] = {-DBL_MAX, -DBL_MAX};
  for (size_t v = 0; v < w->n_verts; ++v) {
    for (int k = 0; k < 2; ++k) {
      const double x = w->verts[2 * v + k];
      if (x < lo[k]) lo[k] = x;
      if (x > hi[k]) hi[k] = x;
    }
  }
  for (int k = 0; k < 2; ++k) {
    const double span = hi[k] - lo[k] + 2.0 * R;
    double cw = span / 2048.0 > R ? span / 2048.0 : R;
    g->nc[k] = (int)ceil(span / cw);
    if (g->nc[k] < 1) g->nc[k] = 1;
    g->lo[k] = lo[k] - R;
    g->inv[k] = 1.0 / cw;
  }
  const size_t nc = (size_t)g->nc[0] * (size_t)g->nc[1];
  g->off = (uint32_t*)calloc(nc + 1, sizeof(uint32_t));
  const double m = R * (1.0 + 1e-6);
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1) {
      for (size_t c = 0; c < nc; ++c) g->off[c + 1] += g->off[c];
      g->idx = (uint32_t*)malloc((g->off[nc] ? g->off[nc] : 1) * sizeof(uint32_t));
    }
    uint32_t* fill = NULL;
    if (pass == 1) {
      fill = (uint32_t*)malloc((nc ? nc : 1) * sizeof(uint32_t));
      memcpy(fill, g->off, nc * sizeof(uint32_t));
    }
    for (size_t f = 0; f < w->n_faces; ++f) {
      const double* a = wall_vert(w, f, 0);
      const double* b = wall_vert(w, f, 1);
      int c0[2], c1[2];
      for (int k = 0; k < 2; ++k) {
        c0[k] = face_cell(g, k, (a[k] < b[k] ? a[k] : b[k]) - m);
        c1[k] = face_cell(g, k, (a[k] < b[k] ? b[k] : a[k]) + m);
      }
      for (int ix = c0[0]; ix <= c1[0]; ++ix) {
        for (int iy = c0[1]; iy <= c1[1]; ++iy) {
          const size_t c = (size_t)ix * g->nc[1] + iy;
          if (pass == 0) {
            g->off[c + 1]++;
          } else {
            g->idx[fill[c]++] = (uint32_t)f;
          }
        }
      }
    }
    free(fill);
  }
  g->w = w;
  g->verts = w->verts;
  g->faces = w->faces;
  g->n_faces = w->n_faces;
  g->R = R;
  return g;
}

/* The candidate range of x: faces idx[*b .. *e). */
static void face_range(const face_grid* g, const double* x, uint32_t* b, uint32_t* e) {
  const size_t c = (size_t)face_cell(g, 0, x[0]) * g->nc[1] + face_cell(g, 1, x[1]);
  *b = g->off[c];
  *e = g->off[c + 1];
}

/* compute_gamma (fluid_equations.hpp:168-193) for every particle, over the
 * faces within the kernel radius in ascending order (ParticleMesh::search_,
 * particle_mesh.hpp:150-162). */
int orc_compute_gamma(const orc_params* prm, orc_state* s) {
  const orc_walls* w = prm->walls;
  if (!w || prm->dim != 2 || !s->gamma || !s->grad_gamma) return -1;
  const size_t n = s->n_fluid + s->n_fixed;
  const double h = prm->h, R = orc_kernel_radius(prm->kernel, h);
  const face_grid* fg = faces_near(w, R);
  for (size_t i = 0; i < n; ++i) {
    const double* x = s->r + 2 * i;
    double gg[2] = {0.0, 0.0};
    uint32_t jb, je;
    face_range(fg, x, &jb, &je);
    for (uint32_t j = jb; j < je; ++j) {
      const size_t f = fg->idx[j];
      if (!orc_face_intersects(wall_vert(w, f, 0), wall_vert(w, f, 1), x, R)) continue;
      double fl[2];
      orc_face_flux(prm->kernel, wall_vert(w, f, 0), wall_vert(w, f, 1), x, h, fl);
      gg[0] += fl[0];
      gg[1] += fl[1];
    }
    double gamma = orc_winding(w, x) > 0.5 ? 1.0 : 0.0;
    const double ng = norm(2, gg);
    if (!is_tiny(ng)) {
      const double na[2] = {gg[0] / ng, gg[1] / ng};
      const double sgn = 2.0 * gamma - 1.0;
      const double xs[2] = {x[0] + sgn * na[0] * pow2(h), x[1] + sgn * na[1] * pow2(h)};
      for (uint32_t j = jb; j < je; ++j) {
        const size_t f = fg->idx[j];
        if (!orc_face_intersects(wall_vert(w, f, 0), wall_vert(w, f, 1), x, R)) continue;
        gamma -= orc_face_antigrad_flux(prm->kernel, wall_vert(w, f, 0), wall_vert(w, f, 1), xs, h);
      }
    }
    s->gamma[i] = gamma;
    s->grad_gamma[2 * i] = gg[0];
    s->grad_gamma[2 * i + 1] = gg[1];
  }
  return 0;
}

/* FluidEquations::initialize (fluid_equations.hpp:75-89): gamma, then the
 * fixed particles' masses are multiplied by gamma (:88).  Masses the caller
 * left at zero are the uniform m first. */
int orc_initialize(const orc_params* prm, orc_state* s) {
  if (!s->mass) return -1;
  const size_t n = s->n_fluid + s->n_fixed;
  for (size_t i = 0; i < n; ++i) {
    if (!(s->mass[i] > 0.0)) s->mass[i] = prm->m;
  }
  if (!prm->walls) return 0;
  if (prm->walls->n_verts != s->n_fixed) return -1;
  if (orc_compute_gamma(prm, s)) return -1;
  for (size_t e = s->n_fluid; e < n; ++e) s->mass[e] *= s->gamma[e];
  return 0;
}

/* Mass of particle i, ParticleArray's m (fluid_equations.hpp:41-48): the
 * per-particle value where one is set (positive), else the uniform m. */
static double mass_of(const orc_params* prm, const orc_state* s, size_t i) {
  return (s->mass && s->mass[i] > 0.0) ? s->mass[i] : prm->m;
}

/* ~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~ */
/* Fluid equations (sph/fluid_equations.hpp).  Without walls (prm->walls
 * null): gamma = 1 and grad gamma = 0 everywhere, so the face loops vanish,
 * and fixed particles carry an explicit wall normal in place of
 * normalize(grad_gamma).  With walls (2D, REF terms): the face loops over the
 * wall segments, pair terms divided by gamma_a, per-particle masses. */

/* setup_boundary, fluid_equations.hpp:122-164 */
int orc_setup_boundary(const orc_params* prm, orc_state* s,
                       const uint64_t* off, const uint32_t* idx) {
  const int dim = prm->dim;
  const size_t nf = s->n_fluid, n = s->n_fluid + s->n_fixed;
  for (size_t e = nf; e < n; ++e) {
    for (int k = 0; k < dim; ++k) s->v[e * dim + k] = 0.0; /* no-slip */
    double S_e = 0.0, H_e = 0.0;
    double n_e[3] = {0.0, 0.0, 0.0};
    if (prm->walls) {
      normalize(dim, s->grad_gamma + e * dim, n_e); /* normalize(grad_gamma[e]) */
    } else {
      normalize(dim, s->normal + (e - nf) * dim, n_e);
    }
    for (uint64_t j = off[e]; j < off[e + 1]; ++j) {
      const size_t b = idx[j];
      if (b >= nf) continue; /* fluid neighbours only */
      const double V_b = mass_of(prm, s, b) / s->rho[b];
      double r_be[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < dim; ++k) r_be[k] = s->r[b * dim + k] - s->r[e * dim + k];
      const double W_be = orc_kernel_value(prm->kernel, dim, r_be, prm->h);
      const double H_b = orc_eos_potential(prm, s->rho[b]);
      S_e += V_b * W_be;
      H_e += V_b * (H_b + prm->g * dot(dim, r_be, n_e) * n_e[1]) * W_be;
    }
    s->rho[e] = orc_eos_density_from_potential(prm, is_tiny(S_e) ? 0.0 : H_e / S_e);
  }
  return 0;
}

/* compute_continuity, fluid_equations.hpp:232-260.  The pair loop runs over
 * unique pairs (b < a) of the sorted adjacency with symmetric scatter, as
 * block_for_each over ParticleMesh::block_pairs does, in ascending order. */
/* gamma_a (1 without walls). */
static double gamma_of(const orc_params* prm, const orc_state* s, size_t a) {
  return prm->walls ? s->gamma[a] : 1.0;
}

/* Face terms of compute_continuity (fluid_equations.hpp:243-247): the face's
 * fixed particles are the ones at its vertices, their fields averaged. */
static void wall_continuity(const orc_params* prm, orc_state* s, size_t a) {
  const orc_walls* w = prm->walls;
  const size_t nf = s->n_fluid;
  const double* x = s->r + 2 * a;
  const double R = orc_kernel_radius(prm->kernel, prm->h);
  const face_grid* fg = faces_near(w, R);
  uint32_t jb, je;
  face_range(fg, x, &jb, &je);
  for (uint32_t j = jb; j < je; ++j) {
    const size_t f = fg->idx[j];
    if (!orc_face_intersects(wall_vert(w, f, 0), wall_vert(w, f, 1), x, R)) continue;
    double gg[2];
    orc_face_flux(prm->kernel, wall_vert(w, f, 0), wall_vert(w, f, 1), x, prm->h, gg);
    const size_t s0 = nf + w->faces[2 * f], s1 = nf + w->faces[2 * f + 1];
    const double rho_s = (s->rho[s0] + s->rho[s1]) / 2.0;
    double v_as[2];
    for (int k = 0; k < 2; ++k) v_as[k] = s->v[2 * a + k] - (s->v[2 * s0 + k] + s->v[2 * s1 + k]) / 2.0;
    s->drho_dt[a] -= rho_s * dot(2, v_as, gg) / s->gamma[a];
  }
}

/* Face terms of compute_momentum (fluid_equations.hpp:277-291). */
static void wall_momentum(const orc_params* prm, orc_state* s, size_t a) {
  const orc_walls* w = prm->walls;
  const size_t nf = s->n_fluid;
  const double* x = s->r + 2 * a;
  const double h = prm->h, R = orc_kernel_radius(prm->kernel, h);
  const face_grid* fg = faces_near(w, R);
  uint32_t jb, je;
  face_range(fg, x, &jb, &je);
  for (uint32_t j = jb; j < je; ++j) {
    const size_t f = fg->idx[j];
    if (!orc_face_intersects(wall_vert(w, f, 0), wall_vert(w, f, 1), x, R)) continue;
    double gg[2];
    orc_face_flux(prm->kernel, wall_vert(w, f, 0), wall_vert(w, f, 1), x, h, gg);
    const size_t s0 = nf + w->faces[2 * f], s1 = nf + w->faces[2 * f + 1];
    const double rho_s = (s->rho[s0] + s->rho[s1]) / 2.0;
    const double p_s = (s->p[s0] + s->p[s1]) / 2.0;
    const double P_as = rho_s * (s->p[a] / pow2(s->rho[a]) + p_s / pow2(rho_s));
    double v_as[2], r_as[2], n_s[2], t[2], t_as[2];
    for (int k = 0; k < 2; ++k) {
      v_as[k] = s->v[2 * a + k] - (s->v[2 * s0 + k] + s->v[2 * s1 + k]) / 2.0;
      r_as[k] = s->r[2 * a + k] - (s->r[2 * s0 + k] + s->r[2 * s1 + k]) / 2.0;
    }
    normalize(2, gg, n_s);
    const double vn = dot(2, v_as, n_s);
    for (int k = 0; k < 2; ++k) t[k] = v_as[k] - vn * n_s[k];
    normalize(2, t, t_as);
    const double rn = dot(2, r_as, n_s);
    const double dr_as = h / 2.0 > rn ? h / 2.0 : rn;
    const double pc = 2.0 * prm->mu / (s->rho[a] * dr_as) * dot(2, v_as, t_as);
    const double ng = norm(2, gg);
    for (int k = 0; k < 2; ++k) {
      s->dv_dt[2 * a + k] += (P_as * gg[k] - pc * t_as[k] * ng) / s->gamma[a];
    }
  }
}

int orc_compute_continuity(const orc_params* prm, orc_state* s,
                           const uint64_t* off, const uint32_t* idx) {
  const int dim = prm->dim;
  const size_t n = s->n_fluid + s->n_fixed;
  for (size_t a = 0; a < n; ++a) s->cs[a] = orc_eos_sound_speed(prm, s->rho[a]);
  for (size_t a = 0; a < n; ++a) s->drho_dt[a] = 0.0;
  if (prm->walls) {
    for (size_t a = 0; a < s->n_fluid; ++a) wall_continuity(prm, s, a);
  }
  for (size_t a = 0; a < n; ++a) {
    for (uint64_t j = off[a]; j < off[a + 1]; ++j) {
      const size_t b = idx[j];
      if (b >= a) break; /* sorted: unique pairs b < a */
      double r_ab[3] = {0.0, 0.0, 0.0}, v_ab[3] = {0.0, 0.0, 0.0}, grad[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < dim; ++k) {
        r_ab[k] = s->r[a * dim + k] - s->r[b * dim + k];
        v_ab[k] = s->v[a * dim + k] - s->v[b * dim + k];
      }
      orc_kernel_grad(prm->kernel, dim, r_ab, prm->h, grad);
      const double rho_ab = s->rho[a] - s->rho[b];
      if (prm->flavour == ORC_FLAVOUR_REF) {
        /* Ferrari et al. (2009) diffusion, :253-258 */
        const double cs_ab = s->cs[a] > s->cs[b] ? s->cs[a] : s->cs[b];
        const double rn = norm(dim, r_ab);
        double Psi[3] = {0.0, 0.0, 0.0}, ta[3] = {0.0, 0.0, 0.0}, tb[3] = {0.0, 0.0, 0.0};
        double psi_sign = 1.0;
        if (g_mutation == 12) psi_sign = -1.0; /* planted: anti-diffusive Psi */
        for (int k = 0; k < dim; ++k) Psi[k] = psi_sign * cs_ab * rho_ab * r_ab[k] / rn;
        for (int k = 0; k < dim; ++k) {
          ta[k] = v_ab[k] + Psi[k] / s->rho[b];
          tb[k] = -v_ab[k] + Psi[k] / s->rho[a];
        }
        const int swap = g_mutation == 14; /* planted: the target's own mass */
        s->drho_dt[a] += mass_of(prm, s, swap ? a : b) / gamma_of(prm, s, a) * dot(dim, ta, grad);
        s->drho_dt[b] -= mass_of(prm, s, swap ? b : a) / gamma_of(prm, s, b) * dot(dim, tb, grad);
      } else {
        /* delta-SPH (Molteni & Colagrossi 2009):
         *   psi_ab = 2 (rho_a - rho_b) r_ab / |r_ab|^2,
         *   drho_a/dt += delta h c0 V_b psi_ab . grad W_ab. */
        const double r2 = norm2(dim, r_ab);
        double psi[3] = {0.0, 0.0, 0.0};
        double coef = 2.0 * rho_ab / r2;
        if (g_mutation == 2) coef = -coef; /* planted: anti-diffusive sign */
        for (int k = 0; k < dim; ++k) psi[k] = coef * r_ab[k];
        const double dhc = prm->delta * prm->h * prm->cs0;
        const double vg = dot(dim, v_ab, grad), pg = dot(dim, psi, grad);
        const int swap = g_mutation == 14; /* planted: the target's own mass */
        const double m_a = mass_of(prm, s, swap ? b : a), m_b = mass_of(prm, s, swap ? a : b);
        s->drho_dt[a] += m_b * vg + dhc * (m_b / s->rho[b]) * pg;
        s->drho_dt[b] += m_a * vg - dhc * (m_a / s->rho[a]) * pg;
      }
    }
  }
  return 0;
}

/* compute_momentum, fluid_equations.hpp:267-305 */
int orc_compute_momentum(const orc_params* prm, orc_state* s,
                         const uint64_t* off, const uint32_t* idx) {
  const int dim = prm->dim;
  const size_t nf = s->n_fluid, n = s->n_fluid + s->n_fixed;
  const double m = prm->m;
  for (size_t a = 0; a < n; ++a) {
    s->p[a] = orc_eos_pressure(prm, s->rho[a]);
    s->cs[a] = orc_eos_sound_speed(prm, s->rho[a]);
  }
  for (size_t a = 0; a < n; ++a) {
    for (int k = 0; k < dim; ++k) s->dv_dt[a * dim + k] = 0.0;
    if (a < nf) s->dv_dt[a * dim + 1] = -prm->g; /* unit<1>(r, -g) */
  }
  if (prm->walls && g_mutation != 10) { /* 10 planted: face terms dropped */
    for (size_t a = 0; a < nf; ++a) wall_momentum(prm, s, a);
  }
  for (size_t a = 0; a < n; ++a) {
    for (uint64_t j = off[a]; j < off[a + 1]; ++j) {
      const size_t b = idx[j];
      if (b >= a) break;
      double r_ab[3] = {0.0, 0.0, 0.0}, v_ab[3] = {0.0, 0.0, 0.0}, grad[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < dim; ++k) {
        r_ab[k] = s->r[a * dim + k] - s->r[b * dim + k];
        v_ab[k] = s->v[a * dim + k] - s->v[b * dim + k];
      }
      orc_kernel_grad(prm->kernel, dim, r_ab, prm->h, grad);
      const double P_ab =
          s->p[a] / pow2(s->rho[a]) + s->p[b] / pow2(s->rho[b]);
      double Pi_ab;
      if (prm->flavour == ORC_FLAVOUR_REF) {
        const double two = g_mutation == 13 ? 1.0 : 2.0; /* 13 planted: factor 2 lost */
        Pi_ab = two * prm->mu * dot(dim, v_ab, r_ab) /
                (s->rho[a] * s->rho[b] * norm2(dim, r_ab));
      } else {
        /* Monaghan (1992): Pi = -alpha c_ab h mu_ab / rho_ab for approaching
         * pairs, mu_ab = v_ab.r_ab / (|r_ab|^2 + 0.01 h^2); written in the
         * reference's (Pi - P) sign convention. */
        const double vr = dot(dim, v_ab, r_ab);
        Pi_ab = 0.0;
        if (vr < 0.0 || g_mutation == 3) {
          const double c_ab = 0.5 * (s->cs[a] + s->cs[b]);
          const double rho_m = 0.5 * (s->rho[a] + s->rho[b]);
          const double eta2 = 0.01 * pow2(prm->h);
          Pi_ab = prm->alpha * c_ab * prm->h * vr /
                  (rho_m * (norm2(dim, r_ab) + eta2));
        }
      }
      const int swap = g_mutation == 15; /* planted: the target's own mass */
      double fa = mass_of(prm, s, swap ? a : b) / gamma_of(prm, s, a) * (Pi_ab - P_ab);
      const double fb = mass_of(prm, s, swap ? b : a) / gamma_of(prm, s, b) * (Pi_ab - P_ab);
      if (g_mutation == 1) fa = m * (Pi_ab + P_ab); /* planted: sign slip */
      for (int k = 0; k < dim; ++k) {
        s->dv_dt[a * dim + k] += fa * grad[k];
        s->dv_dt[b * dim + k] -= fb * grad[k];
      }
    }
  }
  return 0;
}

/* compute_time_step, fluid_equations.hpp:199-222 (par::fold min). */
double orc_compute_time_step(const orc_params* prm, const orc_state* s) {
  const int dim = prm->dim;
  double dt = DBL_MAX;
  const double h = prm->h;
  for (size_t a = 0; a < s->n_fluid; ++a) {
    const double dt_ac =
        prm->cfl * h / (orc_eos_sound_speed(prm, s->rho[a]) + norm(dim, s->v + a * dim));
    if (dt_ac < dt) dt = dt_ac;
    if (prm->mu > 0.0) {
      const double dt_visc = prm->c_visc * pow2(h) * s->rho[a] / prm->mu;
      if (dt_visc < dt) dt = dt_visc;
    }
    double an = norm(dim, s->dv_dt + a * dim);
    if (an < prm->g) an = prm->g;
    const double dt_force = prm->c_force * sqrt(h / an);
    if (dt_force < dt) dt = dt_force;
  }
  return dt;
}

/* ~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~ */
/* SSPRK integrator (sph/time_integrator.hpp:140-228). */

/* Weight of the new state in the Shu-Osher combination after stage `stage`
 * (1-based; 0 = no combination), time_integrator.hpp:168-178. */
static double ssprk_weight(int order, int stage) {
  if (order == 2 && stage == 2) return 1.0 / 2.0;
  if (order == 3 && stage == 2) return 1.0 / 4.0;
  if (order == 3 && stage == 3) return 2.0 / 3.0;
  return 0.0;
}

/* lincomb_: out = (1 - w) old + w out, time_integrator.hpp:212-223 */
static void lincomb(size_t n, const double* old, double* out, double w) {
  if (g_mutation == 4) w = 1.0 - w; /* planted: swapped weights */
  for (size_t i = 0; i < n; ++i) out[i] = (1.0 - w) * old[i] + w * out[i];
}

int orc_ssprk_generic(int order, double dt, size_t n, double* u,
                      orc_rhs_fn rhs, void* ctx) {
  if (order != 2 && order != 3) return -1;
  double* old = (double*)malloc(n * sizeof(double));
  double* du = (double*)malloc(n * sizeof(double));
  if (!old || !du) {
    free(old);
    free(du);
    return -1;
  }
  memcpy(old, u, n * sizeof(double));
  for (int stage = 1; stage <= order; ++stage) {
    rhs(ctx, u, du, n);
    for (size_t i = 0; i < n; ++i) u[i] += dt * du[i];
    const double w = ssprk_weight(order, stage);
    if (w != 0.0) lincomb(n, old, u, w);
  }
  free(old);
  free(du);
  return 0;
}

typedef struct step_ws {
  uint64_t* off;
  size_t off_cap;
  ubuf idx;
  double *r_old, *v_old, *rho_old;
  size_t old_cap;
} step_ws;

static step_ws g_ws;

static int ws_reserve(step_ws* w, size_t n, int dim) {
  if (w->off_cap < n + 1) {
    free(w->off);
    w->off = (uint64_t*)malloc((n + 1) * sizeof(uint64_t));
    if (!w->off) return -1;
    w->off_cap = n + 1;
  }
  if (w->old_cap < n * 3) {
    free(w->r_old);
    free(w->v_old);
    free(w->rho_old);
    w->r_old = (double*)malloc(n * 3 * sizeof(double));
    w->v_old = (double*)malloc(n * 3 * sizeof(double));
    w->rho_old = (double*)malloc(n * sizeof(double));
    if (!w->r_old || !w->v_old || !w->rho_old) return -1;
    w->old_cap = n * 3;
  }
  (void)dim;
  return 0;
}

/* FluidEquations::prepare: index + compute_gamma (with walls) +
 * setup_boundary (fluid_equations.hpp:99-105).  The search uses
 * GridSearch{h} with radius 2h as titwcsph does (wcsph.cpp:145-154,
 * fluid_equations.hpp:112-115). */
static int prepare(const orc_params* prm, orc_state* s, step_ws* w) {
  const size_t n = s->n_fluid + s->n_fixed;
  if (build_adjacency(prm->dim, n, s->r, orc_kernel_radius(prm->kernel, prm->h),
                      prm->h, w->off, &w->idx)) {
    return -1;
  }
  if (prm->walls && orc_compute_gamma(prm, s)) return -1;
  return orc_setup_boundary(prm, s, w->off, w->idx.data);
}

/* ~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~~ */
/* Post-integration (fluid_equations.hpp:313-512).  Without walls gamma = 1,
 * the face loops are empty and the velocity correction always applies. */

/* FluidEquations constants (fluid_equations.hpp:518-524). */
static const double kShiftCfl = 0.4;  /* CFL_ */
static const double kShiftC = 0.2;    /* C_shift_ */
static const double kPhiMax = 1.0;    /* phi_max_ */
#define ORC_PHI_MIN DBL_MIN           /* phi_min_ = numeric_limits::min() */
static double k_free_surface(void) { return -log(0.05) / pow2(0.01); }

static int same_bits(double a, double b) { return memcmp(&a, &b, sizeof a) == 0; }

/* lu (core/_mat/fact.hpp:84-107): Doolittle factors of a row-major dim x dim
 * matrix, near-singular when a pivot is tiny; FactLU::inverse (:69-72) then
 * solves L U X = I, forward with the unit lower factor, backward with the
 * upper one.  Returns 0 on success. */
static int lu_inverse(int dim, const double* A, double* inv) {
  double LU[9];
  for (int i = 0; i < dim; ++i) {
    for (int j = 0; j < i; ++j) {
      double x = A[i * dim + j];
      for (int k = 0; k < j; ++k) x -= LU[i * dim + k] * LU[k * dim + j];
      LU[i * dim + j] = x / LU[j * dim + j];
    }
    for (int j = i; j < dim; ++j) {
      double x = A[i * dim + j];
      for (int k = 0; k < i; ++k) x -= LU[i * dim + k] * LU[k * dim + j];
      LU[i * dim + j] = x;
    }
    if (is_tiny(LU[i * dim + i])) return -1;
  }
  for (int c = 0; c < dim; ++c) {
    double y[3];
    for (int i = 0; i < dim; ++i) {
      double x = i == c ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) x -= LU[i * dim + k] * y[k];
      y[i] = x;
    }
    for (int i = dim - 1; i >= 0; --i) {
      double x = y[i];
      for (int k = i + 1; k < dim; ++k) x -= LU[i * dim + k] * y[k];
      y[i] = x / LU[i * dim + i];
    }
    for (int i = 0; i < dim; ++i) inv[i * dim + c] = y[i];
  }
  return 0;
}

int orc_post_integrate(const orc_params* prm, orc_state* s, const uint64_t* off,
                       const uint32_t* idx, orc_shift_out* out) {
  const int dim = prm->dim, dd = dim * dim;
  const size_t nf = s->n_fluid, n = s->n_fluid + s->n_fixed;
  const double R = orc_kernel_radius(prm->kernel, prm->h);
  double* N = (double*)calloc(n * dim, sizeof(double));
  double* L = (double*)calloc(n * dd, sizeof(double));
  double* gv = (double*)calloc(n * dd, sizeof(double));
  double* gr = (double*)calloc(n * dim, sizeof(double));
  double* dr = (double*)calloc(n * dim, sizeof(double));
  double* phi = (double*)malloc(n * sizeof(double));
  double* rho_raw = (double*)malloc(n * sizeof(double));
  if (!N || !L || !gv || !gr || !dr || !phi || !rho_raw) {
    free(N), free(L), free(gv), free(gr), free(dr), free(phi), free(rho_raw);
    return -1;
  }
  const double* r = s->r;
  const double* v = s->v;
  const double* rho = s->rho;

  /* apply_shifts, face terms (:342-349): the face's fixed particles are the
   * ones at its vertices, their fields averaged (11 planted: dropped) */
  if (prm->walls && g_mutation != 11) {
    const orc_walls* w = prm->walls;
    const face_grid* fg = faces_near(w, R);
    for (size_t a = 0; a < nf; ++a) {
      const double* x = r + 2 * a;
      const double ga = s->gamma[a];
      uint32_t jb, je;
      face_range(fg, x, &jb, &je);
      for (uint32_t j = jb; j < je; ++j) {
        const size_t f = fg->idx[j];
        if (!orc_face_intersects(wall_vert(w, f, 0), wall_vert(w, f, 1), x, R)) continue;
        double gg[2];
        orc_face_flux(prm->kernel, wall_vert(w, f, 0), wall_vert(w, f, 1), x, prm->h, gg);
        const size_t s0 = nf + w->faces[2 * f], s1 = nf + w->faces[2 * f + 1];
        double r_sa[2], v_sa[2];
        for (int k = 0; k < 2; ++k) {
          r_sa[k] = (r[2 * s0 + k] + r[2 * s1 + k]) / 2.0 - r[2 * a + k];
          v_sa[k] = (v[2 * s0 + k] + v[2 * s1 + k]) / 2.0 - v[2 * a + k];
        }
        const double rho_sa = (rho[s0] + rho[s1]) / 2.0 - rho[a];
        for (int i = 0; i < 2; ++i) {
          N[a * 2 + i] -= gg[i] / ga;
          for (int q = 0; q < 2; ++q) {
            L[a * 4 + i * 2 + q] -= r_sa[i] * gg[q] / ga;
            gv[a * 4 + i * 2 + q] -= v_sa[i] * gg[q] / ga;
          }
          gr[a * 2 + i] -= rho_sa * gg[i] / ga;
        }
      }
    }
  }
  /* apply_shifts, pair sums over unique pairs b < a (:350-365); V / gamma */
  for (size_t a = 0; a < n; ++a) {
    for (uint64_t j = off[a]; j < off[a + 1]; ++j) {
      const size_t b = idx[j];
      if (b >= a) continue;
      const double V_a = mass_of(prm, s, a) / rho[a] / gamma_of(prm, s, b);
      const double V_b = mass_of(prm, s, b) / rho[b] / gamma_of(prm, s, a);
      double x_ab[3], gW[3];
      for (int k = 0; k < dim; ++k) x_ab[k] = r[a * dim + k] - r[b * dim + k];
      orc_kernel_grad(prm->kernel, dim, x_ab, prm->h, gW);
      for (int i = 0; i < dim; ++i) {
        N[a * dim + i] += V_b * gW[i];
        N[b * dim + i] -= V_a * gW[i];
        const double r_ba = r[b * dim + i] - r[a * dim + i], r_ab = -r_ba;
        const double v_ba = v[b * dim + i] - v[a * dim + i], v_ab = v[a * dim + i] - v[b * dim + i];
        for (int q = 0; q < dim; ++q) {
          L[a * dd + i * dim + q] += V_b * (r_ba * gW[q]);
          L[b * dd + i * dim + q] -= V_a * (r_ab * gW[q]);
          gv[a * dd + i * dim + q] += V_b * (v_ba * gW[q]);
          gv[b * dd + i * dim + q] -= V_a * (v_ab * gW[q]);
        }
        gr[a * dim + i] += V_b * (rho[b] - rho[a]) * gW[i];
        gr[b * dim + i] -= V_a * (rho[a] - rho[b]) * gW[i];
      }
    }
  }
  /* renormalisation (:354-365) */
  for (size_t a = 0; a < n; ++a) {
    double* Na = N + a * dim;
    for (int i = 0; i < dim; ++i) dr[a * dim + i] = Na[i];
    double LT[9], inv[9];
    for (int i = 0; i < dim; ++i) {
      for (int q = 0; q < dim; ++q) LT[i * dim + q] = L[a * dd + q * dim + i];
    }
    if (lu_inverse(dim, LT, inv) == 0) {
      double t[9];
      for (int i = 0; i < dim; ++i) {
        double x = 0.0;
        for (int q = 0; q < dim; ++q) x += inv[i * dim + q] * Na[q];
        t[i] = x;
      }
      for (int i = 0; i < dim; ++i) Na[i] = t[i];
      double* G = gv + a * dd;
      if (g_mutation != 5) {
        for (int i = 0; i < dim; ++i) {
          for (int q = 0; q < dim; ++q) {
            double x = 0.0;
            for (int k = 0; k < dim; ++k) x += G[i * dim + k] * inv[q * dim + k];
            t[i * dim + q] = x;
          }
        }
        for (int k = 0; k < dd; ++k) G[k] = t[k];
      }
      double* g = gr + a * dim;
      for (int i = 0; i < dim; ++i) {
        double x = 0.0;
        for (int q = 0; q < dim; ++q) x += inv[i * dim + q] * g[q];
        t[i] = x;
      }
      for (int i = 0; i < dim; ++i) g[i] = t[i];
    }
    double nn[3];
    normalize(dim, Na, nn);
    for (int i = 0; i < dim; ++i) Na[i] = nn[i];
  }
  /* free-surface indicator (:375-419) */
  for (size_t a = 0; a < n; ++a) phi[a] = a < nf ? ORC_PHI_MIN : kPhiMax;
  const double cos_fov = cos(3.14159265358979323846 / 4.0); /* std::numbers::pi */
  for (size_t a = 0; a < n; ++a) {
    for (uint64_t j = off[a]; j < off[a + 1]; ++j) {
      const size_t b = idx[j];
      if (b >= a) continue;
      double x_ab[3];
      for (int k = 0; k < dim; ++k) x_ab[k] = r[a * dim + k] - r[b * dim + k];
      const double r2_ab = norm2(dim, x_ab);
      const double thr = (R + R) / 2.0;
      if (r2_ab > pow2(thr)) continue;
      const double fov = pow2(cos_fov) * r2_ab;
      if (same_bits(phi[a], ORC_PHI_MIN)) {
        const double n_a = dot(dim, N + a * dim, x_ab);
        const int in_cone = g_mutation == 6 ? n_a < 0.0 : n_a > 0.0;
        if (in_cone && pow2(n_a) >= fov) phi[a] = kPhiMax;
      }
      if (same_bits(phi[b], ORC_PHI_MIN)) {
        const double n_b = dot(dim, N + b * dim, x_ab);
        const int in_cone = g_mutation == 6 ? n_b > 0.0 : n_b < 0.0;
        if (in_cone && pow2(n_b) >= fov) phi[b] = kPhiMax;
      }
    }
  }
  /* splashes (:421-426) */
  const uint64_t cutoff = dim == 2 ? 8 : 26;
  for (size_t a = 0; a < nf; ++a) {
    if (off[a + 1] - off[a] <= cutoff) phi[a] = ORC_PHI_MIN;
  }
  /* near-surface particles (:439-452): the nearest surface neighbour, first
   * in list order among equals */
  for (size_t a = 0; a < nf; ++a) {
    if (!same_bits(phi[a], kPhiMax)) continue;
    size_t best = n;
    double best_d = 0.0;
    for (uint64_t j = off[a]; j < off[a + 1]; ++j) {
      const size_t b = idx[j];
      if (!same_bits(phi[b], ORC_PHI_MIN)) continue;
      double x_ab[3];
      for (int k = 0; k < dim; ++k) x_ab[k] = r[a * dim + k] - r[b * dim + k];
      const double d2 = norm2(dim, x_ab);
      if (best == n || d2 < best_d) {
        best = b;
        best_d = d2;
      }
    }
    if (best < n) {
      double x_ab[3];
      for (int k = 0; k < dim; ++k) x_ab[k] = r[a * dim + k] - r[best * dim + k];
      phi[a] *= fabs(dot(dim, N + best * dim, x_ab)) / R;
    }
  }
  /* shifts and field corrections (:455-470) */
  const double coef = -kShiftCfl * kShiftC * pow2(prm->h);
  for (size_t a = 0; a < nf; ++a) {
    double* da = dr + a * dim;
    if (!same_bits(phi[a], kPhiMax) && g_mutation != 7) {
      for (int i = 0; i < dim; ++i) da[i] = 0.0;
      continue;
    }
    for (int i = 0; i < dim; ++i) da[i] *= coef;
    double dv[3];
    for (int i = 0; i < dim; ++i) {
      double x = 0.0;
      for (int q = 0; q < dim; ++q) x += gv[a * dd + i * dim + q] * da[q];
      dv[i] = x;
    }
    /* the velocity correction only where gamma ~ 1 (:465-468) */
    const int correct_v = is_tiny(gamma_of(prm, s, a) - 1.0);
    for (int i = 0; i < dim; ++i) {
      s->r[a * dim + i] += da[i];
      if (correct_v) s->v[a * dim + i] += dv[i];
    }
    s->rho[a] += dot(dim, gr + a * dim, da);
  }
  for (size_t a = nf; a < n; ++a) {
    for (int i = 0; i < dim; ++i) dr[a * dim + i] = 0.0;
  }
  /* apply_free_surface_correction (:483-512), on the same lists */
  memcpy(rho_raw, s->rho, n * sizeof(double));
  const double K = k_free_surface();
  for (size_t a = 0; a < nf; ++a) {
    if (same_bits(phi[a], kPhiMax)) continue;
    double alpha = 0.0, rho_t = 0.0;
    for (uint64_t j = off[a]; j < off[a + 1]; ++j) {
      const size_t b = idx[j];
      double x_ab[3];
      for (int k = 0; k < dim; ++k) x_ab[k] = s->r[a * dim + k] - s->r[b * dim + k];
      const double W_ab = orc_kernel_value(prm->kernel, dim, x_ab, prm->h);
      alpha += mass_of(prm, s, b) / rho_raw[b] * W_ab;
      rho_t += mass_of(prm, s, b) * W_ab;
    }
    const double ga = gamma_of(prm, s, a);
    const double ratio = fmin(1.0, alpha / ga);
    if (ratio > 0.99) continue;
    const double e = g_mutation == 8 ? K * pow2(ratio - 1.0) : -K * pow2(ratio - 1.0);
    const double beta = exp(e);
    const double corr = beta * ga + (1.0 - beta) * alpha;
    if (!is_tiny(corr)) s->rho[a] = rho_t / corr;
  }
  if (out) {
    if (out->N) memcpy(out->N, N, n * dim * sizeof(double));
    if (out->grad_v) memcpy(out->grad_v, gv, n * dd * sizeof(double));
    if (out->grad_rho) memcpy(out->grad_rho, gr, n * dim * sizeof(double));
    if (out->phi) memcpy(out->phi, phi, n * sizeof(double));
    if (out->dr) memcpy(out->dr, dr, n * dim * sizeof(double));
  }
  free(N), free(L), free(gv), free(gr), free(dr), free(phi), free(rho_raw);
  return 0;
}

/* SSPRKIntegrator::substep_, time_integrator.hpp:188-210 */
static int substep(const orc_params* prm, orc_state* s, step_ws* w,
                   double* dt) {
  const int dim = prm->dim;
  if (prepare(prm, s, w)) return -1;
  if (*dt <= 0.0) *dt = orc_compute_time_step(prm, s);
  orc_compute_continuity(prm, s, w->off, w->idx.data);
  orc_compute_momentum(prm, s, w->off, w->idx.data);
  const double h = *dt;
  for (size_t a = 0; a < s->n_fluid; ++a) {
    for (int k = 0; k < dim; ++k) s->r[a * dim + k] += h * s->v[a * dim + k];
    for (int k = 0; k < dim; ++k) s->v[a * dim + k] += h * s->dv_dt[a * dim + k];
    s->rho[a] += h * s->drho_dt[a];
  }
  return 0;
}

/* SSPRKIntegrator::step, time_integrator.hpp:159-184; post_integrate
 * (fluid_equations.hpp:313-321: prepare, apply_shifts, free-surface
 * correction) when prm->post_integrate is set. */
int orc_ssprk_step(const orc_params* prm, orc_state* s, int order,
                   double* dt_out) {
  if (order != 2 && order != 3) return -1;
  if (prm->walls && (prm->dim != 2 || prm->flavour != ORC_FLAVOUR_REF)) return -1;
  const int dim = prm->dim;
  const size_t n = s->n_fluid + s->n_fixed, nf = s->n_fluid;
  step_ws* w = &g_ws;
  if (ws_reserve(w, n, dim)) return -1;
  memcpy(w->r_old, s->r, nf * dim * sizeof(double));
  memcpy(w->v_old, s->v, nf * dim * sizeof(double));
  memcpy(w->rho_old, s->rho, nf * sizeof(double));
  double dt = 0.0;
  for (int stage = 1; stage <= order; ++stage) {
    if (substep(prm, s, w, &dt)) return -1;
    const double wgt = ssprk_weight(order, stage);
    if (wgt != 0.0) {
      lincomb(nf * dim, w->r_old, s->r, wgt);
      lincomb(nf * dim, w->v_old, s->v, wgt);
      lincomb(nf, w->rho_old, s->rho, wgt);
    }
  }
  if (prm->post_integrate) {
    if (prepare(prm, s, w)) return -1;
    if (orc_post_integrate(prm, s, w->off, w->idx.data, NULL)) return -1;
  }
  if (dt_out) *dt_out = dt;
  return 0;
}
