/*
 * FP64 CPU oracle for the WCSPH step -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C and double precision, the reference
 * algorithm of TitSolver's WCSPH step (neighbour search, smoothing kernels,
 * Tait equation of state, continuity/momentum right-hand sides, time step
 * selection, boundary extrapolation and the SSPRK integrator).  It is the
 * checker that the HIP path is compared against.  Only the test suite, the
 * smoke check and the CPU-baseline leg of the benchmark may load it; the
 * product library never links or calls it.
 *
 * Each function cites the reference file:line it follows (paths relative to
 * the TitSolver source tree).  Vectors use the reference's interleaved layout:
 * component k of particle i lives at x[i * dim + k].  Particles are ordered
 * fluid first, then fixed (source/tit/sph/particle_array.hpp:204-239).
 *
 * Pinning (what fixes each function independently of this restatement) is
 * listed in the module header of tit_oracle.c.
 */
#ifndef TIT_ORACLE_H
#define TIT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Smoothing kernels (source/tit/sph/kernel.hpp:447-454). */
enum { ORC_KERNEL_WENDLAND_C2 = 0, ORC_KERNEL_WENDLAND_C4 = 1 };

/* Term sets.  REF: Ferrari density diffusion + physical viscosity mu
 * (source/tit/sph/fluid_equations.hpp:249-304).  CFG: delta-SPH density
 * diffusion (Molteni & Colagrossi 2009) + Monaghan (1992) artificial
 * viscosity alpha. */
enum { ORC_FLAVOUR_REF = 0, ORC_FLAVOUR_CFG = 1 };

/* Wall boundary of the semi-analytic formulation, 2D (fluid_equations.hpp:
 * 58-69, wcsph.cpp:55-86, 110-113): the tessellated wall polyline, whose
 * vertex k carries fixed particle n_fluid + k, each face's normal
 * normalize(cross(b - a)) pointing into the fluid; and a closed polygon whose
 * winding number decides containment. */
typedef struct orc_walls {
  size_t n_verts;
  const double* verts;    /* n_verts * 2 */
  size_t n_faces;
  const uint32_t* faces;  /* n_faces * 2 vertex indices (a, b) */
  size_t n_cverts;
  const double* cverts;   /* containment polygon */
  size_t n_cfaces;
  const uint32_t* cfaces;
} orc_walls;

typedef struct orc_params {
  int dim;      /* 2 or 3 */
  int kernel;   /* ORC_KERNEL_* */
  int flavour;  /* ORC_FLAVOUR_* */
  int post_integrate; /* 1: orc_ssprk_step ends with orc_post_integrate */
  double h;       /* kernel width (uniform field h) */
  double m;       /* particle mass where the state sets none */
  double rho0;    /* reference density */
  double cs0;     /* reference sound speed */
  double xi;      /* Tait index */
  double g;       /* gravity magnitude, acting along -axis 1 */
  double mu;      /* physical viscosity (REF) */
  double alpha;   /* artificial viscosity (CFG) */
  double delta;   /* delta-SPH coefficient (CFG) */
  double cfl;     /* acoustic CFL number */
  double c_force; /* force time-step coefficient */
  double c_visc;  /* viscous time-step coefficient */
  const orc_walls* walls; /* null: wall-free form (fixed particles carry normals) */
} orc_params;

typedef struct orc_state {
  size_t n_fluid;
  size_t n_fixed;
  double* r;        /* n * dim */
  double* v;        /* n * dim */
  double* rho;      /* n */
  double* drho_dt;  /* n */
  double* dv_dt;    /* n * dim */
  double* p;        /* n */
  double* cs;       /* n */
  const double* normal; /* n_fixed * dim, wall normals of fixed particles */
  double* gamma;        /* n (walls only) */
  double* grad_gamma;   /* n * dim (walls only) */
  double* mass;         /* n: per-particle masses, ParticleArray's m; an entry
                           <= 0 (or a null array) means the uniform m */
} orc_state;

/* ---- smoothing kernel ---------------------------------------------------- */
double orc_kernel_weight(int kernel, int dim);
double orc_kernel_unit_value(int kernel, double q);
double orc_kernel_unit_deriv(int kernel, double q);
double orc_kernel_radius(int kernel, double h);
double orc_kernel_value(int kernel, int dim, const double* x, double h);
void orc_kernel_grad(int kernel, int dim, const double* x, double h,
                     double* out);

/* ---- Tait equation of state ---------------------------------------------- */
double orc_eos_pressure(const orc_params* prm, double rho);
double orc_eos_density_from_pressure(const orc_params* prm, double p);
double orc_eos_sound_speed(const orc_params* prm, double rho);
double orc_eos_potential(const orc_params* prm, double rho);
double orc_eos_density_from_potential(const orc_params* prm, double H);
/* p[i] = pressure(rho[i]) for i < n. */
void orc_eos_pressure_array(const orc_params* prm, size_t n, const double* rho,
                            double* p);

/* ---- neighbour search ---------------------------------------------------- */
/* Reference GridIndex + ParticleMesh::search_: per point, the ascending list of
 * point indices b with norm2(p_b - p_a) <= radius^2 (self included).
 * Writes CSR offsets off[n+1] and indices idx[cap].  Returns the total number
 * of entries, or -(needed) when cap is too small, or -1 on bad arguments. */
long long orc_grid_search(int dim, size_t n, const double* pts, double radius,
                          double size_hint, uint64_t* off, uint32_t* idx,
                          size_t cap);
/* O(N^2) search; strict != 0 uses '<' as the reference's naive test does. */
long long orc_naive_search(int dim, size_t n, const double* pts,
                           double radius, int strict, uint64_t* off,
                           uint32_t* idx, size_t cap);

/* ---- fluid equations ----------------------------------------------------- */
int orc_setup_boundary(const orc_params* prm, orc_state* s,
                       const uint64_t* off, const uint32_t* idx);
int orc_compute_continuity(const orc_params* prm, orc_state* s,
                           const uint64_t* off, const uint32_t* idx);
int orc_compute_momentum(const orc_params* prm, orc_state* s,
                         const uint64_t* off, const uint32_t* idx);
double orc_compute_time_step(const orc_params* prm, const orc_state* s);

/* ---- semi-analytic wall (2D) ------------------------------------------- */
/* KernelGen::unit_flux / unit_antigrad_flux over a segment in unit
 * coordinates (sph/kernel.hpp:252-312). */
double orc_kernel_unit_flux(int kernel, double eta, double z0, double z1);
double orc_kernel_unit_antigrad_flux(int kernel, double eta, double z0, double z1);
/* Kernel::flux and antigrad_flux of segment (a, b) at x (kernel.hpp:193-220). */
void orc_face_flux(int kernel, const double* a, const double* b, const double* x,
                   double h, double* out);
double orc_face_antigrad_flux(int kernel, const double* a, const double* b,
                              const double* x, double h);
/* Segment::intersects(BSphere{c, R}) (geom/segment.hpp:98-112). */
int orc_face_intersects(const double* a, const double* b, const double* c, double R);
/* Exact generalised winding number of the containment polygon at x. */
double orc_winding(const orc_walls* w, const double* x);
/* compute_gamma (fluid_equations.hpp:168-193) for every particle. */
int orc_compute_gamma(const orc_params* prm, orc_state* s);
/* FluidEquations::initialize (fluid_equations.hpp:75-89): masses left at
 * zero become m; with walls gamma, and the fixed particles' masses times
 * gamma. */
int orc_initialize(const orc_params* prm, orc_state* s);
/* ---- post-integration (particle shifting, free-surface correction) ------- */
/* Optional outputs of orc_post_integrate (any pointer may be null). */
typedef struct orc_shift_out {
  double* N;        /* n * dim: normalised, renormalised surface normal N */
  double* grad_v;   /* n * dim * dim: renormalised grad v, [i][j] = dv_i/dx_j */
  double* grad_rho; /* n * dim: renormalised grad rho */
  double* phi;      /* n: free-surface indicator (DBL_MIN on the surface, 1 far) */
  double* dr;       /* n * dim: the shift applied to each particle */
} orc_shift_out;

/* FluidEquations::apply_shifts + apply_free_surface_correction
 * (fluid_equations.hpp:327-512), on the adjacency of the current positions
 * (as post_integrate's prepare builds it); with walls, the face terms and
 * s->gamma (orc_compute_gamma first).  Updates r, v, rho of the fluid particles.  Returns 0 on success. */
int orc_post_integrate(const orc_params* prm, orc_state* s, const uint64_t* off,
                       const uint32_t* idx, orc_shift_out* out);

/* ---- integrator ---------------------------------------------------------- */
/* One SSPRK step (order 2 or 3) with a grid search before every stage, the
 * reference SSPRKIntegrator::step; post_integrate (re-index, boundary,
 * orc_post_integrate) runs at its end when prm->post_integrate is set.
 * Returns 0 on success.  *dt_out receives the step size. */
int orc_ssprk_step(const orc_params* prm, orc_state* s, int order,
                   double* dt_out);

/* Generic SSPRK on a flat state u[n] with right-hand side callback; used to
 * pin the Shu-Osher combination independently of the SPH terms. */
typedef void (*orc_rhs_fn)(void* ctx, const double* u, double* du, size_t n);
int orc_ssprk_generic(int order, double dt, size_t n, double* u,
                      orc_rhs_fn rhs, void* ctx);

/* Test-only: plant a known mistake (1: momentum pressure sign on the a side,
 * 2: anti-diffusive delta-SPH sign, 3: viscosity on receding pairs,
 * 4: swapped SSPRK combination weights, 5: grad v not renormalised,
 * 6: free-surface visibility cone reversed, 7: shifts applied to every
 * fluid particle, 8: free-surface blend exponent with the wrong sign,
 * 9: wall face normals reversed in the kernel flux, 10: wall face terms of
 * the momentum equation dropped, 11: wall face terms of apply_shifts
 * dropped, 12: anti-diffusive Ferrari term, 13: physical viscosity without
 * its factor 2, 14: continuity weighted by the target's own mass instead of
 * the neighbour's, 15: momentum weighted by the target's own mass instead of
 * the neighbour's); 0 restores the faithful oracle. */
void orc_set_mutation(int m);

#ifdef __cplusplus
}
#endif

#endif /* TIT_ORACLE_H */
