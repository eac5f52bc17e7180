/*
 * tit_hip.h -- C ABI of the MI355X (gfx950) WCSPH step.
 *
 * The functions below are the drop-in boundary for TitSolver's WCSPH step.
 * Plain C types only: no exceptions, no C++ or torch types cross it.  Every
 * entry point that can fail returns a tit_status; tit_hip_last_error() gives
 * the message of the last failure on a context.
 *
 * Threading: one context owns one HIP stream and its device buffers; calls on
 * one context must not overlap.  Several contexts may coexist.
 * Ownership: the caller owns every host array it passes; the library copies
 * what it needs and keeps no host pointer after a call returns.
 * Layout: host arrays use the reference layout -- component k of particle i at
 * [i * dim + k], fluid particles first, then fixed ones, tank after tank
 * (source/tit/sph/particle_array.hpp:204-239).
 */
#ifndef TIT_HIP_H
#define TIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tit_hip_ctx tit_hip_ctx;

typedef enum tit_status {
  TIT_OK = 0,
  TIT_E_ARG = 1,     /* invalid argument or unsupported configuration */
  TIT_E_HIP = 2,     /* HIP runtime error (no device, launch failure, ...) */
  TIT_E_OOM = 3,     /* device allocation failed */
  TIT_E_NUMERIC = 4, /* non-finite state or neighbour-skin overrun */
  TIT_E_STATE = 5    /* call out of order (e.g. step before upload) */
} tit_status;

/* Smoothing kernels: QuarticWendlandKernel (C2) and SixthOrderWendlandKernel
 * (C4), source/tit/sph/kernel.hpp:447-451. */
enum { TIT_KERNEL_WENDLAND_C2 = 0, TIT_KERNEL_WENDLAND_C4 = 1 };

/* Term sets.  TIT_FLAVOUR_REF: the reference's Ferrari density diffusion and
 * physical viscosity mu (source/tit/sph/fluid_equations.hpp:249-304).
 * TIT_FLAVOUR_CFG: delta-SPH density diffusion (Molteni & Colagrossi 2009)
 * and Monaghan (1992) artificial viscosity alpha. */
enum { TIT_FLAVOUR_REF = 0, TIT_FLAVOUR_CFG = 1 };

/* Constructor arguments of FluidEquations + TaitEquationOfState + kernel
 * (source/tit/sph/fluid_equations.hpp:61-69, equation_of_state.hpp:27-34,
 * titwcsph/wcsph.cpp:79-90) plus the time-step constants that the reference
 * keeps private (fluid_equations.hpp:518-520). */
typedef struct tit_hip_params {
  int dim;          /* 2 or 3 */
  int kernel;       /* TIT_KERNEL_* */
  int flavour;      /* TIT_FLAVOUR_* */
  int ssprk_order;  /* 2 or 3 (SSPRKOrder, time_integrator.hpp:134-137) */
  double h;         /* kernel width; support radius is 2h */
  double m;         /* particle mass (uniform) */
  double rho0, cs0, xi; /* Tait EOS */
  double g;         /* gravity magnitude along -axis 1 */
  double mu;        /* physical viscosity (REF flavour) */
  double alpha;     /* artificial viscosity (CFG flavour) */
  double delta;     /* delta-SPH coefficient (CFG flavour) */
  double cfl, c_force, c_visc; /* time-step coefficients */
  double skin;      /* neighbour-cell skin in units of the support radius;
                       0 selects the default 0.1 */
  int post_integrate; /* 1: every step ends with FluidEquations::post_integrate
                         (fluid_equations.hpp:313-321): particle shifting and
                         the free-surface density correction */
  int reserved;       /* 0 */
} tit_hip_params;

/* One tank: the box its cell grid covers.  Particles that leave the box are
 * clamped into the border cells (neighbour sets stay exact). */
typedef struct tit_hip_tank {
  double lo[3];
  double hi[3];
} tit_hip_tank;

/* Host view of one tank's ParticleArray fields (reference layout). */
typedef struct tit_hip_host_view {
  const double* r;      /* (n_fluid + n_fixed) * dim */
  const double* v;      /* (n_fluid + n_fixed) * dim, may be NULL (zero) */
  const double* rho;    /* n_fluid + n_fixed */
  const double* dv_dt;  /* (n_fluid + n_fixed) * dim, may be NULL (zero):
                           the previous momentum RHS, read by the first
                           time-step selection only */
  const double* normal; /* n_fixed * dim wall normals of the fixed particles
                           (ignored, may be NULL, with tit_hip_set_walls) */
  size_t n_fluid;
  size_t n_fixed;
  const double* m;      /* n_fluid + n_fixed particle masses, ParticleArray's m
                           (fluid_equations.hpp:41-48), or NULL: params.m for
                           every particle.  Positive and finite.  With
                           tit_hip_set_walls the fixed particles' masses are
                           multiplied by gamma, as FluidEquations::initialize
                           does (fluid_equations.hpp:87-88).  The learned step
                           needs uniform masses (NULL). */
} tit_hip_host_view;

/* Wall boundary of the semi-analytic formulation, 2D: the tessellated wall
 * surface and a closed polygon for containment, as FluidEquations receives
 * them (fluid_equations.hpp:58-69; built in titwcsph/wcsph.cpp:55-86).
 * Face k joins vertices faces[2k], faces[2k+1]; its normal
 * normalize(cross(b - a)), cross(v) = (v1, -v0), points into the fluid.
 * Fixed particle k of the uploaded tank must sit at vertex k
 * (wcsph.cpp:110-113). */
typedef struct tit_hip_walls {
  const double* verts;    /* n_verts * 2 */
  uint32_t n_verts;
  uint32_t n_faces;
  const uint32_t* faces;  /* n_faces * 2 */
  const double* cverts;   /* n_cverts * 2, containment polygon */
  uint32_t n_cverts;
  uint32_t n_cfaces;
  const uint32_t* cfaces; /* n_cfaces * 2 */
} tit_hip_walls;

/* Creates a context on HIP device `device`. */
tit_status tit_hip_create(int device, const tit_hip_params* params,
                          tit_hip_ctx** out);

/* Switches the context to the semi-analytic wall boundary (2D, REF terms,
 * one tank; post_integrate allowed): before tit_hip_upload.  Every prepare then
 * runs compute_gamma (fluid_equations.hpp:168-193), setup_boundary takes
 * normalize(grad gamma) as the wall normal, the fixed particles carry the
 * masses m gamma set by FluidEquations::initialize (:75-89), and the
 * continuity and momentum equations gain the face terms and the 1/gamma
 * factors (:232-305).  Replaces: constructing FluidEquations with a domain
 * and containment function (fluid_equations.hpp:58-69). */
tit_status tit_hip_set_walls(tit_hip_ctx* ctx, const tit_hip_walls* walls);

/* Copies one tank's particles to the device (replaces any previous state).
 * Replaces: the ParticleArray field vectors that SSPRKIntegrator::step
 * reads (source/tit/sph/particle_array.hpp:242-276). */
tit_status tit_hip_upload(tit_hip_ctx* ctx, const tit_hip_host_view* view,
                          const tit_hip_tank* tank);

/* Copies n_tanks independent tanks; all later calls act on every tank, each
 * with its own time step.  No reference counterpart (the reference runs one
 * tank per process, titwcsph/wcsph.cpp:35-196). */
tit_status tit_hip_upload_batch(tit_hip_ctx* ctx,
                                const tit_hip_host_view* views,
                                const tit_hip_tank* tanks, size_t n_tanks);

/* Advances every tank by n_steps SSPRK steps and returns the last step size
 * of tank 0.  Replaces: SSPRKIntegrator::step (time_integrator.hpp:159-184),
 * ending with FluidEquations::post_integrate when params.post_integrate is 1
 * (as the reference's step does, :181); called as in titwcsph/wcsph.cpp:181. */
tit_status tit_hip_step(tit_hip_ctx* ctx, int n_steps, double* dt_last);

/* Copies the state back in reference order.  Any pointer may be NULL.
 * p is the Tait pressure of rho (FluidEquations fills p in compute_momentum,
 * fluid_equations.hpp:272-273). */
tit_status tit_hip_download(tit_hip_ctx* ctx, double* r, double* v,
                            double* rho, double* p);

/* Per-tank simulated time and the step size the next step will use. */
tit_status tit_hip_tank_clock(tit_hip_ctx* ctx, double* time, double* dt_next);

/* FluidEquations::prepare without the wall terms: neighbour index rebuild and
 * setup_boundary (fluid_equations.hpp:99-105, :122-164). */
tit_status tit_hip_prepare(tit_hip_ctx* ctx);

/* FluidEquations::post_integrate (fluid_equations.hpp:313-321) on the
 * current state: prepare, apply_shifts (:327-471) and
 * apply_free_surface_correction (:480-512), with the wall's face terms and
 * gamma when tit_hip_set_walls was called.  Requires a context created
 * with params.post_integrate = 1.  phi (may be NULL) receives the
 * free-surface indicator of every particle in the reference order (DBL_MIN
 * on the surface, 1 far from it).  The next step size is taken again from
 * the new state and the last stage's acceleration (zero right after an
 * upload). */
tit_status tit_hip_post_integrate(tit_hip_ctx* ctx, double* phi);

/* FluidEquations::compute_time_step (fluid_equations.hpp:199-222) on the
 * current state and the last momentum RHS; one value per tank. */
tit_status tit_hip_compute_time_step(tit_hip_ctx* ctx, double* dt);

/* FluidEquations::compute_continuity + compute_momentum
 * (fluid_equations.hpp:232-305) on the current state; results in reference
 * order (fixed particles get zero).  Does not advance the state. */
tit_status tit_hip_compute_rhs(tit_hip_ctx* ctx, double* drho_dt,
                               double* dv_dt);

/* ParticleMesh adjacency (particle_mesh.hpp:123-147): per particle, the
 * ascending list of neighbour indices (reference order, self included) with
 * norm2(r_b - r_a) <= (2h)^2.  off has n+1 entries; idx has cap entries.
 * *total receives the number of entries; TIT_E_ARG if cap is too small. */
tit_status tit_hip_neighbors(tit_hip_ctx* ctx, uint64_t* off, uint32_t* idx,
                             size_t cap, size_t* total);

/* tit_hip_neighbors without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-particle counts and the device copy of the lists live in a static
 * device array of the library, one per GPU, held until the process ends
 * (72 MiB), so at most 1,048,576 particles and 16,777,216 list entries
 * (TIT_E_ARG beyond; *total is set when the entries overflow).  Same
 * results as tit_hip_neighbors.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (several contexts on several host
 * threads) race. */
tit_status tit_hip_neighbors_noalloc(tit_hip_ctx* ctx, uint64_t* off, uint32_t* idx,
                                     size_t cap, size_t* total);

/* Number of neighbours (self excluded) whose pair terms the fused stage
 * kernel summed for each particle in the last right-hand side of the current
 * state, reference order; 0 for fixed particles (test hook: must equal the
 * adjacency sizes of tit_hip_neighbors minus one for fluid particles). */
tit_status tit_hip_pair_counts(tit_hip_ctx* ctx, uint32_t* count);

/* Kernel timing with HIP events on the context's stream.  While enabled,
 * steps run eagerly (no graph replay) with an event pair around every launch.
 * Mirrors TIT_PROFILE_SECTION (source/tit/core/profiler.hpp:42-44). */
typedef struct tit_hip_kernel_stat {
  char name[48];
  uint64_t launches;
  double total_ms;
} tit_hip_kernel_stat;
tit_status tit_hip_set_profiling(tit_hip_ctx* ctx, int enable);
tit_status tit_hip_kernel_stats(tit_hip_ctx* ctx, tit_hip_kernel_stat* out,
                                int cap, int* n_out);

/* Schedule of the fused stage kernel.  Two families:
 *  - column chunks: each workgroup takes a chunk of at most chunk_targets
 *    consecutive particles of one cell column (one per lane); in 2D up to
 *    lds_capacity candidate positions of a chunk are staged in LDS;
 *  - 3D block chunks (lds_capacity > 0 in 3D): a workgroup of 256 lanes takes
 *    the particles of a 4 x 4 block of cell columns over a few rows, with up
 *    to lds_capacity candidate positions of the block's window in LDS.
 * lds_capacity 0 reads positions from global memory (column chunks).  Results
 * are bitwise identical on every schedule.  occupancy is the mean number of
 * particles per occupied cell at upload.  cached (tit_hip_autotune only): 1
 * when the schedule came from the process's tuning cache.  No reference
 * counterpart (the reference partitions pairs for TBB threads instead,
 * particle_mesh.hpp:164-241). */
typedef struct tit_hip_schedule {
  int chunk_targets;
  int lds_capacity;
  double occupancy;
  int cached;
  int reserved;
} tit_hip_schedule;
tit_status tit_hip_get_schedule(tit_hip_ctx* ctx, tit_hip_schedule* out);
/* chunk_targets = 0 keeps the current chunk length. */
tit_status tit_hip_set_schedule(tit_hip_ctx* ctx, const tit_hip_schedule* s);

/* Autotuner of the stage schedule: times every candidate (chunk_targets in
 * {32, 48, 64, 96, 128} x LDS capacity {max, max/2, 0}) over
 * steps_per_candidate real steps of the current state (all schedules give the
 * same bits, so the steps are ordinary progress), keeps the fastest and
 * returns it with its time per step.  The result is kept in a per-process
 * cache keyed by (dimension, kernel, term set, walls, one or many tanks, the
 * particle count's power of two): a later call on a context of the same key
 * applies the cached schedule without timing (best->cached = 1, no steps
 * taken).  No reference counterpart. */
tit_status tit_hip_autotune(tit_hip_ctx* ctx, int steps_per_candidate,
                            tit_hip_schedule* best, double* best_ms_per_step);
/* Empties the autotuner's cache. */
void tit_hip_autotune_clear_cache(void);

/* Number of particles (all tanks), of cell-index rebuilds run so far (the
 * index is rebuilt only when a tank's particles could otherwise leave the
 * skin, or when forced by a hook, a schedule change or an upload) and of
 * plain steps taken since upload. */
tit_status tit_hip_counts(tit_hip_ctx* ctx, size_t* n_particles,
                          uint64_t* n_rebuilds, uint64_t* n_steps);

const char* tit_hip_last_error(const tit_hip_ctx* ctx);
void tit_hip_destroy(tit_hip_ctx* ctx);

/* ---- learned step ----------------------------------------------------------
 * No reference counterpart: a message-passing network (edge MLP 80 -> 128
 * -> 128, node MLP 160 -> 128 -> 128 -> 16, hard-swish activation
 * x min(max(x / 6 + 1/2, 0), 1), two rounds over the 8
 * nearest neighbours within h) predicts each fluid particle's acceleration
 * and density rate; one learned step advances 4 dt:
 *   v' = v + 4dt a,  r' = r + 4dt (v + v') / 2,  rho' = rho + 4dt drho/dt.
 * It stands in for the pair sums of FluidEquations (fluid_equations.hpp:
 * 232-305) inside SSPRKIntegrator::step (time_integrator.hpp:159-184).
 * Parameters are FP32 (kept in BF16 for the MFMA GEMMs). */

/* Creates the network with seeded random weights (replaces any earlier
 * one; recorded samples are dropped). */
tit_status tit_hip_learned_init(tit_hip_ctx* ctx, uint64_t seed);
/* Parameter vector (row-major weight blocks, per round W1 b1 W2 b2 W3 b3 W4
 * b4 W5 b5).  With w = NULL only *n is written. */
tit_status tit_hip_learned_params(tit_hip_ctx* ctx, float* w, size_t cap,
                                  size_t* n);
tit_status tit_hip_learned_set_params(tit_hip_ctx* ctx, const float* w,
                                      size_t n);
/* Scales: [v_s, rho_s, a_x, a_y, a_z, drho/dt].  Node features use v / v_s
 * and (rho / rho0 - 1) / rho_s; outputs are multiplied by the last four. The
 * output scales are set from the first recorded sample unless set here. */
tit_status tit_hip_learned_scales(tit_hip_ctx* ctx, double* scales6);
tit_status tit_hip_learned_set_scales(tit_hip_ctx* ctx,
                                      const double* scales6);

/* Training sample from the current state: its graph and features, then
 * `plain_steps` plain steps (the context's state advances); the targets are
 * (v_after - v) / T and (rho_after - rho) / T with T the time advanced. */
tit_status tit_hip_learned_record(tit_hip_ctx* ctx, int plain_steps,
                                  int* sample_index);
tit_status tit_hip_learned_clear_samples(tit_hip_ctx* ctx);
/* Sample export (test hook): n nodes; nbr n x 8 (sample-local indices, self
 * for a missing edge), node_feat n x 16, edge_feat n x 8 x 16 (the BF16
 * values as float), target n x 4, fluid n.  NULL pointers are skipped. */
tit_status tit_hip_learned_sample(tit_hip_ctx* ctx, int index, size_t* n,
                                  uint32_t* nbr, float* node_feat,
                                  float* edge_feat, float* target,
                                  uint8_t* fluid);
/* Mean-squared loss over fluid particles and used output channels of one
 * sample, and its gradient with respect to the parameters (test hook). */
tit_status tit_hip_learned_loss_grad(tit_hip_ctx* ctx, int index,
                                     double* loss, float* grad);
/* Adam (beta 0.9 / 0.999) over the recorded samples in turn; losses[it]
 * (may be NULL) receives each iteration's loss before its update. */
tit_status tit_hip_learned_train(tit_hip_ctx* ctx, int iters, double lr,
                                 double* losses);
/* Graph and features of the current state, in reference order (test hook):
 * nbr_ref n x 8 particle indices, node_feat n x 16, edge_feat n x 8 x 16. */
tit_status tit_hip_learned_graph(tit_hip_ctx* ctx, uint32_t* nbr_ref,
                                 float* node_feat, float* edge_feat);
/* Network output (n x 16, reference order) on the current state. */
tit_status tit_hip_learned_forward(tit_hip_ctx* ctx, float* out);
/* Inference path: 1 (default) runs each round as one fused kernel with the
 * weights resident in LDS; 0 runs it as separate GEMM launches (the training
 * path's forward).  Both give the same network to BF16 rounding. */
tit_status tit_hip_learned_set_fused(tit_hip_ctx* ctx, int enable);

typedef struct tit_hip_learned_report {
  int learned_steps;    /* learned steps taken, judged ones included */
  int judged;           /* comparisons with the plain step */
  int rejected;         /* comparisons above the tolerance */
  int plain_steps;      /* plain SSPRK steps (judge references, fallbacks) */
  double max_rho_err;   /* largest max_i |rho_L - rho_P| / rho_P seen */
  double last_rho_err;
  double max_front_err; /* largest |front_L - front_P| / |front_P| seen */
  double max_dv_err;    /* largest |v_L - v_P| / |v_P - v_S| (L2 over fluid,
                           S the judged step's start): error of the step's
                           velocity change */
  double last_dv_err;
} tit_hip_learned_report;
/* n_steps learned steps.  The last step of every group of judge_every is
 * judged: it is taken from state S and compared with 4 plain steps from S;
 * the plain result is kept.  The group is rejected when the max relative
 * density error over fluid particles exceeds tol or the error of the step's
 * velocity change (max_dv_err's measure) exceeds dv_tol: the state goes back
 * to the last judged state and the group's interval is covered by plain
 * steps. */
tit_status tit_hip_learned_run(tit_hip_ctx* ctx, int n_steps, int judge_every,
                               double tol, double dv_tol,
                               tit_hip_learned_report* report);

/* The network's BF16 MFMA GEMM on the current device (test hook, no
 * context): A is M x K (K x M with transA), B is K x N (N x K with transB),
 * row-major host FP32, rounded to BF16; C (M x N) = act(op(A) op(B) + bias)
 * with act 0 (none) or 1 (hard-swish), FP32 accumulation.  k_splits > 1 adds
 * op(A) op(B) into C over that many K slices instead (no bias or act). */
tit_status tit_hip_nn_gemm(int M, int N, int K, const float* A, int transA,
                           const float* B, int transB, const float* bias,
                           int act, int k_splits, float* C);

/* ---- host-side scene set-up (no device needed) ---------------------------
 * 2D/3D dam break on a lattice as titwcsph builds it (wcsph.cpp:104-142):
 * fluid at dr*(i+1, j+1[, k+1]) for a column of column[0] x column[1]
 * (x column[2]), hydrostatic density from the reference's Fourier series,
 * and `layers` layers of fixed particles at offsets 0, -dr, -2dr, ... outside
 * the tank walls x = 0, x = tank[0], y = 0 (and z = 0, z = tank[2] in 3D),
 * walls rising to tank[1].  Returns the particle counts; the fill call writes
 * r, rho and the fixed particles' wall normals. */
tit_status tit_dam_break_count(int dim, double dr, const double* tank,
                               const double* column, int layers,
                               size_t* n_fluid, size_t* n_fixed);
tit_status tit_dam_break_fill(int dim, double dr, const double* tank,
                              const double* column, int layers, double rho0,
                              double g, double cs0, double* r, double* rho,
                              double* normal);

#ifdef __cplusplus
}
#endif

#endif /* TIT_HIP_H */
