// C ABI of the MI355X WCSPH step: context, buffers, step schedule.
//
// Buffers of one SSPRK3 step (state in A at the step start):
//   begin    pending dt -> dt; per tank: rebuild if the cell index could
//            stop covering every neighbour by the end of the step
//   rebuild  (marked tanks only) keys(A) -> stable radix sort -> permute A
//            into T and copy back -> cell offsets -> chunk list
//   stage 1  boundary(A); B = A + dt L(A)
//   stage 2  boundary(B); C = 3/4 A + 1/4 (B + dt L(B))
//   stage 3  boundary(C); A = 1/3 A + 2/3 (C + dt L(C))  (+ next dt, in place:
//            each particle reads and writes only its own slot of A)
// No buffer changes roles, and the rebuild decision is taken on the device,
// so one captured hipGraph (two steps) serves every step.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/tit_hip.h"
#include "context.hpp"

using namespace tit_amd;

namespace tit_amd {

// ---- profiler ----------------------------------------------------------------

hipEvent_t Profiler::get_event() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void Profiler::begin(hipStream_t s, const char* name) {
  Rec r{name, get_event(), get_event()};
  (void)hipEventRecord(r.a, s);
  pending_.push_back(r);
}

void Profiler::end(hipStream_t s) {
  if (!pending_.empty()) (void)hipEventRecord(pending_.back().b, s);
}

void Profiler::collect() {
  for (Rec& r : pending_) {
    (void)hipEventSynchronize(r.b);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    auto it = std::find_if(totals.begin(), totals.end(),
                           [&](const Total& t) { return t.name == r.name; });
    if (it == totals.end()) {
      totals.push_back(Total{r.name, 0, 0.0});
      it = totals.end() - 1;
    }
    it->launches += 1;
    it->total_ms += ms;
    pool_.push_back(r.a);
    pool_.push_back(r.b);
  }
  pending_.clear();
}

void Profiler::clear() {
  collect();
  totals.clear();
}

Profiler::~Profiler() {
  for (Rec& r : pending_) {
    pool_.push_back(r.a);
    pool_.push_back(r.b);
  }
  for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
}

}  // namespace tit_amd


namespace tit_amd {

tit_status fail(tit_hip_ctx* c, tit_status s, const std::string& msg) {
  if (c) c->err = msg;
  return s;
}



void drop_graph(tit_hip_ctx* c) {
  for (int k = 0; k < 2; ++k) {
    if (c->graph[k]) {
      (void)hipGraphExecDestroy(c->graph[k]);
      c->graph[k] = nullptr;
    }
    if (c->graph_def[k]) {
      (void)hipGraphDestroy(c->graph_def[k]);
      c->graph_def[k] = nullptr;
    }
  }
}

void free_all(tit_hip_ctx* c) {
  drop_graph(c);
  for (void* p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  c->uploaded = false;
}

hipError_t alloc_state(tit_hip_ctx* c, DevState& s, uint32_t n, int dim) {
  hipError_t e = hipSuccess;
  for (int k = 0; k < 3; ++k) s.x[k] = s.v[k] = nullptr;
  for (int k = 0; k < dim && e == hipSuccess; ++k) {
    e = dalloc(c, &s.x[k], n);
    if (e == hipSuccess) e = dalloc(c, &s.v[k], n);
  }
  if (e == hipSuccess) e = dalloc(c, &s.rho, n);
  if (e == hipSuccess) e = dalloc(c, &s.aux, n);
  if (e == hipSuccess) e = dalloc(c, &s.pv, n);
  s.vol = nullptr;
  if (e == hipSuccess && dim == 3) e = dalloc(c, &s.vol, n);
  return e;
}

bool is_small_int(double x, int lo, int hi) {
  return x == std::floor(x) && x >= lo && x <= hi;
}

double kernel_omega(int kernel, int dim) {
  const double pi = 3.14159265358979323846;
  if (kernel == TIT_KERNEL_WENDLAND_C2) {
    return dim == 2 ? 7.0 / (4.0 * pi) : 21.0 / (16.0 * pi);
  }
  return dim == 2 ? 9.0 / (4.0 * pi) : 495.0 / (256.0 * pi);
}

// Picks the pair-term flavour (walls, per-particle masses, or the
// parameters' own) and folds its uniform factors (DevParams::mg, kd, kv).
void set_flavour(tit_hip_ctx* c, bool masses) {
  const tit_hip_params& p = c->prm;
  DevParams& d = c->dp;
  if (c->walls || (masses && p.flavour == TIT_FLAVOUR_REF)) {
    d.flavour = tit_amd::kFlavourWall;
  } else if (masses) {
    d.flavour = tit_amd::kFlavourCfgM;
  } else {
    d.flavour = p.flavour;
  }
  const double omega = kernel_omega(p.kernel, p.dim);
  const double c_w = p.kernel == TIT_KERNEL_WENDLAND_C2 ? -5.0 : -14.0 / 3.0;
  const double gsc = c_w * omega / std::pow(p.h, p.dim + 2);
  d.mg = static_cast<float>(p.m * gsc);
  if (p.flavour == TIT_FLAVOUR_CFG) {
    d.kd = static_cast<float>(2.0 * p.delta * p.h * p.cs0 * gsc);
    d.kv = static_cast<float>(masses ? 0.5 * p.alpha * p.h : 0.5 * p.alpha * p.h * p.m * gsc);
  } else {
    d.kd = static_cast<float>(gsc);
    d.kv = static_cast<float>(2.0 * p.mu * p.m * gsc);
  }
}

tit_status fill_params(tit_hip_ctx* c, const tit_hip_params& p) {
  if (p.dim != 2 && p.dim != 3) return fail(c, TIT_E_ARG, "dim must be 2 or 3");
  if (p.kernel != TIT_KERNEL_WENDLAND_C2 && p.kernel != TIT_KERNEL_WENDLAND_C4) {
    return fail(c, TIT_E_ARG, "unknown kernel");
  }
  if (p.flavour != TIT_FLAVOUR_REF && p.flavour != TIT_FLAVOUR_CFG) {
    return fail(c, TIT_E_ARG, "unknown flavour");
  }
  if (p.ssprk_order != 2 && p.ssprk_order != 3) {
    return fail(c, TIT_E_ARG, "ssprk_order must be 2 or 3");
  }
  if (!(p.h > 0.0) || !(p.m > 0.0) || !(p.rho0 > 0.0) || !(p.cs0 > 0.0) ||
      !(p.xi > 1.0) || !(p.g >= 0.0) || !(p.mu >= 0.0) || !(p.cfl > 0.0) ||
      !(p.c_force > 0.0) || !(p.skin >= 0.0)) {
    return fail(c, TIT_E_ARG, "non-physical parameter");
  }
  if (p.flavour == TIT_FLAVOUR_REF && p.mu > 0.0 && !(p.c_visc > 0.0)) {
    return fail(c, TIT_E_ARG, "c_visc must be positive when mu > 0");
  }
  if (p.post_integrate != 0 && p.post_integrate != 1) {
    return fail(c, TIT_E_ARG, "post_integrate must be 0 or 1");
  }
  c->prm = p;
  DevParams& d = c->dp;
  d = DevParams{};
  d.dim = p.dim;
  d.kernel = p.kernel;
  d.flavour = p.flavour;
  d.order = p.ssprk_order;
  d.xi_int = is_small_int(p.xi, 2, 15) ? static_cast<int>(p.xi) : 0;
  d.h = p.h;
  d.m = p.m;
  d.rho0 = p.rho0;
  d.cs0 = p.cs0;
  d.xi = p.xi;
  d.g = p.g;
  d.mu = p.mu;
  d.alpha = p.alpha;
  d.delta = p.delta;
  d.cfl = p.cfl;
  d.c_force = p.c_force;
  d.c_visc = p.c_visc;
  const double radius = 2.0 * p.h;
  d.R2 = radius * radius;
  d.eos_B = p.rho0 * (p.cs0 * p.cs0) / p.xi;
  d.inv_rho0 = 1.0 / p.rho0;
  const double skin = (p.skin > 0.0 ? p.skin : 0.1) * radius;
  d.skin_half = 0.5 * skin;
  const double omega = kernel_omega(p.kernel, p.dim);
  const double c_w = p.kernel == TIT_KERNEL_WENDLAND_C2 ? -5.0 : -14.0 / 3.0;
  d.inv_h = static_cast<float>(1.0 / p.h);
  d.gsc = static_cast<float>(c_w * omega / std::pow(p.h, p.dim + 2));
  d.wsc = static_cast<float>(omega / std::pow(p.h, p.dim));
  d.m_f = static_cast<float>(p.m);
  d.dhc2 = static_cast<float>(2.0 * p.delta * p.h * p.cs0);
  d.ahc = static_cast<float>(p.alpha * p.h);
  d.eta2 = static_cast<float>(0.01 * p.h * p.h);
  d.two_mu = static_cast<float>(2.0 * p.mu);
  d.inv_rho0_f = static_cast<float>(1.0 / p.rho0);
  d.cs0_f = static_cast<float>(p.cs0);
  d.rho0_f = static_cast<float>(p.rho0);
  d.half_inv_h = static_cast<float>(0.5 / p.h);
  set_flavour(c, false);
  return TIT_OK;
}

// Cell size: (support radius + skin) / stencil half-width, a hair larger so
// that rounding in the cell index cannot break the coverage argument.
double cell_size(const tit_hip_ctx* c) {
  const double radius = 2.0 * c->prm.h;
  return (radius + 2.0 * c->dp.skin_half) / kStencil * (1.0 + 1e-9);
}

// ---- step pieces ----------------------------------------------------------------




StageArgs stage_args(tit_hip_ctx* c);

RebuildCtl rebuild_ctl(tit_hip_ctx* c) {
  return RebuildCtl{c->rebuild_ctl, c->tank_need, c->rebuild_count};
}

// Cell rebuild: keys -> stable radix sort -> permute into T and copy back ->
// cell offsets -> chunk list.  Buffers keep their roles; on the device the
// whole sequence runs only for the tanks marked before it (by step_begin,
// or every tank with kRebuildForce, or every tank only when a forced
// rebuild is pending with kRebuildPending).
void enqueue_rebuild(tit_hip_ctx* c, RebuildMode mode) {
  hipStream_t s = c->stream;
  const RebuildCtl rc = rebuild_ctl(c);
  if (mode == kRebuildForce) launch_force_rebuild(s, c->n_tanks, rc);
  if (mode == kRebuildPending) launch_pending_rebuild(s, c->n_tanks, rc);
  {
    ProfScope ps(c->prof, s, "cell_keys");
    launch_cell_keys(s, c->dp, c->tanks, c->n, c->A, c->tag, c->key, c->counter + 1, rc);
  }
  radix_sort_pairs(s, c->prof, c->n, c->key_bits, c->key, c->key_alt, c->val, c->val_alt,
                   c->sort, rc.ctl);
  {
    ProfScope ps(c->prof, s, "permute");
    launch_permute(s, c->dp, c->n, c->val, c->A, c->T, c->id, c->id_alt, c->tag, c->tag_alt,
                   c->fixed_list, c->counter, rc.ctl);
    launch_restore_order(s, c->prm.dim, c->n, c->val, c->T, c->A, c->id_alt, c->id,
                         c->tag_alt, c->tag, rc.ctl);
  }
  {
    ProfScope ps(c->prof, s, "cell_offsets");
    launch_cell_offsets(s, c->n_cells, c->n, c->key, c->tag, c->cell_off, c->cell_fluid,
                        c->cell_gaps, c->cell_n_gaps, rc.ctl);
  }
  {
    ProfScope ps(c->prof, s, "fixed_active");
    launch_fixed_active(s, boundary_args(c, c->A), rc.ctl);
  }
  {
    ProfScope ps(c->prof, s, "build_chunks");
    StageArgs a = stage_args(c);
    launch_build_chunks(s, a, c->block_tot, rc.ctl);
  }
  launch_reset_disp(s, c->n_tanks, c->disp, rc);
}

BoundaryArgs boundary_args(tit_hip_ctx* c, const DevState& st) {
  BoundaryArgs a{};
  a.P = c->dp;
  a.tanks = c->tanks;
  a.s = st;
  a.key = c->key;
  a.tag = c->tag;
  a.id = c->id;
  a.cell_off = c->cell_off;
  a.fixed_list = c->fixed_list;
  a.active = c->fixed_active;
  a.n_active = c->counter + 2;
  a.normal = c->normal;
  a.cell_fluid = c->cell_fluid;
  a.n_fixed = c->n_fixed;
  if (c->walls) {
    a.wall_mu = c->wall_mu;
    a.wall_rho = c->wall_rho;
    a.wall_id0 = c->wall_id0;
  }
  return a;
}

void enqueue_boundary(tit_hip_ctx* c, const DevState& st, bool for_post_integrate) {
  const BoundaryArgs a = boundary_args(c, st);
  {
    ProfScope ps(c->prof, c->stream, "boundary");
    launch_boundary(c->stream, a);
  }
  if (c->walls) {  // compute_gamma + the face terms of this state
    tit_amd::WallArgs w{};
    w.P = c->dp;
    w.n = c->n;
    w.s = st;
    w.tag = c->tag;
    w.faces = c->wall_face_d;
    w.cell_faces_off = c->wall_cell_off;
    w.cell_faces = c->wall_cell_faces;
    w.cell_inside = c->wall_cell_inside;
    for (int k = 0; k < 2; ++k) {
      w.lo[k] = c->wall_lo[k];
      w.inv_cw[k] = c->wall_inv_cw[k];
      w.nc[k] = c->wall_nc[k];
    }
    w.cedges = c->wall_cedges;
    w.n_cedges = c->wall_n_cedges;
    w.vert_rho = c->wall_rho;
    w.out = c->wall_out;
    w.near = c->wall_near;
    w.near_count = c->wall_near_count;
    if (for_post_integrate) {
      w.near_idx = c->wall_near_idx;
      w.pi_face = c->wall_pi_face;
    }
    ProfScope ps(c->prof, c->stream, "wall");
    tit_amd::launch_wall(c->stream, w, c->n_cu);
  }
}

void forget(tit_hip_ctx* c, void* q) {
  if (!q) return;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), q);
  if (it != c->allocs.end()) c->allocs.erase(it);
  (void)hipFree(q);
}

// Cell columns of every tank (2D: one per x cell, 3D: one per (x, y) cell)
// and chunk buffers sized for the current chunk length.
tit_status build_schedule(tit_hip_ctx* c) {
  const int dim = c->prm.dim;
  std::vector<uint4> cols;
  for (size_t t = 0; t < c->host_tanks.size(); ++t) {
    const DevTank& T = c->host_tanks[t];
    const int nz = T.nc[dim - 1];
    const int ny = dim == 3 ? T.nc[1] : 1;
    for (int ix = 0; ix < T.nc[0]; ++ix) {
      for (int iy = 0; iy < ny; ++iy) {
        const uint32_t col_row =
            T.cell_off + (static_cast<uint32_t>(ix) * ny + static_cast<uint32_t>(iy)) * nz;
        cols.push_back(make_uint4(static_cast<uint32_t>(t), ix, iy, col_row));
      }
    }
  }
  if (cols.size() >= (1ull << 31)) return fail(c, TIT_E_ARG, "too many cell columns");
  // equal parts: n / chunk_targets + one per column at most; cuts by span:
  // (nz - 1) / kChunkRows per column at most
  int nz_max = 1;
  for (const DevTank& T : c->host_tanks) nz_max = std::max(nz_max, T.nc[dim - 1]);
  const size_t need = c->n / c->chunk_targets +
                      cols.size() * (1 + static_cast<size_t>(nz_max - 1) / kChunkRows) + 1;
  if (!c->cols || cols.size() != c->n_cols) {
    forget(c, c->cols);
    forget(c, c->block_tot);
    c->cols = nullptr;
    c->block_tot = nullptr;
    TIT_HIP_TRY(c, dalloc(c, &c->cols, cols.size()));
    // per workgroup of the chunk builder (16 columns each)
    TIT_HIP_TRY(c, dalloc(c, &c->block_tot, cols.size() / 16 + 2));
    TIT_HIP_TRY(c, hipMemcpy(c->cols, cols.data(), cols.size() * sizeof(uint4),
                             hipMemcpyHostToDevice));
    c->n_cols = static_cast<uint32_t>(cols.size());
  }
  if (need > c->chunk_cap) {
    forget(c, c->chunks);
    forget(c, c->overflow);
    c->chunks = nullptr;
    c->overflow = nullptr;
    TIT_HIP_TRY(c, dalloc(c, &c->chunks, need));
    TIT_HIP_TRY(c, dalloc(c, &c->overflow, need));
    c->chunk_cap = need;
  }
  if (!c->chunk_counts) TIT_HIP_TRY(c, dalloc(c, &c->chunk_counts, 4));
  if (dim == 3 && !c->blocks) {
    // blocks of kBlk3 x kBlk3 columns; a block chunk holds at least one
    // particle and one non-empty row of its block
    std::vector<uint4> blocks;
    for (size_t t = 0; t < c->host_tanks.size(); ++t) {
      const DevTank& T = c->host_tanks[t];
      for (int bx = 0; bx < T.nc[0]; bx += kBlk3) {
        for (int by = 0; by < T.nc[1]; by += kBlk3) {
          blocks.push_back(make_uint4(static_cast<uint32_t>(t), bx, by, 0u));
        }
      }
    }
    const size_t cap =
        std::min(static_cast<size_t>(c->n), blocks.size() * static_cast<size_t>(nz_max)) + 1;
    TIT_HIP_TRY(c, dalloc(c, &c->blocks, blocks.size()));
    TIT_HIP_TRY(c, dalloc(c, &c->bcount, blocks.size() + 1));
    TIT_HIP_TRY(c, dalloc(c, &c->bchunks, cap));
    TIT_HIP_TRY(c, dalloc(c, &c->boverflow, cap));
    TIT_HIP_TRY(c, hipMemcpy(c->blocks, blocks.data(), blocks.size() * sizeof(uint4),
                             hipMemcpyHostToDevice));
    c->n_blocks = static_cast<uint32_t>(blocks.size());
    c->bchunk_cap = cap;
  }
  StageGrids g;
  stage_grids(c->dp, c->n_cu, &g);
  c->grid_lds = g.lds;
  c->grid_global = g.global;
  c->grid_lds3 = g.lds3;
  c->grid_block_global = g.block_global;
  drop_graph(c);
  if (c->rebuild_ctl) {
    // the chunk list follows the schedule: recut it from the current cell
    // offsets (gated on ctl[2], so nothing runs before the first rebuild)
    StageArgs a = stage_args(c);
    launch_build_chunks(c->stream, a, c->block_tot, c->rebuild_ctl + 2);
    TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  return TIT_OK;
}

StageArgs stage_args(tit_hip_ctx* c) {
  StageArgs a{};
  a.cols = c->cols;
  a.n_cols = c->n_cols;
  a.chunk_targets = c->chunk_targets;
  a.lds_cap = c->lds_cap;
  a.chunks = c->chunks;
  a.overflow = c->overflow;
  a.counts = c->chunk_counts;
  a.grid_lds = c->grid_lds;
  a.grid_global = c->grid_global;
  a.blocks = c->blocks;
  a.n_blocks = c->n_blocks;
  a.bchunks = c->bchunks;
  a.boverflow = c->boverflow;
  a.bcount = c->bcount;
  a.grid_lds3 = c->grid_lds3;
  a.grid_block_global = c->grid_block_global;
  a.P = c->dp;
  a.tanks = c->tanks;
  a.key = c->key;
  a.tag = c->tag;
  a.id = c->id;
  a.cell_off = c->cell_off;
  a.dt = c->dt;
  a.n = c->n;
  a.dt_min = c->dt_min;
  a.vmax = c->vmax;
  a.rhs_drho = c->rhs_drho;
  a.rhs_count = c->rhs_count;
  for (int k = 0; k < 3; ++k) a.rhs_dv[k] = c->rhs_dv[k];
  for (int k = 0; k < 3; ++k) a.dv_last[k] = c->prm.post_integrate ? c->rhs_dv[k] : nullptr;
  a.wall = c->wall_out;
  return a;
}

void enqueue_stage(tit_hip_ctx* c, const DevState& in, const DevState& old,
                   const DevState& out, int kind, double w, bool final) {
  StageArgs a = stage_args(c);
  a.in = in;
  a.old = old;
  a.out = out;
  a.w = w;
  ProfScope ps(c->prof, c->stream, "stage");
  launch_stage(c->stream, a, kind, final);
}

// One SSPRK step, SSPRKIntegrator::step (time_integrator.hpp:159-184).
void enqueue_step(tit_hip_ctx* c) {
  hipStream_t s = c->stream;
  launch_step_begin(s, c->n_tanks, c->dt_min, c->dt, c->vmax, c->disp, c->dp.skin_half,
                    rebuild_ctl(c));
  enqueue_rebuild(c, kRebuildMarked);
  enqueue_boundary(c, c->A);
  enqueue_stage(c, c->A, c->A, c->B, kEuler, 1.0, false);
  enqueue_boundary(c, c->B);
  if (c->prm.ssprk_order == 2) {
    enqueue_stage(c, c->B, c->A, c->A, kCombine, 1.0 / 2.0, true);
  } else {
    enqueue_stage(c, c->B, c->A, c->C, kCombine, 1.0 / 4.0, false);
    enqueue_boundary(c, c->C);
    enqueue_stage(c, c->C, c->A, c->A, kCombine, 2.0 / 3.0, true);
  }
  launch_step_end(s, c->n_tanks, c->dt, c->vmax, c->time, c->disp,
                  c->dp.skin_half, c->counter + 1);
  if (c->prm.post_integrate) {
    enqueue_boundary(c, c->A, true);  // prepare (fluid_equations.hpp:99-105)
    enqueue_post_integrate(c);
  }
}

// FluidEquations::post_integrate (fluid_equations.hpp:313-321) on the state
// A: the index is the step's (it covers the post-step positions by the skin
// bound); the boundary values are refreshed by the caller.
void enqueue_post_integrate(tit_hip_ctx* c) {
  const int dim = c->prm.dim;
  const size_t n = c->n;
  PiArgs a{};
  a.P = c->dp;
  a.tanks = c->tanks;
  a.n = c->n;
  a.a = c->A;
  a.b = c->B;
  a.c = c->C;
  a.key = c->key;
  a.tag = c->tag;
  a.id = c->id;
  a.cell_off = c->cell_off;
  double* p = c->pi_buf;
  for (int k = 0; k < dim; ++k, p += n) a.N[k] = p;
  for (int k = 0; k < dim; ++k, p += n) a.dr0[k] = p;
  for (int k = 0; k < dim * dim; ++k, p += n) a.grad_v[k] = p;
  for (int k = 0; k < dim; ++k, p += n) a.grad_rho[k] = p;
  a.phi = p;
  a.phi_near = p + n;
  a.vol = p + 2 * n;
  if (c->pi_pack) {
    a.pk_x = c->pi_pack;
    a.pk_v = c->pi_pack + n;
    a.pk_rv = c->pi_pack + 2 * n;
  }
  a.count = c->pi_count;
  a.cell_mark = c->pi_cells;
  a.cell_near = c->pi_cells + c->n_cells + 1;
  a.n_cells = c->n_cells;
  a.shift_max = c->pi_shift;
  for (int k = 0; k < dim; ++k) a.dv_last[k] = c->rhs_dv[k];
  a.dt_min = c->dt_min;
  a.vmax = c->vmax;
  if (c->walls) {
    a.wall = c->wall_out;
    a.wall_mu = c->wall_mu;
    a.wall_id0 = c->wall_id0;
    a.near_idx = c->wall_near_idx;
    a.pi_face = c->wall_pi_face;
  }
  ProfScope ps(c->prof, c->stream, "post_integrate");
  launch_post_integrate(c->stream, a, c->n_tanks, c->disp, c->dp.skin_half,
                        rebuild_ctl(c).ctl, c->n_cu);
}

tit_status check_launch(tit_hip_ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    return fail(c, TIT_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  }
  return TIT_OK;
}

tit_status read_flags(tit_hip_ctx* c) {
  uint32_t flags = 0;
  TIT_HIP_TRY(c, hipMemcpyAsync(&flags, c->counter + 1, sizeof(uint32_t),
                                hipMemcpyDeviceToHost, c->stream));
  TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (flags & 2u) {
    return fail(c, TIT_E_NUMERIC, "non-finite particle position");
  }
  if (flags & 1u) {
    return fail(c, TIT_E_NUMERIC,
                "particles moved farther than half the neighbour skin within "
                "one step; raise tit_hip_params.skin");
  }
  return TIT_OK;
}

// Scratch of tit_hip_neighbors_noalloc, one copy per GPU for the life of the
// process: per-particle neighbour counts, then list offsets, and the device
// copy of the lists.  Host code never passes the array's name to a kernel:
// scratch_ptr() gives each device's own address (hipGetSymbolAddress at
// tit_hip_create, with the device current).
constexpr size_t kNoAllocParticles = size_t(1) << 20;
constexpr size_t kNoAllocEntries = size_t(1) << 24;
struct NeighborScratch {
  uint64_t off[kNoAllocParticles + 1];
  uint32_t idx[kNoAllocEntries];
};
__device__ __align__(256) NeighborScratch scratch;

constexpr int kMaxDevices = 64;
struct ScratchSlot {
  std::once_flag once;
  NeighborScratch* ptr = nullptr;
  hipError_t err = hipSuccess;
};
ScratchSlot g_scratch[kMaxDevices];

// Address of `scratch` on `device` (the current device on the first call
// for it); null when the device has none.
NeighborScratch* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  ScratchSlot& s = g_scratch[device];
  std::call_once(s.once, [&s] {
    void* p = nullptr;
    s.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (s.err == hipSuccess) s.ptr = static_cast<NeighborScratch*>(p);
  });
  return s.ptr;
}

// The autotuner's per-process cache: the schedule found for a kind of
// context (dimension, kernel, term set, walls, one or many tanks) and size
// (the particle count's power of two).
struct TuneKey {
  int dim, kernel, flavour, walls, batched, log2n;
  bool operator<(const TuneKey& o) const {
    return std::tie(dim, kernel, flavour, walls, batched, log2n) <
           std::tie(o.dim, o.kernel, o.flavour, o.walls, o.batched, o.log2n);
  }
};
struct TuneVal {
  uint32_t chunk_targets, lds_cap;
  double ms_per_step;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;

TuneKey tune_key(const tit_hip_ctx* c) {
  int log2n = 0;
  while ((static_cast<size_t>(1) << (log2n + 1)) <= c->n) ++log2n;
  return TuneKey{c->prm.dim, c->prm.kernel, c->dp.flavour, c->walls ? 1 : 0,
                 c->n_tanks > 1 ? 1 : 0, log2n};
}

}  // namespace

namespace tit_amd {

// n_steps plain SSPRK steps: pairs of steps replay the two-step graph of the
// current rebuild parity (captured on first use), an odd remainder runs
// eagerly.  Synchronises and checks the device flags.
tit_status run_plain(tit_hip_ctx* c, int n_steps) {
  int done = 0;
  if (!c->prof.enabled && n_steps >= 2) {
    const int par = 0;  // buffers keep their roles: one graph serves every step
    if (!c->graph[par]) {
      TIT_HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      enqueue_step(c);
      enqueue_step(c);
      hipError_t e = hipStreamEndCapture(c->stream, &c->graph_def[par]);
      if (e != hipSuccess) {
        c->graph_def[par] = nullptr;
        return fail(c, TIT_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
      }
      TIT_HIP_TRY(c, hipGraphInstantiate(&c->graph[par], c->graph_def[par], nullptr, nullptr, 0));
    }
    for (; done + 2 <= n_steps; done += 2) {
      TIT_HIP_TRY(c, hipGraphLaunch(c->graph[par], c->stream));
    }
  }
  for (; done < n_steps; ++done) enqueue_step(c);
  c->steps += static_cast<uint64_t>(n_steps);
  tit_status st = check_launch(c, "step");
  if (st != TIT_OK) return st;
  if (c->prof.enabled) c->prof.collect();
  return read_flags(c);
}

}  // namespace tit_amd

extern "C" {

tit_status tit_hip_create(int device, const tit_hip_params* params,
                          tit_hip_ctx** out) {
  if (!out || !params) return TIT_E_ARG;
  *out = nullptr;
  auto* c = new (std::nothrow) tit_hip_ctx();
  if (!c) return TIT_E_OOM;
  tit_status st = fill_params(c, *params);
  if (st != TIT_OK) {
    delete c;
    return st;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    delete c;
    return TIT_E_HIP;
  }
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete c;
    return TIT_E_HIP;
  }
  c->n_cu = prop.multiProcessorCount;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return TIT_E_HIP;
  }
  (void)scratch_ptr(device);  // this device's scratch address, once per device
  *out = c;
  return TIT_OK;
}

tit_status tit_hip_set_walls(tit_hip_ctx* c, const tit_hip_walls* w) {
  if (!c || !w) return TIT_E_ARG;
  if (c->uploaded) return fail(c, TIT_E_STATE, "set_walls after upload");
  if (c->prm.dim != 2) return fail(c, TIT_E_ARG, "walls are 2D only");
  if (c->prm.flavour != TIT_FLAVOUR_REF) {
    return fail(c, TIT_E_ARG, "walls need the reference flavour");
  }
  if (!w->verts || w->n_verts < 2 || !w->faces || w->n_faces == 0 || !w->cverts ||
      w->n_cverts < 3 || !w->cfaces || w->n_cfaces < 3) {
    return fail(c, TIT_E_ARG, "walls: empty surface or containment polygon");
  }
  for (uint32_t k = 0; k < 2 * w->n_faces; ++k) {
    if (w->faces[k] >= w->n_verts) return fail(c, TIT_E_ARG, "walls: face vertex out of range");
  }
  for (uint32_t k = 0; k < 2 * w->n_cfaces; ++k) {
    if (w->cfaces[k] >= w->n_cverts) {
      return fail(c, TIT_E_ARG, "walls: containment vertex out of range");
    }
  }
  for (uint32_t k = 0; k < 2 * w->n_verts; ++k) {
    if (!std::isfinite(w->verts[k])) return fail(c, TIT_E_ARG, "walls: non-finite vertex");
  }
  c->wall_verts.assign(w->verts, w->verts + 2 * static_cast<size_t>(w->n_verts));
  c->wall_faces.assign(w->faces, w->faces + 2 * static_cast<size_t>(w->n_faces));
  c->wall_cverts.assign(w->cverts, w->cverts + 2 * static_cast<size_t>(w->n_cverts));
  c->wall_cfaces.assign(w->cfaces, w->cfaces + 2 * static_cast<size_t>(w->n_cfaces));
  c->walls = true;
  set_flavour(c, false);
  return TIT_OK;
}

void tit_hip_destroy(tit_hip_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  destroy_learned(c);
  free_all(c);
  c->prof.clear();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* tit_hip_last_error(const tit_hip_ctx* c) {
  return c ? c->err.c_str() : "null context";
}

tit_status tit_hip_upload(tit_hip_ctx* c, const tit_hip_host_view* view,
                          const tit_hip_tank* tank) {
  return tit_hip_upload_batch(c, view, tank, 1);
}

tit_status tit_hip_upload_batch(tit_hip_ctx* c, const tit_hip_host_view* views,
                                const tit_hip_tank* tanks, size_t n_tanks) {
  if (!c) return TIT_E_ARG;
  if (!views || !tanks || n_tanks == 0 || n_tanks > (1u << 20)) {
    return fail(c, TIT_E_ARG, "need at least one tank");
  }
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
  free_all(c);
  const int dim = c->prm.dim;

  // Sizes and tank grids.
  size_t n_total = 0, n_fixed_total = 0;
  for (size_t t = 0; t < n_tanks; ++t) {
    const tit_hip_host_view& v = views[t];
    if (!v.r || !v.rho) return fail(c, TIT_E_ARG, "r and rho are required");
    if (v.n_fixed > 0 && !v.normal && !c->walls) {
      return fail(c, TIT_E_ARG, "fixed particles need wall normals");
    }
    n_total += v.n_fluid + v.n_fixed;
    n_fixed_total += v.n_fixed;
  }
  if (n_total == 0) return fail(c, TIT_E_ARG, "no particles");
  if (n_total >= (1ull << 31)) return fail(c, TIT_E_ARG, "too many particles");
  const double csz = cell_size(c);
  std::vector<DevTank> ht(n_tanks);
  uint64_t cells = 0;
  for (size_t t = 0; t < n_tanks; ++t) {
    DevTank& d = ht[t];
    uint64_t tc = 1;
    for (int k = 0; k < 3; ++k) {
      d.lo[k] = 0.0;
      d.inv_c[k] = 0.0;
      d.nc[k] = 1;
    }
    for (int k = 0; k < dim; ++k) {
      const double lo = tanks[t].lo[k], hi = tanks[t].hi[k];
      if (!(hi > lo)) return fail(c, TIT_E_ARG, "empty tank box");
      const double nf = std::ceil((hi - lo) / csz);
      if (nf > 1e7) return fail(c, TIT_E_ARG, "tank box too large for h");
      d.nc[k] = std::max(1, static_cast<int>(nf));
      d.lo[k] = lo;
      d.inv_c[k] = 1.0 / csz;
      tc *= static_cast<uint64_t>(d.nc[k]);
    }
    d.cell_off = static_cast<uint32_t>(cells);
    d.n_cells = static_cast<uint32_t>(tc);
    cells += tc;
    if (cells >= (1ull << 31)) return fail(c, TIT_E_ARG, "too many cells");
  }
  c->n = static_cast<uint32_t>(n_total);
  c->n_fixed = static_cast<uint32_t>(n_fixed_total);
  c->n_tanks = static_cast<int>(n_tanks);
  c->dp.n_tanks = c->n_tanks;
  c->n_cells = static_cast<uint32_t>(cells);
  c->key_bits = 1;
  while ((1ull << c->key_bits) < cells) ++c->key_bits;

  // Host staging in reference layout.
  std::vector<double> hr(n_total * dim), hv(n_total * dim, 0.0), hrho(n_total),
      hn(n_total * dim, 0.0), hdv;
  std::vector<uint32_t> htag(n_total);
  bool any_dv = false;
  for (size_t t = 0; t < n_tanks; ++t) any_dv = any_dv || views[t].dv_dt;
  if (any_dv) hdv.assign(n_total * dim, 0.0);
  size_t o = 0;
  for (size_t t = 0; t < n_tanks; ++t) {
    const tit_hip_host_view& v = views[t];
    const size_t nt = v.n_fluid + v.n_fixed;
    std::memcpy(&hr[o * dim], v.r, nt * dim * sizeof(double));
    if (v.v) std::memcpy(&hv[o * dim], v.v, nt * dim * sizeof(double));
    std::memcpy(&hrho[o], v.rho, nt * sizeof(double));
    if (v.dv_dt) std::memcpy(&hdv[o * dim], v.dv_dt, nt * dim * sizeof(double));
    if (v.n_fixed && v.normal) {
      std::memcpy(&hn[(o + v.n_fluid) * dim], v.normal,
                  v.n_fixed * dim * sizeof(double));
    }
    for (size_t a = 0; a < nt; ++a) {
      htag[o + a] = static_cast<uint32_t>(t << 1) | (a >= v.n_fluid ? 1u : 0u);
      // fixed particles start with zero velocity (setup_boundary sets v = 0)
      if (a >= v.n_fluid) {
        for (int k = 0; k < dim; ++k) hv[(o + a) * dim + k] = 0.0;
      }
    }
    o += nt;
  }
  for (double x : hr) {
    if (!std::isfinite(x)) return fail(c, TIT_E_ARG, "non-finite position");
  }
  for (double x : hrho) {
    if (!(x > 0.0) || !std::isfinite(x)) {
      return fail(c, TIT_E_ARG, "density must be positive and finite");
    }
  }
  // Per-particle masses (ParticleArray's m); tanks without them take params.m.
  // Masses that all equal params.m take the uniform-mass terms.  2D only: the
  // pair record's third slot carries V = m / rho there.
  std::vector<double> hm;
  bool any_m = false;
  {
    size_t q = 0;
    for (size_t t = 0; t < n_tanks; ++t) {
      const size_t nt = views[t].n_fluid + views[t].n_fixed;
      if (views[t].m) {
        if (hm.empty()) hm.assign(n_total, c->prm.m);
        std::memcpy(&hm[q], views[t].m, nt * sizeof(double));
      }
      q += nt;
    }
    for (double x : hm) {
      if (!(x > 0.0) || !std::isfinite(x)) {
        return fail(c, TIT_E_ARG, "masses must be positive and finite");
      }
      any_m = any_m || x != c->prm.m;
    }
  }

  // Semi-analytic wall: faces, face grid, and the fixed particles' gamma (their
  // mass fractions) and grad gamma (their wall normals), FluidEquations::
  // initialize (fluid_equations.hpp:75-89).
  tit_amd::WallHost wh;
  std::vector<double> wall_gamma;
  if (c->walls) {
    if (n_tanks != 1) return fail(c, TIT_E_ARG, "walls need a single tank");
    const tit_hip_host_view& v = views[0];
    const uint32_t nv = static_cast<uint32_t>(c->wall_verts.size() / 2);
    if (v.n_fixed != nv) {
      return fail(c, TIT_E_ARG, "walls: one fixed particle per wall vertex is required");
    }
    for (uint32_t k = 0; k < nv; ++k) {
      for (int d = 0; d < 2; ++d) {
        if (hr[(v.n_fluid + k) * 2 + d] != c->wall_verts[2 * k + d]) {
          return fail(c, TIT_E_ARG, "walls: fixed particle k must sit at wall vertex k");
        }
      }
    }
    std::string err;
    if (!tit_amd::wall_build(c->wall_verts.data(), nv, c->wall_faces.data(),
                             static_cast<uint32_t>(c->wall_faces.size() / 2),
                             c->wall_cverts.data(), static_cast<uint32_t>(c->wall_cverts.size() / 2),
                             c->wall_cfaces.data(), static_cast<uint32_t>(c->wall_cfaces.size() / 2),
                             tanks[0].lo, tanks[0].hi, c->prm.h, &wh, &err)) {
      return fail(c, TIT_E_ARG, "walls: " + err);
    }
    wall_gamma.assign(nv, 0.0);
    std::vector<double> grad(2 * static_cast<size_t>(nv), 0.0);
    tit_amd::wall_fixed_gamma(c->dp.kernel, wh, c->prm.h, &hr[v.n_fluid * 2], nv,
                              wall_gamma.data(), grad.data());
    std::memcpy(&hn[v.n_fluid * 2], grad.data(), grad.size() * sizeof(double));
    c->wall_id0 = static_cast<uint32_t>(v.n_fluid);
  }

  // Mean occupancy of the occupied cells (sets the default strip length).
  {
    std::vector<uint64_t> keys(n_total);
    size_t q = 0;
    for (size_t t = 0; t < n_tanks; ++t) {
      const size_t nt = views[t].n_fluid + views[t].n_fixed;
      for (size_t a = 0; a < nt; ++a, ++q) {
        uint64_t flat = 0;
        for (int k = 0; k < dim; ++k) {
          const double f = (hr[q * dim + k] - ht[t].lo[k]) * ht[t].inv_c[k];
          int cc = f <= 0.0 ? 0 : (f >= ht[t].nc[k] - 1 ? ht[t].nc[k] - 1 : static_cast<int>(f));
          flat = flat * static_cast<uint64_t>(ht[t].nc[k]) + static_cast<uint64_t>(cc);
        }
        keys[q] = ht[t].cell_off + flat;
      }
    }
    std::sort(keys.begin(), keys.end());
    const size_t occupied = static_cast<size_t>(std::unique(keys.begin(), keys.end()) - keys.begin());
    c->occupancy = occupied ? static_cast<double>(n_total) / occupied : 1.0;
  }
  c->host_tanks = ht;

  // Device buffers.
  const uint32_t n = c->n;
  TIT_HIP_TRY(c, alloc_state(c, c->A, n, dim));
  TIT_HIP_TRY(c, alloc_state(c, c->T, n, dim));
  TIT_HIP_TRY(c, alloc_state(c, c->B, n, dim));
  TIT_HIP_TRY(c, alloc_state(c, c->C, n, dim));
  TIT_HIP_TRY(c, dalloc(c, &c->key, n));
  TIT_HIP_TRY(c, dalloc(c, &c->key_alt, n));
  TIT_HIP_TRY(c, dalloc(c, &c->val, n));
  TIT_HIP_TRY(c, dalloc(c, &c->val_alt, n));
  TIT_HIP_TRY(c, dalloc(c, &c->id, n));
  TIT_HIP_TRY(c, dalloc(c, &c->id_alt, n));
  TIT_HIP_TRY(c, dalloc(c, &c->tag, n));
  TIT_HIP_TRY(c, dalloc(c, &c->tag_alt, n));
  TIT_HIP_TRY(c, dalloc(c, &c->cell_off, static_cast<size_t>(c->n_cells) + 1));
  TIT_HIP_TRY(c, dalloc(c, &c->cell_fluid, std::max<uint32_t>(c->n_cells, 1)));
  TIT_HIP_TRY(c, dalloc(c, &c->cell_gaps, static_cast<size_t>(n_total) + 1));
  TIT_HIP_TRY(c, dalloc(c, &c->cell_n_gaps, 1));
  TIT_HIP_TRY(c, dalloc(c, &c->fixed_list, std::max<uint32_t>(c->n_fixed, 1)));
  TIT_HIP_TRY(c, dalloc(c, &c->fixed_active, std::max<uint32_t>(c->n_fixed, 1)));
  TIT_HIP_TRY(c, dalloc(c, &c->counter, 4));
  TIT_HIP_TRY(c, dalloc(c, &c->rebuild_ctl, 4));
  TIT_HIP_TRY(c, dalloc(c, &c->tank_need, n_tanks));
  TIT_HIP_TRY(c, dalloc(c, &c->rebuild_count, 1));
  c->sort.max_tiles = radix_sort_tiles(n);
  TIT_HIP_TRY(c, dalloc(c, &c->sort.hist, 2048ull * std::max<uint32_t>(c->sort.max_tiles, 1)));
  TIT_HIP_TRY(c, dalloc(c, &c->sort.tot, 2048));
  TIT_HIP_TRY(c, dalloc(c, &c->tanks, n_tanks));
  TIT_HIP_TRY(c, dalloc(c, &c->normal, static_cast<size_t>(n) * dim));
  TIT_HIP_TRY(c, dalloc(c, &c->dt, n_tanks));
  TIT_HIP_TRY(c, dalloc(c, &c->time, n_tanks));
  TIT_HIP_TRY(c, dalloc(c, &c->disp, n_tanks));
  c->mass = nullptr;
  if (any_m) TIT_HIP_TRY(c, dalloc(c, &c->mass, n));
  // per-tank reduction slots: kRedStripes copies per tank
  const size_t n_red = static_cast<size_t>(n_tanks) * kRedStripes;
  TIT_HIP_TRY(c, dalloc(c, &c->dt_min, n_red));
  TIT_HIP_TRY(c, dalloc(c, &c->vmax, n_red));
  TIT_HIP_TRY(c, dalloc(c, &c->stage_r, static_cast<size_t>(n) * dim));
  TIT_HIP_TRY(c, dalloc(c, &c->stage_v, static_cast<size_t>(n) * dim));
  TIT_HIP_TRY(c, dalloc(c, &c->stage_rho, n));
  TIT_HIP_TRY(c, dalloc(c, &c->stage_p, n));
  TIT_HIP_TRY(c, dalloc(c, &c->rhs_drho, n));
  TIT_HIP_TRY(c, dalloc(c, &c->rhs_count, n));
  for (int k = 0; k < dim; ++k) TIT_HIP_TRY(c, dalloc(c, &c->rhs_dv[k], n));
  if (c->prm.post_integrate) {
    const size_t per = static_cast<size_t>(3 * dim + dim * dim + 3);
    TIT_HIP_TRY(c, dalloc(c, &c->pi_buf, per * n));
    if (dim == 2) TIT_HIP_TRY(c, dalloc(c, &c->pi_pack, 3 * static_cast<size_t>(n)));
    TIT_HIP_TRY(c, dalloc(c, &c->pi_count, n));
    TIT_HIP_TRY(c, dalloc(c, &c->pi_cells, 2 * (static_cast<size_t>(c->n_cells) + 1)));
    TIT_HIP_TRY(c, dalloc(c, &c->pi_shift, n_red));
    TIT_HIP_TRY(c, hipMemsetAsync(c->pi_shift, 0, n_red * sizeof(unsigned long long),
                                  c->stream));
    // the first step's time step selection reads the uploaded dv/dt only
    for (int k = 0; k < dim; ++k) {
      TIT_HIP_TRY(c, hipMemsetAsync(c->rhs_dv[k], 0, n * sizeof(double), c->stream));
    }
  }

  hipStream_t s = c->stream;
  if (c->walls) {
    const size_t nv = wall_gamma.size();
    TIT_HIP_TRY(c, dalloc(c, &c->wall_face_d, std::max<size_t>(wh.faces.size(), 1)));
    TIT_HIP_TRY(c, dalloc(c, &c->wall_cell_off, wh.cell_off.size()));
    TIT_HIP_TRY(c, dalloc(c, &c->wall_cell_faces, std::max<size_t>(wh.cell_faces.size(), 1)));
    TIT_HIP_TRY(c, dalloc(c, &c->wall_cell_inside, wh.cell_inside.size()));
    TIT_HIP_TRY(c, dalloc(c, &c->wall_cedges, std::max<size_t>(wh.cedges.size(), 4)));
    TIT_HIP_TRY(c, dalloc(c, &c->wall_mu, std::max<size_t>(nv, 1)));
    TIT_HIP_TRY(c, dalloc(c, &c->wall_rho, std::max<size_t>(nv, 1)));
    TIT_HIP_TRY(c, dalloc(c, &c->wall_out, 4 * static_cast<size_t>(n)));
    TIT_HIP_TRY(c, dalloc(c, &c->wall_near, static_cast<size_t>(n)));
    if (c->prm.post_integrate) {
      TIT_HIP_TRY(c, dalloc(c, &c->wall_near_idx, static_cast<size_t>(n)));
      TIT_HIP_TRY(c, dalloc(c, &c->wall_pi_face, 12 * static_cast<size_t>(n)));
    }
    TIT_HIP_TRY(c, dalloc(c, &c->wall_near_count, 1));
    TIT_HIP_TRY(c, hipMemcpyAsync(c->wall_face_d, wh.faces.data(),
                                  wh.faces.size() * sizeof(tit_amd::WallFace),
                                  hipMemcpyHostToDevice, s));
    TIT_HIP_TRY(c, hipMemcpyAsync(c->wall_cell_off, wh.cell_off.data(),
                                  wh.cell_off.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    TIT_HIP_TRY(c, hipMemcpyAsync(c->wall_cell_faces, wh.cell_faces.data(),
                                  wh.cell_faces.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    TIT_HIP_TRY(c, hipMemcpyAsync(c->wall_cell_inside, wh.cell_inside.data(),
                                  wh.cell_inside.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    TIT_HIP_TRY(c, hipMemcpyAsync(c->wall_cedges, wh.cedges.data(),
                                  wh.cedges.size() * sizeof(double), hipMemcpyHostToDevice, s));
    TIT_HIP_TRY(c, hipMemcpyAsync(c->wall_mu, wall_gamma.data(), nv * sizeof(double),
                                  hipMemcpyHostToDevice, s));
    TIT_HIP_TRY(c, hipMemsetAsync(c->wall_rho, 0, nv * sizeof(double), s));
    c->wall_n_cedges = static_cast<uint32_t>(wh.cedges.size() / 4);
    for (int k = 0; k < 2; ++k) {
      c->wall_lo[k] = wh.lo[k];
      c->wall_inv_cw[k] = wh.inv_cw[k];
      c->wall_nc[k] = wh.nc[k];
    }
  }
  TIT_HIP_TRY(c, hipMemcpyAsync(c->stage_r, hr.data(), hr.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
  TIT_HIP_TRY(c, hipMemcpyAsync(c->stage_v, hv.data(), hv.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
  TIT_HIP_TRY(c, hipMemcpyAsync(c->stage_rho, hrho.data(), hrho.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
  TIT_HIP_TRY(c, hipMemcpyAsync(c->normal, hn.data(), hn.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
  TIT_HIP_TRY(c, hipMemcpyAsync(c->tag, htag.data(), htag.size() * sizeof(uint32_t),
                                hipMemcpyHostToDevice, s));
  TIT_HIP_TRY(c, hipMemcpyAsync(c->tanks, ht.data(), ht.size() * sizeof(DevTank),
                                hipMemcpyHostToDevice, s));
  if (any_m) {
    TIT_HIP_TRY(c, hipMemcpyAsync(c->mass, hm.data(), hm.size() * sizeof(double),
                                  hipMemcpyHostToDevice, s));
  }
  c->dp.mass = c->mass;
  set_flavour(c, any_m);
  launch_import(s, c->dp, n, c->stage_r, c->stage_v, c->stage_rho, c->A, c->id);
  std::vector<double> zeros(n_tanks, 0.0);
  std::vector<unsigned long long> inf(static_cast<size_t>(n_tanks) * kRedStripes, ~0ull);
  std::vector<uint32_t> z32(4, 0u);
  TIT_HIP_TRY(c, hipMemcpyAsync(c->time, zeros.data(), n_tanks * sizeof(double),
                                hipMemcpyHostToDevice, s));
  TIT_HIP_TRY(c, hipMemcpyAsync(c->disp, zeros.data(), n_tanks * sizeof(double),
                                hipMemcpyHostToDevice, s));
  TIT_HIP_TRY(c, hipMemcpyAsync(c->dt, zeros.data(), n_tanks * sizeof(double),
                                hipMemcpyHostToDevice, s));
  TIT_HIP_TRY(c, hipMemcpyAsync(c->dt_min, inf.data(), inf.size() * sizeof(unsigned long long),
                                hipMemcpyHostToDevice, s));
  TIT_HIP_TRY(c, hipMemsetAsync(c->vmax, 0, inf.size() * sizeof(uint32_t), s));
  TIT_HIP_TRY(c, hipMemcpyAsync(c->counter, z32.data(), 4 * sizeof(uint32_t),
                                hipMemcpyHostToDevice, s));
  // the first step rebuilds every tank
  const uint32_t ctl0[4] = {0u, 1u, 0u, 0u};
  TIT_HIP_TRY(c, hipMemcpyAsync(c->rebuild_ctl, ctl0, sizeof(ctl0), hipMemcpyHostToDevice, s));
  TIT_HIP_TRY(c, hipMemsetAsync(c->tank_need, 0, n_tanks * sizeof(uint32_t), s));
  // every cell may hold fluid until the first rebuild marks them
  TIT_HIP_TRY(c, hipMemsetAsync(c->cell_fluid, 1, std::max<uint32_t>(c->n_cells, 1), s));
  TIT_HIP_TRY(c, hipMemsetAsync(c->rebuild_count, 0, sizeof(unsigned long long), s));
  // First step size from the uploaded state and dv/dt (zero by default).
  double* d_dv = nullptr;
  if (any_dv) {
    TIT_HIP_TRY(c, hipMalloc(&d_dv, hdv.size() * sizeof(double)));
    TIT_HIP_TRY(c, hipMemcpyAsync(d_dv, hdv.data(), hdv.size() * sizeof(double),
                                  hipMemcpyHostToDevice, s));
  }
  launch_time_step(s, c->dp, n, c->A, c->tag, c->id, d_dv, c->dt_min);
  tit_status st = check_launch(c, "upload");
  TIT_HIP_TRY(c, hipStreamSynchronize(s));
  if (d_dv) (void)hipFree(d_dv);
  if (st != TIT_OK) return st;
  c->cols = nullptr;
  c->chunks = nullptr;
  c->overflow = nullptr;
  c->block_tot = nullptr;
  c->chunk_counts = nullptr;
  c->blocks = nullptr;
  c->bchunks = nullptr;
  c->boverflow = nullptr;
  c->bcount = nullptr;
  c->n_blocks = 0;
  c->bchunk_cap = 0;
  c->n_cols = 0;
  c->chunk_cap = 0;
  c->lds_cap = stage_lds_capacity(dim);
  st = build_schedule(c);
  if (st != TIT_OK) return st;
  c->uploaded = true;
  c->steps = 0;
  return TIT_OK;
}

tit_status tit_hip_step(tit_hip_ctx* c, int n_steps, double* dt_last) {
  if (!c) return TIT_E_ARG;
  if (!c->uploaded) return fail(c, TIT_E_STATE, "step before upload");
  if (n_steps < 0) return fail(c, TIT_E_ARG, "negative step count");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  tit_status st = run_plain(c, n_steps);
  if (st != TIT_OK) return st;
  if (dt_last) {
    TIT_HIP_TRY(c, hipMemcpy(dt_last, c->dt, sizeof(double), hipMemcpyDeviceToHost));
  }
  return TIT_OK;
}

tit_status tit_hip_download(tit_hip_ctx* c, double* r, double* v, double* rho,
                            double* p) {
  if (!c) return TIT_E_ARG;
  if (!c->uploaded) return fail(c, TIT_E_STATE, "download before upload");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  const int dim = c->prm.dim;
  const size_t n = c->n;
  launch_export(c->stream, c->dp, c->n, c->A, c->id, c->stage_r, c->stage_v,
                c->stage_rho, c->stage_p);
  tit_status st = check_launch(c, "download");
  if (st != TIT_OK) return st;
  if (r) TIT_HIP_TRY(c, hipMemcpyAsync(r, c->stage_r, n * dim * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (v) TIT_HIP_TRY(c, hipMemcpyAsync(v, c->stage_v, n * dim * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (rho) TIT_HIP_TRY(c, hipMemcpyAsync(rho, c->stage_rho, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (p) TIT_HIP_TRY(c, hipMemcpyAsync(p, c->stage_p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return TIT_OK;
}

tit_status tit_hip_tank_clock(tit_hip_ctx* c, double* time, double* dt_next) {
  if (!c) return TIT_E_ARG;
  if (!c->uploaded) return fail(c, TIT_E_STATE, "no state");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (time) {
    TIT_HIP_TRY(c, hipMemcpy(time, c->time, c->n_tanks * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (dt_next) {
    std::vector<unsigned long long> b(static_cast<size_t>(c->n_tanks) * kRedStripes);
    TIT_HIP_TRY(c, hipMemcpy(b.data(), c->dt_min, b.size() * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost));
    for (int t = 0; t < c->n_tanks; ++t) {
      // the smallest of the tank's copies
      unsigned long long m = ~0ull;
      for (int k = 0; k < kRedStripes; ++k) m = std::min(m, b[static_cast<size_t>(t) * kRedStripes + k]);
      double x = 0.0;
      if (m != ~0ull) std::memcpy(&x, &m, sizeof(double));
      dt_next[t] = x;
    }
  }
  return TIT_OK;
}

tit_status tit_hip_compute_time_step(tit_hip_ctx* c, double* dt) {
  return tit_hip_tank_clock(c, nullptr, dt);
}

tit_status tit_hip_prepare(tit_hip_ctx* c) {
  if (!c) return TIT_E_ARG;
  if (!c->uploaded) return fail(c, TIT_E_STATE, "prepare before upload");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  enqueue_rebuild(c, kRebuildPending);
  enqueue_boundary(c, c->A);
  tit_status st = check_launch(c, "prepare");
  if (st != TIT_OK) return st;
  if (c->prof.enabled) c->prof.collect();
  return read_flags(c);
}

tit_status tit_hip_post_integrate(tit_hip_ctx* c, double* phi) {
  if (!c) return TIT_E_ARG;
  if (!c->uploaded) return fail(c, TIT_E_STATE, "post_integrate before upload");
  if (!c->prm.post_integrate) {
    return fail(c, TIT_E_STATE, "context created without post_integrate");
  }
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  enqueue_rebuild(c, kRebuildPending);
  enqueue_boundary(c, c->A, true);
  enqueue_post_integrate(c);
  const size_t n = c->n, dim = static_cast<size_t>(c->prm.dim);
  if (phi) {  // phi_near follows N, dr0, grad v, grad rho, phi in pi_buf
    launch_export_f64(c->stream, c->n, c->id, c->pi_buf + (3 * dim + dim * dim + 1) * n,
                      c->stage_rho);
  }
  tit_status st = check_launch(c, "post_integrate");
  if (st != TIT_OK) return st;
  if (phi) {
    TIT_HIP_TRY(c, hipMemcpyAsync(phi, c->stage_rho, n * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
  }
  if (c->prof.enabled) c->prof.collect();
  return read_flags(c);
}

tit_status tit_hip_compute_rhs(tit_hip_ctx* c, double* drho_dt, double* dv_dt) {
  if (!c) return TIT_E_ARG;
  if (!c->uploaded) return fail(c, TIT_E_STATE, "compute_rhs before upload");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  enqueue_rebuild(c, kRebuildPending);
  if (c->walls) enqueue_boundary(c, c->A);  // gamma and face terms of this state
  enqueue_stage(c, c->A, c->A, c->A, kRhsOnly, 1.0, false);
  // reuse the staging buffers: drho -> stage_rho, dv -> stage_v
  launch_export_rhs(c->stream, c->prm.dim, c->n, c->id, c->rhs_drho, c->rhs_dv,
                    c->stage_rho, c->stage_v);
  tit_status st = check_launch(c, "compute_rhs");
  if (st != TIT_OK) return st;
  const size_t n = c->n;
  if (drho_dt) TIT_HIP_TRY(c, hipMemcpyAsync(drho_dt, c->stage_rho, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (dv_dt) TIT_HIP_TRY(c, hipMemcpyAsync(dv_dt, c->stage_v, n * c->prm.dim * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (c->prof.enabled) c->prof.collect();
  return read_flags(c);
}

tit_status tit_hip_pair_counts(tit_hip_ctx* c, uint32_t* count) {
  if (!c || !count) return TIT_E_ARG;
  if (!c->uploaded) return fail(c, TIT_E_STATE, "pair_counts before upload");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  enqueue_rebuild(c, kRebuildPending);
  if (c->walls) enqueue_boundary(c, c->A);
  enqueue_stage(c, c->A, c->A, c->A, kRhsOnly, 1.0, false);
  uint32_t* out = reinterpret_cast<uint32_t*>(c->stage_rho);  // n doubles >= n u32
  launch_export_u32(c->stream, c->n, c->id, c->rhs_count, out);
  tit_status st = check_launch(c, "pair_counts");
  if (st != TIT_OK) return st;
  TIT_HIP_TRY(c, hipMemcpyAsync(count, out, c->n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                c->stream));
  TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return read_flags(c);
}

tit_status tit_hip_neighbors(tit_hip_ctx* c, uint64_t* off, uint32_t* idx,
                             size_t cap, size_t* total) {
  if (!c || !off) return TIT_E_ARG;
  if (!c->uploaded) return fail(c, TIT_E_STATE, "neighbors before upload");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  enqueue_rebuild(c, kRebuildPending);
  const size_t n = c->n;
  uint64_t* d_cnt = nullptr;
  TIT_HIP_TRY(c, hipMalloc(&d_cnt, (n + 1) * sizeof(uint64_t)));
  launch_neighbor_count(c->stream, c->dp, c->tanks, c->n, c->A, c->key, c->tag,
                        c->id, c->cell_off, d_cnt);
  std::vector<uint64_t> cnt(n + 1, 0);
  hipError_t e = hipMemcpyAsync(cnt.data(), d_cnt, n * sizeof(uint64_t),
                                hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    (void)hipFree(d_cnt);
    return fail(c, TIT_E_HIP, hipGetErrorString(e));
  }
  off[0] = 0;
  for (size_t a = 0; a < n; ++a) off[a + 1] = off[a] + cnt[a];
  if (total) *total = off[n];
  if (off[n] > cap || !idx) {
    (void)hipFree(d_cnt);
    return fail(c, TIT_E_ARG, "neighbour capacity too small");
  }
  uint32_t* d_idx = nullptr;
  e = hipMemcpyAsync(d_cnt, off, (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice,
                     c->stream);
  if (e == hipSuccess) e = hipMalloc(&d_idx, std::max<uint64_t>(off[n], 1) * sizeof(uint32_t));
  if (e == hipSuccess) {
    launch_neighbor_fill(c->stream, c->dp, c->tanks, c->n, c->A, c->key, c->tag,
                         c->id, c->cell_off, d_cnt, d_idx);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    e = hipMemcpyAsync(idx, d_idx, off[n] * sizeof(uint32_t), hipMemcpyDeviceToHost,
                       c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_cnt);
  if (d_idx) (void)hipFree(d_idx);
  if (e != hipSuccess) return fail(c, TIT_E_HIP, hipGetErrorString(e));
  return read_flags(c);
}

tit_status tit_hip_neighbors_noalloc(tit_hip_ctx* c, uint64_t* off, uint32_t* idx,
                                     size_t cap, size_t* total) {
  if (!c || !off) return TIT_E_ARG;
  if (!c->uploaded) return fail(c, TIT_E_STATE, "neighbors before upload");
  const size_t n = c->n;
  if (n > kNoAllocParticles) {
    return fail(c, TIT_E_ARG, "tit_hip_neighbors_noalloc holds at most 1048576 particles");
  }
  NeighborScratch* sp = scratch_ptr(c->device);
  if (!sp) return fail(c, TIT_E_HIP, "no scratch address on this device");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  enqueue_rebuild(c, kRebuildPending);
  // counts by reference index into sp->off[0, n) (every entry written), then
  // the offsets on the host, in off itself
  launch_neighbor_count(c->stream, c->dp, c->tanks, c->n, c->A, c->key, c->tag, c->id,
                        c->cell_off, sp->off);
  hipError_t e = hipMemcpyAsync(off + 1, sp->off, n * sizeof(uint64_t),
                                hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, TIT_E_HIP, hipGetErrorString(e));
  off[0] = 0;
  for (size_t a = 0; a < n; ++a) off[a + 1] += off[a];
  if (total) *total = off[n];
  if (off[n] > kNoAllocEntries) {
    return fail(c, TIT_E_ARG, "tit_hip_neighbors_noalloc holds at most 16777216 list entries");
  }
  if (off[n] > cap || !idx) return fail(c, TIT_E_ARG, "neighbour capacity too small");
  e = hipMemcpyAsync(sp->off, off, (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice,
                     c->stream);
  if (e == hipSuccess) {
    launch_neighbor_fill(c->stream, c->dp, c->tanks, c->n, c->A, c->key, c->tag, c->id,
                         c->cell_off, sp->off, sp->idx);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    e = hipMemcpyAsync(idx, sp->idx, off[n] * sizeof(uint32_t), hipMemcpyDeviceToHost,
                       c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, TIT_E_HIP, hipGetErrorString(e));
  return read_flags(c);
}

tit_status tit_hip_set_profiling(tit_hip_ctx* c, int enable) {
  if (!c) return TIT_E_ARG;
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->prof.clear();
  c->prof.enabled = enable != 0;
  return TIT_OK;
}

tit_status tit_hip_kernel_stats(tit_hip_ctx* c, tit_hip_kernel_stat* out, int cap,
                                int* n_out) {
  if (!c) return TIT_E_ARG;
  c->prof.collect();
  const int n = static_cast<int>(c->prof.totals.size());
  if (n_out) *n_out = n;
  for (int k = 0; k < n && k < cap && out; ++k) {
    std::memset(out[k].name, 0, sizeof(out[k].name));
    std::strncpy(out[k].name, c->prof.totals[k].name.c_str(), sizeof(out[k].name) - 1);
    out[k].launches = c->prof.totals[k].launches;
    out[k].total_ms = c->prof.totals[k].total_ms;
  }
  return TIT_OK;
}

tit_status tit_hip_get_schedule(tit_hip_ctx* c, tit_hip_schedule* out) {
  if (!c || !out) return TIT_E_ARG;
  out->chunk_targets = static_cast<int>(c->chunk_targets);
  out->lds_capacity = static_cast<int>(c->lds_cap);
  out->occupancy = c->occupancy;
  return TIT_OK;
}

tit_status tit_hip_set_schedule(tit_hip_ctx* c, const tit_hip_schedule* sch) {
  if (!c || !sch) return TIT_E_ARG;
  if (!c->uploaded) return fail(c, TIT_E_STATE, "schedule before upload");
  if (sch->chunk_targets < 0 || sch->chunk_targets > kChunkMax || sch->lds_capacity < 0 ||
      sch->lds_capacity > static_cast<int>(stage_lds_capacity(c->prm.dim))) {
    return fail(c, TIT_E_ARG, "chunk_targets must be in [0, 128] and lds_capacity "
                              "within the kernel's LDS capacity");
  }
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (sch->chunk_targets > 0) c->chunk_targets = static_cast<uint32_t>(sch->chunk_targets);
  c->lds_cap = static_cast<uint32_t>(sch->lds_capacity);
  return build_schedule(c);
}

tit_status tit_hip_autotune(tit_hip_ctx* c, int steps_per_candidate,
                            tit_hip_schedule* best, double* best_ms_per_step) {
  if (!c) return TIT_E_ARG;
  if (!c->uploaded) return fail(c, TIT_E_STATE, "autotune before upload");
  if (steps_per_candidate < 2) return fail(c, TIT_E_ARG, "need at least 2 steps per candidate");
  if (c->prof.enabled) return fail(c, TIT_E_STATE, "autotune with profiling enabled");
  TIT_HIP_TRY(c, hipSetDevice(c->device));
  const TuneKey key = tune_key(c);
  TuneVal hit{};
  bool cached = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      hit = it->second;
      cached = true;
    }
  }
  if (cached) {
    TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->chunk_targets = hit.chunk_targets;
    c->lds_cap = hit.lds_cap;
    const tit_status st = build_schedule(c);
    if (st != TIT_OK) return st;
    if (best) {
      best->chunk_targets = static_cast<int>(hit.chunk_targets);
      best->lds_capacity = static_cast<int>(hit.lds_cap);
      best->occupancy = c->occupancy;
      best->cached = 1;
    }
    if (best_ms_per_step) *best_ms_per_step = hit.ms_per_step;
    return TIT_OK;
  }
  const int timed = steps_per_candidate / 2 * 2;
  const uint32_t lds_max = stage_lds_capacity(c->prm.dim);
  std::vector<uint32_t> chunks = {32, 48, 64, 96, 128};
  std::vector<uint32_t> caps = {lds_max};
  if (lds_max > 0) {
    caps.push_back(lds_max / 2);
    caps.push_back(0);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  TIT_HIP_TRY(c, hipEventCreate(&e0));
  TIT_HIP_TRY(c, hipEventCreate(&e1));
  double best_ms = 1e300;
  uint32_t best_chunk = c->chunk_targets, best_cap = c->lds_cap;
  tit_status st = TIT_OK;
  for (uint32_t ch : chunks) {
    for (uint32_t cap : caps) {
      c->chunk_targets = ch;
      c->lds_cap = cap;
      st = build_schedule(c);
      if (st == TIT_OK) st = run_plain(c, 2);  // captures this parity's graph
      if (st != TIT_OK) break;
      hipError_t e = hipEventRecord(e0, c->stream);
      if (e == hipSuccess) st = run_plain(c, timed);
      if (e == hipSuccess && st == TIT_OK) e = hipEventRecord(e1, c->stream);
      if (e == hipSuccess && st == TIT_OK) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess && st == TIT_OK) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) st = fail(c, TIT_E_HIP, std::string("autotune: ") + hipGetErrorString(e));
      if (st != TIT_OK) break;
      const double per = ms / timed;
      if (per < best_ms) {
        best_ms = per;
        best_chunk = ch;
        best_cap = cap;
      }
    }
    if (st != TIT_OK) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != TIT_OK) return st;
  c->chunk_targets = best_chunk;
  c->lds_cap = best_cap;
  st = build_schedule(c);
  if (st != TIT_OK) return st;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = TuneVal{best_chunk, best_cap, best_ms};
  }
  if (best) {
    best->chunk_targets = static_cast<int>(best_chunk);
    best->lds_capacity = static_cast<int>(best_cap);
    best->occupancy = c->occupancy;
    best->cached = 0;
  }
  if (best_ms_per_step) *best_ms_per_step = best_ms;
  return TIT_OK;
}

void tit_hip_autotune_clear_cache(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
}

tit_status tit_hip_counts(tit_hip_ctx* c, size_t* n_particles, uint64_t* n_rebuilds,
                          uint64_t* n_steps) {
  if (!c) return TIT_E_ARG;
  if (n_particles) *n_particles = c->n;
  if (n_steps) *n_steps = c->steps;
  if (n_rebuilds) {
    *n_rebuilds = 0;
    if (c->uploaded) {
      unsigned long long r = 0;
      TIT_HIP_TRY(c, hipSetDevice(c->device));
      TIT_HIP_TRY(c, hipStreamSynchronize(c->stream));
      TIT_HIP_TRY(c, hipMemcpy(&r, c->rebuild_count, sizeof(r), hipMemcpyDeviceToHost));
      *n_rebuilds = r;
    }
  }
  return TIT_OK;
}

}  // extern "C"
