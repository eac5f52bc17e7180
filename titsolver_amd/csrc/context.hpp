// Context of the C ABI (tit_hip_ctx) and the step pieces shared by the
// translation units that implement the ABI.
#pragma once

#include <string>
#include <vector>

#include "../../include/tit_hip.h"
#include "sph_kernels.hpp"
#include "wall.hpp"

namespace tit_amd {
struct Learned;
}

using tit_amd::DevParams;
using tit_amd::DevState;
using tit_amd::DevTank;
using tit_amd::Profiler;
using tit_amd::SortScratch;
using tit_amd::kChunkMax;

struct tit_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  tit_hip_params prm{};
  DevParams dp{};
  std::string err;
  Profiler prof;

  // sizes
  uint32_t n = 0;
  uint32_t n_fixed = 0;
  int n_tanks = 0;
  uint32_t n_cells = 0;
  int key_bits = 1;
  bool uploaded = false;
  uint64_t steps = 0;  // plain steps taken since upload

  // device memory
  std::vector<void*> allocs;
  DevState A{}, T{}, B{}, C{};
  uint32_t *key = nullptr, *key_alt = nullptr, *val = nullptr, *val_alt = nullptr;
  uint32_t *id = nullptr, *id_alt = nullptr, *tag = nullptr, *tag_alt = nullptr;
  uint32_t* cell_off = nullptr;
  uint32_t* fixed_list = nullptr;
  uint32_t* fixed_active = nullptr;  // fixed particles with fluid in their stencil
  uint32_t* counter = nullptr;  // [0] fixed count, [1] flags, [2] fixed_active count
  uint32_t* rebuild_ctl = nullptr;       // RebuildCtl::ctl
  uint32_t* tank_need = nullptr;         // RebuildCtl::need
  unsigned long long* rebuild_count = nullptr;
  SortScratch sort;
  DevTank* tanks = nullptr;
  double* normal = nullptr;
  uint8_t* cell_fluid = nullptr;  // per cell: holds fluid (rebuilt with the index)
  uint4* cell_gaps = nullptr;     // scratch of the cell offsets: long empty runs
  uint32_t* cell_n_gaps = nullptr;
  double *dt = nullptr, *time = nullptr, *disp = nullptr;
  unsigned long long* dt_min = nullptr;
  uint32_t* vmax = nullptr;
  double *stage_r = nullptr, *stage_v = nullptr, *stage_rho = nullptr,
         *stage_p = nullptr;
  double* rhs_drho = nullptr;
  uint32_t* rhs_count = nullptr;  // pair counts of the last right-hand side
  double* rhs_dv[3] = {nullptr, nullptr, nullptr};
  // post-integration scratch (allocated when params.post_integrate is set);
  // rhs_dv holds the last stage's acceleration between steps then
  double* pi_buf = nullptr;  // N, dr0, grad v, grad rho, phi, phi_near (SoA)
  double2* pi_pack = nullptr;  // 2D: packed x, v, (rho, vol), n each
  uint32_t* pi_count = nullptr;
  uint32_t* pi_cells = nullptr;  // per cell: surface mark, then window mark
  unsigned long long* pi_shift = nullptr;  // per tank
  // semi-analytic wall (tit_hip_set_walls; one 2D tank, REF terms)
  bool walls = false;
  std::vector<double> wall_verts, wall_cverts;
  std::vector<uint32_t> wall_faces, wall_cfaces;
  tit_amd::WallFace* wall_face_d = nullptr;
  uint32_t* wall_cell_off = nullptr;
  uint32_t* wall_cell_faces = nullptr;
  uint32_t* wall_cell_inside = nullptr;
  double* wall_cedges = nullptr;
  uint32_t wall_n_cedges = 0;
  double wall_lo[2] = {0.0, 0.0}, wall_inv_cw[2] = {0.0, 0.0};
  int wall_nc[2] = {0, 0};
  double* wall_mu = nullptr;   // per vertex: gamma at set-up (mass fraction)
  double* mass = nullptr;      // per particle by id: uploaded masses (null: uniform m)
  double* wall_rho = nullptr;  // per vertex: density of this stage
  double* wall_out = nullptr;  // per slot: face terms / gamma, gamma
  uint32_t* wall_near = nullptr;  // per slot: list of particles near a wall
  uint32_t* wall_near_count = nullptr;
  uint32_t* wall_near_idx = nullptr;  // post-integration: per slot, near-list index
  double* wall_pi_face = nullptr;     // post-integration: 12 face terms per near particle
  uint32_t wall_id0 = 0;       // id of the fixed particle at vertex 0

  // chunk schedule of the stage kernel
  std::vector<DevTank> host_tanks;
  uint4* cols = nullptr;
  uint32_t n_cols = 0;
  uint4* chunks = nullptr;
  uint32_t* overflow = nullptr;
  uint32_t* block_tot = nullptr;
  uint32_t* chunk_counts = nullptr;
  size_t chunk_cap = 0;
  uint32_t chunk_targets = kChunkMax;
  uint32_t lds_cap = 0;
  uint32_t grid_lds = 0, grid_global = 0;
  // 3D block chunks of the LDS stage kernel
  uint4* blocks = nullptr;
  uint32_t n_blocks = 0;
  uint4* bchunks = nullptr;
  uint32_t* boverflow = nullptr;
  uint32_t* bcount = nullptr;
  size_t bchunk_cap = 0;
  uint32_t grid_lds3 = 0, grid_block_global = 0;
  int n_cu = 0;
  double occupancy = 1.0;  // mean particles per occupied cell at upload

  // graphs of two steps, one per rebuild parity (every rebuild swaps the
  // state, key and id/tag pointers)
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  hipGraph_t graph_def[2] = {nullptr, nullptr};

  // learned step (surrogate.hip), created on first use
  tit_amd::Learned* learned = nullptr;
};

namespace tit_amd {

tit_status fail(tit_hip_ctx* c, tit_status s, const std::string& msg);

#define TIT_HIP_TRY(ctx, call)                                              \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      return fail((ctx), e_ == hipErrorOutOfMemory ? TIT_E_OOM : TIT_E_HIP, \
                  std::string(#call) + ": " + hipGetErrorString(e_));       \
    }                                                                       \
  } while (0)

template <class T>
hipError_t dalloc(tit_hip_ctx* c, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e == hipSuccess) c->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return e;
}

void forget(tit_hip_ctx* c, void* q);
void drop_graph(tit_hip_ctx* c);
RebuildCtl rebuild_ctl(tit_hip_ctx* c);
enum RebuildMode {
  kRebuildMarked,   // the tanks the last step_begin marked (nothing if none)
  kRebuildPending,  // every tank if a forced rebuild is pending, else none
  kRebuildForce     // every tank
};
void enqueue_rebuild(tit_hip_ctx* c, RebuildMode mode);
// setup_boundary (and with walls compute_gamma and the face terms; before
// post-integration also apply_shifts' face terms).
void enqueue_boundary(tit_hip_ctx* c, const DevState& st, bool for_post_integrate = false);
// The boundary extrapolation's arguments on state st.
BoundaryArgs boundary_args(tit_hip_ctx* c, const DevState& st);
void enqueue_step(tit_hip_ctx* c);
void enqueue_post_integrate(tit_hip_ctx* c);
tit_status check_launch(tit_hip_ctx* c, const char* what);
tit_status read_flags(tit_hip_ctx* c);
// n_steps plain SSPRK steps (graph replay where possible); synchronises.
tit_status run_plain(tit_hip_ctx* c, int n_steps);
hipError_t alloc_state(tit_hip_ctx* c, DevState& s, uint32_t n, int dim);
// Frees the learned-step state of a context (surrogate.hip).
void destroy_learned(tit_hip_ctx* c);

}  // namespace tit_amd
