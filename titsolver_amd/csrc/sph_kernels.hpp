// Launch interfaces of the WCSPH kernels (sph_kernels.hip).
#pragma once

#include "internal.hpp"

namespace tit_amd {

// Largest chunk (targets per workgroup) and LDS position capacity of the
// 2D stage kernel.
constexpr int kChunkMax = 128;
constexpr int kStageLdsCap = 1024;
// Cells a chunk's targets span along a column at most: a column's equal
// parts are cut further where they span more (a column whose particles are
// split by empty cells, such as fluid below and a wall's vertices at the
// top), so that every chunk's window fits the LDS kernel's cell rows.
constexpr int kChunkRows = 92;
// Window columns of a 2D chunk and LDS cell-offset rows of its window.
constexpr int kCols2 = 2 * kStencil + 1;
constexpr int kRowsMax = kChunkRows + 2 * kStencil;
// 3D block chunks: the cell columns of a tank in blocks of kBlk3 x kBlk3;
// a chunk holds the particles of one block in a range of at most
// kChunkRows3 rows, at most kStage3Threads of them (a row with more goes to
// the global-memory kernel whole).  Its window is the block grown by the
// stencil (kWin3 x kWin3 columns, kRows3 rows), FP64 positions of at most
// kStage3LdsCap candidates staged in LDS.
constexpr int kBlk3 = 4;
constexpr int kWin3 = kBlk3 + 2 * kStencil;
constexpr int kChunkRows3 = 6;
constexpr int kRows3 = kChunkRows3 + 2 * kStencil;
constexpr int kWords3 = 6;
constexpr int kStage3Threads = 256;
// (a chunk of 256 targets in the 3D dam break spans 5-6 rows of 16 columns,
// so its window holds about 64 x 9.5 x 2.9 = 1,800 candidates, and one of 6
// rows at most 64 x 10 x 2.9 = 1,860; 2,048 candidates keep the workgroup's
// LDS under 53 KB: three workgroups of four waves per CU)
constexpr int kStage3LdsCap = 2048;

struct StageArgs {
  DevParams P;
  const DevTank* tanks;
  DevState in;       // state the right-hand side is evaluated on
  DevState old;      // step-start state (combination stages)
  DevState out;      // state written by the stage
  const uint32_t* key;
  const uint32_t* tag;       // tank << 1 | fixed
  const uint32_t* id;        // particle ids (per-particle masses are by id)
  const uint32_t* cell_off;  // CSR over all cells of all tanks
  const double* dt;          // per tank
  double w;                  // weight of the new state, (1-w) old + w new
  uint32_t n;
  // outputs of the right-hand-side-only form (sorted order)
  double* rhs_drho;
  double* rhs_dv[3];
  uint32_t* rhs_count;       // accepted neighbours per target (may be null)
  // reductions (kRedStripes copies per tank, device_common.hpp)
  unsigned long long* dt_min;  // per tank, FP64 bits
  uint32_t* vmax;              // per tank, FP32 bits
  // chunk schedule: cell columns (tank, ix, iy, first cell) and, rebuilt on
  // the device after every cell rebuild, chunks (column, first particle,
  // end particle) of at most chunk_targets consecutive particles of a column
  // spanning at most kChunkRows cells
  const uint4* cols;
  uint32_t n_cols;
  uint32_t chunk_targets;
  uint32_t lds_cap;            // candidate positions staged in LDS at most
  uint4* chunks;
  uint32_t* overflow;          // 2D chunks whose candidates exceed lds_cap
  // [0] chunks, [1] overflow chunks, [2] 3D block chunks, [3] block chunks
  // that take the global-memory kernel
  uint32_t* counts;
  uint32_t grid_lds, grid_global;  // persistent grid sizes
  // 3D with LDS staging (lds_cap > 0): cell-column blocks (tank, first ix,
  // first iy, 0), rebuilt on the device after every cell rebuild: block
  // chunks (block, first row, end row, targets), per-block chunk counts
  // (scanned in place), and the block chunks that do not fit
  const uint4* blocks;
  uint32_t n_blocks;
  uint4* bchunks;
  uint32_t* boverflow;
  uint32_t* bcount;
  uint32_t grid_lds3, grid_block_global;
  double* dv_last[3];          // final stage: its acceleration (post-integration only)
  // semi-analytic wall (kFlavourWall): per slot, the face terms of drho/dt
  // and dv/dt already divided by gamma, then gamma
  const double* wall;
};

// Post-integration passes (post_integrate.hip): state a (post-step, its
// positions give the neighbour sets), b (shifted), c (corrected density),
// FP64 scratch per particle, SoA.
struct PiArgs {
  DevParams P;
  const DevTank* tanks;
  uint32_t n;
  DevState a, b, c;
  const uint32_t* key;
  const uint32_t* tag;
  const uint32_t* id;
  const uint32_t* cell_off;
  double* N[3];          // renormalised, normalised N
  double* dr0[3];        // raw N (the unscaled shift)
  double* grad_v[9];     // renormalised grad v, [p * dim + q] = dv_p / dx_q
  double* grad_rho[3];   // renormalised grad rho
  uint32_t* count;       // neighbours, self included
  double* phi;           // free-surface indicator after the cone test
  double* phi_near;      // after the near-surface classification
  uint32_t* cell_mark;   // per cell: holds a surface particle
  uint32_t* cell_near;   // per cell: its 5^d window holds a surface particle
                         // (cell_mark + n_cells + 1: the two are cleared together)
  double* vol;           // m / rho per particle
  // 2D: (x, y), (v_x, v_y), (rho, m / rho) per particle, packed by the volume
  // pass so that a neighbour's fields come in three 16-byte loads
  double2* pk_x;
  double2* pk_v;
  double2* pk_rv;
  uint32_t n_cells;
  unsigned long long* shift_max;  // per tank, FP64 bits (kRedStripes copies per tank)
  double* dv_last[3];    // the last stage's acceleration
  unsigned long long* dt_min;
  uint32_t* vmax;
  // semi-analytic wall (null without): gamma at [4 slot + 3], the fixed
  // particles' mass fractions by id - wall_id0, the face terms of apply_shifts
  const double* wall;
  const double* wall_mu;
  uint32_t wall_id0;
  const uint32_t* near_idx;
  const double* pi_face;
};

struct BoundaryArgs {
  DevParams P;
  const DevTank* tanks;
  DevState s;
  const uint32_t* key;
  const uint32_t* tag;
  const uint32_t* id;
  const uint32_t* cell_off;
  const uint32_t* fixed_list;
  // the fixed particles with fluid in their stencil at the last rebuild
  // (k_fixed_active lists them; the others keep the state it wrote) and
  // their count
  uint32_t* active;
  uint32_t* n_active;
  const double* normal;  // reference order, dim per particle (fixed only)
  const uint8_t* cell_fluid;  // per cell: holds a fluid particle (last rebuild)
  uint32_t n_fixed;
  // semi-analytic wall: fixed particle id0 + k sits at wall vertex k; its
  // mass fraction gamma_k (mass m gamma_k) and its density, written here for
  // the face terms
  const double* wall_mu;
  double* wall_rho;
  uint32_t wall_id0;
};

// Device-side control of the cell rebuild.  ctl[0]: the rebuild of this step
// runs (some tank needs it); ctl[1]: a rebuild of every tank is pending
// (after upload); ctl[2]: the cell index is valid.  need[t]: tank t is
// re-sorted (other tanks keep their keys, hence their order).  count: rebuilds
// run so far.  Every rebuild kernel reads ctl[0] and returns at once when it
// is clear, so a captured step holds the rebuild unconditionally.
struct RebuildCtl {
  uint32_t* ctl;
  uint32_t* need;
  unsigned long long* count;
};

// Kernel kinds of the stage launch.
enum StageKind { kRhsOnly = 0, kEuler = 1, kCombine = 2 };

// Reference-order arrays into the state (also its pair records).
void launch_import(hipStream_t s, const DevParams& P, uint32_t n, const double* r,
                   const double* v, const double* rho, DevState st,
                   uint32_t* id);
void launch_cell_keys(hipStream_t s, const DevParams& P, const DevTank* tanks,
                      uint32_t n, DevState st, const uint32_t* tag,
                      uint32_t* key, uint32_t* flags, RebuildCtl rc);
void launch_permute(hipStream_t s, const DevParams& P, uint32_t n, const uint32_t* perm,
                    DevState src, DevState dst, const uint32_t* id_src,
                    uint32_t* id_dst, const uint32_t* tag_src,
                    uint32_t* tag_dst, uint32_t* fixed_list, uint32_t* fixed_count,
                    const uint32_t* ctl);
// Copies the permuted state (src) back into the working buffers (dst) for
// the slots whose particle moved (perm[i] != i; the others were not written).
void launch_restore_order(hipStream_t s, int dim, uint32_t n, const uint32_t* perm,
                          DevState src, DevState dst, const uint32_t* id_src,
                          uint32_t* id_dst, const uint32_t* tag_src, uint32_t* tag_dst,
                          const uint32_t* ctl);
// Cell offsets, and per cell whether it holds a fluid particle; gaps
// (n + 1 entries) and n_gaps are scratch for the long runs of empty cells.
void launch_cell_offsets(hipStream_t s, uint32_t n_cells, uint32_t n,
                         const uint32_t* key, const uint32_t* tag, uint32_t* cell_off,
                         uint8_t* fluid, uint4* gaps, uint32_t* n_gaps, const uint32_t* ctl);
// After a rebuild (ctl[0]): the fixed particles with fluid in their stencil
// into a.active / a.n_active, the others' state written into a.s.
void launch_fixed_active(hipStream_t s, const BoundaryArgs& a, const uint32_t* ctl);
// Extrapolation of the listed fixed particles into a.s.
void launch_boundary(hipStream_t s, const BoundaryArgs& a);

// Semi-analytic wall (wall.hip): per fluid particle of the stage input state,
// gamma, grad gamma and the face terms, into out (4 per slot: face drho/dt and
// dv/dt already divided by gamma, then gamma).  With pi_face set (before
// post-integration), also the face terms of apply_shifts of each particle
// near a wall (2D: N, L, grad v, grad rho = 12 values, already divided by
// gamma) at its index in the near list, near_idx[slot] (~0: none).
struct WallFace;
struct WallArgs {
  DevParams P;
  uint32_t n;
  DevState s;
  const uint32_t* tag;
  const WallFace* faces;
  const uint32_t* cell_faces_off;  // CSR of ascending face indices per grid cell
  const uint32_t* cell_faces;
  const uint32_t* cell_inside;     // cells without faces: wholly contained
  double lo[2], inv_cw[2];
  int nc[2];
  const double* cedges;            // containment polygon, 4 per edge
  uint32_t n_cedges;
  const double* vert_rho;          // density of the fixed particle at each vertex
  double* out;
  uint32_t* near;                  // scratch: slots of the particles near a wall
  uint32_t* near_count;
  uint32_t* near_idx;              // post-integration: per slot, index in near (~0: none)
  double* pi_face;                 // post-integration: 12 per near-list index
};
void launch_wall(hipStream_t s, const WallArgs& a, int n_cu);
// post_integrate.hip: the passes in order, the shift into disp[] (a rebuild
// of every tank marked pending in ctl[1] when it passes skin_half), the next
// step size into dt_min.
void launch_post_integrate(hipStream_t s, const PiArgs& a, int n_tanks, double* disp,
                           double skin_half, uint32_t* ctl, int n_cu);
// dst[id[i]] = src[i] (FP64).
void launch_export_f64(hipStream_t s, uint32_t n, const uint32_t* id, const double* src,
                       double* dst);
void launch_stage(hipStream_t s, const StageArgs& a, int kind, bool final);
uint32_t stage_lds_capacity(int dim);
// Persistent grid sizes of the stage kernels (multiples of 8 workgroups).
struct StageGrids {
  uint32_t lds, global, lds3, block_global;
};
void stage_grids(const DevParams& P, int n_cu, StageGrids* g);
// Rebuilds the chunk list from the cell offsets (after a cell rebuild).
void launch_build_chunks(hipStream_t s, const StageArgs& a, uint32_t* block_tot,
                         const uint32_t* ctl);
void launch_time_step(hipStream_t s, const DevParams& P, uint32_t n,
                      DevState st, const uint32_t* tag, const uint32_t* id,
                      const double* dv_ref, unsigned long long* dt_min);
// The per-tank dt_min / vmax / shift_max buffers hold kRedStripes copies per
// tank (n_tanks x kRedStripes values); the launchers below fold them.
// Takes each tank's pending step size dt and decides its rebuild for a step
// of span x dt: needed when forced or when disp + 2 span dt vmax (vmax of the
// previous step) could pass 0.9 of half the skin by the end of the step.
void launch_step_begin(hipStream_t s, int n_tanks, unsigned long long* dt_min,
                       double* dt, uint32_t* vmax, const double* disp,
                       double skin_half, RebuildCtl rc, double span = 1.0);
// Marks every tank for the next rebuild sequence.
void launch_force_rebuild(hipStream_t s, int n_tanks, RebuildCtl rc);
// Makes a forced rebuild pending (after moves the displacement bound does not
// cover, such as a learned step).
void launch_mark_pending(hipStream_t s, RebuildCtl rc);
// Marks every tank if a forced rebuild is pending, none otherwise.
void launch_pending_rebuild(hipStream_t s, int n_tanks, RebuildCtl rc);
// time += span dt; disp += span dt vmax, flagging a skin overrun.
void launch_step_end(hipStream_t s, int n_tanks, const double* dt,
                     uint32_t* vmax, double* time, double* disp,
                     double skin_half, uint32_t* flags, double span = 1.0);
// Ends a rebuild sequence: zero displacement of the re-sorted tanks, count.
void launch_reset_disp(hipStream_t s, int n_tanks, double* disp, RebuildCtl rc);
void launch_export(hipStream_t s, const DevParams& P, uint32_t n, DevState st,
                   const uint32_t* id, double* r, double* v, double* rho,
                   double* p);
// dst[id[i]] = src[i].
void launch_export_u32(hipStream_t s, uint32_t n, const uint32_t* id,
                       const uint32_t* src, uint32_t* dst);
void launch_export_rhs(hipStream_t s, int dim, uint32_t n, const uint32_t* id,
                       const double* drho, double* const* dv, double* drho_ref,
                       double* dv_ref);
void launch_neighbor_count(hipStream_t s, const DevParams& P,
                           const DevTank* tanks, uint32_t n, DevState st,
                           const uint32_t* key, const uint32_t* tag,
                           const uint32_t* id, const uint32_t* cell_off,
                           uint64_t* count_ref);
void launch_neighbor_fill(hipStream_t s, const DevParams& P,
                          const DevTank* tanks, uint32_t n, DevState st,
                          const uint32_t* key, const uint32_t* tag,
                          const uint32_t* id, const uint32_t* cell_off,
                          const uint64_t* off_ref, uint32_t* idx_ref);

}  // namespace tit_amd
