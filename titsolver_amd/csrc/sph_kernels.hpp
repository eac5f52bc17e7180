// Launch interfaces of the WCSPH kernels (sph_kernels.hip).
#pragma once

#include "internal.hpp"

namespace tit_amd {

// Longest strip (cells along the fast axis) and LDS candidate capacity of
// the 2D strip kernel.
constexpr int kStripMax = 64;
constexpr int kStripLdsCap = 768;

struct StageArgs {
  DevParams P;
  const DevTank* tanks;
  DevState in;       // state the right-hand side is evaluated on
  DevState old;      // step-start state (combination stages)
  DevState out;      // state written by the stage
  const uint32_t* key;
  const uint32_t* tag;       // tank << 1 | fixed
  const uint32_t* cell_off;  // CSR over all cells of all tanks
  const double* dt;          // per tank
  double w;                  // weight of the new state, (1-w) old + w new
  uint32_t n;
  // outputs of the right-hand-side-only form (sorted order)
  double* rhs_drho;
  double* rhs_dv[3];
  // reductions
  unsigned long long* dt_min;  // per tank, FP64 bits
  uint32_t* vmax;              // per tank, FP32 bits
  // strip schedule: work items (tank, column x, column y, first cell z)
  const uint4* work;
  uint32_t n_work;
  uint32_t strip_len;          // cells per strip along the fast axis
  uint32_t lds_cap;            // candidates staged in LDS at most
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
  const double* normal;  // reference order, dim per particle (fixed only)
  uint32_t n_fixed;
};

// Kernel kinds of the stage launch.
enum StageKind { kRhsOnly = 0, kEuler = 1, kCombine = 2 };

void launch_import(hipStream_t s, int dim, uint32_t n, const double* r,
                   const double* v, const double* rho, DevState st,
                   uint32_t* id);
void launch_cell_keys(hipStream_t s, const DevParams& P, const DevTank* tanks,
                      uint32_t n, DevState st, const uint32_t* tag,
                      uint32_t* key, uint32_t* flags);
void launch_permute(hipStream_t s, int dim, uint32_t n, const uint32_t* perm,
                    DevState src, DevState dst, const uint32_t* id_src,
                    uint32_t* id_dst, const uint32_t* tag_src,
                    uint32_t* tag_dst);
void launch_cell_offsets(hipStream_t s, uint32_t n_cells, uint32_t n,
                         const uint32_t* key, uint32_t* cell_off);
void launch_compact_fixed(hipStream_t s, uint32_t n, const uint32_t* tag,
                          uint32_t* fixed_list, uint32_t* counter);
void launch_boundary(hipStream_t s, const BoundaryArgs& a);
void launch_stage(hipStream_t s, const StageArgs& a, int kind, bool final);
int strip_threads(int dim);
uint32_t strip_lds_capacity(int dim);
void launch_time_step(hipStream_t s, const DevParams& P, uint32_t n,
                      DevState st, const uint32_t* tag, const uint32_t* id,
                      const double* dv_ref, unsigned long long* dt_min);
void launch_step_begin(hipStream_t s, int n_tanks, unsigned long long* dt_min,
                       double* dt, uint32_t* vmax);
void launch_step_end(hipStream_t s, int n_tanks, const double* dt,
                     const uint32_t* vmax, double* time, double* disp,
                     double skin_half, uint32_t* flags);
void launch_reset_disp(hipStream_t s, int n_tanks, double* disp);
void launch_export(hipStream_t s, const DevParams& P, uint32_t n, DevState st,
                   const uint32_t* id, double* r, double* v, double* rho,
                   double* p);
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
