// Learned step: an edge/node-MLP message-passing network that predicts the
// acceleration and the density rate of every fluid particle for a step of
// 4 dt.  It has no counterpart in the reference; it replaces the pair sums of
// FluidEquations (fluid_equations.hpp:232-305) inside an explicit step.
//
// Graph: every particle a (fluid and fixed) is joined to its kKnn nearest
// other particles b with |r_b - r_a|^2 <= h^2 (FP64, the reference's
// inclusive predicate form), ordered by (distance^2, particle id); missing
// edges point back to a and carry valid = 0.
//
// Per round r = 0, 1 with node input H_r = [nf | u_{r-1}] (u_{-1} = 0):
//   edge   x_e = [ef_e | H_a | H_b]                         (kEdgeIn)
//          m_e = act(act(x_e W1 + b1) W2 + b2)              (kHidden)
//   node   M_a = sum over the kKnn edges of a of m_e
//          y_a = [H_a | M_a]                                (kNodeCat)
//          u_a = act(act(y_a W3 + b3) W4 + b4) W5 + b5      (kOut)
// The last round's u holds (a_x, a_y, a_z, drho/dt) divided by their scales.
// GEMM operands are BF16, accumulation FP32.
#pragma once

#include "nn_kernels.hpp"

namespace tit_amd {

constexpr int kKnn = 8;
constexpr int kNodeFeat = 16;
constexpr int kLatent = 16;
constexpr int kNodeIn = kNodeFeat + kLatent;          // 32
constexpr int kEdgeFeat = 16;
constexpr int kEdgeIn = kEdgeFeat + 2 * kNodeIn;      // 80
constexpr int kHidden = 128;
constexpr int kNodeCat = kNodeIn + kHidden;           // 160
constexpr int kOut = kLatent;                          // 16
constexpr int kRounds = 2;

// Offsets of one round's parameters in the flat FP32 parameter vector; W are
// row-major [in x out].
struct RoundLayout {
  static constexpr int W1 = 0;
  static constexpr int b1 = W1 + kEdgeIn * kHidden;
  static constexpr int W2 = b1 + kHidden;
  static constexpr int b2 = W2 + kHidden * kHidden;
  static constexpr int W3 = b2 + kHidden;
  static constexpr int b3 = W3 + kNodeCat * kHidden;
  static constexpr int W4 = b3 + kHidden;
  static constexpr int b4 = W4 + kHidden * kHidden;
  static constexpr int W5 = b4 + kHidden;
  static constexpr int b5 = W5 + kHidden * kOut;
  static constexpr int size = b5 + kOut;
};
constexpr int kParams = kRounds * RoundLayout::size;

// Feature and output scales.
//   nf = [v / v_s (3), (rho / rho0 - 1) / rho_s, fixed, n_wall (3), 0...]
//   ef = [(r_b - r_a) / h (3), |r_b - r_a| / h, w(q) / w(0),
//         (v_b - v_a) / v_s (3), (rho_b - rho_a) / (rho0 rho_s), valid, 0...]
// Output channel c (0..2 acceleration, 3 density rate) is out_scale[c] u[c].
struct LearnedScales {
  float v_s;
  float rho_s;
  float out_scale[4];
};

// Fused inference (surrogate_fused.hip): MFMA operand fragments of both
// rounds' weights, packed from the BF16 parameter vector, and one launch per
// round.  H (n x kNodeIn) is the round's node input; Hn (may be null)
// receives [nf | u] for the next round and u (may be null) the outputs.
size_t fused_fragment_elems();
void pack_fragments(hipStream_t s, const bf16* w16, bf16* frag);
void launch_round_fused(hipStream_t s, int n_cu, uint32_t n, int round,
                        const bf16* frag, const float* w32, const uint32_t* nbr,
                        const bf16* ef, const bf16* H, bf16* Hn, float* u);

}  // namespace tit_amd
