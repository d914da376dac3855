// swe_common.h -- constants and small helpers shared by the HIP kernels and
// the host side of the MI355X SWE block library.
#pragma once

#include <cstdint>

namespace swe {

// Solver ids follow the reference's WAVE_PROPAGATION_SOLVER switch
// (src/blocks/SWE_DimensionalSplittingMpi.hh:42-50: 0 HLLE, 1 f-wave,
// 2 augmented Riemann), with 3 = Hybrid (src/blocks/SWE_DimensionalSplitting.hh:39).
enum Solver : int { SOLVER_HLLE = 0, SOLVER_FWAVE = 1, SOLVER_AUGRIE = 2, SOLVER_HYBRID = 3 };
// AugRie whose Newton middle state starts from a learned guess (internal id:
// selected when a block has a surrogate network attached).
constexpr int SOLVER_AUGRIE_MLP = 4;
// AugRie with the learned middle state, hard edges deferred: the sweep keeps
// the edges the linearised guess settles and queues the others for a second
// launch that runs the network (fused strategy of a single block only).
constexpr int SOLVER_AUGRIE_DEFER = 5;

// Waves per SIMD the flavour-U sweeps of each solver are compiled for
// (its amdgpu_waves_per_eu floor; the build fails where the compiler cannot
// meet it).  A SIMD lane has 512 registers: 7 waves need at most 72 per lane,
// 6 at most 80.  Hybrid and the deferring learned sweep fit 7 (71-72
// registers, no spills).  AugRie's fused kernel keeps a quiet-tile record
// and loads each column one column ahead (8 registers): 86 registers at 5
// waves.  Measured on the 2048^2 islands against 6 waves without the early
// loads (79 registers; the loads spill at 6): 2.5% faster with quiet tiles
// skipped, 12% faster with every tile computed, 2.6% slower on the disturbed
// islands, where every edge runs the full solver.  HLLE (77) and f-wave (78)
// keep 6 with the early loads; the in-sweep network call holds 136 vector
// and 32 accumulation registers (3 waves).
constexpr int fused_waves_per_simd(int solver) {
    return (solver == SOLVER_HYBRID || solver == SOLVER_AUGRIE_DEFER) ? 7
           : solver == SOLVER_AUGRIE_MLP                             ? 3
           : solver == SOLVER_AUGRIE                                 ? 5
                                                                     : 6;
}

// Step flavours: U = every parallel block of the reference (all edges from
// one state, unsplit update with dry clamp), S = swe_simple's split sweep.
enum Flavour : int { FLAVOUR_U = 0, FLAVOUR_S = 1 };

// src/types/Boundary.hh:5-28
enum Boundary : int { BND_LEFT = 0, BND_RIGHT = 1, BND_BOTTOM = 2, BND_TOP = 3 };
enum BoundaryType : int { OUTFLOW = 0, WALL = 1, INFLOW = 2, CONNECT = 3, CONNECT_WITHIN_RANK = 4, PASSIVE = 5 };
enum GhostlayerState : int { GL_DONE = 0, GL_UNVALID = 1, GL_INTER = 2, GL_SYNC = 3, GL_NEXT = 4 };

// Solver constants (the reference's solver submodule is not vendored; these
// are the values this library states and its CPU oracle uses).
constexpr float kGravity = 9.81f;
constexpr float kSqrtGravity = 0x1.90e864p+1f;  // sqrtf(9.81f), correctly rounded: celerity = kSqrtGravity sqrt(h)
constexpr float kDryTol = 0.01f;
constexpr float kZeroTol = 1e-7f;
constexpr float kNewtonTol = 1e-6f;
constexpr int kMaxNewton = 10;
constexpr float kCriticalTol = 1e-6f;
constexpr float kHybridTol = 1.0f;
// still edges of the augmented solver: both depths >= kStillTol and within a
// factor kStillRatio of each other
constexpr float kStillTol = 0.1f;
constexpr float kStillRatio = 2.0f;

// Block-level constants (src/Constants.hh:4-6, SWE_DimensionalSplittingMpi.cpp:448-456, 502-515).
constexpr float kCfl = 0.4f;
constexpr float kMinSpeed = 0.00001f;
constexpr float kClampTol = 0.1f;

}  // namespace swe
