// residual_launch.hpp -- launch-variant interface of the fused residual kernel.
//
// Tunable launch variants of the FP64 value kernel (the directional-derivative
// kernel always uses variant 2's shape with 2 waves):
//   0: elements per block = 256/Np, 3 waves per SIMD
//   1: elements per block = 256/Np, 4 waves per SIMD
//   2: elements per block = 128/Np, 3 waves per SIMD
//   3: elements per block = 128/Np, 4 waves per SIMD
// The waves-per-SIMD bound trades registers for occupancy; the choice is
// made per context by the autotuner (pdes_autotune) or from the default
// table in residual.hip.
#pragma once
#include "ctx.hpp"

namespace pdes {
constexpr int NUM_RES_VARIANTS = 4;
template <class Op, class T>
int launch_shape(pdes_ctx* c, const double* q, const double* v, double* res, double* jv, int32_t* flag,
                 const RKStage* rk, int variant);
}  // namespace pdes
