// residual_launch.hpp -- launch-variant interface of the fused residual kernel.
//
// Tunable launch variants of the FP64 value kernel (the directional-derivative
// kernel always uses variant 2's shape with 2 waves):
//   0: elements per block = 256/Np, 3 waves per SIMD
//   1: elements per block = 256/Np, 4 waves per SIMD
//   2: elements per block = 128/Np, 3 waves per SIMD
//   3: elements per block = 128/Np, 4 waves per SIMD
//   4-7: the weak form only: variants 0-3 with the volume term Q^T F on the
//        FP64 matrix cores (blocks rounded up to whole waves); the split form
//        runs variant (v & 3) for these
// The waves-per-SIMD bound trades registers for occupancy; the choice is
// made per context by the autotuner (pdes_autotune) or from the default
// table in residual.hip.
#pragma once
#include "ctx.hpp"

namespace pdes {
constexpr int NUM_RES_VARIANTS = 8;
// variants that differ for this context (the split form has no matrix-core variants)
inline int res_variant_count(const pdes_ctx* c) { return c->phys.volume_integral_type == 1 ? 8 : 4; }
template <class Op, class T>
int launch_shape(pdes_ctx* c, const double* q, const double* v, double* res, double* jv, int32_t* flag,
                 const RKStage* rk, int variant);
}  // namespace pdes
