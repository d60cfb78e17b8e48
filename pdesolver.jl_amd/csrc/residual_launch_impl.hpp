// residual_launch_impl.hpp -- instantiation bodies of launch_shape, included by
// the per-shape translation units (res_*.hip) so the kernel variants compile
// in parallel.
#pragma once
#include "residual_launch.hpp"

namespace pdes {

template <class Op, int VOL, int FLUX, class T, int EPB, int WAVES, bool MF = false>
static int launch_cfg(pdes_ctx* c, const double* q, const double* v, double* res, double* jv, int32_t* flag,
                      const RKStage* rk) {
  // the matrix-core variants run whole waves (the extra lanes only stage
  // tables and join the matrix products)
  const int threads = MF ? (int)ceil_div(EPB * Op::NP, 64) * 64 : EPB * Op::NP;
  if (c->active_count() == 0) return PDES_OK;
  const int blocks = (int)ceil_div(c->active_count(), EPB);
  if constexpr (std::is_same<T, double>::value) {
    if (rk) {
      hipLaunchKernelGGL((euler_residual_kernel<Op, VOL, FLUX, EPB, T, true, WAVES, MF>), dim3(blocks), dim3(threads), 0,
                         c->stream, c->devmesh(), c->devsbp(), c->pc, q, v, res, jv, flag, *rk);
      PDES_HIP(hipGetLastError());
      return PDES_OK;
    }
  }
  hipLaunchKernelGGL((euler_residual_kernel<Op, VOL, FLUX, EPB, T, false, WAVES, MF>), dim3(blocks), dim3(threads), 0,
                     c->stream, c->devmesh(), c->devsbp(), c->pc, q, v, res, jv, flag, RKStage{});
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

template <class Op, int VOL, int FLUX, class T>
static int launch_one(pdes_ctx* c, const double* q, const double* v, double* res, double* jv, int32_t* flag,
                      const RKStage* rk, int variant) {
  // the entropy-stable face integral stages the neighbours' stencil data in
  // LDS: half the elements per block
  constexpr int FULL = (FLUX == PDES_FLUX_ESLF ? 128 : 256) / Op::NP;
  constexpr int HALF = (FLUX == PDES_FLUX_ESLF ? 64 : 128) / Op::NP;
  if constexpr (!std::is_same<T, double>::value) {
    return launch_cfg<Op, VOL, FLUX, T, HALF, 2>(c, q, v, res, jv, flag, nullptr);
  } else {
    if constexpr (VOL == VOL_WEAK) {
      switch (variant) {
        case 4: return launch_cfg<Op, VOL, FLUX, T, FULL, 3, true>(c, q, v, res, jv, flag, rk);
        case 5: return launch_cfg<Op, VOL, FLUX, T, FULL, 4, true>(c, q, v, res, jv, flag, rk);
        case 6: return launch_cfg<Op, VOL, FLUX, T, HALF, 3, true>(c, q, v, res, jv, flag, rk);
        case 7: return launch_cfg<Op, VOL, FLUX, T, HALF, 4, true>(c, q, v, res, jv, flag, rk);
        default: break;
      }
    }
    switch (variant & 3) {
      case 0: return launch_cfg<Op, VOL, FLUX, T, FULL, 3>(c, q, v, res, jv, flag, rk);
      case 1: return launch_cfg<Op, VOL, FLUX, T, FULL, 4>(c, q, v, res, jv, flag, rk);
      case 2: return launch_cfg<Op, VOL, FLUX, T, HALF, 3>(c, q, v, res, jv, flag, rk);
      default: return launch_cfg<Op, VOL, FLUX, T, HALF, 4>(c, q, v, res, jv, flag, rk);
    }
  }
}

template <class Op, class T>
int launch_shape(pdes_ctx* c, const double* q, const double* v, double* res, double* jv, int32_t* flag,
                 const RKStage* rk, int variant) {
  const int vol = c->phys.volume_integral_type, flux = c->phys.face_flux;
  if (vol == 1 && flux == PDES_FLUX_ROE)
    return launch_one<Op, VOL_WEAK, PDES_FLUX_ROE, T>(c, q, v, res, jv, flag, rk, variant);
  if constexpr (Op::DIAGE && Op::DIM == 2) {
    if (vol == 2 && flux == PDES_FLUX_IRSLF)
      return launch_one<Op, VOL_SPLIT, PDES_FLUX_IRSLF, T>(c, q, v, res, jv, flag, rk, variant);
  }
  if constexpr (!Op::DIAGE) {
    if (vol == 2 && flux == PDES_FLUX_ESLF)
      return launch_one<Op, VOL_SPLIT, PDES_FLUX_ESLF, T>(c, q, v, res, jv, flag, rk, variant);
  }
  set_error("unsupported combination volume_integral_type=%d face_flux=%d for this operator", vol, flux);
  return PDES_ERR_UNSUPPORTED;
}

}  // namespace pdes

#define PDES_INSTANTIATE_SHAPE(D, NP, NF, ST, DE)                                                           \
  template int pdes::launch_shape<pdes::OpShape<D, NP, NF, ST, DE>, double>(                                \
      pdes_ctx*, const double*, const double*, double*, double*, int32_t*, const pdes::RKStage*, int);     \
  template int pdes::launch_shape<pdes::OpShape<D, NP, NF, ST, DE>, pdes::dual<1>>(                         \
      pdes_ctx*, const double*, const double*, double*, double*, int32_t*, const pdes::RKStage*, int);
