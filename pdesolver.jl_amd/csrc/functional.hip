// functional.hip -- boundary functionals (calcBndryFunctional,
// src/solver/euler/boundary_functional.jl:82-126) on the device state.
//
// Per boundary face node: q_b = R q, then the integrand of
//   BoundaryForceData (lift/drag, boundary_functional.jl:340-401): Euler flux of
//     q with its normal momentum removed, dotted with (facx, facy[, facz]);
//   EntropyFluxData (:417-436): (rho u . n) * U, U = -rho s / (gamma-1)
//     (the reference's `unet/q[1]` is a discarded expression, so unet is the
//     normal momentum, not the normal velocity; kept as is);
//   MassFlowData (:404-414): rho u . n.
// integratefunctional! sums wface_k * integrand_k over the faces of the listed
// BC groups; the sum is taken on the host in face order (deterministic).
#include <vector>

#include "ctx.hpp"

namespace pdes {

template <int DIM, int NF, int ST>
__global__ void functional_integrand_kernel(DevMesh mesh, DevSBP sbp, PhysConst phys, const double* __restrict__ q,
                                            int Np, const int32_t* __restrict__ bfaces,
                                            const int32_t* __restrict__ bidx, int nface, double* __restrict__ out) {
  constexpr int ND = DIM + 2, NFACES = DIM + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nface * NF) return;
  const int i = t / NF, k = t - i * NF;
  const int b = bidx[i];
  const int64_t e = bfaces[2 * b];
  const int f = bfaces[2 * b + 1];
  double qb[ND];
  for (int c = 0; c < ND; ++c) qb[c] = 0.0;
  for (int s = 0; s < ST; ++s) {
    const double r = sbp.interp[s * NF + k];
    const int node = sbp.perm[s * NFACES + f];
    for (int c = 0; c < ND; ++c) qb[c] += r * q[(e * Np + node) * ND + c];
  }
  const double* nrm = mesh.nrm_bndry + ((int64_t)b * NF + k) * DIM;
  const double w = sbp.wface[k];
  // force
  double nn = 0.0;
  for (int d = 0; d < DIM; ++d) nn += nrm[d] * nrm[d];
  const double fac = 1.0 / sqrt(nn);
  double nh[DIM];
  double normal_momentum = 0.0;
  for (int d = 0; d < DIM; ++d) { nh[d] = nrm[d] * fac; normal_momentum += nh[d] * qb[1 + d]; }
  double qg[ND];
  for (int c = 0; c < ND; ++c) qg[c] = qb[c];
  for (int d = 0; d < DIM; ++d) qg[1 + d] -= nh[d] * normal_momentum;
  double F[ND];
  euler_flux<DIM>(phys, qg, nrm, F);
  double* o = out + (int64_t)t * 5;
  o[0] = w * F[1];
  o[1] = w * F[2];
  o[2] = DIM == 3 ? w * F[3] : 0.0;
  // entropy flux
  double unet = 0.0;
  for (int d = 0; d < DIM; ++d) unet += qb[d + 1] * nrm[d];
  const double p = pressure<DIM>(phys, qb);
  const double s = log(p) - phys.gamma * log(qb[0]);
  const double U = -qb[0] * s / phys.gamma_1;
  o[3] = w * unet * U;
  // mass flow
  o[4] = w * unet;
}

}  // namespace pdes

using namespace pdes;

extern "C" {

// kind: 0 force vector (dim values), 1 lift force, 2 drag force,
//       3 entropy flux, 4 mass flow, 5 lift coefficient, 6 drag coefficient
int pdes_eval_functional(pdes_ctx* c, int kind, const int32_t* bcnums, int nbc, double* val) {
  PDES_CHECK(c && bcnums && val && nbc > 0 && kind >= 0 && kind <= 6, PDES_ERR_ARG, "bad argument");
  std::vector<int32_t> bidx;
  for (int i = 0; i < nbc; ++i) {
    PDES_CHECK(bcnums[i] >= 0 && bcnums[i] < c->numBC, PDES_ERR_ARG, "BC number %d out of range", bcnums[i]);
    for (int b = c->h_bndry_offsets[bcnums[i]]; b < c->h_bndry_offsets[bcnums[i] + 1]; ++b) bidx.push_back(b);
  }
  const int nface = (int)bidx.size();
  const int Nf = c->Nf;
  std::vector<double> host((size_t)nface * Nf * 5, 0.0);
  if (nface > 0) {
    int32_t *d_bf = nullptr, *d_bidx = nullptr;
    double* d_out = nullptr;
    PDES_HIP(hipMalloc((void**)&d_bf, c->h_bndryfaces.size() * sizeof(int32_t)));
    PDES_HIP(hipMalloc((void**)&d_bidx, bidx.size() * sizeof(int32_t)));
    PDES_HIP(hipMalloc((void**)&d_out, host.size() * sizeof(double)));
    PDES_HIP(hipMemcpyAsync(d_bf, c->h_bndryfaces.data(), c->h_bndryfaces.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    PDES_HIP(hipMemcpyAsync(d_bidx, bidx.data(), bidx.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    const int threads = 128, blocks = (int)ceil_div((int64_t)nface * Nf, threads);
    bool launched = false;
#define FUNC_LAUNCH(D, NFV, STV)                                                                                    \
    if (!launched && c->dim == D && Nf == NFV && c->stencil == STV) {                                              \
      hipLaunchKernelGGL((functional_integrand_kernel<D, NFV, STV>), dim3(blocks), dim3(threads), 0, c->stream,     \
                         c->devmesh(), c->devsbp(), c->pc, c->d_q, c->Np, d_bf, d_bidx, nface, d_out);            \
      launched = true;                                                                                             \
    }
    FUNC_LAUNCH(2, 2, 3) FUNC_LAUNCH(2, 3, 6) FUNC_LAUNCH(2, 4, 10) FUNC_LAUNCH(2, 5, 15)
    FUNC_LAUNCH(2, 3, 3) FUNC_LAUNCH(2, 4, 4) FUNC_LAUNCH(2, 5, 5) FUNC_LAUNCH(2, 6, 6)
    FUNC_LAUNCH(3, 3, 4) FUNC_LAUNCH(3, 6, 10)
#undef FUNC_LAUNCH
    hipError_t le = hipGetLastError();
    if (launched && le == hipSuccess)
      le = hipMemcpyAsync(host.data(), d_out, host.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (le == hipSuccess) le = hipStreamSynchronize(c->stream);
    hipFree(d_bf);
    hipFree(d_bidx);
    hipFree(d_out);
    PDES_CHECK(launched, PDES_ERR_UNSUPPORTED, "no functional kernel for this operator");
    PDES_CHECK(le == hipSuccess, PDES_ERR_HIP, "functional kernel failed: %s", hipGetErrorString(le));
  }
  double acc[5] = {0, 0, 0, 0, 0};
  for (size_t t = 0; t < (size_t)nface * Nf; ++t)
    for (int v = 0; v < 5; ++v) acc[v] += host[t * 5 + v];
  const double aoa = c->phys.aoa;
  double lift, drag;
  if (c->dim == 2) {
    lift = -std::sin(aoa) * acc[0] + std::cos(aoa) * acc[1];
    drag = std::cos(aoa) * acc[0] + std::sin(aoa) * acc[1];
  } else {
    lift = -std::sin(aoa) * acc[0] + std::cos(aoa) * acc[2];
    drag = std::cos(aoa) * acc[0] + std::sin(aoa) * acc[2];
  }
  const double qdyn = 0.5 * c->phys.rho_free * c->phys.Ma * c->phys.Ma;
  switch (kind) {
    case 0: for (int d = 0; d < c->dim; ++d) val[d] = acc[d]; break;
    case 1: val[0] = lift; break;
    case 2: val[0] = drag; break;
    case 3: val[0] = acc[3]; break;
    case 4: val[0] = acc[4]; break;
    case 5: val[0] = lift / qdyn; break;
    case 6: val[0] = drag / qdyn; break;
  }
  return PDES_OK;
}

}  // extern "C"
