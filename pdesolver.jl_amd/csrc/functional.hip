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
#include <algorithm>
#include <cmath>
#include <vector>

#include "ctx.hpp"

namespace pdes {

// 1/sqrt(x), for the normal's length
__device__ __forceinline__ double inv_len(double x) { return 1.0 / sqrt(x); }
template <int N> __device__ __forceinline__ dual<N> inv_len(const dual<N>& a) {
  dual<N> r;
  r.v = 1.0 / sqrt(a.v);
  const double f = -0.5 * r.v / a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * f;
  return r;
}

// the five integrands at one face node: force x, y, z (weighted by wface),
// entropy flux, mass flow; T is double, or a dual type for their derivatives
// with respect to the face-node state (TN double) or to the normal (TN = T)
template <int DIM, class T, class TN = double>
__device__ __forceinline__ void face_integrands(const PhysConst& phys, const T* qb, const TN* nrm, double w,
                                                T* o) {
  constexpr int ND = DIM + 2;
  // force
  TN nn = 0.0;
  for (int d = 0; d < DIM; ++d) nn = nn + nrm[d] * nrm[d];
  const TN fac = inv_len(nn);
  TN nh[DIM];
  T normal_momentum = 0.0;
  for (int d = 0; d < DIM; ++d) { nh[d] = nrm[d] * fac; normal_momentum = normal_momentum + nh[d] * qb[1 + d]; }
  T qg[ND];
  for (int c = 0; c < ND; ++c) qg[c] = qb[c];
  for (int d = 0; d < DIM; ++d) qg[1 + d] = qg[1 + d] - nh[d] * normal_momentum;
  T F[ND];
  euler_flux<DIM>(phys, qg, nrm, F);
  o[0] = w * F[1];
  o[1] = w * F[2];
  o[2] = DIM == 3 ? w * F[3] : T(0.0);
  // entropy flux
  T unet = 0.0;
  for (int d = 0; d < DIM; ++d) unet = unet + qb[d + 1] * nrm[d];
  const T p = pressure<DIM>(phys, qb);
  const T s = dlog(p) - phys.gamma * dlog(qb[0]);
  const T U = -qb[0] * s / phys.gamma_1;
  o[3] = w * unet * U;
  // mass flow
  o[4] = w * unet;
}

// q_b = R q at face node k of boundary face b
template <int DIM, int NF, int ST>
__device__ __forceinline__ void face_state(const DevSBP& sbp, const double* q, int Np, int64_t e, int f, int k,
                                           double* qb) {
  constexpr int ND = DIM + 2, NFACES = DIM + 1;
  for (int c = 0; c < ND; ++c) qb[c] = 0.0;
  for (int s = 0; s < ST; ++s) {
    const double r = sbp.interp[s * NF + k];
    const int node = sbp.perm[s * NFACES + f];
    for (int c = 0; c < ND; ++c) qb[c] += r * q[(e * Np + node) * ND + c];
  }
}

template <int DIM, int NF, int ST>
__global__ void functional_integrand_kernel(DevMesh mesh, DevSBP sbp, PhysConst phys, const double* __restrict__ q,
                                            int Np, const int32_t* __restrict__ bfaces,
                                            const int32_t* __restrict__ bidx, int nface, double* __restrict__ out) {
  constexpr int ND = DIM + 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nface * NF) return;
  const int i = t / NF, k = t - i * NF;
  const int b = bidx[i];
  double qb[ND];
  face_state<DIM, NF, ST>(sbp, q, Np, bfaces[2 * b], bfaces[2 * b + 1], k, qb);
  face_integrands<DIM>(phys, qb, mesh.nrm_bndry + ((int64_t)b * NF + k) * DIM, sbp.wface[k], out + (int64_t)t * 5);
}

// evalFunctionalDeriv_q (src/interface2.jl:259-283) for a scalar functional
// sum_k wts . integrands_k: at each face node of the selected boundary faces,
// g_k = sum_v wts[v] d(integrand_v)/dq_b (forward mode); zero elsewhere
template <int DIM, int NF, int ST>
__global__ void functional_face_grad_kernel(DevMesh mesh, DevSBP sbp, PhysConst phys, const double* __restrict__ q,
                                            int Np, const int32_t* __restrict__ bfaces, int nbndry,
                                            const int32_t* __restrict__ sel, const double* __restrict__ wts,
                                            double* __restrict__ grad) {
  constexpr int ND = DIM + 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbndry * NF) return;
  const int b = t / NF, k = t - b * NF;
  double* gk = grad + (int64_t)t * ND;
  if (!sel[b]) {
    for (int c = 0; c < ND; ++c) gk[c] = 0.0;
    return;
  }
  double qv[ND];
  face_state<DIM, NF, ST>(sbp, q, Np, bfaces[2 * b], bfaces[2 * b + 1], k, qv);
  dual<ND> qb[ND], o[5];
  for (int c = 0; c < ND; ++c) { qb[c] = dual<ND>(qv[c]); qb[c].d[c] = 1.0; }
  face_integrands<DIM>(phys, qb, mesh.nrm_bndry + ((int64_t)b * NF + k) * DIM, sbp.wface[k], o);
  for (int c = 0; c < ND; ++c) {
    double g = 0.0;
    for (int v = 0; v < 5; ++v) g += wts[v] * o[v].d[c];
    gk[c] = g;
  }
}

// evalFunctionalDeriv_m (src/interface2.jl:205-227 -> _evalFunctionalDeriv_m,
// boundary_functional.jl:61-79) for a scalar functional sum_k wts .
// integrands_k: the integrands depend on the metrics through the boundary
// normals only, so nrm_bndry_bar[b][k][x] = sum_v wts[v] d(integrand_v)/dn_x
// (forward mode over the normal) at the selected faces, zero elsewhere
template <int DIM, int NF, int ST>
__global__ void functional_nrm_grad_kernel(DevMesh mesh, DevSBP sbp, PhysConst phys, const double* __restrict__ q,
                                           int Np, const int32_t* __restrict__ bfaces, int nbndry,
                                           const int32_t* __restrict__ sel, const double* __restrict__ wts,
                                           double* __restrict__ nbar) {
  constexpr int ND = DIM + 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbndry * NF) return;
  const int b = t / NF, k = t - b * NF;
  double* gk = nbar + (int64_t)t * DIM;
  if (!sel[b]) {
    for (int x = 0; x < DIM; ++x) gk[x] = 0.0;
    return;
  }
  double qv[ND];
  face_state<DIM, NF, ST>(sbp, q, Np, bfaces[2 * b], bfaces[2 * b + 1], k, qv);
  dual<DIM> qb[ND], nd[DIM], o[5];
  for (int c = 0; c < ND; ++c) qb[c] = dual<DIM>(qv[c]);
  const double* nr = mesh.nrm_bndry + ((int64_t)b * NF + k) * DIM;
  for (int x = 0; x < DIM; ++x) {
    nd[x] = dual<DIM>(nr[x]);
    nd[x].d[x] = 1.0;
  }
  face_integrands<DIM>(phys, qb, nd, sbp.wface[k], o);
  for (int x = 0; x < DIM; ++x) {
    double g = 0.0;
    for (int v = 0; v < 5; ++v) g += wts[v] * o[v].d[x];
    gk[x] = g;
  }
}

// dJ/dq at volume node i of element e: sum over the element's boundary faces f
// that contain node i (stencil position s) of sum_k R[s][k] g_k; one thread
// per (element, node), so every entry has one writer
template <int DIM, int NF, int ST>
__global__ void functional_gather_kernel(DevMesh mesh, DevSBP sbp, int Np, const double* __restrict__ grad,
                                         double* __restrict__ dJdq) {
  constexpr int ND = DIM + 2, NFACES = DIM + 1;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)mesh.numEl * Np) return;
  const int64_t e = t / Np;
  const int i = (int)(t - e * Np);
  double acc[ND];
  for (int c = 0; c < ND; ++c) acc[c] = 0.0;
  for (int f = 0; f < NFACES; ++f) {
    const ElFace r = mesh.elface[e * NFACES + f];
    if (r.x >= 0) continue;
    const int s = sbp.nodepos[i * NFACES + f];
    if (s < 0) continue;
    for (int k = 0; k < NF; ++k) {
      const double w = sbp.interp[s * NF + k];
      const double* gk = grad + ((int64_t)r.y * NF + k) * ND;
      for (int c = 0; c < ND; ++c) acc[c] += w * gk[c];
    }
  }
  for (int c = 0; c < ND; ++c) dJdq[t * ND + c] = acc[c];
}

// the boundary faces of the BC groups bcnums (sel[b] = 1) and the weights of
// the five integrands that make up the scalar functional `kind` (1..6)
static int functional_setup(const pdes_ctx* c, int kind, const int32_t* bcnums, int nbc, std::vector<int32_t>& sel,
                            double* wts) {
  PDES_CHECK(c && bcnums && nbc > 0 && kind >= 1 && kind <= 6, PDES_ERR_ARG,
             "bad argument (the derivative needs a scalar functional, kind 1..6)");
  const int nb = (int)(c->h_bndryfaces.size() / 2);
  sel.assign(std::max(1, nb), 0);
  for (int i = 0; i < nbc; ++i) {
    PDES_CHECK(bcnums[i] >= 0 && bcnums[i] < c->numBC, PDES_ERR_ARG, "BC number %d out of range", bcnums[i]);
    for (int b = c->h_bndry_offsets[bcnums[i]]; b < c->h_bndry_offsets[bcnums[i] + 1]; ++b) sel[b] = 1;
  }
  const double aoa = c->phys.aoa, qdyn = 0.5 * c->phys.rho_free * c->phys.Ma * c->phys.Ma;
  const int up = c->dim == 2 ? 1 : 2;
  for (int v = 0; v < 5; ++v) wts[v] = 0.0;
  switch (kind) {
    case 1: case 5: wts[0] = -std::sin(aoa); wts[up] = std::cos(aoa); break;
    case 2: case 6: wts[0] = std::cos(aoa); wts[up] = std::sin(aoa); break;
    case 3: wts[3] = 1.0; break;
    case 4: wts[4] = 1.0; break;
  }
  if (kind >= 5)
    for (int v = 0; v < 5; ++v) wts[v] /= qdyn;
  return PDES_OK;
}

// dJ/dq of the scalar functional `kind` (1..6) over the BC groups bcnums,
// into the device vector d_out [numDof]
int functional_grad_device(pdes_ctx* c, int kind, const int32_t* bcnums, int nbc, double* d_out) {
  PDES_CHECK(c && d_out, PDES_ERR_ARG, "null argument");
  const int nb = (int)(c->h_bndryfaces.size() / 2), Nf = c->Nf, dim = c->dim;
  std::vector<int32_t> sel;
  double wts[5];
  int st0 = functional_setup(c, kind, bcnums, nbc, sel, wts);
  if (st0) return st0;
  int32_t *d_bf = nullptr, *d_sel = nullptr;
  double *d_w = nullptr, *d_grad = nullptr;
  const size_t ngrad = (size_t)std::max(1, nb) * Nf * c->nd;
  hipError_t he = hipMalloc((void**)&d_bf, std::max<size_t>(1, c->h_bndryfaces.size()) * sizeof(int32_t));
  if (he == hipSuccess) he = hipMalloc((void**)&d_sel, sel.size() * sizeof(int32_t));
  if (he == hipSuccess) he = hipMalloc((void**)&d_w, sizeof(wts));
  if (he == hipSuccess) he = hipMalloc((void**)&d_grad, ngrad * sizeof(double));
  if (he == hipSuccess && nb > 0)
    he = hipMemcpyAsync(d_bf, c->h_bndryfaces.data(), c->h_bndryfaces.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                        c->stream);
  if (he == hipSuccess)
    he = hipMemcpyAsync(d_sel, sel.data(), sel.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
  if (he == hipSuccess) he = hipMemcpyAsync(d_w, wts, sizeof(wts), hipMemcpyHostToDevice, c->stream);
  bool launched = false;
  if (he == hipSuccess) {
    const int gb = (int)ceil_div((int64_t)std::max(1, nb) * Nf, 128);
    const int ge = (int)ceil_div((int64_t)c->numEl * c->Np, 128);
#define GRAD_LAUNCH(D, NFV, STV)                                                                                   \
    if (!launched && dim == D && Nf == NFV && c->stencil == STV) {                                                \
      if (nb > 0)                                                                                                  \
        hipLaunchKernelGGL((functional_face_grad_kernel<D, NFV, STV>), dim3(gb), dim3(128), 0, c->stream,          \
                           c->devmesh(), c->devsbp(), c->pc, c->d_q, c->Np, d_bf, nb, d_sel, d_w, d_grad);        \
      hipLaunchKernelGGL((functional_gather_kernel<D, NFV, STV>), dim3(ge), dim3(128), 0, c->stream, c->devmesh(), \
                         c->devsbp(), c->Np, d_grad, d_out);                                                       \
      launched = true;                                                                                             \
    }
    GRAD_LAUNCH(2, 2, 3) GRAD_LAUNCH(2, 3, 6) GRAD_LAUNCH(2, 4, 10) GRAD_LAUNCH(2, 5, 15)
    GRAD_LAUNCH(2, 3, 3) GRAD_LAUNCH(2, 4, 4) GRAD_LAUNCH(2, 5, 5) GRAD_LAUNCH(2, 6, 6)
    GRAD_LAUNCH(3, 3, 4) GRAD_LAUNCH(3, 6, 10)
#undef GRAD_LAUNCH
    he = hipGetLastError();
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
  }
  hipFree(d_bf);
  hipFree(d_sel);
  hipFree(d_w);
  hipFree(d_grad);
  PDES_CHECK(he == hipSuccess, PDES_ERR_HIP, "functional derivative failed: %s", hipGetErrorString(he));
  PDES_CHECK(launched, PDES_ERR_UNSUPPORTED, "no functional kernel for this operator");
  return PDES_OK;
}

}  // namespace pdes

using namespace pdes;

extern "C" {

// evalFunctionalDeriv_q: dJ/dq [numDof] of a scalar functional (kind 1..6)
int pdes_functional_deriv(pdes_ctx* c, int kind, const int32_t* bcnums, int nbc, double* dJdq) {
  PDES_CHECK(c && dJdq, PDES_ERR_ARG, "null argument");
  int st = functional_grad_device(c, kind, bcnums, nbc, c->d_work);
  if (st) return st;
  PDES_HIP(hipMemcpyAsync(dJdq, c->d_work, c->numDof * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

// evalFunctionalDeriv_m: dJ/d(nrm_bndry) [numBoundaryFaces][Nf][dim] of a
// scalar functional (kind 1..6); the functionals depend on no other metric
int pdes_functional_deriv_m(pdes_ctx* c, int kind, const int32_t* bcnums, int nbc, double* nrm_bndry_bar) {
  PDES_CHECK(c && nrm_bndry_bar, PDES_ERR_ARG, "null argument");
  std::vector<int32_t> sel;
  double wts[5];
  int st = functional_setup(c, kind, bcnums, nbc, sel, wts);
  if (st) return st;
  const int nb = (int)(c->h_bndryfaces.size() / 2), Nf = c->Nf, dim = c->dim;
  if (nb == 0) return PDES_OK;
  int32_t *d_bf = nullptr, *d_sel = nullptr;
  double *d_w = nullptr, *d_bar = nullptr;
  const size_t nbar = (size_t)nb * Nf * dim;
  hipError_t he = hipMalloc((void**)&d_bf, c->h_bndryfaces.size() * sizeof(int32_t));
  if (he == hipSuccess) he = hipMalloc((void**)&d_sel, sel.size() * sizeof(int32_t));
  if (he == hipSuccess) he = hipMalloc((void**)&d_w, sizeof(wts));
  if (he == hipSuccess) he = hipMalloc((void**)&d_bar, nbar * sizeof(double));
  if (he == hipSuccess)
    he = hipMemcpyAsync(d_bf, c->h_bndryfaces.data(), c->h_bndryfaces.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                        c->stream);
  if (he == hipSuccess)
    he = hipMemcpyAsync(d_sel, sel.data(), sel.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
  if (he == hipSuccess) he = hipMemcpyAsync(d_w, wts, sizeof(wts), hipMemcpyHostToDevice, c->stream);
  bool launched = false;
  if (he == hipSuccess) {
    const int gb = (int)ceil_div((int64_t)nb * Nf, 128);
#define NGRAD_LAUNCH(D, NFV, STV)                                                                           \
    if (!launched && dim == D && Nf == NFV && c->stencil == STV) {                                         \
      hipLaunchKernelGGL((functional_nrm_grad_kernel<D, NFV, STV>), dim3(gb), dim3(128), 0, c->stream,      \
                         c->devmesh(), c->devsbp(), c->pc, c->d_q, c->Np, d_bf, nb, d_sel, d_w, d_bar);    \
      launched = true;                                                                                      \
    }
    NGRAD_LAUNCH(2, 2, 3) NGRAD_LAUNCH(2, 3, 6) NGRAD_LAUNCH(2, 4, 10) NGRAD_LAUNCH(2, 5, 15)
    NGRAD_LAUNCH(2, 3, 3) NGRAD_LAUNCH(2, 4, 4) NGRAD_LAUNCH(2, 5, 5) NGRAD_LAUNCH(2, 6, 6)
    NGRAD_LAUNCH(3, 3, 4) NGRAD_LAUNCH(3, 6, 10)
#undef NGRAD_LAUNCH
    he = hipGetLastError();
    if (launched && he == hipSuccess)
      he = hipMemcpyAsync(nrm_bndry_bar, d_bar, nbar * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
  }
  hipStreamSynchronize(c->stream);
  hipFree(d_bf);
  hipFree(d_sel);
  hipFree(d_w);
  hipFree(d_bar);
  PDES_CHECK(he == hipSuccess, PDES_ERR_HIP, "functional metric derivative failed: %s", hipGetErrorString(he));
  PDES_CHECK(launched, PDES_ERR_UNSUPPORTED, "no functional kernel for this operator");
  return PDES_OK;
}

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
