// revm.hip -- reverse mode of the residual with respect to the mesh metrics
// (evalResidual_revm, src/interface2.jl:329-372 -> src/solver/euler/evaldRdm.jl:19-70):
// for a seed psi (res_bar), the products psi^T dR/dm for
//   dxidx        (volume terms: calcVolumeIntegrals_nopre_revm and
//                 calcVolumeIntegralsSplitForm_revm, evaldRdm.jl:144-163),
//   nrm_face     (interface fluxes: calcFaceIntegral_nopre_revm and
//                 getFaceElementIntegral_revm, :207-230),
//   nrm_bndry and coords_bndry (boundary fluxes and the boundary states:
//                 getBCFluxes_revm, :176-184).
//
// Every metric entry enters a few flux evaluations of one or two elements, so
// each product is formed where those evaluations are, with one writer per
// output entry (no atomics, deterministic):
//   * the volume fluxes are linear in the metric: the Euler flux F(q, n) =
//     sum_x n_x F^x(q) and the Ismail-Roe flux F(qa, qb, n) = sum_x n_x F^x(qa, qb),
//     so dxidx_bar is a sum of Cartesian fluxes weighted by psi moved through Q
//     (weak form) or by 2 S and psi differences (split form);
//   * the face and boundary fluxes are not linear in the normal (the Roe
//     dissipation and the entropy penalty take |U_n| and |n|): each face node
//     evaluates its flux once with the normal, and at boundaries the node
//     coordinates that fix the boundary state, as dual numbers, and dots the
//     derivatives with the flux's weight psi_hat (the interpolated seed of
//     the element(s) the flux enters, times the face quadrature weight).
#include <algorithm>
#include <cmath>

#include "ctx.hpp"
#include "shapes.hpp"

namespace pdes {

// weak form (weakdifferentiate! with Q^T, calcVolumeIntegrals_nopre):
// res_i += sum_d sum_j Q_d[j][i] F(q_j, dxidx[d, :, j]), so
// dxidx_bar[e][j][x][d] = sum_c (sum_i Q_d[j][i] psi_i)_c F^x_c(q_j)
template <int DIM, int NP>
__global__ void revm_volume_weak_kernel(DevMesh mesh, DevSBP sbp, PhysConst phys, const double* __restrict__ q,
                                        const double* __restrict__ psi, double* __restrict__ bar) {
  constexpr int ND = DIM + 2;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)mesh.numEl * NP) return;
  const int64_t e = t / NP;
  const int j = (int)(t - e * NP);
  double qj[ND];
  for (int c = 0; c < ND; ++c) qj[c] = q[(e * NP + j) * ND + c];
  double Psi[DIM][ND];
  for (int d = 0; d < DIM; ++d) {
    for (int c = 0; c < ND; ++c) Psi[d][c] = 0.0;
    for (int i = 0; i < NP; ++i) {
      const double Qji = sbp.Q[(d * NP + j) * NP + i];
      for (int c = 0; c < ND; ++c) Psi[d][c] += Qji * psi[(e * NP + i) * ND + c];
    }
  }
  for (int x = 0; x < DIM; ++x) {
    double dir[DIM];
    for (int y = 0; y < DIM; ++y) dir[y] = y == x ? 1.0 : 0.0;
    double F[ND];
    euler_flux<DIM>(phys, qj, dir, F);
    for (int d = 0; d < DIM; ++d) {
      double acc = 0.0;
      for (int c = 0; c < ND; ++c) acc += Psi[d][c] * F[c];
      bar[((e * NP + j) * DIM + x) * DIM + d] = acc;
    }
  }
}

// split form (calcVolumeIntegralsSplitFormLinear, euler_funcs.jl:240-288):
// for each pair k < j, F_d = F_IR(q_j, q_k, dxidx[d, :, j]),
// res_j -= 2 S_d[j][k] F_d and res_k += 2 S_d[j][k] F_d, so
// dxidx_bar[e][j][x][d] = sum_{k<j} 2 S_d[j][k] sum_c (psi_k - psi_j)_c F^x_c(q_j, q_k)
// with 2 S_d = Q_d - Q_d^T
template <int DIM, int NP>
__global__ void revm_volume_split_kernel(DevMesh mesh, DevSBP sbp, PhysConst phys, const double* __restrict__ q,
                                         const double* __restrict__ psi, double* __restrict__ bar) {
  constexpr int ND = DIM + 2;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)mesh.numEl * NP) return;
  const int64_t e = t / NP;
  const int j = (int)(t - e * NP);
  double qj[ND], zj[ND], pj[ND];
  for (int c = 0; c < ND; ++c) {
    qj[c] = q[(e * NP + j) * ND + c];
    pj[c] = psi[(e * NP + j) * ND + c];
  }
  ir_zvars<DIM>(phys, qj, zj);
  double dir[DIM * DIM];
  for (int x = 0; x < DIM; ++x)
    for (int y = 0; y < DIM; ++y) dir[x * DIM + y] = x == y ? 1.0 : 0.0;
  double acc[DIM][DIM];
  for (int x = 0; x < DIM; ++x)
    for (int d = 0; d < DIM; ++d) acc[x][d] = 0.0;
  for (int k = 0; k < j; ++k) {
    double qk[ND], zk[ND], dp[ND];
    for (int c = 0; c < ND; ++c) {
      qk[c] = q[(e * NP + k) * ND + c];
      dp[c] = psi[(e * NP + k) * ND + c] - pj[c];
    }
    ir_zvars<DIM>(phys, qk, zk);
    double F[DIM * ND];   // [x][c]
    ir_flux_z<DIM, DIM>(phys, zj, zk, dir, F);
    for (int x = 0; x < DIM; ++x) {
      double w = 0.0;
      for (int c = 0; c < ND; ++c) w += dp[c] * F[x * ND + c];
      for (int d = 0; d < DIM; ++d) {
        const double s2 = sbp.Q[(d * NP + j) * NP + k] - sbp.Q[(d * NP + k) * NP + j];
        acc[x][d] += s2 * w;
      }
    }
  }
  for (int x = 0; x < DIM; ++x)
    for (int d = 0; d < DIM; ++d) bar[((e * NP + j) * DIM + x) * DIM + d] = acc[x][d];
}

// calcIsentropicVortex (common_funcs.jl:25-76) with the coordinates as dual
// numbers: sin(theta) = y/r and cos(theta) = x/r in place of atan2
template <class T>
__device__ __forceinline__ void isentropic_vortex_dual(const PhysConst& p, const T* xy, T* sol) {
  const double gamma = p.gamma;
  const double r_in = 1, rho_in = 2;
  const double M_in = p.vortex_Min, p_in = 1 / gamma;
  const T x = xy[0], y = xy[1];
  const T r2 = x * x + y * y;
  const T r = dsqrt(r2);
  const double tmp1 = ((gamma - 1) / 2) * M_in * M_in;
  const T rho_r = rho_in * dpow(1.0 + tmp1 * (1.0 - (r_in * r_in) / r2), 1 / (gamma - 1));
  const T p_r = p_in * dpow(rho_r / rho_in, gamma);
  const T a_r = dsqrt(gamma * p_r / rho_r);
  const T M_r = dsqrt((2 / (gamma - 1)) * dpow(rho_in / rho_r, gamma - 1) * (1 + tmp1) - 2 / (gamma - 1));
  const T U_r = M_r * a_r;
  const T rinv = 1.0 / r;
  const T u_r = U_r * (y * rinv);
  const T v_r = -(U_r * (x * rinv));
  const T e_r = p.cv * p_r / (rho_r * p.R);
  const T E_r = rho_r * e_r + 0.5 * rho_r * U_r * U_r;
  sol[0] = rho_r;
  sol[1] = rho_r * u_r;
  sol[2] = rho_r * v_r;
  sol[3] = E_r;
}

// One thread per (element, face, face node).  Interior faces are taken from
// their elementL side (each interface node once); the flux there is the one
// the residual uses (Roe, IRSLF, or the ESLF entropy penalty plus the
// entropy-conservative face integral), with weight
// w_k (psi_hat_R(kR) - psi_hat_L(k)): the face integral subtracts
// R^T B F from elementL and adds it to elementR.  Boundary faces: Roe flux
// against the boundary state, weight -w_k psi_hat_L(k).  Dual slots
// 0..DIM-1: the normal; DIM..2 DIM-1: the boundary node's coordinates.
template <int DIM, int NP, int NF, int ST, int FLUX>
__global__ void revm_face_kernel(DevMesh mesh, DevSBP sbp, PhysConst phys, const double* __restrict__ q,
                                 const double* __restrict__ psi, double* __restrict__ nrm_face_bar,
                                 double* __restrict__ nrm_bndry_bar, double* __restrict__ coords_bndry_bar) {
  constexpr int ND = DIM + 2, NFACES = DIM + 1, K = 2 * DIM;
  using D = dual<K>;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)mesh.numEl * NFACES * NF) return;
  const int64_t ef = t / NF;
  const int k = (int)(t - ef * NF);
  const int64_t e = ef / NFACES;
  const int f = (int)(ef - e * NFACES);
  const ElFace rec = mesh.elface[e * NFACES + f];
  const bool bnd = rec.x < 0;
  if (!bnd && ((rec.z >> 16) & 1)) return;   // elementR side: its elementL takes the node
  double qL[ND], pL[ND];
  for (int c = 0; c < ND; ++c) { qL[c] = 0.0; pL[c] = 0.0; }
  for (int s = 0; s < ST; ++s) {
    const double r = sbp.interp[s * NF + k];
    const int node = sbp.perm[s * NFACES + f];
    for (int c = 0; c < ND; ++c) {
      qL[c] += r * q[(e * NP + node) * ND + c];
      pL[c] += r * psi[(e * NP + node) * ND + c];
    }
  }
  D qLd[ND], qRd[ND], nd[DIM], Fd[ND];
  for (int c = 0; c < ND; ++c) qLd[c] = D(qL[c]);
  double fbar[ND];
  const double w = sbp.wface[k];
  const double* nsrc = (bnd ? mesh.nrm_bndry : mesh.nrm_face) + ((int64_t)rec.y * NF + k) * DIM;
  for (int x = 0; x < DIM; ++x) {
    nd[x] = D(nsrc[x]);
    nd[x].d[x] = 1.0;
  }
  if (bnd) {
    const PhysConst bp = bc_phys(mesh, phys, rec);
    bool vortex = false;
    if constexpr (DIM == 2) vortex = bc_kind(rec) == PDES_BC_ISENTROPIC_VORTEX;
    if (vortex) {
      D xy[2];
      for (int x = 0; x < 2; ++x) {
        xy[x] = D(mesh.coords_bndry[((int64_t)rec.y * NF + k) * DIM + x]);
        xy[x].d[DIM + x] = 1.0;
      }
      isentropic_vortex_dual(bp, xy, qRd);
    } else {
      double qg[ND];
      free_stream<DIM>(bp, qg);
      for (int c = 0; c < ND; ++c) qRd[c] = D(qg[c]);
    }
    roe_flux<DIM>(phys, qLd, qRd, nd, Fd);
    for (int c = 0; c < ND; ++c) fbar[c] = -w * pL[c];
  } else {
    const int64_t nb = rec.x;
    const int fN = rec.z & 0xff;
    const int o = (rec.z >> 8) & 0xff;
    const int kR = sbp.nbrperm[k * sbp.norient + o];
    double qR[ND], pR[ND];
    for (int c = 0; c < ND; ++c) { qR[c] = 0.0; pR[c] = 0.0; }
    for (int s = 0; s < ST; ++s) {
      const double r = sbp.interp[s * NF + kR];
      const int node = sbp.perm[s * NFACES + fN];
      for (int c = 0; c < ND; ++c) {
        qR[c] += r * q[(nb * NP + node) * ND + c];
        pR[c] += r * psi[(nb * NP + node) * ND + c];
      }
    }
    if constexpr (FLUX == PDES_FLUX_ESLF) {
      // ESLFFaceIntegral (faceElementIntegrals.jl:611-630): the Lax-Friedrichs
      // entropy penalty at face node k (calcEntropyPenaltyIntegral :199-290)
      // from the interpolated IR entropy variables of both sides, with the
      // face weight inside the flux
      double wl[ND], wr[ND];
      for (int c = 0; c < ND; ++c) { wl[c] = 0.0; wr[c] = 0.0; }
      for (int s = 0; s < ST; ++s) {
        const int nodeL = sbp.perm[s * NFACES + f], nodeR = sbp.perm[s * NFACES + fN];
        double qn[ND], wv[ND];
        for (int c = 0; c < ND; ++c) qn[c] = q[(e * NP + nodeL) * ND + c];
        to_ir_vars<DIM>(phys, qn, wv);
        for (int c = 0; c < ND; ++c) wl[c] += sbp.interp[s * NF + k] * wv[c];
        for (int c = 0; c < ND; ++c) qn[c] = q[(nb * NP + nodeR) * ND + c];
        to_ir_vars<DIM>(phys, qn, wv);
        for (int c = 0; c < ND; ++c) wr[c] += sbp.interp[s * NF + kR] * wv[c];
      }
      double qa[ND], qb[ND];
      ir_to_conservative<DIM>(phys, wl, qa);
      ir_to_conservative<DIM>(phys, wr, qb);
      D qad[ND], dwd[ND];
      for (int c = 0; c < ND; ++c) {
        qad[c] = D(0.5 * (qa[c] + qb[c]));
        dwd[c] = D(wl[c] - wr[c]);
        Fd[c] = D(0.0);
      }
      lf_entropy_apply<DIM>(phys, qad, dwd, nd, w, Fd);
      for (int c = 0; c < ND; ++c) fbar[c] = pR[c] - pL[c];
    } else {
      for (int c = 0; c < ND; ++c) qRd[c] = D(qR[c]);
      if constexpr (FLUX == PDES_FLUX_ROE) roe_flux<DIM>(phys, qLd, qRd, nd, Fd);
      else irslf_flux<DIM>(phys, qLd, qRd, nd, Fd);
      for (int c = 0; c < ND; ++c) fbar[c] = w * (pR[c] - pL[c]);
    }
  }
  double* out = (bnd ? nrm_bndry_bar : nrm_face_bar) + ((int64_t)rec.y * NF + k) * DIM;
  double acc[DIM];
  for (int x = 0; x < DIM; ++x) {
    acc[x] = 0.0;
    for (int c = 0; c < ND; ++c) acc[x] += fbar[c] * Fd[c].d[x];
  }
  if constexpr (FLUX == PDES_FLUX_ESLF) {
    if (!bnd) {
      // the entropy-conservative part (calcECFaceIntegral :58-120): elementL's
      // stencil node a against elementR's stencil node b, with
      // E_ab = sum_k R[a][k] R[b][kR] w_k n[k], subtracts F_IR(q_a, q_b, E_ab)
      // from node a and adds it to node b; F_IR is linear in E, so node k's
      // normal gets R[a][k] R[b][kR] w_k (psi_b - psi_a) . F_IR^x(q_a, q_b)
      const int64_t nb = rec.x;
      const int fN = rec.z & 0xff;
      const int o = (rec.z >> 8) & 0xff;
      const int kR = sbp.nbrperm[k * sbp.norient + o];
      double dir[DIM * DIM];
      for (int x = 0; x < DIM; ++x)
        for (int y = 0; y < DIM; ++y) dir[x * DIM + y] = x == y ? 1.0 : 0.0;
      for (int a = 0; a < ST; ++a) {
        const double ra = sbp.interp[a * NF + k];
        if (ra == 0.0) continue;
        const int nodeL = sbp.perm[a * NFACES + f];
        double qa[ND], za[ND], pa[ND];
        for (int c = 0; c < ND; ++c) {
          qa[c] = q[(e * NP + nodeL) * ND + c];
          pa[c] = psi[(e * NP + nodeL) * ND + c];
        }
        ir_zvars<DIM>(phys, qa, za);
        for (int bb = 0; bb < ST; ++bb) {
          const double coef = ra * sbp.interp[bb * NF + kR] * w;
          if (coef == 0.0) continue;
          const int nodeR = sbp.perm[bb * NFACES + fN];
          double qb[ND], zb[ND], dp[ND], F[DIM * ND];
          for (int c = 0; c < ND; ++c) {
            qb[c] = q[(nb * NP + nodeR) * ND + c];
            dp[c] = psi[(nb * NP + nodeR) * ND + c] - pa[c];
          }
          ir_zvars<DIM>(phys, qb, zb);
          ir_flux_z<DIM, DIM>(phys, za, zb, dir, F);
          for (int x = 0; x < DIM; ++x) {
            double s2 = 0.0;
            for (int c = 0; c < ND; ++c) s2 += dp[c] * F[x * ND + c];
            acc[x] += coef * s2;
          }
        }
      }
    }
  }
  for (int x = 0; x < DIM; ++x) out[x] = acc[x];
  if (bnd) {
    double* oc = coords_bndry_bar + ((int64_t)rec.y * NF + k) * DIM;
    for (int x = 0; x < DIM; ++x) {
      double acc = 0.0;
      for (int c = 0; c < ND; ++c) acc += fbar[c] * Fd[c].d[DIM + x];
      oc[x] = acc;
    }
  }
}

template <int DIM, int NP, int NF, int ST>
static int revm_launch(pdes_ctx* c, const double* psi, double* dxidx_bar, double* nrm_face_bar, double* nrm_bndry_bar,
                       double* coords_bndry_bar) {
  const int64_t nv = (int64_t)c->numEl * NP, nfn = (int64_t)c->numEl * (DIM + 1) * NF;
  if (c->phys.volume_integral_type == 1)
    hipLaunchKernelGGL((revm_volume_weak_kernel<DIM, NP>), dim3((unsigned)ceil_div(nv, 128)), dim3(128), 0,
                       c->stream, c->devmesh(), c->devsbp(), c->pc, c->d_q, psi, dxidx_bar);
  else
    hipLaunchKernelGGL((revm_volume_split_kernel<DIM, NP>), dim3((unsigned)ceil_div(nv, 128)), dim3(128), 0,
                       c->stream, c->devmesh(), c->devsbp(), c->pc, c->d_q, psi, dxidx_bar);
  if (c->phys.face_flux == PDES_FLUX_ROE)
    hipLaunchKernelGGL((revm_face_kernel<DIM, NP, NF, ST, PDES_FLUX_ROE>), dim3((unsigned)ceil_div(nfn, 128)),
                       dim3(128), 0, c->stream, c->devmesh(), c->devsbp(), c->pc, c->d_q, psi, nrm_face_bar,
                       nrm_bndry_bar, coords_bndry_bar);
  else if (c->phys.face_flux == PDES_FLUX_ESLF)
    hipLaunchKernelGGL((revm_face_kernel<DIM, NP, NF, ST, PDES_FLUX_ESLF>), dim3((unsigned)ceil_div(nfn, 128)),
                       dim3(128), 0, c->stream, c->devmesh(), c->devsbp(), c->pc, c->d_q, psi, nrm_face_bar,
                       nrm_bndry_bar, coords_bndry_bar);
  else
    hipLaunchKernelGGL((revm_face_kernel<DIM, NP, NF, ST, PDES_FLUX_IRSLF>), dim3((unsigned)ceil_div(nfn, 128)),
                       dim3(128), 0, c->stream, c->devmesh(), c->devsbp(), c->pc, c->d_q, psi, nrm_face_bar,
                       nrm_bndry_bar, coords_bndry_bar);
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

}  // namespace pdes

using namespace pdes;

extern "C" {

int pdes_eval_residual_revm(pdes_ctx* c, const double* psi, double* dxidx_bar, double* nrm_face_bar,
                            double* nrm_bndry_bar, double* coords_bndry_bar) {
  PDES_CHECK(c && psi, PDES_ERR_ARG, "null argument");
  PDES_CHECK(c->nlist < 0 && c->ninst == 1, PDES_ERR_UNSUPPORTED, "pdes_eval_residual_revm takes a single-solve context");
  const int dim = c->dim, Np = c->Np, Nf = c->Nf;
  const size_t n_dx = (size_t)c->numEl * Np * dim * dim, n_if = (size_t)std::max(1, c->numIface) * Nf * dim,
               n_bf = (size_t)std::max(1, c->numBndry) * Nf * dim;
  double* d = nullptr;
  const size_t total = (size_t)c->numDof + n_dx + n_if + 2 * n_bf;
  PDES_CHECK(hipMalloc((void**)&d, total * sizeof(double)) == hipSuccess, PDES_ERR_NOMEM,
             "device allocation for the metric reverse mode failed");
  double* d_psi = d;
  double* d_dx = d_psi + c->numDof;
  double* d_if = d_dx + n_dx;
  double* d_bn = d_if + n_if;
  double* d_bc = d_bn + n_bf;
  hipError_t he = hipMemcpyAsync(d_psi, psi, c->numDof * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (he == hipSuccess) he = hipMemsetAsync(d_if, 0, (n_if + 2 * n_bf) * sizeof(double), c->stream);
  int st = he == hipSuccess ? PDES_ERR_UNSUPPORTED : PDES_ERR_HIP;
  if (he == hipSuccess) {
#define REVM_LAUNCH(D_, NP_, NF_, ST_, DE_)                                                            \
    if (st == PDES_ERR_UNSUPPORTED && dim == D_ && Np == NP_ && (c->diag_e != 0) == DE_)               \
      st = revm_launch<D_, NP_, NF_, ST_>(c, d_psi, d_dx, d_if, d_bn, d_bc);
    PDES_FOR_EACH_SHAPE(REVM_LAUNCH)
#undef REVM_LAUNCH
    if (st == PDES_ERR_UNSUPPORTED) set_error("no metric reverse-mode kernel for this operator");
  }
  if (st == PDES_OK) {
    struct { double* dst; const double* src; size_t n; } outs[] = {
        {dxidx_bar, d_dx, n_dx},
        {nrm_face_bar, d_if, (size_t)c->numIface * Nf * dim},
        {nrm_bndry_bar, d_bn, (size_t)c->numBndry * Nf * dim},
        {coords_bndry_bar, d_bc, (size_t)c->numBndry * Nf * dim}};
    for (auto& o : outs)
      if (o.dst && o.n && he == hipSuccess)
        he = hipMemcpyAsync(o.dst, o.src, o.n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) {
      set_error("metric reverse mode failed: %s", hipGetErrorString(he));
      st = PDES_ERR_HIP;
    }
  }
  hipStreamSynchronize(c->stream);
  hipFree(d);
  return st;
}

}  // extern "C"
