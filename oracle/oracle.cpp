// oracle.cpp -- TEST INFRASTRUCTURE ONLY: the CPU parity oracle.
//
// An FP64 CPU restatement of the reference's Euler residual path, used by
// tests/, the smoke check and bench.py's CPU-baseline leg to check and time the
// HIP library.  It is never linked into, or called by, the product library.
//
// Restated reference functions (file:line, relative to the reference root):
//   evalResidual            src/solver/euler/euler.jl:111-175
//   dataPrep                src/solver/euler/euler.jl:441-519
//   getEulerFlux            src/solver/euler/euler_funcs.jl:23-58
//   evalVolumeIntegrals     src/solver/euler/euler.jl:628-658 (weakdifferentiate!, trans=true)
//   calcVolumeIntegralsSplitFormLinear  src/solver/euler/euler_funcs.jl:240-288
//   interpolateFace / calcFaceFlux      src/solver/euler/flux.jl:613-641, 37-64
//   calcFaceIntegral_nopre  src/solver/euler/flux.jl:79-125
//   evalFaceIntegrals       src/solver/euler/euler.jl:770-802 (interiorfaceintegrate!, Subtract)
//   interpolateBoundary / getBCFluxes / calcBoundaryFlux  src/solver/euler/bc.jl:162-175, 49-80, 251-284
//   evalBoundaryIntegrals   src/solver/euler/euler.jl:669-690 (boundaryintegrate!, Subtract)
//   checkDensity / checkPressure  src/solver/euler/euler.jl:543-611
//   calcJacobianSparse (complex step, distance-2 coloring)  src/jacobian/jacobian.jl:365-446
//   SparseMatrixCSC(mesh, Tv, INVISCID, face_type)           src/jacobian/sparse.jl:250-495
//   calcHomotopyDiss        src/solver/euler/homotopy.jl:51-249
// The SBP application routines (weakdifferentiate!, interiorfaceinterpolate!,
// interiorfaceintegrate!, boundaryinterpolate!, boundaryintegrate!) live in
// the external SummationByParts.jl; they are restated from their documented
// definitions (res += Q_d^T F_d; u_face = R u; res -= R^T B F).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

#include "../include/pdes.h"
#include "euler_oracle.hpp"

using namespace oracle;

namespace {

Params make_params(const pdes_phys_desc* ph, int dim) {
  Params p;
  p.dim = dim;
  p.gamma = ph->gamma;
  p.gamma_1 = ph->gamma - 1;
  p.R = ph->R_gas;
  p.cv = p.R / p.gamma_1;
  p.Ma = ph->Ma;
  p.aoa = ph->aoa;
  p.rho_free = ph->rho_free;
  p.p_free = ph->p_free;
  p.E_free = ph->p_free / p.gamma_1 + 0.5 * ph->rho_free * ph->Ma * ph->Ma;  // types.jl:252
  p.vortex_Min = ph->vortex_Min;
  return p;
}

// accumulate helper: in |.| mode the magnitude of each contribution is summed
// (gives the floating-point scale of every residual entry)
template <class T> inline void acc(T& dst, const T& v, bool absmode) {
  if (absmode) dst += T(std::abs(v)); else dst += v;
}

// Which terms residual() adds (test selector, see oracle_set_parts): the
// reference's own entropy-stability tests evaluate the terms one at a time
// (test/euler/test_ESS.jl:445-578 with addVolumeIntegrals /
// addBoundaryIntegrals off, test_functionals.jl:160-175).
enum { PART_VOLUME = 1, PART_BOUNDARY = 2, PART_FACE = 4, PART_PENALTY = 8, PART_ALL = 15 };
int g_parts = PART_ALL;
int g_iface_only = -1;   // >= 0: the interface loop visits this interface only
inline bool iface_skipped(int i) { return g_iface_only >= 0 && i != g_iface_only; }

// Threads of the CPU timing baseline (oracle_set_threads).  With one thread
// every loop runs in the reference's order; with more, elements are split
// over OpenMP threads and faces run in groups whose faces touch disjoint
// elements, so no residual entry is updated by two threads at once (the
// order of the contributions to an entry then differs in the last bits).
int g_threads = 1;

// face index groups: one group 0..n-1 when serial, else a greedy colouring of
// the faces by the elements they touch (els(i, a, b): elements of face i, b
// = -1 for a boundary face)
template <class Els>
std::vector<std::vector<int>> face_groups(int n, int numEl, Els els) {
  std::vector<std::vector<int>> groups;
  if (g_threads <= 1) {
    groups.emplace_back(n);
    for (int i = 0; i < n; ++i) groups[0][i] = i;
    return groups;
  }
  std::vector<uint64_t> used(numEl, 0);
  for (int i = 0; i < n; ++i) {
    int a, b;
    els(i, a, b);
    const uint64_t busy = used[a] | (b >= 0 ? used[b] : 0);
    int c = 0;
    while (busy >> c & 1) ++c;
    if ((int)groups.size() <= c) groups.resize(c + 1);
    groups[c].push_back(i);
    used[a] |= uint64_t(1) << c;
    if (b >= 0) used[b] |= uint64_t(1) << c;
  }
  return groups;
}

template <class T>
int residual(const pdes_mesh_desc* m, const pdes_sbp_desc* s, const pdes_phys_desc* ph,
             const T* q, T* res, bool absmode) {
  const int dim = m->dim, nd = dim + 2, Np = s->Np, Nf = s->Nf, nfaces = dim + 1;
  const int numEl = m->numEl;
  Params p = make_params(ph, dim);
  const size_t ndofs = (size_t)numEl * Np * nd;
  for (size_t i = 0; i < ndofs; ++i) res[i] = 0.0;

  // checkDensity / checkPressure (euler.jl:543-611)
  if (ph->check_positivity && !absmode) {
    for (size_t n = 0; n < (size_t)numEl * Np; ++n) {
      const T* qn = q + n * nd;
      if (re(qn[0]) <= 0) return PDES_ERR_PHYS;
      if (re(calcPressure(p, qn)) <= 0) return PDES_ERR_PHYS;
    }
  }
  auto dxidx = [&](int e, int j, int x, int xi) { return m->dxidx[(((size_t)e * Np + j) * dim + x) * dim + xi]; };
  auto Q = [&](int d, int i, int j) { return s->Q[((size_t)d * Np + i) * Np + j]; };

  // ---- volume integrals
  if (!(g_parts & PART_VOLUME)) {
  } else if (ph->volume_integral_type == 1) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int e = 0; e < numEl; ++e) {
      T F[27 * 3 * 5];   // Np * dim * nd
      // getEulerFlux: flux_parametric[:, j, e, d] with nrm = dxidx[d, :, j, e]
      for (int j = 0; j < Np; ++j)
        for (int d = 0; d < dim; ++d) {
          double nrm[3];
          for (int x = 0; x < dim; ++x) nrm[x] = dxidx(e, j, x, d);
          calcEulerFlux(p, q + ((size_t)e * Np + j) * nd, nrm, &F[((size_t)j * dim + d) * nd]);
        }
      // weakdifferentiate!(sbp, d, F_d, res, trans=true): res[:, i] += sum_j Q_d[j, i] F_d[:, j]
      for (int d = 0; d < dim; ++d)
        for (int i = 0; i < Np; ++i)
          for (int j = 0; j < Np; ++j) {
            double qji = Q(d, j, i);
            for (int k = 0; k < nd; ++k)
              acc(res[((size_t)e * Np + i) * nd + k], qji * F[((size_t)j * dim + d) * nd + k], absmode);
          }
    }
  } else {
    // calcVolumeIntegralsSplitFormLinear with IRFlux (multi-direction)
    std::vector<double> S((size_t)Np * Np * dim);
    for (int d = 0; d < dim; ++d)
      for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) S[((size_t)i * Np + j) * dim + d] = 0.5 * (Q(d, i, j) - Q(d, j, i));
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int e = 0; e < numEl; ++e)
      for (int j = 0; j < Np; ++j) {
        T Fd[15];
        double nrmD[9];
        const T* qj = q + ((size_t)e * Np + j) * nd;
        for (int k = 0; k < j; ++k) {
          const T* qk = q + ((size_t)e * Np + k) * nd;
          // nrmD[p, d] = dxidx[d, p, j, i]
          for (int d = 0; d < dim; ++d)
            for (int x = 0; x < dim; ++x) nrmD[x * dim + d] = dxidx(e, j, x, d);
          calcEulerFlux_IR(p, qj, qk, nrmD, dim, Fd);
          for (int d = 0; d < dim; ++d) {
            double sjk = S[((size_t)j * Np + k) * dim + d];
            for (int pp = 0; pp < nd; ++pp) {
              T v = 2 * sjk * Fd[d * nd + pp];
              acc(res[((size_t)e * Np + j) * nd + pp], T(-v), absmode);
              acc(res[((size_t)e * Np + k) * nd + pp], v, absmode);
            }
          }
        }
      }
  }

  // ---- boundary integrals (precompute_boundary_flux: getBCFluxes then boundaryintegrate!)
  if (g_parts & PART_BOUNDARY) {
    const int nb = m->bndry_offsets[m->numBC] - m->bndry_offsets[0];
    std::vector<int> bc_of(nb);
    for (int bc = 0; bc < m->numBC; ++bc)
      for (int b = m->bndry_offsets[bc]; b < m->bndry_offsets[bc + 1]; ++b) bc_of[b - m->bndry_offsets[0]] = bc;
    auto groups = face_groups(nb, numEl, [&](int i, int& a, int& c) {
      a = m->bndryfaces[2 * (i + m->bndry_offsets[0])];
      c = -1;
    });
    for (const auto& grp : groups) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
      for (size_t gi = 0; gi < grp.size(); ++gi) {
        const int b = grp[gi] + m->bndry_offsets[0], bc = bc_of[grp[gi]];
        T qb[6 * 5], fb[6 * 5];   // Nf * nd
        int e = m->bndryfaces[2 * b], f = m->bndryfaces[2 * b + 1];
        // boundaryinterpolate!
        for (int kf = 0; kf < Nf; ++kf)
          for (int k = 0; k < nd; ++k) {
            T v = 0.0;
            for (int st = 0; st < s->stencil; ++st)
              v += s->interp[st * Nf + kf] * q[((size_t)e * Np + s->perm[st * nfaces + f]) * nd + k];
            qb[kf * nd + k] = v;
          }
        for (int kf = 0; kf < Nf; ++kf) {
          const double* coords = m->coords_bndry + ((size_t)b * Nf + kf) * dim;
          const double* nrm = m->nrm_bndry + ((size_t)b * Nf + kf) * dim;
          double qg_d[5];
          if (m->bc_type[bc] == PDES_BC_ISENTROPIC_VORTEX) calcIsentropicVortex(p, coords, qg_d);
          else calcFreeStream(p, qg_d);
          T qg[5];
          for (int k = 0; k < nd; ++k) qg[k] = qg_d[k];
          RoeSolver(p, &qb[kf * nd], qg, nrm, &fb[kf * nd]);
        }
        // boundaryintegrate!(..., Subtract): res[:, perm[st,f]] -= sum_k interp[st,k] wface[k] F[:, k]
        for (int st = 0; st < s->stencil; ++st) {
          int node = s->perm[st * nfaces + f];
          for (int kf = 0; kf < Nf; ++kf) {
            double fac = s->interp[st * Nf + kf] * s->wface[kf];
            for (int k = 0; k < nd; ++k)
              acc(res[((size_t)e * Np + node) * nd + k], T(-fac * fb[kf * nd + k]), absmode);
          }
        }
      }
    }
  }

  // ---- face integrals
  if (ph->face_flux == PDES_FLUX_ESLF) {
    // face_integral_type 2 with ESLFFaceIntegral on a dense face
    // (getFaceElementIntegral flux.jl:132-160 -> calcESFaceIntegral
    // faceElementIntegrals.jl:151-170): calcECFaceIntegral (:58-120) plus
    // calcEntropyPenaltyIntegral with the LF kernel (:199-290, :455-470)
    const int st_n = s->stencil;
    double nrmD[9];
    for (int a = 0; a < dim * dim; ++a) nrmD[a] = 0.0;
    for (int d = 0; d < dim; ++d) nrmD[d * dim + d] = 1.0;
    auto groups = face_groups(m->numInterfaces, numEl, [&](int i, int& a, int& c) {
      a = m->interfaces[5 * i];
      c = m->interfaces[5 * i + 1];
    });
    for (const auto& grp : groups) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (size_t gi = 0; gi < grp.size(); ++gi) {
      const int i = grp[gi];
      if (iface_skipped(i)) continue;
      T fluxD[15], wL[15 * 5], wR[15 * 5], wLi[5], wRi[5], qLi[5], qRi[5], qavg[5], dw[5], A0[25], flux[5];
      const int* it = m->interfaces + 5 * i;
      int eL = it[0], eR = it[1], fL = it[2], fR = it[3], o = it[4];
      const double* nrm_xy = m->nrm_face + (size_t)i * Nf * dim;   // [k][d]
      // entropy conservative part
      for (int ii = 0; ii < st_n && (g_parts & PART_FACE); ++ii) {
        int p_i = s->perm[ii * nfaces + fL];
        const T* qi = q + ((size_t)eL * Np + p_i) * nd;
        for (int jj = 0; jj < st_n; ++jj) {
          int p_j = s->perm[jj * nfaces + fR];
          const T* qj = q + ((size_t)eR * Np + p_j) * nd;
          calcEulerFlux_IR(p, qi, qj, nrmD, dim, fluxD);
          for (int d = 0; d < dim; ++d) {
            double Eij = 0.0;
            for (int k = 0; k < Nf; ++k) {
              int kR = s->nbrperm[k * s->norient + o];
              Eij += s->interp[ii * Nf + k] * s->interp[jj * Nf + kR] * s->wface[k] * nrm_xy[k * dim + d];
            }
            for (int pp = 0; pp < nd; ++pp) {
              acc(res[((size_t)eL * Np + p_i) * nd + pp], T(-Eij * fluxD[d * nd + pp]), absmode);
              acc(res[((size_t)eR * Np + p_j) * nd + pp], T(Eij * fluxD[d * nd + pp]), absmode);
            }
          }
        }
      }
      // entropy penalty: IR variables at the stencil nodes
      for (int ii = 0; ii < st_n; ++ii) {
        convertToIR(p, q + ((size_t)eL * Np + s->perm[ii * nfaces + fL]) * nd, &wL[ii * nd]);
        convertToIR(p, q + ((size_t)eR * Np + s->perm[ii * nfaces + fR]) * nd, &wR[ii * nd]);
      }
      for (int k = 0; k < Nf && (g_parts & PART_PENALTY); ++k) {
        int nk = s->nbrperm[k * s->norient + o];
        const double* dir = nrm_xy + (size_t)k * dim;
        for (int c = 0; c < nd; ++c) { wLi[c] = 0.0; wRi[c] = 0.0; }
        for (int jj = 0; jj < st_n; ++jj) {
          double iL = s->interp[jj * Nf + k], iR = s->interp[jj * Nf + nk];
          for (int c = 0; c < nd; ++c) { wLi[c] += iL * wL[jj * nd + c]; wRi[c] += iR * wR[jj * nd + c]; }
        }
        convertToConservativeFromIR(p, wLi, qLi);
        convertToConservativeFromIR(p, wRi, qRi);
        for (int c = 0; c < nd; ++c) { qavg[c] = 0.5 * (qLi[c] + qRi[c]); dw[c] = wLi[c] - wRi[c]; }
        getIRA0(p, qavg, A0);
        T lambda_max = getLambdaMax(p, qavg, dir);
        for (int r = 0; r < nd; ++r) {
          T a = 0.0;
          for (int c = 0; c < nd; ++c) a += A0[c * nd + r] * dw[c];
          flux[r] = a * lambda_max * s->wface[k];
        }
        for (int jj = 0; jj < st_n; ++jj) {
          int jL = s->perm[jj * nfaces + fL], jR = s->perm[jj * nfaces + fR];
          for (int c = 0; c < nd; ++c) {
            acc(res[((size_t)eL * Np + jL) * nd + c], T(-s->interp[jj * Nf + k] * flux[c]), absmode);
            acc(res[((size_t)eR * Np + jR) * nd + c], T(s->interp[jj * Nf + nk] * flux[c]), absmode);
          }
        }
      }
    }
    }
  } else if (g_parts & PART_FACE) {
    auto groups = face_groups(m->numInterfaces, numEl, [&](int i, int& a, int& c) {
      a = m->interfaces[5 * i];
      c = m->interfaces[5 * i + 1];
    });
    for (const auto& grp : groups) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (size_t gi = 0; gi < grp.size(); ++gi) {
      const int i = grp[gi];
      if (iface_skipped(i)) continue;
      T qL[6 * 5], qR[6 * 5], ff[6 * 5];   // Nf * nd
      const int* it = m->interfaces + 5 * i;
      int eL = it[0], eR = it[1], fL = it[2], fR = it[3], o = it[4];
      // interiorfaceinterpolate!
      for (int kf = 0; kf < Nf; ++kf) {
        int kR = s->nbrperm[kf * s->norient + o];
        for (int k = 0; k < nd; ++k) {
          T vl = 0.0, vr = 0.0;
          for (int st = 0; st < s->stencil; ++st) {
            vl += s->interp[st * Nf + kf] * q[((size_t)eL * Np + s->perm[st * nfaces + fL]) * nd + k];
            vr += s->interp[st * Nf + kR] * q[((size_t)eR * Np + s->perm[st * nfaces + fR]) * nd + k];
          }
          qL[kf * nd + k] = vl;
          qR[kf * nd + k] = vr;
        }
      }
      for (int kf = 0; kf < Nf; ++kf) {
        const double* nrm = m->nrm_face + ((size_t)i * Nf + kf) * dim;
        if (ph->face_flux == PDES_FLUX_ROE) RoeSolver(p, &qL[kf * nd], &qR[kf * nd], nrm, &ff[kf * nd]);
        else calcEulerFlux_IRSLF(p, &qL[kf * nd], &qR[kf * nd], nrm, &ff[kf * nd]);
      }
      // interiorfaceintegrate!(..., Subtract)
      for (int st = 0; st < s->stencil; ++st) {
        int nodeL = s->perm[st * nfaces + fL];
        int nodeR = s->perm[st * nfaces + fR];
        for (int kf = 0; kf < Nf; ++kf) {
          int kR = s->nbrperm[kf * s->norient + o];
          double facL = s->interp[st * Nf + kf] * s->wface[kf];
          double facR = s->interp[st * Nf + kR] * s->wface[kR];
          for (int k = 0; k < nd; ++k) {
            acc(res[((size_t)eL * Np + nodeL) * nd + k], T(-facL * ff[kf * nd + k]), absmode);
            acc(res[((size_t)eR * Np + nodeR) * nd + k], T(facR * ff[kf * nd + k]), absmode);
          }
        }
      }
    }
    }
  }
  return PDES_OK;
}

// element adjacency through interfaces
// calcHomotopyDiss (src/solver/euler/homotopy.jl:51-249), the first-order
// dissipation of the homotopy globalisation (bndry: the boundary term of
// homotopy_addBoundaryIntegrals, default false, src/input/read_input.jl:438):
//   volume, per reference direction d: t1 = H^-1 Q_d q (differentiateElement!),
//     t2_j = lambda_max(q_j, dxidx[d, :, j]) t1_j, res += Q_d^T t2
//     (weakDifferentiateElement!, trans = true);
//   interfaces: flux_k = lambda_max(avg(qL_k, qR_k), nrm_k) (qL_k - qR_k) / 2 at
//     each face node (getLambdaMaxSimple, IR_stab.jl:310-325), resL += R_L^T B
//     flux, resR -= R_R^T B flux (interiorFaceIntegrate!, Add);
//   boundary faces (bndry only): the same flux between the face state and the
//     free stream (calcFreeStream, whatever the BC), res += R^T B flux
//     (boundaryFaceIntegrate!, Add);
//   then res = -res.
template <class T>
void homotopy_diss(const pdes_mesh_desc* m, const pdes_sbp_desc* s, const pdes_phys_desc* ph, const T* q,
                   T* res, bool bndry) {
  const int dim = m->dim, nd = dim + 2, Np = s->Np, Nf = s->Nf, nfaces = dim + 1;
  Params p = make_params(ph, dim);
  const size_t n = (size_t)m->numEl * Np * nd;
  for (size_t i = 0; i < n; ++i) res[i] = 0.0;
  auto Q = [&](int d, int i, int j) { return s->Q[((size_t)d * Np + i) * Np + j]; };
  std::vector<T> t2((size_t)Np * nd);
  for (int e = 0; e < m->numEl; ++e) {
    const T* qe = q + (size_t)e * Np * nd;
    T* re_ = res + (size_t)e * Np * nd;
    for (int d1 = 0; d1 < dim; ++d1) {
      for (int j = 0; j < Np; ++j) {
        double dir[3];
        for (int x = 0; x < dim; ++x) dir[x] = m->dxidx[(((size_t)e * Np + j) * dim + x) * dim + d1];
        const T lam = getLambdaMax(p, qe + (size_t)j * nd, dir);
        for (int k = 0; k < nd; ++k) {
          T t1 = 0.0;
          for (int jj = 0; jj < Np; ++jj) t1 += Q(d1, j, jj) * qe[(size_t)jj * nd + k];
          t2[(size_t)j * nd + k] = lam * (t1 / s->w[j]);
        }
      }
      for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) {
          const double qji = Q(d1, j, i);
          for (int k = 0; k < nd; ++k) re_[(size_t)i * nd + k] += qji * t2[(size_t)j * nd + k];
        }
    }
  }
  std::vector<T> qL((size_t)Nf * nd), qR((size_t)Nf * nd), ff((size_t)Nf * nd);
  for (int i = 0; i < m->numInterfaces; ++i) {
    const int* it = m->interfaces + 5 * i;
    const int eL = it[0], eR = it[1], fL = it[2], fR = it[3], o = it[4];
    for (int kf = 0; kf < Nf; ++kf) {
      const int kR = s->nbrperm[kf * s->norient + o];
      for (int k = 0; k < nd; ++k) {
        T vl = 0.0, vr = 0.0;
        for (int st = 0; st < s->stencil; ++st) {
          vl += s->interp[st * Nf + kf] * q[((size_t)eL * Np + s->perm[st * nfaces + fL]) * nd + k];
          vr += s->interp[st * Nf + kR] * q[((size_t)eR * Np + s->perm[st * nfaces + fR]) * nd + k];
        }
        qL[kf * nd + k] = vl;
        qR[kf * nd + k] = vr;
      }
    }
    for (int kf = 0; kf < Nf; ++kf) {
      T qa[5];
      for (int k = 0; k < nd; ++k) qa[k] = 0.5 * (qL[kf * nd + k] + qR[kf * nd + k]);
      const T lam = getLambdaMax(p, qa, m->nrm_face + ((size_t)i * Nf + kf) * dim);
      for (int k = 0; k < nd; ++k) ff[kf * nd + k] = 0.5 * lam * (qL[kf * nd + k] - qR[kf * nd + k]);
    }
    for (int st = 0; st < s->stencil; ++st) {
      const int nodeL = s->perm[st * nfaces + fL], nodeR = s->perm[st * nfaces + fR];
      for (int kf = 0; kf < Nf; ++kf) {
        const int kR = s->nbrperm[kf * s->norient + o];
        const double facL = s->interp[st * Nf + kf] * s->wface[kf];
        const double facR = s->interp[st * Nf + kR] * s->wface[kR];
        for (int k = 0; k < nd; ++k) {
          res[((size_t)eL * Np + nodeL) * nd + k] += facL * ff[kf * nd + k];
          res[((size_t)eR * Np + nodeR) * nd + k] -= facR * ff[kf * nd + k];
        }
      }
    }
  }
  if (bndry) {
    double qg_d[5];
    calcFreeStream(p, qg_d);
    for (int bi = 0; bi < m->numBoundaryFaces; ++bi) {
      const int e = m->bndryfaces[2 * bi], f = m->bndryfaces[2 * bi + 1];
      for (int kf = 0; kf < Nf; ++kf) {
        T qa[5];
        for (int k = 0; k < nd; ++k) {
          T vl = 0.0;
          for (int st = 0; st < s->stencil; ++st)
            vl += s->interp[st * Nf + kf] * q[((size_t)e * Np + s->perm[st * nfaces + f]) * nd + k];
          qL[kf * nd + k] = vl;
          qa[k] = 0.5 * (vl + qg_d[k]);
        }
        const T lam = getLambdaMax(p, qa, m->nrm_bndry + ((size_t)bi * Nf + kf) * dim);
        for (int k = 0; k < nd; ++k) ff[kf * nd + k] = 0.5 * lam * (qL[kf * nd + k] - qg_d[k]);
      }
      for (int st = 0; st < s->stencil; ++st) {
        const int node = s->perm[st * nfaces + f];
        for (int kf = 0; kf < Nf; ++kf) {
          const double fac = s->interp[st * Nf + kf] * s->wface[kf];
          for (int k = 0; k < nd; ++k) res[((size_t)e * Np + node) * nd + k] += fac * ff[kf * nd + k];
        }
      }
    }
  }
  for (size_t i = 0; i < n; ++i) res[i] = -res[i];
}

std::vector<std::vector<int>> element_neighbors(const pdes_mesh_desc* m) {
  std::vector<std::vector<int>> nb(m->numEl);
  for (int i = 0; i < m->numInterfaces; ++i) {
    nb[m->interfaces[5 * i]].push_back(m->interfaces[5 * i + 1]);
    nb[m->interfaces[5 * i + 1]].push_back(m->interfaces[5 * i]);
  }
  return nb;
}

}  // namespace

extern "C" {

// Test selector: the terms the residual adds from now on (bits: 1 volume,
// 2 boundary, 4 interface flux / entropy-conservative face integral,
// 8 entropy penalty of the ESLF face integral; 15 = the full residual) and
// the one interface the interface loop visits (-1: all).
void oracle_set_parts(int parts, int iface_only) {
  g_parts = parts;
  g_iface_only = iface_only;
}

// Threads of residual evaluations from now on (1: serial, the reference's
// summation order); returns the number of host cores OpenMP sees.
int oracle_set_threads(int threads) {
  g_threads = threads < 1 ? 1 : threads;
  return omp_get_num_procs();
}

int oracle_residual(const pdes_mesh_desc* m, const pdes_sbp_desc* s, const pdes_phys_desc* ph,
                    const double* q, double* res) {
  return residual<double>(m, s, ph, q, res, false);
}

// per-entry floating-point scale: the residual with every contribution replaced by its magnitude
int oracle_residual_scale(const pdes_mesh_desc* m, const pdes_sbp_desc* s, const pdes_phys_desc* ph,
                          const double* q, double* res) {
  return residual<double>(m, s, ph, q, res, true);
}

int oracle_residual_complex(const pdes_mesh_desc* m, const pdes_sbp_desc* s, const pdes_phys_desc* ph,
                            const double* q_re, const double* q_im, double* res_re, double* res_im) {
  size_t n = (size_t)m->numEl * s->Np * (m->dim + 2);
  std::vector<cplx> q(n), r(n);
  for (size_t i = 0; i < n; ++i) q[i] = cplx(q_re[i], q_im[i]);
  int st = residual<cplx>(m, s, ph, q.data(), r.data(), false);
  for (size_t i = 0; i < n; ++i) { res_re[i] = r[i].real(); res_im[i] = r[i].imag(); }
  return st;
}

// evaldRdqProduct (src/interface2.jl:454-498): b = Im R(q + i h v) / h
int oracle_jv_complex_step(const pdes_mesh_desc* m, const pdes_sbp_desc* s, const pdes_phys_desc* ph,
                           const double* q, const double* v, double h, double* b) {
  size_t n = (size_t)m->numEl * s->Np * (m->dim + 2);
  std::vector<cplx> qc(n), r(n);
  for (size_t i = 0; i < n; ++i) qc[i] = cplx(q[i], h * v[i]);
  int st = residual<cplx>(m, s, ph, qc.data(), r.data(), false);
  for (size_t i = 0; i < n; ++i) b[i] = r[i].imag() / h;
  return st;
}

int oracle_homotopy(const pdes_mesh_desc* m, const pdes_sbp_desc* s, const pdes_phys_desc* ph, const double* q,
                    int bndry, double* res) {
  homotopy_diss<double>(m, s, ph, q, res, bndry != 0);
  return PDES_OK;
}

// directional derivative of calcHomotopyDiss by the complex step
int oracle_homotopy_jv(const pdes_mesh_desc* m, const pdes_sbp_desc* s, const pdes_phys_desc* ph,
                       const double* q, const double* v, double h, int bndry, double* b) {
  size_t n = (size_t)m->numEl * s->Np * (m->dim + 2);
  std::vector<cplx> qc(n), r(n);
  for (size_t i = 0; i < n; ++i) qc[i] = cplx(q[i], h * v[i]);
  homotopy_diss<cplx>(m, s, ph, qc.data(), r.data(), bndry != 0);
  for (size_t i = 0; i < n; ++i) b[i] = r[i].imag() / h;
  return PDES_OK;
}

// Block sparsity of SparseMatrixCSC(mesh, Tv, INVISCID, face_type) (sparse.jl:250-495),
// node level (block size ndof).  First call with colind == NULL returns
// rowptr only (nrows+1 entries); second call fills sorted column lists.
int oracle_sparsity_bsr(const pdes_mesh_desc* m, const pdes_sbp_desc* s, int64_t* rowptr, int32_t* colind) {
  const int Np = s->Np, Nf = s->Nf, nfaces = m->dim + 1;
  const int64_t nrows = (int64_t)m->numEl * Np;
  std::vector<std::vector<int32_t>> rows(nrows);
  // volume terms: every node of an element couples to all nodes of that element
  for (int e = 0; e < m->numEl; ++e)
    for (int j = 0; j < Np; ++j)
      for (int pp = 0; pp < Np; ++pp) rows[(int64_t)e * Np + j].push_back(e * Np + pp);
  // interface terms
  for (int i = 0; i < m->numInterfaces; ++i) {
    const int* it = m->interfaces + 5 * i;
    int eL = it[0], eR = it[1], fL = it[2], fR = it[3], o = it[4];
    if (!s->diag_e) {  // face_type 1: all permL <-> all permR
      for (int j = 0; j < s->stencil; ++j) {
        int nL = s->perm[j * nfaces + fL], nR = s->perm[j * nfaces + fR];
        for (int pp = 0; pp < s->stencil; ++pp) {
          rows[(int64_t)eL * Np + nL].push_back(eR * Np + s->perm[pp * nfaces + fR]);
          rows[(int64_t)eR * Np + nR].push_back(eL * Np + s->perm[pp * nfaces + fL]);
        }
      }
    } else {  // face_type 2: node <-> nbrperm partner
      for (int j = 0; j < Nf; ++j) {
        int nodeL = s->perm[j * nfaces + fL];
        int pnbr = s->nbrperm[j * s->norient + o];
        int nodeR = s->perm[pnbr * nfaces + fR];
        rows[(int64_t)eL * Np + nodeL].push_back(eR * Np + nodeR);
        rows[(int64_t)eR * Np + nodeR].push_back(eL * Np + nodeL);
      }
    }
  }
  rowptr[0] = 0;
  for (int64_t r = 0; r < nrows; ++r) rowptr[r + 1] = rowptr[r] + (int64_t)rows[r].size();
  if (colind) {
    for (int64_t r = 0; r < nrows; ++r) {
      std::sort(rows[r].begin(), rows[r].end());
      std::memcpy(colind + rowptr[r], rows[r].data(), rows[r].size() * sizeof(int32_t));
    }
  }
  return PDES_OK;
}

// Distance-2 element coloring (greedy, element order): no two elements of a
// color share an element-face neighbour, so one complex residual per
// (color, node, dof) separates every column.
int oracle_color_elements(const pdes_mesh_desc* m, int32_t* color, int32_t* ncolors) {
  auto nb = element_neighbors(m);
  int nc = 0;
  std::vector<int> mark;
  for (int e = 0; e < m->numEl; ++e) {
    mark.assign(nc + 1, 0);
    auto visit = [&](int x) { if (x < e) mark[color[x]] = 1; };
    for (int a : nb[e]) {
      visit(a);
      for (int b : nb[a]) if (b != e) visit(b);
    }
    int c = 0;
    while (c < nc && mark[c]) ++c;
    color[e] = c;
    if (c == nc) ++nc;
  }
  *ncolors = nc;
  return PDES_OK;
}

// pertNeighborEls (docs/src/interfaces.md:566-569): [numEl][numColors], the
// element whose perturbation reaches element i when colour c is perturbed
// (i itself or one of its face neighbours), -1 for none.
int oracle_pert_neighbor_els(const pdes_mesh_desc* m, const int32_t* color, int32_t ncolors, int32_t* pert) {
  auto nb = element_neighbors(m);
  std::fill(pert, pert + (size_t)m->numEl * ncolors, -1);
  for (int e = 0; e < m->numEl; ++e) {
    pert[(size_t)e * ncolors + color[e]] = e;
    for (int a : nb[e]) {
      int32_t& slot = pert[(size_t)e * ncolors + color[a]];
      if (slot >= 0 && slot != a) return PDES_ERR_ARG;   // not a distance-2 colouring
      slot = a;
    }
  }
  return PDES_OK;
}

// calcJacobianSparse (jacobian.jl:365-446): complex-step Jacobian through
// element coloring, values scattered into the BSR pattern from
// oracle_sparsity_bsr (block = [ndof][ndof] row-major, row = residual dof).
int oracle_jacobian_cs(const pdes_mesh_desc* m, const pdes_sbp_desc* s, const pdes_phys_desc* ph,
                       const double* q, const int64_t* rowptr, const int32_t* colind, double* val) {
  const int nd = m->dim + 2, Np = s->Np;
  const size_t n = (size_t)m->numEl * Np * nd;
  const double h = 1e-20;
  std::vector<int32_t> color(m->numEl);
  int32_t nc = 0;
  oracle_color_elements(m, color.data(), &nc);
  auto nb = element_neighbors(m);
  std::fill(val, val + (size_t)rowptr[(size_t)m->numEl * Np] * nd * nd, 0.0);
  std::vector<cplx> qc(n), r(n);
  for (int c = 0; c < nc; ++c)
    for (int j = 0; j < Np; ++j)
      for (int k = 0; k < nd; ++k) {
        for (size_t i = 0; i < n; ++i) qc[i] = q[i];
        for (int e = 0; e < m->numEl; ++e)
          if (color[e] == c) qc[((size_t)e * Np + j) * nd + k] += cplx(0, h);
        int st = residual<cplx>(m, s, ph, qc.data(), r.data(), false);
        if (st) return st;
        for (int e = 0; e < m->numEl; ++e) {
          if (color[e] != c) continue;
          int32_t col = e * Np + j;
          std::vector<int> rowsel = nb[e];
          rowsel.push_back(e);
          for (int er : rowsel)
            for (int i = 0; i < Np; ++i) {
              int64_t row = (int64_t)er * Np + i;
              const int32_t* b = std::lower_bound(colind + rowptr[row], colind + rowptr[row + 1], col);
              bool present = (b != colind + rowptr[row + 1] && *b == col);
              for (int kr = 0; kr < nd; ++kr) {
                double d = r[((size_t)er * Np + i) * nd + kr].imag() / h;
                if (!present) {
                  if (d != 0.0) return PDES_ERR_ARG;   // entry outside the sparsity pattern
                  continue;
                }
                val[((size_t)(b - colind) * nd + kr) * nd + k] = d;
              }
            }
        }
      }
  return PDES_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// point functions, for the known-answer tests of the reference's test suite
// ---------------------------------------------------------------------------
namespace {
Params point_params(int dim, const pdes_phys_desc* ph) { return make_params(ph, dim); }
}

extern "C" {

double oracle_pressure(int dim, const pdes_phys_desc* ph, const double* q) {
  Params p = point_params(dim, ph);
  return calcPressure(p, q);
}

void oracle_euler_flux(int dim, const pdes_phys_desc* ph, const double* q, const double* dir, double* F) {
  Params p = point_params(dim, ph);
  calcEulerFlux(p, q, dir, F);
}

void oracle_roe(int dim, const pdes_phys_desc* ph, const double* qL, const double* qR, const double* nrm,
                double* F) {
  Params p = point_params(dim, ph);
  RoeSolver(p, qL, qR, nrm, F);
}

void oracle_roe_complex(int dim, const pdes_phys_desc* ph, const double* qL_re, const double* qL_im,
                        const double* qR_re, const double* qR_im, const double* nrm, double* F_re, double* F_im) {
  Params p = point_params(dim, ph);
  cplx qL[5], qR[5], F[5];
  for (int i = 0; i < dim + 2; ++i) { qL[i] = cplx(qL_re[i], qL_im[i]); qR[i] = cplx(qR_re[i], qR_im[i]); }
  RoeSolver(p, qL, qR, nrm, F);
  for (int i = 0; i < dim + 2; ++i) { F_re[i] = F[i].real(); F_im[i] = F[i].imag(); }
}

// dir: [x][ndir]
void oracle_ir_flux(int dim, const pdes_phys_desc* ph, const double* qL, const double* qR, const double* dir,
                    int ndir, double* F) {
  Params p = point_params(dim, ph);
  calcEulerFlux_IR(p, qL, qR, dir, ndir, F);
}

void oracle_irslf_flux(int dim, const pdes_phys_desc* ph, const double* qL, const double* qR,
                       const double* nrm, double* F) {
  Params p = point_params(dim, ph);
  calcEulerFlux_IRSLF(p, qL, qR, nrm, F);
}

double oracle_logavg(double aL, double aR) { return logavg(aL, aR); }

void oracle_logavg_complex(double aL_re, double aL_im, double aR_re, double aR_im, double* out) {
  cplx r = logavg(cplx(aL_re, aL_im), cplx(aR_re, aR_im));
  out[0] = r.real();
  out[1] = r.imag();
}

void oracle_vortex(const pdes_phys_desc* ph, const double* coords, double* sol) {
  Params p = point_params(2, ph);
  calcIsentropicVortex(p, coords, sol);
}

void oracle_ir_vars(int dim, const pdes_phys_desc* ph, const double* q, double* v) {
  Params p = point_params(dim, ph);
  convertToIR(p, q, v);
}

}  // extern "C"
