// jacobian_kernels.hpp -- explicit Jacobian blocks of the DG-SBP Euler residual.
//
// evalJacobian (src/solver/euler/evalJacobian.jl:6-72) computes, per element,
// the volume block dR_e/dq_e (calcVolumeIntegrals_nopre_diff,
// euler_funcs_diff.jl:54-104, or the split-form calcVolumeIntegralsSplitForm_diff
// :330-430) and per face the blocks jacLL/jacLR/jacRL/jacRR
// (calcFaceIntegral_nopre_diff, flux_diff.jl:6-88) and per boundary face the
// block dR_e/dq_e (calcBoundaryFlux_nopre_diff, bc_diff.jl:124-212), and
// scatters them with ADD_VALUES into a block matrix (assembleElement /
// assembleInterface / assembleBoundary, src/jacobian/jacobian.jl:822-1759).
//
// Here one workgroup owns one element and writes every block of that
// element's rows (its own Np x Np block plus the blocks coupling it to each
// neighbour) directly into the BSR value array: each block is written by
// exactly one workgroup, no atomics and no assembly pass.  Point derivatives
// come from forward-mode dual numbers with 2*ndof directions (dF/dqL, dF/dqR),
// the analytic counterpart of the reference's hand-differentiated *_diff
// functions.
#pragma once
#include "residual_kernels.hpp"

namespace pdes {

template <class Op, int VOL, int FLUX>
struct JacSmem {
  static constexpr int NP = Op::NP, ND = Op::ND, DIM = Op::DIM, NF = Op::NF, NFACES = Op::NFACES;
  double Q[DIM][NP][NP];
  double interp[Op::ST][NF];
  int32_t perm[Op::ST][NFACES];
  int32_t nodepos[NP][NFACES];
  double wface[NF];
  double q[NP][ND];
  double A[(VOL == VOL_WEAK) ? NP * DIM : 1][ND][ND];   // dF_d/dq at node j
  double dFo[NFACES][NF][ND][ND];   // dF_k/dq_own   (k: elementL face node numbering)
  double dFn[NFACES][NF][ND][ND];   // dF_k/dq_nbr
  ElFace rec[NFACES];
  int32_t kown[NFACES][NF];         // own face node of slot k
  int32_t knbr[NFACES][NF];         // neighbour face node of slot k
  double sgn[NFACES];
};

template <class BT> __device__ __forceinline__ BT to_block(double x) { return (BT)x; }

// position of block column `col` in BSR row `row` (sorted colind), -1 if absent
__device__ __forceinline__ int64_t bsr_find(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
                                            int64_t row, int32_t col) {
  int64_t lo = rowptr[row], hi = rowptr[row + 1] - 1;
  while (lo <= hi) {
    int64_t mid = (lo + hi) >> 1;
    int32_t c = colind[mid];
    if (c == col) return mid;
    if (c < col) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

// dw/dq of the IR entropy variables at state q (convertToIR_ differentiated)
template <int DIM>
__device__ __forceinline__ void ir_vars_jacobian(const PhysConst& p, const double* q, double (*W)[DIM + 2]) {
  constexpr int ND = DIM + 2;
  dual<ND> qd[ND], w[ND];
#pragma unroll
  for (int c = 0; c < ND; ++c) { qd[c] = dual<ND>(q[c]); qd[c].d[c] = 1.0; }
  to_ir_vars<DIM>(p, qd, w);
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int c = 0; c < ND; ++c) W[a][c] = w[a].d[c];
}

// E_tu[d] = sum_k R[iL][k] R[jR][kR(k)] B[k] n_d[k] of the entropy-conservative
// face integral (calcECFaceIntegral, faceElementIntegrals.jl:58-120), with
// (iL, jR) = (t, u) when this element is elementL and (u, t) otherwise
template <class Op, class Sm>
__device__ __forceinline__ void ec_weights(const Sm& sm, const DevSBP& sbp, const DevMesh& mesh, const ElFace& r,
                                           int t, int u, double* E) {
  constexpr int DIM = Op::DIM, NF = Op::NF;
  const int o = (r.z >> 8) & 0xff, side = (r.z >> 16) & 1;
  const int iL = side == 0 ? t : u, jR = side == 0 ? u : t;
  const double* nr = mesh.nrm_face + (int64_t)r.y * NF * DIM;
#pragma unroll
  for (int d = 0; d < DIM; ++d) E[d] = 0.0;
  for (int k = 0; k < NF; ++k) {
    const int kRk = sbp.nbrperm[k * sbp.norient + o];
    const double base = sm.interp[iL][k] * sm.interp[jR][kRk] * sm.wface[k];
#pragma unroll
    for (int d = 0; d < DIM; ++d) E[d] += base * nr[k * DIM + d];
  }
}

// d/dq of F_IR(q_L, q_R) . E for the pair (own stencil node t, neighbour
// stencil node u) of face f, with respect to the own state (wrt_own) or the
// neighbour's; the sum over directions of E_d F_d is the flux in direction E
template <class Op, class Sm>
__device__ __forceinline__ void ec_pair_derivative(const Sm& sm, const DevSBP& sbp, const DevMesh& mesh,
                                                   const PhysConst& phys, const double* qin, int f, int t, int u,
                                                   bool wrt_own, double (*D)[Op::ND]) {
  constexpr int DIM = Op::DIM, NP = Op::NP, ND = Op::ND;
  const ElFace r = sm.rec[f];
  const int fN = r.z & 0xff, side = (r.z >> 16) & 1;
  const double* qn = qin + ((int64_t)r.x * NP + sm.perm[u][fN]) * ND;
  dual<ND> qo[ND], qb[ND], zo[ND], zb[ND], zL[ND], zR[ND], F[ND];
#pragma unroll
  for (int c = 0; c < ND; ++c) {
    qo[c] = dual<ND>(sm.q[sm.perm[t][f]][c]);
    qb[c] = dual<ND>(qn[c]);
    if (wrt_own) qo[c].d[c] = 1.0;
    else qb[c].d[c] = 1.0;
  }
  ir_zvars<DIM>(phys, qo, zo);
  ir_zvars<DIM>(phys, qb, zb);
#pragma unroll
  for (int c = 0; c < ND; ++c) {
    zL[c] = side == 0 ? zo[c] : zb[c];
    zR[c] = side == 0 ? zb[c] : zo[c];
  }
  double E[DIM];
  ec_weights<Op>(sm, sbp, mesh, r, t, u, E);
  ir_flux_z<DIM, 1>(phys, zL, zR, E, F);
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int c = 0; c < ND; ++c) D[a][c] = F[a].d[c];
}

// ESLF interior face f, own block (i, j): the entropy penalty through the
// interpolated own entropy variables, sum_k sg R[si][ko] R[sj][ko] dP_k/dw_own
// times dw/dq(q_j); for i == j also the entropy-conservative pairs of node si
template <class Op, class Sm>
__device__ __forceinline__ void eslf_own_block(const Sm& sm, const DevSBP& sbp, const DevMesh& mesh,
                                               const PhysConst& phys, const double* qin, int f, int i, int j,
                                               int si, int sj, double (*B)[Op::ND]) {
  constexpr int DIM = Op::DIM, NF = Op::NF, ST = Op::ST, ND = Op::ND;
  const double sg = sm.sgn[f];
  double M[ND][ND], W[ND][ND];
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int c = 0; c < ND; ++c) M[a][c] = 0.0;
  for (int k = 0; k < NF; ++k) {
    const int m = sm.kown[f][k];
    const double wgt = sg * sm.interp[si][m] * sm.interp[sj][m];
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int c = 0; c < ND; ++c) M[a][c] += wgt * sm.dFo[f][k][a][c];
  }
  ir_vars_jacobian<DIM>(phys, sm.q[j], W);
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int c = 0; c < ND; ++c) {
      double acc = 0.0;
#pragma unroll
      for (int b = 0; b < ND; ++b) acc += M[a][b] * W[b][c];
      B[a][c] += acc;
    }
  if (i != j) return;
  for (int u = 0; u < ST; ++u) {
    double D[ND][ND];
    ec_pair_derivative<Op>(sm, sbp, mesh, phys, qin, f, si, u, true, D);
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int c = 0; c < ND; ++c) B[a][c] += sg * D[a][c];
  }
}

// ESLF interior face f, neighbour block (row node with stencil position si,
// neighbour node jn with stencil position sj on the neighbour's face)
template <class Op, class Sm>
__device__ __forceinline__ void eslf_nbr_block(const Sm& sm, const DevSBP& sbp, const DevMesh& mesh,
                                               const PhysConst& phys, const double* qin, int f, int jn,
                                               int si, int sj, double (*B)[Op::ND]) {
  constexpr int DIM = Op::DIM, NP = Op::NP, NF = Op::NF, ND = Op::ND;
  const double sg = sm.sgn[f];
  double M[ND][ND], W[ND][ND];
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int c = 0; c < ND; ++c) M[a][c] = 0.0;
  for (int k = 0; k < NF; ++k) {
    const double wgt = sg * sm.interp[si][sm.kown[f][k]] * sm.interp[sj][sm.knbr[f][k]];
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int c = 0; c < ND; ++c) M[a][c] += wgt * sm.dFn[f][k][a][c];
  }
  ir_vars_jacobian<DIM>(phys, qin + ((int64_t)sm.rec[f].x * NP + jn) * ND, W);
  double D[ND][ND];
  ec_pair_derivative<Op>(sm, sbp, mesh, phys, qin, f, si, sj, false, D);
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int c = 0; c < ND; ++c) {
      double acc = 0.0;
#pragma unroll
      for (int b = 0; b < ND; ++b) acc += M[a][b] * W[b][c];
      B[a][c] += acc + sg * D[a][c];
    }
}

template <class Op, int VOL, int FLUX, class BT>
__global__ void __launch_bounds__(256)
euler_jacobian_kernel(DevMesh mesh, DevSBP sbp, PhysConst phys, const double* __restrict__ qin,
                      const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
                      BT* __restrict__ vals, int32_t* __restrict__ err,
                      const double* __restrict__ ptc_diag, const double* __restrict__ ptc_scale) {
  constexpr int DIM = Op::DIM, NP = Op::NP, NF = Op::NF, ST = Op::ST, ND = Op::ND, NFACES = Op::NFACES;
  constexpr int NB = ND * ND;
  using Smem = JacSmem<Op, VOL, FLUX>;
  __shared__ Smem sm;
  const int tid = threadIdx.x;
  const int64_t e = element_of(mesh, blockIdx.x);
  for (int x = tid; x < DIM * NP * NP; x += blockDim.x) (&sm.Q[0][0][0])[x] = sbp.Q[x];
  for (int x = tid; x < ST * NF; x += blockDim.x) (&sm.interp[0][0])[x] = sbp.interp[x];
  for (int x = tid; x < ST * NFACES; x += blockDim.x) (&sm.perm[0][0])[x] = sbp.perm[x];
  for (int x = tid; x < NP * NFACES; x += blockDim.x) (&sm.nodepos[0][0])[x] = sbp.nodepos[x];
  for (int x = tid; x < NF; x += blockDim.x) sm.wface[x] = sbp.wface[x];
  for (int x = tid; x < NP * ND; x += blockDim.x) (&sm.q[0][0])[x] = qin[e * NP * ND + x];
  if (tid < NFACES) {
    ElFace r = mesh.elface[e * NFACES + tid];
    sm.rec[tid] = r;
    const int side = (r.x >= 0) ? ((r.z >> 16) & 1) : 0;
    sm.sgn[tid] = side == 0 ? -1.0 : 1.0;
  }
  __syncthreads();
  for (int t = tid; t < NFACES * NF; t += blockDim.x) {
    const int f = t / NF, k = t - f * NF;
    const ElFace r = sm.rec[f];
    int ko = k, kn = k;
    if (r.x >= 0) {
      const int o = (r.z >> 8) & 0xff;
      const int kR = sbp.nbrperm[k * sbp.norient + o];
      const int side = (r.z >> 16) & 1;
      ko = side == 0 ? k : kR;
      kn = side == 0 ? kR : k;
    }
    sm.kown[f][k] = ko;
    sm.knbr[f][k] = kn;
  }

  // ---- point derivatives
  if constexpr (VOL == VOL_WEAK) {
    for (int t = tid; t < NP * DIM; t += blockDim.x) {
      const int j = t / DIM, d = t - j * DIM;
      dual<ND> qd[ND];
#pragma unroll
      for (int c = 0; c < ND; ++c) { qd[c] = dual<ND>(sm.q[j][c]); qd[c].d[c] = 1.0; }
      double nrm[DIM];
#pragma unroll
      for (int x = 0; x < DIM; ++x) nrm[x] = metric<DIM, NP>(mesh, e, j)[x * DIM + d];
      dual<ND> F[ND];
      euler_flux<DIM>(phys, qd, nrm, F);
#pragma unroll
      for (int r = 0; r < ND; ++r)
#pragma unroll
        for (int c = 0; c < ND; ++c) sm.A[j * DIM + d][r][c] = F[r].d[c];
    }
  }
  __syncthreads();
  for (int t = tid; t < NFACES * NF; t += blockDim.x) {
    const int f = t / NF, k = t - f * NF;
    const ElFace r = sm.rec[f];
    dual<2 * ND> qL[ND], qR[ND], F[ND];
    if (r.x < 0) {
      // boundary: seed the interior state (directions 0..ND-1)
#pragma unroll
      for (int c = 0; c < ND; ++c) {
        double v = 0.0;
        for (int s = 0; s < ST; ++s) v += sm.interp[s][k] * sm.q[sm.perm[s][f]][c];
        qL[c] = dual<2 * ND>(v);
        qL[c].d[c] = 1.0;
      }
      double qg[ND];
      bc_state<DIM>(mesh, r, k, NF, qg);
#pragma unroll
      for (int c = 0; c < ND; ++c) qR[c] = dual<2 * ND>(qg[c]);
      roe_flux<DIM>(phys, qL, qR, mesh.nrm_bndry + ((int64_t)r.y * NF + k) * DIM, F);
#pragma unroll
      for (int a = 0; a < ND; ++a)
#pragma unroll
        for (int c = 0; c < ND; ++c) { sm.dFo[f][k][a][c] = F[a].d[c]; sm.dFn[f][k][a][c] = 0.0; }
    } else if constexpr (FLUX == PDES_FLUX_ESLF) {
      // entropy penalty at face node k (calcEntropyPenaltyIntegral,
      // faceElementIntegrals.jl:199-290) as a function of the interpolated
      // IR entropy variables wl, wr: dP/dw_own and dP/dw_nbr; the chain
      // through dw/dq of each stencil node is applied in the block tasks
      const int fN = r.z & 0xff;
      const int side = (r.z >> 16) & 1;
      const int o = (r.z >> 8) & 0xff;
      const int kR = sbp.nbrperm[k * sbp.norient + o];
      const double* qnb = qin + r.x * NP * ND;
      double wlv[ND], wrv[ND];
#pragma unroll
      for (int c = 0; c < ND; ++c) { wlv[c] = 0.0; wrv[c] = 0.0; }
      for (int sn = 0; sn < ST; ++sn) {
        double qo[ND], qn[ND], wo[ND], wn[ND];
#pragma unroll
        for (int c = 0; c < ND; ++c) {
          qo[c] = sm.q[sm.perm[sn][f]][c];
          qn[c] = qnb[sm.perm[sn][fN] * ND + c];
        }
        to_ir_vars<DIM>(phys, qo, wo);
        to_ir_vars<DIM>(phys, qn, wn);
        const double rl = sm.interp[sn][k], rr = sm.interp[sn][kR];
#pragma unroll
        for (int c = 0; c < ND; ++c) {
          wlv[c] += rl * (side == 0 ? wo[c] : wn[c]);
          wrv[c] += rr * (side == 0 ? wn[c] : wo[c]);
        }
      }
      dual<2 * ND> wl[ND], wr[ND], qa[ND], qlr[ND], dw[ND];
#pragma unroll
      for (int c = 0; c < ND; ++c) {
        wl[c] = dual<2 * ND>(wlv[c]);
        wl[c].d[c] = 1.0;
        wr[c] = dual<2 * ND>(wrv[c]);
        wr[c].d[ND + c] = 1.0;
      }
      ir_to_conservative<DIM>(phys, wl, qa);
      ir_to_conservative<DIM>(phys, wr, qlr);
#pragma unroll
      for (int c = 0; c < ND; ++c) {
        qa[c] = 0.5 * (qa[c] + qlr[c]);
        dw[c] = wl[c] - wr[c];
        F[c] = dual<2 * ND>(0.0);
      }
      lf_entropy_apply<DIM>(phys, qa, dw, mesh.nrm_face + ((int64_t)r.y * NF + k) * DIM, sm.wface[k], F);
      const int oo = side == 0 ? 0 : ND, on = side == 0 ? ND : 0;
#pragma unroll
      for (int a = 0; a < ND; ++a)
#pragma unroll
        for (int c = 0; c < ND; ++c) { sm.dFo[f][k][a][c] = F[a].d[oo + c]; sm.dFn[f][k][a][c] = F[a].d[on + c]; }
    } else {
      const int64_t nb = r.x;
      const int fN = r.z & 0xff;
      const int side = (r.z >> 16) & 1;
      const int kR = sm.kown[f][k] == k ? sm.knbr[f][k] : sm.kown[f][k];
      const int fL = side == 0 ? f : fN, fR = side == 0 ? fN : f;
      const double* qnb = qin + nb * NP * ND;
#pragma unroll
      for (int c = 0; c < ND; ++c) {
        double vl = 0.0, vr = 0.0;
        // own side from LDS, neighbour side from global memory (separate
        // branches: no pointer select across address spaces)
        if (side == 0) {
          for (int s = 0; s < ST; ++s) {
            vl += sm.interp[s][k] * sm.q[sm.perm[s][fL]][c];
            vr += sm.interp[s][kR] * qnb[sm.perm[s][fR] * ND + c];
          }
        } else {
          for (int s = 0; s < ST; ++s) {
            vl += sm.interp[s][k] * qnb[sm.perm[s][fL] * ND + c];
            vr += sm.interp[s][kR] * sm.q[sm.perm[s][fR]][c];
          }
        }
        qL[c] = dual<2 * ND>(vl);
        qL[c].d[c] = 1.0;
        qR[c] = dual<2 * ND>(vr);
        qR[c].d[ND + c] = 1.0;
      }
      const double* nrm = mesh.nrm_face + ((int64_t)r.y * NF + k) * DIM;
      if constexpr (FLUX == PDES_FLUX_ROE) roe_flux<DIM>(phys, qL, qR, nrm, F);
      else irslf_flux<DIM>(phys, qL, qR, nrm, F);
      const int oo = side == 0 ? 0 : ND, on = side == 0 ? ND : 0;
#pragma unroll
      for (int a = 0; a < ND; ++a)
#pragma unroll
        for (int c = 0; c < ND; ++c) { sm.dFo[f][k][a][c] = F[a].d[oo + c]; sm.dFn[f][k][a][c] = F[a].d[on + c]; }
    }
  }
  __syncthreads();

  // ---- blocks of this element's rows
  // tasks: own blocks (i, j), then per face f the neighbour blocks
  constexpr int NBFACE = Op::DIAGE ? NF : NP * NP;
  const int ntask = NP * NP + NFACES * NBFACE;
  for (int t = tid; t < ntask; t += blockDim.x) {
    double B[ND][ND];
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int c = 0; c < ND; ++c) B[a][c] = 0.0;
    int64_t row, col;
    if (t < NP * NP) {
      const int i = t / NP, j = t - i * NP;
      row = e * NP + i;
      col = e * NP + j;
      if constexpr (VOL == VOL_WEAK) {
        // d/dq_j of sum_d Q_d[j][i] F_d(q_j)
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const double qji = sm.Q[d][j][i];
#pragma unroll
          for (int a = 0; a < ND; ++a)
#pragma unroll
            for (int c = 0; c < ND; ++c) B[a][c] += qji * sm.A[j * DIM + d][a][c];
        }
      } else {
        // split form: res_i -= sum_k 2 S_ik F*(q_i, q_k) (reference orientation F(q_max, q_min))
        for (int k = 0; k < NP; ++k) {
          if (k == i) continue;
          if (i != j && k != j) continue;
          const int hi = i > k ? i : k, lo = i > k ? k : i;
          dual<2 * ND> qa[ND], qb[ND];
#pragma unroll
          for (int c = 0; c < ND; ++c) {
            qa[c] = dual<2 * ND>(sm.q[hi][c]); qa[c].d[c] = 1.0;
            qb[c] = dual<2 * ND>(sm.q[lo][c]); qb[c].d[ND + c] = 1.0;
          }
          dual<2 * ND> za[ND], zb[ND], Fd[DIM * ND];
          ir_zvars<DIM>(phys, qa, za);
          ir_zvars<DIM>(phys, qb, zb);
          double nrmD[DIM * DIM];
#pragma unroll
          for (int x = 0; x < DIM; ++x)
#pragma unroll
            for (int d = 0; d < DIM; ++d) nrmD[x * DIM + d] = metric<DIM, NP>(mesh, e, hi)[x * DIM + d];
          ir_flux_z<DIM, DIM>(phys, za, zb, nrmD, Fd);
          // derivative w.r.t. q_j: directions of hi or lo
          const int off = (j == hi) ? 0 : ND;
#pragma unroll
          for (int d = 0; d < DIM; ++d) {
            const double s_ik = 0.5 * (sm.Q[d][i][k] - sm.Q[d][k][i]);
#pragma unroll
            for (int a = 0; a < ND; ++a)
#pragma unroll
              for (int c = 0; c < ND; ++c) B[a][c] -= 2 * s_ik * Fd[d * ND + a].d[off + c];
          }
        }
      }
      // face and boundary terms with both indices on this element
#pragma unroll
      for (int f = 0; f < NFACES; ++f) {
        const int si = sm.nodepos[i][f], sj = sm.nodepos[j][f];
        if (si < 0 || sj < 0) continue;
        const double sg = sm.sgn[f];
        if constexpr (Op::DIAGE) {
          if (i != j) continue;
          // own face node si; find the slot k with kown == si
          int k = 0;
          for (int kk = 0; kk < NF; ++kk) if (sm.kown[f][kk] == si) k = kk;
          const double wgt = sg * sm.wface[si];
#pragma unroll
          for (int a = 0; a < ND; ++a)
#pragma unroll
            for (int c = 0; c < ND; ++c) B[a][c] += wgt * sm.dFo[f][k][a][c];
        } else if (FLUX == PDES_FLUX_ESLF && sm.rec[f].x >= 0) {
          if constexpr (FLUX == PDES_FLUX_ESLF) eslf_own_block<Op>(sm, sbp, mesh, phys, qin, f, i, j, si, sj, B);
        } else {
          for (int k = 0; k < NF; ++k) {
            const int m = sm.kown[f][k];
            const double wgt = sg * sm.wface[m] * sm.interp[si][m] * sm.interp[sj][m];
#pragma unroll
            for (int a = 0; a < ND; ++a)
#pragma unroll
              for (int c = 0; c < ND; ++c) B[a][c] += wgt * sm.dFo[f][k][a][c];
          }
        }
      }
      // implicit-Euler globalization (applyEuler, globalization.jl): -M/tau_vec on the diagonal
      if (ptc_scale && i == j) {
        const double s = ptc_scale[mesh.el_inst ? mesh.el_inst[e] : 0];
        const double d = ptc_diag[e * NP + i] * s;
#pragma unroll
        for (int a = 0; a < ND; ++a) B[a][a] -= d;
      }
    } else {
      const int tt = t - NP * NP;
      const int f = tt / NBFACE, u = tt - f * NBFACE;
      const ElFace r = sm.rec[f];
      if (r.x < 0) continue;  // boundary face: no neighbour block
      const int fN = r.z & 0xff;
      const double sg = sm.sgn[f];
      if constexpr (Op::DIAGE) {
        // slot k couples own node perm[kown][f] with neighbour node perm[knbr][fN]
        const int k = u;
        const int m = sm.kown[f][k], mn = sm.knbr[f][k];
        row = e * NP + sm.perm[m][f];
        col = (int64_t)r.x * NP + sm.perm[mn][fN];
        const double wgt = sg * sm.wface[m];
#pragma unroll
        for (int a = 0; a < ND; ++a)
#pragma unroll
          for (int c = 0; c < ND; ++c) B[a][c] = wgt * sm.dFn[f][k][a][c];
      } else {
        const int i = u / NP, jn = u - i * NP;
        row = e * NP + i;
        col = (int64_t)r.x * NP + jn;
        const int si = sm.nodepos[i][f], sj = sbp.nodepos[jn * NFACES + fN];
        if constexpr (FLUX == PDES_FLUX_ESLF) {
          if (si >= 0 && sj >= 0) eslf_nbr_block<Op>(sm, sbp, mesh, phys, qin, f, jn, si, sj, B);
        } else if (si >= 0 && sj >= 0) {
          for (int k = 0; k < NF; ++k) {
            const int m = sm.kown[f][k], mn = sm.knbr[f][k];
            const double wgt = sg * sm.wface[m] * sm.interp[si][m] * sm.interp[sj][mn];
#pragma unroll
            for (int a = 0; a < ND; ++a)
#pragma unroll
              for (int c = 0; c < ND; ++c) B[a][c] += wgt * sm.dFn[f][k][a][c];
          }
        }
      }
    }
    const int64_t pos = bsr_find(rowptr, colind, row, (int32_t)col);
    if (pos < 0) { err[0] = 1; continue; }
    BT* dst = vals + pos * NB;
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int c = 0; c < ND; ++c) dst[a * ND + c] = to_block<BT>(B[a][c]);
  }
}

// y = J x (BSR, block rows of ndof, FP32 or FP64 blocks, FP64 vectors):
// one thread per output dof; the threads of a block row read the block
// rows [a][0..ND) in turn.
template <int ND, class BT>
__global__ void __launch_bounds__(256)
bsr_spmv_kernel(int64_t nrows, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
                const BT* __restrict__ vals, const double* __restrict__ x, double* __restrict__ y,
                double alpha, const double* __restrict__ yadd, double beta,
                const int32_t* __restrict__ el_list, int np,
                const int32_t* __restrict__ el_inst, const int32_t* __restrict__ mask) {
  const int64_t tt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t rr = tt / ND;
  const int a = (int)(tt - rr * ND);
  if (rr >= nrows) return;
  // with an element list, launch row rr is node (rr % np) of element el_list[rr / np]
  const int64_t row = el_list ? (int64_t)el_list[rr / np] * np + (rr % np) : rr;
  // packed batches: rows of instances whose mask entry is 0 are left untouched
  if (mask && !mask[el_inst[row / np]]) return;
  const int64_t t = row * ND + a;
  double acc = 0.0;
  const int64_t b0 = rowptr[row], b1 = rowptr[row + 1];
  for (int64_t b = b0; b < b1; ++b) {
    const BT* blk = vals + b * (ND * ND) + a * ND;
    const double* xc = x + (int64_t)colind[b] * ND;
#pragma unroll
    for (int c = 0; c < ND; ++c) acc += (double)blk[c] * xc[c];
  }
  double v = alpha * acc;
  if (yadd) v += beta * yadd[t];
  y[t] = v;
}

}  // namespace pdes
