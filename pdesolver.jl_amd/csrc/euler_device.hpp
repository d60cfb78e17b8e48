// euler_device.hpp -- Euler point functions on the device, templated on the
// space dimension and on the scalar type (double, or dual<N> for Jacobian
// blocks and Jacobian-vector products).  Each mirrors the reference function
// cited beside it; the arithmetic sequence is kept so that results agree with
// the reference to rounding.  Divisions and square roots of the hot flux
// functions use drcp / dsqrt (Newton-refined hardware estimates, within about
// one ulp) instead of the correctly rounded IEEE sequences.
#pragma once
#include <cstddef>
#include <type_traits>
#include "pdes_common.hpp"

namespace pdes {

struct PhysConst {
  double gamma, gamma_1, R, cv;
  double Ma, aoa, rho_free, E_free;
  double vortex_Min;
};

// calcPressure  src/solver/euler/euler_funcs.jl:856-863 (2D), 897-903 (3D)
template <int DIM, class T>
__device__ __forceinline__ T pressure(const PhysConst& p, const T* q) {
  if constexpr (DIM == 2) return p.gamma_1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) * drcp(q[0]));
  else return p.gamma_1 * (q[4] - 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) * drcp(q[0]));
}

// calcEulerFlux  euler_funcs.jl:512-536 (2D), 749-774 (3D)
template <int DIM, class T, class TN = double>
__device__ __forceinline__ void euler_flux(const PhysConst& p, const T* q, const TN* dir, T* F) {
  T press = pressure<DIM>(p, q);
  if constexpr (DIM == 2) {
    T U = (q[1] * dir[0] + q[2] * dir[1]) * drcp(q[0]);
    F[0] = q[0] * U;
    F[1] = q[1] * U + dir[0] * press;
    F[2] = q[2] * U + dir[1] * press;
    F[3] = (q[3] + press) * U;
  } else {
    T U = (q[1] * dir[0] + q[2] * dir[1] + q[3] * dir[2]) * drcp(q[0]);
    F[0] = q[0] * U;
    F[1] = q[1] * U + dir[0] * press;
    F[2] = q[2] * U + dir[1] * press;
    F[3] = q[3] * U + dir[2] * press;
    F[4] = (q[4] + press) * U;
  }
}

// RoeSolver + calcSAT  src/solver/euler/bc_solvers.jl:29-103 / 207-310 (2D),
// 111-187 / 313-420 (3D).  flux = SAT(qL - qR) + F_euler(qL).n
// (TN: the normal's type, double, or the dual type of T for derivatives with
// respect to the normal)
template <int DIM, class T, class TN = double>
__device__ __forceinline__ void roe_flux(const PhysConst& p, const T* q, const T* qg, const TN* nrm, T* flux) {
  constexpr int ND = DIM + 2;
  const double gamma = p.gamma, gami = p.gamma_1;
  const double sat_Vn = 0.025, sat_Vl = 0.025, tau = 1.0;
  T u, v, w, H;
  {
    T fac = drcp(q[0]);
    T uL = q[1] * fac, vL = q[2] * fac;
    T wL = 0.0;
    T phi;
    if constexpr (DIM == 2) phi = 0.5 * (uL * uL + vL * vL);
    else { wL = q[3] * fac; phi = 0.5 * (uL * uL + vL * vL + wL * wL); }
    T HL = gamma * q[ND - 1] * fac - gami * phi;
    fac = drcp(qg[0]);
    T uR = qg[1] * fac, vR = qg[2] * fac;
    T wR = 0.0;
    if constexpr (DIM == 2) phi = 0.5 * (uR * uR + vR * vR);
    else { wR = qg[3] * fac; phi = 0.5 * (uR * uR + vR * vR + wR * wR); }
    T HR = gamma * qg[ND - 1] * fac - gami * phi;
    T sqL = dsqrt(q[0]);
    T sqR = dsqrt(qg[0]);
    fac = drcp(sqL + sqR);
    u = (sqL * uL + sqR * uR) * fac;
    v = (sqL * vL + sqR * vR) * fac;
    if constexpr (DIM == 3) w = (sqL * wL + sqR * wR) * fac;
    H = (sqL * HL + sqR * HR) * fac;
  }
  T dq[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) dq[i] = q[i] - qg[i];
  // calcSAT
  TN nx = nrm[0], ny = nrm[1], nz = TN(0.0);
  TN dA;
  T phi, Un;
  if constexpr (DIM == 2) {
    dA = dsqrt(nx * nx + ny * ny);
    Un = u * nx + v * ny;
    phi = 0.5 * (u * u + v * v);
  } else {
    nz = nrm[2];
    dA = dsqrt(nx * nx + ny * ny + nz * nz);
    phi = 0.5 * (u * u + v * v + w * w);
    Un = u * nx + v * ny + w * nz;
  }
  T a = dsqrt(gami * (H - phi));
  T lambda1 = Un + dA * a;
  T lambda2 = Un - dA * a;
  T lambda3 = Un;
  T rhoA = absvalue(Un) + dA * a;
  lambda1 = 0.5 * (tau * jmax(absvalue(lambda1), sat_Vn * rhoA) - lambda1);
  lambda2 = 0.5 * (tau * jmax(absvalue(lambda2), sat_Vn * rhoA) - lambda2);
  lambda3 = 0.5 * (tau * jmax(absvalue(lambda3), sat_Vl * rhoA) - lambda3);
  T sat[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) sat[i] = lambda3 * dq[i];
  T E1dq[ND], E2dq[ND];
  T vel[3] = {u, v, w};
  TN n3[3] = {nx, ny, nz};
  // E1 dq, E2 dq
  {
    T e1 = phi * dq[0] - u * dq[1] - v * dq[2];
    if constexpr (DIM == 3) e1 = e1 - w * dq[3];
    e1 = e1 + dq[ND - 1];
    E1dq[0] = e1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) E1dq[1 + d] = e1 * vel[d];
    E1dq[ND - 1] = e1 * H;
    T e2 = -Un * dq[0] + nx * dq[1] + ny * dq[2];
    if constexpr (DIM == 3) e2 = e2 + nz * dq[3];
    E2dq[0] = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) E2dq[1 + d] = e2 * n3[d];
    E2dq[ND - 1] = e2 * Un;
    T tmp1 = 0.5 * (lambda1 + lambda2) - lambda3;
    T tmp2 = gami * drcp(a * a);
    TN tmp3 = drcp(dA * dA);
#pragma unroll
    for (int i = 0; i < ND; ++i) sat[i] = sat[i] + tmp1 * (tmp2 * E1dq[i] + tmp3 * E2dq[i]);
    // E3 dq, E4 dq
    E1dq[0] = e2;
#pragma unroll
    for (int d = 0; d < DIM; ++d) E1dq[1 + d] = e2 * vel[d];
    E1dq[ND - 1] = e2 * H;
    E2dq[0] = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) E2dq[1 + d] = e1 * n3[d];
    E2dq[ND - 1] = e1 * Un;
    tmp1 = 0.5 * (lambda1 - lambda2) * drcp(dA * a);
#pragma unroll
    for (int i = 0; i < ND; ++i) sat[i] = sat[i] + tmp1 * (E1dq[i] + gami * E2dq[i]);
  }
  T ef[ND];
  euler_flux<DIM>(p, q, nrm, ef);
#pragma unroll
  for (int i = 0; i < ND; ++i) flux[i] = sat[i] + ef[i];
}

// Ismail-Roe parameter vector of one state: z1 = sqrt(rho/p), z_{1+d} = z1 u_d,
// z_last = sqrt(rho p)  (the per-state part of calcEulerFlux_IR, bc_solvers.jl:776-805)
template <int DIM, class T>
__device__ __forceinline__ void ir_zvars(const PhysConst& p, const T* q, T* z) {
  T pr = pressure<DIM>(p, q);
  T z1 = dsqrt(q[0] * drcp(pr));
  z[0] = z1;
  const T rrho = drcp(q[0]);
#pragma unroll
  for (int d = 0; d < DIM; ++d) z[1 + d] = z1 * q[1 + d] * rrho;
  z[DIM + 1] = dsqrt(q[0] * pr);
}

// The series/log factor of logavg for aL, aR with r = 1/(aL + aR):
// logavg = (aL + aR) / (2 F) and 1/logavg = 2 F r (bc_solvers.jl:942-956,
// same branch test on u = f^2).
// When logs of both arguments are at hand (lgL, lgR, e.g. staged per node),
// log(aL/aR) is taken as lgL - lgR (the branch is only taken for
// |f| >= 0.0316, where the difference loses no accuracy that matters).
template <class T, class LP>
__device__ __forceinline__ T logavg_factor(const T& aL, const T& aR, const T& r, LP lgL, LP lgR) {
  const T f = (aL - aR) * r;
  const T u = f * f;
  if (re(u) < 1e-3) {
    // 1 + u/3 + u^2/5 + u^3/7 + u^4/9, evaluated as (1 + u/3) + u^2 ((1/5 + u/7) + u^2/9)
    const T u2 = u * u;
    return (1.0 + u * (1.0 / 3.0)) + u2 * ((1.0 / 5.0 + u * (1.0 / 7.0)) + u2 * (1.0 / 9.0));
  }
  if constexpr (std::is_same<LP, std::nullptr_t>::value) return dlog(aL / aR) * (0.5 * drcp(f));
  else return (*lgL - *lgR) * (0.5 * drcp(f));
}

// calcEulerFlux_IR, multi-direction  bc_solvers.jl:776-805 (2D), 842-874 (3D),
// from precomputed z vectors.  dir: [x][ndir] ; F: [ndir][ndof].
// The reference's averages with its divisions shared and its products
// simplified: with s0 = zL1 + zR1, s3 = zL4 + zR4, r0 = 1/s0, r3 = 1/s3 and
// the logavg factors F0, F3 (logavg(z1) = s0/(2 F0), logavg(z4) = s3/(2 F3)),
//   rho_hat = z1_avg logavg(z4) = s0 s3 / (4 F3),   p1_hat = s3 r0,
//   p2_hat = p1_hat (c1 F0/F3 + c2)   (c1 = (gamma+1)/(2 gamma), c2 = (gamma-1)/(2 gamma)),
//   h_hat = gamma/(gamma-1) p2_hat/rho_hat + |v|^2/2 = 4 gamma/(gamma-1) r0^2 (c1 F0 + c2 F3) + |v|^2/2.
// r0 and r3 come from one drcp of s0 s3 and 1/F3 from another (Newton-refined
// hardware estimates).
// lzL / lzR: optional logs {log z1, log z4} of the two states (pointers to
// log z1, with log z4 LS elements further; read only on the log branch of
// logavg), or nullptr.  dl: optional output {log(zL1/zR1), log(zL4/zR4)},
// which the logavg factors give without a log: log(aL/aR) = 2 F f.
template <int DIM, int NDIR, class T, class LP = std::nullptr_t, int LS = 1, class TN = double>
__device__ __forceinline__ void ir_flux_z(const PhysConst& p, const T* zL, const T* zR,
                                          const TN* dir, T* F, LP lzL = nullptr, LP lzR = nullptr,
                                          T* dl = nullptr) {
  constexpr int ND = DIM + 2;
  const double gamma = p.gamma, gamma_1 = p.gamma_1;
  const T s0 = zL[0] + zR[0];
  const T s3 = zL[ND - 1] + zR[ND - 1];
  // one refined reciprocal for both: 1/s0 = s3 t, 1/s3 = s0 t with t = 1/(s0 s3)
  const T t03 = drcp(s0 * s3);
  const T r0 = s3 * t03;
  const T r3 = s0 * t03;
  T F0, F3;
  if constexpr (std::is_same<LP, std::nullptr_t>::value) {
    F0 = logavg_factor(zL[0], zR[0], r0, nullptr, nullptr);
    F3 = logavg_factor(zL[ND - 1], zR[ND - 1], r3, nullptr, nullptr);
  } else {
    F0 = logavg_factor(zL[0], zR[0], r0, lzL, lzR);
    F3 = logavg_factor(zL[ND - 1], zR[ND - 1], r3, lzL + LS, lzR + LS);
  }
  if (dl) {
    dl[0] = (2.0 * F0) * ((zL[0] - zR[0]) * r0);
    dl[1] = (2.0 * F3) * ((zL[ND - 1] - zR[ND - 1]) * r3);
  }
  const T rho_hat = (0.25 * s0) * (s3 * drcp(F3));
  T vh[3];
#pragma unroll
  for (int d = 0; d < DIM; ++d) vh[d] = (zL[1 + d] + zR[1 + d]) * r0;
  const T p1_hat = s3 * r0;
  T ke = vh[0] * vh[0] + vh[1] * vh[1];
  if constexpr (DIM == 3) ke = ke + vh[2] * vh[2];
  const double c1 = (gamma + 1) / (2 * gamma), c2 = gamma_1 / (2 * gamma);
  const T h_hat = ((4.0 * gamma / gamma_1) * (r0 * r0)) * (c1 * F0 + c2 * F3) + 0.5 * ke;
#pragma unroll
  for (int i = 0; i < NDIR; ++i) {
    T un = dir[0 * NDIR + i] * vh[0] + dir[1 * NDIR + i] * vh[1];
    if constexpr (DIM == 3) un = un + dir[2 * NDIR + i] * vh[2];
    T mv_n = rho_hat * un;
    F[i * ND + 0] = mv_n;
#pragma unroll
    for (int d = 0; d < DIM; ++d) F[i * ND + 1 + d] = mv_n * vh[d] + dir[d * NDIR + i] * p1_hat;
    F[i * ND + ND - 1] = mv_n * h_hat;
  }
}

// convertToIR_  src/solver/euler/conversion.jl:160-205.  The entropy
// s = log(p) - gamma log(rho) enters only v[0], as -s/(gamma-1).  SMODE 0
// takes it from two logs of the state; SMODE 1 from the logs of the IR
// parameter vector, log rho = log z1 + log z4 and log p = log z4 - log z1
// (lz1 and lz4 given); SMODE 2 leaves it out (s = 0), for callers that add
// the entropy jump of two states themselves.
template <int DIM, class T, int SMODE = 0>
__device__ __forceinline__ void to_ir_vars(const PhysConst& p, const T* qc, T* qe, T lz1 = 0.0, T lz4 = 0.0) {
  constexpr int ND = DIM + 2;
  const double gamma = p.gamma, gamma_1 = p.gamma_1, gamma_1i = 1.0 / gamma_1;
  T k1 = qc[1] * qc[1] + qc[2] * qc[2];
  if constexpr (DIM == 3) k1 = k1 + qc[3] * qc[3];
  k1 = 0.5 * k1 * drcp(qc[0]);
  T rho_int = qc[ND - 1] - k1;
  // log(gamma_1 rho_int / rho^gamma) taken as log(gamma_1 rho_int) - gamma log(rho)
  T s;
  if constexpr (SMODE == 1) s = (lz4 - lz1) - gamma * (lz1 + lz4);
  else if constexpr (SMODE == 2) s = 0.0;
  else s = dlog(gamma_1 * rho_int) - gamma * dlog(qc[0]);
  T fac = drcp(rho_int);
  T tmp1 = -qc[0] * fac * gamma_1i;
  qe[0] = ((rho_int * (gamma + 1.0 - s) - qc[ND - 1]) * fac) * gamma_1i;
#pragma unroll
  for (int d = 0; d < DIM; ++d) qe[1 + d] = qc[1 + d] * fac * gamma_1i;
  qe[ND - 1] = tmp1;
}

// LFKernel applyEntropyKernel (faceElementIntegrals.jl:455-470):
// F += scale * lambda_max(qa, nrm) * A0(qa) dv, with getIRA0 (IR_stab.jl:15-110)
// applied in factored form: with m the momentum, E the energy,
// h = (E + p)/rho and s = m . dv_m,
//   row 0:    rho dv_0 + s + E dv_E
//   row d:    m_d (dv_0 + s/rho + h dv_E) + p dv_d
//   row E:    E dv_0 + h s + (rho h^2 - a^2 p/(gamma-1)) dv_E
// and getLambdaMax (euler_funcs.jl:1887-1910) = |U_n| + |n| a.
template <int DIM, class T, class TN = double>
__device__ __forceinline__ void lf_entropy_apply(const PhysConst& p, const T* qa, const T* dv, const TN* nrm,
                                                 double scale, T* F) {
  constexpr int ND = DIM + 2;
  const double gamma = p.gamma, gamma_1 = p.gamma_1;
  T pr = pressure<DIM>(p, qa);
  T rho = qa[0];
  T rhoinv = drcp(rho);
  T h = (qa[ND - 1] + pr) * rhoinv;
  T a2 = gamma * pr * rhoinv;
  T Un = 0.0;
  TN dA = 0.0;
#pragma unroll
  for (int i = 0; i < DIM; ++i) {
    Un = Un + nrm[i] * qa[i + 1] * rhoinv;
    dA = dA + nrm[i] * nrm[i];
  }
  dA = dsqrt(dA);
  const T lambda_max = scale * (absvalue3(Un) + dA * dsqrt(a2));
  T sm = 0.0;
#pragma unroll
  for (int d = 1; d <= DIM; ++d) sm = sm + qa[d] * dv[d];
  const T dvE = dv[ND - 1];
  const T common = dv[0] + sm * rhoinv + h * dvE;
  F[0] = F[0] + lambda_max * (rho * dv[0] + sm + qa[ND - 1] * dvE);
#pragma unroll
  for (int d = 1; d <= DIM; ++d) F[d] = F[d] + lambda_max * (qa[d] * common + pr * dv[d]);
  F[ND - 1] = F[ND - 1] + lambda_max * (qa[ND - 1] * dv[0] + h * sm + (rho * h * h - a2 * pr / gamma_1) * dvE);
}

// IRSLF face flux: calcEulerFlux_IRSLF bc_solvers.jl:898-909 ->
// applyEntropyKernel_diagE faceElementIntegrals.jl:510-575 -> LFKernel :455-470
// with getIRA0 IR_stab.jl:15-110 and getLambdaMax euler_funcs.jl:1887-1910.
template <int DIM, class T, class TN = double>
__device__ __forceinline__ void irslf_flux(const PhysConst& p, const T* qL, const T* qR,
                                           const TN* nrm, T* F) {
  constexpr int ND = DIM + 2;
  T zL[ND], zR[ND];
  ir_zvars<DIM>(p, qL, zL);
  ir_zvars<DIM>(p, qR, zR);
  ir_flux_z<DIM, 1>(p, zL, zR, nrm, F);
  // entropy-variable jump dv = v(qL) - v(qR)
  T dv[ND], vR[ND];
  to_ir_vars<DIM>(p, qL, dv);
  to_ir_vars<DIM>(p, qR, vR);
#pragma unroll
  for (int i = 0; i < ND; ++i) dv[i] = dv[i] - vR[i];
  T qa[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) qa[i] = 0.5 * (qL[i] + qR[i]);
  lf_entropy_apply<DIM>(p, qa, dv, nrm, 1.0, F);
}

// irslf_flux from the IR parameter vectors of both states, for kernels that
// hold them per node already: no square roots of the own state, and no logs.
// The entropy-variable jump is v(qL) - v(qR) without the entropy terms, plus
// -(s_L - s_R)/(gamma-1) with s_L - s_R = log(p_L/p_R) - gamma log(rho_L/rho_R)
// = (1 - gamma) log(zL4/zR4) - (1 + gamma) log(zL1/zR1), log ratios that the
// IR flux's logavg factors provide.
template <int DIM, class T>
__device__ __forceinline__ void irslf_flux_z(const PhysConst& p, const T* qL, const T* qR, const T* zL,
                                             const T* zR, const double* nrm, T* F) {
  constexpr int ND = DIM + 2;
  const double gamma = p.gamma;
  T dl[2];
  ir_flux_z<DIM, 1>(p, zL, zR, nrm, F, nullptr, nullptr, dl);
  T dv[ND], vR[ND];
  to_ir_vars<DIM, T, 2>(p, qL, dv);
  to_ir_vars<DIM, T, 2>(p, qR, vR);
#pragma unroll
  for (int i = 0; i < ND; ++i) dv[i] = dv[i] - vR[i];
  dv[0] = dv[0] - ((1.0 - gamma) * dl[1] - (1.0 + gamma) * dl[0]) * (1.0 / p.gamma_1);
  T qa[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) qa[i] = 0.5 * (qL[i] + qR[i]);
  lf_entropy_apply<DIM>(p, qa, dv, nrm, 1.0, F);
}

// convertToConservativeFromIR_ (src/solver/euler/conversion.jl:225-259):
// rho_int = gamma_1 exp(-s/gamma_1) (gamma_1 / (-gamma_1 w_E)^gamma)^(1/gamma_1),
// taken as gamma_1 exp((log(gamma_1) - gamma log(-gamma_1 w_E) - s) / gamma_1)
template <int DIM, class T>
__device__ __forceinline__ void ir_to_conservative(const PhysConst& p, const T* w, T* q) {
  constexpr int ND = DIM + 2;
  const double gamma = p.gamma, gamma_1 = p.gamma_1;
  T k1 = w[1] * w[1] + w[2] * w[2];
  if constexpr (DIM == 3) k1 = k1 + w[3] * w[3];
  k1 = 0.5 * gamma_1 * k1 / w[ND - 1];
  const T s = gamma - gamma_1 * w[0] + k1;
  const T rho_int = gamma_1 * dexp((log(gamma_1) - gamma * dlog(-gamma_1 * w[ND - 1]) - s) / gamma_1);
  q[0] = -w[ND - 1] * rho_int;
#pragma unroll
  for (int d = 1; d <= DIM; ++d) q[d] = w[d] * rho_int;
  q[ND - 1] = (1.0 - k1) * rho_int / gamma_1;
}

// calcIsentropicVortex  src/solver/euler/common_funcs.jl:25-76 (inflow Mach a parameter)
__device__ __forceinline__ void isentropic_vortex(const PhysConst& p, const double* coords, double* sol) {
  double x = coords[0], y = coords[1];
  double gamma = p.gamma;
  const double r_in = 1, rho_in = 2;
  double M_in = p.vortex_Min, p_in = 1 / gamma;
  double r = sqrt(x * x + y * y);
  double theta = atan2(y, x);
  double tmp1 = ((gamma - 1) / 2) * M_in * M_in;
  double rho_r = rho_in * pow(1 + tmp1 * (1 - (r_in * r_in) / (r * r)), 1 / (gamma - 1));
  double p_r = p_in * pow(rho_r / rho_in, gamma);
  double a_r = sqrt(gamma * p_r / rho_r);
  double M_r = sqrt((2 / (gamma - 1)) * pow(rho_in / rho_r, gamma - 1) * (1 + tmp1) - 2 / (gamma - 1));
  double U_r = M_r * a_r;
  double u_r = U_r * sin(theta);
  double v_r = -U_r * cos(theta);
  double e_r = p.cv * p_r / (rho_r * p.R);
  double E_r = rho_r * e_r + 0.5 * rho_r * U_r * U_r;
  sol[0] = rho_r;
  sol[1] = rho_r * u_r;
  sol[2] = rho_r * v_r;
  sol[3] = E_r;
}

// calcFreeStream  common_funcs.jl:312-347
template <int DIM>
__device__ __forceinline__ void free_stream(const PhysConst& p, double* sol) {
  double rho = p.rho_free;
  sol[0] = rho;
  sol[DIM + 1] = p.E_free;
  if constexpr (DIM == 2) {
    sol[1] = rho * p.Ma * cos(p.aoa);
    sol[2] = rho * p.Ma * sin(p.aoa);
  } else {
    sol[1] = rho * p.Ma * cos(p.aoa);
    sol[2] = 0.0;
    sol[3] = -rho * p.Ma * sin(p.aoa);
  }
}

}  // namespace pdes
