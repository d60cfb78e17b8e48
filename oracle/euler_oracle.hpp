// euler_oracle.hpp -- TEST INFRASTRUCTURE ONLY (the parity oracle).
//
// CPU restatement of the reference's Euler point functions, templated on the
// scalar type so that std::complex<double> gives the complex-step semantics of
// the reference (src/Utils/complexify.jl).  Only tests/, the smoke check and
// bench.py's CPU-baseline leg use it; the product library never links it.
//
// Each function cites the reference file:line it restates.
#pragma once
#include <cmath>
#include <complex>

namespace oracle {

using cplx = std::complex<double>;

inline double re(double x) { return x; }
inline double re(const cplx& x) { return x.real(); }

// absvalue: src/Utils/complexify.jl:25-45 (flipsign by the real part)
inline double absvalue(double x) { return std::fabs(x); }
inline cplx absvalue(const cplx& x) { return x.real() < 0 ? -x : x; }

// Julia max(x, y) = isless(y, x) ? x : y, with complex ordered by real part
template <class T> inline T jmax(const T& x, const T& y) { return re(y) < re(x) ? x : y; }

// absvalue3: src/Utils/complexify.jl:161-172 (Harten entropy fix, delta=1e-7)
template <class T> inline T absvalue3(const T& val) {
  const double delta = 1e-7;
  T val1 = absvalue(val);
  if (re(val1) > delta) return val1;
  return ((val * val) / delta + delta) / 2.0;
}

struct Params {
  int dim;
  double gamma, gamma_1, R, cv;
  double Ma, aoa, rho_free, p_free, E_free;
  double vortex_Min;
};

// calcPressure 2D/3D: src/solver/euler/euler_funcs.jl:856-863, 897-903
template <class T> inline T calcPressure(const Params& p, const T* q) {
  if (p.dim == 2) return p.gamma_1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
  return p.gamma_1 * (q[4] - 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / q[0]);
}

// calcEulerFlux 2D :512-536, 3D :749-774 (euler_funcs.jl)
template <class T, class Tm> inline void calcEulerFlux(const Params& p, const T* q, const Tm* dir, T* F) {
  T press = calcPressure(p, q);
  if (p.dim == 2) {
    T U = (q[1] * dir[0] + q[2] * dir[1]) / q[0];
    F[0] = q[0] * U;
    F[1] = q[1] * U + dir[0] * press;
    F[2] = q[2] * U + dir[1] * press;
    F[3] = (q[3] + press) * U;
  } else {
    T U = (q[1] * dir[0] + q[2] * dir[1] + q[3] * dir[2]) / q[0];
    F[0] = q[0] * U;
    F[1] = q[1] * U + dir[0] * press;
    F[2] = q[2] * U + dir[1] * press;
    F[3] = q[3] * U + dir[2] * press;
    F[4] = (q[4] + press) * U;
  }
}

// calcSAT 2D: src/solver/euler/bc_solvers.jl:207-310 ; 3D :313-420
template <class T> inline void calcSAT(const Params& p, const T* roe_vars, const T* dq, const double* nrm, T* sat) {
  const double sat_Vn = 0.025, sat_Vl = 0.025, tau = 1.0;
  const double gami = p.gamma_1;
  if (p.dim == 2) {
    T u = roe_vars[0], v = roe_vars[1], H = roe_vars[2];
    double nx = nrm[0], ny = nrm[1];
    double dA = std::sqrt(nx * nx + ny * ny);
    T Un = u * nx + v * ny;
    T phi = 0.5 * (u * u + v * v);
    T a = std::sqrt(gami * (H - phi));
    T lambda1 = Un + dA * a;
    T lambda2 = Un - dA * a;
    T lambda3 = Un;
    T rhoA = absvalue(Un) + dA * a;
    lambda1 = 0.5 * (tau * jmax(absvalue(lambda1), sat_Vn * rhoA) - lambda1);
    lambda2 = 0.5 * (tau * jmax(absvalue(lambda2), sat_Vn * rhoA) - lambda2);
    lambda3 = 0.5 * (tau * jmax(absvalue(lambda3), sat_Vl * rhoA) - lambda3);
    T dq1 = dq[0], dq2 = dq[1], dq3 = dq[2], dq4 = dq[3];
    sat[0] = lambda3 * dq1; sat[1] = lambda3 * dq2; sat[2] = lambda3 * dq3; sat[3] = lambda3 * dq4;
    T E1dq[4], E2dq[4];
    E1dq[0] = phi * dq1 - u * dq2 - v * dq3 + dq4;
    E1dq[1] = E1dq[0] * u; E1dq[2] = E1dq[0] * v; E1dq[3] = E1dq[0] * H;
    E2dq[0] = 0.0;
    E2dq[1] = -Un * dq1 + nx * dq2 + ny * dq3;
    E2dq[2] = E2dq[1] * ny; E2dq[3] = E2dq[1] * Un; E2dq[1] = E2dq[1] * nx;
    T tmp1 = 0.5 * (lambda1 + lambda2) - lambda3;
    T tmp2 = gami / (a * a);
    double tmp3 = 1.0 / (dA * dA);
    for (int i = 0; i < 4; ++i) sat[i] = sat[i] + tmp1 * (tmp2 * E1dq[i] + tmp3 * E2dq[i]);
    E1dq[0] = -Un * dq1 + nx * dq2 + ny * dq3;
    E1dq[1] = E1dq[0] * u; E1dq[2] = E1dq[0] * v; E1dq[3] = E1dq[0] * H;
    E2dq[0] = 0.0;
    E2dq[1] = phi * dq1 - u * dq2 - v * dq3 + dq4;
    E2dq[2] = E2dq[1] * ny; E2dq[3] = E2dq[1] * Un; E2dq[1] = E2dq[1] * nx;
    tmp1 = 0.5 * (lambda1 - lambda2) / (dA * a);
    for (int i = 0; i < 4; ++i) sat[i] = sat[i] + tmp1 * (E1dq[i] + gami * E2dq[i]);
  } else {
    T u = roe_vars[0], v = roe_vars[1], w = roe_vars[2], H = roe_vars[3];
    double nx = nrm[0], ny = nrm[1], nz = nrm[2];
    double dA = std::sqrt(nx * nx + ny * ny + nz * nz);
    T phi = 0.5 * (u * u + v * v + w * w);
    T a = std::sqrt(gami * (H - phi));
    T Un = u * nx + v * ny + w * nz;
    T lambda1 = Un + dA * a;
    T lambda2 = Un - dA * a;
    T lambda3 = Un;
    T rhoA = absvalue(Un) + dA * a;
    lambda1 = 0.5 * (tau * jmax(absvalue(lambda1), sat_Vn * rhoA) - lambda1);
    lambda2 = 0.5 * (tau * jmax(absvalue(lambda2), sat_Vn * rhoA) - lambda2);
    lambda3 = 0.5 * (tau * jmax(absvalue(lambda3), sat_Vl * rhoA) - lambda3);
    T dq1 = dq[0], dq2 = dq[1], dq3 = dq[2], dq4 = dq[3], dq5 = dq[4];
    sat[0] = lambda3 * dq1; sat[1] = lambda3 * dq2; sat[2] = lambda3 * dq3;
    sat[3] = lambda3 * dq4; sat[4] = lambda3 * dq5;
    T E1dq[5], E2dq[5];
    E1dq[0] = phi * dq1 - u * dq2 - v * dq3 - w * dq4 + dq5;
    E1dq[1] = E1dq[0] * u; E1dq[2] = E1dq[0] * v; E1dq[3] = E1dq[0] * w; E1dq[4] = E1dq[0] * H;
    E2dq[0] = 0.0;
    E2dq[1] = -Un * dq1 + nx * dq2 + ny * dq3 + nz * dq4;
    E2dq[2] = E2dq[1] * ny; E2dq[3] = E2dq[1] * nz; E2dq[4] = E2dq[1] * Un; E2dq[1] = E2dq[1] * nx;
    T tmp1 = 0.5 * (lambda1 + lambda2) - lambda3;
    T tmp2 = gami / (a * a);
    double tmp3 = 1.0 / (dA * dA);
    for (int i = 0; i < 5; ++i) sat[i] = sat[i] + tmp1 * (tmp2 * E1dq[i] + tmp3 * E2dq[i]);
    E1dq[0] = -Un * dq1 + nx * dq2 + ny * dq3 + nz * dq4;
    E1dq[1] = E1dq[0] * u; E1dq[2] = E1dq[0] * v; E1dq[3] = E1dq[0] * w; E1dq[4] = E1dq[0] * H;
    E2dq[0] = 0.0;
    E2dq[1] = phi * dq1 - u * dq2 - v * dq3 - w * dq4 + dq5;
    E2dq[2] = E2dq[1] * ny; E2dq[3] = E2dq[1] * nz; E2dq[4] = E2dq[1] * Un; E2dq[1] = E2dq[1] * nx;
    tmp1 = 0.5 * (lambda1 - lambda2) / (dA * a);
    for (int i = 0; i < 5; ++i) sat[i] = sat[i] + tmp1 * (E1dq[i] + gami * E2dq[i]);
  }
}

// RoeSolver 2D: src/solver/euler/bc_solvers.jl:29-103 ; 3D :111-187
template <class T> inline void RoeSolver(const Params& p, const T* q, const T* qg, const double* nrm, T* flux) {
  const double gamma = p.gamma, gami = p.gamma_1;
  const int nd = p.dim + 2;
  T roe_vars[4], dq[5], sat[5], euler_flux[5];
  if (p.dim == 2) {
    T fac = 1.0 / q[0];
    T uL = q[1] * fac, vL = q[2] * fac;
    T phi = 0.5 * (uL * uL + vL * vL);
    T HL = gamma * q[3] * fac - gami * phi;
    fac = 1.0 / qg[0];
    T uR = qg[1] * fac, vR = qg[2] * fac;
    phi = 0.5 * (uR * uR + vR * vR);
    T HR = gamma * qg[3] * fac - gami * phi;
    T sqL = std::sqrt(q[0]);
    T sqR = std::sqrt(qg[0]);
    fac = 1.0 / (sqL + sqR);
    roe_vars[0] = (sqL * uL + sqR * uR) * fac;
    roe_vars[1] = (sqL * vL + sqR * vR) * fac;
    roe_vars[2] = (sqL * HL + sqR * HR) * fac;
  } else {
    T fac = 1.0 / q[0];
    T uL = q[1] * fac, vL = q[2] * fac, wL = q[3] * fac;
    T phi = 0.5 * (uL * uL + vL * vL + wL * wL);
    T HL = gamma * q[4] * fac - gami * phi;
    fac = 1.0 / qg[0];
    T uR = qg[1] * fac, vR = qg[2] * fac, wR = qg[3] * fac;
    phi = 0.5 * (uR * uR + vR * vR + wR * wR);
    T HR = gamma * qg[4] * fac - gami * phi;
    T sqL = std::sqrt(q[0]);
    T sqR = std::sqrt(qg[0]);
    fac = 1.0 / (sqL + sqR);
    roe_vars[0] = (sqL * uL + sqR * uR) * fac;
    roe_vars[1] = (sqL * vL + sqR * vR) * fac;
    roe_vars[2] = (sqL * wL + sqR * wR) * fac;
    roe_vars[3] = (sqL * HL + sqR * HR) * fac;
  }
  for (int i = 0; i < nd; ++i) dq[i] = q[i] - qg[i];
  calcSAT(p, roe_vars, dq, nrm, sat);
  calcEulerFlux(p, q, nrm, euler_flux);
  for (int i = 0; i < nd; ++i) flux[i] = sat[i] + euler_flux[i];
}

// logavg: src/solver/euler/bc_solvers.jl:942-956
template <class T> inline T logavg(const T& aL, const T& aR) {
  T xi = aL / aR;
  T f = (xi - 1.0) / (xi + 1.0);
  T u = f * f;
  const double eps = 1e-3;
  T F;
  if (re(u) < eps) {
    // @evalpoly(u, 1, 1/3, 1/5, 1/7, 1/9) (Horner)
    F = 1.0 + u * (1.0 / 3.0 + u * (1.0 / 5.0 + u * (1.0 / 7.0 + u * (1.0 / 9.0))));
  } else {
    F = (std::log(xi) / 2.0) / f;
  }
  return (aL + aR) / (2.0 * F);
}

// calcEulerFlux_IR multi-direction 2D: bc_solvers.jl:776-805 ; 3D :842-874
// dir: [dim(x)][ndir] column d = direction d ; F: [ndir][ndof]
template <class T> inline void calcEulerFlux_IR(const Params& p, const T* qL, const T* qR,
                                                const double* dir, int ndir, T* F) {
  const double gamma = p.gamma, gamma_1 = p.gamma_1;
  T pL = calcPressure(p, qL), pR = calcPressure(p, qR);
  T z1L = std::sqrt(qL[0] / pL), z1R = std::sqrt(qR[0] / pR);
  if (p.dim == 2) {
    T z2L = z1L * qL[1] / qL[0], z2R = z1R * qR[1] / qR[0];
    T z3L = z1L * qL[2] / qL[0], z3R = z1R * qR[2] / qR[0];
    T z4L = std::sqrt(qL[0] * pL), z4R = std::sqrt(qR[0] * pR);
    T rho_hat = 0.5 * (z1L + z1R) * logavg(z4L, z4R);
    T u_hat = (z2L + z2R) / (z1L + z1R);
    T v_hat = (z3L + z3R) / (z1L + z1R);
    T p1_hat = (z4L + z4R) / (z1L + z1R);
    T p2_hat = ((gamma + 1) / (2 * gamma)) * logavg(z4L, z4R) / logavg(z1L, z1R) +
               (gamma_1 / (2 * gamma)) * (z4L + z4R) / (z1L + z1R);
    T h_hat = gamma * p2_hat / (rho_hat * gamma_1) + 0.5 * (u_hat * u_hat + v_hat * v_hat);
    for (int i = 0; i < ndir; ++i) {
      double d0 = dir[0 * ndir + i], d1 = dir[1 * ndir + i];
      T mv_n = rho_hat * (d0 * u_hat + d1 * v_hat);
      F[i * 4 + 0] = mv_n;
      F[i * 4 + 1] = mv_n * u_hat + d0 * p1_hat;
      F[i * 4 + 2] = mv_n * v_hat + d1 * p1_hat;
      F[i * 4 + 3] = mv_n * h_hat;
    }
  } else {
    T z2L = z1L * qL[1] / qL[0], z2R = z1R * qR[1] / qR[0];
    T z3L = z1L * qL[2] / qL[0], z3R = z1R * qR[2] / qR[0];
    T z4L = z1L * qL[3] / qL[0], z4R = z1R * qR[3] / qR[0];
    T z5L = std::sqrt(qL[0] * pL), z5R = std::sqrt(qR[0] * pR);
    T rho_hat = 0.5 * (z1L + z1R) * logavg(z5L, z5R);
    T u_hat = (z2L + z2R) / (z1L + z1R);
    T v_hat = (z3L + z3R) / (z1L + z1R);
    T w_hat = (z4L + z4R) / (z1L + z1R);
    T p1_hat = (z5L + z5R) / (z1L + z1R);
    T p2_hat = ((gamma + 1) / (2 * gamma)) * logavg(z5L, z5R) / logavg(z1L, z1R) +
               (gamma_1 / (2 * gamma)) * (z5L + z5R) / (z1L + z1R);
    T h_hat = gamma * p2_hat / (rho_hat * gamma_1) +
              0.5 * (u_hat * u_hat + v_hat * v_hat + w_hat * w_hat);
    for (int i = 0; i < ndir; ++i) {
      double d0 = dir[0 * ndir + i], d1 = dir[1 * ndir + i], d2 = dir[2 * ndir + i];
      T mv_n = rho_hat * (d0 * u_hat + d1 * v_hat + d2 * w_hat);
      F[i * 5 + 0] = mv_n;
      F[i * 5 + 1] = mv_n * u_hat + d0 * p1_hat;
      F[i * 5 + 2] = mv_n * v_hat + d1 * p1_hat;
      F[i * 5 + 3] = mv_n * w_hat + d2 * p1_hat;
      F[i * 5 + 4] = mv_n * h_hat;
    }
  }
}

// convertToIR_ 2D: src/solver/euler/conversion.jl:160-182 ; 3D :184-205
template <class T> inline void convertToIR(const Params& p, const T* qc, T* qe) {
  const double gamma = p.gamma, gamma_1 = p.gamma_1, gamma_1i = 1.0 / gamma_1;
  if (p.dim == 2) {
    T k1 = 0.5 * (qc[1] * qc[1] + qc[2] * qc[2]) / qc[0];
    T rho_int = qc[3] - k1;
    T s = std::log(gamma_1 * rho_int / std::pow(qc[0], gamma));
    T fac = 1.0 / rho_int;
    T tmp1 = -qc[0] * fac * gamma_1i;
    qe[0] = ((rho_int * (gamma + 1.0 - s) - qc[3]) * fac) * gamma_1i;
    qe[1] = qc[1] * fac * gamma_1i;
    qe[2] = qc[2] * fac * gamma_1i;
    qe[3] = tmp1;
  } else {
    T k1 = 0.5 * (qc[1] * qc[1] + qc[2] * qc[2] + qc[3] * qc[3]) / qc[0];
    T rho_int = qc[4] - k1;
    T s = std::log(gamma_1 * rho_int / std::pow(qc[0], gamma));
    T fac = 1.0 / rho_int;
    T tmp1 = -qc[0] * fac * gamma_1i;
    qe[0] = ((rho_int * (gamma + 1.0 - s) - qc[4]) * fac) * gamma_1i;
    qe[1] = qc[1] * fac * gamma_1i;
    qe[2] = qc[2] * fac * gamma_1i;
    qe[3] = qc[3] * fac * gamma_1i;
    qe[4] = tmp1;
  }
}

// convertToConservativeFromIR_ 2D: src/solver/euler/conversion.jl:225-242 ; 3D :244-259
template <class T> inline void convertToConservativeFromIR(const Params& p, const T* qe, T* qc) {
  const int dim = p.dim, nd = dim + 2;
  const double gamma = p.gamma, gamma_1 = p.gamma_1;
  T k1 = qe[1] * qe[1] + qe[2] * qe[2];
  if (dim == 3) k1 += qe[3] * qe[3];
  k1 = 0.5 * gamma_1 * k1 / qe[nd - 1];
  T s = gamma - gamma_1 * qe[0] + k1;
  T rho_int = std::exp(-s / gamma_1) * std::pow(gamma_1 / std::pow(-gamma_1 * qe[nd - 1], gamma), 1.0 / gamma_1);
  rho_int *= gamma_1;
  qc[0] = -qe[nd - 1] * rho_int;
  for (int d = 1; d <= dim; ++d) qc[d] = qe[d] * rho_int;
  qc[nd - 1] = (1.0 - k1) * rho_int / gamma_1;
}

// getIRA0 2D: src/solver/euler/IR_stab.jl:15-56 ; 3D :58-110.  A0 column-major [ndof][ndof] as A0[col*nd+row]
template <class T> inline void getIRA0(const Params& p, const T* q, T* A0) {
  const double gamma = p.gamma, gamma_1 = p.gamma_1;
  T pr = calcPressure(p, q);
  const int nd = p.dim + 2;
  auto A = [&](int r, int c) -> T& { return A0[c * nd + r]; };
  if (p.dim == 2) {
    T rho = q[0], rhou = q[1], rhov = q[2], rhoe = q[3];
    T rhoinv = 1.0 / rho;
    T h = (rhoe + pr) * rhoinv;
    T a2 = gamma * pr * rhoinv;
    A(0, 0) = rho; A(1, 0) = rhou; A(2, 0) = rhov; A(3, 0) = rhoe;
    A(0, 1) = rhou; A(1, 1) = rhou * rhou * rhoinv + pr; A(2, 1) = rhou * rhov * rhoinv; A(3, 1) = rhou * h;
    A(0, 2) = rhov; A(1, 2) = rhou * rhov / rho; A(2, 2) = rhov * rhov * rhoinv + pr; A(3, 2) = rhov * h;
    A(0, 3) = rhoe; A(1, 3) = h * rhou; A(2, 3) = h * rhov; A(3, 3) = rho * h * h - a2 * pr / gamma_1;
  } else {
    T rho = q[0], rhou = q[1], rhov = q[2], rhow = q[3], rhoe = q[4];
    T rhoinv = 1.0 / rho;
    T h = (rhoe + pr) * rhoinv;
    T a2 = gamma * pr * rhoinv;
    A(0, 0) = rho; A(1, 0) = rhou; A(2, 0) = rhov; A(3, 0) = rhow; A(4, 0) = rhoe;
    A(0, 1) = rhou; A(1, 1) = rhou * rhou * rhoinv + pr; A(2, 1) = rhou * rhov * rhoinv;
    A(3, 1) = rhou * rhow * rhoinv; A(4, 1) = rhou * h;
    A(0, 2) = rhov; A(1, 2) = rhou * rhov / rho; A(2, 2) = rhov * rhov * rhoinv + pr;
    A(3, 2) = rhov * rhow * rhoinv; A(4, 2) = rhov * h;
    A(0, 3) = rhow; A(1, 3) = rhow * rhou * rhoinv; A(2, 3) = rhow * rhov * rhoinv;
    A(3, 3) = rhow * rhow * rhoinv + pr; A(4, 3) = rhow * h;
    A(0, 4) = rhoe; A(1, 4) = h * rhou; A(2, 4) = h * rhov; A(3, 4) = h * rhow;
    A(4, 4) = rho * h * h - a2 * pr / gamma_1;
  }
}

// getLambdaMax(params, q, dir): src/solver/euler/euler_funcs.jl:1887-1910
template <class T> inline T getLambdaMax(const Params& p, const T* q, const double* dir) {
  T Un = 0.0;
  double dA = 0.0;
  T rhoLinv = 1.0 / q[0];
  T pL = calcPressure(p, q);
  T aL = std::sqrt(p.gamma * pL * rhoLinv);
  for (int i = 0; i < p.dim; ++i) {
    Un += dir[i] * q[i + 1] * rhoLinv;
    dA += dir[i] * dir[i];
  }
  dA = std::sqrt(dA);
  return absvalue3(Un) + dA * aL;
}

// calcEulerFlux_IRSLF: bc_solvers.jl:898-909 -> applyEntropyKernel_diagE
// faceElementIntegrals.jl:510-525, _inner :546-575, LFKernel :455-470
template <class T> inline void calcEulerFlux_IRSLF(const Params& p, const T* qL, const T* qR,
                                                   const double* dir, T* F) {
  const int nd = p.dim + 2;
  calcEulerFlux_IR(p, qL, qR, dir, 1, F);
  T q_avg[5], vL[5], vR[5], A0[25], F_tmp[5];
  for (int i = 0; i < nd; ++i) q_avg[i] = 0.5 * (qL[i] + qR[i]);
  convertToIR(p, qL, vL);
  convertToIR(p, qR, vR);
  for (int i = 0; i < nd; ++i) vL[i] = vL[i] - vR[i];
  getIRA0(p, q_avg, A0);
  T lambda_max = getLambdaMax(p, q_avg, dir);
  // smallmatvec!(A0, delta_w, flux); fastscale!(flux, lambda_max)
  for (int r = 0; r < nd; ++r) {
    T acc = 0.0;
    for (int c = 0; c < nd; ++c) acc += A0[c * nd + r] * vL[c];
    F_tmp[r] = acc * lambda_max;
  }
  for (int i = 0; i < nd; ++i) F[i] += F_tmp[i];
}

// calcIsentropicVortex 2D: src/solver/euler/common_funcs.jl:25-76
// (M_in parameterized; the reference hard-codes 0.95)
inline void calcIsentropicVortex(const Params& p, const double* coords, double* sol) {
  double x = coords[0], y = coords[1];
  double cv = p.cv, R = p.R, gamma = p.gamma;
  double r_in = 1, rho_in = 2, M_in = p.vortex_Min, p_in = 1 / gamma;
  double r = std::sqrt(x * x + y * y);
  double theta = std::atan2(y, x);
  double tmp1 = ((gamma - 1) / 2) * M_in * M_in;
  double rho_r = rho_in * std::pow(1 + tmp1 * (1 - (r_in * r_in) / (r * r)), 1 / (gamma - 1));
  double p_r = p_in * std::pow(rho_r / rho_in, gamma);
  double a_r = std::sqrt(gamma * p_r / rho_r);
  double M_r = std::sqrt((2 / (gamma - 1)) * std::pow(rho_in / rho_r, gamma - 1) * (1 + tmp1) - 2 / (gamma - 1));
  double U_r = M_r * a_r;
  double u_r = U_r * std::sin(theta);
  double v_r = -U_r * std::cos(theta);
  double e_r = cv * p_r / (rho_r * R);
  double E_r = rho_r * e_r + 0.5 * rho_r * U_r * U_r;
  sol[0] = rho_r;
  sol[1] = rho_r * u_r;
  sol[2] = rho_r * v_r;
  sol[3] = E_r;
}

// calcFreeStream 2D: common_funcs.jl:312-327 ; 3D :329-347
inline void calcFreeStream(const Params& p, double* sol) {
  double rho = p.rho_free;
  if (p.dim == 2) {
    sol[0] = rho; sol[3] = p.E_free;
    sol[1] = rho * p.Ma * std::cos(p.aoa);
    sol[2] = rho * p.Ma * std::sin(p.aoa);
  } else {
    sol[0] = rho; sol[4] = p.E_free;
    sol[1] = rho * p.Ma * std::cos(p.aoa);
    sol[2] = 0.0;
    sol[3] = -rho * p.Ma * std::sin(p.aoa);
  }
}

}  // namespace oracle
