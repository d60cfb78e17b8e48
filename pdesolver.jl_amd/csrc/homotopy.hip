// homotopy.hip -- the first-order dissipation D(q) of the homotopy
// globalisation (calcHomotopyDiss, src/solver/euler/homotopy.jl:51-249) on the
// device, and its directional derivative.
//
// Per element, per reference direction d: t1 = H^-1 Q_d q, t2_j =
// lambda_max(q_j, dxidx[d, :, j]) t1_j (getLambdaMax, euler_funcs.jl:1887-1910),
// D -= Q_d^T t2; per interior face node, flux = lambda_max(avg, nrm)
// (qL - qR) / 2 (getLambdaMaxSimple, IR_stab.jl:310-325), D_L -= R_L^T B flux,
// D_R += R_R^T B flux.  With homotopy_addBoundaryIntegrals (default off,
// src/input/read_input.jl:438) each boundary face node adds the same flux
// between its state and the free stream (calcFreeStream, whatever the BC):
// D -= R^T B flux.
//
// One block of 64 threads per element, thread i owns volume node i; the
// element's states, the scaled derivatives t2 and the face contributions are
// staged in LDS.  Each element pulls its own faces (the neighbour's state is
// read from global memory), so no two blocks write the same node.  The
// homotopy is a globalisation step, not the residual hot path: the kernel
// is written for any operator size (Np <= 64) rather than per shape.
#include "jacobian.hpp"

namespace pdes {

template <int DIM, class T>
__device__ __forceinline__ T hom_lambda_max(const PhysConst& p, const T* q, const double* dir) {
  const T rinv = drcp(q[0]);
  const T a = dsqrt(p.gamma * pressure<DIM>(p, q) * rinv);
  T un = 0.0;
  double dA = 0.0;
#pragma unroll
  for (int x = 0; x < DIM; ++x) {
    un = un + dir[x] * q[x + 1] * rinv;
    dA += dir[x] * dir[x];
  }
  return absvalue3(un) + sqrt(dA) * a;
}

template <class T> struct HomOut;
template <> struct HomOut<double> {
  static __device__ __forceinline__ void put(double* out, double*, int64_t i, double a, double b, double d) {
    out[i] = (a == 0.0 ? 0.0 : a * out[i]) + b * d;
  }
};
template <> struct HomOut<dual<1>> {
  static __device__ __forceinline__ void put(double* out, double* jv, int64_t i, double a, double b,
                                             const dual<1>& d) {
    if (out) out[i] = (a == 0.0 ? 0.0 : a * out[i]) + b * d.v;
    jv[i] = (a == 0.0 ? 0.0 : a * jv[i]) + b * d.d[0];
  }
};

template <int DIM, class T>
__global__ void __launch_bounds__(64)
homotopy_diss_kernel(int np, int nf, int st, DevSBP sbp, DevMesh mesh, const double* __restrict__ w,
                     PhysConst phys, int bndry, const double* __restrict__ q, const double* __restrict__ v,
                     double a, double b, double* __restrict__ out, double* __restrict__ jv) {
  constexpr int ND = DIM + 2, NFACES = DIM + 1;
  extern __shared__ __align__(16) unsigned char hom_smem[];
  T* qs = reinterpret_cast<T*>(hom_smem);
  T* t2 = qs + np * ND;
  T* G = t2 + np * ND;
  const int64_t e = blockIdx.x;
  const int i = threadIdx.x;
  const bool node = i < np;
  for (int idx = i; idx < np * ND; idx += blockDim.x) qs[idx] = Io<T>::load(q, v, e * np * ND + idx);
  __syncthreads();
  T acc[ND];
#pragma unroll
  for (int c = 0; c < ND; ++c) acc[c] = 0.0;
  // volume: acc -= Q_d^T (lambda_d H^-1 Q_d q)
  for (int d1 = 0; d1 < DIM; ++d1) {
    if (node) {
      T t1[ND];
#pragma unroll
      for (int c = 0; c < ND; ++c) t1[c] = 0.0;
      for (int j = 0; j < np; ++j) {
        const double qij = sbp.Q[((int64_t)d1 * np + i) * np + j];
#pragma unroll
        for (int c = 0; c < ND; ++c) t1[c] = t1[c] + qij * qs[j * ND + c];
      }
      double dir[DIM];
#pragma unroll
      for (int x = 0; x < DIM; ++x)
        dir[x] = mesh.metric_per_node ? mesh.dxidx[((e * np + i) * DIM + x) * DIM + d1]
                                      : mesh.dxidx[(e * DIM + x) * DIM + d1];
      const T lam = hom_lambda_max<DIM>(phys, &qs[i * ND], dir);
      const double winv = 1.0 / w[i];
#pragma unroll
      for (int c = 0; c < ND; ++c) t2[i * ND + c] = lam * (t1[c] * winv);
    }
    __syncthreads();
    if (node)
      for (int j = 0; j < np; ++j) {
        const double qji = sbp.Q[((int64_t)d1 * np + j) * np + i];
#pragma unroll
        for (int c = 0; c < ND; ++c) acc[c] = acc[c] - qji * t2[j * ND + c];
      }
    __syncthreads();
  }
  // interior faces: G[f][m] is this element's B-weighted flux at its face node m
  for (int slot = i; slot < NFACES * nf; slot += blockDim.x) {
    const int f = slot / nf;
    const int k = slot - f * nf;
    const ElFace rec = mesh.elface[e * NFACES + f];
    if (rec.x < 0) {
      if (!bndry) {
#pragma unroll
        for (int c = 0; c < ND; ++c) G[(f * nf + k) * ND + c] = 0.0;
        continue;
      }
      T qb[ND], qa[ND];
      double qg[ND];
      free_stream<DIM>(phys, qg);
#pragma unroll
      for (int c = 0; c < ND; ++c) qb[c] = 0.0;
      for (int s = 0; s < st; ++s) {
        const double r = sbp.interp[s * nf + k];
        const int node = sbp.perm[s * NFACES + f];
#pragma unroll
        for (int c = 0; c < ND; ++c) qb[c] = qb[c] + r * qs[node * ND + c];
      }
#pragma unroll
      for (int c = 0; c < ND; ++c) qa[c] = 0.5 * (qb[c] + qg[c]);
      const T lam = hom_lambda_max<DIM>(phys, qa, mesh.nrm_bndry + ((int64_t)rec.y * nf + k) * DIM);
      const double wgt = -sbp.wface[k];
#pragma unroll
      for (int c = 0; c < ND; ++c) G[(f * nf + k) * ND + c] = wgt * (0.5 * lam * (qb[c] - qg[c]));
      continue;
    }
    const int64_t nb = rec.x;
    const int fN = rec.z & 0xff;
    const int o = (rec.z >> 8) & 0xff;
    const int side = (rec.z >> 16) & 1;
    const int kR = sbp.nbrperm[k * sbp.norient + o];
    T qL[ND], qR[ND], qa[ND];
#pragma unroll
    for (int c = 0; c < ND; ++c) { qL[c] = 0.0; qR[c] = 0.0; }
    for (int s = 0; s < st; ++s) {
      const double rl = sbp.interp[s * nf + k], rr = sbp.interp[s * nf + kR];
      if (side == 0) {
        const int nodeL = sbp.perm[s * NFACES + f], nodeR = sbp.perm[s * NFACES + fN];
#pragma unroll
        for (int c = 0; c < ND; ++c) {
          qL[c] = qL[c] + rl * qs[nodeL * ND + c];
          qR[c] = qR[c] + rr * Io<T>::load(q, v, (nb * np + nodeR) * ND + c);
        }
      } else {
        const int nodeL = sbp.perm[s * NFACES + fN], nodeR = sbp.perm[s * NFACES + f];
#pragma unroll
        for (int c = 0; c < ND; ++c) {
          qL[c] = qL[c] + rl * Io<T>::load(q, v, (nb * np + nodeL) * ND + c);
          qR[c] = qR[c] + rr * qs[nodeR * ND + c];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < ND; ++c) qa[c] = 0.5 * (qL[c] + qR[c]);
    const T lam = hom_lambda_max<DIM>(phys, qa, mesh.nrm_face + ((int64_t)rec.y * nf + k) * DIM);
    const int m = side == 0 ? k : kR;
    const double wgt = side == 0 ? -sbp.wface[k] : sbp.wface[kR];
#pragma unroll
    for (int c = 0; c < ND; ++c) G[(f * nf + m) * ND + c] = wgt * (0.5 * lam * (qL[c] - qR[c]));
  }
  __syncthreads();
  if (!node) return;
  for (int f = 0; f < NFACES; ++f) {
    const int s = sbp.nodepos[i * NFACES + f];
    if (s < 0) continue;
    for (int mm = 0; mm < nf; ++mm) {
      const double r = sbp.interp[s * nf + mm];
#pragma unroll
      for (int c = 0; c < ND; ++c) acc[c] = acc[c] + r * G[(f * nf + mm) * ND + c];
    }
  }
#pragma unroll
  for (int c = 0; c < ND; ++c) HomOut<T>::put(out, jv, (e * np + i) * ND + c, a, b, acc[c]);
}

template <int DIM, class T>
static int launch_hom(pdes_ctx* c, const double* q, const double* v, double a, double b, double* out, double* jv) {
  constexpr int ND = DIM + 2;
  const size_t lds = (size_t)(2 * c->Np * ND + (DIM + 1) * c->Nf * ND) * sizeof(T);
  hipLaunchKernelGGL((homotopy_diss_kernel<DIM, T>), dim3((unsigned)c->numEl), dim3(64), lds, c->stream, c->Np,
                     c->Nf, c->stencil, c->devsbp(), c->devmesh(), c->d_w, c->pc, c->hom_bndry, q, v, a, b, out, jv);
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

int launch_homotopy(pdes_ctx* c, const double* q, const double* v, double a, double b, double* out, double* jv) {
  PDES_CHECK(c->ninst == 1 && c->nlist < 0, PDES_ERR_UNSUPPORTED, "the homotopy dissipation is for single contexts");
  PDES_CHECK(c->Np <= 64, PDES_ERR_UNSUPPORTED, "the homotopy dissipation supports up to 64 nodes per element");
  PDES_CHECK(v == nullptr || jv != nullptr, PDES_ERR_ARG, "a direction needs a product vector");
  if (!c->d_w) {
    if (hipMalloc((void**)&c->d_w, c->Np * sizeof(double)) != hipSuccess) {
      c->d_w = nullptr;
      set_error("device allocation for the SBP weights failed");
      return PDES_ERR_NOMEM;
    }
    PDES_HIP(hipMemcpyAsync(c->d_w, c->h_w.data(), c->Np * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  if (c->numEl == 0) return PDES_OK;
  if (v) return c->dim == 2 ? launch_hom<2, dual<1>>(c, q, v, a, b, out, jv) : launch_hom<3, dual<1>>(c, q, v, a, b, out, jv);
  return c->dim == 2 ? launch_hom<2, double>(c, q, nullptr, a, b, out, nullptr)
                     : launch_hom<3, double>(c, q, nullptr, a, b, out, nullptr);
}

}  // namespace pdes

using namespace pdes;

extern "C" {

int pdes_eval_homotopy(pdes_ctx* c, double* res) {
  PDES_CHECK(c && res, PDES_ERR_ARG, "null argument");
  int st = launch_homotopy(c, c->d_q, nullptr, 0.0, 1.0, c->d_work, nullptr);
  if (st) return st;
  PDES_HIP(hipMemcpyAsync(res, c->d_work, c->numDof * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

int pdes_set_homotopy_boundary(pdes_ctx* c, int add_boundary_integrals) {
  PDES_CHECK(c, PDES_ERR_ARG, "null context");
  c->hom_bndry = add_boundary_integrals ? 1 : 0;
  return PDES_OK;
}

int pdes_homotopy_jv(pdes_ctx* c, const double* v, double* b) {
  PDES_CHECK(c && v && b, PDES_ERR_ARG, "null argument");
  PDES_HIP(hipMemcpyAsync(c->d_work, v, c->numDof * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int st = launch_homotopy(c, c->d_q, c->d_work, 0.0, 1.0, nullptr, c->d_work2);
  if (st) return st;
  PDES_HIP(hipMemcpyAsync(b, c->d_work2, c->numDof * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

}  // extern "C"
