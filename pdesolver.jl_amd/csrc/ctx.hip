// ctx.hip -- context lifetime, table upload and state transfer (C ABI).
#include <cstdarg>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "ctx.hpp"

namespace pdes {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

template <class T>
static int upload(T** dst, const T* src, size_t n, hipStream_t s) {
  if (n == 0) { *dst = nullptr; return PDES_OK; }
  hipError_t e = hipMalloc((void**)dst, n * sizeof(T));
  if (e != hipSuccess) { set_error("hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e)); return PDES_ERR_NOMEM; }
  if (src) PDES_HIP(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
  return PDES_OK;
}

int sync_check_flag(pdes_ctx* c) {
  int32_t f = 0;
  PDES_HIP(hipMemcpyAsync(&f, c->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  if (f) {
    PDES_HIP(hipMemsetAsync(c->d_flag, 0, sizeof(int32_t), c->stream));
    set_error("non-positive density or pressure in the state (checkDensity/checkPressure)");
    return PDES_ERR_PHYS;
  }
  return PDES_OK;
}

// boundary states of all boundary face nodes into d_qbc (re-run whenever the
// element-face records or the per-instance physics change)
int bc_state_setup(pdes_ctx* c) {
  const int64_t n = (int64_t)c->numEl * c->nfaces * c->Nf;
  if (n == 0 || !c->d_qbc) return PDES_OK;
  const dim3 grid((unsigned)ceil_div(n, 256));
  if (c->dim == 2)
    hipLaunchKernelGGL(bc_state_kernel<2>, grid, dim3(256), 0, c->stream, c->devmesh(), c->pc, c->nfaces, c->Nf,
                       c->d_qbc);
  else
    hipLaunchKernelGGL(bc_state_kernel<3>, grid, dim3(256), 0, c->stream, c->devmesh(), c->pc, c->nfaces, c->Nf,
                       c->d_qbc);
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}
}  // namespace pdes

using namespace pdes;

extern "C" {

const char* pdes_last_error(void) { return g_err; }
int pdes_version(void) { return 1; }

int pdes_ctx_create(const pdes_mesh_desc* m, const pdes_sbp_desc* s, const pdes_phys_desc* ph,
                    pdes_ctx** out) {
  PDES_CHECK(m && s && ph && out, PDES_ERR_ARG, "null argument");
  PDES_CHECK(m->dim == 2 || m->dim == 3, PDES_ERR_ARG, "dim must be 2 or 3");
  PDES_CHECK(m->numEl > 0 && s->Np > 0 && s->Nf > 0, PDES_ERR_ARG, "empty mesh or operator");
  pdes_ctx* c = new pdes_ctx();
  c->use_graphs = std::getenv("PDES_NO_GRAPHS") == nullptr;
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    set_error("hipGetDevice failed");
    return PDES_ERR_HIP;
  }
  c->dim = m->dim;
  c->nd = m->dim + 2;
  c->nfaces = m->dim + 1;
  c->numEl = m->numEl;
  c->Np = s->Np;
  c->Nf = s->Nf;
  c->stencil = s->stencil;
  c->norient = s->norient;
  c->diag_e = s->diag_e;
  c->numIface = m->numInterfaces;
  c->numBndry = m->numBoundaryFaces;
  c->numBC = m->numBC;
  c->numDof = (int64_t)m->numEl * s->Np * c->nd;
  c->phys = *ph;
  c->pc.gamma = ph->gamma;
  c->pc.gamma_1 = ph->gamma - 1;
  c->pc.R = ph->R_gas;
  c->pc.cv = ph->R_gas / (ph->gamma - 1);
  c->pc.Ma = ph->Ma;
  c->pc.aoa = ph->aoa;
  c->pc.rho_free = ph->rho_free;
  c->pc.E_free = ph->p_free / (ph->gamma - 1) + 0.5 * ph->rho_free * ph->Ma * ph->Ma;  // types.jl:252
  c->pc.vortex_Min = ph->vortex_Min;
  // default launch variants measured on MI355X: the split form at 3 waves per
  // SIMD; the weak form at 4, with Q^T F on the matrix cores where its product
  // lays over the flux planes in LDS ((dim - 1) Np >= nfaces Nf), else on the
  // vector ALUs (the product's own LDS array costs a block per CU)
  if (ph->volume_integral_type == 2)
    c->res_variant[0] = c->res_variant[1] = 0;
  else
    c->res_variant[0] = c->res_variant[1] = ((c->dim - 1) * c->Np >= c->nfaces * c->Nf) ? 5 : 1;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    delete c;
    return PDES_ERR_HIP;
  }
  c->own_stream = true;
  const int dim = c->dim, Np = c->Np, Nf = c->Nf, nfaces = c->nfaces;

  // host copies
  c->h_w.assign(s->w, s->w + Np);
  c->h_Q.assign(s->Q, s->Q + (size_t)dim * Np * Np);
  c->h_interp.assign(s->interp, s->interp + (size_t)s->stencil * Nf);
  c->h_perm.assign(s->perm, s->perm + (size_t)s->stencil * nfaces);
  c->h_nbrperm.assign(s->nbrperm, s->nbrperm + (size_t)Nf * s->norient);
  c->h_wface.assign(s->wface, s->wface + Nf);
  c->h_jac.assign(m->jac, m->jac + (size_t)m->numEl * Np);
  c->h_interfaces.assign(m->interfaces, m->interfaces + (size_t)5 * m->numInterfaces);
  c->h_bndryfaces.assign(m->bndryfaces, m->bndryfaces + (size_t)2 * m->numBoundaryFaces);
  c->h_bndry_offsets.assign(m->bndry_offsets, m->bndry_offsets + m->numBC + 1);
  c->h_bc_type.assign(m->bc_type, m->bc_type + m->numBC);
  c->h_nrm_bndry.assign(m->nrm_bndry, m->nrm_bndry + (size_t)m->numBoundaryFaces * Nf * dim);
  c->h_coords_bndry.assign(m->coords_bndry, m->coords_bndry + (size_t)m->numBoundaryFaces * Nf * dim);

  for (size_t x = 0; x < c->h_nbrperm.size(); ++x)
    if (c->h_nbrperm[x] < 0 || c->h_nbrperm[x] >= Nf) {
      set_error("nbrperm entry out of range");
      pdes_ctx_destroy(c);
      return PDES_ERR_ARG;
    }
  PDES_CHECK(s->stencil == (s->diag_e ? Nf : Np) && s->norient >= 1, PDES_ERR_ARG,
             "stencil must equal Nf (diagonal E) or Np (dense face)");
  // node -> stencil position per face
  c->h_nodepos.assign((size_t)Np * nfaces, -1);
  for (int st = 0; st < s->stencil; ++st)
    for (int f = 0; f < nfaces; ++f) {
      int node = s->perm[st * nfaces + f];
      if (node < 0 || node >= Np) { set_error("perm entry out of range"); pdes_ctx_destroy(c); return PDES_ERR_ARG; }
      c->h_nodepos[(size_t)node * nfaces + f] = st;
    }

  // element -> face connectivity
  c->h_elface.assign((size_t)m->numEl * nfaces, ElFace{-2, -1, 0, 0});
  for (int i = 0; i < m->numInterfaces; ++i) {
    const int32_t* it = m->interfaces + 5 * i;
    int eL = it[0], eR = it[1], fL = it[2], fR = it[3], o = it[4];
    if (eL < 0 || eL >= m->numEl || eR < 0 || eR >= m->numEl || fL < 0 || fL >= nfaces || fR < 0 ||
        fR >= nfaces || o < 0 || o >= s->norient) {
      set_error("interface %d out of range", i);
      pdes_ctx_destroy(c);
      return PDES_ERR_ARG;
    }
    c->h_elface[(size_t)eL * nfaces + fL] = ElFace{eR, i, fR | (o << 8), 0};
    c->h_elface[(size_t)eR * nfaces + fR] = ElFace{eL, i, fL | (o << 8) | (1 << 16), 0};
  }
  for (int bc = 0; bc < m->numBC; ++bc)
    for (int b = m->bndry_offsets[bc]; b < m->bndry_offsets[bc + 1]; ++b) {
      int e = m->bndryfaces[2 * b], f = m->bndryfaces[2 * b + 1];
      if (e < 0 || e >= m->numEl || f < 0 || f >= nfaces) {
        set_error("boundary face %d out of range", b);
        pdes_ctx_destroy(c);
        return PDES_ERR_ARG;
      }
      c->h_elface[(size_t)e * nfaces + f] = ElFace{-1, b, 0, m->bc_type[bc]};
    }
  for (size_t x = 0; x < c->h_elface.size(); ++x)
    if (c->h_elface[x].x == -2) {
      set_error("element %zu face %zu is neither an interface nor a boundary face", x / nfaces, x % nfaces);
      pdes_ctx_destroy(c);
      return PDES_ERR_ARG;
    }

  // inverse mass matrix, calcMassMatrixInverse (src/Utils/mass_matrix.jl:20-55)
  // (per DOF as the reference stores it, and once per node for the RK stage
  // epilogue, which would otherwise read the same value ndof times)
  std::vector<double> Minv((size_t)c->numDof), Minv_node((size_t)m->numEl * Np);
  for (int e = 0; e < m->numEl; ++e)
    for (int j = 0; j < Np; ++j) {
      double v = 1.0 / (s->w[j] / m->jac[(size_t)e * Np + j]);
      Minv_node[(size_t)e * Np + j] = v;
      for (int k = 0; k < c->nd; ++k) Minv[((size_t)e * Np + j) * c->nd + k] = v;
    }

  // globalization diagonal M_i (1 + sqrt(jac_i)) = (w/jac)(1 + sqrt(jac)): with
  // tau_vec = tau/(1 + sqrt(jac)) (calcTauVec, globalization.jl:181-201) the
  // term M/tau_vec is this vector times 1/tau
  std::vector<double> ptc((size_t)m->numEl * Np);
  for (int e = 0; e < m->numEl; ++e)
    for (int j = 0; j < Np; ++j) {
      const double jac = m->jac[(size_t)e * Np + j];
      ptc[(size_t)e * Np + j] = (s->w[j] / jac) * (1.0 + std::sqrt(jac));
    }

  int st = PDES_OK;
  hipStream_t sm = c->stream;
  const size_t nI = m->numInterfaces, nB = m->numBoundaryFaces;
  // affine elements (the metric is bitwise identical at every node): keep one
  // metric block per element in HBM instead of one per node
  {
    const size_t blk = (size_t)dim * dim;
    bool affine = true;
    for (int e = 0; e < m->numEl && affine; ++e)
      for (int j = 1; j < Np && affine; ++j)
        affine = std::memcmp(m->dxidx + ((size_t)e * Np + j) * blk, m->dxidx + (size_t)e * Np * blk,
                             blk * sizeof(double)) == 0;
    c->metric_per_node = affine ? 0 : 1;
    if (affine) {
      std::vector<double> compact((size_t)m->numEl * blk);
      for (int e = 0; e < m->numEl; ++e)
        std::memcpy(&compact[(size_t)e * blk], m->dxidx + (size_t)e * Np * blk, blk * sizeof(double));
      if (!st) st = upload(&c->d_dxidx, compact.data(), compact.size(), sm);
      if (!st) st = (hipStreamSynchronize(sm) == hipSuccess) ? PDES_OK : PDES_ERR_HIP;
    } else {
      if (!st) st = upload(&c->d_dxidx, m->dxidx, (size_t)m->numEl * Np * blk, sm);
    }
  }
  if (!st) st = upload(&c->d_elface, c->h_elface.data(), c->h_elface.size(), sm);
  if (s->diag_e) {
    // face slot records (FaceSlot): the neighbour's face node as a global node
    // index and the normal row, so the residual's face phase does not walk
    // record -> nbrperm -> perm -> state on the device
    std::vector<FaceSlot> slots((size_t)m->numEl * nfaces * Nf);
    for (int e = 0; e < m->numEl; ++e)
      for (int f = 0; f < nfaces; ++f) {
        const ElFace& r = c->h_elface[(size_t)e * nfaces + f];
        for (int k = 0; k < Nf; ++k) {
          FaceSlot& sl = slots[((size_t)e * nfaces + f) * Nf + k];
          if (r.x < 0) {
            const int64_t row = (int64_t)r.y * Nf + k;
            sl = FaceSlot{(int32_t)(-1 - row), slot_info(row, k, 0)};
          } else {
            const int fN = r.z & 0xff, o = (r.z >> 8) & 0xff, side = (r.z >> 16) & 1;
            const int kR = s->nbrperm[k * s->norient + o];
            const int nodeN = s->perm[(side == 0 ? kR : k) * nfaces + fN];
            sl = FaceSlot{(int32_t)((int64_t)r.x * Np + nodeN), slot_info((int64_t)r.y * Nf + k, side == 0 ? k : kR, side)};
          }
        }
      }
    if (!((int64_t)m->numEl * Np < ((int64_t)1 << 31) && (int64_t)std::max(nI, nB) * Nf < ((int64_t)1 << 27))) {
      set_error("mesh too large for the face slot records");
      st = PDES_ERR_ARG;
    }
    if (!st) st = upload(&c->d_slot, slots.data(), slots.size(), sm);
    if (!st) st = (hipStreamSynchronize(sm) == hipSuccess) ? PDES_OK : PDES_ERR_HIP;
  }
  if (!st) st = upload(&c->d_nrm_face, m->nrm_face, nI * Nf * dim, sm);
  if (!st) st = upload(&c->d_nrm_bndry, m->nrm_bndry, nB * Nf * dim, sm);
  if (!st) st = upload(&c->d_coords_bndry, m->coords_bndry, nB * Nf * dim, sm);
  if (!st) st = upload(&c->d_Q, s->Q, (size_t)dim * Np * Np, sm);
  {
    // 2 S_d[a][a + dd mod Np] = Q_d[a][b] - Q_d[b][a] for dd = 0..Np/2, the
    // order in which the split-form pair rounds read it
    const int ns = Np / 2 + 1;
    std::vector<double> s2c((size_t)dim * ns * Np);
    for (int d = 0; d < dim; ++d)
      for (int dd = 0; dd < ns; ++dd)
        for (int a = 0; a < Np; ++a) {
          const int b = (a + dd) % Np;
          s2c[((size_t)d * ns + dd) * Np + a] = s->Q[((size_t)d * Np + a) * Np + b] - s->Q[((size_t)d * Np + b) * Np + a];
        }
    if (!st) st = upload(&c->d_S2c, s2c.data(), s2c.size(), sm);
    if (!st) st = (hipStreamSynchronize(sm) == hipSuccess) ? PDES_OK : PDES_ERR_HIP;
  }
  if (!st) st = upload(&c->d_interp, s->interp, (size_t)s->stencil * Nf, sm);
  if (!st) st = upload(&c->d_perm, s->perm, (size_t)s->stencil * nfaces, sm);
  if (!st) st = upload(&c->d_nbrperm, s->nbrperm, (size_t)Nf * s->norient, sm);
  if (!st) st = upload(&c->d_nodepos, c->h_nodepos.data(), c->h_nodepos.size(), sm);
  if (!st) st = upload(&c->d_wface, s->wface, (size_t)Nf, sm);
  if (!st) st = upload(&c->d_Minv, Minv.data(), Minv.size(), sm);
  if (!st) st = upload(&c->d_Minv_node, Minv_node.data(), Minv_node.size(), sm);
  if (!st) st = upload(&c->d_ptc_diag, ptc.data(), ptc.size(), sm);
  if (!st) st = upload<double>(&c->d_ptc_scale, nullptr, 1, sm);
  if (!st) st = upload<double>(&c->d_q, nullptr, (size_t)c->numDof, sm);
  if (!st) st = upload<double>(&c->d_res, nullptr, (size_t)c->numDof, sm);
  if (!st) st = upload<double>(&c->d_work, nullptr, (size_t)c->numDof, sm);
  if (!st) st = upload<double>(&c->d_work2, nullptr, (size_t)c->numDof, sm);
  if (!st) st = upload<double>(&c->d_partial, nullptr, 4096, sm);
  if (!st) st = upload<double>(&c->d_scalar, nullptr, 64, sm);
  if (!st) st = upload<int32_t>(&c->d_flag, nullptr, 1, sm);
  if (!st && hipMemsetAsync(c->d_flag, 0, sizeof(int32_t), sm) != hipSuccess) st = PDES_ERR_HIP;
  if (!st) st = upload<double>(&c->d_qbc, nullptr, std::max<size_t>(1, nB * Nf * c->nd), sm);
  if (!st) st = bc_state_setup(c);
  if (!st && hipStreamSynchronize(sm) != hipSuccess) { set_error("upload failed"); st = PDES_ERR_HIP; }
  if (st) {
    pdes_ctx_destroy(c);
    return st;
  }
  *out = c;
  return PDES_OK;
}

// estimateMem (src/estimateMem.jl:1-74) for the device: the bytes one steady
// solve allocates, summed from the mesh and operator counts exactly as the
// allocations are sized (context arrays, BSR Jacobian, preconditioner, GMRES
// basis and Newton vectors; the per-instance share of a packed batch when
// batched is set).  The metric is counted per node (the compact per-element
// form is only known after the bitwise affinity check).
int pdes_estimate_bytes(const pdes_mesh_desc* m, const pdes_sbp_desc* s, int fp64_blocks, int pc_kind,
                        int gmres_restart, int batched, int64_t* bytes) {
  PDES_CHECK(m && s && bytes && gmres_restart > 0, PDES_ERR_ARG, "bad argument");
  const int64_t dim = m->dim, nd = dim + 2, nfaces = dim + 1, Np = s->Np, Nf = s->Nf;
  const int64_t nel = m->numEl, nI = m->numInterfaces, nB = m->numBoundaryFaces;
  const int64_t ndof = nel * Np * nd;
  int64_t b = 0;
  // context
  b += nel * Np * dim * dim * 8;                       // dxidx
  b += nel * nfaces * 16;                              // element -> face records
  b += (nI + 2 * nB) * Nf * dim * 8;                   // face normals, boundary normals and coordinates
  b += (dim * Np * Np + s->stencil * Nf + Nf) * 8 + (s->stencil * nfaces + Nf * s->norient + Np * nfaces) * 4;
  b += ndof * 8 * 5 + nel * Np * 8;                    // Minv (per DOF and per node), q, res, two work vectors
  b += nel * Np * 8;                                   // implicit-Euler diagonal
  b += (4096 + 64 + 8) * 8;                            // reduction scratch, scalars, flags
  // BSR Jacobian: element blocks plus the face couplings of every interface
  const int64_t cpl = s->diag_e ? Nf : (int64_t)s->stencil * s->stencil;
  const int64_t nnzb = nel * Np * Np + 2 * nI * cpl;
  b += (nel * Np + 1) * 8 + nnzb * 4 + nnzb * nd * nd * (fp64_blocks ? 8 : 4) + 4;
  // preconditioner: one FP32 inverse per element (block-Jacobi and ILU(0))
  if (pc_kind != PDES_PC_NONE) b += nel * (Np * nd) * (Np * nd) * 4;
  if (pc_kind == PDES_PC_ILU0) {
    b += 2 * nel * 4 + 2 * ndof * 8;   // level lists, graph staging
    // own-block offsets per node row, the padded chunk lists (a level padded
    // to whole chunks of at most 4 elements: at most 4 entries per element)
    // and the dependency-driven sweeps' vector
    b += nel * Np * 4 + 2 * 4 * nel * 4 + ndof * 8;
  }
  // GMRES basis and vectors (multi-dot partials: 2048 workgroups per basis
  // vector), Newton step and right-hand side
  b += (gmres_restart + 1) * ndof * 8 + 2 * ndof * 8 + (int64_t)(gmres_restart + 2) * 2048 * 8;
  b += 2 * ndof * 8;
  if (batched) b += 2 * nel * 4 + 2 * ndof * 8 + (ndof / 4096 + 1) * (8 + 4 + 4 + 64 * 8);
  *bytes = b;
  return PDES_OK;
}

int pdes_ctx_destroy(pdes_ctx* c) {
  if (!c) return PDES_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  jacobian_free(c);
  precond_free(c);
  void* ptrs[] = {c->d_dxidx, c->d_elface, c->d_slot, c->d_nrm_face, c->d_nrm_bndry, c->d_coords_bndry, c->d_Q, c->d_S2c,
                  c->d_interp, c->d_perm, c->d_nbrperm, c->d_nodepos, c->d_wface, c->d_Minv, c->d_Minv_node, c->d_q,
                  c->d_res, c->d_work, c->d_work2, c->d_partial, c->d_scalar, c->d_flag,
                  c->d_el_list, c->d_el_inst, c->d_inst_phys, c->d_ptc_diag, c->d_ptc_scale,
                  c->d_inst_active, c->d_qbc, c->d_w};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
  delete c;
  return PDES_OK;
}

int pdes_ctx_numdof(const pdes_ctx* c, int64_t* numdof) {
  PDES_CHECK(c && numdof, PDES_ERR_ARG, "null argument");
  *numdof = c->numDof;
  return PDES_OK;
}

int pdes_set_state(pdes_ctx* c, const double* q) {
  PDES_CHECK(c && q, PDES_ERR_ARG, "null argument");
  PDES_HIP(hipMemcpyAsync(c->d_q, q, c->numDof * sizeof(double), hipMemcpyHostToDevice, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

int pdes_get_state(pdes_ctx* c, double* q) {
  PDES_CHECK(c && q, PDES_ERR_ARG, "null argument");
  PDES_HIP(hipMemcpyAsync(q, c->d_q, c->numDof * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

int pdes_get_residual(pdes_ctx* c, double* res) {
  PDES_CHECK(c && res, PDES_ERR_ARG, "null argument");
  PDES_HIP(hipMemcpyAsync(res, c->d_res, c->numDof * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return sync_check_flag(c);
}

int pdes_state_ptr(pdes_ctx* c, double** q_dev, double** res_dev) {
  PDES_CHECK(c, PDES_ERR_ARG, "null context");
  if (q_dev) *q_dev = c->d_q;
  if (res_dev) *res_dev = c->d_res;
  return PDES_OK;
}

int pdes_sync(pdes_ctx* c) {
  PDES_CHECK(c, PDES_ERR_ARG, "null context");
  PDES_HIP(hipStreamSynchronize(c->stream));
  return sync_check_flag(c);
}

}  // extern "C"
