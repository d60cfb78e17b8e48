// jacobian.hip -- BSR pattern, explicit Jacobian assembly and BSR products (C ABI).
#include <algorithm>

#include "jacobian.hpp"

namespace pdes {

int jacobian_free(pdes_ctx* c) {
  if (!c->jac) return PDES_OK;
  pdes_jacobian* j = c->jac;
  if (j->d_rowptr) hipFree(j->d_rowptr);
  if (j->d_colind) hipFree(j->d_colind);
  if (j->d_vals) hipFree(j->d_vals);
  if (j->d_err) hipFree(j->d_err);
  if (j->d_color_el) hipFree(j->d_color_el);
  if (j->d_seed) hipFree(j->d_seed);
  if (j->d_jv) hipFree(j->d_jv);
  if (j->d_tpos) hipFree(j->d_tpos);
  delete j;
  c->jac = nullptr;
  return PDES_OK;
}

// Block sparsity of SparseMatrixCSC(mesh, Tv, INVISCID, face_type)
// (src/jacobian/sparse.jl:250-495): every node couples to all nodes of its
// element; across an interface, DenseFace couples all stencil nodes of both
// sides, SparseFace couples each face node with its nbrperm partner.
static void build_pattern(const pdes_ctx* c, std::vector<int64_t>& rowptr, std::vector<int32_t>& colind) {
  const int Np = c->Np, Nf = c->Nf, nfaces = c->nfaces;
  const int64_t nrows = (int64_t)c->numEl * Np;
  // inverse nbrperm per orientation
  std::vector<int32_t> inv((size_t)Nf * c->norient);
  for (int o = 0; o < c->norient; ++o)
    for (int k = 0; k < Nf; ++k) inv[(size_t)c->h_nbrperm[(size_t)k * c->norient + o] * c->norient + o] = k;
  rowptr.assign(nrows + 1, 0);
  std::vector<int32_t> cols;
  cols.reserve(Np * (nfaces + 1));
  std::vector<std::vector<int32_t>> rows_tmp;  // unused; rows are built in order
  colind.clear();
  for (int e = 0; e < c->numEl; ++e) {
    for (int i = 0; i < Np; ++i) {
      cols.clear();
      for (int j = 0; j < Np; ++j) cols.push_back(e * Np + j);
      for (int f = 0; f < nfaces; ++f) {
        const ElFace r = c->h_elface[(size_t)e * nfaces + f];
        if (r.x < 0) continue;
        const int pos = c->h_nodepos[(size_t)i * nfaces + f];
        if (pos < 0) continue;
        const int fN = r.z & 0xff, o = (r.z >> 8) & 0xff, side = (r.z >> 16) & 1;
        if (!c->diag_e) {
          for (int s = 0; s < c->stencil; ++s) cols.push_back(r.x * Np + c->h_perm[(size_t)s * nfaces + fN]);
        } else {
          // own face node pos; partner face node on the neighbour
          int kn;
          if (side == 0) kn = c->h_nbrperm[(size_t)pos * c->norient + o];
          else kn = inv[(size_t)pos * c->norient + o];
          cols.push_back(r.x * Np + c->h_perm[(size_t)kn * nfaces + fN]);
        }
      }
      std::sort(cols.begin(), cols.end());
      const int64_t row = (int64_t)e * Np + i;
      rowptr[row + 1] = rowptr[row] + (int64_t)cols.size();
      colind.insert(colind.end(), cols.begin(), cols.end());
    }
  }
}

template <class Op, int VOL, int FLUX, class BT>
static int launch_jac(pdes_ctx* c) {
  pdes_jacobian* j = c->jac;
  if (c->active_count() == 0) return PDES_OK;
  hipLaunchKernelGGL((euler_jacobian_kernel<Op, VOL, FLUX, BT>), dim3((unsigned)c->active_count()), dim3(256), 0, c->stream,
                     c->devmesh(), c->devsbp(), c->pc, c->d_q, j->d_rowptr, j->d_colind,
                     (BT*)j->d_vals, j->d_err, c->d_ptc_diag, c->ptc_on ? c->d_ptc_scale : nullptr);
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

template <class Op, class BT>
static int launch_jac_shape(pdes_ctx* c) {
  const int vol = c->phys.volume_integral_type, flux = c->phys.face_flux;
  if (vol == 1 && flux == PDES_FLUX_ROE) return launch_jac<Op, VOL_WEAK, PDES_FLUX_ROE, BT>(c);
  if constexpr (Op::DIAGE && Op::DIM == 2) {
    if (vol == 2 && flux == PDES_FLUX_IRSLF) return launch_jac<Op, VOL_SPLIT, PDES_FLUX_IRSLF, BT>(c);
  }
  if constexpr (!Op::DIAGE) {
    if (vol == 2 && flux == PDES_FLUX_ESLF) return launch_jac<Op, VOL_SPLIT, PDES_FLUX_ESLF, BT>(c);
  }
  set_error("unsupported physics combination for the explicit Jacobian");
  return PDES_ERR_UNSUPPORTED;
}

int launch_assemble(pdes_ctx* c) {
#define PDES_DISPATCH(D, NP, NF, ST, DE)                                                        \
  if (c->dim == D && c->Np == NP && c->Nf == NF && c->stencil == ST && (c->diag_e != 0) == DE) { \
    using Op = OpShape<D, NP, NF, ST, DE>;                                                      \
    if (c->jac->fp64) return launch_jac_shape<Op, double>(c);                                   \
    return launch_jac_shape<Op, float>(c);                                                      \
  }
  PDES_FOR_EACH_SHAPE(PDES_DISPATCH)
#undef PDES_DISPATCH
  set_error("no compiled Jacobian kernel for this operator");
  return PDES_ERR_UNSUPPORTED;
}

template <int ND>
static int launch_spmv_nd(pdes_ctx* c, const double* x, double* y, double alpha, const double* yadd, double beta) {
  pdes_jacobian* j = c->jac;
  const int64_t nrows = c->active_count() * c->Np;
  if (nrows == 0) return PDES_OK;
  const int64_t n = nrows * ND;
  const int threads = 256;
  const int64_t blocks = ceil_div(n, threads);
  if (j->fp64)
    hipLaunchKernelGGL((bsr_spmv_kernel<ND, double>), dim3((unsigned)blocks), dim3(threads), 0, c->stream, nrows,
                       j->d_rowptr, j->d_colind, (const double*)j->d_vals, x, y, alpha, yadd, beta,
                       c->active_list(), c->Np, c->d_el_inst, c->d_mask);
  else
    hipLaunchKernelGGL((bsr_spmv_kernel<ND, float>), dim3((unsigned)blocks), dim3(threads), 0, c->stream, nrows,
                       j->d_rowptr, j->d_colind, (const float*)j->d_vals, x, y, alpha, yadd, beta,
                       c->active_list(), c->Np, c->d_el_inst, c->d_mask);
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

// y -= s_inst * ptc_diag[node] * x over the active elements (the implicit-Euler
// term of a matrix-free operator, applyEuler for vectors in globalization.jl)
__global__ void ptc_shift_kernel(const double* __restrict__ x, double* __restrict__ y, const double* __restrict__ ptc_diag,
                                 const double* __restrict__ ptc_scale, const int32_t* __restrict__ el_inst,
                                 const int32_t* __restrict__ el_list, int64_t nel, int np, int nd) {
  const int64_t n = nel * np * nd;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const int64_t s = t / (np * nd);
    const int64_t e = el_list ? (int64_t)el_list[s] : s;
    const int64_t k = e * np * nd + (t - s * np * nd);
    const double sc = ptc_scale[el_inst ? el_inst[e] : 0];
    y[k] -= sc * ptc_diag[k / nd] * x[k];
  }
}

int set_ptc_scale(pdes_ctx* c, const double* tau, int n) {
  std::vector<double> inv(n);
  for (int i = 0; i < n; ++i) inv[i] = 1.0 / tau[i];
  PDES_HIP(hipMemcpyAsync(c->d_ptc_scale, inv.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  c->ptc_on = true;
  return PDES_OK;
}

int launch_ptc_shift(pdes_ctx* c, const double* x, double* y) {
  if (!c->ptc_on) return PDES_OK;
  const int64_t n = c->active_count() * c->Np * c->nd;
  if (n == 0) return PDES_OK;
  hipLaunchKernelGGL(ptc_shift_kernel, dim3((unsigned)std::min<int64_t>(16384, ceil_div(n, 256))), dim3(256), 0,
                     c->stream, x, y, c->d_ptc_diag, c->d_ptc_scale, c->d_el_inst, c->active_list(),
                     c->active_count(), c->Np, c->nd);
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

// y = J^T x: row r of J^T holds, at the pattern position of (r, c), the
// transpose of the block stored at (c, r)
template <int ND, class BT>
__global__ void __launch_bounds__(256)
bsr_spmv_t_kernel(int64_t nrows, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
                  const int64_t* __restrict__ tpos, const BT* __restrict__ vals, const double* __restrict__ x,
                  double* __restrict__ y) {
  const int64_t tt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = tt / ND;
  const int a = (int)(tt - row * ND);
  if (row >= nrows) return;
  double acc = 0.0;
  for (int64_t b = rowptr[row]; b < rowptr[row + 1]; ++b) {
    const BT* blk = vals + tpos[b] * (ND * ND) + a;   // column a of the block at (c, r)
    const double* xc = x + (int64_t)colind[b] * ND;
#pragma unroll
    for (int c = 0; c < ND; ++c) acc += (double)blk[c * ND] * xc[c];
  }
  y[tt] = acc;
}

// J <- J^T: the block at p and the block at its transposed position swap,
// each transposed; one thread per pair (the lower position), so no conflicts
template <int ND, class BT>
__global__ void bsr_transpose_kernel(int64_t nnzb, const int64_t* __restrict__ tpos, BT* __restrict__ vals) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nnzb) return;
  const int64_t q = tpos[p];
  if (q < p) return;
  BT* A = vals + p * (ND * ND);
  BT* B = vals + q * (ND * ND);
  if (q == p) {
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int c = a + 1; c < ND; ++c) {
        const BT t = A[a * ND + c];
        A[a * ND + c] = A[c * ND + a];
        A[c * ND + a] = t;
      }
    return;
  }
  BT ta[ND * ND];
#pragma unroll
  for (int k = 0; k < ND * ND; ++k) ta[k] = A[k];
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int c = 0; c < ND; ++c) A[a * ND + c] = B[c * ND + a];
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int c = 0; c < ND; ++c) B[a * ND + c] = ta[c * ND + a];
}

static int ensure_tpos(pdes_ctx* c) {
  pdes_jacobian* j = c->jac;
  if (j->d_tpos) return PDES_OK;
  std::vector<int64_t> tpos((size_t)j->nnzb);
  for (int64_t r = 0; r < j->nrows; ++r)
    for (int64_t b = j->h_rowptr[r]; b < j->h_rowptr[r + 1]; ++b) {
      const int64_t cc = j->h_colind[b];
      const auto beg = j->h_colind.begin() + j->h_rowptr[cc], end = j->h_colind.begin() + j->h_rowptr[cc + 1];
      const auto it = std::lower_bound(beg, end, (int32_t)r);
      PDES_CHECK(it != end && *it == (int32_t)r, PDES_ERR_UNSUPPORTED, "block pattern is not structurally symmetric");
      tpos[(size_t)b] = (int64_t)(it - j->h_colind.begin());
    }
  PDES_CHECK(hipMalloc((void**)&j->d_tpos, std::max<int64_t>(1, j->nnzb) * sizeof(int64_t)) == hipSuccess,
             PDES_ERR_NOMEM, "device allocation of the transposed block positions failed");
  PDES_HIP(hipMemcpyAsync(j->d_tpos, tpos.data(), tpos.size() * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

int launch_spmv_transpose(pdes_ctx* c, const double* x, double* y) {
  PDES_CHECK(c->jac && c->jac->assembled, PDES_ERR_ARG, "Jacobian not assembled");
  int st = ensure_tpos(c);
  if (st) return st;
  pdes_jacobian* j = c->jac;
  const int64_t n = j->nrows * c->nd;
  const dim3 grid((unsigned)ceil_div(n, 256));
  if (c->nd == 4) {
    if (j->fp64)
      hipLaunchKernelGGL((bsr_spmv_t_kernel<4, double>), grid, dim3(256), 0, c->stream, j->nrows, j->d_rowptr,
                         j->d_colind, j->d_tpos, (const double*)j->d_vals, x, y);
    else
      hipLaunchKernelGGL((bsr_spmv_t_kernel<4, float>), grid, dim3(256), 0, c->stream, j->nrows, j->d_rowptr,
                         j->d_colind, j->d_tpos, (const float*)j->d_vals, x, y);
  } else {
    if (j->fp64)
      hipLaunchKernelGGL((bsr_spmv_t_kernel<5, double>), grid, dim3(256), 0, c->stream, j->nrows, j->d_rowptr,
                         j->d_colind, j->d_tpos, (const double*)j->d_vals, x, y);
    else
      hipLaunchKernelGGL((bsr_spmv_t_kernel<5, float>), grid, dim3(256), 0, c->stream, j->nrows, j->d_rowptr,
                         j->d_colind, j->d_tpos, (const float*)j->d_vals, x, y);
  }
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

int transpose_jacobian(pdes_ctx* c) {
  PDES_CHECK(c->jac && c->jac->assembled, PDES_ERR_ARG, "Jacobian not assembled");
  int st = ensure_tpos(c);
  if (st) return st;
  pdes_jacobian* j = c->jac;
  const dim3 grid((unsigned)ceil_div(std::max<int64_t>(1, j->nnzb), 256));
  if (c->nd == 4) {
    if (j->fp64)
      hipLaunchKernelGGL((bsr_transpose_kernel<4, double>), grid, dim3(256), 0, c->stream, j->nnzb, j->d_tpos,
                         (double*)j->d_vals);
    else
      hipLaunchKernelGGL((bsr_transpose_kernel<4, float>), grid, dim3(256), 0, c->stream, j->nnzb, j->d_tpos,
                         (float*)j->d_vals);
  } else {
    if (j->fp64)
      hipLaunchKernelGGL((bsr_transpose_kernel<5, double>), grid, dim3(256), 0, c->stream, j->nnzb, j->d_tpos,
                         (double*)j->d_vals);
    else
      hipLaunchKernelGGL((bsr_transpose_kernel<5, float>), grid, dim3(256), 0, c->stream, j->nnzb, j->d_tpos,
                         (float*)j->d_vals);
  }
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

int launch_spmv(pdes_ctx* c, const double* x, double* y, double alpha, const double* yadd, double beta) {
  PDES_CHECK(c->jac && c->jac->assembled, PDES_ERR_ARG, "Jacobian not assembled");
  if (c->nd == 4) return launch_spmv_nd<4>(c, x, y, alpha, yadd, beta);
  return launch_spmv_nd<5>(c, x, y, alpha, yadd, beta);
}

int jacobian_setup(pdes_ctx* c, int fp64) {
  if (c->jac && c->jac->fp64 == (fp64 != 0)) return PDES_OK;
  jacobian_free(c);
  pdes_jacobian* j = new pdes_jacobian();
  c->jac = j;
  j->fp64 = fp64 != 0;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> colind;
  build_pattern(c, rowptr, colind);
  j->nrows = (int64_t)c->numEl * c->Np;
  j->nnzb = rowptr.back();
  j->h_rowptr = rowptr;
  j->h_colind = colind;
  const size_t blk = (size_t)c->nd * c->nd * (j->fp64 ? 8 : 4);
  hipError_t e1 = hipMalloc((void**)&j->d_rowptr, rowptr.size() * sizeof(int64_t));
  hipError_t e2 = hipMalloc((void**)&j->d_colind, std::max<size_t>(1, colind.size()) * sizeof(int32_t));
  hipError_t e3 = hipMalloc(&j->d_vals, std::max<size_t>(1, (size_t)j->nnzb) * blk);
  hipError_t e4 = hipMalloc((void**)&j->d_err, sizeof(int32_t));
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
    set_error("device allocation of the BSR matrix (%lld blocks) failed", (long long)j->nnzb);
    jacobian_free(c);
    return PDES_ERR_NOMEM;
  }
  PDES_HIP(hipMemcpyAsync(j->d_rowptr, rowptr.data(), rowptr.size() * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  PDES_HIP(hipMemcpyAsync(j->d_colind, colind.data(), colind.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  PDES_HIP(hipMemsetAsync(j->d_err, 0, sizeof(int32_t), c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

// ---------------------------------------------------------------------------
// calcJacobianSparse (src/jacobian/jacobian.jl:365-446): with a distance-2
// colouring of the elements (no two elements of a colour share a neighbour),
// seeding dof (j, a) of every element of one colour gives, in one derivative
// evaluation of the residual, column (e, j, a) of the Jacobian in the rows of
// e and of its neighbours for every element e of that colour.  The derivative
// is the forward-mode (dual number) residual, the device counterpart of the
// complex step (applyPerturbation / calcJacCol, jacobian.jl:474-681).
// ---------------------------------------------------------------------------
__global__ void seed_kernel(const int32_t* __restrict__ els, int64_t n, int np, int nd, int j, int a, double val,
                            double* __restrict__ v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) v[((int64_t)els[t] * np + j) * nd + a] = val;
}

template <int ND, class BT>
__global__ void coloring_scatter_kernel(const int32_t* __restrict__ els, int64_t n, int np, int nfaces,
                                        const ElFace* __restrict__ elface, const int64_t* __restrict__ rowptr,
                                        const int32_t* __restrict__ colind, BT* __restrict__ vals,
                                        const double* __restrict__ jv, int j, int a, bool add, double scale) {
  // thread -> (element slot, row element r in {e, face neighbours}, row node, row dof)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)(nfaces + 1) * np * ND;
  if (t >= n * per) return;
  const int64_t s = t / per;
  int rem = (int)(t - s * per);
  const int r = rem / (np * ND);
  rem -= r * np * ND;
  const int i = rem / ND, rr = rem - i * ND;
  const int64_t e = els[s];
  const int64_t re = r == 0 ? e : (int64_t)elface[e * nfaces + r - 1].x;
  if (re < 0) return;
  const int64_t row = re * np + i;
  const int64_t pos = bsr_find(rowptr, colind, row, (int32_t)(e * np + j));
  if (pos < 0) return;
  BT* dst = vals + pos * ND * ND + rr * ND + a;
  *dst = add ? (BT)((double)*dst + scale * jv[row * ND + rr]) : (BT)jv[row * ND + rr];
}

template <class BT>
__global__ void bsr_scale_kernel(int64_t n, BT* __restrict__ vals, double s) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) vals[t] = (BT)(s * (double)vals[t]);
}

// -M/tau_vec on the diagonal blocks (the implicit-Euler term of applyEuler)
template <int ND, class BT>
__global__ void ptc_diag_add_kernel(int64_t nrows, const int64_t* __restrict__ rowptr,
                                    const int32_t* __restrict__ colind, BT* __restrict__ vals,
                                    const double* __restrict__ ptc_diag, const double* __restrict__ ptc_scale,
                                    const int32_t* __restrict__ el_inst, int np) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nrows) return;
  const int64_t pos = bsr_find(rowptr, colind, row, (int32_t)row);
  if (pos < 0) return;
  const double d = ptc_diag[row] * ptc_scale[el_inst ? el_inst[row / np] : 0];
#pragma unroll
  for (int a = 0; a < ND; ++a) vals[pos * ND * ND + a * ND + a] = (BT)((double)vals[pos * ND * ND + a * ND + a] - d);
}

static int coloring_setup(pdes_ctx* c) {
  pdes_jacobian* J = c->jac;
  if (J->d_color_el) return PDES_OK;
  const int nel = c->numEl, nf = c->nfaces;
  auto nbr = [&](int e, int f) { return c->h_elface[(size_t)e * nf + f].x; };
  // greedy distance-2 colouring in element order
  std::vector<int32_t> color(nel, -1);
  std::vector<int> mark;
  int ncolor = 0;
  for (int e = 0; e < nel; ++e) {
    mark.assign(ncolor + 1, 0);
    for (int f = 0; f < nf; ++f) {
      const int a = nbr(e, f);
      if (a < 0) continue;
      if (color[a] >= 0) mark[color[a]] = 1;
      for (int g = 0; g < nf; ++g) {
        const int b = nbr(a, g);
        if (b >= 0 && b != e && color[b] >= 0) mark[color[b]] = 1;
      }
    }
    int col = 0;
    while (col < ncolor && mark[col]) ++col;
    color[e] = col;
    ncolor = std::max(ncolor, col + 1);
  }
  J->h_color = color;
  J->color_ptr.assign(ncolor + 1, 0);
  for (int e = 0; e < nel; ++e) J->color_ptr[color[e] + 1]++;
  for (int k = 0; k < ncolor; ++k) J->color_ptr[k + 1] += J->color_ptr[k];
  std::vector<int32_t> els(nel);
  std::vector<int64_t> pos(J->color_ptr.begin(), J->color_ptr.end() - 1);
  for (int e = 0; e < nel; ++e) els[pos[color[e]]++] = e;
  if (hipMalloc((void**)&J->d_color_el, std::max(1, nel) * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void**)&J->d_seed, c->numDof * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&J->d_jv, c->numDof * sizeof(double)) != hipSuccess) {
    set_error("device allocation for the coloring Jacobian failed");
    return PDES_ERR_NOMEM;
  }
  PDES_HIP(hipMemcpyAsync(J->d_color_el, els.data(), nel * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  PDES_HIP(hipMemsetAsync(J->d_seed, 0, c->numDof * sizeof(double), c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

// hom: differentiate the homotopy dissipation instead of the residual and add
// scale * its derivative to the stored values (no implicit-Euler term)
template <int ND, class BT>
static int coloring_run(pdes_ctx* c, bool hom = false, double scale = 1.0) {
  pdes_jacobian* J = c->jac;
  const int np = c->Np, nf = c->nfaces;
  // the residual runs over the whole mesh (the work list is not used here)
  const int32_t nlist_saved = c->nlist;
  c->nlist = -1;
  int st = PDES_OK;
  for (size_t k = 0; k + 1 < J->color_ptr.size() && !st; ++k) {
    const int64_t n = J->color_ptr[k + 1] - J->color_ptr[k];
    const int32_t* els = J->d_color_el + J->color_ptr[k];
    for (int j = 0; j < np && !st; ++j)
      for (int a = 0; a < ND && !st; ++a) {
        hipLaunchKernelGGL(seed_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, c->stream, els, n, np, ND, j,
                           a, 1.0, J->d_seed);
        st = hom ? launch_homotopy(c, c->d_q, J->d_seed, 0.0, 1.0, nullptr, J->d_jv)
                 : launch_residual(c, c->d_q, J->d_seed, nullptr, J->d_jv, false);
        if (st) break;
        const int64_t nt = n * (nf + 1) * np * ND;
        hipLaunchKernelGGL((coloring_scatter_kernel<ND, BT>), dim3((unsigned)ceil_div(nt, 256)), dim3(256), 0,
                           c->stream, els, n, np, nf, c->d_elface, J->d_rowptr, J->d_colind, (BT*)J->d_vals, J->d_jv,
                           j, a, hom, scale);
        hipLaunchKernelGGL(seed_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, c->stream, els, n, np, ND, j,
                           a, 0.0, J->d_seed);
      }
  }
  c->nlist = nlist_saved;
  if (st) return st;
  if (c->ptc_on && !hom) {
    const int64_t nrows = (int64_t)c->numEl * np;
    hipLaunchKernelGGL((ptc_diag_add_kernel<ND, BT>), dim3((unsigned)ceil_div(nrows, 256)), dim3(256), 0, c->stream,
                       nrows, J->d_rowptr, J->d_colind, (BT*)J->d_vals, c->d_ptc_diag, c->d_ptc_scale, c->d_el_inst,
                       np);
  }
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

int assemble_coloring(pdes_ctx* c) {
  PDES_CHECK(c->jac, PDES_ERR_ARG, "call pdes_jacobian_setup first");
  PDES_CHECK(c->nlist < 0, PDES_ERR_UNSUPPORTED, "coloring assembly of a packed batch is not available");
  int st = coloring_setup(c);
  if (st) return st;
  PDES_HIP(hipMemsetAsync(c->jac->d_vals, 0, (size_t)c->jac->nnzb * c->nd * c->nd * (c->jac->fp64 ? 8 : 4),
                          c->stream));
  if (c->nd == 4) st = c->jac->fp64 ? coloring_run<4, double>(c) : coloring_run<4, float>(c);
  else st = c->jac->fp64 ? coloring_run<5, double>(c) : coloring_run<5, float>(c);
  if (st) return st;
  c->jac->assembled = true;
  return PDES_OK;
}

// the Jacobian of predictorCorrectorHomotopy's system (1 - lambda) R + lambda D
// (HomotopyMatLO calcLinearOperator, src/NonlinearSolvers/predictor_corrector.jl:
// the physics Jacobian scaled by 1 - lambda, then lambda dD/dq added), with
// dD/dq by colouring over the dual-number dissipation
int assemble_homotopy(pdes_ctx* c, double lambda) {
  PDES_CHECK(c->jac, PDES_ERR_ARG, "call pdes_jacobian_setup first");
  PDES_CHECK(lambda >= 0.0 && lambda <= 1.0, PDES_ERR_ARG, "homotopy parameter outside [0, 1]");
  int st = launch_assemble(c);
  if (st) return st;
  pdes_jacobian* J = c->jac;
  const int64_t nv = J->nnzb * c->nd * c->nd;
  if (lambda != 0.0) {
    if (J->fp64)
      hipLaunchKernelGGL(bsr_scale_kernel<double>, dim3((unsigned)ceil_div(nv, 256)), dim3(256), 0, c->stream, nv,
                         (double*)J->d_vals, 1.0 - lambda);
    else
      hipLaunchKernelGGL(bsr_scale_kernel<float>, dim3((unsigned)ceil_div(nv, 256)), dim3(256), 0, c->stream, nv,
                         (float*)J->d_vals, 1.0 - lambda);
    PDES_HIP(hipGetLastError());
    if ((st = coloring_setup(c))) return st;
    if (c->nd == 4) st = J->fp64 ? coloring_run<4, double>(c, true, lambda) : coloring_run<4, float>(c, true, lambda);
    else st = J->fp64 ? coloring_run<5, double>(c, true, lambda) : coloring_run<5, float>(c, true, lambda);
    if (st) return st;
  }
  J->assembled = true;
  return PDES_OK;
}

int assemble(pdes_ctx* c) {
  PDES_CHECK(c->jac, PDES_ERR_ARG, "call pdes_jacobian_setup first");
  int st = launch_assemble(c);
  if (st) return st;
  c->jac->assembled = true;
  return PDES_OK;
}

}  // namespace pdes

using namespace pdes;

extern "C" {

int pdes_jacobian_setup(pdes_ctx* c, int fp64_blocks) {
  PDES_CHECK(c, PDES_ERR_ARG, "null context");
  return jacobian_setup(c, fp64_blocks);
}

int pdes_assemble_jacobian_coloring(pdes_ctx* c) {
  PDES_CHECK(c, PDES_ERR_ARG, "null context");
  int st = assemble_coloring(c);
  if (st) return st;
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

int pdes_get_coloring(pdes_ctx* c, int32_t* color, int32_t* ncolors, int32_t* pert_neighbor_els) {
  PDES_CHECK(c && c->jac, PDES_ERR_ARG, "call pdes_jacobian_setup first");
  int st = coloring_setup(c);
  if (st) return st;
  const pdes_jacobian* J = c->jac;
  const int nel = c->numEl, nf = c->nfaces, nc = (int)J->color_ptr.size() - 1;
  if (ncolors) *ncolors = nc;
  if (color) std::copy(J->h_color.begin(), J->h_color.end(), color);
  if (pert_neighbor_els) {
    // pertNeighborEls[i, c]: the element of colour c among i and its face
    // neighbours (at most one, the colouring being distance-2), or -1
    std::fill(pert_neighbor_els, pert_neighbor_els + (size_t)nel * nc, -1);
    for (int e = 0; e < nel; ++e) {
      pert_neighbor_els[(size_t)e * nc + J->h_color[e]] = e;
      for (int f = 0; f < nf; ++f) {
        const int a = c->h_elface[(size_t)e * nf + f].x;
        if (a >= 0) pert_neighbor_els[(size_t)e * nc + J->h_color[a]] = a;
      }
    }
  }
  return PDES_OK;
}

int pdes_assemble_homotopy_jacobian(pdes_ctx* c, double lambda) {
  PDES_CHECK(c, PDES_ERR_ARG, "null context");
  int st = assemble_homotopy(c, lambda);
  if (st) return st;
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

int pdes_jacobian_nnzb(pdes_ctx* c, int64_t* nrows_blk, int64_t* nnzb) {
  PDES_CHECK(c && c->jac, PDES_ERR_ARG, "call pdes_jacobian_setup first");
  if (nrows_blk) *nrows_blk = c->jac->nrows;
  if (nnzb) *nnzb = c->jac->nnzb;
  return PDES_OK;
}

int pdes_get_bsr_pattern(pdes_ctx* c, int64_t* rowptr, int32_t* colind) {
  PDES_CHECK(c && c->jac && rowptr && colind, PDES_ERR_ARG, "bad argument");
  std::copy(c->jac->h_rowptr.begin(), c->jac->h_rowptr.end(), rowptr);
  std::copy(c->jac->h_colind.begin(), c->jac->h_colind.end(), colind);
  return PDES_OK;
}

int pdes_assemble_jacobian(pdes_ctx* c) {
  PDES_CHECK(c && c->jac, PDES_ERR_ARG, "call pdes_jacobian_setup first");
  int st = assemble(c);
  if (st) return st;
  int32_t err = 0;
  PDES_HIP(hipMemcpyAsync(&err, c->jac->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  PDES_CHECK(err == 0, PDES_ERR_ARG, "a Jacobian block fell outside the sparsity pattern");
  return PDES_OK;
}

int pdes_get_bsr_values(pdes_ctx* c, void* values) {
  PDES_CHECK(c && c->jac && c->jac->assembled && values, PDES_ERR_ARG, "Jacobian not assembled");
  const size_t bytes = (size_t)c->jac->nnzb * c->nd * c->nd * (c->jac->fp64 ? 8 : 4);
  PDES_HIP(hipMemcpyAsync(values, c->jac->d_vals, bytes, hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

int pdes_bsr_matvec_transpose(pdes_ctx* c, const double* x, double* b) {
  PDES_CHECK(c && x && b, PDES_ERR_ARG, "null argument");
  PDES_HIP(hipMemcpyAsync(c->d_work, x, c->numDof * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int st = launch_spmv_transpose(c, c->d_work, c->d_work2);
  if (st) return st;
  PDES_HIP(hipMemcpyAsync(b, c->d_work2, c->numDof * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

int pdes_bsr_matvec(pdes_ctx* c, const double* x, double* b) {
  PDES_CHECK(c && x && b, PDES_ERR_ARG, "null argument");
  PDES_HIP(hipMemcpyAsync(c->d_work, x, c->numDof * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int st = launch_spmv(c, c->d_work, c->d_work2, 1.0, nullptr, 0.0);
  if (st) return st;
  PDES_HIP(hipMemcpyAsync(b, c->d_work2, c->numDof * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

}  // extern "C"
