// solver.hip -- preconditioners, GMRES and Newton on the device (C ABI).
//
// NewtonBDiagPC (src/NonlinearSolvers/preconditioning_types.jl:41-110,
// preconditioning.jl:7-215): each element's diagonal (Np*ndof)^2 block of the
// Jacobian is inverted (the reference factors it with getrf and applies getrs;
// an explicit inverse makes the apply one coalesced product per element).
// Setup and apply run in FP32.
// GMRES(restart) with right preconditioning replaces the PETSc KSP of
// src/linearsolvers/ls_standard.jl:321-370 (KSP options read_input.jl:560-581);
// the Arnoldi basis, Hessenberg and update stay FP64.  Orthogonalisation is
// classical Gram-Schmidt applied twice (CGS2): two multi-dot and two
// multi-axpy passes per step, each reading the basis once.
// Newton follows newtonInner (src/NonlinearSolvers/newton.jl:137-304).
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "jacobian.hpp"

struct pdes_precond {
  int kind = 0;
  int n = 0;                 // block size Np*nd
  float* d_inv = nullptr;    // [numEl][n][n] column-major inverse of each (modified) element block
  bool factored = false;
  // ILU(0): elements grouped by dependency level of the forward (lower
  // neighbours first) and backward (upper neighbours first) sweeps
  bool ilu_ready = false;
  std::vector<int64_t> fptr, bptr;
  int32_t* d_fel = nullptr;
  int32_t* d_bel = nullptr;
  int32_t* d_own = nullptr;   // per node row: offset of the element's own block run within the row
  // dependency-driven sweeps: the forward sweep's result (a vector of the
  // unknowns that holds ILU_PENDING between applications) and the ticket /
  // error / poll-limit words
  double* d_y = nullptr;
  int32_t* d_sync = nullptr;
  // sweep choice: 0 by the cost model, 1 always level by level
  // (PDES_ILU_LEVELS set), 2 always dependency-driven (PDES_ILU_DEP set)
  int force = 0;
  std::vector<int32_t> h_fl, h_bl;   // forward / backward level of each element
  // the dependency-driven sweeps' level-ordered lists, every level padded
  // with -1 to whole chunks of `chunk` elements, and their chunk counts
  int chunk = 1;
  int32_t* d_fpad = nullptr;
  int32_t* d_bpad = nullptr;
  int64_t fpad_n = 0, bpad_n = 0;
  // the chunks holding elements of selected instances (packed batches), their
  // counts, the selection flags and the select's work space
  int32_t* d_fsel = nullptr;
  int32_t* d_bsel = nullptr;
  int64_t sel_nf = 0, sel_nb = 0;
  int32_t* d_inst_sel = nullptr;
  int32_t* d_nsel = nullptr;
  void* d_select_tmp = nullptr;
  size_t select_tmp_bytes = 0;
  int sel_ninst = 0;
  double blk_per_el = 0;             // neighbour blocks of an element on one side (mean)
  // the 2 x levels sweep launches of one apply, captured once per (mask,
  // matrix) into a HIP graph over fixed staging vectors and replayed
  double* d_stage_in = nullptr;
  double* d_stage_out = nullptr;
  struct Graph {
    const int32_t* mask;
    const void* vals;
    bool fp64;
    const float* inv;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
};

namespace pdes {

struct GmresWork;
void gmres_release(pdes_ctx* c);

int precond_free(pdes_ctx* c) {
  gmres_release(c);
  if (!c->pc_obj) return PDES_OK;
  if (c->pc_obj->d_inv) hipFree(c->pc_obj->d_inv);
  if (c->pc_obj->d_fel) hipFree(c->pc_obj->d_fel);
  if (c->pc_obj->d_bel) hipFree(c->pc_obj->d_bel);
  if (c->pc_obj->d_own) hipFree(c->pc_obj->d_own);
  if (c->pc_obj->d_y) hipFree(c->pc_obj->d_y);
  if (c->pc_obj->d_fpad) hipFree(c->pc_obj->d_fpad);
  if (c->pc_obj->d_bpad) hipFree(c->pc_obj->d_bpad);
  if (c->pc_obj->d_fsel) hipFree(c->pc_obj->d_fsel);
  if (c->pc_obj->d_bsel) hipFree(c->pc_obj->d_bsel);
  if (c->pc_obj->d_inst_sel) hipFree(c->pc_obj->d_inst_sel);
  if (c->pc_obj->d_nsel) hipFree(c->pc_obj->d_nsel);
  if (c->pc_obj->d_select_tmp) hipFree(c->pc_obj->d_select_tmp);
  if (c->pc_obj->d_sync) hipFree(c->pc_obj->d_sync);
  if (c->pc_obj->d_stage_in) hipFree(c->pc_obj->d_stage_in);
  if (c->pc_obj->d_stage_out) hipFree(c->pc_obj->d_stage_out);
  for (auto& g : c->pc_obj->graphs) hipGraphExecDestroy(g.exec);
  delete c->pc_obj;
  c->pc_obj = nullptr;
  return PDES_OK;
}

// ---------------------------------------------------------------------------
// block-Jacobi setup: gather the element block from the BSR matrix into LDS
// (FP32) and invert it in place by Gauss-Jordan elimination with partial
// (row) pivoting; the inverse is stored column-major so that the apply is a
// coalesced matrix-vector product.
// ---------------------------------------------------------------------------
template <int N>
__device__ void gauss_jordan_inverse(float (*A)[N + 1], int* perm, float* fcol, int* pivrow, int32_t* err) {
  const int tid = threadIdx.x;
  for (int k = 0; k < N; ++k) {
    // pivot: max |A[r][k]| over r >= k, lowest row on ties (first wave,
    // butterfly reduction of (value, row) pairs)
    if (tid < 64) {
      float best = -1.0f;
      int bi = N;
      for (int r = k + tid; r < N; r += 64) {
        const float v = fabsf(A[r][k]);
        if (v > best) { best = v; bi = r; }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const float ov = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
      if (tid == 0) {
        *pivrow = bi;
        perm[k] = bi;
        if (!(best > 0.0f)) err[0] = 1;
      }
    }
    __syncthreads();
    const int p = *pivrow;
    if (p != k)
      for (int cc = tid; cc < N; cc += blockDim.x) {
        const float t0 = A[k][cc];
        A[k][cc] = A[p][cc];
        A[p][cc] = t0;
      }
    __syncthreads();
    const float pinv = 1.0f / A[k][k];
    __syncthreads();
    // column k of the identity takes the pivot's place, then row k is scaled
    if (tid == 0) A[k][k] = 1.0f;
    for (int r = tid; r < N; r += blockDim.x) fcol[r] = (r == k) ? 0.0f : A[r][k];
    __syncthreads();
    for (int cc = tid; cc < N; cc += blockDim.x) A[k][cc] *= pinv;
    for (int r = tid; r < N; r += blockDim.x)
      if (r != k) A[r][k] = 0.0f;
    __syncthreads();
    for (int t = tid; t < N * N; t += blockDim.x) {
      const int r = t / N, cc = t - r * N;
      if (r != k) A[r][cc] -= fcol[r] * A[k][cc];
    }
    __syncthreads();
  }
  // undo the row interchanges as column interchanges, last first
  for (int k = N - 1; k >= 0; --k) {
    const int p = perm[k];
    if (p != k)
      for (int r = tid; r < N; r += blockDim.x) {
        const float t0 = A[r][k];
        A[r][k] = A[r][p];
        A[r][p] = t0;
      }
    __syncthreads();
  }
}

template <int NP, int ND, class BT>
__global__ void __launch_bounds__(256)
bjacobi_factor_kernel(const int32_t* __restrict__ el_list, const int64_t* __restrict__ rowptr,
                      const int32_t* __restrict__ colind, const BT* __restrict__ vals, float* __restrict__ inv,
                      int32_t* __restrict__ err) {
  constexpr int N = NP * ND;
  __shared__ float A[N][N + 1];
  __shared__ float fcol[N];
  __shared__ int perm[N];
  __shared__ int pivrow;
  const int tid = threadIdx.x;
  const int64_t e = el_list ? (int64_t)el_list[blockIdx.x] : (int64_t)blockIdx.x;
  for (int t = tid; t < NP * NP; t += blockDim.x) {
    const int i = t / NP, j = t - i * NP;
    const int64_t row = e * NP + i;
    const int64_t pos = bsr_find(rowptr, colind, row, (int32_t)(e * NP + j));
    for (int a = 0; a < ND; ++a)
      for (int b = 0; b < ND; ++b)
        A[i * ND + a][j * ND + b] = pos >= 0 ? (float)vals[pos * ND * ND + a * ND + b] : 0.0f;
  }
  __syncthreads();
  gauss_jordan_inverse<N>(A, perm, fcol, &pivrow, err);
  float* out = inv + e * N * N;
  for (int t = tid; t < N * N; t += blockDim.x) {
    const int cc = t / N, r = t - cc * N;
    out[t] = A[r][cc];
  }
}

// block-Jacobi apply: x_e = inv_e b_e; EPB = 256 / N elements per block, thread
// (element slot, row r) forms one row of the product from the column-major
// inverse (consecutive rows read consecutive floats).
template <int NP, int ND>
__global__ void __launch_bounds__(256)
bjacobi_apply_kernel(const int32_t* __restrict__ el_list, int64_t nel, const float* __restrict__ inv,
                     const double* __restrict__ b, double* __restrict__ x, const int32_t* __restrict__ el_inst,
                     const int32_t* __restrict__ mask) {
  constexpr int N = NP * ND;
  constexpr int EPB = 256 / N;
  __shared__ float v[EPB][N];
  const int slot = threadIdx.x / N;
  const int r = threadIdx.x - slot * N;
  const int64_t s = (int64_t)blockIdx.x * EPB + slot;
  bool act = slot < EPB && s < nel;
  int64_t e = 0;
  if (act) {
    e = el_list ? (int64_t)el_list[s] : s;
    if (mask && !mask[el_inst[e]]) act = false;
  }
  if (act) v[slot][r] = (float)b[e * N + r];
  __syncthreads();
  if (!act) return;
  const float* M = inv + e * N * N;
  float acc = 0.0f;
#pragma unroll
  for (int cc = 0; cc < N; ++cc) acc += M[cc * N + r] * v[slot][cc];
  x[e * N + r] = (double)acc;
}

// ---------------------------------------------------------------------------
// Element-block ILU(0) in the natural element order.  With E the element
// blocks of the Jacobian, L/U its strictly lower/upper element blocks and
//   D'_e = A_ee - sum_{k in lower(e)} A_ek D'_k^-1 A_ke,
// the preconditioner is M = (D' + L) D'^-1 (D' + U); no fill outside the
// element-block pattern is kept (ILU(0) on element blocks).  A_ek couples only
// the face-stencil nodes of e and of its neighbour k, so the update of D'_e
// touches only the (stencil x stencil) part of the block:
//   C = S_ek [D'_k^-1]_(Rk,Rk) S_ke,  Rk / Ce = stencil nodes of the face.
// Elements are processed level by level (a level holds elements whose lower
// neighbours are all in earlier levels), each level in one launch.
// ---------------------------------------------------------------------------
template <int NP, int ND, int NS, class BT>
__global__ void __launch_bounds__(256)
ilu_factor_level_kernel(const int32_t* __restrict__ lev_el, const int64_t* __restrict__ rowptr,
                        const int32_t* __restrict__ colind, const BT* __restrict__ vals,
                        const pdes::ElFace* __restrict__ elface, int nfaces, const int32_t* __restrict__ perm,
                        float* __restrict__ inv, int32_t* __restrict__ err, const int32_t* __restrict__ el_inst,
                        const int32_t* __restrict__ inst_active) {
  constexpr int N = NP * ND;
  constexpr int M = NS * ND;
  __shared__ float A[N][N + 1];
  __shared__ float B1[M][M + 1];
  __shared__ float B2[M][M + 1];
  __shared__ float B3[M][M + 1];
  __shared__ float fcol[N];
  __shared__ int permk[N];
  __shared__ int pivrow;
  __shared__ int Ce[NS], Rk[NS];
  const int tid = threadIdx.x;
  const int64_t e = lev_el[blockIdx.x];
  if (inst_active && !inst_active[el_inst[e]]) return;
  for (int t = tid; t < NP * NP; t += blockDim.x) {
    const int i = t / NP, j = t - i * NP;
    const int64_t pos = bsr_find(rowptr, colind, e * NP + i, (int32_t)(e * NP + j));
    for (int a = 0; a < ND; ++a)
      for (int b = 0; b < ND; ++b)
        A[i * ND + a][j * ND + b] = pos >= 0 ? (float)vals[pos * ND * ND + a * ND + b] : 0.0f;
  }
  __syncthreads();
  for (int f = 0; f < nfaces; ++f) {
    const pdes::ElFace r = elface[e * nfaces + f];
    if (r.x < 0 || r.x >= e) continue;   // boundary face or upper neighbour (uniform over the block)
    const int64_t k = r.x;
    const int fk = r.z & 0xff;
    if (tid < NS) {
      Ce[tid] = perm[tid * nfaces + f];
      Rk[tid] = perm[tid * nfaces + fk];
    }
    __syncthreads();
    // B1 = [D'_k^-1](Rk,Rk), B2 = S_ke = A(k rows Rk, e cols Ce)
    const float* Ik = inv + k * N * N;
    for (int t = tid; t < M * M; t += blockDim.x) {
      const int r1 = t / M, c1 = t - r1 * M;
      const int s1 = r1 / ND, a = r1 - s1 * ND, s2 = c1 / ND, b = c1 - s2 * ND;
      B1[r1][c1] = Ik[(Rk[s2] * ND + b) * N + Rk[s1] * ND + a];
      const int64_t pos = bsr_find(rowptr, colind, k * NP + Rk[s1], (int32_t)(e * NP + Ce[s2]));
      B2[r1][c1] = pos >= 0 ? (float)vals[pos * ND * ND + a * ND + b] : 0.0f;
    }
    __syncthreads();
    for (int t = tid; t < M * M; t += blockDim.x) {
      const int r1 = t / M, c1 = t - r1 * M;
      float acc = 0.0f;
      for (int l = 0; l < M; ++l) acc += B1[r1][l] * B2[l][c1];
      B3[r1][c1] = acc;
    }
    __syncthreads();
    // B1 = S_ek = A(e rows Ce, k cols Rk)
    for (int t = tid; t < M * M; t += blockDim.x) {
      const int r1 = t / M, c1 = t - r1 * M;
      const int s1 = r1 / ND, a = r1 - s1 * ND, s2 = c1 / ND, b = c1 - s2 * ND;
      const int64_t pos = bsr_find(rowptr, colind, e * NP + Ce[s1], (int32_t)(k * NP + Rk[s2]));
      B1[r1][c1] = pos >= 0 ? (float)vals[pos * ND * ND + a * ND + b] : 0.0f;
    }
    __syncthreads();
    for (int t = tid; t < M * M; t += blockDim.x) {
      const int r1 = t / M, c1 = t - r1 * M;
      float acc = 0.0f;
      for (int l = 0; l < M; ++l) acc += B1[r1][l] * B3[l][c1];
      const int s1 = r1 / ND, a = r1 - s1 * ND, s2 = c1 / ND, b = c1 - s2 * ND;
      A[Ce[s1] * ND + a][Ce[s2] * ND + b] -= acc;
    }
    __syncthreads();
  }
  gauss_jordan_inverse<N>(A, permk, fcol, &pivrow, err);
  float* out = inv + e * N * N;
  for (int t = tid; t < N * N; t += blockDim.x) {
    const int cc = t / N, r = t - cc * N;
    out[t] = A[r][cc];
  }
}

// One level of a sweep.  Forward: x_e = D'_e^-1 (b_e - sum_{k<e} A_ek x_k).
// Backward (in place on x): x_e -= D'_e^-1 sum_{k>e} A_ek x_k.  The neighbour
// products walk the element's BSR rows and keep the blocks whose column
// element is below (forward) or above (backward) e.
template <int NP, int ND, class BT>
__global__ void __launch_bounds__(256)
ilu_sweep_level_kernel(const int32_t* __restrict__ lev_el, int64_t count, const int64_t* __restrict__ rowptr,
                       const int32_t* __restrict__ colind, const int32_t* __restrict__ own,
                       const BT* __restrict__ vals, const float* __restrict__ inv,
                       const double* __restrict__ b, double* __restrict__ x, int forward,
                       const int32_t* __restrict__ el_inst, const int32_t* __restrict__ mask,
                       const int32_t* __restrict__ inst_active) {
  constexpr int N = NP * ND;
  constexpr int EPB = 256 / N;
  // entries of this row of D'_e^-1 loaded before the neighbour products, so
  // their latency overlaps with that dependent chain (all of them up to 72
  // per row, the first 32 beyond)
  constexpr int PF = N <= 72 ? N : 32;
  __shared__ float v[EPB][N];
  const int slot = threadIdx.x / N;
  const int r = threadIdx.x - slot * N;
  const int64_t s = (int64_t)blockIdx.x * EPB + slot;
  bool act = slot < EPB && s < count;
  int64_t e = 0;
  if (act) {
    e = lev_el[s];
    const int inst = el_inst ? el_inst[e] : 0;
    if ((mask && !mask[inst]) || (inst_active && !inst_active[inst])) act = false;
  }
  float mrow[PF];
  if (act) {
    const float* Mi = inv + e * N * N;
#pragma unroll
    for (int cc = 0; cc < PF; ++cc) mrow[cc] = Mi[cc * N + r];
  }
  if (act) {
    const int i = r / ND, a = r - i * ND;
    const int64_t row = e * NP + i;
    double acc = forward ? b[e * N + r] : 0.0;
    // columns are sorted: lower-element blocks precede the element's own run
    // of NP blocks (starting at own[row] within the row), upper-element
    // blocks follow it; fixed trip counts so the loads pipeline
    const int64_t p0 = rowptr[row], p1 = rowptr[row + 1];
    const int64_t q0 = forward ? p0 : p0 + own[row] + NP;
    const int64_t q1 = forward ? p0 + own[row] : p1;
    double t = 0.0;
#pragma unroll 4
    for (int64_t p = q0; p < q1; ++p) {
      const int32_t col = colind[p];
      const BT* blk = vals + p * ND * ND + a * ND;
      const double* xc = x + (int64_t)col * ND;
#pragma unroll
      for (int c = 0; c < ND; ++c) t += (double)blk[c] * xc[c];
    }
    acc = forward ? acc - t : acc + t;
    v[slot][r] = (float)acc;
  }
  __syncthreads();
  if (!act) return;
  const float* Mi = inv + e * N * N;
  float y = 0.0f;
#pragma unroll
  for (int cc = 0; cc < PF; ++cc) y += mrow[cc] * v[slot][cc];
#pragma unroll
  for (int cc = PF; cc < N; ++cc) y += Mi[cc * N + r] * v[slot][cc];
  if (forward) x[e * N + r] = (double)y;
  else x[e * N + r] -= (double)y;
}

// The same level with LPR lanes per row, for dense face stencils (a row has
// tens of neighbour blocks): lane l of row r takes the blocks q0 + l,
// q0 + l + LPR, ... and the columns cc = l, l + LPR, ... of D'_e^-1; the
// partial sums meet in LDS and are added in lane order, so the bits do not
// depend on the level's width or on the other instances of a batch.
template <int NP, int ND, class BT, int LPR>
__global__ void __launch_bounds__(256)
ilu_sweep_level_ml_kernel(const int32_t* __restrict__ lev_el, int64_t count, const int64_t* __restrict__ rowptr,
                          const int32_t* __restrict__ colind, const int32_t* __restrict__ own,
                          const BT* __restrict__ vals, const float* __restrict__ inv,
                          const double* __restrict__ b, double* __restrict__ x, int forward,
                          const int32_t* __restrict__ el_inst, const int32_t* __restrict__ mask,
                          const int32_t* __restrict__ inst_active) {
  constexpr int N = NP * ND;
  constexpr int TPE = N * LPR;
  static_assert(TPE <= 256, "one element's rows must fit a workgroup");
  constexpr int EPB = 256 / TPE;
  __shared__ double tp[EPB][N][LPR];
  __shared__ float yp[EPB][N][LPR];
  __shared__ float v[EPB][N];
  const int slot = threadIdx.x / TPE;
  const int t = threadIdx.x - slot * TPE;
  const int r = t / LPR, l = t - r * LPR;
  const int64_t s = (int64_t)blockIdx.x * EPB + slot;
  bool act = slot < EPB && s < count;
  int64_t e = 0;
  if (act) {
    e = lev_el[s];
    const int inst = el_inst ? el_inst[e] : 0;
    if ((mask && !mask[inst]) || (inst_active && !inst_active[inst])) act = false;
  }
  // this lane's entries of D'_e^-1 (columns l, l + LPR, ...) and the
  // right-hand side are loaded before the neighbour products, so their
  // latency overlaps with that dependent chain
  constexpr int NC = (N + LPR - 1) / LPR;
  float mrow[NC];
  double bval = 0.0;
  if (act) {
    const float* Mi = inv + e * N * N;
#pragma unroll
    for (int k = 0; k < NC; ++k)
      if (l + k * LPR < N) mrow[k] = Mi[(l + k * LPR) * N + r];
    if (forward && l == 0) bval = b[e * N + r];
  }
  if (act) {
    const int i = r / ND, a = r - i * ND;
    const int64_t row = e * NP + i;
    const int64_t p0 = rowptr[row], p1 = rowptr[row + 1];
    const int64_t q0 = forward ? p0 : p0 + own[row] + NP;
    const int64_t q1 = forward ? p0 + own[row] : p1;
    double acc = 0.0;
#pragma unroll 2
    for (int64_t p = q0 + l; p < q1; p += LPR) {
      const int32_t col = colind[p];
      const BT* blk = vals + p * ND * ND + a * ND;
      const double* xc = x + (int64_t)col * ND;
#pragma unroll
      for (int c = 0; c < ND; ++c) acc += (double)blk[c] * xc[c];
    }
    tp[slot][r][l] = acc;
  }
  __syncthreads();
  if (act && l == 0) {
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < LPR; ++k) sum += tp[slot][r][k];
    v[slot][r] = (float)(forward ? bval - sum : sum);
  }
  __syncthreads();
  if (act) {
    float y = 0.0f;
#pragma unroll
    for (int k = 0; k < NC; ++k)
      if (l + k * LPR < N) y += mrow[k] * v[slot][l + k * LPR];
    yp[slot][r][l] = y;
  }
  __syncthreads();
  if (act && l == 0) {
    float y = 0.0f;
#pragma unroll
    for (int k = 0; k < LPR; ++k) y += yp[slot][r][k];
    if (forward) x[e * N + r] = (double)y;
    else x[e * N + r] -= (double)y;
  }
}

// ---------------------------------------------------------------------------
// Dependency-driven sweeps: one launch per triangle instead of one per level.
// The level-ordered element list is cut into chunks of ILU_CHUNK elements of
// one level each (every level padded to whole chunks with -1), so the
// elements of a chunk never depend on each other.  Each workgroup takes one
// chunk, in ticket order: an atomic counter hands out chunk positions, so
// every element that a chunk's elements depend on (their lower neighbours in
// the forward sweep, their upper ones in the backward sweep) sits in a chunk
// with a smaller ticket, which already runs on the device or is done.  A
// sweep's output holds ILU_PENDING (a signalling-NaN bit pattern, which no
// arithmetic produces: a NaN result is quiet) wherever it is not yet
// computed, so a workgroup waits for its neighbours by loading their values
// until none is pending, then computes x_e as the level kernels do (same
// lanes, same order of operations: the same bits) and stores it.  Values are
// loaded and stored with device-scope atomics (an 8-byte store is seen whole
// or not at all), so no stale copy in another XCD's L2 is ever read.  The
// forward sweep writes y = L^-1 b into d_y and sets its elements' rows of x
// pending (after reading b: x may be b); the backward sweep reads its own y,
// writes x and sets its rows of y pending again, so d_y is pending between
// applications (each factorization sets it so).  A wait gives up after a bounded number of
// polls, sets the error word and carries on, so every launch ends.
// ---------------------------------------------------------------------------
constexpr int ILU_WAIT_POLLS = 1 << 18;
constexpr unsigned long long ILU_PENDING = 0x7FF4DEADBEEF0001ull;
// sync words: ticket counter, error word, poll limit (0: ILU_WAIT_POLLS)
enum { ILU_SYNC_TICKET = 0, ILU_SYNC_ERR = 1, ILU_SYNC_POLLS = 2, ILU_SYNC_WORDS = 3 };

// elements per chunk for a row block of n unknowns: as many element row
// blocks as fill a workgroup of at most 256 threads, at most 4
__host__ __device__ constexpr int ilu_dep_lanes(int n, bool diage) {
  return diage ? 1 : (256 / n >= 8 ? 8 : 256 / n);
}
__host__ __device__ constexpr int ilu_dep_chunk(int n, bool diage) {
  return 256 / (n * ilu_dep_lanes(n, diage)) >= 4 ? 4 : 256 / (n * ilu_dep_lanes(n, diage));
}

__global__ void ilu_sweep_begin(int32_t* __restrict__ sync) { sync[ILU_SYNC_TICKET] = 0; }

__global__ void ilu_pending_fill(double* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = __longlong_as_double((long long)ILU_PENDING);
}

__device__ __forceinline__ bool ilu_is_pending(double v) {
  return (unsigned long long)__double_as_longlong(v) == ILU_PENDING;
}

__device__ __forceinline__ double x_load(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void x_store(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// in: the right-hand side (forward: b; backward: y); out: the result, which
// the neighbours' values are read from (forward: y; backward: x); reset: the
// vector whose rows of this element become pending after use (forward: x;
// backward: y).  in may alias reset.
template <int NP, int ND, int NS, bool DIAGE, class BT, int LPR, int CH>
__global__ void __launch_bounds__(256)
ilu_sweep_dep_kernel(const int32_t* __restrict__ order, const int32_t* __restrict__ chunks, int64_t nchunks,
                     const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
                     const int32_t* __restrict__ own, const BT* __restrict__ vals, const float* __restrict__ inv,
                     const double* in, double* out, double* reset, int forward,
                     const int32_t* __restrict__ el_inst, const int32_t* __restrict__ mask,
                     const int32_t* __restrict__ inst_active, int32_t* __restrict__ sync) {
  constexpr int N = NP * ND;
  constexpr int TPE = N * LPR;
  static_assert(CH * TPE <= 256, "a chunk's rows must fit a workgroup");
  // neighbour blocks of one row on one side: at most one per face with
  // diagonal-E faces, the face stencil per face otherwise; this lane's share
  // is loaded before the neighbours' values are awaited
  constexpr int NFACES = (ND == 4) ? 3 : 4;
  constexpr int MAXB = DIAGE ? NFACES : (NFACES * NS + LPR - 1) / LPR;
  __shared__ double tp[CH][N][LPR];
  __shared__ float yp[CH][N][LPR];
  __shared__ float v[CH][N];
  __shared__ int64_t sh_e[CH];
  __shared__ int sh_any;
  const int t = threadIdx.x;
  const int slot = t / TPE;
  const int tt = t - slot * TPE;
  const int r = tt / LPR, l = tt - r * LPR;
  const int32_t max_polls = sync[ILU_SYNC_POLLS] > 0 ? sync[ILU_SYNC_POLLS] : ILU_WAIT_POLLS;
  if (t == 0) {
    const int64_t tk = __hip_atomic_fetch_add(&sync[ILU_SYNC_TICKET], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int64_t ch = tk < nchunks ? (chunks ? (int64_t)chunks[tk] : tk) : -1;
    int any = 0;
    for (int sl = 0; sl < CH; ++sl) {
      int64_t e = ch >= 0 ? (int64_t)order[ch * CH + sl] : -1;
      if (e >= 0) {
        const int inst = el_inst ? el_inst[e] : 0;
        if ((mask && !mask[inst]) || (inst_active && !inst_active[inst])) e = -1;   // no element waits on it
      }
      sh_e[sl] = e;
      any |= e >= 0;
    }
    sh_any = any;
  }
  __syncthreads();
  if (!sh_any) return;   // uniform over the block
  const int64_t e = slot < CH ? sh_e[slot] : -1;
  const bool act = e >= 0;
  // everything that does not depend on the neighbours' values: this lane's
  // entries of D'_e^-1 (columns l, l + LPR, ...; the first 32 when there are
  // more than 72), the right-hand side, and this lane's neighbour blocks and
  // their columns
  constexpr int NC = (N + LPR - 1) / LPR;
  constexpr int PF = NC <= 72 ? NC : 32;
  float mrow[PF];
  BT pre[MAXB][ND];
  int32_t pcol[MAXB];
  double bval = 0.0;
  const float* Mi = inv + (act ? e : 0) * N * N;
  int64_t q0 = 0, q1 = 0;
  if (act) {
#pragma unroll
    for (int k = 0; k < PF; ++k)
      if (l + k * LPR < N) mrow[k] = Mi[(l + k * LPR) * N + r];
    if (l == 0) bval = in[e * N + r];
    const int i = r / ND, a = r - i * ND;
    const int64_t row = e * NP + i;
    const int64_t p0 = rowptr[row], p1 = rowptr[row + 1];
    q0 = forward ? p0 : p0 + own[row] + NP;
    q1 = forward ? p0 + own[row] : p1;
#pragma unroll
    for (int u = 0; u < MAXB; ++u) {
      const int64_t p = q0 + l + (int64_t)u * LPR;
      pcol[u] = p < q1 ? colind[p] : 0;
#pragma unroll
      for (int c = 0; c < ND; ++c) pre[u][c] = p < q1 ? vals[p * ND * ND + a * ND + c] : BT(0);
    }
  }
  double acc = 0.0;
  if (act) {
    const int a = r % ND;
    // the neighbours' values: all loads issued at once, then the pending
    // ones polled again until every one has arrived
    double xv[MAXB][ND];
#pragma unroll
    for (int u = 0; u < MAXB; ++u)
#pragma unroll
      for (int c = 0; c < ND; ++c)
        xv[u][c] = q0 + l + (int64_t)u * LPR < q1 ? x_load(out + (int64_t)pcol[u] * ND + c) : 0.0;
    int polls = 0;
    while (true) {
      bool pending = false;
#pragma unroll
      for (int u = 0; u < MAXB; ++u)
#pragma unroll
        for (int c = 0; c < ND; ++c) pending |= ilu_is_pending(xv[u][c]);
      if (!pending) break;
      if (++polls > max_polls) {
        __hip_atomic_store(&sync[ILU_SYNC_ERR], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
#pragma unroll
      for (int u = 0; u < MAXB; ++u)
#pragma unroll
        for (int c = 0; c < ND; ++c)
          if (ilu_is_pending(xv[u][c])) xv[u][c] = x_load(out + (int64_t)pcol[u] * ND + c);
    }
#pragma unroll
    for (int u = 0; u < MAXB; ++u) {
      if (q0 + l + (int64_t)u * LPR < q1) {
#pragma unroll
        for (int c = 0; c < ND; ++c) acc += (double)pre[u][c] * xv[u][c];
      }
    }
    for (int64_t p = q0 + l + (int64_t)MAXB * LPR; p < q1; p += LPR) {
      const int32_t col = colind[p];
      const BT* blk = vals + p * ND * ND + a * ND;
      const double* xc = out + (int64_t)col * ND;
#pragma unroll
      for (int c = 0; c < ND; ++c) {
        double xcv = x_load(xc + c);
        while (ilu_is_pending(xcv)) {
          if (++polls > max_polls) {
            __hip_atomic_store(&sync[ILU_SYNC_ERR], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          xcv = x_load(xc + c);
        }
        acc += (double)blk[c] * xcv;
      }
    }
    if constexpr (LPR > 1) tp[slot][r][l] = acc;
  }
  if constexpr (LPR == 1) {
    // one lane per row: the row sums need no reduction over lanes, so one
    // barrier (for the rows of v) instead of three; the sums keep the
    // lane-reduced form 0 + acc, 0 + y, so the bits are the same
    if (act) {
      const double sum = 0.0 + acc;
      v[slot][r] = (float)(forward ? bval - sum : sum);
    }
    __syncthreads();
    if (act) {
      float y = 0.0f;
#pragma unroll
      for (int k = 0; k < PF; ++k)
        if (k < N) y += mrow[k] * v[slot][k];
#pragma unroll
      for (int k = PF; k < NC; ++k)
        if (k < N) y += Mi[k * N + r] * v[slot][k];
      const float ys = 0.0f + y;
      x_store(out + e * N + r, forward ? (double)ys : bval - (double)ys);
      // this thread read the row of `in` (which `reset` may alias) above
      x_store(reset + e * N + r, __longlong_as_double((long long)ILU_PENDING));
    }
  } else {
    __syncthreads();
    if (act && l == 0) {
      double sum = 0.0;
#pragma unroll
      for (int k = 0; k < LPR; ++k) sum += tp[slot][r][k];
      v[slot][r] = (float)(forward ? bval - sum : sum);
    }
    __syncthreads();
    if (act) {
      float y = 0.0f;
#pragma unroll
      for (int k = 0; k < PF; ++k)
        if (l + k * LPR < N) y += mrow[k] * v[slot][l + k * LPR];
#pragma unroll
      for (int k = PF; k < NC; ++k)
        if (l + k * LPR < N) y += Mi[(l + k * LPR) * N + r] * v[slot][l + k * LPR];
      yp[slot][r][l] = y;
    }
    __syncthreads();
    if (act && l == 0) {
      float y = 0.0f;
#pragma unroll
      for (int k = 0; k < LPR; ++k) y += yp[slot][r][k];
      x_store(out + e * N + r, forward ? (double)y : bval - (double)y);
      // this thread read the row of `in` (which `reset` may alias) above
      x_store(reset + e * N + r, __longlong_as_double((long long)ILU_PENDING));
    }
  }
}

// lanes per row of the sweeps: 1 for diagonal-E operators (a row meets at
// most one neighbour node per face), up to 8 for dense face stencils
template <int N, bool DIAGE>
constexpr int ilu_lanes() {
  return DIAGE ? 1 : (256 / N >= 8 ? 8 : 256 / N);
}

// dependency levels of the natural-order sweeps (host, once per context)
static int ilu_structure(pdes_ctx* c, pdes_precond* p) {
  const int nel = c->numEl, nf = c->nfaces;
  std::vector<int32_t> fl(nel, 0), bl(nel, 0);
  int fmax = 0, bmax = 0;
  for (int e = 0; e < nel; ++e) {
    int lv = 0;
    for (int f = 0; f < nf; ++f) {
      const int k = c->h_elface[(size_t)e * nf + f].x;
      if (k >= 0 && k < e) lv = std::max(lv, fl[k] + 1);
    }
    fl[e] = lv;
    fmax = std::max(fmax, lv);
  }
  for (int e = nel - 1; e >= 0; --e) {
    int lv = 0;
    for (int f = 0; f < nf; ++f) {
      const int k = c->h_elface[(size_t)e * nf + f].x;
      if (k > e) lv = std::max(lv, bl[k] + 1);
    }
    bl[e] = lv;
    bmax = std::max(bmax, lv);
  }
  auto bucket = [&](const std::vector<int32_t>& lv, int nl, std::vector<int64_t>& ptr, std::vector<int32_t>& el) {
    ptr.assign(nl + 2, 0);
    for (int e = 0; e < nel; ++e) ptr[lv[e] + 1]++;
    for (int l = 0; l <= nl; ++l) ptr[l + 1] += ptr[l];
    el.assign(nel, 0);
    std::vector<int64_t> pos(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < nel; ++e) el[pos[lv[e]]++] = e;
  };
  std::vector<int32_t> fel, bel;
  bucket(fl, fmax, p->fptr, fel);
  bucket(bl, bmax, p->bptr, bel);
  // offset of the own-element block run in every row of the BSR pattern
  const pdes_jacobian* J = c->jac;
  const int64_t nrows = (int64_t)nel * c->Np;
  std::vector<int32_t> own(nrows, 0);
  for (int64_t r = 0; r < nrows; ++r) {
    const int32_t first = (int32_t)((r / c->Np) * c->Np);
    int64_t q = J->h_rowptr[r];
    while (q < J->h_rowptr[r + 1] && J->h_colind[q] < first) ++q;
    own[r] = (int32_t)(q - J->h_rowptr[r]);
  }
  if (hipMalloc((void**)&p->d_own, std::max<int64_t>(1, nrows) * sizeof(int32_t)) != hipSuccess) {
    set_error("device allocation of the ILU row offsets failed");
    return PDES_ERR_NOMEM;
  }
  PDES_HIP(hipMemcpyAsync(p->d_own, own.data(), nrows * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  if (hipMalloc((void**)&p->d_fel, std::max(1, nel) * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void**)&p->d_bel, std::max(1, nel) * sizeof(int32_t)) != hipSuccess) {
    set_error("device allocation of the ILU level lists failed");
    return PDES_ERR_NOMEM;
  }
  PDES_HIP(hipMemcpyAsync(p->d_fel, fel.data(), nel * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  PDES_HIP(hipMemcpyAsync(p->d_bel, bel.data(), nel * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  {
    // chunked, padded copies for the dependency-driven sweeps
    const int C = ilu_dep_chunk(c->Np * c->nd, c->diag_e != 0);
    auto pad = [&](const std::vector<int64_t>& ptr, const std::vector<int32_t>& el, std::vector<int32_t>& out) {
      out.clear();
      for (size_t lv = 0; lv + 1 < ptr.size(); ++lv) {
        for (int64_t q = ptr[lv]; q < ptr[lv + 1]; ++q) out.push_back(el[q]);
        while (out.size() % C) out.push_back(-1);
      }
    };
    std::vector<int32_t> fpad, bpad;
    pad(p->fptr, fel, fpad);
    pad(p->bptr, bel, bpad);
    p->chunk = C;
    p->fpad_n = (int64_t)fpad.size() / C;
    p->bpad_n = (int64_t)bpad.size() / C;
    if (hipMalloc((void**)&p->d_fpad, std::max<size_t>(1, fpad.size()) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void**)&p->d_bpad, std::max<size_t>(1, bpad.size()) * sizeof(int32_t)) != hipSuccess) {
      set_error("device allocation of the ILU chunk lists failed");
      return PDES_ERR_NOMEM;
    }
    PDES_HIP(hipMemcpyAsync(p->d_fpad, fpad.data(), fpad.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    PDES_HIP(hipMemcpyAsync(p->d_bpad, bpad.data(), bpad.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    PDES_HIP(hipStreamSynchronize(c->stream));   // the host vectors go out of scope
  }
  p->force = std::getenv("PDES_ILU_LEVELS") ? 1 : std::getenv("PDES_ILU_DEP") ? 2 : 0;
  p->h_fl = fl;
  p->h_bl = bl;
  p->blk_per_el = (double)(J->nnzb - (int64_t)nel * c->Np * c->Np) / 2.0 / std::max(1, nel);
  const int64_t ny = std::max<int64_t>(1, c->numDof);
  // d_y is set pending with each factorization (pc_setup)
  if (hipMalloc((void**)&p->d_y, ny * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&p->d_sync, ILU_SYNC_WORDS * sizeof(int32_t)) != hipSuccess) {
    set_error("device allocation of the ILU sweep vector failed");
    return PDES_ERR_NOMEM;
  }
  PDES_HIP(hipMemsetAsync(p->d_sync, 0, ILU_SYNC_WORDS * sizeof(int32_t), c->stream));
  // PDES_ILU_WAIT_POLLS: a smaller bound on the dependency waits (for tests
  // of the give-up path; the default bound is never reached by a sweep that
  // runs normally)
  const char* polls_env = std::getenv("PDES_ILU_WAIT_POLLS");
  const int32_t poll_limit = polls_env ? (int32_t)std::max(1L, std::strtol(polls_env, nullptr, 10)) : 0;
  if (poll_limit)
    PDES_HIP(hipMemcpyAsync(p->d_sync + ILU_SYNC_POLLS, &poll_limit, sizeof(int32_t), hipMemcpyHostToDevice,
                            c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  p->ilu_ready = true;
  return PDES_OK;
}

// Which sweeps an application runs (both give the same bits).  A level
// launch costs at least ILU_LEVEL_FLOOR_US and streams its elements' bytes at
// ILU_LEVEL_BW; the dependency-driven sweep takes at least ILU_DEP_STEP_US per
// level of its critical path, ILU_DEP_TICKET_US per chunk, and streams at
// ILU_DEP_BW (measured on MI355X:
// 2D diagonal-E p=2 sweeps over 246 levels of 123 elements take 4.8 us per
// level launch and 1.7 us per dependency step; 3D Omega p=2 sweeps over 96
// levels of 864 tets stream 2.1 TB/s level by level, 1.7 TB/s dependency-
// driven).  The elements and levels are those of the instances being solved
// when the batched solver says so, else the whole mesh.
constexpr double ILU_LEVEL_FLOOR_US = 4.5, ILU_LEVEL_BW = 2.1e6;   // bytes per us
constexpr double ILU_DEP_STEP_US = 1.7, ILU_DEP_BW = 1.7e6;
// and one ticket (a device-scope atomic on one word) per chunk: about 21 ns
// each, measured as the per-element rate of one-element chunks
constexpr double ILU_DEP_TICKET_US = 0.021;

static bool ilu_use_dep(const pdes_ctx* c, const pdes_precond* p) {
  if (p->force) return p->force == 2;
  const int nd = c->nd, n = c->Np * nd;
  const double be = p->blk_per_el * nd * nd * (c->jac->fp64 ? 8.0 : 4.0) + (double)n * n * 4.0 + 3.0 * n * 8.0;
  const bool hint = c->ilu_hint_el >= 0;
  const double nel = hint ? (double)c->ilu_hint_el : (double)c->numEl;
  double lev = 0.0, dep = 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    const double L = std::max(1, hint ? (pass ? c->ilu_hint_lb : c->ilu_hint_lf)
                                      : (int)(pass ? p->bptr.size() : p->fptr.size()) - 1);
    lev += L * std::max(ILU_LEVEL_FLOOR_US, nel / L * be / ILU_LEVEL_BW);
    dep += std::max(std::max(L * ILU_DEP_STEP_US, (nel / p->chunk + L) * ILU_DEP_TICKET_US), nel * be / ILU_DEP_BW);
  }
  return dep < lev;
}

struct ChunkSelected {
  const int32_t* order;
  int chunk;
  const int32_t* el_inst;
  const int32_t* sel;
  __host__ __device__ bool operator()(const int32_t& ch) const {
    for (int k = 0; k < chunk; ++k) {
      const int32_t e = order[(int64_t)ch * chunk + k];
      if (e >= 0 && sel[el_inst[e]]) return true;
    }
    return false;
  }
};

// The batched solver's hint for the ILU(0) applications of its next Arnoldi
// steps: they touch only the elements of the instances with inst_sel[i] != 0
// (nel of them, over lf forward and lb backward levels).  When the cost model
// then picks the dependency-driven sweeps, the level-ordered element lists are
// compacted to those instances (order kept), so a sweep dispatches and
// tickets only their elements.  ilu_hint_clear ends the hint.
int ilu_hint(pdes_ctx* c, const int32_t* inst_sel, int ninst, int64_t nel, int lf, int lb) {
  pdes_precond* p = c->pc_obj;
  c->ilu_hint_el = nel;
  c->ilu_hint_lf = lf;
  c->ilu_hint_lb = lb;
  c->ilu_sel = false;
  if (!p || p->kind != PDES_PC_ILU0 || !c->d_el_inst || !ilu_use_dep(c, p)) return PDES_OK;
  const int64_t nf = p->fpad_n, nb = p->bpad_n;
  hipcub::CountingInputIterator<int32_t> ids(0);
  if (!p->d_fsel) {
    if (hipMalloc((void**)&p->d_fsel, std::max<int64_t>(1, nf) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void**)&p->d_bsel, std::max<int64_t>(1, nb) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void**)&p->d_nsel, 2 * sizeof(int32_t)) != hipSuccess) {
      set_error("device allocation of the ILU selection lists failed");
      return PDES_ERR_NOMEM;
    }
    size_t bytes = 0;
    PDES_HIP(hipcub::DeviceSelect::If(nullptr, bytes, ids, p->d_fsel, p->d_nsel, (int)std::max(nf, nb),
                                      ChunkSelected{p->d_fpad, p->chunk, c->d_el_inst, nullptr}, c->stream));
    if (hipMalloc(&p->d_select_tmp, std::max<size_t>(bytes, 16)) != hipSuccess) {
      set_error("device allocation of the ILU selection work space failed");
      return PDES_ERR_NOMEM;
    }
    p->select_tmp_bytes = bytes;
  }
  if (p->sel_ninst < ninst) {
    if (p->d_inst_sel) hipFree(p->d_inst_sel);
    p->d_inst_sel = nullptr;
    if (hipMalloc((void**)&p->d_inst_sel, ninst * sizeof(int32_t)) != hipSuccess) {
      set_error("device allocation of the ILU instance selection failed");
      return PDES_ERR_NOMEM;
    }
    p->sel_ninst = ninst;
  }
  PDES_HIP(hipMemcpyAsync(p->d_inst_sel, inst_sel, ninst * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  size_t bytes = p->select_tmp_bytes;
  PDES_HIP(hipcub::DeviceSelect::If(p->d_select_tmp, bytes, ids, p->d_fsel, p->d_nsel, (int)nf,
                                    ChunkSelected{p->d_fpad, p->chunk, c->d_el_inst, p->d_inst_sel}, c->stream));
  bytes = p->select_tmp_bytes;
  PDES_HIP(hipcub::DeviceSelect::If(p->d_select_tmp, bytes, ids, p->d_bsel, p->d_nsel + 1, (int)nb,
                                    ChunkSelected{p->d_bpad, p->chunk, c->d_el_inst, p->d_inst_sel}, c->stream));
  int32_t cnt[2] = {0, 0};
  PDES_HIP(hipMemcpyAsync(cnt, p->d_nsel, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  p->sel_nf = cnt[0];
  p->sel_nb = cnt[1];
  c->ilu_sel = true;
  return PDES_OK;
}

void ilu_hint_clear(pdes_ctx* c) {
  c->ilu_hint_el = -1;
  c->ilu_sel = false;
}

// levels spanned by the elements [el0, el1) (one instance of a packed batch)
int ilu_range_levels(const pdes_ctx* c, int64_t el0, int64_t el1, int* lf, int* lb) {
  const pdes_precond* p = c->pc_obj;
  *lf = *lb = 0;
  if (!p || p->h_fl.empty()) return PDES_ERR_ARG;
  for (int64_t e = el0; e < el1; ++e) {
    *lf = std::max(*lf, p->h_fl[e] + 1);
    *lb = std::max(*lb, p->h_bl[e] + 1);
  }
  return PDES_OK;
}

template <int NP, int ND, int NS>
static int ilu_factor(pdes_ctx* c) {
  pdes_jacobian* j = c->jac;
  pdes_precond* p = c->pc_obj;
  for (size_t l = 0; l + 1 < p->fptr.size(); ++l) {
    const int64_t cnt = p->fptr[l + 1] - p->fptr[l];
    if (cnt == 0) continue;
    const int32_t* lev = p->d_fel + p->fptr[l];
    if (j->fp64)
      hipLaunchKernelGGL((ilu_factor_level_kernel<NP, ND, NS, double>), dim3((unsigned)cnt), dim3(256), 0, c->stream,
                         lev, j->d_rowptr, j->d_colind, (const double*)j->d_vals, c->d_elface, c->nfaces, c->d_perm,
                         p->d_inv, j->d_err, c->d_el_inst, c->d_inst_active);
    else
      hipLaunchKernelGGL((ilu_factor_level_kernel<NP, ND, NS, float>), dim3((unsigned)cnt), dim3(256), 0, c->stream,
                         lev, j->d_rowptr, j->d_colind, (const float*)j->d_vals, c->d_elface, c->nfaces, c->d_perm,
                         p->d_inv, j->d_err, c->d_el_inst, c->d_inst_active);
  }
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

template <int NP, int ND, int NS, bool DIAGE>
static int ilu_apply(pdes_ctx* c, const double* b, double* x) {
  pdes_jacobian* j = c->jac;
  pdes_precond* p = c->pc_obj;
  constexpr int N = NP * ND;
  constexpr int LPR = ilu_lanes<N, DIAGE>();
  constexpr int EPB = 256 / (N * LPR);
  if (ilu_use_dep(c, p)) {
    constexpr int CH = ilu_dep_chunk(N, DIAGE);
    static_assert(ilu_dep_lanes(N, DIAGE) == LPR, "the dependency-driven and level sweeps split rows alike");
    if (p->chunk != CH) {
      set_error("ILU chunk size mismatch");
      return PDES_ERR_ARG;
    }
    const bool sel = c->ilu_sel && c->ilu_hint_el >= 0;
    for (int pass = 0; pass < 2; ++pass) {
      const int32_t* order = pass == 0 ? p->d_fpad : p->d_bpad;
      const int32_t* chunks = sel ? (pass == 0 ? p->d_fsel : p->d_bsel) : nullptr;
      const int64_t nch = sel ? (pass == 0 ? p->sel_nf : p->sel_nb) : (pass == 0 ? p->fpad_n : p->bpad_n);
      if (nch == 0) continue;
      // forward: b -> y, x set pending; backward: y -> x, y set pending again
      const double* in = pass == 0 ? b : p->d_y;
      double* out = pass == 0 ? p->d_y : x;
      double* reset = pass == 0 ? x : p->d_y;
      hipLaunchKernelGGL(ilu_sweep_begin, dim3(1), dim3(1), 0, c->stream, p->d_sync);
      if (j->fp64)
        hipLaunchKernelGGL((ilu_sweep_dep_kernel<NP, ND, NS, DIAGE, double, LPR, CH>), dim3((unsigned)nch),
                           dim3(CH * N * LPR), 0, c->stream, order, chunks, nch, j->d_rowptr, j->d_colind, p->d_own,
                           (const double*)j->d_vals, p->d_inv, in, out, reset, pass == 0, c->d_el_inst, c->d_mask,
                           c->d_inst_active, p->d_sync);
      else
        hipLaunchKernelGGL((ilu_sweep_dep_kernel<NP, ND, NS, DIAGE, float, LPR, CH>), dim3((unsigned)nch),
                           dim3(CH * N * LPR), 0, c->stream, order, chunks, nch, j->d_rowptr, j->d_colind, p->d_own,
                           (const float*)j->d_vals, p->d_inv, in, out, reset, pass == 0, c->d_el_inst, c->d_mask,
                           c->d_inst_active, p->d_sync);
    }
    PDES_HIP(hipGetLastError());
    return PDES_OK;
  }
  for (int pass = 0; pass < 2; ++pass) {
    const std::vector<int64_t>& ptr = pass == 0 ? p->fptr : p->bptr;
    const int32_t* els = pass == 0 ? p->d_fel : p->d_bel;
    for (size_t l = 0; l + 1 < ptr.size(); ++l) {
      const int64_t cnt = ptr[l + 1] - ptr[l];
      if (cnt == 0) continue;
      const dim3 grid((unsigned)ceil_div(cnt, EPB));
      if constexpr (LPR > 1) {
        const dim3 block((unsigned)(EPB * N * LPR));
        if (j->fp64)
          hipLaunchKernelGGL((ilu_sweep_level_ml_kernel<NP, ND, double, LPR>), grid, block, 0, c->stream, els + ptr[l],
                             cnt, j->d_rowptr, j->d_colind, p->d_own, (const double*)j->d_vals, p->d_inv, b, x,
                             pass == 0, c->d_el_inst, c->d_mask, c->d_inst_active);
        else
          hipLaunchKernelGGL((ilu_sweep_level_ml_kernel<NP, ND, float, LPR>), grid, block, 0, c->stream, els + ptr[l],
                             cnt, j->d_rowptr, j->d_colind, p->d_own, (const float*)j->d_vals, p->d_inv, b, x,
                             pass == 0, c->d_el_inst, c->d_mask, c->d_inst_active);
        continue;
      }
      if (j->fp64)
        hipLaunchKernelGGL((ilu_sweep_level_kernel<NP, ND, double>), grid, dim3(256), 0, c->stream, els + ptr[l], cnt,
                           j->d_rowptr, j->d_colind, p->d_own, (const double*)j->d_vals, p->d_inv, b, x, pass == 0,
                           c->d_el_inst, c->d_mask, c->d_inst_active);
      else
        hipLaunchKernelGGL((ilu_sweep_level_kernel<NP, ND, float>), grid, dim3(256), 0, c->stream, els + ptr[l], cnt,
                           j->d_rowptr, j->d_colind, p->d_own, (const float*)j->d_vals, p->d_inv, b, x, pass == 0,
                           c->d_el_inst, c->d_mask, c->d_inst_active);
    }
  }
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

template <int NP, int ND>
static int bj_factor(pdes_ctx* c) {
  pdes_jacobian* j = c->jac;
  pdes_precond* p = c->pc_obj;
  if (c->active_count() == 0) return PDES_OK;
  const dim3 grid((unsigned)c->active_count());
  if (j->fp64)
    hipLaunchKernelGGL((bjacobi_factor_kernel<NP, ND, double>), grid, dim3(256), 0, c->stream, c->active_list(),
                       j->d_rowptr, j->d_colind, (const double*)j->d_vals, p->d_inv, j->d_err);
  else
    hipLaunchKernelGGL((bjacobi_factor_kernel<NP, ND, float>), grid, dim3(256), 0, c->stream, c->active_list(),
                       j->d_rowptr, j->d_colind, (const float*)j->d_vals, p->d_inv, j->d_err);
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

template <int NP, int ND>
static int bj_apply(pdes_ctx* c, const double* b, double* x) {
  if (c->active_count() == 0) return PDES_OK;
  constexpr int EPB = 256 / (NP * ND);
  const int64_t nel = c->active_count();
  hipLaunchKernelGGL((bjacobi_apply_kernel<NP, ND>), dim3((unsigned)ceil_div(nel, EPB)), dim3(256), 0, c->stream,
                     c->active_list(), nel, c->pc_obj->d_inv, b, x, c->d_el_inst, c->d_mask);
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

static int pc_factor(pdes_ctx* c) {
  if (c->dim == 2) {
    switch (c->Np) {
      case 3: return bj_factor<3, 4>(c);
      case 6: return bj_factor<6, 4>(c);
      case 10: return bj_factor<10, 4>(c);
      case 15: return bj_factor<15, 4>(c);
      case 7: return bj_factor<7, 4>(c);
      case 12: return bj_factor<12, 4>(c);
      case 18: return bj_factor<18, 4>(c);
      case 27: return bj_factor<27, 4>(c);
    }
  } else {
    switch (c->Np) {
      case 4: return bj_factor<4, 5>(c);
      case 10: return bj_factor<10, 5>(c);
    }
  }
  set_error("no block-Jacobi kernel for this operator");
  return PDES_ERR_UNSUPPORTED;
}

__global__ void copy_kernel(const double* __restrict__ a, double* __restrict__ b, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

static int ilu_factor_dispatch(pdes_ctx* c) {
#define PDES_ILU_F(D_, NP_, NF_, NS_, DE_) \
  if (c->dim == D_ && c->Np == NP_ && (c->diag_e != 0) == DE_) return ilu_factor<NP_, D_ + 2, NS_>(c);
  PDES_FOR_EACH_SHAPE(PDES_ILU_F)
#undef PDES_ILU_F
  set_error("no ILU(0) kernel for this operator");
  return PDES_ERR_UNSUPPORTED;
}

static int ilu_apply_dispatch(pdes_ctx* c, const double* b, double* x) {
#define PDES_ILU_A(D_, NP_, NF_, NS_, DE_) \
  if (c->dim == D_ && c->Np == NP_ && (c->diag_e != 0) == DE_) return ilu_apply<NP_, D_ + 2, NS_, DE_>(c, b, x);
  PDES_FOR_EACH_SHAPE(PDES_ILU_A)
#undef PDES_ILU_A
  set_error("no ILU(0) kernel for this operator");
  return PDES_ERR_UNSUPPORTED;
}

// ILU(0) apply through a captured graph: b is copied into the staging input,
// the graph (all level launches) is replayed, the staging output is copied to
// x.  Falls back to direct launches if the capture is refused.
static int ilu_apply_graphed(pdes_ctx* c, const double* b, double* x) {
  pdes_precond* p = c->pc_obj;
  const size_t bytes = (size_t)c->numDof * sizeof(double);
  if (!p->d_stage_in) {
    if (hipMalloc((void**)&p->d_stage_in, bytes) != hipSuccess ||
        hipMalloc((void**)&p->d_stage_out, bytes) != hipSuccess) {
      set_error("ILU staging allocation failed");
      return PDES_ERR_NOMEM;
    }
  }
  hipGraphExec_t exec = nullptr;
  for (auto& g : p->graphs)
    if (g.mask == c->d_mask && g.vals == c->jac->d_vals && g.fp64 == c->jac->fp64 && g.inv == p->d_inv)
      exec = g.exec;
  if (!exec) {
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      return ilu_apply_dispatch(c, b, x);
    }
    const int st = ilu_apply_dispatch(c, p->d_stage_in, p->d_stage_out);
    const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    if (st) {
      if (graph) hipGraphDestroy(graph);
      return st;
    }
    if (ec != hipSuccess || !graph || hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      if (graph) hipGraphDestroy(graph);
      (void)hipGetLastError();
      return ilu_apply_dispatch(c, b, x);
    }
    hipGraphDestroy(graph);
    p->graphs.push_back({c->d_mask, c->jac->d_vals, c->jac->fp64, p->d_inv, exec});
  }
  PDES_HIP(hipMemcpyAsync(p->d_stage_in, b, bytes, hipMemcpyDeviceToDevice, c->stream));
  PDES_HIP(hipGraphLaunch(exec, c->stream));
  PDES_HIP(hipMemcpyAsync(x, p->d_stage_out, bytes, hipMemcpyDeviceToDevice, c->stream));
  return PDES_OK;
}

int pc_apply(pdes_ctx* c, const double* b, double* x) {
  // the dependency-driven sweeps are 4 launches; the level-by-level ones are
  // replayed from a captured graph
  if (c->pc_obj && c->pc_obj->kind == PDES_PC_ILU0)
    return c->use_graphs && !c->capturing && !ilu_use_dep(c, c->pc_obj) ? ilu_apply_graphed(c, b, x)
                                                                        : ilu_apply_dispatch(c, b, x);
  if (!c->pc_obj || c->pc_obj->kind == PDES_PC_NONE) {
    const int64_t n = c->numDof;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)std::min<int64_t>(8192, ceil_div(n, 256))), dim3(256), 0,
                       c->stream, b, x, n);
    PDES_HIP(hipGetLastError());
    return PDES_OK;
  }
  if (c->dim == 2) {
    switch (c->Np) {
      case 3: return bj_apply<3, 4>(c, b, x);
      case 6: return bj_apply<6, 4>(c, b, x);
      case 10: return bj_apply<10, 4>(c, b, x);
      case 15: return bj_apply<15, 4>(c, b, x);
      case 7: return bj_apply<7, 4>(c, b, x);
      case 12: return bj_apply<12, 4>(c, b, x);
      case 18: return bj_apply<18, 4>(c, b, x);
      case 27: return bj_apply<27, 4>(c, b, x);
    }
  } else {
    switch (c->Np) {
      case 4: return bj_apply<4, 5>(c, b, x);
      case 10: return bj_apply<10, 5>(c, b, x);
    }
  }
  set_error("no block-Jacobi kernel for this operator");
  return PDES_ERR_UNSUPPORTED;
}

// What a captured sequence of preconditioner applications depends on besides
// device data: the kind, the factor and matrix buffers, and for ILU(0) the
// sweep kind and, for the dependency-driven sweeps over compacted chunk
// lists, their lengths (launch sizes).
void pc_graph_key(const pdes_ctx* c, int64_t key[6]) {
  const pdes_precond* p = c->pc_obj;
  for (int i = 0; i < 6; ++i) key[i] = 0;
  if (!p) return;
  key[0] = p->kind;
  key[1] = (int64_t)(intptr_t)p->d_inv;
  key[2] = c->jac ? (int64_t)(intptr_t)c->jac->d_vals : 0;
  key[3] = c->jac ? (int64_t)c->jac->fp64 : 0;
  if (p->kind == PDES_PC_ILU0) {
    const bool dep = ilu_use_dep(c, p);
    const bool sel = dep && c->ilu_sel && c->ilu_hint_el >= 0;
    key[4] = dep ? 1 + (sel ? 1 : 0) : 0;
    key[5] = sel ? p->sel_nf * (int64_t(1) << 32) + p->sel_nb : 0;
  }
}

// the error word of the dependency-driven sweeps (a wait that gave up): the
// preconditioner applications since the last check are not to be trusted
int pc_check(pdes_ctx* c) {
  pdes_precond* p = c->pc_obj;
  if (!p || p->kind != PDES_PC_ILU0 || !p->d_sync) return PDES_OK;
  int32_t err = 0;
  PDES_HIP(hipMemcpyAsync(&err, p->d_sync + ILU_SYNC_ERR, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  if (err) {
    PDES_HIP(hipMemsetAsync(p->d_sync + ILU_SYNC_ERR, 0, sizeof(int32_t), c->stream));
    set_error("an ILU(0) sweep gave up waiting for a neighbour element (its result is invalid)");
    return PDES_ERR_HIP;
  }
  return PDES_OK;
}

int pc_setup(pdes_ctx* c, int kind) {
  PDES_CHECK(kind == PDES_PC_NONE || kind == PDES_PC_BJACOBI || kind == PDES_PC_ILU0, PDES_ERR_UNSUPPORTED,
             "preconditioner kind %d is not available", kind);
  if (!c->pc_obj) c->pc_obj = new pdes_precond();
  pdes_precond* p = c->pc_obj;
  p->kind = kind;
  if (kind == PDES_PC_NONE) { p->factored = true; return PDES_OK; }
  PDES_CHECK(c->jac && c->jac->assembled, PDES_ERR_ARG, "assemble the Jacobian before the preconditioner");
  const int n = c->Np * c->nd;
  if (!p->d_inv || p->n != n) {
    if (p->d_inv) hipFree(p->d_inv);
    for (auto& g : p->graphs) hipGraphExecDestroy(g.exec);
    p->graphs.clear();
    p->n = n;
    if (hipMalloc((void**)&p->d_inv, (size_t)c->numEl * n * n * sizeof(float)) != hipSuccess) {
      set_error("device allocation of the block-Jacobi factors failed");
      return PDES_ERR_NOMEM;
    }
  }
  int st;
  if (kind == PDES_PC_ILU0) {
    if (!p->ilu_ready && (st = ilu_structure(c, p)) != 0) return st;
    st = ilu_factor_dispatch(c);
    // the dependency-driven sweeps' vector holds ILU_PENDING between
    // applications; set it again with each factorization, so that an
    // application cut short between its two sweeps cannot leave values
    // behind for the next factorization's sweeps
    if (!st) {
      const int64_t ny = std::max<int64_t>(1, c->numDof);
      hipLaunchKernelGGL(ilu_pending_fill, dim3((unsigned)ceil_div(ny, 256)), dim3(256), 0, c->stream, p->d_y, ny);
      PDES_HIP(hipGetLastError());
    }
  } else {
    st = pc_factor(c);
  }
  if (st) return st;
  p->factored = true;
  return PDES_OK;
}

// ---------------------------------------------------------------------------
// GMRES building blocks
// ---------------------------------------------------------------------------
constexpr int MD_THREADS = 256;
constexpr int MD_BLOCKS = 2048;

// partial[i][b] = sum over block b's slice of V_i . w.  MD_GROUP basis
// vectors per pass over the slice, so w is read once per group; each vector's
// terms are summed in the same order, and reduced by the same tree, as one
// vector at a time would be.
constexpr int MD_GROUP = 8;
__global__ void __launch_bounds__(MD_THREADS)
multidot_partial(const double* __restrict__ V, int m, int64_t ldv, const double* __restrict__ w, int64_t n,
                 double* __restrict__ partial) {
  __shared__ double sh[MD_GROUP][MD_THREADS];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int i0 = 0; i0 < m; i0 += MD_GROUP) {
    const int g_n = m - i0 < MD_GROUP ? m - i0 : MD_GROUP;
    double acc[MD_GROUP];
#pragma unroll
    for (int g = 0; g < MD_GROUP; ++g) acc[g] = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
      const double wt = w[t];
#pragma unroll
      for (int g = 0; g < MD_GROUP; ++g)
        if (g < g_n) acc[g] += V[(int64_t)(i0 + g) * ldv + t] * wt;
    }
#pragma unroll
    for (int g = 0; g < MD_GROUP; ++g) sh[g][threadIdx.x] = acc[g];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) {
#pragma unroll
        for (int g = 0; g < MD_GROUP; ++g) sh[g][threadIdx.x] += sh[g][threadIdx.x + s];
      }
      __syncthreads();
    }
    if ((int)threadIdx.x < g_n) partial[(int64_t)(i0 + threadIdx.x) * gridDim.x + blockIdx.x] = sh[threadIdx.x][0];
    __syncthreads();
  }
}

__global__ void multidot_reduce(const double* __restrict__ partial, int nb, double* __restrict__ out) {
  __shared__ double sh[256];
  const int i = blockIdx.x;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += partial[(int64_t)i * nb + b];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = sh[0];
}

// w -= sum_i h[i] V_i
__global__ void multiaxpy(const double* __restrict__ V, int m, int64_t ldv, const double* __restrict__ h,
                          double* __restrict__ w, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    double acc = w[t];
    for (int i = 0; i < m; ++i) acc -= h[i] * V[(int64_t)i * ldv + t];
    w[t] = acc;
  }
}

// y = a*x (+ y if accumulate)
__global__ void scale_kernel(const double* __restrict__ x, double* __restrict__ y, double a, int accumulate, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
    y[t] = accumulate ? y[t] + a * x[t] : a * x[t];
}

// y = sum_i c[i] V_i  (c on the host, passed through a small device buffer)
__global__ void combine_kernel(const double* __restrict__ V, int m, int64_t ldv, const double* __restrict__ cf,
                               double* __restrict__ y, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    double acc = 0.0;
    for (int i = 0; i < m; ++i) acc += cf[i] * V[(int64_t)i * ldv + t];
    y[t] = acc;
  }
}

static unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(8192, ceil_div(n, 256))); }

struct GmresWork;
static void gmres_free(GmresWork& W);

struct GmresWork {
  double* V = nullptr;      // (restart+1) x n
  double* w = nullptr;
  double* z = nullptr;
  double* partial = nullptr;
  double* coef = nullptr;   // device small buffer
  double* hcoef = nullptr;  // pinned host buffer
  int restart = 0;
  int64_t n = 0;
};

static int gmres_alloc(pdes_ctx* c, GmresWork& W, int restart) {
  if (c->gmres_ws) {
    GmresWork* cached = (GmresWork*)c->gmres_ws;
    if (cached->restart == restart && cached->n == c->numDof) { W = *cached; return PDES_OK; }
    gmres_free(*cached);
    delete cached;
    c->gmres_ws = nullptr;
  }
  W.restart = restart;
  W.n = c->numDof;
  const size_t n = (size_t)c->numDof;
  if (hipMalloc((void**)&W.V, (size_t)(restart + 1) * n * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&W.w, n * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&W.z, n * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&W.partial, (size_t)(restart + 2) * MD_BLOCKS * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&W.coef, (size_t)(2 * restart + 8) * sizeof(double)) != hipSuccess ||
      hipHostMalloc((void**)&W.hcoef, (size_t)(2 * restart + 8) * sizeof(double)) != hipSuccess) {
    set_error("GMRES workspace allocation failed");
    return PDES_ERR_NOMEM;
  }
  return PDES_OK;
}

static void gmres_keep(pdes_ctx* c, GmresWork& W) {
  if (!c->gmres_ws) c->gmres_ws = new GmresWork(W);
}

static void gmres_free(GmresWork& W) {
  if (W.V) hipFree(W.V);
  if (W.w) hipFree(W.w);
  if (W.z) hipFree(W.z);
  if (W.partial) hipFree(W.partial);
  if (W.coef) hipFree(W.coef);
  if (W.hcoef) hipHostFree(W.hcoef);
  W = GmresWork();
}

void gmres_release(pdes_ctx* c) {
  if (c->gmres_ws) {
    gmres_free(*(GmresWork*)c->gmres_ws);
    delete (GmresWork*)c->gmres_ws;
    c->gmres_ws = nullptr;
  }
}

// dots of w with V_0..V_{m-1} into W.coef[off..off+m)
static int multidot(pdes_ctx* c, GmresWork& W, int m, const double* w, int off) {
  const int nb = (int)std::min<int64_t>(MD_BLOCKS, std::max<int64_t>(1, ceil_div(W.n, MD_THREADS)));
  hipLaunchKernelGGL(multidot_partial, dim3(nb), dim3(MD_THREADS), 0, c->stream, W.V, m, W.n, w, W.n, W.partial);
  hipLaunchKernelGGL(multidot_reduce, dim3(m), dim3(256), 0, c->stream, W.partial, nb, W.coef + off);
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

static int norm2_dev(pdes_ctx* c, GmresWork& W, const double* x, int off) {
  return launch_wnorm2(c, x, nullptr, W.n, W.coef + off);
}

static int apply_op(pdes_ctx* c, int mode, const double* x, double* y) {
  if (mode == 0) return launch_spmv(c, x, y, 1.0, nullptr, 0.0);
  int st = launch_residual(c, c->d_q, x, nullptr, y, false);
  if (!st && c->hom_lambda >= 0.0)
    st = launch_homotopy(c, c->d_q, x, 1.0 - c->hom_lambda, c->hom_lambda, nullptr, y);
  return st ? st : launch_ptc_shift(c, x, y);
}

// GMRES(restart), right preconditioned, x0 = x on entry (device vectors).
int gmres_device(pdes_ctx* c, const double* b, double* x, double rtol, double atol, int restart, int maxit,
                 int mode, int* iters, double* resnorm, bool x_zero) {
  GmresWork W;
  int st = gmres_alloc(c, W, restart);
  if (st) return st;
  const int64_t n = W.n;
  std::vector<double> H((size_t)(restart + 1) * restart, 0.0), cs(restart), sn(restart), g(restart + 1);
  int total = 0;
  double beta = 0.0, beta0 = -1.0;
  if (x_zero) PDES_HIP(hipMemsetAsync(x, 0, n * sizeof(double), c->stream));
  bool converged = false;
  while (true) {
    // r = b - A x  -> V_0
    if (x_zero && total == 0) {
      hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, b, W.V, n);
    } else {
        if ((st = apply_op(c, mode, x, W.w)) != 0) goto done;
      hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, b, W.V, n);
      hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, W.w, W.V, -1.0, 1, n);
    }
    if ((st = norm2_dev(c, W, W.V, 0)) != 0) goto done;
    PDES_HIP(hipMemcpyAsync(W.hcoef, W.coef, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PDES_HIP(hipStreamSynchronize(c->stream));
    beta = std::sqrt(W.hcoef[0]);
    if (beta0 < 0) beta0 = beta;
    if (beta <= std::max(rtol * beta0, atol) || total >= maxit) { converged = beta <= std::max(rtol * beta0, atol); break; }
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, W.V, W.V, 1.0 / beta, 0, n);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int jdone = 0;
    double res_est = beta;
    for (int j = 0; j < restart && total < maxit; ++j) {
      double* Vj = W.V + (int64_t)j * n;
      double* Vn = W.V + (int64_t)(j + 1) * n;
      if ((st = pc_apply(c, Vj, W.z)) != 0) goto done;
      if ((st = apply_op(c, mode, W.z, Vn)) != 0) goto done;
      // CGS2: h = V^T w ; w -= V h ; twice
      if ((st = multidot(c, W, j + 1, Vn, 0)) != 0) goto done;
      hipLaunchKernelGGL(multiaxpy, dim3(grid_for(n)), dim3(256), 0, c->stream, W.V, j + 1, n, W.coef, Vn, n);
      if ((st = multidot(c, W, j + 1, Vn, restart + 2)) != 0) goto done;
      hipLaunchKernelGGL(multiaxpy, dim3(grid_for(n)), dim3(256), 0, c->stream, W.V, j + 1, n,
                         W.coef + restart + 2, Vn, n);
      if ((st = norm2_dev(c, W, Vn, 2 * restart + 4)) != 0) goto done;
      PDES_HIP(hipMemcpyAsync(W.hcoef, W.coef, (2 * restart + 5) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      PDES_HIP(hipStreamSynchronize(c->stream));
      double hn = std::sqrt(W.hcoef[2 * restart + 4]);
      for (int i = 0; i <= j; ++i) H[(size_t)i * restart + j] = W.hcoef[i] + W.hcoef[restart + 2 + i];
      H[(size_t)(j + 1) * restart + j] = hn;
      if (hn > 0) hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, Vn, Vn, 1.0 / hn, 0, n);
      // Givens
      for (int i = 0; i < j; ++i) {
        double a = H[(size_t)i * restart + j], bb = H[(size_t)(i + 1) * restart + j];
        H[(size_t)i * restart + j] = cs[i] * a + sn[i] * bb;
        H[(size_t)(i + 1) * restart + j] = -sn[i] * a + cs[i] * bb;
      }
      double a = H[(size_t)j * restart + j], bb = H[(size_t)(j + 1) * restart + j];
      double r = std::hypot(a, bb);
      cs[j] = r > 0 ? a / r : 1.0;
      sn[j] = r > 0 ? bb / r : 0.0;
      H[(size_t)j * restart + j] = r;
      H[(size_t)(j + 1) * restart + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      res_est = std::fabs(g[j + 1]);
      ++total;
      jdone = j + 1;
      if (res_est <= std::max(rtol * beta0, atol) || hn == 0.0) break;
    }
    // y = H^-1 g ; x += M^-1 V y
    {
      std::vector<double> y(jdone);
      for (int i = jdone - 1; i >= 0; --i) {
        double s = g[i];
        for (int k = i + 1; k < jdone; ++k) s -= H[(size_t)i * restart + k] * y[k];
        y[i] = s / H[(size_t)i * restart + i];
      }
      for (int i = 0; i < jdone; ++i) W.hcoef[i] = y[i];
      PDES_HIP(hipMemcpyAsync(W.coef, W.hcoef, jdone * sizeof(double), hipMemcpyHostToDevice, c->stream));
      hipLaunchKernelGGL(combine_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, W.V, jdone, n, W.coef, W.w, n);
      if ((st = pc_apply(c, W.w, W.z)) != 0) goto done;
      if (x_zero && total == jdone)
        hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, W.z, x, 1.0, 0, n);
      else
        hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, W.z, x, 1.0, 1, n);
      PDES_HIP(hipStreamSynchronize(c->stream));
    }
    // The Givens estimate res_est is the residual of x + M^-1 V y only when
    // M^-1 is applied exactly; the block preconditioners apply it in FP32, so
    // the loop returns to the top, where the true residual b - A x decides
    // convergence (and, when it is still too large, seeds the next cycle).
    x_zero = false;
  }
done:
  gmres_keep(c, W);
  if (st) return st;
  if ((st = pc_check(c)) != 0) return st;
  if (iters) *iters = total;
  if (resnorm) *resnorm = beta;
  (void)converged;
  return PDES_OK;
}

}  // namespace pdes

using namespace pdes;

extern "C" {

int pdes_pc_setup(pdes_ctx* c, int kind) {
  PDES_CHECK(c, PDES_ERR_ARG, "null context");
  int st = pc_setup(c, kind);
  if (st) return st;
  if (kind != PDES_PC_NONE) {
    int32_t err = 0;
    PDES_HIP(hipMemcpyAsync(&err, c->jac->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    PDES_HIP(hipStreamSynchronize(c->stream));
    if (err) {
      PDES_HIP(hipMemsetAsync(c->jac->d_err, 0, sizeof(int32_t), c->stream));
      set_error("singular element block in the preconditioner factorisation");
      return PDES_ERR_PHYS;
    }
  }
  return PDES_OK;
}

int pdes_pc_apply(pdes_ctx* c, const double* b, double* x) {
  PDES_CHECK(c && b && x, PDES_ERR_ARG, "null argument");
  PDES_HIP(hipMemcpyAsync(c->d_work, b, c->numDof * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int st = pc_apply(c, c->d_work, c->d_work2);
  if (st) return st;
  PDES_HIP(hipMemcpyAsync(x, c->d_work2, c->numDof * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return pc_check(c);
}

int pdes_gmres(pdes_ctx* c, const double* b_host, double* x_host, double rtol, double atol, int restart,
               int maxit, int mode, int* iters, double* final_resnorm) {
  PDES_CHECK(c && b_host && x_host && restart > 0 && maxit >= 0, PDES_ERR_ARG, "bad argument");
  PDES_CHECK(mode == 1 || (c->jac && c->jac->assembled), PDES_ERR_ARG, "assemble the Jacobian for mode 0");
  PDES_HIP(hipMemcpyAsync(c->d_work, b_host, c->numDof * sizeof(double), hipMemcpyHostToDevice, c->stream));
  PDES_HIP(hipMemcpyAsync(c->d_work2, x_host, c->numDof * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int st = gmres_device(c, c->d_work, c->d_work2, rtol, atol, restart, maxit, mode, iters, final_resnorm, false);
  if (st) return st;
  PDES_HIP(hipMemcpyAsync(x_host, c->d_work2, c->numDof * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

// calcAdjoint (src/optimization/steady_adjoint.jl:30-66): J^T psi = -dJ/dq
// at the current state, with J assembled here, its blocks transposed in
// place for the solve (so the chosen preconditioner is built from J^T, as
// linearSolveTranspose applies the transposed preconditioner) and restored
// afterwards, with the preconditioner rebuilt from J.
int pdes_adjoint_solve(pdes_ctx* c, int kind, const int32_t* bcnums, int nbc, int pc_kind, int fp64_blocks,
                       double rtol, double atol, int restart, int maxit, double* psi, int* iters,
                       double* final_resnorm) {
  PDES_CHECK(c && psi && restart > 0 && maxit >= 0, PDES_ERR_ARG, "bad argument");
  const int64_t n = c->numDof;
  int st = jacobian_setup(c, fp64_blocks);
  if (st) return st;
  c->ptc_on = false;
  if ((st = assemble(c))) return st;
  // rhs = -dJ/dq in d_work, psi in d_work2 (x0 = 0)
  if ((st = functional_grad_device(c, kind, bcnums, nbc, c->d_work))) return st;
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, c->d_work, c->d_work, -1.0, 0, n);
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, c->d_work, c->d_work2, 0.0, 0, n);
  PDES_HIP(hipGetLastError());
  if ((st = transpose_jacobian(c))) return st;
  int gst = pc_setup(c, pc_kind);
  if (!gst) gst = gmres_device(c, c->d_work, c->d_work2, rtol, atol, restart, maxit, 0, iters, final_resnorm, true);
  // restore J and its preconditioner whatever the solve returned
  if ((st = transpose_jacobian(c))) return st;
  if ((st = pc_setup(c, pc_kind))) return st;
  if (gst) return gst;
  PDES_HIP(hipMemcpyAsync(psi, c->d_work2, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

int pdes_newton_solve(pdes_ctx* c, const pdes_newton_opts* o, pdes_newton_report* rep) {
  PDES_CHECK(c && o && rep, PDES_ERR_ARG, "null argument");
  PDES_CHECK(!o->globalize_euler || o->euler_tau > 0, PDES_ERR_ARG, "euler_tau must be positive");
  auto t0 = std::chrono::steady_clock::now();
  *rep = pdes_newton_report{};
  const int64_t n = c->numDof;
  int st;
  double* res = c->d_res;
  // every exit, the early returns of the HIP checks included, turns the
  // implicit-Euler term off and frees the workspace
  struct Guard {
    pdes_ctx* c;
    double* dq = nullptr;
    double* rhs = nullptr;
    ~Guard() {
      c->ptc_on = false;
      hipStreamSynchronize(c->stream);
      if (dq) hipFree(dq);
      if (rhs) hipFree(rhs);
    }
  } guard{c};
  PDES_CHECK(hipMalloc((void**)&guard.dq, n * sizeof(double)) == hipSuccess &&
             hipMalloc((void**)&guard.rhs, n * sizeof(double)) == hipSuccess, PDES_ERR_NOMEM, "Newton workspace");
  double* const dq = guard.dq;
  double* const rhs = guard.rhs;
  auto norm_res = [&](double* out) -> int {
    int s2 = launch_residual(c, c->d_q, nullptr, res, nullptr, c->phys.check_positivity != 0);
    if (!s2 && c->hom_lambda >= 0.0)
      s2 = launch_homotopy(c, c->d_q, nullptr, 1.0 - c->hom_lambda, c->hom_lambda, res, nullptr);
    if (s2) return s2;
    s2 = launch_wnorm2(c, res, c->d_Minv, n, c->d_scalar);
    if (s2) return s2;
    double v = 0;
    PDES_HIP(hipMemcpyAsync(&v, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PDES_HIP(hipStreamSynchronize(c->stream));
    s2 = sync_check_flag(c);
    if (s2) return s2;
    *out = std::sqrt(v);
    return PDES_OK;
  };
  double nrm = 0, tau = 0, prev_nrm = 0;
  if ((st = norm_res(&nrm)) != 0) goto fin;
  rep->res_norm0 = nrm;
  rep->res_norm = nrm;
  if (nrm < o->res_abstol) { rep->converged = 1; goto fin; }
  if ((st = jacobian_setup(c, o->fp64_blocks)) != 0) goto fin;
  for (int it = 1; it <= o->itermax; ++it) {
    if (o->globalize_euler) {
      // updateEuler at each Jacobian evaluation (globalization.jl:237-262)
      if (it == 1) {
        tau = o->euler_tau;
        if (o->scale_euler) tau *= 1.0 / nrm;
      } else {
        tau *= prev_nrm / nrm;
      }
      prev_nrm = nrm;
      if ((st = set_ptc_scale(c, &tau, 1)) != 0) goto fin;
    }
    // doRecalculation: Jacobian and preconditioner every iteration (recalc_freq 1)
    if ((st = c->hom_lambda >= 0.0 ? assemble_homotopy(c, c->hom_lambda) : assemble(c)) != 0) goto fin;
    if ((st = pdes_pc_setup(c, o->pc_kind)) != 0) goto fin;
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, res, rhs, -1.0, 0, n);
    int kits = 0;
    double kres = 0;
    if ((st = gmres_device(c, rhs, dq, o->krylov_reltol, o->krylov_abstol, o->gmres_restart, o->krylov_itermax,
                           o->jv_mode, &kits, &kres, true)) != 0)
      goto fin;
    rep->total_krylov_iters += kits;
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, dq, c->d_q, 1.0, 1, n);
    // step norm: norm(delta_q_vec)
    if ((st = launch_wnorm2(c, dq, nullptr, n, c->d_scalar + 1)) != 0) goto fin;
    double sn2 = 0;
    PDES_HIP(hipMemcpyAsync(&sn2, c->d_scalar + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if ((st = norm_res(&nrm)) != 0) goto fin;
    rep->step_norm = std::sqrt(sn2);
    rep->res_norm = nrm;
    rep->iterations = it;
    // checkConvergence (newton.jl:402-455)
    if (nrm < o->res_abstol || nrm / rep->res_norm0 < o->res_reltol || rep->step_norm <= o->step_tol) {
      rep->converged = 1;
      break;
    }
  }
fin:
  rep->elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

// predictorCorrectorHomotopy (src/NonlinearSolvers/predictor_corrector.jl:
// HomotopyData :63-148, the loop :226-326, updateTolerances :381-440,
// calcTangentVector :454-476, calcdHdLambda :495-530, takePredictorStep
// :535-595).  H(q, lambda) = (1 - lambda) R(q) + lambda D(q), lambda from 1 to
// 0: each predictor iteration runs a Newton corrector on H (tolerance
// homotopy_tol = 1e-2 relative while lambda > 0, res_reltol at lambda = 0,
// at most 30 iterations), then solves dH/dq t = dH/dlambda = D - R with the
// corrector's last Jacobian and preconditioner, normalises t with the SBP
// norm, sets the step h from the angle between successive tangents (limit
// psi_max) and the corrector's step length (limit 1), and takes
// lambda -= h, q += h t.  The loop ends on the physics residual:
// ||R|| <= res_abstol or ||R|| / ||R_0|| <= res_reltol, or itermax
// predictor iterations.
int pdes_homotopy_solve(pdes_ctx* c, const pdes_newton_opts* o, double psi_max, pdes_newton_report* rep) {
  PDES_CHECK(c && o && rep, PDES_ERR_ARG, "null argument");
  PDES_CHECK(psi_max > 0, PDES_ERR_ARG, "psi_max must be positive");
  PDES_CHECK(!o->globalize_euler, PDES_ERR_UNSUPPORTED, "homotopy with implicit-Euler globalization is not available");
  auto t0 = std::chrono::steady_clock::now();
  *rep = pdes_newton_report{};
  const int64_t n = c->numDof;
  const double eps = 2.220446049250313e-16;
  std::vector<double> M(n), t(n), t1(n), q(n), q0(n);
  // every exit, the early returns of the HIP checks included, leaves the
  // context solving the physics again and frees the workspace
  struct Guard {
    pdes_ctx* c;
    double* d_t = nullptr;
    double* d_rhs = nullptr;
    ~Guard() {
      c->hom_lambda = -1.0;
      hipStreamSynchronize(c->stream);
      if (d_t) hipFree(d_t);
      if (d_rhs) hipFree(d_rhs);
    }
  } guard{c};
  int st = PDES_OK;
  PDES_CHECK(hipMalloc((void**)&guard.d_t, n * sizeof(double)) == hipSuccess &&
             hipMalloc((void**)&guard.d_rhs, n * sizeof(double)) == hipSuccess, PDES_ERR_NOMEM, "homotopy workspace");
  double* const d_t = guard.d_t;
  double* const d_rhs = guard.d_rhs;
  auto phys_norm = [&](double* out) -> int {
    c->hom_lambda = -1.0;
    int s2 = launch_residual(c, c->d_q, nullptr, c->d_res, nullptr, c->phys.check_positivity != 0);
    if (!s2) s2 = launch_wnorm2(c, c->d_res, c->d_Minv, n, c->d_scalar);
    if (s2) return s2;
    double v = 0;
    PDES_HIP(hipMemcpyAsync(&v, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PDES_HIP(hipStreamSynchronize(c->stream));
    if ((s2 = sync_check_flag(c))) return s2;
    *out = std::sqrt(v);
    return PDES_OK;
  };
  // calcNorm (Utils.jl:427-449) and calcL2InnerProduct (:476-493): M-weighted
  auto mdot = [&](const std::vector<double>& a, const std::vector<double>& b) {
    double acc = 0.0;
    for (int64_t i = 0; i < n; ++i) acc += a[i] * M[i] * b[i];
    return acc;
  };
  double lambda = 1.0, h = 0.05, h_max = 0.1, delta_max = 1.0, hom_tol = 1e-2, tan_norm_1 = 0.0;
  double res_norm = 0.0;
  int iter = 1;
  lambda -= h;   // the Jacobian at lambda = 1 is ill-conditioned; start one step in
  {
    std::vector<double> minv(n);
    PDES_HIP(hipMemcpyAsync(minv.data(), c->d_Minv, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PDES_HIP(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; ++i) M[i] = 1.0 / minv[i];
  }
  if ((st = phys_norm(&res_norm)) != 0) goto done;
  rep->res_norm0 = res_norm;
  while (res_norm > rep->res_norm0 * o->res_reltol && res_norm > o->res_abstol && iter < o->itermax) {
    // updateTolerances
    if (std::fabs(lambda) <= eps) hom_tol = o->res_reltol;
    // corrector
    PDES_HIP(hipMemcpyAsync(q0.data(), c->d_q, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    c->hom_lambda = lambda;
    pdes_newton_opts no = *o;
    no.res_reltol = hom_tol;
    no.itermax = 30;
    pdes_newton_report nrep{};
    if ((st = pdes_newton_solve(c, &no, &nrep)) != 0) goto done;
    rep->total_krylov_iters += nrep.total_krylov_iters;
    if (std::fabs(lambda) > eps) {
      // tangent: dH/dq t = D(q) - R(q), with the corrector's last operator
      if ((st = launch_residual(c, c->d_q, nullptr, c->d_res, nullptr, false)) != 0) goto done;
      if ((st = launch_homotopy(c, c->d_q, nullptr, 0.0, 1.0, d_rhs, nullptr)) != 0) goto done;
      hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, c->stream, c->d_res, d_rhs, -1.0, 1, n);
      if (nrep.iterations == 0) {
        if ((st = jacobian_setup(c, o->fp64_blocks)) != 0) goto done;
        if ((st = assemble_homotopy(c, lambda)) != 0) goto done;
        if ((st = pdes_pc_setup(c, o->pc_kind)) != 0) goto done;
      }
      int kits = 0;
      double kres = 0;
      if ((st = gmres_device(c, d_rhs, d_t, o->krylov_reltol, o->krylov_abstol, o->gmres_restart, o->krylov_itermax,
                             o->jv_mode, &kits, &kres, true)) != 0)
        goto done;
      rep->total_krylov_iters += kits;
      PDES_HIP(hipMemcpyAsync(t.data(), d_t, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      PDES_HIP(hipMemcpyAsync(q.data(), c->d_q, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      PDES_HIP(hipStreamSynchronize(c->stream));
      // takePredictorStep
      double tan_norm = std::sqrt(mdot(t, t) + 1.0);
      for (int64_t i = 0; i < n; ++i) t[i] /= tan_norm;
      for (int64_t i = 0; i < n; ++i) q0[i] = q[i] - q0[i];
      const double delta = std::sqrt(mdot(q0, q0));
      double psi = psi_max;
      if (iter > 1) {
        const double arg = std::min(1.0, std::max(-1.0, mdot(t1, t) + 1.0 / (tan_norm * tan_norm_1)));
        psi = std::acos(arg);
      }
      const double fac = std::max(psi / psi_max, std::sqrt(delta / delta_max));
      h /= fac;
      h = std::min(h, h_max);
      lambda = std::max(0.0, lambda - h);
      if (lambda < eps) lambda = 0.0;
      for (int64_t i = 0; i < n; ++i) q[i] += h * t[i];
      PDES_HIP(hipMemcpyAsync(c->d_q, q.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
      t1.swap(t);
      tan_norm_1 = tan_norm;
    }
    if ((st = phys_norm(&res_norm)) != 0) goto done;
    rep->iterations = iter;
    ++iter;
  }
  rep->res_norm = res_norm;
  rep->converged = (res_norm <= o->res_abstol || res_norm <= rep->res_norm0 * o->res_reltol) ? 1 : 0;
done:
  rep->elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

}  // extern "C"
