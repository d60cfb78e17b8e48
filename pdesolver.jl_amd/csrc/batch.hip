// batch.hip -- many independent steady solves packed into few launches.
//
// Instances that share an SBP operator and physics kind are concatenated into
// one packed context ("group"): one residual / Jacobian / preconditioner
// launch covers every instance of the group, driven by an element work list
// that holds only the instances that have not converged.  Every reduction is
// segmented per instance over chunks aligned to the instance's first dof, and
// the GMRES Hessenberg, Givens rotations and Newton tests are kept per
// instance, so an instance's arithmetic does not depend on which other
// instances share its group: a solve gives the same bits alone and in a mix.
// Groups run on their own streams from their own host threads, so the work of
// different operators overlaps on the device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

#include "jacobian.hpp"

namespace pdes {
int pc_setup(pdes_ctx* c, int kind);
int pc_apply(pdes_ctx* c, const double* b, double* x);
int pc_check(pdes_ctx* c);
int ilu_range_levels(const pdes_ctx* c, int64_t el0, int64_t el1, int* lf, int* lb);
int ilu_hint(pdes_ctx* c, const int32_t* inst_sel, int ninst, int64_t nel, int lf, int lb);
void ilu_hint_clear(pdes_ctx* c);
void pc_graph_key(const pdes_ctx* c, int64_t key[6]);
}  // namespace pdes

namespace {

using namespace pdes;

constexpr int CHUNK = 4096;     // dofs per reduction chunk (aligned to the instance start)
constexpr int SEG_THREADS = 256;

// ---------------------------------------------------------------------------
// kernels over the active dofs: launch index t -> element el_list[t / dpe]
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t active_dof(const int32_t* el_list, int64_t t, int dpe) {
  const int64_t s = t / dpe;
  return (int64_t)el_list[s] * dpe + (t - s * dpe);
}

// partial[c][i] = sum over chunk c of V_i . w  (wt: optional weights, then
// sum of V_i^2 wt).  SEG_GROUP basis vectors per pass over the chunk, so w is
// read once per group; each vector's terms are summed in the same order, and
// reduced by the same tree, as one vector at a time would be.
constexpr int SEG_GROUP = 8;
__global__ void __launch_bounds__(SEG_THREADS)
seg_dot_partial(const double* __restrict__ V, int m, int64_t ldv, const double* __restrict__ w,
                const double* __restrict__ wt, const int64_t* __restrict__ cbeg, const int32_t* __restrict__ clen,
                const int32_t* __restrict__ clist, double* __restrict__ partial, int mstride,
                const int32_t* __restrict__ chunk_inst, const int32_t* __restrict__ mask) {
  __shared__ double sh[SEG_GROUP][SEG_THREADS];
  const int c = clist[blockIdx.x];
  if (mask && !mask[chunk_inst[c]]) return;
  const int64_t b0 = cbeg[c];
  const int len = clen[c];
  for (int i0 = 0; i0 < m; i0 += SEG_GROUP) {
    const int g_n = m - i0 < SEG_GROUP ? m - i0 : SEG_GROUP;
    double acc[SEG_GROUP];
#pragma unroll
    for (int g = 0; g < SEG_GROUP; ++g) acc[g] = 0.0;
    for (int t = threadIdx.x; t < len; t += SEG_THREADS) {
      const int64_t k = b0 + t;
      const double wk = wt ? wt[k] : w[k];
#pragma unroll
      for (int g = 0; g < SEG_GROUP; ++g) {
        if (g < g_n) {
          const double v = V[(int64_t)(i0 + g) * ldv + k];
          acc[g] += wt ? v * v * wk : v * wk;
        }
      }
    }
#pragma unroll
    for (int g = 0; g < SEG_GROUP; ++g) sh[g][threadIdx.x] = acc[g];
    __syncthreads();
    for (int s = SEG_THREADS / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) {
#pragma unroll
        for (int g = 0; g < SEG_GROUP; ++g) sh[g][threadIdx.x] += sh[g][threadIdx.x + s];
      }
      __syncthreads();
    }
    if ((int)threadIdx.x < g_n) partial[(int64_t)c * mstride + i0 + threadIdx.x] = sh[threadIdx.x][0];
    __syncthreads();
  }
}

// out[inst][i] = sum of the instance's chunk partials in chunk order
__global__ void seg_dot_reduce(const double* __restrict__ partial, int m, int mstride,
                               const int32_t* __restrict__ inst_c0, const int32_t* __restrict__ ilist,
                               double* __restrict__ out, const int32_t* __restrict__ mask) {
  const int inst = ilist[blockIdx.x];
  const int i = threadIdx.x;
  if (i >= m || (mask && !mask[inst])) return;
  double acc = 0.0;
  for (int c = inst_c0[inst]; c < inst_c0[inst + 1]; ++c) acc += partial[(int64_t)c * mstride + i];
  out[(int64_t)inst * mstride + i] = acc;
}

// w -= sum_i coef[inst][i] V_i   over the active dofs
__global__ void seg_multiaxpy(const double* __restrict__ V, int m, int64_t ldv, const double* __restrict__ coef,
                              int mstride, double* __restrict__ w, const int32_t* __restrict__ el_list,
                              const int32_t* __restrict__ el_inst, int dpe, int64_t n,
                              const int32_t* __restrict__ mask) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const int64_t k = active_dof(el_list, t, dpe);
    const int inst = el_inst[k / dpe];
    if (mask && !mask[inst]) continue;
    const double* cf = coef + (int64_t)inst * mstride;
    double acc = w[k];
    for (int i = 0; i < m; ++i) acc -= cf[i] * V[(int64_t)i * ldv + k];
    w[k] = acc;
  }
}

// y = sum_i coef[inst][i] V_i  over the active dofs
__global__ void seg_combine(const double* __restrict__ V, int m, int64_t ldv, const double* __restrict__ coef,
                            int mstride, double* __restrict__ y, const int32_t* __restrict__ el_list,
                            const int32_t* __restrict__ el_inst, int dpe, int64_t n,
                            const int32_t* __restrict__ mask) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const int64_t k = active_dof(el_list, t, dpe);
    const int inst = el_inst[k / dpe];
    if (mask && !mask[inst]) continue;
    const double* cf = coef + (int64_t)inst * mstride;
    double acc = 0.0;
    // basis vectors past an instance's own Krylov dimension are stale: skip them
    for (int i = 0; i < m; ++i)
      if (cf[i] != 0.0) acc += cf[i] * V[(int64_t)i * ldv + k];
    y[k] = acc;
  }
}

// y = a[inst] * x (+ y)   over the active dofs
__global__ void seg_scale(const double* __restrict__ x, double* __restrict__ y, const double* __restrict__ a,
                          int astride, int accumulate, const int32_t* __restrict__ el_list,
                          const int32_t* __restrict__ el_inst, int dpe, int64_t n,
                          const int32_t* __restrict__ mask) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const int64_t k = active_dof(el_list, t, dpe);
    const int inst = el_inst[k / dpe];
    if (mask && !mask[inst]) continue;
    const double s = a[(int64_t)inst * astride];
    y[k] = accumulate ? y[k] + s * x[k] : s * x[k];
  }
}

// ---------------------------------------------------------------------------
// device-resident GMRES control: one thread per listed instance
// ---------------------------------------------------------------------------
// per-instance GMRES scalars, [inst][GS_*]
enum { GS_BETA0 = 0, GS_RES = 1, GS_NSCAL = 2 };

// sums of the chunk partials -> coef[inst][i] (the CGS coefficients) and the
// Hessenberg column j (pass 0 stores, pass 1 adds)
__global__ void seg_reduce_hcol(const double* __restrict__ partial, int m, int mstride,
                                const int32_t* __restrict__ inst_c0, const int32_t* __restrict__ ilist,
                                const int32_t* __restrict__ mask, double* __restrict__ coef, double* __restrict__ H,
                                int R, int j, int pass) {
  const int inst = ilist[blockIdx.x];
  const int i = threadIdx.x;
  if (i >= m || !mask[inst]) return;
  double acc = 0.0;
  for (int c = inst_c0[inst]; c < inst_c0[inst + 1]; ++c) acc += partial[(int64_t)c * mstride + i];
  coef[(int64_t)inst * mstride + i] = acc;
  double* h = H + (int64_t)inst * (R + 1) * R + (int64_t)i * R + j;
  *h = pass ? *h + acc : acc;
}

// Arnoldi step j done: ||w|| is in seg[inst][0]; apply the previous Givens
// rotations to column j, form rotation j, update g and the residual estimate,
// and decide whether the instance keeps iterating.
__global__ void gmres_finish_step(int nlist, const int32_t* __restrict__ ilist, int32_t* __restrict__ iter,
                                  int32_t* __restrict__ live, int32_t* __restrict__ total, int32_t* __restrict__ jdone,
                                  const double* __restrict__ seg, double* __restrict__ coef, int mstride,
                                  double* __restrict__ H, double* __restrict__ cs, double* __restrict__ sn,
                                  double* __restrict__ g, double* __restrict__ gs, int R, int j, double rtol,
                                  double atol, int maxit) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlist) return;
  const int inst = ilist[t];
  if (!iter[inst]) return;
  double* Hi = H + (int64_t)inst * (R + 1) * R;
  double* ci = cs + (int64_t)inst * R;
  double* si = sn + (int64_t)inst * R;
  double* gi = g + (int64_t)inst * (R + 1);
  const double hn = sqrt(seg[(int64_t)inst * mstride]);
  Hi[(j + 1) * R + j] = hn;
  for (int k = 0; k < j; ++k) {
    const double a = Hi[k * R + j], b = Hi[(k + 1) * R + j];
    Hi[k * R + j] = ci[k] * a + si[k] * b;
    Hi[(k + 1) * R + j] = -si[k] * a + ci[k] * b;
  }
  const double a = Hi[j * R + j], b = Hi[(j + 1) * R + j];
  const double r = hypot(a, b);
  ci[j] = r > 0 ? a / r : 1.0;
  si[j] = r > 0 ? b / r : 0.0;
  Hi[j * R + j] = r;
  Hi[(j + 1) * R + j] = 0.0;
  gi[j + 1] = -si[j] * gi[j];
  gi[j] = ci[j] * gi[j];
  const double res = fabs(gi[j + 1]);
  gs[(int64_t)inst * GS_NSCAL + GS_RES] = res;
  total[inst] += 1;
  jdone[inst] = j + 1;
  const double tol = fmax(rtol * gs[(int64_t)inst * GS_NSCAL + GS_BETA0], atol);
  const bool conv = res <= tol || hn == 0.0;
  if (conv) live[inst] = 0;
  if (conv || total[inst] >= maxit || j + 1 == R) iter[inst] = 0;
  coef[(int64_t)inst * mstride] = hn > 0 ? 1.0 / hn : 0.0;
}

// end of a cycle: y = H^-1 g for the instances that took part -> coef[inst][0..R)
__global__ void gmres_solve_y(int nlist, const int32_t* __restrict__ ilist, const int32_t* __restrict__ part,
                              const int32_t* __restrict__ jdone, const double* __restrict__ H,
                              const double* __restrict__ g, double* __restrict__ coef, int mstride, int R) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlist) return;
  const int inst = ilist[t];
  double* y = coef + (int64_t)inst * mstride;
  for (int k = 0; k < mstride; ++k) y[k] = 0.0;
  if (!part[inst]) return;
  const double* Hi = H + (int64_t)inst * (R + 1) * R;
  const double* gi = g + (int64_t)inst * (R + 1);
  const int n = jdone[inst];
  for (int k = n - 1; k >= 0; --k) {
    double v = gi[k];
    for (int l = k + 1; l < n; ++l) v -= Hi[k * R + l] * y[l];
    y[k] = v / Hi[k * R + k];
  }
}

// (re)start of a cycle: beta = ||b - A x|| in seg[inst][0]; convergence and
// iteration-limit tests; g = beta e_1; coef[inst][0] = 1/beta
__global__ void gmres_restart(int nlist, const int32_t* __restrict__ ilist, int32_t* __restrict__ iter,
                              int32_t* __restrict__ live, int32_t* __restrict__ part, const int32_t* __restrict__ total,
                              int32_t* __restrict__ jdone, const double* __restrict__ seg, double* __restrict__ coef,
                              int mstride, double* __restrict__ g, double* __restrict__ gs, int R, double rtol,
                              double atol, int maxit) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlist) return;
  const int inst = ilist[t];
  jdone[inst] = 0;
  if (!live[inst]) { iter[inst] = 0; part[inst] = 0; return; }
  const double beta = sqrt(seg[(int64_t)inst * mstride]);
  double* gsi = gs + (int64_t)inst * GS_NSCAL;
  if (gsi[GS_BETA0] < 0) gsi[GS_BETA0] = beta;
  gsi[GS_RES] = beta;
  const double tol = fmax(rtol * gsi[GS_BETA0], atol);
  if (beta <= tol || total[inst] >= maxit) {
    live[inst] = 0;
    iter[inst] = 0;
    part[inst] = 0;
    return;
  }
  iter[inst] = 1;
  part[inst] = 1;
  double* gi = g + (int64_t)inst * (R + 1);
  gi[0] = beta;
  for (int k = 1; k <= R; ++k) gi[k] = 0.0;
  coef[(int64_t)inst * mstride] = 1.0 / beta;
}

__global__ void gmres_init(int nlist, const int32_t* __restrict__ ilist, int32_t* __restrict__ iter,
                           int32_t* __restrict__ live, int32_t* __restrict__ part, int32_t* __restrict__ total,
                           double* __restrict__ gs, double* __restrict__ coef, int mstride) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlist) return;
  const int inst = ilist[t];
  iter[inst] = 0;
  live[inst] = 1;
  part[inst] = 0;
  total[inst] = 0;
  gs[(int64_t)inst * GS_NSCAL + GS_BETA0] = -1.0;
  gs[(int64_t)inst * GS_NSCAL + GS_RES] = 0.0;
  coef[(int64_t)inst * mstride] = 0.0;
}

// y = a - b, y += x   over the active dofs of the instances with mask 1
__global__ void sub_masked(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ y,
                           const int32_t* __restrict__ el_list, const int32_t* __restrict__ el_inst, int dpe,
                           int64_t n, const int32_t* __restrict__ mask) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const int64_t k = active_dof(el_list, t, dpe);
    if (mask[el_inst[k / dpe]]) y[k] = a[k] - b[k];
  }
}

__global__ void add_masked(const double* __restrict__ x, double* __restrict__ y, const int32_t* __restrict__ el_list,
                           const int32_t* __restrict__ el_inst, int dpe, int64_t n, const int32_t* __restrict__ mask) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const int64_t k = active_dof(el_list, t, dpe);
    if (mask[el_inst[k / dpe]]) y[k] += x[k];
  }
}

unsigned grid_n(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(16384, ceil_div(n, 256))); }

struct InstState {
  int64_t el0 = 0, el1 = 0;  // element range in the packed mesh
  bool active = true;        // not yet converged (Newton)
  bool failed = false;
  pdes_newton_report rep{};
  // GMRES state
  double tau = 0, tau_prev_norm = 0;  // implicit-Euler pseudo-time step and the norm it was last updated at
  int nit = 0;               // Newton steps taken
  bool in_solve = false;     // its linear system is being solved (GMRES cycles)
  // set by pdes_batch_import: the instance resumes a solve taken over from
  // another batch (its state, Newton counters and, when in_solve, its GMRES
  // iterate, right-hand side, iteration count and initial residual)
  bool imported = false;
  int32_t g_total = 0, g_live = 1;
  double g_beta0 = -1;
  // ILU(0) dependency levels of the instance's elements (forward, backward; -1: not yet known)
  int ilu_lf = -1, ilu_lb = -1;
};

}  // namespace

struct pdes_batch_group {
  pdes_ctx* c = nullptr;
  std::vector<int> inst_global;      // global instance index of each local instance
  std::vector<InstState> inst;
  std::vector<int32_t> h_el_inst;
  int dpe = 0;                       // dofs per element
  // reduction chunks
  std::vector<int64_t> h_cbeg;
  std::vector<int32_t> h_clen, h_inst_c0;
  int64_t* d_cbeg = nullptr;
  int32_t* d_clen = nullptr;
  int32_t* d_inst_c0 = nullptr;
  int32_t* d_clist = nullptr;        // active chunk list
  int32_t* d_ilist = nullptr;        // active instance list
  double* d_partial = nullptr;       // [nchunk][mstride]
  double* d_seg = nullptr;           // [ninst][mstride]
  double* d_coef = nullptr;          // [ninst][mstride]
  double* h_seg = nullptr;           // pinned
  double* h_coef = nullptr;          // pinned
  int mstride = 0;
  // solver vectors
  double *d_V = nullptr, *d_z = nullptr, *d_w = nullptr, *d_dq = nullptr, *d_rhs = nullptr;
  int restart = 0;
  // device-resident GMRES state, per instance
  int32_t* d_chunk_inst = nullptr;   // instance of each reduction chunk
  int32_t *d_iter = nullptr, *d_live = nullptr, *d_part = nullptr, *d_total = nullptr, *d_jdone = nullptr;
  double *d_H = nullptr, *d_cs = nullptr, *d_sn = nullptr, *d_g = nullptr, *d_gs = nullptr;
  int32_t* h_flags = nullptr;        // pinned readback of d_iter / d_live / d_total
  std::vector<int32_t> h_active;     // host copy of the ctx's per-instance work-list flags
  // captured blocks of Arnoldi steps, replayed while their key matches
  struct ArnoldiGraph {
    std::vector<double> key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<ArnoldiGraph> agraphs;
  std::vector<std::vector<double>> akeys_seen;   // per block slot: the key of its last direct run
  int64_t graph_replays = 0, graph_captures = 0;
  int gmres_R = 0;                   // restart length the H/cs/sn/g buffers are sized for
  // incremental solve (pdes_batch_advance): started after the first pass,
  // finished when no instance has work left
  bool started = false, finished = false;
  std::chrono::steady_clock::time_point t0;
};

struct pdes_batch {
  int ninst = 0;
  std::vector<pdes_batch_group*> groups;
  std::vector<std::pair<int, int>> where;  // global instance -> (group, local)
};

namespace {

void group_free(pdes_batch_group* g) {
  if (!g) return;
  void* dev[] = {g->d_cbeg, g->d_clen, g->d_inst_c0, g->d_clist, g->d_ilist, g->d_partial, g->d_seg,
                 g->d_coef, g->d_V, g->d_z, g->d_w, g->d_dq, g->d_rhs, g->d_chunk_inst, g->d_iter,
                 g->d_live, g->d_part, g->d_total, g->d_jdone, g->d_H, g->d_cs, g->d_sn, g->d_g, g->d_gs};
  for (void* p : dev)
    if (p) hipFree(p);
  if (g->h_flags) hipHostFree(g->h_flags);
  if (g->h_seg) hipHostFree(g->h_seg);
  if (g->h_coef) hipHostFree(g->h_coef);
  for (auto& a : g->agraphs)
    if (a.exec) hipGraphExecDestroy(a.exec);
  if (g->c) pdes_ctx_destroy(g->c);
  delete g;
}

bool same_operator(const pdes_sbp_desc& a, const pdes_sbp_desc& b, int dim) {
  if (a.Np != b.Np || a.Nf != b.Nf || a.stencil != b.stencil || a.norient != b.norient || a.diag_e != b.diag_e)
    return false;
  const int nf = dim + 1;
  return !std::memcmp(a.w, b.w, a.Np * 8) && !std::memcmp(a.Q, b.Q, (size_t)dim * a.Np * a.Np * 8) &&
         !std::memcmp(a.interp, b.interp, (size_t)a.stencil * a.Nf * 8) &&
         !std::memcmp(a.perm, b.perm, (size_t)a.stencil * nf * 4) &&
         !std::memcmp(a.nbrperm, b.nbrperm, (size_t)a.Nf * a.norient * 4) && !std::memcmp(a.wface, b.wface, a.Nf * 8);
}

// set the element work list of a group to the instances selected by pick
template <class Pick>
int set_worklist(pdes_batch_group* g, Pick pick, std::vector<int>& chunks, std::vector<int>& insts) {
  std::vector<int32_t> el;
  chunks.clear();
  insts.clear();
  for (size_t i = 0; i < g->inst.size(); ++i) {
    if (!pick(g->inst[i])) continue;
    insts.push_back((int)i);
    for (int64_t e = g->inst[i].el0; e < g->inst[i].el1; ++e) el.push_back((int32_t)e);
    for (int c = g->h_inst_c0[i]; c < g->h_inst_c0[i + 1]; ++c) chunks.push_back(c);
  }
  pdes_ctx* c = g->c;
  c->nlist = (int32_t)el.size();
  for (size_t i = 0; i < g->inst.size(); ++i) g->h_active[i] = 0;
  for (int i : insts) g->h_active[i] = 1;
  PDES_HIP(hipMemcpyAsync(c->d_inst_active, g->h_active.data(), g->inst.size() * 4, hipMemcpyHostToDevice, c->stream));
  if (!el.empty()) PDES_HIP(hipMemcpyAsync(c->d_el_list, el.data(), el.size() * 4, hipMemcpyHostToDevice, c->stream));
  if (!chunks.empty())
    PDES_HIP(hipMemcpyAsync(g->d_clist, chunks.data(), chunks.size() * 4, hipMemcpyHostToDevice, c->stream));
  if (!insts.empty())
    PDES_HIP(hipMemcpyAsync(g->d_ilist, insts.data(), insts.size() * 4, hipMemcpyHostToDevice, c->stream));
  // host vectors are reused by the caller until the next call; wait for the copies
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

// seg[inst][0..m) = V_i . w  (or weighted squares of V_0 when wt) for the listed instances
int seg_dots(pdes_batch_group* g, const double* V, int m, int64_t ldv, const double* w, const double* wt,
             int nchunks, int ninst_list, const int32_t* mask = nullptr) {
  pdes_ctx* c = g->c;
  if (nchunks == 0) return PDES_OK;
  hipLaunchKernelGGL(seg_dot_partial, dim3(nchunks), dim3(SEG_THREADS), 0, c->stream, V, m, ldv, w, wt, g->d_cbeg,
                     g->d_clen, g->d_clist, g->d_partial, g->mstride, g->d_chunk_inst, mask);
  hipLaunchKernelGGL(seg_dot_reduce, dim3(ninst_list), dim3(64), 0, c->stream, g->d_partial, m, g->mstride,
                     g->d_inst_c0, g->d_ilist, g->d_seg, mask);
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

int fetch_seg(pdes_batch_group* g) {
  pdes_ctx* c = g->c;
  PDES_HIP(hipMemcpyAsync(g->h_seg, g->d_seg, g->inst.size() * g->mstride * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

// The host rewrites h_coef right after each upload, so the upload must have
// read the pinned buffer before returning: copy, then wait on the stream.
int push_coef(pdes_batch_group* g) {
  pdes_ctx* c = g->c;
  PDES_HIP(hipMemcpyAsync(g->d_coef, g->h_coef, g->inst.size() * g->mstride * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

int64_t active_dofs(pdes_batch_group* g) { return (int64_t)g->c->nlist * g->dpe; }

int apply_op(pdes_batch_group* g, int mode, const double* x, double* y) {
  pdes_ctx* c = g->c;
  if (mode == 0) return launch_spmv(c, x, y, 1.0, nullptr, 0.0);
  const int st = launch_residual(c, c->d_q, x, nullptr, y, false);
  return st ? st : launch_ptc_shift(c, x, y);
}

// apply the operator with the preconditioner/product mask set to `mask`
int apply_op_masked(pdes_batch_group* g, int mode, const double* x, double* y, const int32_t* mask) {
  g->c->d_mask = mask;
  const int st = apply_op(g, mode, x, y);
  g->c->d_mask = nullptr;
  return st;
}

int pc_apply_masked(pdes_batch_group* g, const double* b, double* x, const int32_t* mask) {
  g->c->d_mask = mask;
  const int st = pc_apply(g->c, b, x);
  g->c->d_mask = nullptr;
  return st;
}

// ||R_i||_{M^-1} for the active instances (work list must select them); writes rep.res_norm
int batch_residual_norms(pdes_batch_group* g, std::vector<int>& chunks, std::vector<int>& insts, bool first) {
  pdes_ctx* c = g->c;
  int st = launch_residual(c, c->d_q, nullptr, c->d_res, nullptr, c->phys.check_positivity != 0);
  if (st) return st;
  // weighted squares: seg_dot_partial with wt uses V_0 = res
  if ((st = seg_dots(g, c->d_res, 1, 0, nullptr, c->d_Minv, (int)chunks.size(), (int)insts.size()))) return st;
  if ((st = fetch_seg(g))) return st;
  std::vector<int32_t> flags(g->inst.size(), 0);
  PDES_HIP(hipMemcpyAsync(flags.data(), c->d_flag, flags.size() * 4, hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  for (int i : insts) {
    InstState& s = g->inst[i];
    if (flags[i]) {
      s.failed = true;
      s.active = false;
      continue;
    }
    const double nrm = std::sqrt(g->h_seg[(int64_t)i * g->mstride]);
    if (first) s.rep.res_norm0 = nrm;
    s.rep.res_norm = nrm;
  }
  return PDES_OK;
}

// Newton for the instances of one group, continuously batched: each
// instance advances its own Newton steps.  Every pass of the loop (1) starts
// a Newton step (Jacobian, preconditioner, right-hand side) for the
// instances that need one, (2) begins a GMRES(m) cycle for every instance
// with a linear system in progress (V_0 = b - A x; an instance that has just
// started has x = 0, so V_0 = b), (3) takes the Newton update of the
// instances whose linear solve has ended there (converged or out of
// iterations) and tests their convergence, and (4) runs the cycle's Arnoldi
// steps for the others.  No instance waits for another's linear solve, and
// each one's arithmetic is the same as alone: every kernel is masked or
// listed per instance and every reduction is instance-aligned.
int group_finish(pdes_batch_group* g);

double group_elapsed(const pdes_batch_group* g) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - g->t0).count();
}

// start of a solve: residual norms of every instance (fresh instances take
// theirs as res_norm0; imported ones keep their counters), the BSR pattern,
// and, for imported instances in the middle of a linear solve, their GMRES
// iterate, right-hand side and counters plus the Jacobian and preconditioner
// of their current Newton step (both functions of the state only, so the
// bits are those of the batch they came from)
int group_init(pdes_batch_group* g, const pdes_newton_opts* o) {
  pdes_ctx* c = g->c;
  g->t0 = std::chrono::steady_clock::now();
  std::vector<int> chunks, insts;
  int st;
  hipStream_t sm = c->stream;
  PDES_HIP(hipMemsetAsync(c->d_flag, 0, g->inst.size() * 4, sm));
  for (auto& s : g->inst)
    if (!s.imported) {
      s.active = true;
      s.failed = false;
      s.rep = pdes_newton_report{};
      s.tau = 0;
      s.nit = 0;
      s.in_solve = false;
    }
  for (int pass = 0; pass < 2; ++pass) {
    const bool fresh = pass == 0;
    if ((st = set_worklist(g, [&](const InstState& s) { return s.active && s.imported != fresh; }, chunks, insts)))
      return st;
    if (insts.empty()) continue;
    if ((st = batch_residual_norms(g, chunks, insts, fresh))) return st;
  }
  for (auto& s : g->inst)
    if (!s.imported && s.active && s.rep.res_norm < o->res_abstol) {
      s.active = false;
      s.rep.converged = 1;
      s.rep.elapsed_s = group_elapsed(g);
    }
  if ((st = jacobian_setup(c, o->fp64_blocks))) return st;
  if ((st = set_worklist(g, [](const InstState& s) { return s.imported && s.active && s.in_solve; }, chunks, insts)))
    return st;
  if (!insts.empty()) {
    const int ms = g->mstride;
    std::vector<int32_t> tot(g->inst.size(), 0), live(g->inst.size(), 0);
    std::vector<double> gs(g->inst.size() * GS_NSCAL, 0.0);
    for (int i : insts) {
      tot[i] = g->inst[i].g_total;
      live[i] = g->inst[i].g_live;
      gs[(size_t)i * GS_NSCAL + GS_BETA0] = g->inst[i].g_beta0;
    }
    PDES_HIP(hipMemcpyAsync(g->d_total, tot.data(), tot.size() * 4, hipMemcpyHostToDevice, sm));
    PDES_HIP(hipMemcpyAsync(g->d_live, live.data(), live.size() * 4, hipMemcpyHostToDevice, sm));
    PDES_HIP(hipMemcpyAsync(g->d_gs, gs.data(), gs.size() * 8, hipMemcpyHostToDevice, sm));
    if (o->globalize_euler) {
      std::vector<double> taus(g->inst.size(), 1.0);
      for (size_t i = 0; i < g->inst.size(); ++i)
        if (g->inst[i].tau > 0) taus[i] = g->inst[i].tau;
      if ((st = set_ptc_scale(c, taus.data(), (int)taus.size()))) return st;
    }
    if ((st = assemble(c))) return st;
    if ((st = pc_setup(c, o->pc_kind))) return st;
    (void)ms;
    PDES_HIP(hipStreamSynchronize(sm));
  }
  for (auto& s : g->inst) s.imported = false;
  g->started = true;
  g->finished = false;
  return PDES_OK;
}

// One pass of the continuously batched Newton loop of a group: each
// instance advances its own Newton steps.  A pass (1) starts a Newton step
// (Jacobian, preconditioner, right-hand side) for the instances that need
// one, (2) begins a GMRES(m) cycle for every instance with a linear system
// in progress (V_0 = b - A x; an instance that has just started has x = 0,
// so V_0 = b), (3) takes the Newton update of the instances whose linear
// solve has ended there (converged or out of iterations) and tests their
// convergence, and (4) runs the cycle's Arnoldi steps for the others.  No
// instance waits for another's linear solve, and each one's arithmetic is
// the same as alone: every kernel is masked or listed per instance and every
// reduction is instance-aligned.  Between passes every instance in a linear
// solve sits at a GMRES restart, whose whole state is its iterate, its
// right-hand side, its iteration count and its initial residual.
int group_pass(pdes_batch_group* g, const pdes_newton_opts* o) {
  pdes_ctx* c = g->c;
  auto elapsed = [&]() { return group_elapsed(g); };
  std::vector<int> chunks, insts;
  int st;
  const int R = o->gmres_restart;
  const int64_t ldv = c->numDof;
  const int ms = g->mstride;
  hipStream_t sm = c->stream;
  {
    // (1) start a Newton step where one is due
    for (auto& s : g->inst)
      if (s.active && !s.in_solve && s.nit >= o->itermax) {
        s.active = false;   // out of Newton iterations, not converged
        s.rep.elapsed_s = elapsed();
      }
    if ((st = set_worklist(g, [](const InstState& s) { return s.active && !s.in_solve; }, chunks, insts))) return st;
    if (!insts.empty()) {
      if (o->globalize_euler) {
        // per-instance updateEuler (globalization.jl:237-262)
        std::vector<double> taus(g->inst.size(), 1.0);
        for (int i : insts) {
          InstState& s = g->inst[i];
          if (s.nit == 0)
            s.tau = o->euler_tau * (o->scale_euler ? 1.0 / s.rep.res_norm : 1.0);
          else
            s.tau *= s.tau_prev_norm / s.rep.res_norm;
          s.tau_prev_norm = s.rep.res_norm;
        }
        for (size_t i = 0; i < g->inst.size(); ++i)
          if (g->inst[i].tau > 0) taus[i] = g->inst[i].tau;
        if ((st = set_ptc_scale(c, taus.data(), (int)taus.size()))) return st;
      }
      for (int i : insts) {
        g->inst[i].nit += 1;
        g->inst[i].rep.iterations = g->inst[i].nit;
        g->inst[i].in_solve = true;
      }
      if ((st = assemble(c))) return st;
      if ((st = pc_setup(c, o->pc_kind))) return st;
      const int64_t n = active_dofs(g);
      // b = -R(q) (the residual of these instances is current in d_res), x = 0
      for (int i : insts) g->h_coef[(int64_t)i * ms] = -1.0;
      if ((st = push_coef(g))) return st;
      hipLaunchKernelGGL(seg_scale, dim3(grid_n(n)), dim3(256), 0, sm, c->d_res, g->d_rhs, g->d_coef, ms, 0,
                         c->d_el_list, c->d_el_inst, g->dpe, n, (const int32_t*)nullptr);
      hipLaunchKernelGGL(gmres_init, dim3((unsigned)ceil_div((int64_t)insts.size(), 64)), dim3(64), 0, sm,
                         (int)insts.size(), g->d_ilist, g->d_iter, g->d_live, g->d_part, g->d_total, g->d_gs,
                         g->d_coef, ms);
      hipLaunchKernelGGL(seg_scale, dim3(grid_n(n)), dim3(256), 0, sm, g->d_rhs, g->d_dq, g->d_coef, ms, 0,
                         c->d_el_list, c->d_el_inst, g->dpe, n, (const int32_t*)nullptr);
      PDES_HIP(hipGetLastError());
    }
    // (2) begin a GMRES cycle for every linear solve in progress
    if ((st = set_worklist(g, [](const InstState& s) { return s.in_solve; }, chunks, insts))) return st;
    if (insts.empty()) return group_finish(g);
    {
      const int64_t n = active_dofs(g);
      if ((st = apply_op_masked(g, o->jv_mode, g->d_dq, g->d_w, g->d_live))) return st;
      hipLaunchKernelGGL(sub_masked, dim3(grid_n(n)), dim3(256), 0, sm, g->d_rhs, g->d_w, g->d_V, c->d_el_list,
                         c->d_el_inst, g->dpe, n, g->d_live);
      if ((st = seg_dots(g, g->d_V, 1, ldv, g->d_V, nullptr, (int)chunks.size(), (int)insts.size(), g->d_live)))
        return st;
      hipLaunchKernelGGL(gmres_restart, dim3((unsigned)ceil_div((int64_t)insts.size(), 64)), dim3(64), 0, sm,
                         (int)insts.size(), g->d_ilist, g->d_iter, g->d_live, g->d_part, g->d_total, g->d_jdone,
                         g->d_seg, g->d_coef, ms, g->d_g, g->d_gs, R, o->krylov_reltol, o->krylov_abstol,
                         o->krylov_itermax);
      // V_0 /= beta for the instances that iterate in this cycle (before the
      // Newton updates below reuse the coefficient buffer)
      hipLaunchKernelGGL(seg_scale, dim3(grid_n(n)), dim3(256), 0, sm, g->d_V, g->d_V, g->d_coef, ms, 0,
                         c->d_el_list, c->d_el_inst, g->dpe, n, g->d_iter);
      PDES_HIP(hipMemcpyAsync(g->h_flags, g->d_iter, g->inst.size() * 4, hipMemcpyDeviceToHost, sm));
      PDES_HIP(hipStreamSynchronize(sm));
      if ((st = pc_check(c))) return st;
    }
    // (3) Newton update of the instances whose linear solve has ended
    std::vector<char> ended(g->inst.size(), 0);
    bool any_end = false, any_iter = false;
    for (int i : insts) {
      if (g->h_flags[i]) any_iter = true;
      else { ended[i] = 1; any_end = true; }
    }
    if (any_end) {
      std::vector<int32_t> tot(g->inst.size(), 0);
      PDES_HIP(hipMemcpyAsync(tot.data(), g->d_total, g->inst.size() * 4, hipMemcpyDeviceToHost, sm));
      PDES_HIP(hipStreamSynchronize(sm));
      std::vector<int> fin_chunks, fin;
      if ((st = set_worklist(g, [&](const InstState& s) { return ended[&s - g->inst.data()] != 0; }, fin_chunks, fin)))
        return st;
      for (int i : fin) {
        g->inst[i].rep.total_krylov_iters += tot[i];
        g->inst[i].in_solve = false;
        g->h_coef[(int64_t)i * ms] = 1.0;
      }
      const int64_t n = active_dofs(g);
      if ((st = push_coef(g))) return st;
      // q += dq, step norms, new residual (and its norms) of these instances
      hipLaunchKernelGGL(seg_scale, dim3(grid_n(n)), dim3(256), 0, sm, g->d_dq, c->d_q, g->d_coef, ms, 1,
                         c->d_el_list, c->d_el_inst, g->dpe, n, (const int32_t*)nullptr);
      if ((st = seg_dots(g, g->d_dq, 1, 0, g->d_dq, nullptr, (int)fin_chunks.size(), (int)fin.size()))) return st;
      if ((st = fetch_seg(g))) return st;
      for (int i : fin) g->inst[i].rep.step_norm = std::sqrt(g->h_seg[(int64_t)i * ms]);
      if ((st = batch_residual_norms(g, fin_chunks, fin, false))) return st;
      for (int i : fin) {
        InstState& s = g->inst[i];
        if (s.failed) { s.rep.elapsed_s = elapsed(); continue; }
        if (s.rep.res_norm < o->res_abstol || s.rep.res_norm / s.rep.res_norm0 < o->res_reltol ||
            s.rep.step_norm <= o->step_tol) {
          s.rep.converged = 1;
          s.active = false;
          s.rep.elapsed_s = elapsed();
        }
      }
      if (!any_iter) return PDES_OK;
      // back to the solves still iterating in this cycle
      if ((st = set_worklist(g, [](const InstState& s) { return s.in_solve; }, chunks, insts))) return st;
    }
    if (!any_iter) return PDES_OK;
    // (4) the Arnoldi steps of this cycle and the update x += M^-1 V y
    const int64_t n = active_dofs(g);
    const int ni = (int)insts.size(), nch = (int)chunks.size();
    const unsigned gi = (unsigned)ceil_div(ni, 64);
    // the preconditioner's sweeps touch only these instances' elements
    struct HintGuard {
      pdes_ctx* c;
      ~HintGuard() { ilu_hint_clear(c); }
    } hint_guard{c};
    if (o->pc_kind == PDES_PC_ILU0) {
      int64_t hel = 0;
      int hlf = 0, hlb = 0;
      std::vector<int32_t> sel(g->inst.size(), 0);
      for (int i : insts) {
        InstState& s = g->inst[i];
        if (s.ilu_lf < 0 && ilu_range_levels(c, s.el0, s.el1, &s.ilu_lf, &s.ilu_lb)) s.ilu_lf = s.ilu_lb = -1;
        hel += s.el1 - s.el0;
        hlf = std::max(hlf, s.ilu_lf);
        hlb = std::max(hlb, s.ilu_lb);
        sel[i] = 1;
      }
      if (hlf > 0 && (st = ilu_hint(c, sel.data(), (int)sel.size(), hel, hlf, hlb))) return st;
    }
    // one Arnoldi step: launches only (no host synchronization, no memset),
    // so a block of steps can be captured into a graph
    auto step = [&](int j) -> int {
      double* Vj = g->d_V + (int64_t)j * ldv;
      double* Vn = g->d_V + (int64_t)(j + 1) * ldv;
      int sst;
      if ((sst = pc_apply_masked(g, Vj, g->d_z, g->d_iter))) return sst;
      if ((sst = apply_op_masked(g, o->jv_mode, g->d_z, Vn, g->d_iter))) return sst;
      // CGS2: two passes of h = V^T w, w -= V h; H column = sum of both passes
      for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(seg_dot_partial, dim3(nch), dim3(SEG_THREADS), 0, sm, g->d_V, j + 1, ldv, Vn,
                           (const double*)nullptr, g->d_cbeg, g->d_clen, g->d_clist, g->d_partial, ms,
                           g->d_chunk_inst, g->d_iter);
        hipLaunchKernelGGL(seg_reduce_hcol, dim3(ni), dim3(64), 0, sm, g->d_partial, j + 1, ms, g->d_inst_c0,
                           g->d_ilist, g->d_iter, g->d_coef, g->d_H, R, j, pass);
        hipLaunchKernelGGL(seg_multiaxpy, dim3(grid_n(n)), dim3(256), 0, sm, g->d_V, j + 1, ldv, g->d_coef, ms, Vn,
                           c->d_el_list, c->d_el_inst, g->dpe, n, g->d_iter);
      }
      if ((sst = seg_dots(g, Vn, 1, ldv, Vn, nullptr, nch, ni, g->d_iter))) return sst;
      hipLaunchKernelGGL(gmres_finish_step, dim3(gi), dim3(64), 0, sm, ni, g->d_ilist, g->d_iter, g->d_live,
                         g->d_total, g->d_jdone, g->d_seg, g->d_coef, ms, g->d_H, g->d_cs, g->d_sn, g->d_g, g->d_gs,
                         R, j, o->krylov_reltol, o->krylov_abstol, o->krylov_itermax);
      // instances that stopped in gmres_finish_step are masked out of the scale
      hipLaunchKernelGGL(seg_scale, dim3(grid_n(n)), dim3(256), 0, sm, Vn, Vn, g->d_coef, ms, 0, c->d_el_list,
                         c->d_el_inst, g->dpe, n, g->d_iter);
      PDES_HIP(hipGetLastError());
      return PDES_OK;
    };
    // Blocks of 8 steps between flag read-backs.  A block is captured once
    // into a HIP graph and replayed in later cycles while everything its
    // launches depend on besides device data is unchanged: the steps, the
    // work-list sizes (launch sizes), the operator and preconditioner
    // configuration and the solver options (the work lists, masks and vectors
    // live at fixed addresses).
    for (int j0 = 0; j0 < R; j0 += 8) {
      const int j1 = std::min(R, j0 + 8);
      bool done_block = false;
      if (c->use_graphs) {
        int64_t pk[6];
        pc_graph_key(c, pk);
        std::vector<double> key = {(double)j0, (double)j1, (double)R, (double)c->nlist, (double)nch, (double)ni,
                                   (double)o->jv_mode, (double)c->ptc_on, o->krylov_reltol, o->krylov_abstol,
                                   (double)o->krylov_itermax};
        for (int k = 0; k < 6; ++k) key.push_back((double)pk[k]);
        hipGraphExec_t exec = nullptr;
        for (auto& a : g->agraphs)
          if (a.key == key) exec = a.exec;
        // capture only a block whose key repeats from its last direct run:
        // while the work lists change every cycle, capturing would cost more
        // than the replays save
        const size_t slot = (size_t)(j0 / 8);
        if (g->akeys_seen.size() <= slot) g->akeys_seen.resize(slot + 1);
        const bool repeat = g->akeys_seen[slot] == key;
        if (!exec && !repeat) g->akeys_seen[slot] = key;
        if (!exec && repeat) {
          // the slot of these steps takes the new graph
          for (auto it = g->agraphs.begin(); it != g->agraphs.end(); ++it)
            if (it->key[0] == (double)j0) {
              hipGraphExecDestroy(it->exec);
              g->agraphs.erase(it);
              break;
            }
          hipGraph_t graph = nullptr;
          if (hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            c->capturing = true;
            int cst = PDES_OK;
            for (int j = j0; j < j1 && !cst; ++j) cst = step(j);
            c->capturing = false;
            const hipError_t ec = hipStreamEndCapture(sm, &graph);
            // a capture that failed runs the block directly (below), which
            // reports the error if it was not the capture's
            if (!cst && ec == hipSuccess && graph &&
                hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
              g->agraphs.push_back({key, exec});
              ++g->graph_captures;
            }
            else
              exec = nullptr;
            if (graph) hipGraphDestroy(graph);
          }
          (void)hipGetLastError();
        }
        if (exec) {
          PDES_HIP(hipGraphLaunch(exec, sm));
          ++g->graph_replays;
          done_block = true;
        }
      }
      if (!done_block)
        for (int j = j0; j < j1; ++j)
          if ((st = step(j))) return st;
      if (j1 < R) {
        PDES_HIP(hipMemcpyAsync(g->h_flags, g->d_iter, g->inst.size() * 4, hipMemcpyDeviceToHost, sm));
        PDES_HIP(hipStreamSynchronize(sm));
        bool still = false;
        for (int i : insts) still = still || g->h_flags[i];
        if (!still) break;
      }
    }
    hipLaunchKernelGGL(gmres_solve_y, dim3(gi), dim3(64), 0, sm, ni, g->d_ilist, g->d_part, g->d_jdone, g->d_H,
                       g->d_g, g->d_coef, ms, R);
    hipLaunchKernelGGL(seg_combine, dim3(grid_n(n)), dim3(256), 0, sm, g->d_V, R, ldv, g->d_coef, ms, g->d_w,
                       c->d_el_list, c->d_el_inst, g->dpe, n, g->d_part);
    if ((st = pc_apply_masked(g, g->d_w, g->d_z, g->d_part))) return st;
    hipLaunchKernelGGL(add_masked, dim3(grid_n(n)), dim3(256), 0, sm, g->d_z, g->d_dq, c->d_el_list, c->d_el_inst,
                       g->dpe, n, g->d_part);
    PDES_HIP(hipGetLastError());

  }
  return PDES_OK;
}

int group_finish(pdes_batch_group* g) {
  const double el = group_elapsed(g);
  for (auto& s : g->inst)
    if (s.rep.converged != 1 && s.rep.elapsed_s == 0) s.rep.elapsed_s = el;
  g->c->nlist = -1;
  g->c->ptc_on = false;
  g->finished = true;
  g->started = false;
  return PDES_OK;
}

// up to max_passes passes of a group (all of them when max_passes <= 0)
int group_newton(pdes_batch_group* g, const pdes_newton_opts* o, int max_passes) {
  int st;
  if (g->finished) return PDES_OK;
  if (!g->started && (st = group_init(g, o))) return st;
  for (int k = 0; max_passes <= 0 || k < max_passes; ++k) {
    if ((st = group_pass(g, o))) return st;
    if (g->finished) break;
  }
  return PDES_OK;
}

}  // namespace

using namespace pdes;

extern "C" {

int pdes_batch_create(int ninst, const pdes_mesh_desc* meshes, const pdes_sbp_desc* sbps,
                      const pdes_phys_desc* phys, pdes_batch** out) {
  PDES_CHECK(ninst > 0 && meshes && sbps && phys && out, PDES_ERR_ARG, "bad argument");
  pdes_batch* B = new pdes_batch();
  B->ninst = ninst;
  B->where.assign(ninst, {-1, -1});
  // group instances by operator and physics kind
  std::vector<std::vector<int>> groups;
  for (int i = 0; i < ninst; ++i) {
    bool placed = false;
    for (auto& gr : groups) {
      const int r = gr[0];
      if (meshes[r].dim == meshes[i].dim && phys[r].volume_integral_type == phys[i].volume_integral_type &&
          phys[r].face_flux == phys[i].face_flux && phys[r].gamma == phys[i].gamma &&
          phys[r].R_gas == phys[i].R_gas && phys[r].check_positivity == phys[i].check_positivity &&
          same_operator(sbps[r], sbps[i], meshes[i].dim)) {
        gr.push_back(i);
        placed = true;
        break;
      }
    }
    if (!placed) groups.push_back({i});
  }
  int st = PDES_OK;
  for (size_t gi = 0; gi < groups.size() && !st; ++gi) {
    const auto& ids = groups[gi];
    const pdes_mesh_desc& m0 = meshes[ids[0]];
    const pdes_sbp_desc& s0 = sbps[ids[0]];
    const int dim = m0.dim, Np = s0.Np, Nf = s0.Nf;
    // pack
    std::vector<double> dxidx, jac, nrm_face, nrm_bndry, coords_bndry;
    std::vector<int32_t> ifaces, bfaces, boff{0}, bct;
    std::vector<int64_t> el_start{0};
    int64_t nel = 0, nif = 0, nbf = 0;
    for (int id : ids) {
      const pdes_mesh_desc& m = meshes[id];
      dxidx.insert(dxidx.end(), m.dxidx, m.dxidx + (size_t)m.numEl * Np * dim * dim);
      jac.insert(jac.end(), m.jac, m.jac + (size_t)m.numEl * Np);
      for (int k = 0; k < m.numInterfaces; ++k) {
        const int32_t* it = m.interfaces + 5 * k;
        ifaces.insert(ifaces.end(), {(int32_t)(it[0] + nel), (int32_t)(it[1] + nel), it[2], it[3], it[4]});
      }
      nrm_face.insert(nrm_face.end(), m.nrm_face, m.nrm_face + (size_t)m.numInterfaces * Nf * dim);
      for (int k = 0; k < m.numBoundaryFaces; ++k)
        bfaces.insert(bfaces.end(), {(int32_t)(m.bndryfaces[2 * k] + nel), m.bndryfaces[2 * k + 1]});
      nrm_bndry.insert(nrm_bndry.end(), m.nrm_bndry, m.nrm_bndry + (size_t)m.numBoundaryFaces * Nf * dim);
      coords_bndry.insert(coords_bndry.end(), m.coords_bndry, m.coords_bndry + (size_t)m.numBoundaryFaces * Nf * dim);
      for (int bc = 0; bc < m.numBC; ++bc) {
        boff.push_back((int32_t)(m.bndry_offsets[bc + 1] + nbf));
        bct.push_back(m.bc_type[bc]);
      }
      nel += m.numEl;
      nif += m.numInterfaces;
      nbf += m.numBoundaryFaces;
      el_start.push_back(nel);
    }
    PDES_CHECK(nel < (int64_t)1 << 31, PDES_ERR_ARG, "packed group too large");
    pdes_mesh_desc pm{dim, (int32_t)nel, (int32_t)nif, (int32_t)nbf, (int32_t)bct.size(), dxidx.data(), jac.data(),
                      ifaces.data(), bfaces.data(), boff.data(), bct.data(), nrm_face.data(), nrm_bndry.data(),
                      coords_bndry.data()};
    pdes_batch_group* g = new pdes_batch_group();
    B->groups.push_back(g);
    st = pdes_ctx_create(&pm, &s0, &phys[ids[0]], &g->c);
    if (st) break;
    pdes_ctx* c = g->c;
    const int ni = (int)ids.size();
    c->ninst = ni;
    g->dpe = Np * c->nd;
    g->inst_global = ids;
    g->inst.resize(ni);
    g->h_el_inst.resize(nel);
    for (int i = 0; i < ni; ++i) {
      g->inst[i].el0 = el_start[i];
      g->inst[i].el1 = el_start[i + 1];
      for (int64_t e = el_start[i]; e < el_start[i + 1]; ++e) g->h_el_inst[e] = i;
      B->where[ids[i]] = {(int)gi, i};
    }
    // per-instance physics and boundary records tagged with the instance
    std::vector<PhysConst> ip(ni);
    for (int i = 0; i < ni; ++i) {
      const pdes_phys_desc& p = phys[ids[i]];
      ip[i] = PhysConst{p.gamma, p.gamma - 1, p.R_gas, p.R_gas / (p.gamma - 1), p.Ma, p.aoa, p.rho_free,
                        p.p_free / (p.gamma - 1) + 0.5 * p.rho_free * p.Ma * p.Ma, p.vortex_Min};
    }
    for (int64_t e = 0; e < nel; ++e)
      for (int f = 0; f < c->nfaces; ++f) {
        ElFace& r = c->h_elface[(size_t)e * c->nfaces + f];
        if (r.x < 0) r.w = (r.w & 0xff) | (g->h_el_inst[e] << 8);
      }
    // reduction chunks aligned to each instance's first dof
    g->h_inst_c0.push_back(0);
    for (int i = 0; i < ni; ++i) {
      const int64_t d0 = el_start[i] * g->dpe, d1 = el_start[i + 1] * g->dpe;
      for (int64_t d = d0; d < d1; d += CHUNK) {
        g->h_cbeg.push_back(d);
        g->h_clen.push_back((int32_t)std::min<int64_t>(CHUNK, d1 - d));
      }
      g->h_inst_c0.push_back((int32_t)g->h_cbeg.size());
    }
    g->restart = 30;
    g->mstride = 64;
    const size_t nchunk = g->h_cbeg.size();
    auto dm = [&](void** p, size_t bytes) {
      if (st) return;
      if (hipMalloc(p, std::max<size_t>(bytes, 8)) != hipSuccess) {
        set_error("batch allocation failed");
        st = PDES_ERR_NOMEM;
      }
    };
    hipFree(c->d_flag);
    c->d_flag = nullptr;
    dm((void**)&c->d_flag, ni * 4);
    dm((void**)&c->d_el_list, nel * 4);
    dm((void**)&c->d_el_inst, nel * 4);
    dm((void**)&c->d_inst_phys, ni * sizeof(PhysConst));
    dm((void**)&c->d_inst_active, ni * 4);
    g->h_active.assign(ni, 1);
    hipFree(c->d_ptc_scale);
    c->d_ptc_scale = nullptr;
    dm((void**)&c->d_ptc_scale, ni * 8);
    dm((void**)&g->d_cbeg, nchunk * 8);
    dm((void**)&g->d_clen, nchunk * 4);
    dm((void**)&g->d_inst_c0, (ni + 1) * 4);
    dm((void**)&g->d_clist, nchunk * 4);
    dm((void**)&g->d_ilist, ni * 4);
    dm((void**)&g->d_partial, nchunk * g->mstride * 8);
    dm((void**)&g->d_seg, (size_t)ni * g->mstride * 8);
    dm((void**)&g->d_coef, (size_t)ni * g->mstride * 8);
    dm((void**)&g->d_z, c->numDof * 8);
    dm((void**)&g->d_w, c->numDof * 8);
    dm((void**)&g->d_dq, c->numDof * 8);
    dm((void**)&g->d_rhs, c->numDof * 8);
    dm((void**)&g->d_chunk_inst, nchunk * 4);
    for (int32_t** f : {&g->d_iter, &g->d_live, &g->d_part, &g->d_total, &g->d_jdone}) dm((void**)f, ni * 4);
    g->gmres_R = g->mstride - 2;
    dm((void**)&g->d_H, (size_t)ni * (g->gmres_R + 1) * g->gmres_R * 8);
    dm((void**)&g->d_cs, (size_t)ni * g->gmres_R * 8);
    dm((void**)&g->d_sn, (size_t)ni * g->gmres_R * 8);
    dm((void**)&g->d_g, (size_t)ni * (g->gmres_R + 1) * 8);
    dm((void**)&g->d_gs, (size_t)ni * GS_NSCAL * 8);
    if (st) break;
    if (hipHostMalloc((void**)&g->h_flags, (size_t)ni * 4) != hipSuccess) {
      set_error("pinned allocation failed");
      st = PDES_ERR_NOMEM;
      break;
    }
    std::vector<int32_t> chunk_inst(nchunk);
    for (int i = 0; i < ni; ++i)
      for (int ch = g->h_inst_c0[i]; ch < g->h_inst_c0[i + 1]; ++ch) chunk_inst[ch] = i;
    if (hipHostMalloc((void**)&g->h_seg, (size_t)ni * g->mstride * 8) != hipSuccess ||
        hipHostMalloc((void**)&g->h_coef, (size_t)ni * g->mstride * 8) != hipSuccess) {
      set_error("pinned allocation failed");
      st = PDES_ERR_NOMEM;
      break;
    }
    std::memset(g->h_coef, 0, (size_t)ni * g->mstride * 8);
    hipStream_t s = c->stream;
    if (hipMemcpyAsync(c->d_elface, c->h_elface.data(), c->h_elface.size() * sizeof(ElFace), hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(c->d_el_inst, g->h_el_inst.data(), nel * 4, hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(c->d_inst_phys, ip.data(), ni * sizeof(PhysConst), hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(g->d_cbeg, g->h_cbeg.data(), nchunk * 8, hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(g->d_clen, g->h_clen.data(), nchunk * 4, hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(g->d_inst_c0, g->h_inst_c0.data(), (ni + 1) * 4, hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(g->d_chunk_inst, chunk_inst.data(), nchunk * 4, hipMemcpyHostToDevice, s) ||
        hipMemsetAsync(c->d_flag, 0, ni * 4, s) || hipStreamSynchronize(s)) {
      set_error("batch upload failed");
      st = PDES_ERR_HIP;
    }
    // boundary states with each instance's own physics (records now tagged)
    if (!st) st = bc_state_setup(c);
    if (!st && hipStreamSynchronize(s) != hipSuccess) {
      set_error("boundary state setup failed");
      st = PDES_ERR_HIP;
    }
  }
  if (st) {
    for (auto* g : B->groups) group_free(g);
    delete B;
    return st;
  }
  *out = B;
  return PDES_OK;
}

int pdes_batch_destroy(pdes_batch* B) {
  if (!B) return PDES_OK;
  for (auto* g : B->groups) group_free(g);
  delete B;
  return PDES_OK;
}

int pdes_batch_num_groups(pdes_batch* B, int* ngroups) {
  PDES_CHECK(B && ngroups, PDES_ERR_ARG, "null argument");
  *ngroups = (int)B->groups.size();
  return PDES_OK;
}

int pdes_batch_graph_stats(pdes_batch* B, int64_t* captures, int64_t* replays) {
  PDES_CHECK(B && captures && replays, PDES_ERR_ARG, "null argument");
  *captures = *replays = 0;
  for (auto* g : B->groups) {
    *captures += g->graph_captures;
    *replays += g->graph_replays;
  }
  return PDES_OK;
}

static int locate(pdes_batch* B, int inst, pdes_batch_group** g, InstState** s) {
  PDES_CHECK(B && inst >= 0 && inst < B->ninst, PDES_ERR_ARG, "instance %d out of range", inst);
  auto w = B->where[inst];
  *g = B->groups[w.first];
  *s = &(*g)->inst[w.second];
  return PDES_OK;
}

int pdes_batch_set_state(pdes_batch* B, int inst, const double* q) {
  pdes_batch_group* g;
  InstState* s;
  int st = locate(B, inst, &g, &s);
  if (st) return st;
  PDES_CHECK(q, PDES_ERR_ARG, "null state");
  const int64_t off = s->el0 * g->dpe, n = (s->el1 - s->el0) * g->dpe;
  PDES_HIP(hipMemcpyAsync(g->c->d_q + off, q, n * 8, hipMemcpyHostToDevice, g->c->stream));
  PDES_HIP(hipStreamSynchronize(g->c->stream));
  return PDES_OK;
}

int pdes_batch_get_state(pdes_batch* B, int inst, double* q) {
  pdes_batch_group* g;
  InstState* s;
  int st = locate(B, inst, &g, &s);
  if (st) return st;
  PDES_CHECK(q, PDES_ERR_ARG, "null state");
  const int64_t off = s->el0 * g->dpe, n = (s->el1 - s->el0) * g->dpe;
  PDES_HIP(hipMemcpyAsync(q, g->c->d_q + off, n * 8, hipMemcpyDeviceToHost, g->c->stream));
  PDES_HIP(hipStreamSynchronize(g->c->stream));
  return PDES_OK;
}

// residual of every instance (one launch per group), instance inst copied out;
// returns PDES_ERR_PHYS if that instance fails checkDensity/checkPressure
int pdes_batch_eval_residual(pdes_batch* B, int inst, double* res) {
  pdes_batch_group* g;
  InstState* s;
  int st = locate(B, inst, &g, &s);
  if (st) return st;
  PDES_CHECK(res, PDES_ERR_ARG, "null output");
  pdes_ctx* c = g->c;
  const int local = B->where[inst].second;
  c->nlist = -1;
  PDES_HIP(hipMemsetAsync(c->d_flag, 0, g->inst.size() * 4, c->stream));
  st = launch_residual(c, c->d_q, nullptr, c->d_res, nullptr, c->phys.check_positivity != 0);
  if (st) return st;
  const int64_t off = s->el0 * g->dpe, n = (s->el1 - s->el0) * g->dpe;
  int32_t flag = 0;
  PDES_HIP(hipMemcpyAsync(res, c->d_res + off, n * 8, hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipMemcpyAsync(&flag, c->d_flag + local, 4, hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  PDES_CHECK(!flag, PDES_ERR_PHYS, "instance %d: non-positive density or pressure", inst);
  return PDES_OK;
}

}  // extern "C"

namespace {

// up to max_passes passes of every group (to the end when max_passes <= 0);
// the groups run concurrently, each from its own host thread on its own stream
int batch_run(pdes_batch* B, const pdes_newton_opts* o, int max_passes, pdes_newton_report* reports, int* all_done) {
  PDES_CHECK(B && o, PDES_ERR_ARG, "null argument");
  PDES_CHECK(o->gmres_restart > 0 && o->gmres_restart < 63, PDES_ERR_ARG, "gmres_restart must be in 1..62");
  PDES_CHECK(!o->globalize_euler || o->euler_tau > 0, PDES_ERR_ARG, "euler_tau must be positive");
  std::vector<int> status(B->groups.size(), 0);
  std::vector<std::string> errs(B->groups.size());
  std::vector<std::thread> th;
  for (size_t gi = 0; gi < B->groups.size(); ++gi) {
    pdes_batch_group* g = B->groups[gi];
    if (g->started) {
      PDES_CHECK(g->restart == o->gmres_restart, PDES_ERR_ARG, "gmres_restart changed during a solve");
    } else if (g->restart != o->gmres_restart || !g->d_V) {
      if (g->d_V) hipFree(g->d_V);
      g->d_V = nullptr;
      g->restart = o->gmres_restart;
      if (hipMalloc((void**)&g->d_V, (size_t)(g->restart + 1) * g->c->numDof * 8) != hipSuccess ||
          hipMemsetAsync(g->d_V, 0, (size_t)(g->restart + 1) * g->c->numDof * 8, g->c->stream) != hipSuccess ||
          hipStreamSynchronize(g->c->stream) != hipSuccess) {
        set_error("GMRES basis allocation failed");
        return PDES_ERR_NOMEM;
      }
    }
    th.emplace_back([&, gi, g]() {
      status[gi] = group_newton(g, o, max_passes);
      if (status[gi]) errs[gi] = pdes_last_error();
    });
  }
  for (auto& t : th) t.join();
  for (size_t gi = 0; gi < B->groups.size(); ++gi)
    if (status[gi]) {
      set_error("%s", errs[gi].c_str());
      return status[gi];
    }
  if (reports)
    for (int i = 0; i < B->ninst; ++i) {
      auto w = B->where[i];
      reports[i] = B->groups[w.first]->inst[w.second].rep;
      if (B->groups[w.first]->inst[w.second].failed) reports[i].converged = -1;
    }
  if (all_done) {
    *all_done = 1;
    for (auto* g : B->groups) *all_done = *all_done && g->finished;
  }
  return PDES_OK;
}

}  // namespace

extern "C" {

// Newton for every instance of the batch, from the start.
int pdes_batch_newton_solve(pdes_batch* B, const pdes_newton_opts* o, pdes_newton_report* reports) {
  PDES_CHECK(B && o, PDES_ERR_ARG, "null argument");
  for (auto* g : B->groups) {
    g->started = false;
    g->finished = false;
  }
  return batch_run(B, o, 0, reports, nullptr);
}

int pdes_batch_advance(pdes_batch* B, const pdes_newton_opts* o, int max_passes, pdes_newton_report* reports,
                       int32_t* active, int* all_done) {
  PDES_CHECK(B && o && max_passes > 0, PDES_ERR_ARG, "bad argument");
  int st = batch_run(B, o, max_passes, reports, all_done);
  if (st) return st;
  if (active)
    for (int i = 0; i < B->ninst; ++i) {
      auto w = B->where[i];
      active[i] = B->groups[w.first]->inst[w.second].active ? 1 : 0;
    }
  return PDES_OK;
}

int pdes_batch_export(pdes_batch* B, int inst, pdes_batch_inst_state* out, double* q, double* dq, double* rhs) {
  pdes_batch_group* g;
  InstState* s;
  int st = locate(B, inst, &g, &s);
  if (st) return st;
  PDES_CHECK(out && q, PDES_ERR_ARG, "null argument");
  pdes_ctx* c = g->c;
  const int local = B->where[inst].second;
  const int64_t off = s->el0 * g->dpe, n = (s->el1 - s->el0) * g->dpe;
  out->active = s->active;
  out->failed = s->failed;
  out->in_solve = s->in_solve;
  out->newton_steps = s->nit;
  out->tau = s->tau;
  out->tau_prev_norm = s->tau_prev_norm;
  out->gmres_iters = 0;
  out->gmres_live = 1;
  out->gmres_beta0 = -1.0;
  out->report = s->rep;
  PDES_HIP(hipMemcpyAsync(q, c->d_q + off, n * 8, hipMemcpyDeviceToHost, c->stream));
  if (s->in_solve) {
    PDES_CHECK(dq && rhs, PDES_ERR_ARG, "an instance in a linear solve needs dq and rhs");
    double gs[GS_NSCAL];
    PDES_HIP(hipMemcpyAsync(dq, g->d_dq + off, n * 8, hipMemcpyDeviceToHost, c->stream));
    PDES_HIP(hipMemcpyAsync(rhs, g->d_rhs + off, n * 8, hipMemcpyDeviceToHost, c->stream));
    PDES_HIP(hipMemcpyAsync(&out->gmres_iters, g->d_total + local, 4, hipMemcpyDeviceToHost, c->stream));
    PDES_HIP(hipMemcpyAsync(&out->gmres_live, g->d_live + local, 4, hipMemcpyDeviceToHost, c->stream));
    PDES_HIP(hipMemcpyAsync(gs, g->d_gs + (size_t)local * GS_NSCAL, sizeof gs, hipMemcpyDeviceToHost, c->stream));
    PDES_HIP(hipStreamSynchronize(c->stream));
    out->gmres_beta0 = gs[GS_BETA0];
  }
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

int pdes_batch_import(pdes_batch* B, int inst, const pdes_batch_inst_state* in, const double* q, const double* dq,
                      const double* rhs) {
  pdes_batch_group* g;
  InstState* s;
  int st = locate(B, inst, &g, &s);
  if (st) return st;
  PDES_CHECK(in && q, PDES_ERR_ARG, "null argument");
  PDES_CHECK(!g->started, PDES_ERR_ARG, "instances are imported before the batch's first pdes_batch_advance");
  PDES_CHECK(in->active, PDES_ERR_ARG, "only an instance that has not finished can be imported");
  pdes_ctx* c = g->c;
  const int64_t off = s->el0 * g->dpe, n = (s->el1 - s->el0) * g->dpe;
  s->imported = true;
  s->active = true;
  s->failed = false;
  s->in_solve = in->in_solve != 0;
  s->nit = in->newton_steps;
  s->tau = in->tau;
  s->tau_prev_norm = in->tau_prev_norm;
  s->g_total = in->gmres_iters;
  s->g_live = in->gmres_live;
  s->g_beta0 = in->gmres_beta0;
  s->rep = in->report;
  PDES_HIP(hipMemcpyAsync(c->d_q + off, q, n * 8, hipMemcpyHostToDevice, c->stream));
  if (s->in_solve) {
    PDES_CHECK(dq && rhs, PDES_ERR_ARG, "an instance in a linear solve needs dq and rhs");
    PDES_HIP(hipMemcpyAsync(g->d_dq + off, dq, n * 8, hipMemcpyHostToDevice, c->stream));
    PDES_HIP(hipMemcpyAsync(g->d_rhs + off, rhs, n * 8, hipMemcpyHostToDevice, c->stream));
  }
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

}  // extern "C"
