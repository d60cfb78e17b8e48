// residual.hip -- operator-shape dispatch of the fused residual kernel and the
// reduction / vector kernels used around it.
#include <algorithm>
#include <vector>

#include "residual_launch.hpp"
#include "shapes.hpp"

namespace pdes {

int launch_residual(pdes_ctx* c, const double* q, const double* v, double* res, double* jv, bool check,
                    const RKStage* rk, int32_t* flag_dev) {
  int32_t* flag = check ? (flag_dev ? flag_dev : c->d_flag) : nullptr;
#define PDES_DISPATCH(D, NP, NF, ST, DE)                                                        \
  if (c->dim == D && c->Np == NP && c->Nf == NF && c->stencil == ST && (c->diag_e != 0) == DE) { \
    using Op = OpShape<D, NP, NF, ST, DE>;                                                      \
    if (v) return launch_shape<Op, dual<1>>(c, q, v, res, jv, flag, nullptr, 0);                \
    return launch_shape<Op, double>(c, q, v, res, jv, flag, rk, c->res_variant[rk ? 1 : 0]);    \
  }
  PDES_FOR_EACH_SHAPE(PDES_DISPATCH)
#undef PDES_DISPATCH
  set_error("no compiled kernel for dim=%d Np=%d Nf=%d stencil=%d diag_e=%d", c->dim, c->Np, c->Nf,
            c->stencil, c->diag_e);
  return PDES_ERR_UNSUPPORTED;
}

// ---------------------------------------------------------------------------
// deterministic weighted sum of squares: fixed per-block partial order, then
// one block sums the partials in index order.
// ---------------------------------------------------------------------------
constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 1024;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  return sh[0];
}

__global__ void wnorm2_partial(const double* __restrict__ x, const double* __restrict__ w, int64_t n,
                               double* __restrict__ partial) {
  __shared__ double sh[RED_THREADS];
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double xi = x[i];
    acc += w ? xi * xi * w[i] : xi * xi;
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// y = x over n2 double2 elements, grid-stride (the copy reference kernel)
__global__ void pdes_copy_kernel(const double2* __restrict__ x, double2* __restrict__ y, int64_t n2) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) y[i] = x[i];
}

__global__ void sum_partials(const double* __restrict__ partial, int n, double* __restrict__ out) {
  __shared__ double sh[RED_THREADS];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[0] = s;
}

int launch_wnorm2(pdes_ctx* c, const double* x, const double* w, int64_t n, double* out_dev) {
  int blocks = (int)std::min<int64_t>(RED_BLOCKS, std::max<int64_t>(1, ceil_div(n, RED_THREADS)));
  hipLaunchKernelGGL(wnorm2_partial, dim3(blocks), dim3(RED_THREADS), 0, c->stream, x, w, n, c->d_partial);
  hipLaunchKernelGGL(sum_partials, dim3(1), dim3(RED_THREADS), 0, c->stream, c->d_partial, blocks, out_dev);
  PDES_HIP(hipGetLastError());
  return PDES_OK;
}

}  // namespace pdes

using namespace pdes;

extern "C" {

int pdes_eval_residual(pdes_ctx* c, double t) {
  (void)t;
  PDES_CHECK(c, PDES_ERR_ARG, "null context");
  return launch_residual(c, c->d_q, nullptr, c->d_res, nullptr, c->phys.check_positivity != 0);
}

int pdes_residual_norm(pdes_ctx* c, double* norm) {
  PDES_CHECK(c && norm, PDES_ERR_ARG, "null argument");
  int st = launch_wnorm2(c, c->d_res, c->d_Minv, c->numDof, c->d_scalar);
  if (st) return st;
  double v = 0;
  PDES_HIP(hipMemcpyAsync(&v, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  st = sync_check_flag(c);
  if (st) return st;
  *norm = std::sqrt(v);
  return PDES_OK;
}

int pdes_jv_complex_step(pdes_ctx* c, const double* v_host, double* b_host) {
  PDES_CHECK(c && v_host && b_host, PDES_ERR_ARG, "null argument");
  PDES_HIP(hipMemcpyAsync(c->d_work, v_host, c->numDof * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int st = launch_residual(c, c->d_q, c->d_work, nullptr, c->d_work2, false);
  if (st) return st;
  PDES_HIP(hipMemcpyAsync(b_host, c->d_work2, c->numDof * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  return PDES_OK;
}

// Device-resident Jv: b = dR/dq v for device vectors (matrix-free linear operator).
int pdes_jv_device(pdes_ctx* c, const double* v_dev, double* b_dev) {
  PDES_CHECK(c && v_dev && b_dev, PDES_ERR_ARG, "null argument");
  return launch_residual(c, c->d_q, v_dev, nullptr, b_dev, false);
}

// Autotuner: times every launch variant of the residual kernel (plain and
// RK-epilogue forms) on this context's mesh and keeps the fastest.
int pdes_autotune(pdes_ctx* c, int iters, int32_t* chosen) {
  PDES_CHECK(c && iters > 0, PDES_ERR_ARG, "bad argument");
  hipEvent_t a, b;
  PDES_HIP(hipEventCreate(&a));
  PDES_HIP(hipEventCreate(&b));
  // every variant is timed in each of AUTOTUNE_ROUNDS interleaved rounds and
  // keeps its best round, so a clock ramp at the start or a one-off stall
  // does not decide the choice
  constexpr int AUTOTUNE_ROUNDS = 3;
  const int nvar = res_variant_count(c);
  for (int rkmode = 0; rkmode < 2; ++rkmode) {
    std::vector<float> best(nvar, 1e30f);
    for (int round = 0; round < AUTOTUNE_ROUNDS; ++round) {
      for (int var = 0; var < nvar; ++var) {
        c->res_variant[rkmode] = var;
        // stage-1 epilogue: reads q and the accumulator, writes scratch buffers only
        RKStage rk{c->d_Minv_node, c->d_q, c->d_work, c->d_work2, 2.0, 0.0, 0.0, 1};
        const RKStage* rkp = rkmode ? &rk : nullptr;
        int st = launch_residual(c, c->d_q, nullptr, c->d_res, nullptr, false, rkp);  // warm-up
        if (st) return st;
        PDES_HIP(hipEventRecord(a, c->stream));
        for (int it = 0; it < iters; ++it) {
          st = launch_residual(c, c->d_q, nullptr, c->d_res, nullptr, false, rkp);
          if (st) return st;
        }
        PDES_HIP(hipEventRecord(b, c->stream));
        PDES_HIP(hipEventSynchronize(b));
        float ms = 0;
        PDES_HIP(hipEventElapsedTime(&ms, a, b));
        best[var] = std::min(best[var], ms);
      }
    }
    const int bestv = (int)(std::min_element(best.begin(), best.end()) - best.begin());
    c->res_variant[rkmode] = bestv;
    if (chosen) chosen[rkmode] = bestv;
  }
  PDES_HIP(hipEventDestroy(a));
  PDES_HIP(hipEventDestroy(b));
  return PDES_OK;
}

int pdes_set_launch_variant(pdes_ctx* c, int plain_variant, int rk_variant) {
  PDES_CHECK(c && plain_variant >= 0 && plain_variant < NUM_RES_VARIANTS && rk_variant >= 0 &&
             rk_variant < NUM_RES_VARIANTS, PDES_ERR_ARG, "launch variant out of range");
  c->res_variant[0] = plain_variant;
  c->res_variant[1] = rk_variant;
  return PDES_OK;
}

// Times `iters` back-to-back residual evaluations on the context stream with
// HIP events; *avg_ms is the average per evaluation (one kernel launch each).
int pdes_copy_bandwidth(int64_t nbytes, int iters, double* avg_ms) {
  PDES_CHECK(avg_ms && iters > 0 && nbytes >= 16 && nbytes % 16 == 0, PDES_ERR_ARG, "bad argument");
  const int64_t n2 = nbytes / 16;
  double2 *x = nullptr, *y = nullptr;
  if (hipMalloc((void**)&x, nbytes) != hipSuccess) { set_error("copy buffer"); return PDES_ERR_NOMEM; }
  if (hipMalloc((void**)&y, nbytes) != hipSuccess) { hipFree(x); set_error("copy buffer"); return PDES_ERR_NOMEM; }
  hipEvent_t a = nullptr, b = nullptr;
  float ms = 0;
  const unsigned grid = (unsigned)std::min<int64_t>(8192, (n2 + 255) / 256);
  int st = PDES_OK;
  if (hipMemset(x, 0, nbytes) != hipSuccess || hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    set_error("copy setup");
    st = PDES_ERR_HIP;
  }
  if (!st) {
    hipLaunchKernelGGL(pdes_copy_kernel, dim3(grid), dim3(256), 0, nullptr, x, y, n2);  // warm
    hipEventRecord(a, nullptr);
    for (int it = 0; it < iters; ++it) hipLaunchKernelGGL(pdes_copy_kernel, dim3(grid), dim3(256), 0, nullptr, x, y, n2);
    hipEventRecord(b, nullptr);
    if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) {
      set_error("copy timing");
      st = PDES_ERR_HIP;
    }
  }
  if (a) hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  hipFree(x);
  hipFree(y);
  *avg_ms = ms / iters;
  return st;
}

int pdes_device_memory(int64_t* free_bytes, int64_t* total_bytes) {
  PDES_CHECK(free_bytes && total_bytes, PDES_ERR_ARG, "null argument");
  size_t f = 0, t = 0;
  PDES_HIP(hipMemGetInfo(&f, &t));
  *free_bytes = (int64_t)f;
  *total_bytes = (int64_t)t;
  return PDES_OK;
}

int pdes_time_residual(pdes_ctx* c, int iters, double* avg_ms) {
  PDES_CHECK(c && avg_ms && iters > 0, PDES_ERR_ARG, "bad argument");
  hipEvent_t a, b;
  PDES_HIP(hipEventCreate(&a));
  PDES_HIP(hipEventCreate(&b));
  PDES_HIP(hipEventRecord(a, c->stream));
  for (int it = 0; it < iters; ++it) {
    int st = launch_residual(c, c->d_q, nullptr, c->d_res, nullptr, false);
    if (st) return st;
  }
  PDES_HIP(hipEventRecord(b, c->stream));
  PDES_HIP(hipEventSynchronize(b));
  float ms = 0;
  PDES_HIP(hipEventElapsedTime(&ms, a, b));
  PDES_HIP(hipEventDestroy(a));
  PDES_HIP(hipEventDestroy(b));
  *avg_ms = ms / iters;
  return PDES_OK;
}

// rk4 (src/NonlinearSolvers/rk4.jl:144-344): nsteps classical RK4 steps of
// dq/dt = Minv .* R(q) on the device state; 4 residual evaluations per step.
// If kernel_ms is not null, the residual launches of the whole run are
// bracketed by events: *elapsed_ms gets the wall time of all steps.
int pdes_rk4(pdes_ctx* c, int nsteps, double h, double* elapsed_ms) {
  PDES_CHECK(c && nsteps >= 0, PDES_ERR_ARG, "bad argument");
  double* q_old = c->d_q;
  double* acc = c->d_work;
  double* stage_a = c->d_work2;
  double* stage_b = c->d_res;   // residual storage doubles as the second stage buffer
  hipEvent_t a = nullptr, b = nullptr;
  if (elapsed_ms) {
    PDES_HIP(hipEventCreate(&a));
    PDES_HIP(hipEventCreate(&b));
    PDES_HIP(hipEventRecord(a, c->stream));
  }
  const double cst[4] = {h / 2, h / 2, h, 0.0};
  const double wst[4] = {1.0, 2.0, 2.0, 1.0};
  // stage inputs/outputs: q_old -> A -> B -> A -> q_old (in place)
  const double* qin[4] = {q_old, stage_a, stage_b, stage_a};
  double* qout[4] = {stage_a, stage_b, stage_a, q_old};
  for (int s = 0; s < nsteps; ++s) {
    for (int stage = 0; stage < 4; ++stage) {
      RKStage rk{c->d_Minv_node, q_old, acc, qout[stage], wst[stage], cst[stage], h / 6, stage};
      int st = launch_residual(c, qin[stage], nullptr, nullptr, nullptr, c->phys.check_positivity != 0, &rk);
      if (st) return st;
    }
  }
  PDES_HIP(hipGetLastError());
  if (elapsed_ms) {
    PDES_HIP(hipEventRecord(b, c->stream));
    PDES_HIP(hipEventSynchronize(b));
    float ms = 0;
    PDES_HIP(hipEventElapsedTime(&ms, a, b));
    *elapsed_ms = ms;
    PDES_HIP(hipEventDestroy(a));
    PDES_HIP(hipEventDestroy(b));
  }
  PDES_HIP(hipStreamSynchronize(c->stream));
  return sync_check_flag(c);
}

}  // extern "C"
