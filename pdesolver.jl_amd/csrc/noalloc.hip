// noalloc.hip -- the RK4 driver for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  pdes_rk4 keeps its stage accumulator and stage states in
// buffers the library allocates with the context; pdes_rk4_noalloc runs the
// same kernels with those buffers, and the positivity flag, in a fixed
// file-scope device array instead, so a call allocates nothing.
#include <mutex>

#include "ctx.hpp"

namespace pdes {

// largest state (numDof doubles) the array serves: stage accumulator and
// two stage states of NOALLOC_MAX_DOF doubles each, then the flag word
constexpr int64_t NOALLOC_MAX_DOF = int64_t(1) << 21;
constexpr int64_t NOALLOC_DOUBLES = 3 * NOALLOC_MAX_DOF + 32;

// One copy per GPU, held until the process ends (48 MiB + 256 B).  Host code
// never uses this name as an address: it is a host-side stand-in; the device
// address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static double* g_scratch_ptr[MAX_DEVICES] = {};

// the device address of `scratch` on device dev, looked up on the first call
// on that device with dev current
static int scratch_on(int dev, double** out) {
  PDES_CHECK(dev >= 0 && dev < MAX_DEVICES, PDES_ERR_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    int prev = 0;
    PDES_HIP(hipGetDevice(&prev));
    PDES_HIP(hipSetDevice(dev));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    PDES_HIP(hipSetDevice(prev));
    if (e != hipSuccess || !ptr) {
      set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      return PDES_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return PDES_OK;
}

__global__ void clear_flag_kernel(int32_t* flag) { *flag = 0; }

}  // namespace pdes

using namespace pdes;

extern "C" {

int pdes_rk4_noalloc(pdes_ctx* c, double* q_dev, int nsteps, double h) {
  PDES_CHECK(c && q_dev && nsteps >= 0, PDES_ERR_ARG, "bad argument");
  PDES_CHECK(c->nlist < 0 && c->ninst == 1, PDES_ERR_UNSUPPORTED, "pdes_rk4_noalloc takes a single-solve context");
  PDES_CHECK(c->numDof <= NOALLOC_MAX_DOF, PDES_ERR_ARG,
             "state of %lld doubles exceeds the %lld that pdes_rk4_noalloc serves without allocating",
             (long long)c->numDof, (long long)NOALLOC_MAX_DOF);
  double* base = nullptr;
  int st = scratch_on(c->device, &base);
  if (st) return st;
  double* acc = base;
  double* stage_a = base + NOALLOC_MAX_DOF;
  double* stage_b = base + 2 * NOALLOC_MAX_DOF;
  int32_t* flag = reinterpret_cast<int32_t*>(base + 3 * NOALLOC_MAX_DOF);
  // the array keeps the previous call's contents: the flag must start at zero
  // (a kernel, not a memset, so the call can sit inside a captured graph)
  hipLaunchKernelGGL(clear_flag_kernel, dim3(1), dim3(1), 0, c->stream, flag);
  const bool check = c->phys.check_positivity != 0;
  const double cst[4] = {h / 2, h / 2, h, 0.0};
  const double wst[4] = {1.0, 2.0, 2.0, 1.0};
  // stage inputs/outputs as in pdes_rk4: q -> A -> B -> A -> q (in place)
  const double* qin[4] = {q_dev, stage_a, stage_b, stage_a};
  double* qout[4] = {stage_a, stage_b, stage_a, q_dev};
  for (int s = 0; s < nsteps; ++s) {
    for (int stage = 0; stage < 4; ++stage) {
      RKStage rk{c->d_Minv_node, q_dev, acc, qout[stage], wst[stage], cst[stage], h / 6, stage};
      st = launch_residual(c, qin[stage], nullptr, nullptr, nullptr, check, &rk, flag);
      if (st) return st;
    }
  }
  PDES_HIP(hipGetLastError());
  int32_t f = 0;
  PDES_HIP(hipMemcpyAsync(&f, flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  PDES_HIP(hipStreamSynchronize(c->stream));
  PDES_CHECK(f == 0, PDES_ERR_PHYS, "non-positive density or pressure in the state (checkDensity/checkPressure)");
  return PDES_OK;
}

}  // extern "C"
