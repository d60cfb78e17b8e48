// jacobian.hpp -- device BSR Jacobian owned by a context.
#pragma once
#include "ctx.hpp"
#include "jacobian_kernels.hpp"
#include "shapes.hpp"

struct pdes_jacobian {
  bool fp64 = false;
  bool assembled = false;
  int64_t nrows = 0;   // block rows (nodes)
  int64_t nnzb = 0;
  int64_t* d_rowptr = nullptr;
  int32_t* d_colind = nullptr;
  void* d_vals = nullptr;    // float or double blocks [nnzb][nd][nd], row-major
  int32_t* d_err = nullptr;
  std::vector<int64_t> h_rowptr;
  std::vector<int32_t> h_colind;
  // complex-step coloring assembly (calcJacobianSparse): elements grouped by
  // a distance-2 colouring, the seed and product vectors
  std::vector<int64_t> color_ptr;
  std::vector<int32_t> h_color;   // colour of each element
  int32_t* d_color_el = nullptr;
  double* d_seed = nullptr;
  double* d_jv = nullptr;
  // position of block (c, r) for the block at (r, c): the pattern is
  // structurally symmetric (built on first use by a transposed product)
  int64_t* d_tpos = nullptr;
};

namespace pdes {
int jacobian_setup(pdes_ctx* c, int fp64);
int assemble(pdes_ctx* c);
// BSR values from directional derivatives of the residual, one per (colour,
// node, dof) of a distance-2 element colouring
int assemble_coloring(pdes_ctx* c);
// y = alpha J x + beta yadd (yadd may be null)
int launch_spmv(pdes_ctx* c, const double* x, double* y, double alpha, const double* yadd, double beta);
// y = J^T x (applyLinearOperatorTranspose), every block row
int launch_spmv_transpose(pdes_ctx* c, const double* x, double* y);
// J <- J^T in place (blocks swapped with their transposed positions and
// transposed), so that the products and preconditioners act on J^T
int transpose_jacobian(pdes_ctx* c);
// dJ/dq [numDof] of the scalar boundary functional `kind` (1..6) over the BC
// groups bcnums, into a device vector (functional.hip)
int functional_grad_device(pdes_ctx* c, int kind, const int32_t* bcnums, int nbc, double* d_out);
// out = a out + b D(q) with D the homotopy dissipation (homotopy.hip); with v
// set, jv = a jv + b (dD/dq) v and out (when not null) takes the value part
int launch_homotopy(pdes_ctx* c, const double* q, const double* v, double a, double b, double* out, double* jv);
// BSR values of (1 - lambda) dR/dq + lambda dD/dq (element-wise dR/dq, dD/dq
// by colouring)
int assemble_homotopy(pdes_ctx* c, double lambda);
// y -= (M / tau_vec) x for the implicit-Euler globalization (no-op when off)
int launch_ptc_shift(pdes_ctx* c, const double* x, double* y);
// per-instance pseudo-time steps tau[0..n) (stored as 1/tau); turns the term on
int set_ptc_scale(pdes_ctx* c, const double* tau, int n);
}  // namespace pdes
