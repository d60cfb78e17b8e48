// ctx.hpp -- the library context (one per solve instance / HIP stream).
#pragma once
#include <vector>

#include "residual_kernels.hpp"

struct pdes_jacobian;
struct pdes_precond;

struct pdes_ctx {
  // problem shape
  int dim = 0, nd = 0, numEl = 0, Np = 0, Nf = 0, stencil = 0, norient = 0, nfaces = 0;
  int diag_e = 0, numIface = 0, numBndry = 0, numBC = 0;
  int metric_per_node = 1;    // 0 when every element's dxidx is the same at all its nodes
  int64_t numDof = 0;
  pdes_phys_desc phys{};
  pdes::PhysConst pc{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;             // HIP device current at creation (owns the stream and buffers)
  // device tables
  double* d_dxidx = nullptr;
  pdes::ElFace* d_elface = nullptr;
  pdes::FaceSlot* d_slot = nullptr;   // diagonal-E face slots [e][f * Nf + k]
  double* d_nrm_face = nullptr;
  double* d_nrm_bndry = nullptr;
  double* d_coords_bndry = nullptr;
  double* d_Q = nullptr;
  double* d_S2c = nullptr;    // 2 S by circulant distance (split-form volume)
  double* d_interp = nullptr;
  int32_t* d_perm = nullptr;
  int32_t* d_nbrperm = nullptr;
  int32_t* d_nodepos = nullptr;
  double* d_wface = nullptr;
  double* d_Minv = nullptr;
  double* d_Minv_node = nullptr;   // one value per volume node
  // host copies used by assembly / functionals
  std::vector<double> h_w, h_Q, h_interp, h_wface, h_jac, h_nrm_bndry, h_coords_bndry;
  std::vector<int32_t> h_perm, h_nbrperm, h_nodepos, h_interfaces, h_bndryfaces, h_bndry_offsets, h_bc_type;
  std::vector<pdes::ElFace> h_elface;
  // state
  double* d_q = nullptr;
  double* d_res = nullptr;
  double* d_work = nullptr;     // numDof scratch (RK4 / Jv direction)
  double* d_work2 = nullptr;    // numDof scratch
  double* d_partial = nullptr;  // reduction partials
  double* d_scalar = nullptr;   // reduction results
  int32_t* d_flag = nullptr;    // positivity flag, one per instance
  // packed batches: element work list (active instances), element -> instance,
  // per-instance physics; single-instance contexts leave these unset
  int32_t* d_el_list = nullptr;
  int32_t nlist = -1;           // < 0: no list, every element is processed
  int32_t* d_el_inst = nullptr;
  pdes::PhysConst* d_inst_phys = nullptr;
  int ninst = 1;
  // implicit-Euler globalization: per node M (1 + sqrt(jac)), per instance 1/tau
  double* d_ptc_diag = nullptr;
  double* d_ptc_scale = nullptr;
  bool ptc_on = false;
  // homotopy (predictorCorrectorHomotopy): when hom_lambda >= 0 the Newton
  // system is H = (1 - lambda) R + lambda D with D the first-order dissipation
  double hom_lambda = -1.0;
  int hom_bndry = 0;          // homotopy_addBoundaryIntegrals
  double* d_w = nullptr;      // SBP weights H (uploaded on first homotopy use)
  // packed batches: per-instance 0/1 mask honoured by the BSR product and the
  // preconditioner apply (instances with 0 are skipped); null = all
  const int32_t* d_mask = nullptr;
  // packed batches: 0/1 per instance, 1 while the instance is in the Newton
  // work list (kernels that walk fixed element lists skip the others)
  int32_t* d_inst_active = nullptr;
  double* d_qbc = nullptr;   // boundary states [numBoundaryFaces][Nf][ndof]
  // replay multi-launch sequences (ILU sweeps) from captured HIP graphs;
  // off when the environment variable PDES_NO_GRAPHS is set at creation
  bool use_graphs = true;
  // the elements an ILU(0) application will touch and the dependency levels
  // they span (set by the batched solver around its Arnoldi steps; -1: all);
  // ilu_sel: the dependency-driven sweeps run over the level-ordered lists
  // compacted to those elements
  int64_t ilu_hint_el = -1;
  int ilu_hint_lf = 0, ilu_hint_lb = 0;
  bool ilu_sel = false;
  // a caller is capturing this context's stream into a graph: launch
  // everything directly (no nested graph replays)
  bool capturing = false;
  int partial_n = 0;
  // residual launch variants (see residual_launch.hpp) for the plain and the
  // RK-epilogue kernel, chosen by the autotuner or the default table
  int res_variant[2] = {1, 1};
  // solver state
  pdes_jacobian* jac = nullptr;
  pdes_precond* pc_obj = nullptr;
  void* gmres_ws = nullptr;     // cached GMRES workspace

  pdes::DevMesh devmesh() const {
    return pdes::DevMesh{numEl, metric_per_node, nlist >= 0 ? d_el_list : nullptr, nlist, d_el_inst, d_inst_phys,
                         d_dxidx, d_elface, d_nrm_face, d_nrm_bndry, d_coords_bndry, d_qbc, d_slot};
  }
  // elements processed by a launch (the work list when one is set)
  int64_t active_count() const { return nlist >= 0 ? nlist : numEl; }
  const int32_t* active_list() const { return nlist >= 0 ? d_el_list : nullptr; }
  pdes::DevSBP devsbp() const {
    return pdes::DevSBP{d_Q, d_interp, d_perm, d_nbrperm, d_nodepos, d_wface, norient, d_S2c};
  }
};

namespace pdes {
// residual launch (value run, or directional derivative when v != nullptr)
// (flag, when set, replaces the context's positivity flag word)
int launch_residual(pdes_ctx* c, const double* q, const double* v, double* res, double* jv,
                    bool check, const RKStage* rk = nullptr, int32_t* flag = nullptr);
// deterministic sum of x_i^2 * w_i (w may be null) into *out_dev
int launch_wnorm2(pdes_ctx* c, const double* x, const double* w, int64_t n, double* out_dev);
int sync_check_flag(pdes_ctx* c);
int jacobian_free(pdes_ctx* c);
int precond_free(pdes_ctx* c);
int bc_state_setup(pdes_ctx* c);
}  // namespace pdes
