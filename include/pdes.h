/*
 * pdes.h -- C ABI of the MI355X (gfx950) Euler residual / Newton-Krylov library.
 *
 * Every entry point returns an int status: 0 on success, a negative PDES_ERR_*
 * code on failure; pdes_last_error() returns a message for the calling thread.
 * Handles are opaque.  All array arguments are plain host pointers laid out as
 * the reference's Julia arrays (column-major), i.e. a C array whose axes are the
 * Julia axes reversed:  q[ndof, Np, numEl] (Julia)  ==  q[numEl][Np][ndof] (C).
 * Element, face, node and dof indices are 0-based.
 *
 * Threading: one context owns one HIP stream; calls on one context are not
 * re-entrant, calls on distinct contexts may run on different host threads.
 */
#ifndef PDES_H
#define PDES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDES_OK 0
#define PDES_ERR_ARG -1        /* bad argument / shape */
#define PDES_ERR_HIP -2        /* HIP runtime error */
#define PDES_ERR_NOMEM -3      /* device allocation failed */
#define PDES_ERR_PHYS -4       /* non-positive density or pressure (checkDensity/checkPressure) */
#define PDES_ERR_UNSUPPORTED -5
#define PDES_ERR_NOCONV -6     /* Newton / GMRES did not converge */

/* boundary condition functors (reference BCDict, src/solver/euler/bc.jl:2342-2369) */
#define PDES_BC_ISENTROPIC_VORTEX 0   /* isentropicVortexBC  bc.jl:554-567 */
#define PDES_BC_FREESTREAM 1          /* FreeStreamBC        bc.jl:1573-1586 */

/* numerical face fluxes (reference FluxDict, src/solver/euler/flux.jl:1229-1242) */
#define PDES_FLUX_ROE 0               /* RoeFlux    flux.jl:783-795 */
#define PDES_FLUX_IRSLF 1             /* IRSLFFlux  flux.jl:964-977 */
/* face_integral_type 2 with ESLFFaceIntegral (faceElementIntegrals.jl:611-630):
 * entropy-conservative element face integral plus the Lax-Friedrichs entropy
 * penalty, for dense-face (Omega) operators with volume_integral_type 2 */
#define PDES_FLUX_ESLF 2

/* Mesh description: the fields of the reference's AbstractDGMesh that the
 * residual reads (docs/src/interfaces.md:191-560).  Layouts (C order):
 *   dxidx        [numEl][Np][dim(x)][dim(xi)]   (Julia dxidx[xi, x, Np, numEl])
 *   jac          [numEl][Np]
 *   interfaces   [numInterfaces][5]  elementL, elementR, faceL, faceR, orient
 *   bndryfaces   [numBoundaryFaces][2] element, face
 *   bndry_offsets[numBC+1], faces of BC i are bndry_offsets[i] .. bndry_offsets[i+1]-1
 *   bc_type      [numBC]  PDES_BC_*
 *   nrm_face     [numInterfaces][Nf][dim]  scaled normal, outward from elementL
 *   nrm_bndry, coords_bndry [numBoundaryFaces][Nf][dim]
 */
typedef struct {
  int32_t dim;
  int32_t numEl;
  int32_t numInterfaces;
  int32_t numBoundaryFaces;
  int32_t numBC;
  const double* dxidx;
  const double* jac;
  const int32_t* interfaces;
  const int32_t* bndryfaces;
  const int32_t* bndry_offsets;
  const int32_t* bc_type;
  const double* nrm_face;
  const double* nrm_bndry;
  const double* coords_bndry;
} pdes_mesh_desc;

/* SBP operator tables (SummationByParts.jl fields used by the reference):
 *   w[Np]                     H diagonal
 *   Q[dim][Np][Np]            Q_d(i,j)
 *   interp[stencil][Nf]       face interpolation (identity for diagonal-E)
 *   perm[stencil][dim+1]      volume node of stencil entry s on face f
 *   nbrperm[Nf][norient]      elementR face node matching elementL face node k
 *   wface[Nf]                 face cubature weights (B)
 */
typedef struct {
  int32_t Np;
  int32_t Nf;
  int32_t stencil;
  int32_t norient;
  int32_t diag_e;              /* 1: SparseFace (R is a node selection) */
  const double* w;
  const double* Q;
  const double* interp;
  const int32_t* perm;
  const int32_t* nbrperm;
  const double* wface;
} pdes_sbp_desc;

/* Physics options (ParamType fields, src/solver/euler/types.jl:57-345, and the
 * option keys volume_integral_type / Flux_name / Volume_flux_name). */
typedef struct {
  double gamma;                /* 1.4 */
  double R_gas;                /* 287.058 */
  double Ma;                   /* freestream Mach */
  double aoa;                  /* angle of attack, radians */
  double rho_free;             /* 1.0 */
  double p_free;               /* 1/gamma */
  double vortex_Min;           /* isentropic vortex inflow Mach; 0.95 in the reference */
  int32_t volume_integral_type;/* 1: weak form Q^T F, 2: entropy-stable split form (IR) */
  int32_t face_flux;           /* PDES_FLUX_* */
  int32_t check_positivity;    /* checkDensity / checkPressure, euler.jl:543-611 */
} pdes_phys_desc;

typedef struct pdes_ctx pdes_ctx;

/* estimateMem (src/estimateMem.jl:1-74) for the device: bytes of HBM that one
 * steady solve of this mesh/operator allocates (context, BSR Jacobian in FP32
 * or FP64 blocks, preconditioner of kind pc_kind, GMRES(gmres_restart) basis,
 * Newton vectors; batched adds the per-instance share of a packed batch).
 * Used for admission by footprint. */
int pdes_estimate_bytes(const pdes_mesh_desc* mesh, const pdes_sbp_desc* sbp, int fp64_blocks,
                        int pc_kind, int gmres_restart, int batched, int64_t* bytes);

const char* pdes_last_error(void);
int pdes_version(void);

/* Context: copies mesh/SBP tables to HBM and allocates q/res element-major.
 * Replaces the construction of EulerData_ (src/solver/euler/types.jl:401-723)
 * with createMeshAndOperator (src/solver/common.jl:26-82). */
int pdes_ctx_create(const pdes_mesh_desc* mesh, const pdes_sbp_desc* sbp,
                    const pdes_phys_desc* phys, pdes_ctx** out);
int pdes_ctx_destroy(pdes_ctx* ctx);
int pdes_ctx_numdof(const pdes_ctx* ctx, int64_t* numdof);

/* State transfer, eqn.q / eqn.res in the reference layout [ndof, Np, numEl]. */
int pdes_set_state(pdes_ctx* ctx, const double* q_host);
int pdes_get_state(pdes_ctx* ctx, double* q_host);
int pdes_get_residual(pdes_ctx* ctx, double* res_host);

/* Device-pointer variants: q/res already resident in HBM (numDof doubles). */
int pdes_state_ptr(pdes_ctx* ctx, double** q_dev, double** res_dev);

/* evalResidual(mesh, sbp, eqn, opts, t)  src/solver/euler/euler.jl:111-175:
 * reads the device q, writes the device res.  Returns PDES_ERR_PHYS if
 * check_positivity is set and a node has rho <= 0 or p <= 0. */
int pdes_eval_residual(pdes_ctx* ctx, double t);

/* Launch tuning (no counterpart in the reference): the residual kernel has
 * launch variants 0-3 (elements per workgroup 256/Np or 128/Np x 3 or 4 waves
 * per SIMD) and, for the weak form, 4-7: the same shapes with the volume term
 * Q^T F on the FP64 matrix cores (the split form runs variant v & 3 for
 * these); pdes_autotune times each on this context's mesh (iters launches,
 * best of three interleaved rounds), for the plain and the RK4-stage form, and
 * keeps the fastest (chosen[0], chosen[1]).
 * pdes_set_launch_variant (0..7) fixes them. */
int pdes_autotune(pdes_ctx* ctx, int iters, int32_t* chosen);
int pdes_set_launch_variant(pdes_ctx* ctx, int plain_variant, int rk_variant);

/* Timing helpers: average residual-kernel duration over `iters` launches, and
 * nsteps RK4 steps (rk4, src/NonlinearSolvers/rk4.jl:144-344) on the device
 * state with 4 residual evaluations per step, h the time step; elapsed_ms (if
 * not null) receives the HIP-event time of all steps. */
int pdes_time_residual(pdes_ctx* ctx, int iters, double* avg_ms);
/* Measurement helper with no counterpart in the reference: copies nbytes of
 * doubles between two fresh device buffers with a plain grid-stride kernel
 * (pdes_copy_kernel) `iters` times; avg_ms receives the average kernel time.
 * Used as the attainable-bandwidth reference and to scale HBM byte counters. */
int pdes_copy_bandwidth(int64_t nbytes, int iters, double* avg_ms);
/* Free and total device memory in bytes (hipMemGetInfo), for footprint checks. */
int pdes_device_memory(int64_t* free_bytes, int64_t* total_bytes);
int pdes_rk4(pdes_ctx* ctx, int nsteps, double h, double* elapsed_ms);
/* No counterpart in the reference: pdes_rk4 for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Runs nsteps RK4 steps on the caller's device state
 * q_dev (numDof doubles, updated in place) with the context's operator, mesh
 * and launch variants, keeping the stage accumulator, the two stage states and
 * the positivity flag in a fixed per-GPU device array of the library (48 MiB,
 * held until the process ends) instead of context buffers: the call allocates
 * nothing.  Contexts with numDof above 2^21 (2,097,152) are rejected with
 * PDES_ERR_ARG; batch contexts with PDES_ERR_UNSUPPORTED.  Like pdes_rk4 it
 * waits for the context's stream and returns PDES_ERR_PHYS on a non-positive
 * density or pressure.  Every call on one GPU shares that GPU's array, so
 * calls that overlap in time on one GPU (from several threads or streams)
 * race and must be serialised by the caller. */
int pdes_rk4_noalloc(pdes_ctx* ctx, double* q_dev, int nsteps, double h);
/* b_dev = (dR/dq) v_dev at the device state, device vectors of numDof doubles
 * (the device-pointer form of pdes_jv_complex_step). */
int pdes_jv_device(pdes_ctx* ctx, const double* v_dev, double* b_dev);
/* Waits for the context's stream. */
int pdes_sync(pdes_ctx* ctx);

/* calcNorm(eqn, res, strongres=true)  src/Utils/Utils.jl:427-449 on the device
 * residual: sqrt(sum res_i^2 Minv_i). */
int pdes_residual_norm(pdes_ctx* ctx, double* norm);

/* evaldRdqProduct  src/interface2.jl:454-498: b = Im R(q + i h v)/h on the
 * device, h = 1e-20 (epsilon, src/input/read_input.jl:477-479).  v and b are
 * host arrays of numDof doubles. */
int pdes_jv_complex_step(pdes_ctx* ctx, const double* v_host, double* b_host);

/* Explicit Jacobian: evalJacobian  src/solver/euler/evalJacobian.jl:6-72 with
 * the assembler of src/jacobian/jacobian.jl:703-1759 targeting an on-device
 * BSR matrix, block size ndof, sparsity of SparseMatrixCSC(mesh, ..., INVISCID,
 * face_type)  src/jacobian/sparse.jl:250-495.  Values are stored in FP32 (or FP64
 * when fp64_blocks != 0). */
int pdes_jacobian_setup(pdes_ctx* ctx, int fp64_blocks);
int pdes_jacobian_nnzb(pdes_ctx* ctx, int64_t* nrows_blk, int64_t* nnzb);
int pdes_get_bsr_pattern(pdes_ctx* ctx, int64_t* rowptr, int32_t* colind);
int pdes_assemble_jacobian(pdes_ctx* ctx);
/* calcJacobianSparse (src/jacobian/jacobian.jl:365-446): the same BSR values
 * from directional derivatives of the residual over a distance-2 element
 * colouring (numColours x Np x ndof derivative evaluations).  An independent
 * check of the element-wise Jacobian kernels (pdes_assemble_jacobian), and
 * the derivative path of the homotopy Jacobian. */
int pdes_assemble_jacobian_coloring(pdes_ctx* ctx);
/* The distance-2 element colouring that pdes_assemble_jacobian_coloring
 * perturbs by, as the reference's mesh fields numColors, color and
 * pertNeighborEls (docs/src/interfaces.md:535-570, used by
 * calcJacobianSparse, src/jacobian/jacobian.jl:365-446): color[numEl]
 * (0-based), *ncolors, and pert_neighbor_els[numEl][ncolors], the element of
 * colour c among element i and its face neighbours, or -1.  Greedy in element
 * order: element e takes the lowest colour not used by an element within two
 * face-neighbour steps of it.  Any output pointer may be null (call once with
 * pert_neighbor_els null to learn ncolors).  Needs pdes_jacobian_setup. */
int pdes_get_coloring(pdes_ctx* ctx, int32_t* color, int32_t* ncolors, int32_t* pert_neighbor_els);
int pdes_get_bsr_values(pdes_ctx* ctx, void* values_host);
/* b = J x with the assembled BSR matrix (applyLinearOperator of PetscMatLO,
 * src/linearsolvers/lo_petscmat.jl), host vectors. */
int pdes_bsr_matvec(pdes_ctx* ctx, const double* x_host, double* b_host);

/* Preconditioner (AbstractPC, src/linearsolvers/LinearSolvers.jl:185-400):
 * PDES_PC_NONE, PDES_PC_BJACOBI = element block-Jacobi (NewtonBDiagPC,
 * preconditioning_types.jl:41), PDES_PC_ILU0 = element-block ILU(0) on the
 * BSR matrix.  Factor and apply in FP32. */
#define PDES_PC_NONE 0
#define PDES_PC_BJACOBI 1
#define PDES_PC_ILU0 2
int pdes_pc_setup(pdes_ctx* ctx, int kind);
int pdes_pc_apply(pdes_ctx* ctx, const double* b_host, double* x_host);

/* GMRES(restart) with right preconditioning (PETSc KSP options of
 * src/input/read_input.jl:560-581) on the device; classical Gram-Schmidt
 * with one reorthogonalisation, convergence decided on the true residual
 * ||b - A x|| <= max(rtol ||b - A x0||, atol).  mode 0: assembled BSR matvec,
 * mode 1: matrix-free complex-step Jv. */
int pdes_gmres(pdes_ctx* ctx, const double* b_host, double* x_host, double rtol,
               double atol, int restart, int maxit, int mode, int* iters,
               double* final_resnorm);

/* Newton (newtonInner, src/NonlinearSolvers/newton.jl:137-304) on the device
 * state: J dq = -R, q += dq until ||R||_{M^-1} <= res_abstol or reltol.
 * With globalize_euler (opts "newton_globalize_euler", implicit-Euler
 * pseudo-time stepping of src/NonlinearSolvers/globalization.jl:181-350) the
 * operator is J - M/tau_vec, tau_vec[node] = tau/(1 + sqrt(jac[node])),
 * tau = euler_tau at the first Jacobian and tau *= ||R_prev||/||R|| at each
 * later one; scale_euler ("newton_scale_euler") makes the first update use
 * ||R_prev|| = 1. */
typedef struct {
  double res_abstol;
  double res_reltol;
  double step_tol;
  int32_t itermax;
  double krylov_reltol;        /* inexact-Newton forcing, default 1e-2 */
  double krylov_abstol;
  int32_t krylov_itermax;
  int32_t gmres_restart;       /* 30 */
  int32_t pc_kind;             /* PDES_PC_* */
  int32_t fp64_blocks;         /* store the Jacobian blocks in FP64 instead of FP32; the preconditioner
                                  factors (block inverses, ILU(0)) are FP32 either way */
  int32_t jv_mode;             /* 0 BSR, 1 matrix-free complex step */
  int32_t globalize_euler;     /* implicit-Euler globalization on/off */
  double euler_tau;            /* initial pseudo-time step (1.0) */
  int32_t scale_euler;         /* newton_scale_euler */
} pdes_newton_opts;

typedef struct {
  int32_t iterations;
  int32_t converged;
  int32_t total_krylov_iters;
  double res_norm0;
  double res_norm;
  double step_norm;
  double elapsed_s;
} pdes_newton_report;

int pdes_newton_solve(pdes_ctx* ctx, const pdes_newton_opts* opts,
                      pdes_newton_report* report);

/* evalHomotopy with homotopy_function "FirstOrderDissipation"
 * (src/solver/euler/homotopy.jl:11-31 -> calcHomotopyDiss :51-249):
 * res[numDof] = D(q) at the device state.  Single (unbatched) contexts,
 * Np <= 64. */
int pdes_eval_homotopy(pdes_ctx* ctx, double* res);

/* opts "homotopy_addBoundaryIntegrals" (src/input/read_input.jl:438, default
 * 0) for this context's homotopy dissipation, its products, Jacobian and
 * solve: the boundary faces add the dissipation against the free stream. */
int pdes_set_homotopy_boundary(pdes_ctx* ctx, int add_boundary_integrals);

/* b = (dD/dq) v at the device state (the homotopy term of the complex-step
 * product of HomotopyPetscMatFreeLO, predictor_corrector.jl). */
int pdes_homotopy_jv(pdes_ctx* ctx, const double* v, double* b);

/* Jacobian of the homotopy system (1 - lambda) R + lambda D into the BSR
 * matrix (HomotopyMatLO calcLinearOperator, predictor_corrector.jl: the
 * physics Jacobian scaled by 1 - lambda plus lambda dD/dq). */
int pdes_assemble_homotopy_jacobian(pdes_ctx* ctx, double lambda);

/* predictorCorrectorHomotopy (src/NonlinearSolvers/predictor_corrector.jl:
 * 226-326) on the device state: lambda from 1 to 0, Newton corrector on
 * (1 - lambda) R + lambda D with the options' PC and Krylov settings, tangent
 * predictor with step control by psi_max (opts "homotopy_psi_max", default
 * 10 degrees) and the corrector step length.  report.iterations counts
 * predictor iterations; converged when ||R|| <= res_abstol or
 * ||R||/||R_0|| <= res_reltol within itermax. */
int pdes_homotopy_solve(pdes_ctx* ctx, const pdes_newton_opts* opts, double psi_max,
                        pdes_newton_report* report);

/* Boundary force functionals (calcBndryFunctional,
 * src/solver/euler/boundary_functional.jl:82-126, functionals.jl:19-160):
 * kind 0 = force vector (dim values), 1 = lift, 2 = drag, 3 = entropy flux,
 * 4 = mass flow, 5 = lift coefficient, 6 = drag coefficient (lift and drag
 * over 0.5 rho_free Ma^2); over the BC groups listed in bcnums. */
int pdes_eval_functional(pdes_ctx* ctx, int kind, const int32_t* bcnums, int nbc,
                         double* val);

/* evalFunctionalDeriv_q (src/interface2.jl:259-283, the boundary-functional
 * derivative of src/solver/euler/boundary_functional.jl): dJdq[numDof], the
 * derivative of the scalar functional `kind` (1..6 of pdes_eval_functional)
 * with respect to q at the current state, in the layout of q. */
int pdes_functional_deriv(pdes_ctx* ctx, int kind, const int32_t* bcnums, int nbc,
                          double* dJdq);

/* evalFunctionalDeriv_m (src/interface2.jl:205-227 -> _evalFunctionalDeriv_m,
 * src/solver/euler/boundary_functional.jl:61-79): the derivative of the scalar
 * functional `kind` (1..6) with respect to the boundary normals,
 * nrm_bndry_bar[numBoundaryFaces][Nf][dim] (the layout of nrm_bndry; zero on
 * faces outside bcnums), overwritten.  The functionals depend on no other
 * metric array. */
int pdes_functional_deriv_m(pdes_ctx* ctx, int kind, const int32_t* bcnums, int nbc,
                            double* nrm_bndry_bar);

/* evalResidual_revm (src/interface2.jl:329-372 -> src/solver/euler/evaldRdm.jl:
 * 19-70): with the seed psi[numDof] (res_bar, host), the products
 * psi^T dR/dm at the context's state for the mesh metrics, overwritten:
 *   dxidx_bar[numEl][Np][dim][dim]          (the layout of dxidx),
 *   nrm_face_bar[numInterfaces][Nf][dim]    (the layout of nrm_face),
 *   nrm_bndry_bar[numBoundaryFaces][Nf][dim],
 *   coords_bndry_bar[numBoundaryFaces][Nf][dim]  (through the boundary states).
 * The reference adds into the mesh's _bar fields, which its caller zeroes
 * first.  Any output may be NULL.  Every volume form and face flux of the
 * residual; PDES_ERR_UNSUPPORTED for packed batches. */
int pdes_eval_residual_revm(pdes_ctx* ctx, const double* psi, double* dxidx_bar, double* nrm_face_bar,
                            double* nrm_bndry_bar, double* coords_bndry_bar);

/* applyLinearOperatorTranspose (src/linearsolvers/LinearSolvers.jl:533-541,
 * the matrix-explicit method src/linearsolvers/lo_petscmat.jl:143-161):
 * b = J^T x with the BSR matrix of the last pdes_assemble_jacobian. */
int pdes_bsr_matvec_transpose(pdes_ctx* ctx, const double* x, double* b);

/* calcAdjoint (src/optimization/steady_adjoint.jl:30-66): solves
 * J^T psi = -dJ/dq at the current state with GMRES(restart), right
 * preconditioned by pc_kind built from J^T (linearSolveTranspose).  The
 * Jacobian is assembled here (FP32 or FP64 blocks); on return the context
 * holds J and its preconditioner again.  psi[numDof] in the layout of q. */
int pdes_adjoint_solve(pdes_ctx* ctx, int kind, const int32_t* bcnums, int nbc, int pc_kind,
                       int fp64_blocks, double rtol, double atol, int restart, int maxit,
                       double* psi, int* iters, double* final_resnorm);

/* Batched steady solves (no single counterpart in the reference, which runs
 * one solve per process): ninst independent problems, each described as for
 * pdes_ctx_create.  Instances that share an SBP operator and physics kind are
 * packed into one group whose kernels cover all its unconverged instances in
 * one launch; reductions, Krylov scalars and convergence tests are per
 * instance, so a result does not depend on the other instances in the batch.
 * Groups run concurrently on separate streams.  reports[i].converged is 1 on
 * convergence, 0 if itermax was reached, -1 if checkDensity/checkPressure
 * failed for that instance (the other instances continue). */
typedef struct pdes_batch pdes_batch;
int pdes_batch_create(int ninst, const pdes_mesh_desc* meshes, const pdes_sbp_desc* sbps,
                      const pdes_phys_desc* phys, pdes_batch** out);
int pdes_batch_destroy(pdes_batch* batch);
int pdes_batch_num_groups(pdes_batch* batch, int* ngroups);
/* Inspection (no counterpart in the reference): the Arnoldi blocks (8 GMRES
 * steps between flag read-backs) captured into HIP graphs and the replays of
 * those graphs so far, over all groups of the batch. */
int pdes_batch_graph_stats(pdes_batch* batch, int64_t* captures, int64_t* replays);
int pdes_batch_set_state(pdes_batch* batch, int inst, const double* q_host);
int pdes_batch_get_state(pdes_batch* batch, int inst, double* q_host);
int pdes_batch_newton_solve(pdes_batch* batch, const pdes_newton_opts* opts,
                            pdes_newton_report* reports);
/* residual of instance inst at its current state (res_host: its numDof doubles) */
int pdes_batch_eval_residual(pdes_batch* batch, int inst, double* res_host);

/* Incremental form of pdes_batch_newton_solve, for serving loops (no
 * counterpart in the reference): runs at most max_passes passes of each
 * group's continuously batched Newton loop (a pass starts due Newton steps,
 * begins a GMRES cycle for every linear solve in progress, applies the Newton
 * updates of the solves that ended and runs the cycle's Arnoldi steps), then
 * returns.  The first call starts the solve; later calls continue it.
 * reports[ninst] (may be null) receives every instance's report so far,
 * active[ninst] (may be null) 1 for instances still being solved, and
 * *all_done (may be null) 1 once no instance has work left.  Between calls
 * every instance in a linear solve sits at a GMRES restart. */
int pdes_batch_advance(pdes_batch* batch, const pdes_newton_opts* opts, int max_passes,
                       pdes_newton_report* reports, int32_t* active, int* all_done);

/* State of one instance between pdes_batch_advance calls, to move an
 * unfinished solve into another batch (a serving loop repacks its running
 * group when requests join or finished ones leave).  With in_solve set the
 * instance is at a GMRES restart of its current Newton step: its iterate dq,
 * right-hand side rhs, iteration count and initial residual norm carry over;
 * the Jacobian and preconditioner are rebuilt from the state, so the solve
 * continues with the bits it would have had in its old batch. */
typedef struct {
  int32_t active;              /* 1: not finished */
  int32_t failed;
  int32_t in_solve;            /* in the middle of a linear solve */
  int32_t newton_steps;        /* Newton steps started */
  double tau, tau_prev_norm;   /* implicit-Euler pseudo-time step and the norm of its last update */
  int32_t gmres_iters;         /* GMRES iterations of the current linear solve */
  int32_t gmres_live;          /* 1 while that solve has not met its tolerance */
  double gmres_beta0;          /* its initial residual norm */
  pdes_newton_report report;
} pdes_batch_inst_state;

/* q, dq, rhs: the instance's numDof doubles (dq and rhs used when in_solve). */
int pdes_batch_export(pdes_batch* batch, int inst, pdes_batch_inst_state* state, double* q, double* dq,
                      double* rhs);
/* Before the batch's first pdes_batch_advance: the instance continues the
 * exported solve instead of starting a new one. */
int pdes_batch_import(pdes_batch* batch, int inst, const pdes_batch_inst_state* state, const double* q,
                      const double* dq, const double* rhs);

#ifdef __cplusplus
}
#endif
#endif
