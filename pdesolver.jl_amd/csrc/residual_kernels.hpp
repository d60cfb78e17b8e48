// residual_kernels.hpp -- the fused DG-SBP Euler residual kernel for gfx950.
//
// One launch evaluates the whole of evalResidual (src/solver/euler/euler.jl:
// 111-175) for one operator:
//   * volume term: weak form res += sum_d Q_d^T F_d (euler.jl:628-658), or the
//     entropy-stable split form res_j -= sum_k 2 S_jk F*(q_j, q_k) with the
//     Ismail-Roe flux (euler_funcs.jl:240-288);
//   * face term: R q on both sides, numerical flux, res -= R^T B F
//     (flux.jl:37-125, euler.jl:770-802);
//   * boundary term: R q, BC state, Roe flux, res -= R^T B F (bc.jl:49-328).
//
// Layout: one thread per (element, volume node); a block holds EPB whole
// elements.  The element state, volume fluxes and face contributions are
// staged in LDS.  Faces are evaluated by the element that owns the residual
// rows ("owner pull"): each element recomputes the flux of each of its faces
// with the elementL orientation, so both neighbours obtain bit-identical
// +/- contributions without atomics and without a second pass over HBM.
#pragma once
#include "euler_device.hpp"

namespace pdes {

// element -> face connectivity record
//   x: neighbour element (>= 0) or -1 for a boundary face
//   y: interface index (interior) or boundary face index
//   z: neighbour face | orient << 8 | side << 16  (side 0: this element is elementL)
//   w: BC functor of a boundary face (PDES_BC_*) | instance << 8 (packed batches)
struct ElFace { int32_t x, y, z, w; };

// diagonal-E face slot s = f * Nf + k of an element (k in elementL's face-node
// numbering), built once per context so that the face phase is one dependent
// load away from the neighbour's state:
//   x: global node index (neighbour element * Np + node) of the neighbour's
//      face node, or -1 - row for a boundary face;
//   y: row << 4 | m << 1 | side, with row = interface * Nf + k (nrm_face) or
//      boundary face * Nf + k (nrm_bndry, boundary states), m the face node in
//      this element's numbering and side 1 when this element is elementR.
struct FaceSlot { int32_t x, y; };
__host__ __device__ inline int32_t slot_info(int64_t row, int m, int side) {
  return (int32_t)((row << 4) | (m << 1) | side);
}

struct PhysConst;

struct DevMesh {
  int32_t numEl;
  int32_t metric_per_node;                // 1: dxidx[e][j][x][xi]; 0: affine, dxidx[e][x][xi]
  // work queue: when el_list is set, slot s of the launch processes element
  // el_list[s], s < nlist (packed batches run only the unconverged instances)
  const int32_t* __restrict__ el_list;
  int32_t nlist;
  // packed batches: instance of each element (positivity flags) and the
  // per-instance physics used by boundary states (index in ElFace.w >> 8)
  const int32_t* __restrict__ el_inst;
  const PhysConst* __restrict__ inst_phys;
  const double* __restrict__ dxidx;
  const ElFace* __restrict__ elface;      // [e][f]
  const double* __restrict__ nrm_face;    // [iface][k][dim]
  const double* __restrict__ nrm_bndry;   // [b][k][dim]
  const double* __restrict__ coords_bndry;// [b][k][dim]
  const double* __restrict__ qbc;         // [b][k][ndof] boundary states (functions of the coordinates only)
  const FaceSlot* __restrict__ slot;      // [e][f * Nf + k] (diagonal-E operators)
};

struct DevSBP {
  const double* __restrict__ Q;        // [d][i][j]
  const double* __restrict__ interp;   // [s][k]
  const int32_t* __restrict__ perm;    // [s][f]
  const int32_t* __restrict__ nbrperm; // [k][o]
  const int32_t* __restrict__ nodepos; // [i][f]: stencil position of node i on face f, -1 if none
  const double* __restrict__ wface;    // [k]
  int32_t norient;
  const double* __restrict__ S2c;      // [d][dd][a] = Q_d[a][b] - Q_d[b][a], b = a + dd mod Np, dd <= Np/2
};

__device__ __forceinline__ int32_t launch_count(const DevMesh& m) { return m.el_list ? m.nlist : m.numEl; }
__device__ __forceinline__ int64_t element_of(const DevMesh& m, int64_t slot) {
  return m.el_list ? (int64_t)m.el_list[slot] : slot;
}
// boundary functor and its physics (per instance in a packed batch)
__device__ __forceinline__ int bc_kind(const ElFace& r) { return r.w & 0xff; }
template <class P>
__device__ __forceinline__ P bc_phys(const DevMesh& m, const P& phys, const ElFace& r) {
  if (m.inst_phys) return m.inst_phys[r.w >> 8];
  return phys;
}

// boundary state of face node k of boundary face rec.y, as stored by
// bc_state_kernel (the BC state depends only on the coordinates and the
// instance's physics, never on q)
template <int DIM>
__device__ __forceinline__ void bc_state(const DevMesh& m, const ElFace& rec, int k, int nf, double* qg) {
  const double* s = m.qbc + ((int64_t)rec.y * nf + k) * (DIM + 2);
#pragma unroll
  for (int c = 0; c < DIM + 2; ++c) qg[c] = s[c];
}

// the DIM x DIM metric block [x][xi] of node j of element e
template <int DIM, int NP>
__device__ __forceinline__ const double* metric(const DevMesh& m, int64_t e, int j) {
  return m.dxidx + (m.metric_per_node ? (e * NP + j) : e) * (DIM * DIM);
}

template <int DIM_, int NP_, int NF_, int ST_, bool DIAGE_>
struct OpShape {
  static constexpr int DIM = DIM_, NP = NP_, NF = NF_, ST = ST_;
  static constexpr bool DIAGE = DIAGE_;
  static constexpr int ND = DIM_ + 2, NFACES = DIM_ + 1;
};

// scalar loads/stores for value (double) and directional-derivative (dual<1>) runs
template <class T> struct Io;
template <> struct Io<double> {
  static __device__ __forceinline__ double load(const double* q, const double*, int64_t i) { return q[i]; }
  static __device__ __forceinline__ void store(double* r, double*, int64_t i, double v) { r[i] = v; }
};
template <> struct Io<dual<1>> {
  static __device__ __forceinline__ dual<1> load(const double* q, const double* v, int64_t i) {
    dual<1> a; a.v = q[i]; a.d[0] = v[i]; return a;
  }
  static __device__ __forceinline__ void store(double* r, double* jv, int64_t i, const dual<1>& a) {
    if (r) r[i] = a.v;
    jv[i] = a.d[0];
  }
};

// state and normal of a diagonal-E face slot: the neighbour's face-node state
// (or the boundary state, a constant) and the scaled normal of the face node
template <int DIM, class T>
__device__ __forceinline__ void load_slot(const DevMesh& m, const double* qin, const double* vin, FaceSlot s, T* qn,
                                          double* nrm) {
  const int64_t row = s.y >> 4;
  const double* nr;
  if (s.x >= 0) {
#pragma unroll
    for (int c = 0; c < DIM + 2; ++c) qn[c] = Io<T>::load(qin, vin, (int64_t)s.x * (DIM + 2) + c);
    nr = m.nrm_face + row * DIM;
  } else {
#pragma unroll
    for (int c = 0; c < DIM + 2; ++c) qn[c] = T(m.qbc[row * (DIM + 2) + c]);
    nr = m.nrm_bndry + row * DIM;
  }
#pragma unroll
  for (int x = 0; x < DIM; ++x) nrm[x] = nr[x];
}

enum { VOL_WEAK = 1, VOL_SPLIT = 2 };

// RK4 stage epilogue (rk4.jl:238-319 with pde_post_func :446-460): instead of
// storing res, k = Minv .* res updates the stage accumulator and writes the
// next stage state (or, in the last stage, the new solution in place).
struct RKStage {
  const double* Minv;            // inverse mass matrix, one value per volume node
  const double* q_old;           // aliases q_next in the last stage (in-place update)
  double* acc;
  double* q_next;                // stage state written (stages 1-3) / q_old updated (stage 4)
  double wacc, cstage, h6;
  int32_t stage;                 // 0..3
};

template <class Op, int VOL, int FLUX, int EPB, class T, bool MF = false>
struct ResidualSmem {
  static constexpr int NP = Op::NP, ND = Op::ND, DIM = Op::DIM, NF = Op::NF, NFACES = Op::NFACES;
  // weak form: Q_d[j][i]; split form: 2 S by circulant distance, [d][dd][i]
  double Q[(VOL == VOL_SPLIT) ? DIM * (NP / 2 + 1) * NP : DIM * NP * NP];
  double interp[Op::ST][NF];
  int32_t perm[Op::ST][NFACES];
  int32_t nodepos[NP][NFACES];
  double wface[NF];
  // the element states, for the face phase of dense-face operators (which
  // interpolates over all stencil nodes); diagonal-E face slots read their
  // one own node from global memory instead (an L2 hit: the element was just
  // read), which leaves room for one more workgroup per CU
  static constexpr bool Q_LDS = !Op::DIAGE;
  T q[Q_LDS ? EPB : 1][Q_LDS ? NP : 1][ND];
  // weak form: volume fluxes; split form: IR z-variables and partner contributions
  // (the F and C arrays are planes of one value per (element, node),
  //  [component][element][node], so that the lanes of a pair round, which
  //  read node i + dd, touch consecutive words)
  T F[EPB][(VOL == VOL_WEAK) ? DIM * NP * ND : NP * ND];
  // split form: partner contributions, two buffers used in alternate rounds
  T C[(VOL == VOL_SPLIT) ? 2 : 1][EPB][(VOL == VOL_SPLIT) ? NP * ND : 1];
  // entropy-stable element face integral (dense faces): own IR entropy
  // variables, the neighbours' stencil z and entropy variables per face, and
  // the entropy-conservative sums per (face, own stencil node)
  static constexpr bool ES = (FLUX == PDES_FLUX_ESLF);
  T W[EPB][ES ? NP : 1][ND];
  T Zn[EPB][ES ? NFACES : 1][ES ? Op::ST : 1][ND];
  T Wn[EPB][ES ? NFACES : 1][ES ? Op::ST : 1][ND];
  T H[EPB][ES ? NFACES : 1][ES ? Op::ST : 1][ND];
  // face contributions; the weak form keeps them in the flux planes F, which
  // are dead once the volume term is summed (one more barrier, less LDS); the
  // split form keeps them, where they fit, in the partner buffer that the last
  // pair round does not use (its last reads precede that round's barrier)
  static constexpr bool G_IN_F = (VOL == VOL_WEAK);
  static constexpr bool G_IN_C = (VOL == VOL_SPLIT) && (NP >= NFACES * NF);
  static constexpr int G_C_BUF = ((NP / 2) + 1) & 1;
  static_assert(!G_IN_F || DIM * NP >= NFACES * NF, "face contributions must fit the weak-form flux planes");
  T G[(G_IN_F || G_IN_C) ? 1 : EPB][NFACES][NF][ND];
  // weak form on the matrix cores: the volume term of every (element, node)
  // as planes [component][element][node], laid over the flux planes above
  // the face contributions where they fit there (the flux planes are dead
  // once the product is formed), else in an array of their own
  static constexpr bool RV_IN_F = (DIM - 1) * NP >= NFACES * NF;
  static constexpr int RV_OFF = ND * EPB * NFACES * NF;
  double Rv[(MF && !RV_IN_F) ? ND : 1][(MF && !RV_IN_F) ? EPB * NP : 1];
};

// Weak-form volume term res_i = sum_d sum_j Q_d[j][i] F_d[j] of all EPB
// elements of a block as one product on the FP64 matrix cores,
//   R (NP x ND*EPB) = A (NP x DIM*NP) . B (DIM*NP x ND*EPB),
//   A[i][(d, j)] = Q_d[j][i] (shared by every element of the block),
//   B[(d, j)][(k, el)] = component k of element el's flux in direction d at node j,
// in 16 x 16 output tiles with V_MFMA_F64_16X16X4_F64, k-steps of 4 in the
// order (d, j) of the vector loop (the instruction accumulates its four
// products in order, each fused, so R has the vector loop's bits).  Lane l of
// a wave supplies A[l % 16][l / 16] and B[l / 16][l % 16] of each step and
// holds D[4 r + l / 16][l % 16], r < 4, of the tile.  Every thread of the
// block calls compute() (whole waves: the matrix instruction reads all 64
// lanes); tile t goes to wave t % NW.  store() writes the tiles as planes
// [component][element][node]; the caller puts a barrier between the two when
// the planes overlay the flux planes that compute() reads.
template <int DIM, int NP, int ND, int EPB>
struct WeakVolumeMfma {
  typedef double v4d __attribute__((ext_vector_type(4)));
  static constexpr int PL = EPB * NP;
  static constexpr int KK = DIM * NP, KS = (KK + 3) / 4;
  static constexpr int NCOL = ND * EPB, MT = (NP + 15) / 16, NT = (NCOL + 15) / 16;
  static constexpr int NW = (EPB * NP + 63) / 64;           // waves per block
  static constexpr int TPW = (MT * NT + NW - 1) / NW;       // tiles per wave (at most)
  v4d acc[TPW];

  __device__ __forceinline__ void compute(const double* __restrict__ Q, const double* __restrict__ fpl, int tid) {
    const int lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, q4 = lane >> 4;
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      acc[u] = v4d{0.0, 0.0, 0.0, 0.0};
      const int t = wave + u * NW;
      if (t >= MT * NT) continue;                 // wave-uniform
      const int mt = t / NT, nt = t - mt * NT;
      const int row = mt * 16 + l16;             // A row this lane supplies
      const int col = nt * 16 + l16;             // B column this lane supplies
      const int kc = col / EPB, ec = col - kc * EPB;
      const bool rok = row < NP, cok = col < NCOL;
#pragma unroll 2
      for (int ks = 0; ks < KS; ++ks) {
        const int kk = ks * 4 + q4;
        const int d = kk / NP, j = kk - d * NP;
        double a = 0.0, b = 0.0;
        if (kk < KK) {
          if (rok) a = Q[(d * NP + j) * NP + row];
          if (cok) b = fpl[(d * ND + kc) * PL + ec * NP + j];
        }
        acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[u], 0, 0, 0);
      }
    }
  }

  __device__ __forceinline__ void store(double* __restrict__ rpl, int tid) const {
    const int lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, q4 = lane >> 4;
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int t = wave + u * NW;
      if (t >= MT * NT) continue;
      const int mt = t / NT, nt = t - mt * NT;
      const int col = nt * 16 + l16;
      if (col >= NCOL) continue;
      const int kc = col / EPB, ec = col - kc * EPB;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int node = mt * 16 + 4 * r + q4;
        if (node < NP) rpl[kc * PL + ec * NP + node] = acc[u][r];
      }
    }
  }
};

template <class Op, int VOL, int FLUX, int EPB, class T, bool RK = false, int WAVES = 2, bool MF = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES)))
euler_residual_kernel(DevMesh mesh, DevSBP sbp, PhysConst phys,
                      const double* __restrict__ qin, const double* __restrict__ vin,
                      double* __restrict__ res, double* __restrict__ jv, int32_t* __restrict__ flag,
                      RKStage rk) {
  constexpr int DIM = Op::DIM, NP = Op::NP, NF = Op::NF, ST = Op::ST, ND = Op::ND, NFACES = Op::NFACES;
  using Smem = ResidualSmem<Op, VOL, FLUX, EPB, T, MF>;
  __shared__ Smem sm;
  static_assert(!MF || (VOL == VOL_WEAK && std::is_same<T, double>::value),
                "the matrix-core volume term is the FP64 weak form");

  const int tid = threadIdx.x;
  const int el = tid / NP;
  const int i = tid - el * NP;
  // LDS planes: value c of node n of this block's element el at
  // [c * PL + el * NP + n] (weak form: component k of the flux in parametric
  // direction d as c = d * ND + k; split form: IR parameter vector z)
  constexpr int PL = EPB * NP;
  T* const zpl = &sm.F[0][0];
  const int me = el * NP;
  // face contributions, planes [c][element][face][face node]
  T* const gpl = Smem::G_IN_F ? &sm.F[0][0] : Smem::G_IN_C ? &sm.C[Smem::G_C_BUF][0][0] : &sm.G[0][0][0][0];
  // consecutive element batches on one XCD, so that an element's face
  // neighbours (the adjacent batches, or a mesh row away) were most likely
  // just read into the same L2
  const int64_t eslot = xcd_contiguous_batch(blockIdx.x, gridDim.x) * EPB + el;
  const bool active = (el < EPB) && (eslot < launch_count(mesh));
  const int64_t e = active ? element_of(mesh, eslot) : 0;

  // diagonal-E: the slot record of this thread's first face slot (f, k) =
  // (i / NF, i % NF), then the neighbour's face-node state (or the boundary
  // state) and the normal it names, are issued first, so that these two
  // dependent loads overlap with the element's own loads and the volume phase
  FaceSlot slot0{0, 0};
  T qn0[ND];
  double nrm0[DIM];
  if constexpr (Op::DIAGE) {
    if (active && i < NFACES * NF) {
      slot0 = mesh.slot[e * (NFACES * NF) + i];
      load_slot<DIM, T>(mesh, qin, vin, slot0, qn0, nrm0);
    }
  }

  T qi[ND];
  if (active) {
#pragma unroll
    for (int k = 0; k < ND; ++k) {
      qi[k] = Io<T>::load(qin, vin, (e * NP + i) * ND + k);
      if constexpr (Smem::Q_LDS) sm.q[el][i][k] = qi[k];
    }
    // checkDensity / checkPressure (euler.jl:543-611)
    if (flag) {
      if (!(re(qi[0]) > 0.0) || !(re(pressure<DIM>(phys, qi)) > 0.0)) flag[mesh.el_inst ? mesh.el_inst[e] : 0] = 1;
    }
  }
  double dxi[DIM][DIM];  // dxidx[e][i][x][xi]
  if (active) {
#pragma unroll
    for (int x = 0; x < DIM; ++x)
#pragma unroll
      for (int d = 0; d < DIM; ++d) dxi[x][d] = metric<DIM, NP>(mesh, e, i)[x * DIM + d];
  }
  // stage operator tables
  if constexpr (VOL == VOL_SPLIT) {
    // the split form only needs the antisymmetric part Q - Q^T = 2 S, and
    // only in the order the pair rounds read it: entry [d][dd][a] holds
    // 2 S_d[a][a+dd mod NP], so the lanes of one round read consecutive words
    // (the table is built once per context, sbp.S2c)
    for (int x = tid; x < DIM * (NP / 2 + 1) * NP; x += blockDim.x) sm.Q[x] = sbp.S2c[x];
  } else {
    for (int x = tid; x < DIM * NP * NP; x += blockDim.x) sm.Q[x] = sbp.Q[x];
  }
  for (int x = tid; x < ST * NF; x += blockDim.x) (&sm.interp[0][0])[x] = sbp.interp[x];
  for (int x = tid; x < ST * NFACES; x += blockDim.x) (&sm.perm[0][0])[x] = sbp.perm[x];
  for (int x = tid; x < NP * NFACES; x += blockDim.x) (&sm.nodepos[0][0])[x] = sbp.nodepos[x];
  for (int x = tid; x < NF; x += blockDim.x) sm.wface[x] = sbp.wface[x];


  T ri[ND];
#pragma unroll
  for (int k = 0; k < ND; ++k) ri[k] = 0.0;
  // matrix-core weak form: the product's tiles, and where they are stored
  // (read back into ri after the face phase's barrier)
  WeakVolumeMfma<DIM, NP, ND, EPB> mfma;
  double* rvpl = nullptr;
  if constexpr (MF) rvpl = Smem::RV_IN_F ? &sm.F[0][0] + Smem::RV_OFF : &sm.Rv[0][0];

  if constexpr (VOL == VOL_WEAK) {
    // getEulerFlux: flux in parametric direction d uses nrm = dxidx[d, :, j, e]
    if (active) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        double nrm[DIM];
#pragma unroll
        for (int x = 0; x < DIM; ++x) nrm[x] = dxi[x][d];
        T Fd[ND];
        euler_flux<DIM>(phys, qi, nrm, Fd);
#pragma unroll
        for (int k = 0; k < ND; ++k) zpl[(d * ND + k) * PL + me + i] = Fd[k];
      }
    }
    __syncthreads();
    // weakdifferentiate!(trans=true): res_i += sum_d sum_j Q_d[j][i] F_d[j]
    if constexpr (MF) {
      mfma.compute(sm.Q, zpl, tid);
    } else if (active) {
#pragma unroll
      for (int d = 0; d < DIM; ++d)
#pragma unroll 4
        for (int j = 0; j < NP; ++j) {
          double qji = sm.Q[(d * NP + j) * NP + i];
#pragma unroll
          for (int k = 0; k < ND; ++k) ri[k] = ri[k] + qji * zpl[(d * ND + k) * PL + me + j];
        }
    }
    __syncthreads();   // the face phase writes its contributions into the flux planes
    if constexpr (MF) mfma.store(rvpl, tid);
  } else {
    // split form (calcVolumeIntegralsSplitFormLinear): per-node IR z-variables in LDS,
    // each unordered node pair evaluated once by the node that starts its circulant distance.
    // this node's z stays in registers for the pair rounds
    T zi[ND];
    if (active) {
      ir_zvars<DIM>(phys, qi, zi);
#pragma unroll
      for (int k = 0; k < ND; ++k) zpl[k * PL + me + i] = zi[k];
    }
    __syncthreads();
    // circulant distance dd: node i evaluates the pair (i, i+dd mod NP); the
    // partner's share goes through one LDS slot per node, one round per dd.
    // The slots alternate between two buffers, so one barrier per round
    // suffices: a thread writes a buffer again only after the next round's
    // barrier, which every thread passes after its reads of that buffer.
    constexpr int NPAIR = NP / 2;
    const double* const s2pl = sm.Q;   // 2 S by circulant distance, [d][dd][i]
    for (int dd = 1; dd <= NPAIR; ++dd) {
      const bool half = (2 * dd == NP);            // even NP: distance NP/2 owned by i < NP/2
      T* const cbuf = &sm.C[dd & 1][0][0];
      if (active) {
        T g[ND];
#pragma unroll
        for (int k = 0; k < ND; ++k) g[k] = 0.0;
        if (!(half && i >= NP / 2)) {
          int kk = i + dd;
          if (kk >= NP) kk -= NP;
          // the reference evaluates F(q_j, q_k) with j > k and the direction
          // rows of dxidx at node j.  The IR flux is symmetric in its two
          // states (on the logavg series branch to the last bit: the sums
          // commute and f changes sign only), so this node's own z is the
          // left state and only the partner's is read from LDS.  The
          // direction takes node j's metric.
          const bool i_is_j = i > kk;
          // sum_d 2 S_d[i][kk] F_d is the IR flux in the one direction
          // n = sum_d 2 S_d[i][kk] dxidx[d, :] of node j (the flux is linear in
          // the direction); node j is this node, or the metric is the
          // element's constant one, except for per-node metrics with j = kk
          double s2[DIM], n[DIM];
#pragma unroll
          for (int d = 0; d < DIM; ++d) s2[d] = s2pl[(d * (NPAIR + 1) + dd) * NP + i];
#pragma unroll
          for (int x = 0; x < DIM; ++x) {
            n[x] = s2[0] * dxi[x][0];
#pragma unroll
            for (int d = 1; d < DIM; ++d) n[x] = n[x] + s2[d] * dxi[x][d];
          }
          if (mesh.metric_per_node && !i_is_j) {
            const double* mk = metric<DIM, NP>(mesh, e, kk);
#pragma unroll
            for (int x = 0; x < DIM; ++x) {
              n[x] = s2[0] * mk[x * DIM];
#pragma unroll
              for (int d = 1; d < DIM; ++d) n[x] = n[x] + s2[d] * mk[x * DIM + d];
            }
          }
          T zR[ND];
#pragma unroll
          for (int k = 0; k < ND; ++k) zR[k] = zpl[k * PL + me + kk];
          ir_flux_z<DIM, 1>(phys, zi, zR, n, g);
          // res_i -= 2 S[i][kk] F, and the partner gets -2 S[kk][i] F = +g
#pragma unroll
          for (int k = 0; k < ND; ++k) ri[k] = ri[k] - g[k];
        }
#pragma unroll
        for (int k = 0; k < ND; ++k) cbuf[k * PL + me + i] = g[k];
      }
      __syncthreads();
      if (active) {
        int src = i - dd;
        if (src < 0) src += NP;
#pragma unroll
        for (int k = 0; k < ND; ++k) ri[k] = ri[k] + cbuf[k * PL + me + src];
      }
    }
  }

  // ---- faces: slot (f, k) with k in elementL's face-node numbering
  if constexpr (Op::DIAGE) {
    // SparseFace: R selects face nodes, so each slot needs one own node and
    // one neighbour node
    if (active) {
      for (int slot = i; slot < NFACES * NF; slot += NP) {
        const int f = slot / NF;
        FaceSlot sr;
        T qn[ND];
        double nrm[DIM];
        if (slot == i) {
          sr = slot0;
#pragma unroll
          for (int c = 0; c < ND; ++c) qn[c] = qn0[c];
#pragma unroll
          for (int x = 0; x < DIM; ++x) nrm[x] = nrm0[x];
        } else {
          sr = mesh.slot[e * (NFACES * NF) + slot];
          load_slot<DIM, T>(mesh, qin, vin, sr, qn, nrm);
        }
        // m: the face node in this element's numbering (k as elementL, the
        // neighbour-permuted kR as elementR)
        const int m = (sr.y >> 1) & 7, side = sr.y & 1;
        const int node = sm.perm[m][f];
        T qL[ND], qR[ND], Fk[ND];
        double sign;
        if (sr.x < 0) {
          // boundary: own face node against the boundary state, Roe flux
          sign = -1.0;
#pragma unroll
          for (int c = 0; c < ND; ++c) {
            qL[c] = Io<T>::load(qin, vin, (e * NP + node) * ND + c);
            qR[c] = qn[c];
          }
          roe_flux<DIM>(phys, qL, qR, nrm, Fk);
        } else {
          sign = side == 0 ? -1.0 : 1.0;
#pragma unroll
          for (int c = 0; c < ND; ++c) {
            const T qo = Io<T>::load(qin, vin, (e * NP + node) * ND + c);
            qL[c] = side == 0 ? qo : qn[c];
            qR[c] = side == 0 ? qn[c] : qo;
          }
          if constexpr (FLUX == PDES_FLUX_ROE) {
            roe_flux<DIM>(phys, qL, qR, nrm, Fk);
          } else if constexpr (VOL == VOL_SPLIT) {
            // the own node's z-variables are in the LDS planes already; the
            // neighbour's are computed here by the same function, so the two
            // elements of the face evaluate the same bits
            T zn[ND], zL[ND], zR[ND];
            ir_zvars<DIM>(phys, qn, zn);
#pragma unroll
            for (int c = 0; c < ND; ++c) {
              const T zo = zpl[c * PL + me + node];
              zL[c] = side == 0 ? zo : zn[c];
              zR[c] = side == 0 ? zn[c] : zo;
            }
            irslf_flux_z<DIM>(phys, qL, qR, zL, zR, nrm, Fk);
          } else {
            irslf_flux<DIM>(phys, qL, qR, nrm, Fk);
          }
        }
        const double wgt = sign * sm.wface[m];
#pragma unroll
        for (int c = 0; c < ND; ++c) gpl[((c * EPB + el) * NFACES + f) * NF + m] = wgt * Fk[c];
      }
    }
  } else if constexpr (FLUX == PDES_FLUX_ESLF) {
    // getFaceElementIntegral with ESLFFaceIntegral (flux.jl:132-160,
    // faceElementIntegrals.jl:58-290, 611-630) on dense faces, owner pull:
    // every element evaluates the whole integral of each of its faces with
    // the elementL state first, so both neighbours get bit-identical
    // contributions.  Stage the IR entropy variables of the own nodes and the
    // z / entropy variables of the neighbours' face-stencil nodes.
    if (active) {
      T w[ND];
      to_ir_vars<DIM>(phys, qi, w);
#pragma unroll
      for (int c = 0; c < ND; ++c) sm.W[el][i][c] = w[c];
      for (int t = i; t < NFACES * ST; t += NP) {
        const int f = t / ST, sn = t - f * ST;
        const ElFace rec = mesh.elface[e * NFACES + f];
        if (rec.x < 0) continue;
        const int nodeN = sm.perm[sn][rec.z & 0xff];
        T qn[ND], zn[ND], wn[ND];
#pragma unroll
        for (int c = 0; c < ND; ++c) qn[c] = Io<T>::load(qin, vin, ((int64_t)rec.x * NP + nodeN) * ND + c);
        ir_zvars<DIM>(phys, qn, zn);
        to_ir_vars<DIM>(phys, qn, wn);
#pragma unroll
        for (int c = 0; c < ND; ++c) {
          sm.Zn[el][f][sn][c] = zn[c];
          sm.Wn[el][f][sn][c] = wn[c];
        }
      }
    }
    __syncthreads();
    if (active) {
      // boundary faces: BC flux at the face nodes; interior faces: the
      // entropy penalty (calcEntropyPenaltyIntegral :199-290) at face node k
      for (int slot = i; slot < NFACES * NF; slot += NP) {
        const int f = slot / NF;
        const int k = slot - f * NF;
        const ElFace rec = mesh.elface[e * NFACES + f];
        T Fk[ND];
        int m = k;
        double sign = -1.0;
        if (rec.x < 0) {
          T qL[ND], qR[ND];
#pragma unroll
          for (int c = 0; c < ND; ++c) qL[c] = 0.0;
          for (int sn = 0; sn < ST; ++sn) {
            const double r = sm.interp[sn][k];
            const int node = sm.perm[sn][f];
#pragma unroll
            for (int c = 0; c < ND; ++c) qL[c] = qL[c] + r * sm.q[el][node][c];
          }
          double qg[ND];
          bc_state<DIM>(mesh, rec, k, NF, qg);
#pragma unroll
          for (int c = 0; c < ND; ++c) qR[c] = qg[c];
          roe_flux<DIM>(phys, qL, qR, mesh.nrm_bndry + ((int64_t)rec.y * NF + k) * DIM, Fk);
#pragma unroll
          for (int c = 0; c < ND; ++c) Fk[c] = sm.wface[k] * Fk[c];
        } else {
          const int o = (rec.z >> 8) & 0xff, side = (rec.z >> 16) & 1;
          const int kR = sbp.nbrperm[k * sbp.norient + o];
          T wl[ND], wr[ND];
#pragma unroll
          for (int c = 0; c < ND; ++c) { wl[c] = 0.0; wr[c] = 0.0; }
          for (int sn = 0; sn < ST; ++sn) {
            const double rl = sm.interp[sn][k], rr = sm.interp[sn][kR];
            const int own = sm.perm[sn][f];
#pragma unroll
            for (int c = 0; c < ND; ++c) {
              if (side == 0) {
                wl[c] = wl[c] + rl * sm.W[el][own][c];
                wr[c] = wr[c] + rr * sm.Wn[el][f][sn][c];
              } else {
                wl[c] = wl[c] + rl * sm.Wn[el][f][sn][c];
                wr[c] = wr[c] + rr * sm.W[el][own][c];
              }
            }
          }
          T qa[ND], qlr[ND], dw[ND];
          ir_to_conservative<DIM>(phys, wl, qa);
          ir_to_conservative<DIM>(phys, wr, qlr);
#pragma unroll
          for (int c = 0; c < ND; ++c) {
            qa[c] = 0.5 * (qa[c] + qlr[c]);
            dw[c] = wl[c] - wr[c];
            Fk[c] = 0.0;
          }
          lf_entropy_apply<DIM>(phys, qa, dw, mesh.nrm_face + ((int64_t)rec.y * NF + k) * DIM, sm.wface[k], Fk);
          if (side == 1) { m = kR; sign = 1.0; }
        }
#pragma unroll
        for (int c = 0; c < ND; ++c) gpl[((c * EPB + el) * NFACES + f) * NF + m] = sign * Fk[c];
      }
      // entropy-conservative part (calcECFaceIntegral :58-120): stencil node
      // t of this element on face f against every stencil node of the
      // neighbour, E_ij^d = sum_k R[i][k] R[j][kR] B[k] n_d[k], IR flux in
      // the x, y (, z) directions
      for (int tt = i; tt < NFACES * ST; tt += NP) {
        const int f = tt / ST, t = tt - f * ST;
        const ElFace rec = mesh.elface[e * NFACES + f];
        T h[ND];
#pragma unroll
        for (int c = 0; c < ND; ++c) h[c] = 0.0;
        if (rec.x >= 0) {
          const int o = (rec.z >> 8) & 0xff, side = (rec.z >> 16) & 1;
          const double* nr = mesh.nrm_face + (int64_t)rec.y * NF * DIM;
          const int own = sm.perm[t][f];
          T zo[ND];
#pragma unroll
          for (int c = 0; c < ND; ++c) zo[c] = (&sm.F[0][0])[c * (EPB * NP) + el * NP + own];
          for (int u = 0; u < ST; ++u) {
            const int iL = side == 0 ? t : u, jR = side == 0 ? u : t;
            double E[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) E[d] = 0.0;
            for (int k = 0; k < NF; ++k) {
              const int kRk = sbp.nbrperm[k * sbp.norient + o];
              const double base = sm.interp[iL][k] * sm.interp[jR][kRk] * sm.wface[k];
#pragma unroll
              for (int d = 0; d < DIM; ++d) E[d] += base * nr[k * DIM + d];
            }
            // sum_d E[d] F_d is the IR flux in the direction E
            T zL[ND], zR[ND], Fe[ND];
#pragma unroll
            for (int c = 0; c < ND; ++c) {
              const T zn = sm.Zn[el][f][u][c];
              zL[c] = side == 0 ? zo[c] : zn;
              zR[c] = side == 0 ? zn : zo[c];
            }
            ir_flux_z<DIM, 1>(phys, zL, zR, E, Fe);
#pragma unroll
            for (int c = 0; c < ND; ++c) h[c] = h[c] + Fe[c];
          }
          if (side == 0) {
#pragma unroll
            for (int c = 0; c < ND; ++c) h[c] = -h[c];
          }
        }
#pragma unroll
        for (int c = 0; c < ND; ++c) sm.H[el][f][t][c] = h[c];
      }
    }
  } else
  {
  if (active) {
    for (int slot = i; slot < NFACES * NF; slot += NP) {
      const int f = slot / NF;
      const int k = slot - f * NF;
      const ElFace rec = mesh.elface[e * NFACES + f];
      T qL[ND], qR[ND], Fk[ND];
      int m;
      double sign;
      if (rec.x < 0) {
        // boundary: q_b = R q, state from the BC functor, RoeSolver
        m = k;
        sign = -1.0;
        if constexpr (Op::DIAGE) {
          const int node = sm.perm[k][f];
#pragma unroll
          for (int c = 0; c < ND; ++c) qL[c] = sm.q[el][node][c];
        } else {
#pragma unroll
          for (int c = 0; c < ND; ++c) qL[c] = 0.0;
          for (int s = 0; s < ST; ++s) {
            double r = sm.interp[s][k];
            const int node = sm.perm[s][f];
#pragma unroll
            for (int c = 0; c < ND; ++c) qL[c] = qL[c] + r * sm.q[el][node][c];
          }
        }
        const int64_t b = rec.y;
        double qg[ND];
        bc_state<DIM>(mesh, rec, k, NF, qg);
#pragma unroll
        for (int c = 0; c < ND; ++c) qR[c] = qg[c];
        roe_flux<DIM>(phys, qL, qR, mesh.nrm_bndry + (b * NF + k) * DIM, Fk);
      } else {
        const int64_t nb = rec.x;
        const int fN = rec.z & 0xff;
        const int o = (rec.z >> 8) & 0xff;
        const int side = (rec.z >> 16) & 1;
        const int kR = sbp.nbrperm[k * sbp.norient + o];
#pragma unroll
        for (int c = 0; c < ND; ++c) { qL[c] = 0.0; qR[c] = 0.0; }
        if constexpr (Op::DIAGE) {
          // SparseFace: R selects the face nodes
          const int nodeL = sm.perm[k][side == 0 ? f : fN];
          const int nodeR = sm.perm[kR][side == 0 ? fN : f];
          if (side == 0) {
            m = k;
            sign = -1.0;
#pragma unroll
            for (int c = 0; c < ND; ++c) {
              qL[c] = sm.q[el][nodeL][c];
              qR[c] = Io<T>::load(qin, vin, (nb * NP + nodeR) * ND + c);
            }
          } else {
            m = kR;
            sign = 1.0;
#pragma unroll
            for (int c = 0; c < ND; ++c) {
              qL[c] = Io<T>::load(qin, vin, (nb * NP + nodeL) * ND + c);
              qR[c] = sm.q[el][nodeR][c];
            }
          }
        } else if (side == 0) {
          // this element is elementL
          m = k;
          sign = -1.0;
          for (int s = 0; s < ST; ++s) {
            const double rl = sm.interp[s][k], rr = sm.interp[s][kR];
            const int nodeL = sm.perm[s][f], nodeR = sm.perm[s][fN];
#pragma unroll
            for (int c = 0; c < ND; ++c) {
              qL[c] = qL[c] + rl * sm.q[el][nodeL][c];
              qR[c] = qR[c] + rr * Io<T>::load(qin, vin, (nb * NP + nodeR) * ND + c);
            }
          }
        } else {
          m = kR;
          sign = 1.0;
          for (int s = 0; s < ST; ++s) {
            const double rl = sm.interp[s][k], rr = sm.interp[s][kR];
            const int nodeL = sm.perm[s][fN], nodeR = sm.perm[s][f];
#pragma unroll
            for (int c = 0; c < ND; ++c) {
              qL[c] = qL[c] + rl * Io<T>::load(qin, vin, (nb * NP + nodeL) * ND + c);
              qR[c] = qR[c] + rr * sm.q[el][nodeR][c];
            }
          }
        }
        const double* nrm = mesh.nrm_face + ((int64_t)rec.y * NF + k) * DIM;
        if constexpr (FLUX == PDES_FLUX_ROE) roe_flux<DIM>(phys, qL, qR, nrm, Fk);
        else irslf_flux<DIM>(phys, qL, qR, nrm, Fk);
      }
      const double wgt = sign * sm.wface[m];
#pragma unroll
      for (int c = 0; c < ND; ++c) gpl[((c * EPB + el) * NFACES + f) * NF + m] = wgt * Fk[c];
    }
  }
  }
  __syncthreads();
  if (active) {
    if constexpr (MF) {
#pragma unroll
      for (int k = 0; k < ND; ++k) ri[k] = rvpl[k * PL + me + i];
    }
    // res_i += sum_f sum_m interp[s(i,f)][m] G[f][m]
#pragma unroll
    for (int f = 0; f < NFACES; ++f) {
      const int s = sm.nodepos[i][f];
      if (s < 0) continue;
      if constexpr (Op::DIAGE) {
#pragma unroll
        for (int c = 0; c < ND; ++c) ri[c] = ri[c] + gpl[((c * EPB + el) * NFACES + f) * NF + s];
      } else {
#pragma unroll
        for (int mm = 0; mm < NF; ++mm) {
          const double r = sm.interp[s][mm];
#pragma unroll
          for (int c = 0; c < ND; ++c) ri[c] = ri[c] + r * gpl[((c * EPB + el) * NFACES + f) * NF + mm];
        }
        if constexpr (FLUX == PDES_FLUX_ESLF) {
#pragma unroll
          for (int c = 0; c < ND; ++c) ri[c] = ri[c] + sm.H[el][f][s][c];
        }
      }
    }
    if constexpr (RK) {
      const int64_t base = (e * NP + i) * ND;
      const double minv = rk.Minv[e * NP + i];
#pragma unroll
      for (int c = 0; c < ND; ++c) {
        const double k = re(ri[c]) * minv;
        // stage 0: q_old is this stage's input, still staged in LDS where
        // the operator stages it (reading it there lets the entry registers
        // qi die after the volume phase)
        double qo;
        if constexpr (Smem::Q_LDS) qo = rk.stage == 0 ? re(sm.q[el][i][c]) : rk.q_old[base + c];
        else qo = rk.q_old[base + c];
        if (rk.stage == 3) {
          rk.q_next[base + c] = qo + rk.h6 * (rk.acc[base + c] + rk.wacc * k);
        } else {
          const double a = rk.stage == 0 ? rk.wacc * k : rk.acc[base + c] + rk.wacc * k;
          rk.acc[base + c] = a;
          rk.q_next[base + c] = qo + rk.cstage * k;
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < ND; ++c) Io<T>::store(res, jv, (e * NP + i) * ND + c, ri[c]);
    }
  }
}

// boundary states of every boundary face node: isentropicVortexBC
// (bc.jl:554-567, calcIsentropicVortex) or FreeStreamBC (bc.jl:1573-1586),
// with the instance's physics; one thread per (element face, face node)
template <int DIM>
__global__ void bc_state_kernel(DevMesh mesh, PhysConst phys, int nfaces, int nf, double* __restrict__ qbc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)mesh.numEl * nfaces * nf) return;
  const int64_t ef = t / nf;
  const int k = (int)(t - ef * nf);
  const ElFace rec = mesh.elface[ef];
  if (rec.x >= 0) return;
  double qg[DIM + 2];
  if constexpr (DIM == 2) {
    if (bc_kind(rec) == PDES_BC_ISENTROPIC_VORTEX)
      isentropic_vortex(bc_phys(mesh, phys, rec), mesh.coords_bndry + ((int64_t)rec.y * nf + k) * DIM, qg);
    else
      free_stream<DIM>(bc_phys(mesh, phys, rec), qg);
  } else {
    free_stream<DIM>(bc_phys(mesh, phys, rec), qg);
  }
  double* out = qbc + ((int64_t)rec.y * nf + k) * (DIM + 2);
#pragma unroll
  for (int c = 0; c < DIM + 2; ++c) out[c] = qg[c];
}

}  // namespace pdes
