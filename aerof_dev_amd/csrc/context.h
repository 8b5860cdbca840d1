// context.h -- the per-mesh context behind the C ABI (internal).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "aerofx.h"
#include "device_physics.h"
#include "device_viscous.h"

namespace aerofx {

constexpr int kBlock = 256;       // threads per block of the edge/entry kernels
constexpr int kMaxRows = 64;      // node rows per block (LDS staging bound)
constexpr int kNumXcd = 8;        // MI355X: 8 XCDs, blocks dealt round-robin
constexpr int kPad = kBlock + 1;  // LDS row stride (odd: conflict-free)
// threads (= edges) per block of the flux pass; also the largest number of
// upper edges one node may have (measured at 128^3: 256 -> 0.645 ms,
// 128 -> 0.629 ms, 64 -> 0.618 ms for the edge-flux kernel).  A mesh whose
// nodes all have at most kFluxBlockSmall upper edges gets the small blocks.
constexpr int kFluxBlock = 128;
constexpr int kFluxBlockSmall = 64;
// LDS-staged residual passes: strips of consecutive nodes (gradients) or
// rows (fluxes) whose touched nodes are loaded into LDS once per block.
//   gradient strip: <= kGradStrip nodes (one lane each), <= kGradHalo staged
//                   nodes of 8 doubles (V, X): 61,440 B of LDS
//   flux strip:     <= kFluxStrip edges (one lane each), <= kFluxHalo staged
//                   nodes of 23 doubles (V, X, gradients): 35,328 B of LDS
constexpr int kGradStrip = 128;
constexpr int kGradHalo = 960;
constexpr int kFluxStrip = 256;
constexpr int kFluxHalo = 192;

// Faces of every wall code take the wall flux (FluxFcn.h:270-276): slip,
// adiabatic and isothermal walls, fixed or moving.
__host__ __device__ inline bool is_wall_code(int code) {
  const int a = code < 0 ? -code : code;
  return a == AEROFX_BC_SLIP_WALL || a == AEROFX_BC_ADIABATIC_WALL || a == AEROFX_BC_ISOTHERMAL_WALL;
}

// Block b of a launch runs on XCD b % 8; remap so that each XCD sweeps one
// contiguous range of chunks (its L2 then holds that range's neighbours).
__device__ __forceinline__ int xcd_chunk(int b, int nchunks) {
  const int q = nchunks / kNumXcd, r = nchunks % kNumXcd;
  const int x = b % kNumXcd, k = b / kNumXcd;
  return x * q + min(x, r) + k;
}

// 5x5 block storage interleaved in groups of 64 blocks: element q of block e
// at (e / 64) * 1600 + q * 64 + e % 64, so that 64 lanes reading element q of
// 64 consecutive blocks touch one contiguous 512 B (FP64) segment.
__host__ __device__ inline size_t blk64_base(size_t e) { return (e >> 6) * (64 * 25) + (e & 63); }
constexpr int kBlkStride = 64;
inline size_t blk64_count(size_t nblocks) { return (nblocks + 63) / 64 * 64 * 25; }

// Device-resident mesh tables.  All arrays are laid out for the kernel that
// streams them: contiguous per node row or per edge.
struct DevTables {
  int N, E, F;
  const double *X4;      // [N][4] coordinates padded to 32 B
  const int *fs_mask;    // [N] failsafe first-order nodes (gradients zeroed), or null
  const double *lsq6;    // [N][6] (1/r11, 1/r22, 1/r33, r12/r11, r23/r22, psi)
  const double *invcv;   // [N] 1/ctrlVol
  const double *cv;      // [N] ctrlVol
  // gradient pass: node adjacency, rows sorted ascending (= edge-id order)
  int ngchunks;
  const int *gchunk;     // [ngchunks+1] row starts; a row of more than kBlock entries is a chunk alone
  int nglong;
  const int *glong;      // [nglong] the rows of more than kBlock entries
  const int *rowptr;     // [N+1]
  const int *adj;        // [2E]
  const int *eidx;       // [2E] edge id of each adjacency entry
  const int *mirror;     // [2E] entry (m, n) of the entry (n, m)
  // flux pass A: edges grouped by their lower node i (edge order)
  int nfchunks;
  int fblock;            // edges per flux block: kFluxBlockSmall or kFluxBlock
  const int4 *fchunk4;   // [nfchunks] (first row, end row, first edge, end edge) of each chunk;
                         // a row of more than fblock upper edges is cut into row-less pieces
  const int *eup;        // [N+1] first edge whose lower node is i
  const int2 *eij;       // [E] edge endpoints (i, j), i < j

  const double *normal;  // [E][3] dual-face area vector oriented i -> j
  // flux pass B: per node, lower-edge slots and boundary faces
  const int *lrow;       // [N+1] lower-list pointers
  const int *ledge;      // [E] edges (m, n), m < n, of node n, ascending m
  const int *frow;       // [N+1] face-list pointers
  const int *fidx;       // [3F] faces touching each node, ascending
  const int *fcode;      // [F] boundary code
  const double *fnorm3;  // [F][3] (1/3) face normal
  const double *fnorm;   // [F][3] face normal
  // LDS-staged passes (ngb == 0: the mesh does not fit, per-lane kernels run)
  int ngb;               // gradient strips
  const int4 *gb4;       // [ngb] (first node, end node, first halo entry, end halo entry)
  const int *gown;       // [ngb] halo position of the strip's first node
  const int *ghalo;      // staged node ids, ascending within each strip
  const uint16_t *gloc;  // [2E] halo position of each adjacency entry's neighbour
  int nfb;               // flux strips
  const int4 *fb4;       // [nfb] (first row, end row, first edge, end edge)
  const int2 *fbh;       // [nfb] (first halo entry, end halo entry)
  const int *fhalo;      // staged node ids, ascending within each strip
  const uint32_t *floc;  // [E] halo position of i | halo position of j << 16
};

struct FluxConsts {
  Gas gas;
  double beta;
  double ub_in[5], ub_out[5];
  int low_mach;  // 1: spatial low-Mach preconditioning (ROEFLUX5 and its ROEJAC6)
  LowMach lm;
};

int set_hip_error(hipError_t e, const char *what);
struct Krylov;
struct Adjoint;
struct BlockSolver;
struct Tangent;
struct Ilu;
struct FgWork;

// Element tables of the geometry derivatives (device copies, built on first
// use): the adjoint and the tangent passes sum per-element contributions per
// node or per edge in element order.
struct ElemTables {
  int *tets = nullptr;    // [T][4]
  int *tedge = nullptr;   // [T][6] edge of each local tet edge
  int *faces = nullptr;   // [F][3]
  int *ntrow = nullptr;   // [N+1]
  int *ntslot = nullptr;  // [4T] 4 t + local vertex, ascending t per node
  int *nfslot = nullptr;  // [3F] 3 f + local vertex, in the order of DevTables::fidx
  int *etrow = nullptr;   // [E+1]
  int *etslot = nullptr;  // [6T] 6 t + local edge, ascending t per edge
};

}  // namespace aerofx

struct aerofx_ctx {
  int N = 0, T = 0, F = 0, E = 0;
  aerofx_params p{};
  aerofx::Gas gas{};
  aerofx::FluxConsts fc{};
  hipStream_t own_stream = nullptr, stream = nullptr;
  // host copies of the set-up tables
  std::vector<int> ptr;
  std::vector<double> cv, normals, fnorm, lsq;
  std::vector<int> tets, faces, tet_edges;  // element tables for the adjoint geometry terms
  std::vector<int> fcodes;                  // face codes
  // device
  aerofx::DevTables tab{};
  std::vector<void *> allocs;
  double *d_V = nullptr;     // [N][5] primitive state of the last residual
  double *d_grad = nullptr;  // [N][16] LSQ gradients, [d*5+k]
  double *d_Fm = nullptr;    // [E][5] edge fluxes in edge order
  double *d_P = nullptr;     // [N][5] per-node sums over upper edges
  int *d_err = nullptr;      // [0] reconstruction errors, [1] singular blocks, [2] bad states
  int residual_kernels = AEROFX_RESIDUAL_PER_LANE;  // AEROFX_RESIDUAL_PER_LANE or _STAGED
  // Newton.FailSafe: 0 off, 1 on, 2 always on; d_fs [3N] = tag0, tag1, mask
  // (DistSVec<int,2> tag of Domain and the node tags of DistNodalGrad)
  int failsafe = 0;
  int *d_fs = nullptr;
  bool fs_any = false;  // some node is first order by the failsafe
  int check_final_res = 0;  // KspData checkFinalRes: GMRES reports -999 above its target
  int max_degree = 0;        // most neighbours of a node
  // laminar Navier-Stokes terms (viscous.hip): constants, node types, the
  // wall nodes (node, type) and the viscous time-step scratch
  int viscous = 0;
  aerofx::NsConsts ns{};
  std::vector<int> node_type;
  int nwall = 0, wall_cap = 0;
  int2 *d_wall = nullptr;
  double *d_idtv = nullptr;
  double *d_visc_c = nullptr;  // [4T][5] element contributions of the Galerkin term, in node-slot order
  double *d_adj_tmp = nullptr;  // [N][5] the adjoint with the wall rows masked
  // viscous wall force: element of each face, per node the wall-face element
  // slots (4 f + k), per slot the derivatives [4F][5] / [4F][3]
  int *d_face_elem = nullptr, *d_nfe_row = nullptr, *d_nfe_slot = nullptr;
  double *d_force_u = nullptr, *d_force_x = nullptr;
  int *d_visc_pos = nullptr;   // [4T] node-slot position of element slot 4 t + j
  // device address of this GPU's static scratch array of
  // aerofx_residual_noalloc (looked up at ctx_create; null if that failed,
  // with the reason in noalloc_err)
  char *noalloc_scratch = nullptr;
  std::string noalloc_err;
  int *h_err = nullptr;
  // implicit operators (jacobian.hip) and Krylov workspace (krylov.hip)
  int jac_precision = -1;    // 0: FP64 blocks, 1: FP32 blocks
  void *d_jacD = nullptr, *d_jacB = nullptr, *d_jacDi = nullptr;  // current precision's
  void *jac_store[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};  // [precision][D, B, Di]
  long long jac_version = 0;  // bumped at every assembly
  // off-diagonal blocks in mirror order (slot e holds the block of the
  // mirror entry), so the transposed apply reads them as the forward one
  // does; rebuilt before transposed use when the assembly is newer
  void *jac_mirror[2] = {nullptr, nullptr};
  long long mirror_version = -1;
  int mirror_precision = -1;
  double *d_fjac = nullptr;    // [3F][25] boundary-face Jacobians, per node face slot
  // MatVecProd of aerofx_gmres and the Newton driver (AEROFX_MVP_*) and the
  // scratch of the finite-difference products, allocated on first use:
  // Q, Fa, Fb, F0 [5N] each, reduction partials [kDotBlocks], scalars [16]
  int mvp_type = AEROFX_MVP_H1, mvp_order = 1;
  const double *mvp_Q = nullptr, *mvp_F0 = nullptr, *mvp_dt = nullptr;
  double *d_fd = nullptr;
  double *d_dt = nullptr;
  aerofx::Krylov *ksp = nullptr;
  aerofx::Adjoint *adj = nullptr;
  aerofx::ElemTables *elem = nullptr;
  aerofx::Tangent *tan = nullptr;
  // preconditioner of the Krylov solves (AEROFX_PREC_*), taking effect at the
  // next Jacobian assembly; the RAS / ILU(0) subdomain size in nodes
  int prec_type = AEROFX_PREC_JACOBI;
  int prec_subdomain = 256;
  aerofx::Ilu *ilu = nullptr;
  aerofx::FgWork *fgm = nullptr;  // flexible GMRES workspace (fgmres.hip)
  aerofx::BlockSolver *blk = nullptr;
  int blk_tune[2] = {0, 0};  // autotuned waves per workgroup of the block GEMMs
  // per-kernel timers (the reference's Timer buckets, Timer.h:12-15)
  bool prof = false;
  struct Pending {
    int bucket;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  double timer_ms[AEROFX_NUM_TIMERS] = {0};
  long long timer_n[AEROFX_NUM_TIMERS] = {0};
};

namespace aerofx {

// Brackets one kernel launch with events on the context stream when
// profiling is on.
struct ProfScope {
  aerofx_ctx *c;
  int bucket;
  hipEvent_t a = nullptr;
  ProfScope(aerofx_ctx *ctx, int b);
  ~ProfScope();
};

// device allocation owned by the context, and its early release
int ctx_alloc(aerofx_ctx *c, size_t bytes, void **out);
void ctx_release(aerofx_ctx *c, void *p);

__global__ void k_zero_int(int *p, int n, const int *done);

// laminar Navier-Stokes terms (viscous.hip), no-ops with the terms off:
// R += Galerkin term (/ V if scale); H1 += its blocks, then the wall
// conditions (U: the state, for isothermal walls); F's wall rows; the
// viscous time-step term (idtv = NULL when not used)
int launch_visc_residual(aerofx_ctx *c, const double *V, double *R, bool scale, const int *done);
// the element pass alone into the contribution buffer (the node pass of the
// residual sums it)
int launch_visc_elements(aerofx_ctx *c, const double *V, const int *done);
int launch_visc_jacobian(aerofx_ctx *c, const double *U, int precision);
int launch_bcs_residual(aerofx_ctx *c, const double *U, double *F, const int *done);
int launch_visc_idtv(aerofx_ctx *c, double **idtv);
// dR += (d Galerkin / dU) dU / V at the primitive state V; the wall rows of
// a product y = A p
int launch_visc_tangent(aerofx_ctx *c, const double *V, const double *dU, const double *dX, double *dR,
                        const int *done);
// the viscous wall force: Jf[f] += -(tau n_f) . dir on the wall faces (and
// the slot derivatives when derivs), then dJdU / dJdX += their node sums
int launch_visc_force(aerofx_ctx *c, const double dir[3], double *Jf, bool derivs);
int launch_visc_force_gather(aerofx_ctx *c, double *dJdU, double *dJdX);
// TX[t][12] += the Galerkin part of the element's coordinate adjoint
int launch_visc_adjoint_x(aerofx_ctx *c, const double *lam, double *TX);
int launch_bcs_product(aerofx_ctx *c, const double *U, const double *p, double *y, const int *done);
// Ub += (d Galerkin / dU)^T lam (the residual's 1/V included), at c->d_V
int launch_visc_adjoint(aerofx_ctx *c, const double *lam, double *Ub);
// the wall rows of the transposed Jacobian A^T psi = (dR/dU)^T P psi + B^T psi:
// pre: Ppsi = P psi; post: y += B^T psi
int launch_bcs_adjoint(aerofx_ctx *c, const double *U, const double *psi, double *Ppsi, double *y, bool pre);
// AEROFX_ERR_STATE when the Navier-Stokes terms are on
int refuse_viscous(aerofx_ctx *c, const char *who);

// residual passes (residual.hip); all enqueue on c->stream; every kernel
// returns at once when done != nullptr and *done != 0 (device-resident
// Krylov loops with a matrix-free operator)
int launch_residual(aerofx_ctx *c, const double *U, double *R, bool scale, const int *done = nullptr);
// the same passes with their scratch (primitive state V [N][5], gradients
// [N][16], upper-edge sums P [N][5], edge fluxes Fm [E][5]) given explicitly;
// the error count still goes to c->d_err[0]
struct ResidualBufs {
  double *V, *grad, *Fm, *P;
  int *tag0 = nullptr;  // failsafe: sides of non-physical reconstructions are tagged here
};
// The residual with Newton.FailSafe (Domain.C:864-917): with the failsafe on
// and a non-physical reconstruction in the pass, the nodes are tagged
// (finalizeTags), their gradients zeroed from then on (DistNodalGrad::fix and
// limit) and the pass re-run.  Synchronises; *count = the non-physical
// reconstructions of the pass that stands.
int residual_failsafe(aerofx_ctx *c, const ResidualBufs &b, const double *U, double *R, bool scale, int *count);
// AEROFX_ERR_STATE while failsafe first-order nodes are set (the derivative
// passes differentiate the full second-order residual)
int refuse_failsafe_nodes(aerofx_ctx *c, const char *who);
int launch_residual_with(aerofx_ctx *c, const ResidualBufs &b, const double *U, double *R, bool scale,
                         const int *done = nullptr);
int launch_gradients_v(aerofx_ctx *c, const double *V, double *ddx, double *ddy, double *ddz);
int launch_primitive(aerofx_ctx *c, const double *U, double *V);

// implicit operators (jacobian.hip)
int launch_timestep(aerofx_ctx *c, double cfl, double *dt);
// U: the state of the wall conditions (needed with the Navier-Stokes terms)
int launch_jacobian(aerofx_ctx *c, const double *dt, int precision, const double *U = nullptr);
// trans: apply the transposed operator (adjoint solves)
int launch_h1_apply(aerofx_ctx *c, const double *p, double *y, bool trans = false);
// builds the mirrored off-diagonal blocks of the current assembly if they are
// older than it (enqueued on c->stream; call outside graph capture)
int ensure_mirror(aerofx_ctx *c);
int launch_prec_apply(aerofx_ctx *c, const double *y, double *x, bool trans = false);
// variants that return at once when *done != 0 (device-resident Krylov loops)
int launch_h1_apply_flag(aerofx_ctx *c, const double *p, double *y, const int *done, bool trans = false);
int launch_prec_apply_flag(aerofx_ctx *c, const double *y, double *x, const int *done, bool trans = false);
// V = (src - sum_{k < nsub} h_k Vb_k) * (*scale) then x = P V (the Arnoldi step's
// last Gram-Schmidt update and normalisation fused with the preconditioner)
int launch_prec_norm_flag(aerofx_ctx *c, const double *src, const double *scale, const double *Vb, size_t ldv,
                          int nsub, const double *h, double *V, double *x, const int *done, bool trans = false);
// Gram-Schmidt update pass of Arnoldi step j fused with the preconditioner
// (delayed reorthogonalisation; coefficients in coef, layout in krylov.hip)
constexpr int kMaxBasisCoef = 32;
int launch_dcgs_update(aerofx_ctx *c, double *V, size_t ldv, int j, const double *w, const double *coef,
                       const int *done, const int *jlast, double *x, bool trans);
void free_krylov(aerofx_ctx *c);
void free_adjoint(aerofx_ctx *c);
void free_tangent(aerofx_ctx *c);
// RAS / ILU(0) (ilu.hip): tables for subdomains of B nodes, numericILU of the
// current H1 (zero pivots counted in d_err[1]), and the solve x = M^-1 y
// (returning at once when done != nullptr and *done != 0)
void free_ilu(aerofx_ctx *c);
// flexible GMRES over nrhs right-hand sides rhs, x [nrhs][5N] (fgmres.hip):
// apply(x, y, nq): y_q = A x_q for q < nq; prec(v, z, nq, &inner): z_q ~
// M^-1 v_q (any, possibly varying, approximate inverse), adding its inner
// iterations.  Stops when every ||rhs_q - A x_q|| <= eps ||rhs_q|| or after
// max_outer steps; outer restart mo (<= 31).
struct FgOps {
  std::function<int(const double *, double *, int)> apply;
  std::function<int(const double *, double *, int, int *)> prec;
};
struct FgResult {
  int converged = 0, outer_its = 0, inner_its = 0;
  double res0 = 0.0, res = 0.0;  // the worst right-hand side: ||rhs||, ||rhs - A x||
};
int fgmres_solve(aerofx_ctx *c, int nrhs, const double *rhs, double *x, int mo, double eps, int max_outer,
                 const FgOps &ops, FgResult *res);
void free_fgmres(aerofx_ctx *c);
int ensure_ilu(aerofx_ctx *c, int B);
int launch_ilu_factor(aerofx_ctx *c);
int launch_ilu_solve(aerofx_ctx *c, const double *y, double *x, const int *done);
// c->elem (adjoint.hip)
int ensure_elem_tables(aerofx_ctx *c);
// dR = (dR/dU) dU + (dR/dX) dX at U (tangent.hip; dU, dX may be NULL),
// enqueued on c->stream; the forward residual it runs leaves R (raw flux
// sums) in the tangent workspace
int launch_tangent(aerofx_ctx *c, const double *U, const double *dU, const double *dX, double *dR);
// the exact product operator (tangent.hip): select keeps the forward pass at
// Q; apply gives y = (dR/dU) p + p / dt (dt may be NULL), every kernel
// returning at once when done != nullptr and *done != 0
int launch_h2_select(aerofx_ctx *c, const double *Q);
int launch_h2_apply(aerofx_ctx *c, const double *Q, const double *p, const double *dt, double *y,
                    const int *done);
// (dR/dU)^T lam -> Ub and/or (dR/dX)^T lam -> Xout (adjoint.hip), enqueued on c->stream
int launch_adjoint(aerofx_ctx *c, const double *U, const double *lam, double *Ub, double *Xout);
void free_block(aerofx_ctx *c);
// LINPACK inverse of FP32 5x5 blocks on a given stream (zero-pivot count in
// d_err[1], cleared first)
int launch_invert_f32(aerofx_ctx *c, hipStream_t stream, const float *D, float *Di);

}  // namespace aerofx
