/*
 * aerofx.h -- C ABI of the MI355X-native edge-based finite-volume operator.
 *
 * One opaque context per (mesh, parameters).  All compute entry points take
 * DEVICE pointers, enqueue work on the context's HIP stream and return
 * without synchronising unless stated.  Every entry point returns an int:
 *   0   success
 *   >0  count of non-physical reconstructed states (the reference's ierr)
 *   <0  error (AEROFX_ERR_*); aerofx_last_error() describes it.
 * No entry point throws.  Contexts are not re-entrant: calls on one context
 * must be serialised by the caller (as the reference's SpaceOperator is).
 *
 * Array layouts follow the reference's SVec: row-major [n][dim] doubles.
 * Node, edge and face numbering is the caller's; the edge table is the
 * reference's (EdgeCore.C:55-119): unique (i,j), i<j, sorted
 * lexicographically.
 */
#ifndef AEROFX_H
#define AEROFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AEROFX_OK 0
#define AEROFX_ERR_ARG (-1)
#define AEROFX_ERR_HIP (-2)
#define AEROFX_ERR_MESH (-3)
#define AEROFX_ERR_NOMEM (-4)
#define AEROFX_ERR_STATE (-5)

/* Boundary codes (BcDef.h:4-28). */
#define AEROFX_BC_INTERNAL 0
#define AEROFX_BC_ISOTHERMAL_WALL 1 /* no-slip wall at a fixed temperature (viscous) */
#define AEROFX_BC_SLIP_WALL 2
#define AEROFX_BC_ADIABATIC_WALL 3  /* no-slip adiabatic wall (viscous) */
#define AEROFX_BC_INLET 4
#define AEROFX_BC_OUTLET 5

typedef struct aerofx_ctx aerofx_ctx;

/* Mesh handed to aerofx_ctx_create: tets [T][4] (positively oriented),
 * boundary triangles [F][3] with outward normal 0.5*(x2-x0)^(x1-x0), one
 * code per triangle.  Host pointers; copied. */
typedef struct {
  int num_nodes, num_tets, num_faces;
  const double *X;
  const int *tets;
  const int *faces;
  const int *face_codes;
} aerofx_mesh;

/* Flow and scheme parameters (IoData defaults, IoDataCore.C:2418-2429,
 * 2662; DistBcData.C:652-706 for the freestream). Angles in radians. */
typedef struct {
  double gamma;        /* specific heat ratio (1.4) */
  double pstiff;       /* pressure constant (0: perfect gas) */
  double beta;         /* MUSCL beta (1/3) */
  int second_order;    /* 1: LSQ gradients + Van Albada; 0: first order */
  double rho_inf, p_inf, mach, alpha, yaw;
} aerofx_params;

/* Synthetic Kuhn-split box (no reference counterpart: the reference reads
 * pre-partitioned SOWER meshes, GeoSource.C:246-356). */
typedef struct {
  int nx, ny, nz;            /* hexes per axis */
  double lo[3], hi[3];       /* extent */
  double jitter;             /* interior nodes moved by U(-j,j)*h per axis */
  uint64_t seed;             /* Philox4x32-10 key */
  int face_code[6];          /* x-, x+, y-, y+, z-, z+ */
  int nbumps;                /* Gaussian bumps on the z- wall, blended */
  const double *bump_amp, *bump_center, *bump_width;  /* x-positions */
} aerofx_box_spec;

const char *aerofx_last_error(void);
int aerofx_version(void);

/* ---- host-side mesh preparation (no GPU needed) ------------------------ */
int aerofx_mesh_box_sizes(const aerofx_box_spec *spec, int *num_nodes,
                          int *num_tets, int *num_faces);
int aerofx_mesh_box(const aerofx_box_spec *spec, double *X, int *tets,
                    int *faces, int *face_codes);
/* Shape derivative of the synthetic box's coordinates (no reference
 * counterpart: the design variables of the sensitivity workloads): dX/dp
 * [N][3] for bump `bump`'s amplitude (param 0) or centre (param 1), in the
 * generator's node order.  The bump adds a exp(-((x - c)/w)^2)(1 - zeta) to
 * z, zeta the node's height fraction before the bumps. */
int aerofx_mesh_box_bump_derivative(const aerofx_box_spec *spec, int bump, int param, double *dX);
/* Edge count and table (EdgeSet::find/createPointers, EdgeCore.C:55-119):
 * ptr must hold 2*num_edges ints (query with ptr == NULL). */
int aerofx_mesh_edges(int num_nodes, int num_tets, const int *tets,
                      int *num_edges, int *ptr);
/* Reference RCM (Connectivity::renumByComponent with renumRCM,
 * ConnectivityCore.C:311-635): renum[old] = new. */
int aerofx_mesh_rcm(int num_nodes, int num_tets, const int *tets, int *renum);
/* Relabels a mesh in place by renum[old] = new. */
int aerofx_mesh_renumber(int num_nodes, int num_tets, int num_faces,
                         const int *renum, double *X, int *tets, int *faces);
/* Edge colouring rule (no reference counterpart): returns colours. */
int aerofx_mesh_color_edges(int num_nodes, int num_edges, const int *ptr,
                            int *color, int *num_colors);
/* The same two orderings computed on the GPU, bit-identical to the host
 * versions: aerofx_mesh_color_edges_gpu runs Jones-Plassmann rounds whose
 * result is the greedy colouring in decreasing key order; aerofx_mesh_rcm_gpu
 * builds each breadth-first level structure of Connectivity::rootLS /
 * findPseudoDiam / renumRCM (ConnectivityCore.C:324-635) on the GPU, in the
 * sequential append order.  Host arrays in and out. */
int aerofx_mesh_color_edges_gpu(int num_nodes, int num_edges, const int *ptr,
                                int *color);
int aerofx_mesh_rcm_gpu(int num_nodes, int num_tets, const int *tets, int *renum);

/* Synthetic states: conservative U[N][5] on host.
 *  kind 0: uniform freestream
 *  kind 1: freestream with multiplicative Gaussian noise of std `amp` on
 *          density and pressure (Philox Box-Muller, key = seed)
 *  kind 2: freestream with density * (1 + amp*sin(2pi(x+a))sin(2pi(y+b))
 *          sin(2pi(z+c))), phases a,b,c from Philox(seed) */
int aerofx_state_init(const aerofx_params *p, int kind, double amp,
                      uint64_t seed, int num_nodes, const double *X,
                      double *U);

/* ---- context ----------------------------------------------------------- */
int aerofx_ctx_create(const aerofx_mesh *mesh, const aerofx_params *params,
                      aerofx_ctx **out);
int aerofx_ctx_destroy(aerofx_ctx *ctx);
int aerofx_ctx_sizes(const aerofx_ctx *ctx, int *num_nodes, int *num_edges,
                     int *num_faces);
/* Enqueues all later work on `stream` (a hipStream_t; NULL is the null
 * stream), or on the context's own stream if use_context_stream != 0. */
int aerofx_ctx_set_stream(aerofx_ctx *ctx, void *stream, int use_context_stream);
int aerofx_ctx_sync(aerofx_ctx *ctx);
/* Host copies of the set-up tables (for parity checks): edge table [E][2],
 * control volumes [N], edge normals [E][3], face normals [F][3] (full
 * triangle normal, FaceTria::computeNormal), LSQ factors [N][6]. Any may be
 * NULL. */
int aerofx_ctx_tables(const aerofx_ctx *ctx, int *ptr, double *ctrl_vol,
                      double *edge_normals, double *face_normals,
                      double *lsq);

/* Per-kernel timers, named after the reference's Timer buckets
 * (Timer.h:12-15).  When enabled, each kernel launch is bracketed by HIP
 * events on the context stream; aerofx_ctx_timers synchronises and returns
 * accumulated milliseconds and launch counts per bucket. Enabling resets. */
#define AEROFX_TIMER_NODAL_GRAD 0   /* LSQ gradient pass */
#define AEROFX_TIMER_FV_TERM 1      /* edge flux pass */
#define AEROFX_TIMER_FV_NODES 2     /* node assembly of fluxes + faces */
#define AEROFX_TIMER_FV_JAC 3       /* Jacobian assembly */
#define AEROFX_TIMER_PREC_SETUP 4   /* block-Jacobi inversion */
#define AEROFX_TIMER_JAC_APPLY 5    /* Jacobian-vector product */
#define AEROFX_TIMER_PREC_APPLY 6   /* preconditioner apply */
#define AEROFX_TIMER_GRAM_SCHMIDT 7 /* Krylov orthogonalisation */
#define AEROFX_TIMER_FE_TERM 8      /* Navier-Stokes Galerkin term */
#define AEROFX_TIMER_FE_JAC 9       /* its H1 blocks */
#define AEROFX_NUM_TIMERS 10
int aerofx_ctx_profile(aerofx_ctx *ctx, int enable);
int aerofx_ctx_timers(aerofx_ctx *ctx, double *ms, long long *counts);

/* Kernel family of the residual passes (no counterpart in the reference; the
 * results are bit-identical either way):
 *   AEROFX_RESIDUAL_PER_LANE  one lane per node or edge gathering from HBM
 *                             (the default, and what AEROFX_RESIDUAL_AUTO
 *                             selects: the faster family on the 128^3 box)
 *   AEROFX_RESIDUAL_STAGED    LDS-staged strips, when the mesh fits them
 *                             (each node with its neighbours within 960
 *                             staged nodes; each node with at most 256 upper
 *                             edges, touching at most 191 other nodes through
 *                             them); otherwise per-lane
 * Returns the family the next residual will use (AEROFX_RESIDUAL_STAGED or
 * AEROFX_RESIDUAL_PER_LANE), or < 0. */
#define AEROFX_RESIDUAL_AUTO 0
#define AEROFX_RESIDUAL_PER_LANE 1
#define AEROFX_RESIDUAL_STAGED 2
int aerofx_ctx_set_residual_kernels(aerofx_ctx *ctx, int family);

/* Spatial low-Mach preconditioning: SpatialLowMachPrec (LowMachPrec.h) with
 * Problem.Prec = Preconditioned and no time preconditioner (prec tag 1).
 * The interior edge flux becomes the Roe-Turkel flux ROEFLUX5
 * (FluxFcnSGApprJacRoeEuler3D::compute, FluxFcnDescSG.C:125-139) with
 * beta = min((1 + sqrt(irey)) max(slope M_roe, min_mach)
 *            + shock_reducer |p_i - p_j| / ((p_i + p_j) |X_j - X_i|), cutoff)
 * per edge (length |X_j - X_i|, Edge.C:400; irey the edge's 1/Re, 0 for
 * Euler flow), and the H1 edge blocks its ROEJAC6 (roejacappr3Dgas,
 * FluxFcnDescSG.C:157-175, Edge.C:4472-4526).  The time step keeps beta = 1
 * (TimeLowMachPrec without Ts.Prec).  lm = NULL turns it off.  The residual
 * runs the per-lane kernels while it is on; the forward and reverse
 * derivative passes (aerofx_residual_derivative, aerofx_adjoint_residual,
 * AEROFX_MVP_H2, aerofx_adjoint_solve, aerofx_direct_solve) return
 * AEROFX_ERR_STATE: they differentiate ROEFLUX7. */
typedef struct {
  double min_mach;       /* Preconditioner.Mach (beta floor) */
  double slope;          /* Preconditioner.k (slope on the local Mach number) */
  double cutoff;         /* Preconditioner.CutOffMach (beta ceiling) */
  double shock_reducer;  /* Preconditioner.ShockReducer */
  double irey;           /* 1/Re of the edges (0: inviscid) */
} aerofx_low_mach;
/* (not with the Navier-Stokes terms of aerofx_ctx_set_viscous) */
int aerofx_ctx_set_low_mach(aerofx_ctx *ctx, const aerofx_low_mach *lm);

/* Laminar Navier-Stokes terms (Equations.Type = NavierStokes without a
 * turbulence model, full wall integration; the reference's
 * FemEquationTermNS + BcFcnNS, SpaceOperator.C:76-123, 246-265):
 *   - the residual (aerofx_residual*) adds the P1 Galerkin term of every
 *     tetrahedron, vol (r . grad phi_j) with r the viscous stress and heat
 *     flux at the element's mean state (ElemTet::computeGalerkinTerm,
 *     ElemTet.C:24-120; FemEquationTermNS::computeVolumeTerm,
 *     FemEquationTermDesc.C:98-158), before the 1/V scaling;
 *   - aerofx_jacobian_setup adds its exact Jacobian blocks to H1
 *     (ElemTet::computeJacobianGalerkinTerm, ElemTet.C:1912-1957;
 *     FemEquationTermNS::computeJacobianVolumeTerm, FemEquationTermDesc.C:
 *     338-449) and then applies the wall conditions to H1
 *     (SpaceOperator::applyBCsToJacobian, SpaceOperator.C:2421-2433);
 *   - the Newton driver and the finite-difference products apply the wall
 *     conditions to every F(Q) (ImplicitTsDesc::computeFunction,
 *     ImplicitTsDesc.C:125-141; MatVecProd.C:88-118, 262-300);
 *   - the time step adds viscous_cst times the elemental viscous term
 *     (ElemSet::computeTimeStep, Elem.C:25-83; Domain.C:91).
 * Walls are the faces of code AEROFX_BC_ADIABATIC_WALL (no-slip) and
 * AEROFX_BC_ISOTHERMAL_WALL (no-slip, temperature wall_temperature); a node
 * takes the code of the highest-priority face touching it
 * (Domain::setNodeType, DomainCore.C:604-670).  Walls are fixed (zero wall
 * velocity, so the adiabatic-wall face term vanishes).  The transport laws
 * are ViscoFcn.h's and ThermalCondFcn.h's with the AERO-F non-dimensional
 * scaling (alpha = gamma (gamma - 1) M^2).  v = NULL turns the terms off.
 * The transposed product aerofx_adjoint_residual(dRdU_T_lambda) includes
 * the Galerkin term's transposed element derivative, and
 * aerofx_adjoint_solve solves with the wall rows (A = P dR/dU + B, A^T psi =
 * (dR/dU)^T P psi + B^T psi); aerofx_residual_derivative (dU and dX), the
 * exact products (AEROFX_MVP_H2, with the wall rows) and dRdX_T_lambda
 * include its element derivatives; aerofx_direct_solve solves
 * (P dR/dU + B) x = P rhs (the wall rows of rhs become the derivative of the
 * fixed-wall conditions, zero).  Not with the low-Mach preconditioner. */
#define AEROFX_VISC_CONSTANT 0
#define AEROFX_VISC_SUTHERLAND 1
#define AEROFX_VISC_PRANDTL 2
#define AEROFX_COND_CONSTANT_PRANDTL 0
#define AEROFX_COND_CONSTANT 1
typedef struct {
  double reynolds;          /* ReferenceState.Reynolds (ref.reynolds_mu) */
  int viscosity_model;      /* AEROFX_VISC_* (ViscosityModel.Type) */
  double bulk_viscosity;    /* constant model: lambda = bulk - 2/3 mu */
  double sutherland_ts;     /* SutherlandReferenceTemperature / ReferenceState.Temperature */
  int conductivity_model;   /* AEROFX_COND_* (ThermalConductivityModel.Type) */
  double prandtl;           /* constant-Prandtl model */
  double conductivity;      /* constant model (non-dimensional) */
  double wall_temperature;  /* isothermal walls: non-dimensional T = p / ((gamma - 1) rho) */
  double viscous_cst;       /* Time.ViscosityParameter (0) */
} aerofx_viscous;
int aerofx_ctx_set_viscous(aerofx_ctx *ctx, const aerofx_viscous *v);
/* Node types of the wall conditions [N] (the face codes by priority, 0
 * inside); needs the viscous terms on. */
int aerofx_ctx_node_types(const aerofx_ctx *ctx, int *types);
/* SpaceOperator::applyBCsToResidual (SpaceOperator.C:2389-2395;
 * BcFcnNS::applyToResidualTerm, BcFcnCore.C:230-245): on the wall nodes
 * F[1..3] = U[1..3] (isothermal walls also F[4] = (E - rho T_w) rho -
 * |m|^2 / 2).  Device pointers [N][5]; nothing without the viscous terms. */
int aerofx_apply_bcs_residual(aerofx_ctx *ctx, const double *U, double *F);
/* SpaceOperator::applyBCsToSolutionVector (BcFcnCore.C:212-226): zero
 * momentum on the wall nodes (and the wall temperature on isothermal
 * walls).  Device pointer [N][5]. */
int aerofx_apply_bcs_solution(aerofx_ctx *ctx, double *U);

/* Newton.FailSafe (NewtonData::failsafe, IoData.C:10-35): 0 Off (default),
 * 1 On, 2 AlwaysOn; second-order schemes only, as the reference turns it off
 * for constant reconstruction (IoDataCore.C:6010-6015).  With it on, a
 * synchronous residual whose pass meets a non-physical reconstruction tags
 * the failing sides, finalizes the tags over the edges (finalizeTags: each
 * node takes tag0[i] + tag0[j] of its last edge), makes the tagged nodes
 * first order from then on (their gradients zeroed, DistNodalGrad::fix and
 * limit) and re-runs the pass (Domain::computeFiniteVolumeTerm,
 * Domain.C:864-917); On clears the tags after each fix, AlwaysOn keeps them.
 * aerofx_steady_solve also redoes a Newton step whose update leaves
 * non-physical states (NewtonSolver.h:193-215, up to 5 redos per step).
 * First-order nodes stay until aerofx_ctx_reset_failsafe (the Newton driver
 * resets them after a step that redid, with FailSafe = On); while any is set,
 * the derivative passes return AEROFX_ERR_STATE.  The asynchronous residual
 * (aerofx_residual_ex flag 2) and the residuals inside captured Krylov cycles
 * use the first-order nodes but do not tag. */
int aerofx_ctx_set_failsafe(aerofx_ctx *ctx, int mode);
/* KspData checkFinalRes for aerofx_gmres / aerofx_gmres_transpose: after
 * the solve, ||b - A x|| is recomputed and an iteration count of -999 is
 * returned when it misses the target (KspSolver.C:455-469). */
int aerofx_ctx_set_krylov_check(aerofx_ctx *ctx, int check_final_res);
int aerofx_ctx_reset_failsafe(aerofx_ctx *ctx);
/* tags: host [3][N] = tag0 (failing sides), tag1 (finalized), first-order mask */
int aerofx_ctx_failsafe_nodes(aerofx_ctx *ctx, int *tags);

/* ---- SpaceOperator<5> -------------------------------------------------- */
/* SpaceOperator<5>::computeResidual (SpaceOperator.C:524-634), NONDESCRIPTOR
 * form: R = (sum of edge and face fluxes)/V.  U, R: device [N][5].
 * Returns 0, or the count of non-physical reconstructions (>0; R is then not
 * meaningful: the reference aborts, Domain.C:876-882), or <0.
 * Synchronises to read the error count. */
int aerofx_residual(aerofx_ctx *ctx, const double *U, double *R);
/* As aerofx_residual; flags bit 0: skip the 1/V scaling (raw flux sums);
 * bit 1: do not synchronise or read the error count (returns 0; use
 * aerofx_error_count afterwards). */
int aerofx_residual_ex(aerofx_ctx *ctx, const double *U, double *R, int flags);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  As aerofx_residual_ex (same flags, same kernels, results
 * bit-identical), but the passes' scratch (V, gradients, upper-edge sums,
 * edge fluxes: about 8 (26 N + 5 E) bytes) lives in one static device array
 * of 64 MiB per GPU instead of the context's buffers, and the call allocates
 * nothing.  A mesh that needs more than 64 MiB is rejected with
 * AEROFX_ERR_ARG (about 137k nodes of a Kuhn mesh).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (several contexts,
 * several streams) race, so the caller serialises them.  The error count
 * goes to the context as for aerofx_residual_ex. */
int aerofx_residual_noalloc(aerofx_ctx *ctx, const double *U, double *R, int flags);
int aerofx_error_count(aerofx_ctx *ctx, int *count);
/* DistNodalGrad::compute with LeastSquares (DistNodalGrad.C:790-806,
 * SubDomain.C:297-374) on primitive V [N][5]: ddx, ddy, ddz [N][5]. */
int aerofx_gradients(aerofx_ctx *ctx, const double *V, double *ddx,
                     double *ddy, double *ddz);
/* VarFcnSGEuler::conservativeToPrimitive over all nodes. */
int aerofx_primitive(aerofx_ctx *ctx, const double *U, double *V);

/* ---- MatVecProd / KspPrec / KspSolver ---------------------------------- */
/* Local pseudo-time step dt_i = cfl V_i / sum(-min(lambda, 0) S)
 * (Edge.C:87-143, Face.C:242-290, Domain.C:32-100) at state U: dt [N]. */
int aerofx_timestep(aerofx_ctx *ctx, const double *U, double cfl, double *dt);
/* MatVecProdH1::evaluate + JacobiPrec::setup (MatVecProd.C:1075-1100,
 * KspPrec.C:9-124): assembles the first-order Roe Jacobian H1 at U
 * (SpaceOperator.C:2088-2150) plus alpha/dt_i on the diagonal (dt may be
 * NULL: no time term), then inverts its 5x5 diagonal blocks (LINPACK,
 * DiagMatrix.C:44-82).  precision 0: FP64 blocks; 1: FP32 blocks (the
 * mixed-precision JVP and preconditioner). */
int aerofx_jacobian_setup(aerofx_ctx *ctx, const double *U, const double *dt,
                          int precision);
/* Host copies of the assembled H1: D [N][25] and, per edge, (A_ij, A_ji)
 * [E][2][25] (the reference's block order, SubDomain.C:3929-3947). */
int aerofx_jacobian_blocks(aerofx_ctx *ctx, double *D, double *Aedge);
/* MatVecProdH1::apply (MatVecProd.C:1169-1187): Jp = H1 p. */
int aerofx_jvp(aerofx_ctx *ctx, const double *p, double *Jp);
/* JacobiPrec::apply (KspPrec.C:87-111): x = D^-1 y. */
int aerofx_prec_apply(aerofx_ctx *ctx, const double *y, double *x);
/* MatVecProdH1::applyTranspose (MatVecProd.C:1189-1206,
 * SubDomain.C:3958-3990): JTp = H1^T p, per node in the reference's order.
 * The first transposed use after an assembly (here, in
 * aerofx_gmres_transpose or aerofx_adjoint_solve) copies the off-diagonal
 * blocks into mirror order, into one more buffer of the blocks' size in the
 * current precision, allocated on first use and kept with the context. */
int aerofx_jvp_transpose(aerofx_ctx *ctx, const double *p, double *JTp);
/* Transposed block-Jacobi: x = (D^-1)^T y (the preconditioner of the
 * transposed solve, KspSolver::solveT, KspSolver.C:754). */
int aerofx_prec_apply_transpose(aerofx_ctx *ctx, const double *y, double *x);
/* MatVecProdFD::evaluate + apply + computeEpsilon (MatVecProd.C:88-118,
 * 262-372, 879-929): Jp = (F(U + eps p) - F(U)) / eps (order 1) or
 * (F(U + eps p) - F(U - eps p)) / (2 eps) (order 2), F = R + (Q - U)/dt
 * (TimeState::add_dAW_dt, backward Euler, Un = U; dt NULL: R only),
 * eps = sum(1e-6 |U| + 1e-6) / (5N ||p||) (1e-6 when ||p|| < 1e-14).
 * Device U, p, Jp [N][5], dt [N].  Its scratch (four [N][5] vectors) is
 * allocated on the first call.  Synchronises; returns the summed
 * reconstruction error counts of its two residual evaluations (0 when
 * valid; Jp is then not meaningful) or < 0. */
int aerofx_jvp_fd(aerofx_ctx *ctx, const double *U, const double *p, double *Jp,
                  int order, const double *dt);
/* The MatVecProd of aerofx_gmres (ImplicitCoupledTsDesc.C:33-44 picks it
 * from ImplicitData::mvp):
 *   AEROFX_MVP_H1  MatVecProdH1: the assembled first-order Jacobian of the
 *                  last aerofx_jacobian_setup (the default)
 *   AEROFX_MVP_FD  MatVecProdFD::evaluate (MatVecProd.C:88-118) about Q:
 *                  products (F(Q + eps p) - F0) / eps (fd_order 1) or
 *                  (F(Q + eps p) - F(Q - eps p)) / (2 eps) (fd_order 2), eps
 *                  of computeEpsilon (MatVecProd.C:879-929), with
 *                  F(Q') = R(Q') + (Q' - Q) / dt (dt NULL: R only); F0 =
 *                  F(Q) may be passed (the Newton function value the caller
 *                  already holds) or NULL (evaluated here).  Q, F0 and dt
 *                  (device [N][5], [N][5], [N]) must stay valid and unchanged
 *                  while the operator is selected.  One residual per product
 *                  at fd_order 1, two at fd_order 2.
 *   AEROFX_MVP_H2  MatVecProdH2::evaluate/apply (MatVecProd.C:1560-1670)
 *                  about Q: the exact derivative of the second-order
 *                  residual, (dR/dU) p + p / dt (dt NULL: no time term),
 *                  formed matrix-free by the forward-mode kernels of
 *                  aerofx_residual_derivative (the forward pass at Q is
 *                  kept at selection; F0 and fd_order are ignored).  Q and
 *                  dt must stay valid and unchanged while it is selected.
 * The preconditioner stays the block-Jacobi of the last aerofx_jacobian_setup;
 * the transposed solves always use H1^T.  Returns the count of non-physical
 * reconstructions in the evaluation of F0 (AEROFX_MVP_FD) or of the
 * residual at Q (AEROFX_MVP_H2), 0 when valid, or < 0. */
#define AEROFX_MVP_H1 0
#define AEROFX_MVP_FD 1
#define AEROFX_MVP_H2 2
int aerofx_mvp_select(aerofx_ctx *ctx, int type, const double *Q, const double *F0,
                      const double *dt, int fd_order);
/* The preconditioner of aerofx_prec_apply, aerofx_gmres and the Newton
 * driver (PcData::type, IoDataCore.C:2744), taking effect at the next
 * aerofx_jacobian_setup (or Newton step):
 *   AEROFX_PREC_JACOBI    JacobiPrec: the inverted 5x5 diagonal blocks of
 *                         H1 (the default)
 *   AEROFX_PREC_RAS_ILU0  IluPrec with fill 0, type RAS (KspPrec.C:127-320):
 *                         the ILU(0) (SparseMat::numericILU / lusol,
 *                         SparseMatrix.C:353-466) of each subdomain of
 *                         `subdomain_nodes` consecutive nodes (1..1024), the
 *                         solution kept on the subdomain's own nodes.  The
 *                         node numbering is the subdomains' order: number the
 *                         mesh by RCM (aerofx_mesh_rcm) as the reference's
 *                         IluPrec renumbers its subdomains.  Factors in the
 *                         Jacobian's precision.  Transposed solves, the
 *                         adjoint and direct solves and block GMRES keep
 *                         block-Jacobi.
 * Returns 0 or < 0. */
#define AEROFX_PREC_JACOBI 0
#define AEROFX_PREC_RAS_ILU0 1
int aerofx_prec_select(aerofx_ctx *ctx, int type, int subdomain_nodes);
/* MatVecProd::apply (MatVecProd.h:54-80) with the selected operator:
 * prod = A p.  Synchronises; returns the count of non-physical
 * reconstructions in its residual evaluations (0 when valid), or < 0. */
int aerofx_mvp_apply(aerofx_ctx *ctx, const double *p, double *prod);
/* GmresSolver::solve (KspSolver.C:353-492): right-preconditioned GMRES(m)
 * with the selected MatVecProd (aerofx_mvp_select; H1 by default) and the
 * block-Jacobi of the last aerofx_jacobian_setup, x0 = 0, stop at
 * ||r|| <= eps ||b|| or maxits.  Orthogonalisation is CGS2 in FP64, the
 * reorthogonalisation of each vector delayed to the next step.  Returns 0,
 * or with AEROFX_MVP_FD the count of non-physical reconstructions met in its
 * products (> 0; x is then not meaningful), or < 0. */
int aerofx_gmres(aerofx_ctx *ctx, const double *b, double *x, int m,
                 double eps, int maxits, int *its, double *final_res);

/* GmresSolver::solveT (KspSolver.C:754): GMRES(m) on the transposed
 * operators H1^T and (D^-1)^T, as aerofx_gmres. */
int aerofx_gmres_transpose(aerofx_ctx *ctx, const double *b, double *x, int m,
                           double eps, int maxits, int *its, double *final_res);

/* ---- batched right-hand sides ------------------------------------------ */
typedef struct {
  int cycles;            /* restart cycles (each m block steps) */
  int inner_its;         /* block steps in total */
  int converged;
  double res_max;        /* max_k ||b_k - A x_k|| / ||b_k|| (FP64 operator) */
  double seconds;
  int tune_waves_dots, tune_waves_update;  /* autotuned launch shapes */
} aerofx_block_stats;
/* Block GMRES(m) for nrhs <= 16 systems A x_k = b_k with the operators of
 * the last aerofx_jacobian_setup (which must be FP64, precision 0): H1 with
 * right block-Jacobi.  B, X device [nrhs][N][5].  precision 0: FP64 inner
 * operators; 1: FP32 copies of H1 and of the Jacobi inverse inside the
 * cycles, FP64 residual B - A X at each restart (iterative refinement);
 * the Krylov basis and its block Gram-Schmidt (FP64 MFMA tall-skinny GEMMs)
 * are FP64 either way.  use_graph != 0 replays each cycle as one captured
 * hipGraph.  Stops when every ||b_k - A x_k|| <= eps ||b_k|| or after
 * max_cycles.  Nodes may have at most 64 neighbours (AEROFX_ERR_MESH
 * otherwise); a block space of 16 m vectors that exhausts a small system
 * makes the block rank-deficient (AEROFX_ERR_STATE): use a smaller m.
 * No counterpart in the reference: its multi-RHS
 * GmresSolver::solve(VecSet&, VecSet&) (KspSolver.C:988-1135) is dead code. */
int aerofx_block_gmres(aerofx_ctx *ctx, int nrhs, const double *B, double *X,
                       int m, double eps, int max_cycles, int precision,
                       int use_graph, aerofx_block_stats *stats);

/* ---- adjoint ----------------------------------------------------------- */
/* Transposed derivatives of the residual of aerofx_residual at U (device
 * [N][5]) applied to lambda (device [N][5]):
 *   dRdU_T_lambda [N][5] = (dR/dU)^T lambda   (may be NULL)
 *   dRdX_T_lambda [N][3] = (dR/dX)^T lambda   (may be NULL; X enters through
 *   the dual normals, control volumes, face normals, LSQ factors and the
 *   reconstruction distances).
 * Replaces SpaceOperator::computeTransposeDerivativeOfResidual and the
 * transposed sensitivity operators (SpaceOperator.C:862-910, 1235-1304;
 * Edge.C:1452-1728).  The MUSCL gradients are recomputed in the call.
 * Synchronises before returning. */
int aerofx_adjoint_residual(aerofx_ctx *ctx, const double *U, const double *lambda,
                            double *dRdU_T_lambda, double *dRdX_T_lambda);

/* Directional derivative of the residual of aerofx_residual at U (device
 * [N][5]): dR = (dR/dU) dU + (dR/dX) dX for a state perturbation dU (device
 * [N][5], may be NULL: zero) and a mesh perturbation dX (device [N][3], may
 * be NULL: zero); dR device [N][5].  Forward mode through every term the
 * residual computes from U and X (dual normals, control volumes, face
 * normals, LSQ factors, gradients, reconstruction, fluxes).  Replaces
 * SpaceOperator::computeDerivativeOfResidual (SpaceOperator.C:917-1070,
 * EdgeSet::computeDerivativeOfFiniteVolumeTerm, Edge.C:1494-1573) for the
 * direct shape sensitivity (dX = dX/ds) and, with dX NULL, is the exact
 * product of the second-order Jacobian (MatVecProdH2::apply,
 * MatVecProd.C:1293+).  The MUSCL gradients are recomputed.  Synchronises;
 * returns the non-physical reconstruction count (0 when valid) or < 0. */
int aerofx_residual_derivative(aerofx_ctx *ctx, const double *U, const double *dU,
                               const double *dX, double *dR);

/* Force on the wall faces along dir (lift for dir = e_z):
 * J = sum_{wall faces f} sum_{v in f} p_v (n_f / 3) . dir, plus, with the
 * Navier-Stokes terms on, the viscous stress -(tau_e n_f) . dir of each wall
 * face's element (PostFcnNS::computeViscousForce, PostFcn.C:1578-1612), with
 * dJdU (device [N][5], may be NULL) and dJdX (device [N][3], may be NULL).
 * The objective of the shape sensitivity (FluidShapeOptimizationHandler.C:880
 * evaluates the forces through PostOperator). */
int aerofx_wall_force(aerofx_ctx *ctx, const double *U, const double dir[3],
                      double *J, double *dJdU, double *dJdX);

typedef struct {
  int precision;        /* 0: FP64 H1/Jacobi in the inner solves; 1: FP32 */
  double cfl;           /* pseudo-time term 1/dt of the inner operator (0: none) */
  int max_its;          /* outer (defect-correction) iterations */
  double eps;           /* stop at ||rhs - (dR/dU)^T psi|| <= eps ||rhs|| */
  int gmres_m, gmres_maxits;
  double gmres_eps;
} aerofx_adjoint_opts;
typedef struct {
  int its, converged, krylov_its;
  double res0, res, seconds;
} aerofx_adjoint_stats;
/* Adjoint solve (dR/dU)^T psi = rhs at U (the sensitivity handler's
 * ksp->solveT, FluidShapeOptimizationHandler.C:1594-1624): outer FP64
 * iterations on the exact transposed residual derivative
 * (aerofx_adjoint_residual), each correcting psi with GMRES on the
 * transposed H1 + diag(1/dt) and block-Jacobi in the chosen precision. */
int aerofx_adjoint_solve(aerofx_ctx *ctx, const double *U, const double *rhs,
                         double *psi, const aerofx_adjoint_opts *opts,
                         aerofx_adjoint_stats *stats);

/* Direct sensitivity solve (dR/dU) x_k = rhs_k, k < nrhs <= 16, at U (the
 * sensitivity handler's ksp->solve(dFdS, dUdS), FluidShapeOptimizationHandler.C:
 * 1369-1430; rhs_k = -(dR/dX) dX/ds_k from aerofx_residual_derivative):
 * outer FP64 iterations on the exact second-order products of
 * aerofx_residual_derivative, each correcting every x_k by GMRES (nrhs 1)
 * or one block GMRES (nrhs > 1) on H1 + diag(1/dt) with block-Jacobi in the
 * chosen precision.  rhs, x device [nrhs][N][5].  Stops when every
 * ||rhs_k - (dR/dU) x_k|| <= eps ||rhs_k||; stats->res0 and res are the
 * norms of the worst right-hand side and its residual. */
int aerofx_direct_solve(aerofx_ctx *ctx, const double *U, int nrhs, const double *rhs,
                        double *x, const aerofx_adjoint_opts *opts,
                        aerofx_adjoint_stats *stats);

/* ---- steady pseudo-transient Newton-GMRES ------------------------------ */
typedef struct {
  int precision;        /* 0: all FP64; 1: FP32 Jacobian/JVP/block-Jacobi */
  int max_steps;        /* pseudo-time steps */
  double eps;           /* stop at ||R||_2 <= eps ||R_0||_2 */
  double cfl0, cfl_max, ser, cfl_coef1, cfl_coef2;  /* CflLaw Old */
  int gmres_m, gmres_maxits;
  double gmres_eps;     /* relative Krylov tolerance */
  int mvp;              /* AEROFX_MVP_H1 (assembled H1), AEROFX_MVP_FD (matrix-free
                           finite differences) or AEROFX_MVP_H2 (exact products) */
  int fd_order;         /* 1 or 2 with AEROFX_MVP_FD */
  int prec;             /* AEROFX_PREC_JACOBI or AEROFX_PREC_RAS_ILU0 (for the solve) */
  int prec_subdomain;   /* RAS / ILU(0) subdomain size in nodes */
  int eps_formula;      /* Krylov tolerance: 0 constant gmres_eps, 1 Eisenstat
                           (KspConvCriterion.C:28-74: gamma (|F|/|F_prev|)^alpha,
                           gamma 0.1, alpha 2, safeguard delta 0.1, capped at gmres_eps) */
  int failsafe;         /* Newton.FailSafe for the solve: 0 off, 1 on, 2 always on
                           (aerofx_ctx_set_failsafe) */
  int check_final_res;  /* KspData checkFinalRes: recompute ||b - A x|| after each
                           Krylov solve and count the solves above their target
                           (GmresSolver returns -999 for them, KspSolver.C:455-469) */
} aerofx_newton_opts;
typedef struct {
  int steps, converged, krylov_its;
  double res0, res, final_cfl, seconds;
  int failsafe_redos;      /* Newton steps redone by the failsafe */
  int saturated;           /* Krylov solves that hit gmres_maxits (ErrorHandler::SATURATED_LS) */
  int final_res_failures;  /* Krylov solves whose true residual missed the target */
} aerofx_newton_stats;
/* TsSolver::resolve + NewtonSolver::solve with one Newton iteration per
 * pseudo-time step (NewtonSolver.h:70-231, TsDesc.C:697-730, 1178-1278):
 * U (device [N][5]) is advanced in place; history (optional, hist_cap
 * entries) receives ||R||_2 per step. */
int aerofx_steady_solve(aerofx_ctx *ctx, double *U, const aerofx_newton_opts *opts,
                        aerofx_newton_stats *stats, double *history,
                        int hist_cap);

/* ---- SOWER mesh and solution files (host memory, no GPU) --------------- */
/* GeoSource::getSubDomain (GeoSource.C:246-356): one subdomain of a SOWER
 * cluster, from its geometry (.geo) and decomposition (.dec) files in the
 * BinFileHandler format (BinFileHandler.h; either byte order): node, element
 * and face ranges (getRangeInfo, GeoSource.C:204-240), NodeSet::read
 * (Node.C:7-50), ElemSet::read (ElemCore.C:135-200, tets only), FaceSet::read
 * (FaceCore.C:270-347, triangles; code 6 is read as BC_SYMMETRY = 11), with
 * element and face nodes renumbered from cluster to subdomain numbering
 * (GeoSource.C:360-379). */
typedef struct {
  int num_nodes, num_tets, num_faces;  /* of the subdomain */
  int num_clus_nodes;                  /* of the cluster (solution files hold these) */
  int num_sub_cluster;                 /* subdomains in the cluster */
  int num_neighbours;                  /* subdomains sharing nodes with this one */
  int num_bcs;                         /* BCondSet entries of the cluster */
  double version;                      /* BinFileHandler header version */
} aerofx_sower_header;
int aerofx_sower_info(const char *geo, const char *dec, int sub, aerofx_sower_header *info);
/* X [N][3], tets [T][4], faces [F][3], codes [F] (host); node_glob [N]
 * (global node numbers) and node_clus [N] (cluster node numbers, the rows of
 * the cluster's solution files) may be NULL. */
int aerofx_sower_read(const char *geo, const char *dec, int sub, double *X, int *tets, int *faces, int *codes,
                      int *node_glob, int *node_clus);
/* No counterpart in the reference (SOWER writes these files): a cluster of
 * one subdomain holding the whole mesh, in the layout aerofx_sower_read and
 * GeoSource::getSubDomain read (global numbers = indices, ids 0, no
 * neighbours, no boundary-condition records). */
int aerofx_sower_write(const char *geo, const char *dec, int N, int T, int F, const double *X, const int *tets,
                       const int *faces, const int *codes);
/* Solution / restart files (SubDomain.C:5609-5810): {cluster nodes, dim,
 * solutions} then per solution a double tag and [cluster nodes][dim]
 * values.  aerofx_solution_write: step 0 creates the file, a later step
 * updates it in place and sets the count to step + 1 (openFileForWriting,
 * writeTagToFile, writeVectorToFile with a cluster of one subdomain).
 * aerofx_solution_read: the rows node_clus[0..n) (NULL: all n, which must be
 * the cluster size) of solution `step`, min(dim, file dim) values per row
 * (readVectorFromFile), and its tag. */
int aerofx_solution_info(const char *path, int *num_clus_nodes, int *dim, int *nsol);
int aerofx_solution_write(const char *path, int step, double tag, int n, int dim, const double *U);
int aerofx_solution_read(const char *path, int step, double *tag, int n, int dim, const int *node_clus,
                         double *U);

/* ---- physics routines, one case per thread ----------------------------- */
/* No counterpart in the reference: evaluates the device routines that the
 * residual, Jacobian and preconditioner kernels inline, on n cases in device
 * arrays, so that tests can compare them case by case with stored outputs of
 * the reference's routines.  Runs on the null stream and synchronises.
 *   AEROFX_PHYS_ROEFLUX7     ROEFLUX7 (roeflux7.f), scalar = gamma:
 *                            a normal [n][3], b V_L [n][5], c V_R [n][5] -> out [n][5]
 *   AEROFX_PHYS_BOUNDFLUX5   BOUNDFLUX5 (boundflux5.f), scalar = gamma:
 *                            a normal, b V [n][5], c U_b [n][5] (conservative) -> out [n][5]
 *   AEROFX_PHYS_VANALBADA    RecFcn::vanalbada per variable (RecFcn.h:433-446),
 *                            scalar = beta: a V_i, b ddV_ij, c V_j, d ddV_ji
 *                            [n][5] -> out [n][10] (V_ij, V_ji)
 *   AEROFX_PHYS_ROEJAC6_SW1/2  ROEJAC6 type 1 or 2 (roejac6.f), scalar = gamma:
 *                            a normal, b U_g [n][5], c U_d [n][5] -> out [n][25] row-major
 *   AEROFX_PHYS_INVERT5_F64/F32  LINPACK dgefa + dgedi (DenseMatrixOps.h:1006-1181)
 *                            in FP64 or FP32: a blocks [n][25] -> out [n][25]
 *                            (FP32 results widened to FP64)
 *   AEROFX_PHYS_ROEFLUX5     ROEFLUX5 (roeflux5.f, Roe-Turkel low-Mach dissipation,
 *                            static face), scalar = gamma: a normal, b V_L, c V_R,
 *                            d [n][7] = (min Mach, slope, cut-off Mach, shock
 *                            reducer, 1/Re, edge length, prec flag) -> out [n][5]
 *   AEROFX_PHYS_ROEJAC6_LM_SW1/2  ROEJAC6 type 1 or 2 with the low-Mach constants
 *                            (roejac6.f:130-160), scalar = gamma: a normal,
 *                            b U_g, c U_d, d [n][7] as for ROEFLUX5 -> out [n][25]
 * Returns the number of singular blocks (inversions), 0, or < 0. */
#define AEROFX_PHYS_ROEFLUX7 0
#define AEROFX_PHYS_BOUNDFLUX5 1
#define AEROFX_PHYS_VANALBADA 2
#define AEROFX_PHYS_ROEJAC6_SW1 3
#define AEROFX_PHYS_ROEJAC6_SW2 4
#define AEROFX_PHYS_INVERT5_F64 5
#define AEROFX_PHYS_INVERT5_F32 6
#define AEROFX_PHYS_ROEFLUX5 7
#define AEROFX_PHYS_ROEJAC6_LM_SW1 8
#define AEROFX_PHYS_ROEJAC6_LM_SW2 9
int aerofx_physics_eval(int kind, int n, double scalar, const double *a, const double *b,
                        const double *c, const double *d, double *out);

#ifdef __cplusplus
}
#endif
#endif
