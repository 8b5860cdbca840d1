"""Python mirror of the reference's SpaceOperator<5> surface over the C ABI.

Device arrays are torch tensors (float64, contiguous, on the GPU); torch is
only the allocator and stream here.  Every call goes to the HIP kernels in
libaerofx.so -- there is no other code path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import Mesh as _CMesh
from ._abi import (AdjointOpts, AdjointStats, BlockStats, NewtonOpts, NewtonStats, c_double_p, c_int_p, check,
                   lib)
from .mesh import FlowParams, Mesh


def _ptr(t, numel: int) -> int:
    """Device address of a contiguous float64 tensor of exactly numel elements.

    Every entry point of the C ABI reads and writes implicit extents ([N][5],
    [N][3], [N] or [nrhs][N][5]); a tensor of another size is rejected here,
    before any launch could touch memory past its end."""
    if not t.is_cuda:
        raise ValueError("device tensor expected")
    if not t.is_contiguous():
        raise ValueError("contiguous tensor expected")
    import torch
    if t.dtype != torch.float64:
        raise ValueError("float64 tensor expected")
    if t.numel() != numel:
        raise ValueError(f"tensor of {numel} elements expected, got {t.numel()} (shape {tuple(t.shape)})")
    return t.data_ptr()


def _optr(t, numel: int):
    """_ptr for an optional argument (None -> NULL)."""
    return None if t is None else _ptr(t, numel)


class SpaceOperator:
    """SpaceOperator<5> (SpaceOperator.h:171-173) bound to one mesh.

    The context owns the mesh tables (edges, dual normals, control volumes,
    LSQ factors), as the reference's SpaceOperator owns its geometry state.
    """

    def __init__(self, mesh: Mesh, params: FlowParams, stream=None):
        self.mesh = mesh
        self.params = params
        X = np.ascontiguousarray(mesh.X, dtype=np.float64)
        tets = np.ascontiguousarray(mesh.tets, dtype=np.int32)
        faces = np.ascontiguousarray(mesh.faces, dtype=np.int32)
        codes = np.ascontiguousarray(mesh.codes, dtype=np.int32)
        cm = _CMesh(mesh.num_nodes, mesh.num_tets, mesh.num_faces, X.ctypes.data_as(c_double_p),
                    tets.ctypes.data_as(c_int_p), faces.ctypes.data_as(c_int_p),
                    codes.ctypes.data_as(c_int_p))
        p = params.c()
        h = C.c_void_p()
        check(lib().aerofx_ctx_create(C.byref(cm), C.byref(p), C.byref(h)), "ctx_create")
        self._h = h
        N, E, F = C.c_int(), C.c_int(), C.c_int()
        check(lib().aerofx_ctx_sizes(h, C.byref(N), C.byref(E), C.byref(F)), "ctx_sizes")
        self.num_nodes, self.num_edges, self.num_faces = N.value, E.value, F.value
        self.n5, self.n3 = 5 * self.num_nodes, 3 * self.num_nodes  # element counts of [N][5], [N][3]
        self.set_stream(stream)
        self.set_low_mach(params.low_mach)
        self.set_viscous(getattr(params, "viscous", None))

    def set_failsafe(self, mode: int):
        """Newton.FailSafe for the synchronous residual: 0 off, 1 on, 2 always on."""
        check(lib().aerofx_ctx_set_failsafe(self._h, int(mode)), "set_failsafe")

    def reset_failsafe(self):
        check(lib().aerofx_ctx_reset_failsafe(self._h), "reset_failsafe")

    def failsafe_nodes(self):
        """(tag0, tag1, first-order mask), host int arrays [N]."""
        t = np.zeros((3, self.num_nodes), dtype=np.int32)
        check(lib().aerofx_ctx_failsafe_nodes(self._h, t.ctypes.data_as(c_int_p)), "failsafe_nodes")
        return t[0], t[1], t[2]

    def set_krylov_check(self, on: bool):
        check(lib().aerofx_ctx_set_krylov_check(self._h, int(bool(on))), "set_krylov_check")

    def set_low_mach(self, lm=None):
        """Spatial low-Mach preconditioning (LowMachPrec) on, or off (None)."""
        from ._abi import LowMachC
        if lm is None:
            check(lib().aerofx_ctx_set_low_mach(self._h, None), "set_low_mach")
        else:
            c = LowMachC(lm.min_mach, lm.slope, lm.cutoff, lm.shock_reducer, lm.irey)
            check(lib().aerofx_ctx_set_low_mach(self._h, C.byref(c)), "set_low_mach")

    def set_viscous(self, v=None):
        """Laminar Navier-Stokes terms (ViscousParams) on, or off (None)."""
        from ._abi import ViscousC
        if v is None:
            check(lib().aerofx_ctx_set_viscous(self._h, None), "set_viscous")
            return
        vm, cm = v.codes()
        c = ViscousC(v.reynolds, vm, v.bulk_viscosity, v.sutherland_ts, cm, v.prandtl, v.conductivity,
                     v.wall_temperature, v.viscous_cst)
        check(lib().aerofx_ctx_set_viscous(self._h, C.byref(c)), "set_viscous")

    def node_types(self):
        """Node types of the wall conditions (host int array [N])."""
        t = np.empty(self.num_nodes, dtype=np.int32)
        check(lib().aerofx_ctx_node_types(self._h, t.ctypes.data_as(c_int_p)), "node_types")
        return t

    def apply_bcs_residual(self, U, F):
        """SpaceOperator::applyBCsToResidual: the wall rows of F (in place)."""
        check(lib().aerofx_apply_bcs_residual(self._h, _ptr(U, self.n5), _ptr(F, self.n5)), "apply_bcs_residual")

    def apply_bcs_solution(self, U):
        """SpaceOperator::applyBCsToSolutionVector (in place)."""
        check(lib().aerofx_apply_bcs_solution(self._h, _ptr(U, self.n5)), "apply_bcs_solution")

    def close(self):
        if getattr(self, "_h", None):
            lib().aerofx_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream=None):
        """Binds the torch stream (default: torch's current stream)."""
        import torch
        if stream is None and torch.cuda.is_available():
            stream = torch.cuda.current_stream()
        if stream is None:
            check(lib().aerofx_ctx_set_stream(self._h, None, 1), "set_stream")
        else:
            check(lib().aerofx_ctx_set_stream(self._h, C.c_void_p(stream.cuda_stream), 0), "set_stream")

    def profile(self, enable: bool = True):
        """Turns per-kernel event timing on/off (resets the buckets)."""
        check(lib().aerofx_ctx_profile(self._h, int(enable)), "ctx_profile")

    def timers(self) -> dict:
        """{bucket: (total ms, launches)} for buckets with launches."""
        from ._abi import TIMER_NAMES
        nt = len(TIMER_NAMES)
        ms = (C.c_double * nt)()
        n = (C.c_longlong * nt)()
        check(lib().aerofx_ctx_timers(self._h, ms, n), "ctx_timers")
        return {TIMER_NAMES[k]: (ms[k], n[k]) for k in range(nt) if n[k]}

    def sync(self):
        check(lib().aerofx_ctx_sync(self._h), "sync")

    def set_residual_kernels(self, family: str = "auto") -> str:
        """'per_lane' (the default; 'auto' selects it) or 'staged' (LDS-staged
        strips when the mesh fits them); returns the family the next residual
        uses ('staged' or 'per_lane')."""
        code = {"auto": 0, "per_lane": 1, "staged": 2}[family]
        rc = check(lib().aerofx_ctx_set_residual_kernels(self._h, code), "set_residual_kernels")
        return {1: "per_lane", 2: "staged"}[rc]

    def tables(self):
        """Host copies: edge table, ctrlVol, edge normals, face normals, LSQ R."""
        N, E, F = self.num_nodes, self.num_edges, self.num_faces
        ptr = np.empty((E, 2), np.int32)
        cv = np.empty(N, np.float64)
        en = np.empty((E, 3), np.float64)
        fn = np.empty((F, 3), np.float64)
        R = np.empty((N, 6), np.float64)
        check(lib().aerofx_ctx_tables(self._h, ptr.ctypes.data_as(c_int_p), cv.ctypes.data_as(c_double_p),
                                      en.ctypes.data_as(c_double_p), fn.ctypes.data_as(c_double_p),
                                      R.ctypes.data_as(c_double_p)), "ctx_tables")
        return dict(ptr=ptr, ctrl_vol=cv, edge_normals=en, face_normals=fn, lsq=R)

    def compute_residual(self, U, R, fluxes_only: bool = False) -> int:
        """computeResidual (SpaceOperator.C:524-634): R = fluxes / V.

        Returns the count of non-physical reconstructions (0 when valid)."""
        return check(lib().aerofx_residual_ex(self._h, _ptr(U, self.n5), _ptr(R, self.n5), 1 if fluxes_only else 0),
                     "residual")

    def compute_residual_async(self, U, R):
        check(lib().aerofx_residual_ex(self._h, _ptr(U, self.n5), _ptr(R, self.n5), 2), "residual")

    def compute_gradients(self, V, ddx, ddy, ddz):
        """DistNodalGrad::compute, least squares, on primitive V."""
        check(lib().aerofx_gradients(self._h, _ptr(V, self.n5), _ptr(ddx, self.n5), _ptr(ddy, self.n5),
                                       _ptr(ddz, self.n5)), "gradients")

    def conservative_to_primitive(self, U, V):
        check(lib().aerofx_primitive(self._h, _ptr(U, self.n5), _ptr(V, self.n5)), "primitive")

    def wall_force(self, U, direction=(0.0, 0.0, 1.0), dJdU=None, dJdX=None) -> float:
        """Force on the walls along direction (pressure, plus the viscous stress
        with the Navier-Stokes terms on); optional dJ/dU [N,5], dJ/dX [N,3]."""
        d = (C.c_double * 3)(*direction)
        J = C.c_double()
        check(lib().aerofx_wall_force(self._h, _ptr(U, self.n5), d, C.byref(J), _optr(dJdU, self.n5),
                                      _optr(dJdX, self.n3)), "wall_force")
        return J.value

    def compute_derivative_of_residual(self, U, dR, dU=None, dX=None) -> int:
        """dR = (dR/dU) dU + (dR/dX) dX at U (SpaceOperator::computeDerivativeOfResidual,
        SpaceOperator.C:917-1070); dU [N,5] and dX [N,3] may be None (zero).
        Returns the count of non-physical reconstructions (0 when valid)."""
        return check(lib().aerofx_residual_derivative(self._h, _ptr(U, self.n5), _optr(dU, self.n5),
                                                      _optr(dX, self.n3), _ptr(dR, self.n5)),
                     "residual_derivative")

    def compute_transpose_derivative_of_residual(self, U, lam, dRdU_T_lam=None, dRdX_T_lam=None):
        """(dR/dU)^T lam -> dRdU_T_lam [N,5] and (dR/dX)^T lam -> dRdX_T_lam [N,3]
        (SpaceOperator::computeTransposeDerivativeOfResidual, SpaceOperator.C:1235-1304).
        Either output may be None.  The MUSCL gradients are recomputed."""
        check(lib().aerofx_adjoint_residual(self._h, _ptr(U, self.n5), _ptr(lam, self.n5),
                                            _optr(dRdU_T_lam, self.n5), _optr(dRdX_T_lam, self.n3)),
              "adjoint_residual")


class MatVecProdH1:
    """MatVecProdH1<5,Scalar,5> (MatVecProd.h:37-259) and JacobiPrec<Scalar,5>
    (KspPrec.h:69-93) over one SpaceOperator context.

    evaluate() assembles the first-order Roe Jacobian at U plus alpha/dt on
    the diagonal and sets up the block-Jacobi inverse; precision 0 stores the
    blocks in FP64, 1 in FP32 (the mixed-precision JVP and preconditioner).
    """

    def __init__(self, op: SpaceOperator):
        self.op = op

    def select_preconditioner(self, kind: str = "jacobi", subdomain: int = 256):
        """'jacobi' (JacobiPrec) or 'ras_ilu0' (IluPrec, fill 0, RAS over
        subdomains of `subdomain` consecutive nodes); effective at the next
        evaluate()."""
        check(lib().aerofx_prec_select(self.op._h, {"jacobi": 0, "ras_ilu0": 1}[kind], int(subdomain)), "prec_select")

    def evaluate(self, U, dt=None, precision: int = 0):
        check(lib().aerofx_jacobian_setup(self.op._h, _ptr(U, self.op.n5), _optr(dt, self.op.num_nodes),
                                          int(precision)), "jacobian_setup")

    def apply(self, p, prod):
        check(lib().aerofx_jvp(self.op._h, _ptr(p, self.op.n5), _ptr(prod, self.op.n5)), "jvp")

    def prec_apply(self, y, x):
        check(lib().aerofx_prec_apply(self.op._h, _ptr(y, self.op.n5), _ptr(x, self.op.n5)), "prec_apply")

    def apply_transpose(self, p, prod):
        """MatVecProdH1::applyTranspose (MatVecProd.C:1190-1206)."""
        check(lib().aerofx_jvp_transpose(self.op._h, _ptr(p, self.op.n5), _ptr(prod, self.op.n5)), "jvp_transpose")

    def prec_apply_transpose(self, y, x):
        check(lib().aerofx_prec_apply_transpose(self.op._h, _ptr(y, self.op.n5), _ptr(x, self.op.n5)), "prec_apply_transpose")

    def blocks(self):
        """Host copies: D [N,5,5] and per edge (A_ij, A_ji) [E,2,5,5]."""
        N, E = self.op.num_nodes, self.op.num_edges
        D = np.empty((N, 5, 5))
        A = np.empty((E, 2, 5, 5))
        check(lib().aerofx_jacobian_blocks(self.op._h, D.ctypes.data_as(c_double_p),
                                           A.ctypes.data_as(c_double_p)), "jacobian_blocks")
        return D, A

    def timestep(self, U, cfl, dt):
        check(lib().aerofx_timestep(self.op._h, _ptr(U, self.op.n5), float(cfl), _ptr(dt, self.op.num_nodes)), "timestep")


class MatVecProdFD:
    """MatVecProdFD<5,5> (MatVecProd.h:83-160): finite-difference JVP of the
    residual plus the backward-Euler time term (dt given) about U.

    evaluate() selects it as the operator of the context's GMRES (the
    block-Jacobi preconditioner stays the one of the last
    MatVecProdH1.evaluate); apply() is one product.  Q, F0 and dt must stay
    alive and unchanged while it is selected.  standalone_apply() is the
    self-contained aerofx_jvp_fd (F(U) recomputed per call)."""

    def __init__(self, op: SpaceOperator, order: int = 1):
        self.op, self.order = op, order
        self.U = self.dt = self.F0 = None

    def evaluate(self, U, dt=None, F0=None) -> int:
        from ._abi import AEROFX_MVP_FD
        self.U, self.dt, self.F0 = U, dt, F0
        return check(lib().aerofx_mvp_select(self.op._h, AEROFX_MVP_FD, _ptr(U, self.op.n5), _optr(F0, self.op.n5),
                                             _optr(dt, self.op.num_nodes), int(self.order)), "mvp_select")

    def release(self):
        """Selects the assembled H1 again as the GMRES operator."""
        from ._abi import AEROFX_MVP_H1
        check(lib().aerofx_mvp_select(self.op._h, AEROFX_MVP_H1, None, None, None, 1), "mvp_select")
        self.U = self.dt = self.F0 = None

    def apply(self, p, prod) -> int:
        return check(lib().aerofx_mvp_apply(self.op._h, _ptr(p, self.op.n5), _ptr(prod, self.op.n5)), "mvp_apply")

    def standalone_apply(self, U, p, prod, dt=None) -> int:
        return check(lib().aerofx_jvp_fd(self.op._h, _ptr(U, self.op.n5), _ptr(p, self.op.n5),
                                         _ptr(prod, self.op.n5), int(self.order),
                                         _optr(dt, self.op.num_nodes)), "jvp_fd")


class MatVecProdH2:
    """MatVecProdH2<5,double,5> (MatVecProd.h, MatVecProd.C:1560-1670): the
    exact derivative of the second-order residual plus the time term, about
    U, as the operator of the context's GMRES.  Formed matrix-free by the
    forward-mode kernels; evaluate() keeps the forward pass at U.  U and dt
    must stay alive and unchanged while it is selected."""

    def __init__(self, op: SpaceOperator):
        self.op = op
        self.U = self.dt = None

    def evaluate(self, U, dt=None) -> int:
        from ._abi import AEROFX_MVP_H2
        self.U, self.dt = U, dt
        return check(lib().aerofx_mvp_select(self.op._h, AEROFX_MVP_H2, _ptr(U, self.op.n5), None,
                                             _optr(dt, self.op.num_nodes), 1), "mvp_select")

    def release(self):
        """Selects the assembled H1 again as the GMRES operator."""
        from ._abi import AEROFX_MVP_H1
        check(lib().aerofx_mvp_select(self.op._h, AEROFX_MVP_H1, None, None, None, 1), "mvp_select")
        self.U = self.dt = None

    def apply(self, p, prod) -> int:
        return check(lib().aerofx_mvp_apply(self.op._h, _ptr(p, self.op.n5), _ptr(prod, self.op.n5)), "mvp_apply")


class GmresSolver:
    """GmresSolver (KspSolver.h:23-143): right-preconditioned GMRES(m)."""

    def __init__(self, op: SpaceOperator, m: int = 30, eps: float = 1e-2, maxits: int = 30):
        self.op, self.m, self.eps, self.maxits = op, m, eps, maxits

    def solve(self, b, x):
        """Returns (iterations, final residual estimate); raises if a
        finite-difference product met a non-physical state."""
        its = C.c_int()
        res = C.c_double()
        bad = check(lib().aerofx_gmres(self.op._h, _ptr(b, self.op.n5), _ptr(x, self.op.n5), self.m, self.eps, self.maxits, C.byref(its),
                                 C.byref(res)), "gmres")
        if bad:
            raise RuntimeError(f"gmres: {bad} non-physical reconstructions in finite-difference products")
        return its.value, res.value

    def solveT(self, b, x):
        """GmresSolver::solveT (KspSolver.C:754): the transposed system."""
        its = C.c_int()
        res = C.c_double()
        check(lib().aerofx_gmres_transpose(self.op._h, _ptr(b, self.op.n5), _ptr(x, self.op.n5), self.m, self.eps, self.maxits,
                                           C.byref(its), C.byref(res)), "gmres_transpose")
        return its.value, res.value


class BlockGmresSolver:
    """Block GMRES(m) over up to 16 right-hand sides sharing the operators of
    the last MatVecProdH1.evaluate (FP64).  precision 1 runs the cycles with
    FP32 copies of H1 and the Jacobi inverse and refines with the FP64
    residual; use_graph replays each cycle as a captured hipGraph."""

    def __init__(self, op: SpaceOperator, m: int = 8, eps: float = 1e-8, max_cycles: int = 200,
                 precision: int = 1, use_graph: bool = True):
        self.op, self.m, self.eps, self.max_cycles = op, m, eps, max_cycles
        self.precision, self.use_graph = precision, use_graph

    def solve(self, B, X) -> dict:
        """B, X: device float64 [nrhs, N, 5], nrhs <= 16."""
        nrhs = int(B.shape[0]) if B.dim() > 0 else 0
        if not 1 <= nrhs <= 16:
            raise ValueError(f"1 to 16 right-hand sides expected, got {nrhs}")
        st = BlockStats()
        check(lib().aerofx_block_gmres(self.op._h, nrhs, _ptr(B, nrhs * self.op.n5), _ptr(X, nrhs * self.op.n5),
                                       self.m, self.eps,
                                       self.max_cycles, self.precision, 1 if self.use_graph else 0, C.byref(st)),
              "block_gmres")
        return dict(cycles=st.cycles, inner_its=st.inner_its, converged=bool(st.converged), res_max=st.res_max,
                    seconds=st.seconds, tune=(st.tune_waves_dots, st.tune_waves_update))


def adjoint_solve(op: SpaceOperator, U, rhs, psi, precision: int = 1, cfl: float = 1e4, max_its: int = 400,
                  eps: float = 1e-10, gmres_m: int = 30, gmres_maxits: int = 30, gmres_eps: float = 1e-2):
    """(dR/dU)^T psi = rhs: FP64 outer iterations on the exact adjoint, inner
    transposed GMRES on H1 + 1/dt with block-Jacobi (FP32 when precision=1)."""
    o = AdjointOpts(int(precision), float(cfl), int(max_its), float(eps), int(gmres_m), int(gmres_maxits),
                    float(gmres_eps))
    st = AdjointStats()
    check(lib().aerofx_adjoint_solve(op._h, _ptr(U, op.n5), _ptr(rhs, op.n5), _ptr(psi, op.n5), C.byref(o),
                                       C.byref(st)),
          "adjoint_solve")
    return dict(its=st.its, converged=bool(st.converged), krylov_its=st.krylov_its, res0=st.res0, res=st.res,
                seconds=st.seconds)


def direct_solve(op: SpaceOperator, U, rhs, x, precision: int = 1, cfl: float = 1e4, max_its: int = 400,
                 eps: float = 1e-10, gmres_m: int = 30, gmres_maxits: int = 30, gmres_eps: float = 1e-2):
    """(dR/dU) x_k = rhs_k for rhs, x [nrhs, N, 5] (or [N, 5]), nrhs <= 16:
    FP64 outer iterations on the exact products, inner GMRES (one RHS) or
    block GMRES (several) on H1 + 1/dt with block-Jacobi (FP32 when
    precision=1)."""
    nrhs = 1 if rhs.dim() == 2 else int(rhs.shape[0])
    o = AdjointOpts(int(precision), float(cfl), int(max_its), float(eps), int(gmres_m), int(gmres_maxits),
                    float(gmres_eps))
    st = AdjointStats()
    check(lib().aerofx_direct_solve(op._h, _ptr(U, op.n5), nrhs, _ptr(rhs, nrhs * op.n5), _ptr(x, nrhs * op.n5),
                                    C.byref(o), C.byref(st)), "direct_solve")
    return dict(its=st.its, converged=bool(st.converged), krylov_its=st.krylov_its, res0=st.res0, res=st.res,
                seconds=st.seconds)


def steady_solve(op: SpaceOperator, U, precision: int = 0, max_steps: int = 200, eps: float = 1e-8,
                 cfl0: float = 10.0, cfl_max: float = 1e4, ser: float = 0.7, gmres_m: int = 30,
                 gmres_maxits: int = 30, gmres_eps: float = 1e-2, mvp: str = "h1", fd_order: int = 1,
                 prec: str = "jacobi", prec_subdomain: int = 256, eps_formula: str = "constant",
                 failsafe: int = 0, check_final_res: bool = False):
    """Pseudo-transient Newton-GMRES to ||R|| <= eps ||R0|| (U updated in place).

    mvp "h1": Krylov products with the assembled first-order Jacobian (FP32
    blocks when precision=1); "fd": matrix-free MatVecProdFD products of the
    FP64 residual (one residual per product); "h2": exact products of the
    second-order residual (MatVecProdH2, forward-mode kernels); block-Jacobi
    in `precision` for all three.  prec "jacobi" (block-Jacobi of H1) or
    "ras_ilu0" (RAS with ILU(0) on subdomains of prec_subdomain nodes).
    eps_formula "constant" (gmres_eps) or "eisenstat" (KspConvCriterion);
    failsafe 0/1/2 (Newton.FailSafe Off/On/AlwaysOn); check_final_res
    counts the Krylov solves whose true residual misses the target."""
    code = {"h1": 0, "fd": 1, "h2": 2}[mvp]
    pcode = {"jacobi": 0, "ras_ilu0": 1}[prec]
    ecode = {"constant": 0, "eisenstat": 1}[eps_formula]
    o = NewtonOpts(int(precision), int(max_steps), float(eps), float(cfl0), float(cfl_max), float(ser), 0.0, 0.0,
                   int(gmres_m), int(gmres_maxits), float(gmres_eps), code, int(fd_order), pcode, int(prec_subdomain),
                   ecode, int(failsafe), int(bool(check_final_res)))
    st = NewtonStats()
    hist = np.zeros(max_steps + 1)
    check(lib().aerofx_steady_solve(op._h, _ptr(U, op.n5), C.byref(o), C.byref(st), hist.ctypes.data_as(c_double_p),
                                    max_steps + 1), "steady_solve")
    n = min(st.steps + 1, max_steps + 1)
    return dict(steps=st.steps, converged=bool(st.converged), krylov_its=st.krylov_its, res0=st.res0,
                res=st.res, final_cfl=st.final_cfl, seconds=st.seconds, history=hist[:n].copy(),
                failsafe_redos=st.failsafe_redos, saturated=st.saturated, final_res_failures=st.final_res_failures)
