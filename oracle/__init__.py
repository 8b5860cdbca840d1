"""ctypes access to the oracle (CPU restatement) and to oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libaerof_ref.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


class OrParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("pstiff", C.c_double), ("beta", C.c_double),
                ("second_order", C.c_int), ("rho_inf", C.c_double), ("p_inf", C.c_double),
                ("mach", C.c_double), ("alpha", C.c_double), ("yaw", C.c_double),
                ("low_mach", C.c_int), ("lm_min_mach", C.c_double), ("lm_slope", C.c_double),
                ("lm_cutoff", C.c_double), ("lm_shock", C.c_double), ("lm_irey", C.c_double),
                ("viscous", C.c_int), ("visc_model", C.c_int), ("cond_model", C.c_int),
                ("reynolds", C.c_double), ("prandtl", C.c_double), ("bulk_viscosity", C.c_double),
                ("sutherland_ts", C.c_double), ("conductivity", C.c_double), ("wall_temperature", C.c_double),
                ("viscous_cst", C.c_double)]


class OrMesh(C.Structure):
    _fields_ = [("N", C.c_int), ("T", C.c_int), ("F", C.c_int), ("E", C.c_int),
                ("X", dp), ("tets", ip), ("faces", ip), ("codes", ip),
                ("ptr", ip), ("tet_edges", ip), ("cv", dp), ("normals", dp), ("normal_vel", dp),
                ("fnorm", dp), ("R", dp),
                ("Uin", C.c_double * 5), ("Uout", C.c_double * 5),
                ("Ub_in", C.c_double * 5), ("Ub_out", C.c_double * 5)]


def build(quiet: bool = True) -> None:
    """Builds liboracle.so (and oracle/_ref when the reference tree exists)."""
    out = subprocess.DEVNULL if quiet else None
    subprocess.check_call(["make", "-s", "-C", HERE], stdout=out)
    if os.path.isdir("/root/reference"):
        subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "ref")], stdout=out)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        L.or_edges.restype = C.c_int
        L.or_setup.restype = C.c_int
        L.or_residual_ex.restype = C.c_int
        L.or_residual.restype = C.c_int
        L.or_jvp_fd.restype = C.c_int
        L.or_color_edges.restype = C.c_int
        L.or_edge_priority.restype = C.c_uint32
        L.or_edge_priority.argtypes = [C.c_uint32]
        for name in ("or_roeflux7", "or_boundflux5", "or_vanalbada5", "or_wallflux"):
            getattr(L, name).restype = None
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """The reference's own kernels compiled from its sources (oracle/_ref)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            raise FileNotFoundError(REF_SO)
        R = C.CDLL(REF_SO)
        for name in ("refh_roeflux7", "refh_boundflux5", "refh_roejac6", "refh_vanalbada5",
                     "refh_apply5_f32", "refh_apply5_f64"):
            getattr(R, name).restype = None
        R.refh_invert5_f32.restype = C.c_int
        R.refh_invert5_f64.restype = C.c_int
        _ref = R
    return _ref


def params(p) -> OrParams:
    """OrParams from an aerof_dev_amd.FlowParams-like object."""
    lm = getattr(p, "low_mach", None)
    if lm is None:
        o = OrParams(p.gamma, p.pstiff, p.beta, int(p.second_order), p.rho_inf, p.p_inf, p.mach,
                     p.alpha, p.yaw, 0, 1.0, 0.0, 1.0, 0.0, 0.0)
    else:
        o = OrParams(p.gamma, p.pstiff, p.beta, int(p.second_order), p.rho_inf, p.p_inf, p.mach,
                     p.alpha, p.yaw, 1, lm.min_mach, lm.slope, lm.cutoff, lm.shock_reducer, lm.irey)
    v = getattr(p, "viscous", None)
    if v is not None:
        o.viscous = 1
        o.visc_model, o.cond_model = v.codes()
        o.reynolds, o.prandtl, o.bulk_viscosity = v.reynolds, v.prandtl, v.bulk_viscosity
        o.sutherland_ts, o.conductivity = v.sutherland_ts, v.conductivity
        o.wall_temperature, o.viscous_cst = v.wall_temperature, v.viscous_cst
    return o


class OracleMesh:
    """or_mesh with its derived tables; keeps the numpy inputs alive."""

    def __init__(self, mesh, p):
        self.X = np.ascontiguousarray(mesh.X, dtype=np.float64)
        self.tets = np.ascontiguousarray(mesh.tets, dtype=np.int32)
        self.faces = np.ascontiguousarray(mesh.faces, dtype=np.int32)
        self.codes = np.ascontiguousarray(mesh.codes, dtype=np.int32)
        self.p = params(p)
        m = OrMesh()
        m.N, m.T, m.F = self.X.shape[0], self.tets.shape[0], self.faces.shape[0]
        m.X, m.tets, m.faces, m.codes = _d(self.X), _i(self.tets), _i(self.faces), _i(self.codes)
        lib().or_setup(C.byref(m), C.byref(self.p))
        self.m = m
        self.N, self.E, self.F = m.N, m.E, m.F

    def __del__(self):
        try:
            lib().or_free(C.byref(self.m))
        except Exception:
            pass

    def tables(self):
        N, E, F = self.N, self.E, self.F
        m = self.m
        return dict(
            ptr=np.ctypeslib.as_array(m.ptr, shape=(E, 2)).copy(),
            ctrl_vol=np.ctypeslib.as_array(m.cv, shape=(N,)).copy(),
            edge_normals=np.ctypeslib.as_array(m.normals, shape=(E, 3)).copy(),
            normal_vel=np.ctypeslib.as_array(m.normal_vel, shape=(E,)).copy(),
            face_normals=np.ctypeslib.as_array(m.fnorm, shape=(F, 3)).copy() if F else np.zeros((0, 3)),
            lsq=np.ctypeslib.as_array(m.R, shape=(N, 6)).copy(),
        )

    def residual(self, U, fluxes_only=False, want_abs=False, want_grad=False):
        U = np.ascontiguousarray(U, dtype=np.float64)
        N = self.N
        R = np.empty((N, 5))
        Rabs = np.empty((N, 5)) if want_abs else None
        V = np.empty((N, 5))
        g = [np.empty((N, 5)) for _ in range(3)] if want_grad else None
        ierr = lib().or_residual_ex(C.byref(self.m), C.byref(self.p), _d(U), _d(R),
                                    _d(Rabs) if want_abs else None, int(fluxes_only), _d(V),
                                    *( [_d(a) for a in g] if want_grad else [None, None, None]))
        out = dict(R=R, ierr=ierr, V=V)
        if want_abs:
            out["Rabs"] = Rabs
        if want_grad:
            out["ddx"], out["ddy"], out["ddz"] = g
        return out

    def residual_fs(self, U, failsafe, tags):
        """The residual with Newton.FailSafe; tags int32 [3, N] (tag0, tag1,
        first-order mask) is updated in place.  Returns (R, count of the
        pass that stands, count of the first pass)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        R = np.empty((self.N, 5))
        first = C.c_int()
        L = lib()
        L.or_residual_fs.restype = C.c_int
        ierr = L.or_residual_fs(C.byref(self.m), C.byref(self.p), _d(U), _d(R), int(failsafe), _i(tags),
                                C.byref(first))
        return R, ierr, first.value

    def residual_par(self, U, nthreads):
        """The residual on nthreads OpenMP threads over node partitions with
        the shared-node assembly (CPU timing baseline)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        R = np.empty((self.N, 5))
        L = lib()
        L.or_residual_par.restype = C.c_int
        ierr = L.or_residual_par(C.byref(self.m), C.byref(self.p), _d(U), _d(R), int(nthreads))
        return dict(R=R, ierr=ierr)

    def jvp_fd(self, U, p, order=1, dt=None):
        """MatVecProdFD: (F(U + eps p) - F(U)) / eps, or central (order 2)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        out = np.empty_like(U)
        dtp = _d(np.ascontiguousarray(dt, dtype=np.float64)) if dt is not None else None
        ierr = lib().or_jvp_fd(C.byref(self.m), C.byref(self.p), _d(U), _d(p), int(order), dtp, _d(out))
        if ierr:
            raise RuntimeError(f"non-physical reconstruction in jvp_fd ({ierr})")
        return out

    def timestep(self, V, cfl):
        V = np.ascontiguousarray(V, dtype=np.float64)
        dt = np.empty(self.N)
        lib().or_timestep(C.byref(self.m), C.byref(self.p), _d(V), C.c_double(cfl), _d(dt))
        return dt

    def jacobian(self, V, dt=None):
        """H1 blocks: D [N,5,5], per edge (A_ij, A_ji) [E,2,5,5]."""
        V = np.ascontiguousarray(V, dtype=np.float64)
        D = np.empty((self.N, 5, 5))
        A = np.empty((self.E, 2, 5, 5))
        dtp = _d(np.ascontiguousarray(dt, dtype=np.float64)) if dt is not None else None
        lib().or_jacobian_h1(C.byref(self.m), C.byref(self.p), _d(V), dtp, _d(D), _d(A))
        return D, A

    def h1_apply(self, D, A, p):
        y = np.empty((self.N, 5))
        D = np.ascontiguousarray(D); A = np.ascontiguousarray(A); p = np.ascontiguousarray(p, dtype=np.float64)
        lib().or_h1_apply(C.byref(self.m), _d(D), _d(A), _d(p), _d(y))
        return y

    def gmres(self, V, dt, b, precision=0, m=30, eps=1e-2, maxits=30, U=None, fd_order=1):
        """GMRES on H1 + 1/dt at V with block-Jacobi; with U given the Krylov
        products are MatVecProdFD about U (order fd_order)."""
        V = np.ascontiguousarray(V, dtype=np.float64)
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        fr = C.c_double()
        ierr = C.c_int()
        L = lib()
        L.or_gmres_mvp.restype = C.c_int
        Uc = np.ascontiguousarray(U, dtype=np.float64) if U is not None else None
        its = L.or_gmres_mvp(C.byref(self.m), C.byref(self.p), _d(V), _d(Uc) if Uc is not None else None, _d(dt),
                             int(precision), 1 if Uc is not None else 0, int(fd_order), _d(b), _d(x), int(m),
                             C.c_double(eps), int(maxits), C.byref(fr), C.byref(ierr))
        if ierr.value:
            raise RuntimeError(f"non-physical reconstruction in a finite-difference product ({ierr.value})")
        return x, its, fr.value

    def steady_solve(self, U, precision=0, max_steps=50, eps=1e-8, cfl0=10.0, cfl_max=1e4, ser=0.7, m=30,
                     maxits=30, geps=1e-2, mvp=0, fd_order=1, eisenstat=False):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        hist = np.zeros(max_steps + 1)
        conv, kits = C.c_int(), C.c_int()
        L = lib()
        L.or_steady_solve_ew.restype = C.c_int
        steps = L.or_steady_solve_ew(C.byref(self.m), C.byref(self.p), _d(U), int(precision), int(mvp),
                                     int(fd_order), int(max_steps), C.c_double(eps), C.c_double(cfl0),
                                     C.c_double(cfl_max), C.c_double(ser), int(m), int(maxits), C.c_double(geps),
                                     int(bool(eisenstat)), _d(hist), C.byref(conv), C.byref(kits))
        return dict(U=U, steps=steps, converged=bool(conv.value), krylov_its=kits.value,
                    history=hist[:steps + 1].copy())

    def galerkin(self, V, want_abs=False):
        """The Navier-Stokes Galerkin term alone (raw sums, no 1/V)."""
        V = np.ascontiguousarray(V, dtype=np.float64)
        R = np.zeros((self.N, 5))
        Rabs = np.zeros((self.N, 5)) if want_abs else None
        lib().or_galerkin_term(C.byref(self.m), C.byref(self.p), _d(V), _d(R), _d(Rabs) if want_abs else None)
        return (R, Rabs) if want_abs else R

    def galerkin_jacobian(self, V):
        """The Galerkin Jacobian alone: unscaled D [N,5,5], (A_ij, A_ji) / V_row [E,2,5,5]."""
        V = np.ascontiguousarray(V, dtype=np.float64)
        D = np.zeros((self.N, 5, 5))
        A = np.zeros((self.E, 2, 5, 5))
        lib().or_galerkin_jacobian(C.byref(self.m), C.byref(self.p), _d(V), _d(D), _d(A))
        return D, A

    def node_types(self):
        t = np.empty(self.N, np.int32)
        lib().or_node_types(C.byref(self.m), _i(t))
        return t

    def apply_bcs_residual(self, U, F):
        U = np.ascontiguousarray(U, dtype=np.float64)
        F = np.ascontiguousarray(F, dtype=np.float64).copy()
        lib().or_apply_bcs_residual(C.byref(self.m), C.byref(self.p), _i(self.node_types()), _d(U), _d(F))
        return F

    def apply_bcs_solution(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        lib().or_apply_bcs_solution(C.byref(self.m), C.byref(self.p), _i(self.node_types()), _d(U))
        return U

    def apply_bcs_jacobian(self, U, D, A):
        U = np.ascontiguousarray(U, dtype=np.float64)
        D = np.ascontiguousarray(D, dtype=np.float64).copy()
        A = np.ascontiguousarray(A, dtype=np.float64).copy()
        lib().or_apply_bcs_jacobian(C.byref(self.m), C.byref(self.p), _i(self.node_types()), _d(U), _d(D), _d(A))
        return D, A

    def gradients(self, V):
        V = np.ascontiguousarray(V, dtype=np.float64)
        g = [np.empty((self.N, 5)) for _ in range(3)]
        lib().or_gradients(self.N, self.E, self.m.X, self.m.ptr, self.m.R, _d(V), *[_d(a) for a in g])
        return g


def p1_gradients(Xt):
    """P1 gradients [4, 3] and signed volume of the tetrahedron Xt [4, 3]."""
    X = np.ascontiguousarray(Xt, dtype=np.float64)
    tet = np.arange(4, dtype=np.int32)
    g = np.empty((4, 3))
    L = lib()
    L.or_p1_gradients.restype = C.c_double
    vol = L.or_p1_gradients(_d(X), _i(tet), _d(g))
    return g, vol


def ns_transport(p, T):
    """(mu, lambda, kappa) / Re at the non-dimensional temperature T."""
    o = params(p)
    mu, lam, kap = C.c_double(), C.c_double(), C.c_double()
    lib().or_ns_transport(C.byref(o), C.c_double(T), C.byref(mu), C.byref(lam), C.byref(kap))
    return mu.value, lam.value, kap.value


def ns_volume_term(p, Xt, Vt):
    """r [3, 5] of one element: node coordinates Xt [4, 3], primitive Vt [4, 5]."""
    g, _ = p1_gradients(Xt)
    Vt = np.ascontiguousarray(Vt, dtype=np.float64)
    rows = (dp * 4)(*[Vt[k].ctypes.data_as(dp) for k in range(4)])
    r = np.empty((3, 5))
    o = params(p)
    lib().or_ns_volume_term(C.byref(o), _d(np.ascontiguousarray(g)), rows, _d(r))
    return r


def ns_volume_jacobian(p, Xt, Vt):
    """dRdU [4, 3, 5, 5] = d r[j][m] / d U_k[c] of one element."""
    g, _ = p1_gradients(Xt)
    Vt = np.ascontiguousarray(Vt, dtype=np.float64)
    rows = (dp * 4)(*[Vt[k].ctypes.data_as(dp) for k in range(4)])
    out = np.empty((4, 3, 5, 5))
    o = params(p)
    lib().or_ns_volume_jacobian(C.byref(o), _d(np.ascontiguousarray(g)), rows, _d(out))
    return out


def rcm(N: int, ptr: np.ndarray) -> np.ndarray:
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    renum = np.empty(N, np.int32)
    lib().or_rcm(N, ptr.shape[0], _i(ptr), _i(renum))
    return renum


def ras_ilu0(N: int, ptr: np.ndarray, D: np.ndarray, A: np.ndarray, starts, y: np.ndarray, single=False):
    """IluPrec, fill 0, RAS over the node ranges starts[s]..starts[s+1]:
    x = M^-1 y with M the per-range ILU(0) of the H1 blocks D [N,5,5],
    A [E,2,5,5] (or_jacobian_h1 layout) on the edge table ptr [E,2]."""
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    D = np.ascontiguousarray(D, dtype=np.float64)
    A = np.ascontiguousarray(A, dtype=np.float64)
    st = np.ascontiguousarray(starts, dtype=np.int32)
    y = np.ascontiguousarray(y, dtype=np.float64)
    x = np.empty_like(y)
    rc = lib().or_ras_ilu0_apply(int(N), int(ptr.shape[0]), _i(ptr), _d(D), _d(A), int(st.size - 1), _i(st),
                                 1 if single else 0, _d(y), _d(x))
    if rc:
        raise RuntimeError(f"zero pivot in ILU(0) at node {rc - 1}")
    return x


def color_edges(N: int, ptr: np.ndarray):
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    color = np.empty(ptr.shape[0], np.int32)
    nc = lib().or_color_edges(N, ptr.shape[0], _i(ptr), _i(color))
    return color, nc


def roeflux7(gam, pstiff, n, vn, Vl, Vr, use_ref=False):
    out = np.empty(5)
    n = np.ascontiguousarray(n, dtype=np.float64)
    Vl = np.ascontiguousarray(Vl, dtype=np.float64)
    Vr = np.ascontiguousarray(Vr, dtype=np.float64)
    if use_ref:
        ref().refh_roeflux7(C.c_double(gam), C.c_double(pstiff), _d(n), C.c_double(vn), _d(Vl), _d(Vr), _d(out))
    else:
        lib().or_roeflux7(C.c_double(gam), C.c_double(pstiff), _d(n), C.c_double(vn), _d(Vl), _d(Vr), _d(out))
    return out


def roeflux5(gamma, gam, pstiff, n, vn, Vl, Vr, mach, k1, cmach, shock, irey, length, prec, use_ref=False):
    """ROEFLUX5 (Roe-Turkel low-Mach dissipation), type 0, Ugr = Ug, Udr = Ud."""
    out = np.empty(5)
    n = np.ascontiguousarray(n, dtype=np.float64)
    Vl = np.ascontiguousarray(Vl, dtype=np.float64)
    Vr = np.ascontiguousarray(Vr, dtype=np.float64)
    fn = ref().refh_roeflux5 if use_ref else lib().or_roeflux5
    fn(C.c_double(gamma), C.c_double(gam), C.c_double(pstiff), _d(n), C.c_double(vn), _d(Vl), _d(Vr),
       C.c_double(mach), C.c_double(k1), C.c_double(cmach), C.c_double(shock), C.c_double(irey), C.c_double(length),
       C.c_int(prec), _d(out))
    return out


def boundflux5(gam, n, sig, ua, ub, use_ref=False):
    out = np.empty(5)
    n = np.ascontiguousarray(n, dtype=np.float64)
    ua = np.ascontiguousarray(ua, dtype=np.float64)
    ub = np.ascontiguousarray(ub, dtype=np.float64)
    fn = ref().refh_boundflux5 if use_ref else lib().or_boundflux5
    fn(C.c_double(gam), _d(n), C.c_double(sig), _d(ua), _d(ub), _d(out))
    return out


def vanalbada5(beta, Vi, ddVij, Vj, ddVji, use_ref=False):
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (Vi, ddVij, Vj, ddVji)]
    Vij, Vji = np.empty(5), np.empty(5)
    fn = ref().refh_vanalbada5 if use_ref else lib().or_vanalbada5
    fn(C.c_double(beta), *[_d(x) for x in a], _d(Vij), _d(Vji))
    return Vij, Vji


def wallflux(n, nvel, V):
    out = np.empty(5)
    n = np.ascontiguousarray(n, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    lib().or_wallflux(_d(n), C.c_double(nvel), _d(V), _d(out))
    return out


def roejac6(gam, pstiff, n, vn, Vl, Vr, sw, gamma_diss=1.0, use_ref=False, low_mach=None):
    """ROEJAC6 5x5 block (row-major).  low_mach = (min Mach, slope, cut-off,
    shock reducer, 1/Re, edge length) turns the Roe-Turkel preconditioning
    on (prec 1)."""
    jac = np.empty(25)
    n = np.ascontiguousarray(n, dtype=np.float64)
    Vl = np.ascontiguousarray(Vl, dtype=np.float64)
    Vr = np.ascontiguousarray(Vr, dtype=np.float64)
    if low_mach is not None:
        fn = ref().refh_roejac6_lm if use_ref else lib().or_roejac6_lm
        fn(C.c_double(gamma_diss), C.c_double(gam), C.c_double(pstiff), _d(n), C.c_double(vn), _d(Vl), _d(Vr),
           C.c_int(sw), *[C.c_double(float(x)) for x in low_mach], C.c_int(1), _d(jac))
        return jac.reshape(5, 5)
    if use_ref:
        ref().refh_roejac6(C.c_double(gamma_diss), C.c_double(gam), C.c_double(pstiff), _d(n), C.c_double(vn),
                           _d(Vl), _d(Vr), C.c_int(sw), _d(jac))
    else:
        lib().or_roejac6(C.c_double(gamma_diss), C.c_double(gam), C.c_double(pstiff), _d(n), C.c_double(vn),
                         _d(Vl), _d(Vr), C.c_int(sw), _d(jac))
    return jac.reshape(5, 5)


def invert5(a, single=False, use_ref=False):
    """LINPACK dgefa/dgedi inverse of a row-major 5x5 block; returns (inv, info)."""
    if single:
        b = np.ascontiguousarray(a, dtype=np.float32).reshape(25).copy()
        fn = ref().refh_invert5_f32 if use_ref else lib().or_invert5f
        info = fn(b.ctypes.data_as(C.POINTER(C.c_float)))
    else:
        b = np.ascontiguousarray(a, dtype=np.float64).reshape(25).copy()
        fn = ref().refh_invert5_f64 if use_ref else lib().or_invert5
        info = fn(_d(b))
    return b.reshape(5, 5), info


def apply5(a, y, single=False, use_ref=False):
    y = np.ascontiguousarray(y, dtype=np.float64)
    x = np.empty(5)
    if single:
        b = np.ascontiguousarray(a, dtype=np.float32).reshape(25)
        fn = ref().refh_apply5_f32 if use_ref else lib().or_apply5f
        fn(b.ctypes.data_as(C.POINTER(C.c_float)), _d(y), _d(x))
    else:
        b = np.ascontiguousarray(a, dtype=np.float64).reshape(25)
        fn = ref().refh_apply5_f64 if use_ref else lib().or_apply5
        fn(_d(b), _d(y), _d(x))
    return x
