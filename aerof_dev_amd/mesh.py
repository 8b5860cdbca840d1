"""Host-side mesh preparation through the C ABI (no GPU needed).

Synthetic Kuhn-split boxes and channels (the reference reads SOWER meshes,
GeoSource.C:246-356; these generators are this package's own inputs), the
reference edge table and RCM ordering, the edge-colouring rule and seeded
synthetic states.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._abi import (AEROFX_BC_ADIABATIC_WALL, AEROFX_BC_INLET, AEROFX_BC_OUTLET, AEROFX_BC_SLIP_WALL, BoxSpec, Params,
                   c_double_p, c_int_p, check, lib)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(c_double_p)


def _ip(a: np.ndarray):
    return a.ctypes.data_as(c_int_p)


@dataclass
class LowMachPrec:
    """Spatial low-Mach preconditioner constants (SpatialLowMachPrec,
    LowMachPrec.h; Preconditioner.Mach / k / CutOffMach / ShockReducer) and
    the edges' 1/Re (0 for inviscid flow)."""
    min_mach: float = 0.1
    slope: float = 1.0
    cutoff: float = 1.0
    shock_reducer: float = 0.0
    irey: float = 0.0


@dataclass
class ViscousParams:
    """Laminar Navier-Stokes terms (Equations.Type = NavierStokes, no
    turbulence model, full wall integration): Reynolds number (based on the
    freestream of the AERO-F non-dimensional scaling, rho = 1, |u| = 1,
    p = 1/(gamma M^2)), viscosity model ("constant", "sutherland",
    "prandtl"; ViscoFcn.h), conductivity model ("prandtl": constant Prandtl
    number, "constant"; ThermalCondFcn.h), the temperature of isothermal walls
    (non-dimensional T = p / ((gamma - 1) rho)) and the weight of the viscous
    time-step term (Time.ViscosityParameter, 0 by default)."""
    reynolds: float = 100.0
    viscosity: str = "sutherland"
    bulk_viscosity: float = 0.0
    sutherland_ts: float = 110.4 / 288.15
    conductivity_model: str = "prandtl"
    prandtl: float = 0.72
    conductivity: float = 0.0
    wall_temperature: float = 0.0
    viscous_cst: float = 0.0

    _VISC = {"constant": 0, "sutherland": 1, "prandtl": 2}
    _COND = {"prandtl": 0, "constant": 1}

    def codes(self):
        return self._VISC[self.viscosity], self._COND[self.conductivity_model]


@dataclass
class FlowParams:
    """Perfect-gas Euler parameters (IoDataCore.C:2418-2429 defaults)."""
    gamma: float = 1.4
    pstiff: float = 0.0
    beta: float = 1.0 / 3.0
    second_order: bool = True
    rho_inf: float = 1.0
    p_inf: float = 1.0 / 1.4
    mach: float = 0.5
    alpha: float = 0.0
    yaw: float = 0.0
    # spatial low-Mach preconditioning (Problem.Prec = Preconditioned): the
    # Roe-Turkel flux ROEFLUX5 and its Jacobian; None: Roe's flux
    low_mach: LowMachPrec | None = None
    # laminar Navier-Stokes terms; None: Euler
    viscous: ViscousParams | None = None

    def c(self) -> Params:
        return Params(self.gamma, self.pstiff, self.beta, int(self.second_order), self.rho_inf,
                      self.p_inf, self.mach, self.alpha, self.yaw)


@dataclass
class Mesh:
    X: np.ndarray        # [N,3] float64
    tets: np.ndarray     # [T,4] int32
    faces: np.ndarray    # [F,3] int32
    codes: np.ndarray    # [F] int32
    meta: dict = field(default_factory=dict)

    @property
    def num_nodes(self) -> int:
        return int(self.X.shape[0])

    @property
    def num_tets(self) -> int:
        return int(self.tets.shape[0])

    @property
    def num_faces(self) -> int:
        return int(self.faces.shape[0])


FARFIELD_BOX = (AEROFX_BC_INLET, AEROFX_BC_OUTLET, AEROFX_BC_OUTLET, AEROFX_BC_OUTLET,
                AEROFX_BC_SLIP_WALL, AEROFX_BC_OUTLET)
CHANNEL = (AEROFX_BC_INLET, AEROFX_BC_OUTLET, AEROFX_BC_SLIP_WALL, AEROFX_BC_SLIP_WALL,
           AEROFX_BC_SLIP_WALL, AEROFX_BC_SLIP_WALL)
# channel with no-slip adiabatic walls top and bottom, slip side walls
VISCOUS_CHANNEL = (AEROFX_BC_INLET, AEROFX_BC_OUTLET, AEROFX_BC_SLIP_WALL, AEROFX_BC_SLIP_WALL,
                   AEROFX_BC_ADIABATIC_WALL, AEROFX_BC_ADIABATIC_WALL)


def _box_spec(nx, ny, nz, lo, hi, jitter, seed, face_codes, bumps):
    amp = np.array([b[0] for b in bumps], dtype=np.float64)
    cen = np.array([b[1] for b in bumps], dtype=np.float64)
    wid = np.array([b[2] for b in bumps], dtype=np.float64)
    spec = BoxSpec(nx, ny, nz, (C.c_double * 3)(*lo), (C.c_double * 3)(*hi), float(jitter),
                   int(seed) & 0xFFFFFFFFFFFFFFFF, (C.c_int * 6)(*face_codes), len(bumps),
                   _dp(amp) if len(bumps) else None, _dp(cen) if len(bumps) else None,
                   _dp(wid) if len(bumps) else None)
    return spec, (amp, cen, wid)  # the arrays must outlive the spec


def kuhn_box(nx: int, ny: int, nz: int, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
             jitter: float = 0.0, seed: int = 0, face_codes=FARFIELD_BOX,
             bumps=()) -> Mesh:
    """Kuhn-split hex grid; bumps = [(amplitude, x_center, width), ...] on z-."""
    spec, _keep = _box_spec(nx, ny, nz, lo, hi, jitter, seed, face_codes, bumps)
    N, T, F = C.c_int(), C.c_int(), C.c_int()
    L = lib()
    check(L.aerofx_mesh_box_sizes(C.byref(spec), C.byref(N), C.byref(T), C.byref(F)), "mesh_box_sizes")
    X = np.empty((N.value, 3), np.float64)
    tets = np.empty((T.value, 4), np.int32)
    faces = np.empty((F.value, 3), np.int32)
    codes = np.empty((F.value,), np.int32)
    check(L.aerofx_mesh_box(C.byref(spec), _dp(X), _ip(tets), _ip(faces), _ip(codes)), "mesh_box")
    return Mesh(X, tets, faces, codes, dict(nx=nx, ny=ny, nz=nz, lo=tuple(lo), hi=tuple(hi),
                                            jitter=jitter, seed=seed, bumps=tuple(bumps),
                                            face_codes=tuple(face_codes)))


def bump_derivative(mesh: Mesh, bump: int, param: str) -> np.ndarray:
    """Closed-form dX/dp [N,3] of a kuhn_box mesh for bump `bump`'s
    'amplitude' or 'center', in the mesh's node order (renumbered meshes
    included: renumber() records the permutation)."""
    md = mesh.meta
    spec, _keep = _box_spec(md["nx"], md["ny"], md["nz"], md["lo"], md["hi"], md["jitter"], md["seed"],
                            md["face_codes"], md["bumps"])
    dX = np.empty((mesh.num_nodes, 3), np.float64)
    check(lib().aerofx_mesh_box_bump_derivative(C.byref(spec), int(bump), {"amplitude": 0, "center": 1}[param],
                                                _dp(dX)), "mesh_box_bump_derivative")
    renum = md.get("renum")
    if renum is not None:
        out = np.empty_like(dX)
        out[renum] = dX
        dX = out
    return dX


def edges(mesh: Mesh) -> np.ndarray:
    """Reference edge table [E,2] (EdgeCore.C:55-119)."""
    L = lib()
    E = C.c_int()
    tets = np.ascontiguousarray(mesh.tets, dtype=np.int32)
    check(L.aerofx_mesh_edges(mesh.num_nodes, mesh.num_tets, _ip(tets), C.byref(E), None), "mesh_edges")
    ptr = np.empty((E.value, 2), np.int32)
    check(L.aerofx_mesh_edges(mesh.num_nodes, mesh.num_tets, _ip(tets), C.byref(E), _ip(ptr)), "mesh_edges")
    return ptr


def rcm(mesh: Mesh) -> np.ndarray:
    """Reference RCM permutation renum[old] = new (ConnectivityCore.C:311-635)."""
    renum = np.empty(mesh.num_nodes, np.int32)
    tets = np.ascontiguousarray(mesh.tets, dtype=np.int32)
    check(lib().aerofx_mesh_rcm(mesh.num_nodes, mesh.num_tets, _ip(tets), _ip(renum)), "mesh_rcm")
    return renum


def rcm_gpu(mesh: Mesh) -> np.ndarray:
    """The same RCM permutation with the breadth-first searches on the GPU."""
    renum = np.empty(mesh.num_nodes, np.int32)
    tets = np.ascontiguousarray(mesh.tets, dtype=np.int32)
    check(lib().aerofx_mesh_rcm_gpu(mesh.num_nodes, mesh.num_tets, _ip(tets), _ip(renum)), "mesh_rcm_gpu")
    return renum


def renumber(mesh: Mesh, renum: np.ndarray) -> Mesh:
    X = np.ascontiguousarray(mesh.X.copy())
    tets = np.ascontiguousarray(mesh.tets.copy())
    faces = np.ascontiguousarray(mesh.faces.copy())
    renum = np.ascontiguousarray(renum, dtype=np.int32)
    check(lib().aerofx_mesh_renumber(mesh.num_nodes, mesh.num_tets, mesh.num_faces, _ip(renum), _dp(X),
                                     _ip(tets), _ip(faces)), "mesh_renumber")
    meta = dict(mesh.meta)
    meta["renumbered"] = True
    prev = meta.get("renum")
    meta["renum"] = renum.copy() if prev is None else renum[prev]
    return Mesh(X, tets, faces, mesh.codes.copy(), meta)


def color_edges(num_nodes: int, ptr: np.ndarray):
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    color = np.empty(ptr.shape[0], np.int32)
    nc = C.c_int()
    check(lib().aerofx_mesh_color_edges(num_nodes, ptr.shape[0], _ip(ptr), _ip(color), C.byref(nc)),
          "color_edges")
    return color, nc.value


def color_edges_gpu(num_nodes: int, ptr: np.ndarray):
    """Edge colouring on the GPU (Jones-Plassmann rounds, same colours as color_edges)."""
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    color = np.empty(ptr.shape[0], np.int32)
    nc = check(lib().aerofx_mesh_color_edges_gpu(num_nodes, ptr.shape[0], _ip(ptr), _ip(color)),
               "color_edges_gpu")
    return color, nc


def init_state(params: FlowParams, X: np.ndarray, kind: str = "freestream", amp: float = 0.0,
               seed: int = 0) -> np.ndarray:
    """Conservative U[N,5]: 'freestream', 'noise' (Gaussian on rho and p) or 'sine'."""
    k = {"freestream": 0, "noise": 1, "sine": 2}[kind]
    X = np.ascontiguousarray(X, dtype=np.float64)
    U = np.empty((X.shape[0], 5), np.float64)
    p = params.c()
    check(lib().aerofx_state_init(C.byref(p), k, float(amp), int(seed) & 0xFFFFFFFFFFFFFFFF, X.shape[0],
                                  _dp(X), _dp(U)), "state_init")
    return U


# ---- the workloads (sizes and seeds) -------------------------------------

def box_mesh(n: int, jitter: float = 0.15, seed: int = 1234, rcm_order: bool = False) -> Mesh:
    """Unit box, n^3 hexes, farfield on 5 faces and a slip wall at z = 0."""
    m = kuhn_box(n, n, n, jitter=jitter, seed=seed)
    if rcm_order:
        m = renumber(m, rcm(m))
    return m


def channel_mesh(nx: int, ny: int, nz: int, bump_height: float = 0.05, bump_width: float = 0.2,
                 jitter: float = 0.0, seed: int = 0, face_codes=CHANNEL) -> Mesh:
    """Channel [0,3]x[0,1]x[0,1] with a Gaussian bump on the bottom wall
    (face_codes=VISCOUS_CHANNEL: no-slip adiabatic top and bottom walls)."""
    return kuhn_box(nx, ny, nz, lo=(0.0, 0.0, 0.0), hi=(3.0, 1.0, 1.0), jitter=jitter, seed=seed,
                    face_codes=face_codes, bumps=((bump_height, 1.5, bump_width),))
