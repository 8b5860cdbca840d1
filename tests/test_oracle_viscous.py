"""Pins of the oracle's laminar Navier-Stokes terms (CPU only).

The reference's own tests hold no vectors for these terms, and its C++
(FemEquationTermDesc.C, NavierStokesTerm.h) needs IoData and the VarFcn
hierarchy to build, so each function is pinned by what the mathematics
fixes:
  * P1 gradients reproduce linear fields exactly and sum to zero;
  * transport coefficients against the closed forms of ViscoFcn.h /
    ThermalCondFcn.h (Sutherland's law gives mu = 1 at the freestream of
    the AERO-F scaling);
  * the volume term against the Newtonian stress and Fourier flux written
    independently with numpy for linear velocity / temperature fields;
  * the assembled Galerkin term: conservation (sums to zero over the mesh)
    and zero interior momentum for a linear velocity field, and the
    P1 stiffness of a quadratic field on a regular Kuhn grid equals
    -V_i mu u'' at interior nodes;
  * the element Jacobian and the assembled Galerkin Jacobian against
    central differences of the volume term / assembled term;
  * node types (priority rule) and the strong wall conditions.
"""
import numpy as np
import pytest

import oracle as O
from aerof_dev_amd import FlowParams, ViscousParams
from aerof_dev_amd import mesh as M

GAM = 1.4


def ns_params(mach=0.3, re=100.0, second_order=True, **kw):
    """AERO-F's non-dimensional freestream: rho = 1, |u| = 1, p = 1/(gam M^2)."""
    return FlowParams(mach=mach, p_inf=1.0 / (GAM * mach * mach), second_order=second_order,
                      viscous=ViscousParams(reynolds=re, **kw))


def rand_tet(rng):
    while True:
        X = rng.standard_normal((4, 3))
        if np.linalg.det(X[1:] - X[0]) > 0.2:
            return X


def prim_from(rho, u, T):
    """Primitive (rho, u, p) with T = p / ((gam - 1) rho)."""
    return np.array([rho, u[0], u[1], u[2], (GAM - 1.0) * rho * T])


def cons_from_prim(V):
    rho, u, p = V[0], V[1:4], V[4]
    return np.array([rho, rho * u[0], rho * u[1], rho * u[2], p / (GAM - 1.0) + 0.5 * rho * u @ u])


def prim_from_cons(U):
    rho = U[0]
    u = U[1:4] / rho
    return np.array([rho, u[0], u[1], u[2], (GAM - 1.0) * (U[4] - 0.5 * rho * u @ u)])


def test_p1_gradients_reproduce_linear_fields():
    rng = np.random.default_rng(1)
    for _ in range(20):
        X = rand_tet(rng)
        g, vol = O.p1_gradients(X)
        assert vol == pytest.approx(np.linalg.det(X[1:] - X[0]) / 6.0, rel=1e-13)
        a, b = rng.standard_normal(), rng.standard_normal(3)
        f = a + X @ b
        assert np.allclose(g.T @ f, b, rtol=0, atol=1e-12 * np.abs(b).max())
        assert np.allclose(g.sum(axis=0), 0.0, atol=1e-12 * np.abs(g).max())
    # a mirrored tetrahedron has the negative volume and the same gradients
    X = rand_tet(rng)
    g, vol = O.p1_gradients(X)
    g2, vol2 = O.p1_gradients(X[[0, 2, 1, 3]])
    assert vol2 == pytest.approx(-vol, rel=1e-13)
    assert np.allclose(g2, g[[0, 2, 1, 3]], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("model", ["constant", "sutherland", "prandtl"])
@pytest.mark.parametrize("cond", ["prandtl", "constant"])
def test_transport_closed_forms(model, cond):
    mach, re, pr, ts, bulk, kc = 0.4, 250.0, 0.7, 0.38, 0.05, 0.9
    p = ns_params(mach, re, viscosity=model, conductivity_model=cond, prandtl=pr, sutherland_ts=ts,
                  bulk_viscosity=bulk, conductivity=kc)
    alpha = GAM * (GAM - 1.0) * mach * mach
    Tinf = 1.0 / alpha  # freestream temperature of the AERO-F scaling
    for T in (Tinf, 1.7 * Tinf, 0.6 * Tinf):
        mu, lam, kap = O.ns_transport(p, T)
        th = alpha * T
        mu0 = {"constant": 1.0, "sutherland": th ** 1.5 * (1.0 + ts) / (th + ts), "prandtl": th}[model]
        lam0 = (bulk - 2.0 / 3.0 * mu0) if model == "constant" else -2.0 / 3.0 * mu0
        kap0 = GAM * mu0 / pr if cond == "prandtl" else kc
        assert mu == pytest.approx(mu0 / re, rel=1e-14)
        assert lam == pytest.approx(lam0 / re, rel=1e-14, abs=1e-16)
        assert kap == pytest.approx(kap0 / re, rel=1e-14)
    if model == "sutherland":
        assert O.ns_transport(p, Tinf)[0] == pytest.approx(1.0 / re, rel=1e-14)


@pytest.mark.parametrize("model", ["constant", "sutherland"])
def test_volume_term_newtonian_stress(model):
    """r for linear u and T against tau = mu (B + B^T) + lambda tr(B) I and
    q = -kappa grad T, at the element's mean velocity and temperature."""
    rng = np.random.default_rng(7)
    p = ns_params(0.5, 80.0, viscosity=model, bulk_viscosity=0.1)
    for _ in range(10):
        X = rand_tet(rng)
        u0, B = rng.standard_normal(3), 0.3 * rng.standard_normal((3, 3))
        T0, a = 2.0 + rng.random(), 0.2 * rng.standard_normal(3)
        rho = 0.8 + 0.4 * rng.random(4)
        u = u0 + X @ B.T
        T = T0 + X @ a
        V = np.array([prim_from(rho[k], u[k], T[k]) for k in range(4)])
        r = O.ns_volume_term(p, X, V)
        mu, lam, kap = O.ns_transport(p, T.mean())
        tau = mu * (B + B.T) + lam * np.trace(B) * np.eye(3)
        q = -kap * a
        ucg = u.mean(axis=0)
        ref = np.zeros((3, 5))
        ref[:, 1:4] = tau  # r[j][1 + i] = tau_ij (symmetric)
        ref[:, 4] = ucg @ tau - q
        scale = np.abs(ref).max()
        assert np.allclose(r, ref, rtol=0, atol=1e-12 * scale)


def test_volume_jacobian_matches_central_differences():
    rng = np.random.default_rng(3)
    for model in ("constant", "sutherland", "prandtl"):
        p = ns_params(0.3, 50.0, viscosity=model)
        X = rand_tet(rng)
        V = np.array([prim_from(1.0 + 0.2 * rng.random(), 0.4 * rng.standard_normal(3), 20.0 + rng.random())
                      for _ in range(4)])
        U = np.array([cons_from_prim(v) for v in V])
        J = O.ns_volume_jacobian(p, X, V)
        for k in range(4):
            for c in range(5):
                h = 1e-6 * max(1.0, abs(U[k, c]))
                Up, Um = U.copy(), U.copy()
                Up[k, c] += h
                Um[k, c] -= h
                rp = O.ns_volume_term(p, X, np.array([prim_from_cons(x) for x in Up]))
                rm = O.ns_volume_term(p, X, np.array([prim_from_cons(x) for x in Um]))
                fd = (rp - rm) / (2 * h)
                scale = max(np.abs(fd).max(), 1e-12)
                assert np.abs(J[k, :, :, c] - fd).max() <= 2e-6 * scale, (model, k, c)


def _viscous_box(n=5, jitter=0.0):
    return M.kuhn_box(n, n, n, jitter=jitter, seed=11, face_codes=M.VISCOUS_CHANNEL)


def _state(m, p, u_fn, T_fn, rho=1.0):
    V = np.empty((m.num_nodes, 5))
    for i, x in enumerate(m.X):
        V[i] = prim_from(rho, u_fn(x), T_fn(x))
    return V


def test_galerkin_conservation_and_linear_fields():
    m = _viscous_box(5, jitter=0.2)
    p = ns_params(0.3, 40.0)
    om = O.OracleMesh(m, p)
    rng = np.random.default_rng(5)
    Tinf = 1.0 / (GAM * (GAM - 1.0) * 0.09)
    # arbitrary smooth field: the term sums to zero over the mesh
    V = _state(m, p, lambda x: np.array([np.sin(x[1]), x[0] * x[2], np.cos(x[0])]),
               lambda x: Tinf * (1.0 + 0.1 * x[2]))
    R, Rabs = om.galerkin(V, want_abs=True)
    assert np.abs(R.sum(axis=0)).max() <= 1e-13 * Rabs.sum()
    # linear velocity, constant temperature: zero momentum at interior nodes
    B = rng.standard_normal((3, 3))
    V = _state(m, p, lambda x: B @ x, lambda x: Tinf)
    R, Rabs = om.galerkin(V, want_abs=True)
    interior = np.all((m.X > 1e-9) & (m.X < 1 - 1e-9), axis=1)
    assert interior.sum() > 0
    assert np.abs(R[interior, 1:4]).max() <= 1e-12 * Rabs[interior, 1:4].max()
    # constant velocity, linear temperature, constant kappa: zero energy at
    # interior nodes
    a = rng.standard_normal(3)
    omc = O.OracleMesh(m, ns_params(0.3, 40.0, viscosity="constant"))
    V = _state(m, p, lambda x: np.array([0.3, -0.1, 0.2]), lambda x: Tinf * (1.0 + 0.05 * (a @ x)))
    R, Rabs = omc.galerkin(V, want_abs=True)
    assert np.abs(R[interior]).max() <= 1e-12 * Rabs[interior].max()


def test_galerkin_quadratic_field_on_regular_grid():
    """P1 stiffness of u_x = y^2 on a regular Kuhn grid, constant mu and rho:
    R_i[1] = int tau_x . grad phi_i = -2 mu V_i at interior nodes."""
    m = _viscous_box(6, jitter=0.0)
    p = ns_params(0.3, 40.0, viscosity="constant")
    om = O.OracleMesh(m, p)
    Tinf = 1.0 / (GAM * (GAM - 1.0) * 0.09)
    V = _state(m, p, lambda x: np.array([x[1] ** 2, 0.0, 0.0]), lambda x: Tinf)
    R = om.galerkin(V)
    mu = O.ns_transport(p, Tinf)[0]
    cv = om.tables()["ctrl_vol"]
    interior = np.all((m.X > 1e-9) & (m.X < 1 - 1e-9), axis=1)
    assert np.allclose(R[interior, 1], -2.0 * mu * cv[interior], rtol=1e-10, atol=0)
    assert np.abs(R[interior, 2:4]).max() <= 1e-12 * mu * cv.max()


def test_galerkin_jacobian_matches_central_differences():
    m = _viscous_box(4, jitter=0.2)
    p = ns_params(0.3, 30.0)
    om = O.OracleMesh(m, p)
    rng = np.random.default_rng(9)
    Tinf = 1.0 / (GAM * (GAM - 1.0) * 0.09)
    V = _state(m, p, lambda x: np.array([1.0 + 0.3 * np.sin(3 * x[1]), 0.2 * x[0], -0.1 * x[2]]),
               lambda x: Tinf * (1.0 + 0.1 * np.cos(2 * x[0])), rho=1.0)
    V[:, 0] *= 1.0 + 0.05 * rng.random(m.num_nodes)
    U = np.array([cons_from_prim(v) for v in V])
    D, A = om.galerkin_jacobian(V)
    cv = om.tables()["ctrl_vol"]
    ptr = om.tables()["ptr"]
    v = rng.standard_normal(U.shape) * np.abs(U).mean(axis=0)
    # J v with the unscaled diagonal and the off-diagonal blocks times V_row
    Jv = np.einsum("nij,nj->ni", D, v)
    for l, (i, j) in enumerate(ptr):
        Jv[i] += cv[i] * A[l, 0] @ v[j]
        Jv[j] += cv[j] * A[l, 1] @ v[i]
    h = 1e-6
    Rp = om.galerkin(np.array([prim_from_cons(x) for x in U + h * v]))
    Rm = om.galerkin(np.array([prim_from_cons(x) for x in U - h * v]))
    fd = (Rp - Rm) / (2 * h)
    assert np.linalg.norm(Jv - fd) <= 1e-6 * np.linalg.norm(fd)


def test_node_types_and_wall_conditions():
    m = _viscous_box(4)
    p = ns_params(0.3, 30.0, wall_temperature=5.0)
    om = O.OracleMesh(m, p)
    t = om.node_types()
    z = m.X[:, 2]
    on_wall = (np.abs(z) < 1e-12) | (np.abs(z - 1.0) < 1e-12)
    # walls (priority 10) win over inlet / outlet / slip faces on the edges
    assert np.all(t[on_wall] == 3)
    # slip side walls (priority 8) win over the inlet (priority 2)
    x0 = (np.abs(m.X[:, 0]) < 1e-12) & ~on_wall
    yb = (np.abs(m.X[:, 1]) < 1e-12) | (np.abs(m.X[:, 1] - 1.0) < 1e-12)
    assert np.all(t[x0 & ~yb] == 4) and np.all(t[x0 & yb] == 2)
    rng = np.random.default_rng(2)
    U = np.abs(rng.standard_normal((m.num_nodes, 5))) + 1.0
    F = rng.standard_normal(U.shape)
    Fb = om.apply_bcs_residual(U, F)
    assert np.array_equal(Fb[on_wall, 1:4], U[on_wall, 1:4])
    assert np.array_equal(Fb[~on_wall], F[~on_wall])
    assert np.array_equal(Fb[:, [0, 4]], F[:, [0, 4]])
    Ub = om.apply_bcs_solution(U)
    assert np.all(Ub[on_wall, 1:4] == 0.0) and np.array_equal(Ub[~on_wall], U[~on_wall])
    D = rng.standard_normal((m.num_nodes, 5, 5))
    A = rng.standard_normal((om.E, 2, 5, 5))
    Db, Ab = om.apply_bcs_jacobian(U, D, A)
    eye = np.eye(5)
    for i in np.nonzero(on_wall)[0]:
        assert np.array_equal(Db[i, 1:4], eye[1:4]) and np.array_equal(Db[i, [0, 4]], D[i, [0, 4]])
    ptr = om.tables()["ptr"]
    for l, (i, j) in enumerate(ptr):
        for s, row in ((0, i), (1, j)):
            if on_wall[row]:
                assert np.all(Ab[l, s, 1:4] == 0.0) and np.array_equal(Ab[l, s, [0, 4]], A[l, s, [0, 4]])
            else:
                assert np.array_equal(Ab[l, s], A[l, s])


def test_isothermal_wall_conditions():
    codes = list(M.VISCOUS_CHANNEL)
    codes[4] = 1  # isothermal bottom wall
    m = M.kuhn_box(4, 4, 4, face_codes=tuple(codes))
    Tw = 7.5
    p = ns_params(0.3, 30.0, wall_temperature=Tw)
    om = O.OracleMesh(m, p)
    t = om.node_types()
    bottom = np.abs(m.X[:, 2]) < 1e-12
    assert np.all(t[bottom] == 1)
    rng = np.random.default_rng(4)
    U = np.abs(rng.standard_normal((m.num_nodes, 5))) + 1.0
    Ub = om.apply_bcs_solution(U)
    Vb = np.array([prim_from_cons(x) for x in Ub[bottom]])
    assert np.allclose(Vb[:, 4] / ((GAM - 1.0) * Vb[:, 0]), Tw, rtol=1e-14)
    # the energy residual of a wall state at the wall temperature vanishes
    F = om.apply_bcs_residual(Ub, np.ones_like(U))
    assert np.abs(F[bottom, 4]).max() <= 1e-13 * np.abs(Ub[bottom, 4] * Ub[bottom, 0]).max()


def test_viscous_residual_includes_galerkin_before_scaling():
    m = _viscous_box(4, jitter=0.1)
    pe = FlowParams(mach=0.3, p_inf=1.0 / (GAM * 0.09))
    pv = ns_params(0.3, 30.0)
    rng = np.random.default_rng(8)
    Tinf = 1.0 / (GAM * (GAM - 1.0) * 0.09)
    V = _state(m, pv, lambda x: np.array([1.0 + 0.1 * np.sin(4 * x[2]), 0.0, 0.05 * x[0]]),
               lambda x: Tinf * (1.0 + 0.02 * rng.standard_normal()))
    U = np.array([cons_from_prim(v) for v in V])
    Re = O.OracleMesh(m, pe).residual(U)["R"]
    omv = O.OracleMesh(m, pv)
    Rv = omv.residual(U)["R"]
    G = omv.galerkin(V)
    cv = omv.tables()["ctrl_vol"]
    assert np.allclose(Rv, Re + G / cv[:, None], rtol=1e-12, atol=1e-12 * np.abs(Rv).max())


def test_viscous_timestep_scaling():
    """The elemental viscous time-step term (ElemSet::computeTimeStep) is
    9 mu / rho |T| |grad phi_j|^2 summed over the elements: it scales like
    the mesh size (|T| ~ h^3, |grad phi|^2 ~ h^-2) and like 1/Re, so the
    viscous time step cfl V / idtv scales like h^2 Re."""
    import dataclasses
    m = _viscous_box(4, jitter=0.2)
    Tinf = 1.0 / (GAM * (GAM - 1.0) * 0.09)

    def idtv(scale, re):
        mm = dataclasses.replace(m, X=m.X * scale)
        pv = ns_params(0.3, re, viscous_cst=1.0)
        pe = FlowParams(mach=0.3, p_inf=1.0 / (GAM * 0.09))
        V = _state(mm, pv, lambda x: np.zeros(3), lambda x: Tinf)
        omv, ome = O.OracleMesh(mm, pv), O.OracleMesh(mm, pe)
        cv = omv.tables()["ctrl_vol"]
        dv, de = omv.timestep(V, 1.0), ome.timestep(V, 1.0)
        return cv / dv - cv / de  # = idtv (the inviscid part cancels)
    a, b, c = idtv(1.0, 50.0), idtv(2.0, 50.0), idtv(1.0, 100.0)
    assert np.all(a > 0)
    assert np.allclose(b, 2.0 * a, rtol=1e-10)
    assert np.allclose(c, 0.5 * a, rtol=1e-10)
