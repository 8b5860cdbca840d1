"""Edge cases on the device: the smallest meshes, signed boundary codes,
state and argument errors of the implicit entries, and that a context keeps
working after rejected calls.  Every rejected call is refused on the host
before any kernel is launched.
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
from aerof_dev_amd import FlowParams, init_state
from aerof_dev_amd import mesh as M
from aerof_dev_amd._abi import AEROFX_BC_INLET, AEROFX_BC_OUTLET, AerofxError

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _residual(op, U, gpu):
    import torch
    Ud = torch.from_numpy(np.ascontiguousarray(U)).to(gpu)
    Rd = torch.empty_like(Ud)
    ierr = op.compute_residual(Ud, Rd)
    torch.cuda.synchronize()
    return ierr, Rd.cpu().numpy()


def _check(Rg, ref):
    R, Rabs = ref["R"], ref["Rabs"]
    assert np.linalg.norm(Rg - R) <= TOL * np.linalg.norm(R)
    assert np.max(np.abs(Rg - R) / np.maximum(Rabs, 1e-300)) <= TOL


@pytest.mark.parametrize("second_order", [True, False])
def test_single_hex_box(gpu, second_order):
    """One hex, six tets, eight nodes, every node on the boundary."""
    from aerof_dev_amd.operators import SpaceOperator
    m = M.box_mesh(1, jitter=0.0)
    assert (m.num_nodes, m.num_tets) == (8, 6)
    p = FlowParams(mach=0.5, alpha=0.1, second_order=second_order)
    U = init_state(p, m.X, "noise", 0.01, seed=1)
    ref = O.OracleMesh(m, p).residual(U, want_abs=True)
    ierr, Rg = _residual(SpaceOperator(m, p), U, gpu)
    assert ierr == ref["ierr"] == 0
    _check(Rg, ref)


def test_thin_channel_inlet_outlet(gpu):
    """A channel one hex high and wide: inlet, outlet and walls meet at every node."""
    from aerof_dev_amd.operators import SpaceOperator
    m = M.channel_mesh(4, 1, 1, bump_height=0.05, bump_width=0.5)
    p = FlowParams(mach=0.675, alpha=0.02)
    U = init_state(p, m.X, "noise", 0.01, seed=2)
    ref = O.OracleMesh(m, p).residual(U, want_abs=True)
    ierr, Rg = _residual(SpaceOperator(m, p), U, gpu)
    assert ierr == 0
    _check(Rg, ref)


def test_signed_face_codes(gpu):
    """A boundary code counts by its magnitude: negating the inlet/outlet codes
    leaves the residual bit-identical."""
    from aerof_dev_amd.operators import SpaceOperator
    m = M.channel_mesh(5, 2, 2)
    neg = np.where((m.codes == AEROFX_BC_INLET) | (m.codes == AEROFX_BC_OUTLET), -m.codes, m.codes)
    m2 = dataclasses.replace(m, codes=neg.astype(np.int32))
    p = FlowParams(mach=0.675)
    U = init_state(p, m.X, "noise", 0.01, seed=3)
    _, R1 = _residual(SpaceOperator(m, p), U, gpu)
    _, R2 = _residual(SpaceOperator(m2, p), U, gpu)
    assert np.array_equal(R1, R2)


def test_isolated_node_rejected(gpu):
    """A node that no tet references has no control volume."""
    from aerof_dev_amd.operators import SpaceOperator
    m = M.box_mesh(2, jitter=0.1, seed=4)
    X = np.vstack([m.X, [[5.0, 5.0, 5.0]]])
    with pytest.raises(AerofxError, match="non-positive control volume"):
        SpaceOperator(dataclasses.replace(m, X=X), FlowParams())


def test_rejected_calls_leave_the_context_working(gpu):
    import torch
    from aerof_dev_amd.operators import (BlockGmresSolver, GmresSolver, MatVecProdFD, MatVecProdH1, SpaceOperator,
                                         adjoint_solve, steady_solve)
    m = M.box_mesh(3, jitter=0.1, seed=6)
    p = FlowParams(mach=0.5)
    U = init_state(p, m.X, "noise", 0.01, seed=6)
    op = SpaceOperator(m, p)
    _, R0 = _residual(op, U, gpu)
    Ud = torch.from_numpy(U).to(gpu)
    v = torch.ones_like(Ud)
    w = torch.empty_like(Ud)
    h1 = MatVecProdH1(op)
    # operators used before they are set up
    with pytest.raises(AerofxError, match="not set up"):
        h1.apply(v, w)
    with pytest.raises(AerofxError, match="not set up"):
        h1.prec_apply(v, w)
    with pytest.raises(AerofxError, match="not set up"):
        GmresSolver(op).solve(v, w)
    # argument ranges
    with pytest.raises(AerofxError, match="precision 0 or 1"):
        h1.evaluate(Ud, precision=2)
    h1.evaluate(Ud, precision=1)
    B = torch.ones((2, m.num_nodes, 5), dtype=torch.float64, device=gpu)
    X = torch.zeros_like(B)
    with pytest.raises(AerofxError, match="needs FP64 operators"):
        BlockGmresSolver(op, m=4).solve(B, X)
    h1.evaluate(Ud, precision=0)
    for bad_m in (0, 32):
        with pytest.raises(AerofxError, match="Krylov dimension"):
            GmresSolver(op, m=bad_m).solve(v, w)
    B17 = torch.ones((17, m.num_nodes, 5), dtype=torch.float64, device=gpu)
    with pytest.raises(ValueError, match="1 to 16 right-hand sides"):
        BlockGmresSolver(op, m=4).solve(B17, torch.zeros_like(B17))
    from aerof_dev_amd._abi import lib as _lib
    st = __import__("aerof_dev_amd._abi", fromlist=["BlockStats"]).BlockStats()
    import ctypes as _C
    with pytest.raises(AerofxError, match="nrhs"):
        from aerof_dev_amd._abi import check as _check
        _check(_lib().aerofx_block_gmres(op._h, 17, B17.data_ptr(), torch.zeros_like(B17).data_ptr(), 4, 1e-8, 10, 0,
                                         0, _C.byref(st)), "block_gmres")
    with pytest.raises(AerofxError, match="1 <= m <= 16"):
        BlockGmresSolver(op, m=17).solve(B, X)
    with pytest.raises(AerofxError, match="precision 0 or 1"):
        BlockGmresSolver(op, m=4, precision=2).solve(B, X)
    fd = MatVecProdFD(op, order=3)
    with pytest.raises(AerofxError, match="order 1 or 2"):
        fd.evaluate(Ud)
    with pytest.raises(AerofxError, match="order 1 or 2"):
        fd.standalone_apply(Ud, v, w)
    with pytest.raises(AerofxError, match="precision 0 or 1"):
        steady_solve(op, Ud.clone(), precision=2, max_steps=2)
    with pytest.raises(AerofxError, match="precision 0 or 1"):
        adjoint_solve(op, Ud, v, torch.zeros_like(v), precision=2)
    # the context still computes what a fresh one computes
    _, R1 = _residual(op, U, gpu)
    assert np.array_equal(R0, R1)
    dt = torch.empty(m.num_nodes, dtype=torch.float64, device=gpu)
    h1.timestep(Ud, 10.0, dt)
    h1.evaluate(Ud, dt, precision=0)
    its, res = GmresSolver(op, m=20, eps=1e-8, maxits=300).solve(v, w)
    assert 0 < its <= 300 and res <= 1e-8 * float(torch.linalg.norm(v))


def _single_tet():
    """One tet (4 nodes, 20 unknowns), its faces oriented outward: two
    farfield faces and two slip walls."""
    X = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 1.0, 0.05], [0.1, 0.2, 1.0]])
    tets = np.array([[0, 1, 2, 3]], np.int32)
    faces = []
    cen = X.mean(axis=0)
    for f in ([1, 2, 3], [0, 3, 2], [0, 1, 3], [0, 2, 1]):
        a, b, c = X[f]
        if np.dot(np.cross(b - a, c - a), (a + b + c) / 3 - cen) < 0:
            f = [f[0], f[2], f[1]]
        faces.append(f)
    codes = np.array([AEROFX_BC_OUTLET, AEROFX_BC_INLET, 2, 2], np.int32)
    return M.Mesh(X, tets, np.array(faces, np.int32), codes)


@pytest.mark.parametrize("precision", [0, 1])
def test_gmres_exhausts_small_krylov_space(gpu, precision):
    """20 unknowns, GMRES(30), tolerance 1e-14: the Krylov space is exhausted
    (or the tolerance met) inside the first cycle; the solution must solve the
    system that the H1 operator applies, built here column by column."""
    import torch
    from aerof_dev_amd.operators import GmresSolver, MatVecProdH1, SpaceOperator
    m = _single_tet()
    p = FlowParams(mach=0.5, alpha=0.1)
    U = init_state(p, m.X, "noise", 0.01, seed=9)
    op = SpaceOperator(m, p)
    Ud = torch.from_numpy(U).to(gpu)
    h1 = MatVecProdH1(op)
    dt = torch.empty(m.num_nodes, dtype=torch.float64, device=gpu)
    h1.timestep(Ud, 10.0, dt)
    h1.evaluate(Ud, dt, precision=precision)
    n = 5 * m.num_nodes
    A = np.empty((n, n))
    for k in range(n):
        e = torch.zeros(n, dtype=torch.float64, device=gpu)
        e[k] = 1.0
        y = torch.empty_like(e)
        h1.apply(e.view(-1, 5), y.view(-1, 5))
        A[:, k] = y.cpu().numpy()
    b = np.random.default_rng(4).standard_normal(n)
    x = torch.empty(n, dtype=torch.float64, device=gpu)
    its, res = GmresSolver(op, m=30, eps=1e-14, maxits=60).solve(torch.from_numpy(b).to(gpu).view(-1, 5),
                                                                x.view(-1, 5))
    xs = np.linalg.solve(A, b)
    xg = x.cpu().numpy()
    assert its <= 30, its
    assert np.linalg.norm(A @ xg - b) <= 1e-10 * np.linalg.norm(b)
    assert np.linalg.norm(xg - xs) <= 1e-8 * np.linalg.norm(xs)


def _cone(k):
    """k tets around the axis from the base centre to the apex (node 0):
    the apex has k + 1 neighbours, all numbered after it."""
    ang = 2.0 * np.pi * np.arange(k) / k
    rim = np.stack([np.cos(ang), np.sin(ang), np.zeros(k)], axis=1)
    X = np.vstack([[0.0, 0.0, 1.0], [0.0, 0.0, 0.0], rim])
    tets, faces = [], []
    for i in range(k):
        a, b = 2 + i, 2 + (i + 1) % k
        t = [0, 1, a, b]
        q = X[t]
        if np.dot(np.cross(q[1] - q[0], q[2] - q[0]), q[3] - q[0]) < 0:
            t = [0, 1, b, a]
        tets.append(t)
        for f in ([0, a, b], [1, b, a]):
            r = X[f]
            if np.dot(np.cross(r[1] - r[0], r[2] - r[0]), (r[0] + r[1] + r[2]) / 3.0 - [0.0, 0.0, 0.4]) < 0:
                f = [f[0], f[2], f[1]]
            faces.append(f)
    return M.Mesh(X, np.array(tets, np.int32), np.array(faces, np.int32),
                  np.full(len(faces), AEROFX_BC_INLET, np.int32))


@pytest.mark.parametrize("second_order", [True, False])
def test_node_with_many_upper_edges(gpu, second_order):
    """A node with 81 upper edges: more than the 64-edge flux blocks hold, so
    the context takes 128-edge blocks; the residual still matches."""
    from aerof_dev_amd.operators import SpaceOperator
    m = _cone(80)
    p = FlowParams(mach=0.5, alpha=0.1, second_order=second_order)
    U = init_state(p, m.X, "noise", 0.01, seed=12)
    ref = O.OracleMesh(m, p).residual(U, want_abs=True)
    assert ref["ierr"] == 0
    ierr, Rg = _residual(SpaceOperator(m, p), U, gpu)
    assert ierr == 0
    _check(Rg, ref)
