"""Nodes with more neighbours than a block holds: a cone of 300 tets around
its axis (the apex and the base centre have 301 neighbours, more than the
256 entries of an H1 / Jacobian block, and 301 and 300 upper edges, more
than the 128 edges of a flux block), numbered with the two hubs in the
middle of the node order.  Long rows go to the long-row kernels (Jacobian,
H1 and its transpose) and are cut over several flux blocks (their flux sum
is formed by the node pass); results match the oracle as on any mesh, the
Newton history tracks the oracle's, and the batched direct solve falls back
to one inner GMRES per right-hand side (the batched H1 apply takes nodes of
at most 64 neighbours)."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from aerof_dev_amd import FlowParams, init_state
from aerof_dev_amd import mesh as M

pytestmark = pytest.mark.gpu
sys.path.insert(0, os.path.dirname(__file__))
from test_gpu_edge_cases import _cone  # noqa: E402


def _mesh():
    m = _cone(300)
    N = m.num_nodes
    # hubs (old 0 and 1) to the middle of the numbering: renum[old] = new
    order = list(range(2, 150)) + [0, 1] + list(range(150, N))
    renum = np.empty(N, dtype=np.int32)
    renum[np.array(order)] = np.arange(N, dtype=np.int32)
    return M.renumber(m, renum)


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("second_order", [True, False])
def test_high_degree_residual(gpu, second_order):
    import torch
    from aerof_dev_amd.operators import SpaceOperator
    m = _mesh()
    p = FlowParams(mach=0.5, alpha=0.1, second_order=second_order)
    U = init_state(p, m.X, "noise", 0.01, seed=12)
    ref = O.OracleMesh(m, p).residual(U, want_abs=True)
    op = SpaceOperator(m, p)
    Ud = torch.from_numpy(U).to(gpu)
    R = torch.empty_like(Ud)
    assert op.compute_residual(Ud, R) == ref["ierr"] == 0
    Rg = R.cpu().numpy()
    assert np.max(np.abs(Rg - ref["R"]) / np.maximum(ref["Rabs"], 1e-300)) <= 1e-12
    op.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_high_degree_jacobian_and_products(gpu, precision):
    import torch
    from aerof_dev_amd.operators import MatVecProdH1, SpaceOperator
    m = _mesh()
    p = FlowParams(mach=0.5, alpha=0.1)
    U = init_state(p, m.X, "noise", 0.01, seed=13)
    om = O.OracleMesh(m, p)
    V = om.residual(U)["V"]
    dt = om.timestep(V, 20.0)
    Dr, Ar = om.jacobian(V, dt)
    op = SpaceOperator(m, p)
    h1 = MatVecProdH1(op)
    h1.evaluate(torch.from_numpy(U).to(gpu), torch.from_numpy(dt).to(gpu), precision)
    D, A = h1.blocks()
    tol = 1e-12 if precision == 0 else 1e-6
    assert np.max(np.abs(D - Dr) / np.abs(Dr).max(axis=(1, 2))[:, None, None]) <= tol
    assert np.max(np.abs(A - Ar) / (np.abs(Ar).max(axis=(2, 3))[:, :, None, None] + 1e-300)) <= tol
    # products with the blocks as stored, against the dense operator
    N = op.num_nodes
    H = np.zeros((5 * N, 5 * N))
    for n in range(N):
        H[5 * n:5 * n + 5, 5 * n:5 * n + 5] = D[n]
    for l, (i, j) in enumerate(om.tables()["ptr"]):
        H[5 * i:5 * i + 5, 5 * j:5 * j + 5] = A[l, 0]
        H[5 * j:5 * j + 5, 5 * i:5 * i + 5] = A[l, 1]
    pv = np.random.default_rng(3).standard_normal((N, 5))
    pd = torch.from_numpy(pv).to(gpu)
    y = torch.empty_like(pd)
    h1.apply(pd, y)
    yt = torch.empty_like(pd)
    h1.apply_transpose(pd, yt)
    torch.cuda.synchronize()
    assert _rel(y.cpu().numpy().ravel(), H @ pv.ravel()) <= 1e-13
    assert _rel(yt.cpu().numpy().ravel(), H.T @ pv.ravel()) <= 1e-13
    op.close()


def test_high_degree_newton_tracks_oracle(gpu):
    import torch
    from aerof_dev_amd.operators import SpaceOperator, steady_solve
    m = _mesh()
    p = FlowParams(mach=0.5, alpha=0.1)
    U0 = init_state(p, m.X, "noise", 0.01, seed=14)
    # the sliver-rich cone takes a gentle CFL ramp (the oracle's solve leaves
    # the physical states after 29 steps of this one)
    ref = O.OracleMesh(m, p).steady_solve(U0, precision=0, max_steps=20, cfl0=1.0, cfl_max=100.0)
    op = SpaceOperator(m, p)
    out = steady_solve(op, torch.from_numpy(U0).to(gpu), precision=0, max_steps=20, cfl0=1.0, cfl_max=100.0)
    h, hr = out["history"], ref["history"]
    n = min(len(h), len(hr))
    assert n == 21
    assert np.max(np.abs(h[:n] - hr[:n]) / hr[:n]) <= 1e-6
    op.close()


def test_high_degree_direct_solve_falls_back(gpu):
    import torch
    from aerof_dev_amd.operators import SpaceOperator, direct_solve
    m = _mesh()
    p = FlowParams(mach=0.5, alpha=0.1, second_order=False)
    op = SpaceOperator(m, p)
    U = torch.from_numpy(init_state(p, m.X, "freestream")).to(gpu)
    rhs = torch.from_numpy(np.random.default_rng(4).standard_normal((2, op.num_nodes, 5))).to(gpu)
    x = torch.empty_like(rhs)
    s = direct_solve(op, U, rhs, x, precision=0, eps=1e-10, max_its=300)
    assert s["converged"], s
    for k in range(2):
        xk = torch.empty_like(U)
        assert direct_solve(op, U, rhs[k].contiguous(), xk, precision=0, eps=1e-10, max_its=300)["converged"]
        assert float(torch.linalg.norm(x[k] - xk) / torch.linalg.norm(xk)) <= 1e-7
    op.close()
