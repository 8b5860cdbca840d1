"""Batched right-hand sides: block GMRES against single-RHS FP64 GMRES.

The 16 systems share H1 + diag(1/dt) with block-Jacobi (FP64 set-up).  Each
block solution must match the corresponding single-RHS FP64 solve to 1e-7
relative (both solve to a relative residual of 1e-10), in FP64 and in the
mixed mode (FP32 inner cycles, FP64 refinement); a captured-graph run must
reproduce the eager run bit for bit on every replay.
"""
import numpy as np
import pytest

import oracle as O
from aerof_dev_amd import FlowParams, init_state
from aerof_dev_amd import mesh as M

pytestmark = pytest.mark.gpu


def _system(gpu, nrhs=16):
    import torch
    from aerof_dev_amd.operators import MatVecProdH1, SpaceOperator
    m = M.channel_mesh(12, 4, 4, bump_height=0.05, bump_width=0.2)
    p = FlowParams(mach=0.675)
    U = init_state(p, m.X, "noise", 0.01, seed=31)
    om = O.OracleMesh(m, p)
    dt = om.timestep(om.residual(U)["V"], 200.0)
    op = SpaceOperator(m, p)
    h1 = MatVecProdH1(op)
    h1.evaluate(torch.from_numpy(U).to(gpu), torch.from_numpy(dt).to(gpu), 0)
    rng = np.random.default_rng(99)
    B = torch.from_numpy(rng.standard_normal((nrhs, m.num_nodes, 5))).to(gpu)
    return op, h1, B


def _single(op, B):
    import torch
    from aerof_dev_amd.operators import GmresSolver
    X = torch.empty_like(B)
    g = GmresSolver(op, 30, 1e-10, 3000)
    for k in range(B.shape[0]):
        xk = torch.empty_like(B[k])
        g.solve(B[k].contiguous(), xk)
        X[k] = xk
    torch.cuda.synchronize()
    return X


def _rel(a, b):
    return float((a - b).norm() / b.norm())


@pytest.mark.parametrize("precision", [0, 1])
def test_block_matches_single_rhs(gpu, precision):
    import torch
    from aerof_dev_amd.operators import BlockGmresSolver
    op, h1, B = _system(gpu)
    Xs = _single(op, B)
    X = torch.empty_like(B)
    out = BlockGmresSolver(op, m=16, eps=1e-10, max_cycles=500, precision=precision, use_graph=False).solve(B, X)
    assert out["converged"], out
    for k in range(B.shape[0]):
        assert _rel(X[k], Xs[k]) <= 1e-7, (k, _rel(X[k], Xs[k]), out)
    # true residual of every column with the FP64 operator
    for k in range(B.shape[0]):
        y = torch.empty_like(B[k])
        h1.apply(X[k].contiguous(), y)
        assert float((B[k] - y).norm() / B[k].norm()) <= 2e-10


def test_block_graph_replays_match_eager(gpu):
    import torch
    from aerof_dev_amd.operators import BlockGmresSolver
    op, h1, B = _system(gpu, nrhs=11)  # padded to 16 columns internally
    Xe = torch.empty_like(B)
    eager = BlockGmresSolver(op, m=12, eps=1e-9, max_cycles=400, precision=1, use_graph=False).solve(B, Xe)
    assert eager["converged"]
    for rep in range(3):
        Xg = torch.empty_like(B)
        g = BlockGmresSolver(op, m=12, eps=1e-9, max_cycles=400, precision=1, use_graph=True).solve(B, Xg)
        torch.cuda.synchronize()
        assert g["cycles"] == eager["cycles"], (rep, g, eager)
        assert torch.equal(Xg, Xe), rep
        # eager work between replays
        y = torch.empty_like(B[0])
        h1.apply(Xg[0].contiguous(), y)
        torch.cuda.synchronize()


def _cone(k=60):
    """A cone of k tets around the axis from the base centre to the apex:
    both have k + 1 neighbours, more than the 32 of the narrow block-apply
    shape, so the solver takes its 64-entry shape.  Farfield all round."""
    from aerof_dev_amd._abi import AEROFX_BC_INLET
    ang = 2.0 * np.pi * np.arange(k) / k
    rim = np.stack([np.cos(ang), np.sin(ang), np.zeros(k)], axis=1)
    X = np.vstack([[0.0, 0.0, 1.0], [0.0, 0.0, 0.0], rim])  # apex 0, centre 1, rim 2..
    tets, faces = [], []
    for i in range(k):
        a, b = 2 + i, 2 + (i + 1) % k
        t = [0, 1, a, b]
        p = X[t]
        if np.dot(np.cross(p[1] - p[0], p[2] - p[0]), p[3] - p[0]) < 0:
            t = [0, 1, b, a]
        tets.append(t)
        for f in ([0, a, b], [1, b, a]):
            q = X[f]
            cen = (q[0] + q[1] + q[2]) / 3.0
            if np.dot(np.cross(q[1] - q[0], q[2] - q[0]), cen - np.array([0.0, 0.0, 0.4])) < 0:
                f = [f[0], f[2], f[1]]
            faces.append(f)
    return M.Mesh(X, np.array(tets, np.int32), np.array(faces, np.int32),
                  np.full(len(faces), AEROFX_BC_INLET, np.int32))


@pytest.mark.parametrize("precision", [0, 1])
def test_block_wide_rows(gpu, precision):
    """Nodes with more than 32 neighbours: the wide block-apply shape gives
    the single-RHS solutions.  310 unknowns: block cycles of m = 8 (128
    vectors) stay inside the space; CFL 5 keeps the small system well
    conditioned."""
    import torch
    from aerof_dev_amd.operators import BlockGmresSolver, MatVecProdH1, SpaceOperator
    m = _cone()
    p = FlowParams(mach=0.5, alpha=0.1)
    U = init_state(p, m.X, "noise", 0.01, seed=5)
    om = O.OracleMesh(m, p)
    dt = om.timestep(om.residual(U)["V"], 5.0)
    op = SpaceOperator(m, p)
    h1 = MatVecProdH1(op)
    h1.evaluate(torch.from_numpy(U).to(gpu), torch.from_numpy(dt).to(gpu), 0)
    B = torch.from_numpy(np.random.default_rng(8).standard_normal((16, m.num_nodes, 5))).to(gpu)
    Xs = _single(op, B)
    X = torch.empty_like(B)
    out = BlockGmresSolver(op, m=8, eps=1e-10, max_cycles=500, precision=precision, use_graph=False).solve(B, X)
    assert out["converged"], out
    for k in range(B.shape[0]):
        assert _rel(X[k], Xs[k]) <= 1e-7, (k, _rel(X[k], Xs[k]), out)
