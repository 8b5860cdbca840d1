"""Shape sensitivities of the wall lift through the direct and the adjoint
route, on a bump channel whose design variables are bump amplitudes and a
bump centre (closed-form dX/ds from the generator).

  direct:  (dR/dU) dU/ds = -(dR/dX) dX/ds,  dJ/ds = dJ/dX . dX/ds + dJ/dU . dU/ds
  adjoint: (dR/dU)^T psi = -dJ/dU,           dJ/ds = dJ/dX . dX/ds + psi . (dR/dX) dX/ds

Both routes solve with the exact second-order derivatives (forward-mode
products on one side, the hand-written adjoint on the other) to 1e-12, so
they agree to 1e-8 relative (duality); at first order the gradient also
matches central differences of re-converged lifts (step 1e-6) to 1e-6.
"""
import numpy as np
import pytest

from aerof_dev_amd import FlowParams, init_state
from aerof_dev_amd import mesh as M

pytestmark = pytest.mark.gpu
ZDIR = (0.0, 0.0, 1.0)
BUMPS = [(0.05, 1.5, 0.2), (0.0, 1.1, 0.25), (0.0, 1.9, 0.25)]
DESIGN = [(1, "amplitude"), (2, "amplitude"), (0, "center")]


def _mesh(bumps):
    return M.kuhn_box(12, 4, 4, lo=(0.0, 0.0, 0.0), hi=(3.0, 1.0, 1.0), face_codes=M.CHANNEL, bumps=bumps)


def _converge(op, U, eps):
    from aerof_dev_amd.operators import steady_solve
    r = steady_solve(op, U, precision=0, max_steps=200, eps=eps, mvp="fd")
    assert r["converged"], r
    return r


@pytest.mark.parametrize("second_order", [True, False])
def test_direct_and_adjoint_gradients_agree(gpu, second_order):
    import torch
    from aerof_dev_amd.operators import SpaceOperator, adjoint_solve, direct_solve
    m = _mesh(BUMPS)
    p = FlowParams(mach=0.675, second_order=second_order)
    op = SpaceOperator(m, p)
    U = torch.from_numpy(init_state(p, m.X, "freestream")).to(gpu)
    _converge(op, U, 1e-11)
    N = m.num_nodes
    dJdU = torch.empty_like(U)
    dJdX = torch.empty((N, 3), dtype=torch.float64, device=gpu)
    op.wall_force(U, ZDIR, dJdU, dJdX)
    # adjoint
    psi = torch.empty_like(U)
    sa = adjoint_solve(op, U, (-dJdU).contiguous(), psi, precision=0, eps=1e-12, max_its=400)
    assert sa["converged"], sa
    gx = torch.empty_like(dJdX)
    op.compute_transpose_derivative_of_residual(U, psi, None, gx)
    # direct, all design variables in one batched solve
    dXs = [torch.from_numpy(M.bump_derivative(m, b, prm)).to(gpu) for b, prm in DESIGN]
    rhs = torch.empty((len(DESIGN), N, 5), dtype=torch.float64, device=gpu)
    for k, dX in enumerate(dXs):
        dR = torch.empty_like(U)
        assert op.compute_derivative_of_residual(U, dR, None, dX) == 0
        rhs[k] = -dR
    x = torch.empty_like(rhs)
    sd = direct_solve(op, U, rhs, x, precision=1, eps=1e-12, max_its=400)
    assert sd["converged"], sd
    g_adj, g_dir = [], []
    for k, dX in enumerate(dXs):
        g_adj.append(float(torch.sum(dJdX * dX) + torch.sum(gx * dX)))
        g_dir.append(float(torch.sum(dJdX * dX) + torch.sum(dJdU * x[k])))
    g_adj, g_dir = np.array(g_adj), np.array(g_dir)
    assert np.all(np.abs(g_adj) > 1e-6)
    assert np.max(np.abs(g_adj - g_dir) / np.abs(g_adj)) <= 1e-8, (g_adj, g_dir)
    # one right-hand side alone (GMRES inside) gives the batched solution
    x0 = torch.empty_like(U)
    s1 = direct_solve(op, U, rhs[0].contiguous(), x0, precision=0, eps=1e-12, max_its=400)
    assert s1["converged"]
    assert float(torch.linalg.norm(x0 - x[0]) / torch.linalg.norm(x0)) <= 1e-9
    if not second_order:
        # central differences of re-converged lifts for the first design variable
        b, prm = DESIGN[0]
        h = 1e-6
        Js = []
        for sgn in (1.0, -1.0):
            bumps = [list(t) for t in BUMPS]
            bumps[b][0 if prm == "amplitude" else 1] += sgn * h
            opk = SpaceOperator(_mesh(bumps), p)
            Uk = U.clone()
            _converge(opk, Uk, 1e-9)
            Js.append(opk.wall_force(Uk, ZDIR))
            opk.close()
        fd = (Js[0] - Js[1]) / (2 * h)
        assert abs(g_adj[0] - fd) <= 1e-6 * abs(fd), (g_adj[0], fd)
    op.close()
