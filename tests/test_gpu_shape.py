"""Adjoint shape gradient of the wall lift (the bump-channel sensitivity case,
reduced to a 24x8x8 channel and three design variables): Gaussian bump
amplitudes a_k with seeded centres on the bottom wall.

  J(a) = pressure force on the slip walls along z at the converged state
  dJ/da_k = (dJ/dX + (dR/dX)^T psi) . dX/da_k,  (dR/dU)^T psi = -dJ/dU

psi comes from the mixed-precision adjoint solve (FP32 transposed H1 and
block-Jacobi inside, exact FP64 adjoint residual outside).  Checked against
central finite differences of J with fully re-converged states (each
perturbed solve polished by FP64 steps at CFL 1e4 down to the round-off floor
of its residual), step 1e-6 on the amplitudes:
  first-order scheme   to 1e-6 relative (J(a) is smooth);
  second-order scheme  to 5e-5 relative.  With the Van Albada limiter J(a) is
    only piecewise smooth (an edge's limited slope switches branch when its
    two slopes change sign), so central differences carry an O(h) error: on
    this case they approach the adjoint value as 8e-4, 1.9e-5, 1.5e-6 for
    h = 1e-5, 1e-6, 1e-7, the last at the round-off floor of the steady solve.
The wall-force derivatives themselves are checked against central
differences of J on single coordinates.
"""
import dataclasses

import numpy as np
import pytest

from aerof_dev_amd import FlowParams, init_state
from aerof_dev_amd import mesh as M

pytestmark = pytest.mark.gpu
NX, NY, NZ = 24, 8, 8
BASE = (0.05, 1.5, 0.2)
H = 1e-6
ZDIR = (0.0, 0.0, 1.0)


def _mesh(amps, centres, width=0.2):
    bumps = [BASE] + [(a, c, width) for a, c in zip(amps, centres)]
    return M.kuhn_box(NX, NY, NZ, lo=(0.0, 0.0, 0.0), hi=(3.0, 1.0, 1.0), face_codes=M.CHANNEL, bumps=bumps)


def _converge(op, U, eps, polish=False):
    from aerof_dev_amd.operators import steady_solve
    r = steady_solve(op, U, precision=0, max_steps=3000, eps=eps, cfl0=10.0, cfl_max=1e4)
    assert r["converged"], r
    if polish:
        # Newton steps at the final CFL until the residual stops falling
        steady_solve(op, U, precision=0, max_steps=40, eps=1e-6, cfl0=1e4, cfl_max=1e4)
    return r


@pytest.fixture(scope="module", params=["first", "second"])
def base(gpu, request):
    import torch
    from aerof_dev_amd.operators import SpaceOperator
    centres = np.random.default_rng(2024).uniform(0.9, 2.1, size=3)
    p = FlowParams(mach=0.675, second_order=request.param == "second")
    m = _mesh([0.0] * 3, centres)
    op = SpaceOperator(m, p)
    U = torch.from_numpy(init_state(p, m.X, "freestream")).to(gpu)
    _converge(op, U, 1e-12)
    return dict(p=p, m=m, op=op, U=U, centres=centres, tol=5e-5 if request.param == "second" else 1e-6)


def test_wall_force_derivatives(gpu, base):
    import torch
    from aerof_dev_amd.operators import SpaceOperator
    p, m, op, U = base["p"], base["m"], base["op"], base["U"]
    dJdU = torch.empty_like(U)
    dJdX = torch.empty((m.num_nodes, 3), dtype=torch.float64, device=gpu)
    J = op.wall_force(U, ZDIR, dJdU, dJdX)
    gu, gx = dJdU.cpu().numpy(), dJdX.cpu().numpy()
    wall = np.where(np.isclose(m.X[:, 2], m.X[:, 2].min()))[0]
    n = int(wall[len(wall) // 2])
    for k in range(5):
        d = torch.zeros_like(U)
        d[n, k] = 1e-4 * max(1.0, abs(float(U[n, k])))
        fd = (op.wall_force(U + d, ZDIR) - op.wall_force(U - d, ZDIR)) / (2 * float(d[n, k]))
        assert abs(gu[n, k] - fd) <= 1e-8 * max(1.0, abs(fd))
    for a in range(3):
        dX = np.zeros_like(m.X)
        dX[n, a] = 1e-6
        Jp = SpaceOperator(dataclasses.replace(m, X=m.X + dX), p).wall_force(U, ZDIR)
        Jm = SpaceOperator(dataclasses.replace(m, X=m.X - dX), p).wall_force(U, ZDIR)
        fd = (Jp - Jm) / 2e-6
        assert abs(gx[n, a] - fd) <= 1e-7 * max(1.0, abs(fd)), (a, gx[n, a], fd)
    assert J != 0.0


def test_adjoint_shape_gradient_vs_fd(gpu, base):
    import torch
    from aerof_dev_amd.operators import SpaceOperator, adjoint_solve
    p, m, op, U, centres = base["p"], base["m"], base["op"], base["U"], base["centres"]
    dJdU = torch.empty_like(U)
    dJdX = torch.empty((m.num_nodes, 3), dtype=torch.float64, device=gpu)
    op.wall_force(U, ZDIR, dJdU, dJdX)
    psi = torch.empty_like(U)
    st = adjoint_solve(op, U, (-dJdU).contiguous(), psi, precision=1, cfl=1e4, eps=1e-11, max_its=3000)
    assert st["converged"], st
    gx = torch.empty_like(dJdX)
    op.compute_transpose_derivative_of_residual(U, psi, None, gx)
    total = (dJdX + gx).cpu().numpy()
    for k in range(3):
        # dX/da_k: the wall displacement is linear in the amplitude
        amps = [0.0] * 3
        amps[k] = 1.0
        dXda = _mesh(amps, centres).X - m.X
        grad = float(np.sum(total * dXda))
        Js = []
        for s in (1.0, -1.0):
            amps = [0.0] * 3
            amps[k] = s * H
            mk = _mesh(amps, centres)
            opk = SpaceOperator(mk, p)
            Uk = U.clone()
            _converge(opk, Uk, 1e-9, polish=True)
            Js.append(opk.wall_force(Uk, ZDIR))
        fd = (Js[0] - Js[1]) / (2 * H)
        assert abs(grad - fd) <= base["tol"] * abs(fd), (k, grad, fd)
