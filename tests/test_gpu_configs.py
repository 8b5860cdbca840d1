"""The sensitivity workloads at their full sizes.

* Shape gradient on the 192x64x64 bump channel (815,425 nodes): wall lift,
  8 Gaussian bump amplitudes with seeded centres.  Steady mixed-precision
  solve, mixed-precision adjoint solve, (dR/dX)^T psi; the same 8
  derivatives through the direct route (forward-mode (dR/dX) dX/da_k, one
  direct solve per amplitude) agree with the adjoint ones to 5e-8 (duality:
  both routes solve the exact derivative systems of the same residual to
  1e-11; measured 9e-9); two of them against central differences of re-converged lifts
  (step 1e-6) to 1e-6.  First-order scheme: at this size the second-order
  limiter keeps switching on a few edges and the steady solve stalls above
  the drop the finite differences need.  The perturbed solves start from the
  base state, whose residual on the perturbed mesh (about 4e-2) is already
  small: they are converged 1e-8 further (about 4e-10) in mixed precision,
  then polished by 60 FP64 steps at CFL 1e4 down to the FP64 floor of this
  mesh (about 6e-11): stopping at 4e-10 leaves a lift error of 2e-5 of the
  difference quotient.
* Batched direct sensitivities on the 128^3 Kuhn box with jittered wall
  bumps (2,146,689 nodes): 16 right-hand sides -(dR/dX) dX/ds_k, one per
  design variable (the amplitude and the centre of each of 8 seeded bumps),
  solved together (flexible GMRES on the exact FP64 products, each step
  preconditioned by one block GMRES with FP32 inner cycles in captured
  graphs); every solution's true residual meets the tolerance (1e-8), and
  four of them agree with single-RHS all-FP64 solves to 1e-6 (two solvers
  stopped at 1e-8).
"""
import time

import numpy as np
import pytest

from aerof_dev_amd import FlowParams, init_state
from aerof_dev_amd import mesh as M

pytestmark = pytest.mark.gpu
ZDIR = (0.0, 0.0, 1.0)


def _log(msg):
    print(msg, flush=True)


def test_shape_gradient_channel192_adjoint_direct_fd(gpu):
    import torch
    from aerof_dev_amd.operators import SpaceOperator, adjoint_solve, direct_solve, steady_solve
    centres = np.random.default_rng(2024).uniform(0.9, 2.1, size=8)

    def mesh(amps):
        bumps = [(0.05, 1.5, 0.2)] + [(a, c, 0.2) for a, c in zip(amps, centres)]
        return M.kuhn_box(192, 64, 64, lo=(0.0, 0.0, 0.0), hi=(3.0, 1.0, 1.0), face_codes=M.CHANNEL, bumps=bumps)

    p = FlowParams(mach=0.675, second_order=False)
    m = mesh([0.0] * 8)
    op = SpaceOperator(m, p)
    N = op.num_nodes
    assert N == 815425
    U = torch.from_numpy(init_state(p, m.X, "freestream")).to(gpu)
    t0 = time.time()
    r = steady_solve(op, U, precision=1, max_steps=3000, eps=1e-11, cfl0=10.0, cfl_max=1e4)
    assert r["converged"], r
    _log(f"[config4] steady {time.time() - t0:.1f} s, {r['steps']} steps")
    dJdU = torch.empty_like(U)
    dJdX = torch.empty((N, 3), dtype=torch.float64, device=gpu)
    op.wall_force(U, ZDIR, dJdU, dJdX)
    t0 = time.time()
    psi = torch.empty_like(U)
    sa = adjoint_solve(op, U, (-dJdU).contiguous(), psi, precision=1, cfl=1e4, eps=1e-11, max_its=3000)
    assert sa["converged"], sa
    gx = torch.empty_like(dJdX)
    op.compute_transpose_derivative_of_residual(U, psi, None, gx)
    _log(f"[config4] adjoint {time.time() - t0:.1f} s, {sa['its']} outer its")
    # dX/da_k in closed form (the generator's bump k+1)
    bumps = m.meta["bumps"]
    dXs = [torch.from_numpy(M.bump_derivative(m, k + 1, "amplitude")).to(gpu) for k in range(8)]
    g_adj = np.array([float(torch.sum((dJdX + gx) * dX)) for dX in dXs])
    assert len(bumps) == 9 and np.all(np.abs(g_adj) > 1e-6), g_adj
    # direct route: (dR/dU) x_k = -(dR/dX) dX/da_k, one right-hand side at a
    # time (FP64 outer iterations on the exact products, FP32 inner GMRES)
    g_dir = []
    for k, dX in enumerate(dXs):
        t0 = time.time()
        rhs = torch.empty_like(U)
        assert op.compute_derivative_of_residual(U, rhs, None, dX) == 0
        rhs = (-rhs).contiguous()
        x = torch.empty_like(U)
        sd = direct_solve(op, U, rhs, x, precision=1, cfl=1e4, eps=1e-11, max_its=600)
        assert sd["converged"], (k, sd)
        g_dir.append(float(torch.sum(dJdX * dX) + torch.sum(dJdU * x)))
        _log(f"[config4] direct {k}: {time.time() - t0:.1f} s, {sd['its']} outer its")
    g_dir = np.array(g_dir)
    dual = np.abs(g_adj - g_dir) / np.abs(g_adj)
    _log(f"[config4] dJ/da {g_adj}, adjoint vs direct {dual.max():.2e}")
    assert dual.max() <= 5e-8, (g_adj, g_dir)
    op.close()
    # central differences of re-converged lifts for two amplitudes
    h = 1e-6
    for k in (0, 5):
        Js = []
        for sgn in (1.0, -1.0):
            amps = [0.0] * 8
            amps[k] = sgn * h
            opk = SpaceOperator(mesh(amps), p)
            Uk = U.clone()
            rk = steady_solve(opk, Uk, precision=1, max_steps=600, eps=1e-8, cfl0=10.0, cfl_max=1e4)
            assert rk["converged"], rk
            rp = steady_solve(opk, Uk, precision=0, max_steps=60, eps=1e-6, cfl0=1e4, cfl_max=1e4)
            _log(f"[config4] amplitude {k} {sgn:+.0f}h: {rk['steps']} mixed steps to {rk['res']:.2e}, "
                 f"FP64 polish to {rp['history'].min():.2e}")
            assert rp["history"].min() <= 1e-10
            Js.append(opk.wall_force(Uk, ZDIR))
            opk.close()
        fd = (Js[0] - Js[1]) / (2 * h)
        _log(f"[config4] amplitude {k}: adjoint {g_adj[k]:.10e} FD {fd:.10e}")
        assert abs(g_adj[k] - fd) <= 1e-6 * abs(fd), (k, g_adj[k], fd)


def test_batched_direct_sensitivities_box128(gpu):
    import torch
    from aerof_dev_amd.operators import SpaceOperator, direct_solve, steady_solve
    rng = np.random.default_rng(5)
    centres = rng.uniform(0.25, 0.75, size=8)
    amps = rng.uniform(0.004, 0.008, size=8)
    bumps = [(a, c, 0.1) for a, c in zip(amps, centres)]
    m = M.kuhn_box(128, 128, 128, jitter=0.15, seed=1234, bumps=bumps)
    m = M.renumber(m, M.rcm_gpu(m))
    p = FlowParams(mach=0.8)
    op = SpaceOperator(m, p)
    N = op.num_nodes
    assert N == 2146689
    U = torch.from_numpy(init_state(p, m.X, "freestream")).to(gpu)
    t0 = time.time()
    r = steady_solve(op, U, precision=1, max_steps=400, eps=1e-8, cfl0=10.0, cfl_max=1e4)
    assert r["converged"], r
    _log(f"[config5] steady {time.time() - t0:.1f} s, {r['steps']} steps")
    design = [(b, prm) for b in range(8) for prm in ("amplitude", "center")]
    rhs = torch.empty((16, N, 5), dtype=torch.float64, device=gpu)
    for k, (b, prm) in enumerate(design):
        dX = torch.from_numpy(M.bump_derivative(m, b, prm)).to(gpu)
        dR = torch.empty_like(U)
        assert op.compute_derivative_of_residual(U, dR, None, dX) == 0
        rhs[k] = -dR
    eps = 1e-8
    x = torch.empty_like(rhs)
    t0 = time.time()
    sb = direct_solve(op, U, rhs, x, precision=1, cfl=1e4, eps=eps, max_its=200)
    assert sb["converged"], sb
    _log(f"[config5] block mixed {time.time() - t0:.1f} s, {sb['its']} outer steps, {sb['krylov_its']} inner")
    # every solution against the exact products
    for k in range(16):
        Ax = torch.empty_like(U)
        assert op.compute_derivative_of_residual(U, Ax, x[k].contiguous(), None) == 0
        res = float(torch.linalg.norm(rhs[k] - Ax) / torch.linalg.norm(rhs[k]))
        assert res <= 2 * eps, (k, res)
    # four of them against single-RHS all-FP64 solves
    worst = 0.0
    t0 = time.time()
    for k in (0, 5, 10, 15):
        xk = torch.empty_like(U)
        s1 = direct_solve(op, U, rhs[k].contiguous(), xk, precision=0, cfl=1e4, eps=eps, max_its=200)
        assert s1["converged"], (k, s1)
        worst = max(worst, float(torch.linalg.norm(x[k] - xk) / torch.linalg.norm(xk)))
        _log(f"[config5] single FP64 {k}: {s1['its']} outer steps")
    _log(f"[config5] 4 single FP64 {time.time() - t0:.1f} s; max solution difference {worst:.2e}")
    assert worst <= 1e-6, worst
    op.close()
