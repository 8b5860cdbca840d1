"""Spatial low-Mach preconditioning (Problem.Prec = Preconditioned, no time
preconditioner) on the GPU against the oracle.

* residual, first and second order, at Mach 0.05 with noise: 1e-12 of the
  residual's magnitude (the per-edge flux is bit-exact against the
  reference's ROEFLUX5 in tests/test_gpu_physics.py);
* H1 blocks with the preconditioned ROEJAC6: 1e-12 of each block's size;
* the pseudo-transient Newton history tracks the oracle's for 20 steps
  (1e-6), and the converged state differs from Roe's;
* the derivative passes refuse the low-Mach flux (they differentiate
  ROEFLUX7).
"""
import numpy as np
import pytest

import oracle as O
from aerof_dev_amd import FlowParams, LowMachPrec, init_state
from aerof_dev_amd import mesh as M

pytestmark = pytest.mark.gpu
LM = LowMachPrec(min_mach=0.02, slope=1.5, cutoff=0.9, shock_reducer=0.3, irey=1e-4)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("second_order", [False, True])
def test_low_mach_residual_matches_oracle(gpu, second_order):
    import torch
    from aerof_dev_amd.operators import SpaceOperator
    m = M.box_mesh(10, jitter=0.15, seed=3)
    p = FlowParams(mach=0.05, second_order=second_order, low_mach=LM)
    U = init_state(p, m.X, "noise", 0.02, seed=5)
    op = SpaceOperator(m, p)
    ref = O.OracleMesh(m, p).residual(U)
    Ud = torch.from_numpy(U).to(gpu)
    R = torch.empty_like(Ud)
    for family in ("per_lane", "staged"):  # the staged family falls back to per-lane
        op.set_residual_kernels(family)
        assert op.compute_residual(Ud, R) == ref["ierr"] == 0
        assert _rel(R.cpu().numpy(), ref["R"]) <= 1e-12, family
    # the same context with the preconditioner off: Roe's residual
    op.set_low_mach(None)
    op.compute_residual(Ud, R)
    roe = O.OracleMesh(m, FlowParams(mach=0.05, second_order=second_order)).residual(U)["R"]
    assert _rel(R.cpu().numpy(), roe) <= 1e-12
    assert _rel(roe, ref["R"]) > 1e-6
    op.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_low_mach_jacobian_matches_oracle(gpu, precision):
    import torch
    from aerof_dev_amd.operators import MatVecProdH1, SpaceOperator
    m = M.channel_mesh(24, 8, 8, bump_height=0.05, bump_width=0.2)
    p = FlowParams(mach=0.05, low_mach=LM)
    U = init_state(p, m.X, "noise", 0.02, seed=6)
    op = SpaceOperator(m, p)
    om = O.OracleMesh(m, p)
    V = om.residual(U)["V"]
    dt = om.timestep(V, 50.0)
    Dr, Ar = om.jacobian(V, dt)
    h1 = MatVecProdH1(op)
    h1.evaluate(torch.from_numpy(U).to(gpu), torch.from_numpy(dt).to(gpu), precision)
    D, A = h1.blocks()
    tol = 1e-12 if precision == 0 else 1e-6
    assert np.max(np.abs(D - Dr) / np.abs(Dr).max(axis=(1, 2))[:, None, None]) <= tol
    assert np.max(np.abs(A - Ar) / (np.abs(Ar).max(axis=(2, 3))[:, :, None, None] + 1e-300)) <= tol
    op.close()


def test_low_mach_newton_tracks_oracle(gpu):
    import torch
    from aerof_dev_amd.operators import SpaceOperator, steady_solve
    m = M.channel_mesh(24, 8, 8, bump_height=0.05, bump_width=0.2)
    p = FlowParams(mach=0.05, low_mach=LM)
    U0 = init_state(p, m.X, "freestream")
    op = SpaceOperator(m, p)
    ref = O.OracleMesh(m, p).steady_solve(U0, precision=0, max_steps=20)
    Ud = torch.from_numpy(U0).to(gpu)
    out = steady_solve(op, Ud, precision=0, max_steps=20)
    h, hr = out["history"], ref["history"]
    n = min(len(h), len(hr))
    print(f"[lowmach] {n} steps, history {h[0]:.3e} -> {h[n - 1]:.3e}, oracle {hr[n - 1]:.3e}", flush=True)
    assert n >= 20
    assert np.max(np.abs(h[:n] - hr[:n]) / hr[:n]) <= 1e-6
    # to convergence, FP64 and mixed: a steady state of the low-Mach residual
    steps = []
    for precision in (0, 1):
        Uc = torch.from_numpy(U0).to(gpu)
        r = steady_solve(op, Uc, precision=precision, max_steps=400, eps=1e-10)
        print(f"[lowmach] precision {precision}: {r['steps']} steps, {r['krylov_its']} Krylov its", flush=True)
        assert r["converged"], r
        steps.append(r["steps"])
    # Roe's flux at Mach 0.05: slower (the oracle needs 331 steps to 1e-8
    # where the preconditioned one needs 135 to 1e-10) to another state
    Uroe = torch.from_numpy(U0).to(gpu)
    op.set_low_mach(None)
    rr = steady_solve(op, Uroe, precision=0, max_steps=400, eps=1e-8)
    print(f"[lowmach] Roe flux: {rr['steps']} steps to 1e-8", flush=True)
    assert rr["converged"] and rr["steps"] > steps[0]
    a, b = Uc.cpu().numpy(), Uroe.cpu().numpy()
    assert np.linalg.norm(a - b) > 1e-8 * np.linalg.norm(b)
    op.close()


def test_low_mach_refused_by_derivative_passes(gpu):
    import torch
    from aerof_dev_amd import AerofxError
    from aerof_dev_amd.operators import MatVecProdH2, SpaceOperator, adjoint_solve, direct_solve
    m = M.box_mesh(4)
    p = FlowParams(mach=0.05, low_mach=LM)
    U = torch.from_numpy(init_state(p, m.X, "freestream")).to(gpu)
    op = SpaceOperator(m, p)
    out = torch.empty_like(U)
    with pytest.raises(AerofxError, match="low-Mach"):
        op.compute_derivative_of_residual(U, out, U, None)
    with pytest.raises(AerofxError, match="low-Mach"):
        op.compute_transpose_derivative_of_residual(U, U, out, None)
    dt = torch.ones(op.num_nodes, dtype=torch.float64, device=gpu)
    with pytest.raises(AerofxError, match="low-Mach"):
        MatVecProdH2(op).evaluate(U, dt)
    with pytest.raises(AerofxError, match="low-Mach"):
        adjoint_solve(op, U, U, out)
    with pytest.raises(AerofxError, match="low-Mach"):
        direct_solve(op, U, U, out)
    op.close()
