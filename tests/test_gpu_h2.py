"""MatVecProdH2: the exact derivative of the second-order residual as the
GMRES operator (matrix-free, forward-mode kernels).

* its products equal the forward-mode residual derivative plus the time term
  (the same kernels on the same state) to rounding;
* against central finite differences of the oracle's residual (the
  restatement of the reference's), 1e-6 relative, on a state with noise in
  every component (where the Van Albada limiter is differentiable at the
  difference step);
* Newton with exact products (MatVecProdH2) reaches the steady state of the
  Newton with the assembled first-order Jacobian, both FP64 and with the FP32
  block-Jacobi.
"""
import numpy as np
import pytest

import oracle as O
from aerof_dev_amd import FlowParams, init_state
from aerof_dev_amd import mesh as M

pytestmark = pytest.mark.gpu


def _case(second_order=True, noise=True):
    m = M.channel_mesh(24, 8, 8, jitter=0.1, seed=5)
    p = FlowParams(mach=0.675, alpha=0.02, second_order=second_order)
    U = init_state(p, m.X, "noise", 0.01 if noise else 0.0, seed=17)
    if noise:
        rng = np.random.default_rng(18)
        U[:, 1:4] += 0.02 * p.mach * U[:, :1] * rng.standard_normal((U.shape[0], 3))
    return m, p, U


@pytest.mark.parametrize("second_order", [True, False])
def test_h2_products(gpu, second_order):
    import torch
    from aerof_dev_amd.operators import MatVecProdH1, MatVecProdH2, SpaceOperator
    m, p, U = _case(second_order)
    op = SpaceOperator(m, p)
    om = O.OracleMesh(m, p)
    Ud = torch.from_numpy(U).to(gpu)
    dt = torch.empty(op.num_nodes, dtype=torch.float64, device=gpu)
    h1 = MatVecProdH1(op)
    h1.timestep(Ud, 50.0, dt)
    h2 = MatVecProdH2(op)
    assert h2.evaluate(Ud, dt) == 0
    rng = np.random.default_rng(4)
    dtn = dt.cpu().numpy()
    for _ in range(3):
        pv = rng.standard_normal(U.shape) * (np.abs(U) + 0.1)
        pd = torch.from_numpy(pv).to(gpu)
        y = torch.empty_like(Ud)
        assert h2.apply(pd, y) == 0
        d = torch.empty_like(Ud)
        assert op.compute_derivative_of_residual(Ud, d, pd, None) == 0
        ref = d.cpu().numpy() + pv / dtn[:, None]
        yh = y.cpu().numpy()
        assert np.linalg.norm(yh - ref) <= 1e-14 * np.linalg.norm(ref)
        fd = om.jvp_fd(U, pv, 2, dtn)
        err = np.linalg.norm(yh - fd) / np.linalg.norm(fd)
        assert err <= 1e-6, err
    h2.release()
    op.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_newton_with_exact_products(gpu, precision):
    import torch
    from aerof_dev_amd.operators import SpaceOperator, steady_solve
    m = M.channel_mesh(24, 8, 8, bump_height=0.05, bump_width=0.2)
    p = FlowParams(mach=0.675)
    U0 = init_state(p, m.X, "freestream")
    op = SpaceOperator(m, p)
    Uh1 = torch.from_numpy(U0).to(gpu)
    r1 = steady_solve(op, Uh1, precision=0, max_steps=400, eps=1e-10)
    Uh2 = torch.from_numpy(U0).to(gpu)
    r2 = steady_solve(op, Uh2, precision=precision, max_steps=400, eps=1e-10, mvp="h2")
    print(f"[h2] H1 Newton {r1['steps']} steps / {r1['krylov_its']} Krylov its; "
          f"exact products {r2['steps']} steps / {r2['krylov_its']} its", flush=True)
    assert r1["converged"] and r2["converged"]
    a, b = Uh1.cpu().numpy(), Uh2.cpu().numpy()
    assert np.linalg.norm(a - b) <= 1e-8 * np.linalg.norm(a)
    # exact products: fewer Newton steps than the first-order Jacobian's
    # defect correction
    assert r2["steps"] < r1["steps"]
    op.close()
