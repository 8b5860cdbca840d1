"""RAS / ILU(0) preconditioner (IluPrec, fill 0, type RAS) on the GPU.

* the preconditioner apply equals the oracle's restatement of
  SparseMat::numericILU + lusol on the same subdomains (node ranges of the
  RCM-numbered mesh): 1e-12 relative with FP64 factors, FP32 rounding with
  FP32 factors;
* GMRES with it solves H1 + 1/dt in fewer iterations than with block-Jacobi;
* Newton with it reaches the steady state of the block-Jacobi Newton.
"""
import numpy as np
import pytest

import oracle as O
from aerof_dev_amd import FlowParams, init_state
from aerof_dev_amd import mesh as M

pytestmark = pytest.mark.gpu


def _case(kind):
    if kind == "box":
        m = M.box_mesh(8, jitter=0.15, seed=3, rcm_order=True)
        p = FlowParams(mach=0.5)
    else:
        m = M.channel_mesh(24, 8, 8, bump_height=0.05, bump_width=0.2)
        m = M.renumber(m, M.rcm(m))
        p = FlowParams(mach=0.675)
    U = init_state(p, m.X, "noise", 0.01, seed=17)
    return m, p, U


@pytest.mark.parametrize("kind", ["box", "channel"])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("B", [64, 1024])
def test_ras_ilu0_apply_matches_oracle(gpu, kind, precision, B):
    import torch
    from aerof_dev_amd.operators import MatVecProdH1, SpaceOperator
    m, p, U = _case(kind)
    N = m.num_nodes
    print(f"[ilu] {kind} N={N} B={B} precision={precision}", flush=True)
    op = SpaceOperator(m, p)
    om = O.OracleMesh(m, p)
    V = om.residual(U)["V"]
    dt = om.timestep(V, 20.0)
    D, A = om.jacobian(V, dt)
    Ud = torch.from_numpy(U).to(gpu)
    dtd = torch.from_numpy(dt).to(gpu)
    h1 = MatVecProdH1(op)
    h1.select_preconditioner("ras_ilu0", B)
    h1.evaluate(Ud, dtd, precision)
    starts = list(range(0, N, B)) + [N]
    rng = np.random.default_rng(11)
    for _ in range(2):
        y = rng.standard_normal((N, 5))
        x = torch.empty_like(Ud)
        h1.prec_apply(torch.from_numpy(y).to(gpu), x)
        torch.cuda.synchronize()
        xg = x.cpu().numpy()
        ref = O.ras_ilu0(N, M.edges(m), D, A, starts, y, single=bool(precision))
        err = np.linalg.norm(xg - ref) / np.linalg.norm(ref)
        assert err <= (1e-12 if precision == 0 else 1e-5), err
    op.close()


def test_gmres_with_ras_ilu0(gpu):
    import torch
    from aerof_dev_amd.operators import GmresSolver, MatVecProdH1, SpaceOperator
    m, p, U = _case("channel")
    op = SpaceOperator(m, p)
    Ud = torch.from_numpy(U).to(gpu)
    dt = torch.empty(op.num_nodes, dtype=torch.float64, device=gpu)
    h1 = MatVecProdH1(op)
    h1.timestep(Ud, 1000.0, dt)
    b = torch.from_numpy(np.random.default_rng(2).standard_normal(U.shape)).to(gpu)
    its = {}
    for kind in ("jacobi", "ras_ilu0"):
        h1.select_preconditioner(kind, 512)
        h1.evaluate(Ud, dt, 0)
        x = torch.empty_like(Ud)
        n, est = GmresSolver(op, 30, 1e-8, 600).solve(b, x)
        y = torch.empty_like(Ud)
        h1.apply(x, y)
        res = float(torch.linalg.norm(b - y) / torch.linalg.norm(b))
        print(f"[ilu] {kind}: {n} its, estimate {est / float(torch.linalg.norm(b)):.3e}, true {res:.3e}, "
              f"x finite {bool(torch.isfinite(x).all())}, y finite {bool(torch.isfinite(y).all())}", flush=True)
        assert res <= 2e-8, (kind, res)
        its[kind] = n
    print(f"[ilu] GMRES(30) to 1e-8 at CFL 1000: {its}", flush=True)
    assert its["ras_ilu0"] < its["jacobi"]
    op.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_newton_with_ras_ilu0(gpu, precision):
    import torch
    from aerof_dev_amd.operators import SpaceOperator, steady_solve
    m = M.channel_mesh(24, 8, 8, bump_height=0.05, bump_width=0.2)
    m = M.renumber(m, M.rcm(m))
    p = FlowParams(mach=0.675)
    U0 = init_state(p, m.X, "freestream")
    op = SpaceOperator(m, p)
    Uj = torch.from_numpy(U0).to(gpu)
    rj = steady_solve(op, Uj, precision=precision, max_steps=400, eps=1e-10)
    Ui = torch.from_numpy(U0).to(gpu)
    ri = steady_solve(op, Ui, precision=precision, max_steps=400, eps=1e-10, prec="ras_ilu0", prec_subdomain=512)
    print(f"[ilu] Newton precision {precision}: block-Jacobi {rj['steps']} steps / {rj['krylov_its']} its, "
          f"RAS-ILU(0) {ri['steps']} steps / {ri['krylov_its']} its", flush=True)
    assert rj["converged"] and ri["converged"]
    a, b = Uj.cpu().numpy(), Ui.cpu().numpy()
    assert np.linalg.norm(a - b) <= 1e-8 * np.linalg.norm(a)
    assert ri["krylov_its"] < rj["krylov_its"]
    op.close()
