"""Forward-mode derivative of the residual, dR = (dR/dU) dU + (dR/dX) dX.

* duality with the hand-written adjoint kernels: <lam, dR(dU, dX)> equals
  <(dR/dU)^T lam, dU> + <(dR/dX)^T lam, dX> to 1e-11 relative (both are the
  exact derivative of the same residual, so only rounding separates them);
* central finite differences of the residual (geometry rebuilt from the
  perturbed coordinates), 1e-6 relative per direction.  The residual is only
  piecewise smooth: a Van Albada evaluation whose two slopes sit within the
  step of a sign change makes the central difference average two one-sided
  derivatives.  Random state directions move every one of a small mesh's
  ~14k Van Albada evaluations, so they take the step 1e-8 (rounding ~1e-8
  relative, truncation below it); coordinate directions take Richardson
  extrapolation at 1e-6.
"""
import dataclasses

import numpy as np
import pytest

from aerof_dev_amd import FlowParams, init_state
from aerof_dev_amd import mesh as M

pytestmark = pytest.mark.gpu
H = 1e-6


def _state(p, m, seed):
    U = init_state(p, m.X, "noise", 0.01, seed=seed)
    rng = np.random.default_rng(seed + 1000)
    U[:, 1:4] += 0.02 * p.mach * U[:, :1] * rng.standard_normal((U.shape[0], 3))
    return U


def _case(kind, second_order, seed=7):
    if kind == "box":
        m = M.box_mesh(5, jitter=0.15, seed=1234 + seed)
        p = FlowParams(mach=0.5, alpha=0.05, yaw=0.03, second_order=second_order)
    else:
        m = M.channel_mesh(9, 3, 3, bump_height=0.05, bump_width=0.2, jitter=0.1, seed=seed)
        p = FlowParams(mach=0.675, alpha=0.02, second_order=second_order)
    return m, p, _state(p, m, seed)


def _interior_dX(m, rng):
    """A coordinate perturbation that keeps boundary nodes on their faces."""
    dX = rng.standard_normal(m.X.shape) * 0.1
    onb = np.zeros(m.num_nodes, bool)
    onb[m.faces.ravel()] = True
    dX[onb] = 0.0
    return dX


def _tangent(op, U, dU, dX, gpu):
    import torch
    t = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
    Ud = t(U)
    dR = torch.empty_like(Ud)
    assert op.compute_derivative_of_residual(Ud, dR, t(dU), t(dX)) == 0
    return dR.cpu().numpy()


def _res(mesh, params, U, gpu):
    import torch
    from aerof_dev_amd.operators import SpaceOperator
    op = SpaceOperator(mesh, params)
    Ud = torch.from_numpy(np.ascontiguousarray(U)).to(gpu)
    R = torch.empty_like(Ud)
    assert op.compute_residual(Ud, R) == 0
    op.close()
    return R.cpu().numpy()


@pytest.mark.parametrize("kind", ["box", "channel"])
@pytest.mark.parametrize("second_order", [True, False])
def test_tangent_adjoint_duality(gpu, kind, second_order):
    import torch
    from aerof_dev_amd.operators import SpaceOperator
    m, p, U = _case(kind, second_order)
    rng = np.random.default_rng(3)
    lam = rng.standard_normal(U.shape)
    dU = rng.standard_normal(U.shape) * np.abs(U).mean(axis=0)
    dX = rng.standard_normal(m.X.shape) * 0.05
    op = SpaceOperator(m, p)
    dR_u = _tangent(op, U, dU, None, gpu)
    dR_x = _tangent(op, U, None, dX, gpu)
    dR_ux = _tangent(op, U, dU, dX, gpu)
    Ud = torch.from_numpy(U).to(gpu)
    gu = torch.empty_like(Ud)
    gx = torch.empty((m.num_nodes, 3), dtype=torch.float64, device=gpu)
    op.compute_transpose_derivative_of_residual(Ud, torch.from_numpy(lam).to(gpu), gu, gx)
    gu, gx = gu.cpu().numpy(), gx.cpu().numpy()
    for fwd, rev in ((np.sum(lam * dR_u), np.sum(gu * dU)), (np.sum(lam * dR_x), np.sum(gx * dX))):
        assert abs(fwd - rev) <= 1e-11 * max(abs(fwd), abs(rev)), (fwd, rev)
    # linearity in (dU, dX)
    assert np.abs(dR_ux - (dR_u + dR_x)).max() <= 1e-12 * np.abs(dR_ux).max()
    op.close()


@pytest.mark.parametrize("kind", ["box", "channel"])
def test_tangent_dX_matches_central_differences(gpu, kind):
    from aerof_dev_amd.operators import SpaceOperator
    m, p, U = _case(kind, True, seed=11)
    rng = np.random.default_rng(5)
    op = SpaceOperator(m, p)
    for k in range(2):
        dX = _interior_dX(m, rng) if k == 0 else rng.standard_normal(m.X.shape) * 0.05
        fwd = _tangent(op, U, None, dX, gpu)

        def fd(h):
            Rp = _res(dataclasses.replace(m, X=m.X + h * dX), p, U, gpu)
            Rm = _res(dataclasses.replace(m, X=m.X - h * dX), p, U, gpu)
            return (Rp - Rm) / (2 * h)

        ref = (4.0 * fd(H / 2) - fd(H)) / 3.0
        err = np.linalg.norm(fwd - ref) / np.linalg.norm(ref)
        assert err <= 1e-6, f"direction {k}: rel {err:.2e}"
    op.close()


@pytest.mark.parametrize("second_order", [True, False])
def test_tangent_dU_matches_central_differences(gpu, second_order):
    """Projections <lam, dR> on random lam, as the adjoint kernels' checks."""
    from aerof_dev_amd.operators import SpaceOperator
    m, p, U = _case("box", second_order, seed=13)
    rng = np.random.default_rng(9)
    op = SpaceOperator(m, p)
    for _ in range(3):
        dU = rng.standard_normal(U.shape) * (np.abs(U) + 0.1)
        lam = rng.standard_normal(U.shape)
        fwd = float(np.sum(lam * _tangent(op, U, dU, None, gpu)))

        h = 1e-8
        ref = float(np.sum(lam * (_res(m, p, U + h * dU, gpu) - _res(m, p, U - h * dU, gpu)))) / (2 * h)
        assert abs(fwd - ref) <= 1e-6 * abs(ref), (fwd, ref)
    op.close()
