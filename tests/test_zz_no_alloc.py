"""aerofx_residual_noalloc (the residual with its scratch in a static device
array, allocating nothing) against aerofx_residual_ex on the same context:
the same kernels over the same inputs, so the results must be bit-identical,
and the error counts equal.  Two calls in a row on two meshes, the second
smaller (the static array keeps the first call's contents), at second and
first order, then a mesh too large for the array, which must be rejected
before any launch."""
import ctypes as C

import numpy as np
import pytest

from aerof_dev_amd import FlowParams, init_state
from aerof_dev_amd import mesh as M

pytestmark = pytest.mark.gpu


def _pair(gpu, mesh, params, U, flags=0):
    import torch
    from aerof_dev_amd._abi import check, lib
    from aerof_dev_amd.operators import SpaceOperator
    op = SpaceOperator(mesh, params)
    Ud = torch.from_numpy(U).to(gpu)
    Ra = torch.full_like(Ud, float("nan"))
    Rb = torch.full_like(Ud, float("nan"))
    ea = check(lib().aerofx_residual_ex(op._h, C.c_void_p(Ud.data_ptr()), C.c_void_p(Ra.data_ptr()), flags),
               "residual_ex")
    eb = check(lib().aerofx_residual_noalloc(op._h, C.c_void_p(Ud.data_ptr()), C.c_void_p(Rb.data_ptr()), flags),
               "residual_noalloc")
    torch.cuda.synchronize()
    op.close()
    return (ea, Ra.cpu().numpy()), (eb, Rb.cpu().numpy())


@pytest.mark.parametrize("second_order", [True, False])
def test_noalloc_matches_residual_two_calls(gpu, second_order):
    p = FlowParams(mach=0.675, alpha=0.03, second_order=second_order)
    cases = [M.channel_mesh(48, 16, 16, jitter=0.1, seed=3),  # 14,161 nodes
             M.box_mesh(12, jitter=0.15, seed=8)]              # 2,197 nodes: smaller
    for k, m in enumerate(cases):
        U = init_state(p, m.X, "noise", 0.01, seed=20 + k)
        print(f"[noalloc] order {2 if second_order else 1} call {k}: N={m.num_nodes}", flush=True)
        (ea, Ra), (eb, Rb) = _pair(gpu, m, p, U)
        assert ea == eb == 0
        assert np.isfinite(Rb).all()
        assert np.array_equal(Ra, Rb), f"call {k}: max |diff| {np.max(np.abs(Ra - Rb)):.3e}"


def test_noalloc_counts_nonphysical_states(gpu):
    m = M.box_mesh(6, jitter=0.0, seed=0)
    p = FlowParams(mach=0.5)
    U = init_state(p, m.X, "freestream")
    U[U.shape[0] // 2, 4] = -2.375  # energy below the kinetic energy: negative pressure
    (ea, _), (eb, _) = _pair(gpu, m, p, U)
    assert ea > 0 and eb == ea


def test_noalloc_rejects_meshes_beyond_the_scratch(gpu):
    import torch
    from aerof_dev_amd._abi import AEROFX_ERR_ARG, lib
    from aerof_dev_amd.operators import SpaceOperator
    m = M.box_mesh(64, jitter=0.15, seed=1)  # 274,625 nodes: about 134 MB of scratch
    p = FlowParams(mach=0.5)
    op = SpaceOperator(m, p)
    U = torch.from_numpy(init_state(p, m.X, "freestream")).to(gpu)
    R = torch.zeros_like(U)
    rc = lib().aerofx_residual_noalloc(op._h, C.c_void_p(U.data_ptr()), C.c_void_p(R.data_ptr()), 0)
    assert rc == AEROFX_ERR_ARG
    assert b"static scratch" in lib().aerofx_last_error()
    torch.cuda.synchronize()
    assert float(R.abs().max()) == 0.0  # nothing was launched
    op.close()
