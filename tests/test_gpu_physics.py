"""The device physics routines themselves, case by case, against outputs of
the reference's own routines (tests/golden/ref_kernels.npz, written from the
reference's roeflux7.f, boundflux5.f, roejac6.f, RecFcn.h Van Albada and
DenseMatrixOps.h LINPACK inverse compiled in the build container).

The kernels follow those routines' operation order and are built without
FMA contraction, with IEEE division and square root, so every case must
match bit for bit: 400 Roe fluxes, 400 Steger-Warming farfield fluxes,
400 x 5 Van Albada reconstructions, 2 x 400 Roe Jacobians and 400 LINPACK
inverses in FP64 and in FP32."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_kernels.npz")


def _eval(gpu, kind, scalar, ins, width):
    import torch
    from aerof_dev_amd._abi import check, lib
    d = [None if a is None else torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(gpu) for a in ins]
    n = d[0].shape[0]
    out = torch.full((n, width), float("nan"), dtype=torch.float64, device=gpu)
    ptrs = [None if t is None else C.c_void_p(t.data_ptr()) for t in d] + [None] * (4 - len(d))
    bad = check(lib().aerofx_physics_eval(kind, n, scalar, *ptrs, C.c_void_p(out.data_ptr())), "physics_eval")
    return bad, out.cpu().numpy()


def _same(got, ref, what):
    got, ref = np.asarray(got), np.asarray(ref)
    if not np.array_equal(got, ref):
        bad = np.flatnonzero(got.ravel() != ref.ravel())
        ulps = np.abs(got.ravel()[bad] - ref.ravel()[bad]) / np.spacing(np.abs(ref.ravel()[bad]))
        pytest.fail(f"{what}: {bad.size} of {ref.size} values differ, up to {ulps.max():.1f} ulp")


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(GOLDEN))


def test_roeflux7_bit_exact(gpu, golden):
    from aerof_dev_amd._abi import AEROFX_PHYS_ROEFLUX7
    g = golden
    _, out = _eval(gpu, AEROFX_PHYS_ROEFLUX7, float(g["gamma"]), [g["normal"], g["VL"], g["VR"]], 5)
    _same(out, g["roeflux7"], "roeflux7")


def test_boundflux5_bit_exact(gpu, golden):
    from aerof_dev_amd._abi import AEROFX_PHYS_BOUNDFLUX5
    g = golden
    _, out = _eval(gpu, AEROFX_PHYS_BOUNDFLUX5, float(g["gamma"]), [g["normal"], g["VL"], g["Ub"]], 5)
    _same(out, g["boundflux5"], "boundflux5")


def test_vanalbada_bit_exact(gpu, golden):
    from aerof_dev_amd._abi import AEROFX_PHYS_VANALBADA
    g = golden
    _, out = _eval(gpu, AEROFX_PHYS_VANALBADA, 1.0 / 3.0, [g["VL"], g["ddVij"], g["VR"], g["ddVji"]], 10)
    _same(out[:, :5], g["va_ij"], "Van Albada V_ij")
    _same(out[:, 5:], g["va_ji"], "Van Albada V_ji")


def test_roejac6_bit_exact(gpu, golden):
    from aerof_dev_amd._abi import AEROFX_PHYS_ROEJAC6_SW1, AEROFX_PHYS_ROEJAC6_SW2
    g = golden
    _, j1 = _eval(gpu, AEROFX_PHYS_ROEJAC6_SW1, float(g["gamma"]), [g["normal"], g["VL"], g["VR"]], 25)
    _same(j1.reshape(-1, 5, 5), g["roejac6_sw1"], "roejac6 type 1")
    _, j2 = _eval(gpu, AEROFX_PHYS_ROEJAC6_SW2, float(g["gamma"]), [g["normal"], g["VR"], g["VL"]], 25)
    _same(j2.reshape(-1, 5, 5), g["roejac6_sw2"], "roejac6 type 2")


def test_linpack_inverse_bit_exact(gpu, golden):
    from aerof_dev_amd._abi import AEROFX_PHYS_INVERT5_F32, AEROFX_PHYS_INVERT5_F64
    g = golden
    blocks = g["blocks"].reshape(-1, 25)
    bad, inv = _eval(gpu, AEROFX_PHYS_INVERT5_F64, 0.0, [blocks], 25)
    assert bad == 0
    _same(inv.reshape(-1, 5, 5), g["inv_f64"], "LINPACK inverse FP64")
    bad, inv32 = _eval(gpu, AEROFX_PHYS_INVERT5_F32, 0.0, [blocks], 25)
    assert bad == 0
    _same(inv32.astype(np.float32).reshape(-1, 5, 5), g["inv_f32"], "LINPACK inverse FP32")


def test_roeflux5_bit_exact(gpu):
    """ROEFLUX5 (roeflux5.f, Roe-Turkel low-Mach dissipation) on 400 seeded
    low-Mach cases, three in four with the preconditioning on
    (tests/golden/ref_lowmach.npz)."""
    from aerof_dev_amd._abi import AEROFX_PHYS_ROEFLUX5
    g = dict(np.load(os.path.join(os.path.dirname(GOLDEN), "ref_lowmach.npz")))
    par = np.column_stack([g["params"], g["prec"].astype(np.float64)])
    _, out = _eval(gpu, AEROFX_PHYS_ROEFLUX5, float(g["gamma"]), [g["normal"], g["VL"], g["VR"], par], 5)
    _same(out, g["roeflux5"], "roeflux5")


def test_roejac6_low_mach_bit_exact(gpu):
    """ROEJAC6 with the Roe-Turkel preconditioning on, both sides, on the
    400 low-Mach cases of tests/golden/ref_lowmach.npz."""
    from aerof_dev_amd._abi import AEROFX_PHYS_ROEJAC6_LM_SW1, AEROFX_PHYS_ROEJAC6_LM_SW2
    g = dict(np.load(os.path.join(os.path.dirname(GOLDEN), "ref_lowmach.npz")))
    par = np.column_stack([g["params"], np.ones(len(g["VL"]))])
    _, j1 = _eval(gpu, AEROFX_PHYS_ROEJAC6_LM_SW1, float(g["gamma"]), [g["normal"], g["VL"], g["VR"], par], 25)
    _, j2 = _eval(gpu, AEROFX_PHYS_ROEJAC6_LM_SW2, float(g["gamma"]), [g["normal"], g["VR"], g["VL"], par], 25)
    _same(j1, g["roejac6_sw1"].reshape(-1, 25), "roejac6 low-Mach sw1")
    _same(j2, g["roejac6_sw2"].reshape(-1, 25), "roejac6 low-Mach sw2")
