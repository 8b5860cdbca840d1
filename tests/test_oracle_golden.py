"""The oracle against the reference's own kernels.

tests/golden/ref_kernels.npz holds input/output vectors produced by the
reference's Fortran (roeflux7.f, boundflux5.f, roejac6.f) and headers
(RecFcnDesc.h Van Albada, DenseMatrixOps.h LINPACK inverse) compiled from
the reference tree (tests/golden/make_golden.py).  The oracle must
reproduce them bit for bit (it is built without FMA contraction, like the
reference).  When oracle/_ref is built the same is checked on fresh random
inputs.
"""
import os

import numpy as np
import pytest

import oracle as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_kernels.npz"))


def test_roeflux7_golden():
    out = np.array([O.roeflux7(float(G["gamma"]), 0.0, G["normal"][k], 0.0, G["VL"][k], G["VR"][k])
                    for k in range(len(G["VL"]))])
    np.testing.assert_array_equal(out, G["roeflux7"])


def test_boundflux5_golden():
    out = np.array([O.boundflux5(float(G["gamma"]), G["normal"][k], 0.0, G["VL"][k], G["Ub"][k])
                    for k in range(len(G["VL"]))])
    np.testing.assert_array_equal(out, G["boundflux5"])


def test_vanalbada_golden():
    for k in range(len(G["VL"])):
        a, b = O.vanalbada5(1.0 / 3.0, G["VL"][k], G["ddVij"][k], G["VR"][k], G["ddVji"][k])
        np.testing.assert_array_equal(a, G["va_ij"][k])
        np.testing.assert_array_equal(b, G["va_ji"][k])


def test_roejac6_golden():
    for k in range(len(G["VL"])):
        j1 = O.roejac6(float(G["gamma"]), 0.0, G["normal"][k], 0.0, G["VL"][k], G["VR"][k], 1)
        j2 = O.roejac6(float(G["gamma"]), 0.0, G["normal"][k], 0.0, G["VR"][k], G["VL"][k], 2)
        np.testing.assert_array_equal(j1, G["roejac6_sw1"][k])
        np.testing.assert_array_equal(j2, G["roejac6_sw2"][k])


def test_linpack_inverse_golden():
    for k in range(len(G["blocks"])):
        a, info = O.invert5(G["blocks"][k])
        assert info == 0
        np.testing.assert_array_equal(a, G["inv_f64"][k])
        af, info = O.invert5(G["blocks"][k], single=True)
        np.testing.assert_array_equal(af, G["inv_f32"][k])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")
def test_oracle_matches_live_reference():
    rng = np.random.default_rng(99)
    for _ in range(300):
        n = rng.normal(size=3)
        Vl = np.array([rng.uniform(.2, 3), *rng.normal(size=3), rng.uniform(.2, 3)])
        Vr = np.array([rng.uniform(.2, 3), *rng.normal(size=3), rng.uniform(.2, 3)])
        np.testing.assert_array_equal(O.roeflux7(1.4, 0.0, n, 0.0, Vl, Vr),
                                      O.roeflux7(1.4, 0.0, n, 0.0, Vl, Vr, use_ref=True))
        np.testing.assert_array_equal(O.boundflux5(1.4, n, 0.0, Vl, Vr),
                                      O.boundflux5(1.4, n, 0.0, Vl, Vr, use_ref=True))
        for sw in (1, 2):
            np.testing.assert_array_equal(O.roejac6(1.4, 0.0, n, 0.0, Vl, Vr, sw),
                                          O.roejac6(1.4, 0.0, n, 0.0, Vl, Vr, sw, use_ref=True))


L = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_lowmach.npz"))


def test_roeflux5_golden():
    """ROEFLUX5 (roeflux5.f, Roe-Turkel low-Mach dissipation) on seeded
    low-Mach states, with and without the preconditioning."""
    out = np.array([O.roeflux5(1.0, float(L["gamma"]), 0.0, L["normal"][k], 0.0, L["VL"][k], L["VR"][k],
                               *L["params"][k], int(L["prec"][k])) for k in range(len(L["VL"]))])
    np.testing.assert_array_equal(out, L["roeflux5"])


def test_roeflux5_pins():
    """What the mathematics fixes: with the preconditioning off (beta = 1)
    the Roe-Turkel flux is Roe's (roeflux7) to rounding; for equal states
    the dissipation vanishes and the flux is the physical F(V).n, with or
    without it."""
    rng = np.random.default_rng(5)
    for _ in range(100):
        n = rng.normal(size=3)
        Vl = np.array([rng.uniform(.2, 3), *(0.3 * rng.normal(size=3)), rng.uniform(.2, 3)])
        Vr = np.array([rng.uniform(.2, 3), *(0.3 * rng.normal(size=3)), rng.uniform(.2, 3)])
        a = O.roeflux5(1.0, 1.4, 0.0, n, 0.0, Vl, Vr, 0.1, 1.0, 1.0, 0.0, 0.0, 0.1, 0)
        b = O.roeflux7(1.4, 0.0, n, 0.0, Vl, Vr)
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
        rho, u, v, w, p = Vl
        un = u * n[0] + v * n[1] + w * n[2]
        E = p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)
        F = np.array([rho * un, rho * u * un + p * n[0], rho * v * un + p * n[1], rho * w * un + p * n[2],
                      (E + p) * un])
        c = O.roeflux5(1.0, 1.4, 0.0, n, 0.0, Vl, Vl, 0.1, 1.0, 1.0, 0.0, 0.0, 0.1, 1)
        assert np.max(np.abs(c - F)) <= 1e-13 * np.max(np.abs(F))
        # preconditioned dissipation differs from Roe's at low Mach
        d = O.roeflux5(1.0, 1.4, 0.0, n, 0.0, Vl, Vr, 0.01, 1.0, 1.0, 0.0, 0.0, 0.1, 1)
        assert np.max(np.abs(d - b)) > 1e-6 * np.max(np.abs(b))


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")
def test_roeflux5_matches_live_reference():
    rng = np.random.default_rng(98)
    for k in range(300):
        n = rng.normal(size=3)
        Vl = np.array([rng.uniform(.2, 3), *(0.2 * rng.normal(size=3)), rng.uniform(.2, 3)])
        Vr = np.array([rng.uniform(.2, 3), *(0.2 * rng.normal(size=3)), rng.uniform(.2, 3)])
        par = (rng.uniform(0.005, 0.2), rng.uniform(0.5, 2), rng.uniform(0.5, 1), rng.uniform(0, 0.5),
               rng.uniform(0, 1e-3), rng.uniform(0.01, 0.2))
        np.testing.assert_array_equal(O.roeflux5(1.0, 1.4, 0.0, n, 0.0, Vl, Vr, *par, k % 2),
                                      O.roeflux5(1.0, 1.4, 0.0, n, 0.0, Vl, Vr, *par, k % 2, use_ref=True))
        for sw in (1, 2):
            np.testing.assert_array_equal(O.roejac6(1.4, 0.0, n, 0.0, Vl, Vr, sw, low_mach=par),
                                          O.roejac6(1.4, 0.0, n, 0.0, Vl, Vr, sw, low_mach=par, use_ref=True))


def test_roejac6_low_mach_golden():
    """ROEJAC6 with the Roe-Turkel preconditioning on (prec 1), both sides."""
    for k in range(len(L["VL"])):
        j1 = O.roejac6(float(L["gamma"]), 0.0, L["normal"][k], 0.0, L["VL"][k], L["VR"][k], 1, low_mach=L["params"][k])
        j2 = O.roejac6(float(L["gamma"]), 0.0, L["normal"][k], 0.0, L["VR"][k], L["VL"][k], 2, low_mach=L["params"][k])
        np.testing.assert_array_equal(j1, L["roejac6_sw1"][k])
        np.testing.assert_array_equal(j2, L["roejac6_sw2"][k])


def test_roejac6_low_mach_linearises_roeflux5():
    """The preconditioned Jacobian is the linearisation of the Roe-Turkel
    flux about equal states: F5(V, V + d) - F5(V, V) = J_sw2(V, V) dU and
    F5(V + d, V) - F5(V, V) = J_sw1(V, V) dU to first order (no shock term)."""
    rng = np.random.default_rng(8)
    gam = 1.4

    def cons(V):
        r, u, v, w, p = V
        return np.array([r, r * u, r * v, r * w, p / (gam - 1) + 0.5 * r * (u * u + v * v + w * w)])

    for _ in range(50):
        n = rng.normal(size=3)
        V = np.array([rng.uniform(.5, 2), *(0.1 * rng.normal(size=3)), rng.uniform(.5, 2)])
        lm = (rng.uniform(0.005, 0.2), rng.uniform(0.5, 2), rng.uniform(0.5, 1), 0.0, rng.uniform(0, 1e-3),
              rng.uniform(0.01, 0.2))
        d = 1e-7 * rng.normal(size=5) * np.abs(V)
        dU = cons(V + d) - cons(V)
        f0 = O.roeflux5(1.0, gam, 0.0, n, 0.0, V, V, *lm, 1)
        for sw, (a, b) in ((2, (V, V + d)), (1, (V + d, V))):
            J = O.roejac6(gam, 0.0, n, 0.0, V, V, sw, low_mach=lm)
            df = O.roeflux5(1.0, gam, 0.0, n, 0.0, a, b, *lm, 1) - f0
            assert np.linalg.norm(df - J @ dU) <= 2e-6 * np.linalg.norm(J @ dU)
