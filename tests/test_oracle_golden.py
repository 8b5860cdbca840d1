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
