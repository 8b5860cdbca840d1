"""Host logic of the product and the C ABI surface (no GPU needed).

The library must load and export every entry point declared in
include/aerofx.h; the host-side mesh preparation (generator, edge table,
RCM, colouring, renumbering, synthetic states) must agree with the oracle.
"""
import ctypes as C
import os
import re

import numpy as np

import oracle as O
from aerof_dev_amd import FlowParams, init_state
from aerof_dev_amd import mesh as M
from aerof_dev_amd._abi import LIB_PATH, SIGNATURES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "aerofx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aerofx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), f"{name} not exported"
        assert name in SIGNATURES, f"{name} has no ctypes signature"


def test_generator_counts_and_orientation():
    m = M.kuhn_box(4, 3, 2, jitter=0.15, seed=5)
    assert m.num_nodes == 5 * 4 * 3 and m.num_tets == 6 * 24 and m.num_faces == 4 * (12 + 6 + 8)
    x = m.X[m.tets]
    vol = np.einsum("ij,ij->i", x[:, 3] - x[:, 0], np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0])) / 6
    assert vol.min() > 0
    # jitter moves interior nodes only, within 0.15 h
    lat = M.kuhn_box(4, 3, 2).X
    d = np.abs(m.X - lat)
    h = np.array([1 / 4, 1 / 3, 1 / 2])
    assert np.all(d <= 0.15 * h + 1e-15)
    on_bnd = np.zeros(m.num_nodes, bool)
    on_bnd[np.unique(m.faces)] = True
    assert np.all(d[on_bnd] == 0)


def test_bump_channel_geometry():
    m = M.channel_mesh(24, 8, 8)
    bottom = m.X[:, 2] < 0.2
    x = m.X[bottom, 0]
    z0 = np.isclose(m.X[:, 2] - 0.05 * np.exp(-((m.X[:, 0] - 1.5) / 0.2) ** 2), 0.0, atol=1e-15)
    assert z0.sum() == 25 * 9
    assert set(np.unique(m.codes)) == {2, 4, 5}
    assert x.size > 0


def test_edges_rcm_colouring_match_oracle():
    m = M.box_mesh(6, jitter=0.15, seed=8)
    t = O.OracleMesh(m, FlowParams()).tables()
    ptr = M.edges(m)
    np.testing.assert_array_equal(ptr, t["ptr"])
    np.testing.assert_array_equal(M.rcm(m), O.rcm(m.num_nodes, ptr))
    c1, n1 = M.color_edges(m.num_nodes, ptr)
    c2, n2 = O.color_edges(m.num_nodes, ptr)
    assert n1 == n2
    np.testing.assert_array_equal(c1, c2)


def test_renumbering_is_consistent():
    m = M.box_mesh(5, jitter=0.1, seed=3)
    r = M.rcm(m)
    m2 = M.renumber(m, r)
    np.testing.assert_array_equal(m2.X[r], m.X)
    np.testing.assert_array_equal(m2.tets, r[m.tets])
    # the residual is equivariant under renumbering
    p = FlowParams(mach=0.5)
    U = init_state(p, m.X, "noise", 0.01, seed=4)
    R1 = O.OracleMesh(m, p).residual(U)["R"]
    U2 = np.empty_like(U)
    U2[r] = U
    R2 = O.OracleMesh(m2, p).residual(U2)["R"]
    np.testing.assert_allclose(R2[r], R1, rtol=1e-11, atol=1e-11 * np.abs(R1).max())


def test_states():
    p = FlowParams(mach=0.8)
    m = M.kuhn_box(3, 3, 3)
    U = init_state(p, m.X, "freestream")
    u = np.sqrt(1.4 * p.p_inf / p.rho_inf) * 0.8
    np.testing.assert_allclose(U[0], [1.0, u, 0, 0, p.p_inf / 0.4 + 0.5 * u * u], rtol=1e-15)
    Un = init_state(p, m.X, "noise", 0.01, seed=1)
    assert 0.0 < np.std(Un[:, 0]) < 0.02
    Us = init_state(p, m.X, "sine", 0.02, seed=1)
    assert np.abs(Us[:, 0] - 1).max() <= 0.02 + 1e-15
    np.testing.assert_array_equal(init_state(p, m.X, "noise", 0.01, seed=1), Un)  # seeded


def test_bump_derivative_closed_form():
    """dX/da and dX/dc of the bump generator against central differences of
    the generator itself (the amplitude enters linearly; the centre through
    exp(-t^2)), also through an RCM renumbering."""
    base = [(0.05, 1.3, 0.2), (0.03, 1.8, 0.3)]
    kw = dict(lo=(0.0, 0.0, 0.0), hi=(3.0, 1.0, 1.0), jitter=0.1, seed=4, face_codes=M.CHANNEL)
    m = M.kuhn_box(12, 4, 4, bumps=base, **kw)
    h = 1e-6
    for b in range(2):
        for k, name in ((0, "amplitude"), (1, "center")):
            bp = [list(x) for x in base]
            bm = [list(x) for x in base]
            bp[b][k] += h
            bm[b][k] -= h
            fd = (M.kuhn_box(12, 4, 4, bumps=bp, **kw).X - M.kuhn_box(12, 4, 4, bumps=bm, **kw).X) / (2 * h)
            dX = M.bump_derivative(m, b, name)
            assert np.abs(dX - fd).max() <= 1e-8 * max(np.abs(fd).max(), 1e-300)
            assert np.abs(dX[:, :2]).max() == 0.0 and np.abs(dX[:, 2]).max() > 0.0
    r = M.renumber(m, M.rcm(m))
    assert np.array_equal(M.bump_derivative(r, 1, "center")[M.rcm(m)], M.bump_derivative(m, 1, "center"))
