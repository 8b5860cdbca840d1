"""Argument and mesh validation of aerofx_ctx_create, on the host.

Every rejected input returns a negative code and a message before the
library touches the device, so these run without a GPU.  The reference reads
its mesh from files that its own tools validated; the library takes raw
arrays, so it checks index ranges and boundary codes itself.
"""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from aerof_dev_amd import FlowParams
from aerof_dev_amd import mesh as M
from aerof_dev_amd._abi import AerofxError, Mesh as CMesh, c_double_p, c_int_p, lib

ERR_ARG, ERR_MESH = -1, -3


def _create(m, p):
    X = np.ascontiguousarray(m.X, dtype=np.float64)
    tets = np.ascontiguousarray(m.tets, dtype=np.int32)
    faces = np.ascontiguousarray(m.faces, dtype=np.int32)
    codes = np.ascontiguousarray(m.codes, dtype=np.int32)
    cm = CMesh(m.num_nodes, m.num_tets, int(faces.shape[0]), X.ctypes.data_as(c_double_p),
               tets.ctypes.data_as(c_int_p), faces.ctypes.data_as(c_int_p), codes.ctypes.data_as(c_int_p))
    h = C.c_void_p()
    pc = p.c()
    rc = lib().aerofx_ctx_create(C.byref(cm), C.byref(pc), C.byref(h))
    return rc, lib().aerofx_last_error().decode(), h


@pytest.fixture(scope="module")
def box():
    return M.box_mesh(2, jitter=0.1, seed=3)


def test_null_arguments():
    rc = lib().aerofx_ctx_create(None, None, None)
    assert rc == ERR_ARG and "null argument" in lib().aerofx_last_error().decode()


@pytest.mark.parametrize("change, code, text", [
    (lambda m: dataclasses.replace(m, tets=np.where(m.tets == 5, m.num_nodes, m.tets)), ERR_MESH,
     "tet node out of range"),
    (lambda m: dataclasses.replace(m, tets=np.where(m.tets == 5, -1, m.tets)), ERR_MESH, "tet node out of range"),
    (lambda m: dataclasses.replace(m, faces=np.where(m.faces == 0, m.num_nodes + 7, m.faces)), ERR_MESH,
     "face node out of range"),
    (lambda m: dataclasses.replace(m, codes=np.where(np.arange(m.codes.size) == 3, 6, m.codes).astype(np.int32)),
     ERR_MESH, "unsupported face code"),
    (lambda m: dataclasses.replace(m, tets=m.tets[:0]), ERR_ARG, "bad mesh"),
])
def test_bad_mesh_rejected(box, change, code, text):
    rc, msg, h = _create(change(box), FlowParams())
    assert rc == code, (rc, msg)
    assert text in msg
    assert not h.value


@pytest.mark.parametrize("params", [FlowParams(gamma=1.0), FlowParams(rho_inf=0.0), FlowParams(p_inf=-1.0)])
def test_bad_parameters_rejected(box, params):
    rc, msg, h = _create(box, params)
    assert rc == ERR_ARG and "bad parameters" in msg
    assert not h.value


def test_python_front_end_raises(box):
    from aerof_dev_amd.operators import SpaceOperator
    with pytest.raises(AerofxError, match="unsupported face code"):
        SpaceOperator(dataclasses.replace(box, codes=np.full_like(box.codes, 9)), FlowParams())
