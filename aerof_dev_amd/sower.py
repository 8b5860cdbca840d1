"""SOWER mesh files and AERO-F solution (restart) files, through the C ABI.

read_mesh: GeoSource::getSubDomain (GeoSource.C:246-356) for one subdomain
of a cluster (.geo + .dec, BinFileHandler format, either byte order);
write_mesh: a cluster of one subdomain holding the whole mesh (what a
one-GPU run reads back with sub=0).  read_solution / write_solution:
SubDomain::readVectorFromFile / writeVectorToFile (SubDomain.C:5609-5810).
Host memory only; device tensors are copied.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._abi import SowerHeader, c_double_p, c_int_p, check, lib
from .mesh import Mesh


def _b(path) -> bytes:
    return os.fsencode(path)


def mesh_info(geo, dec, sub: int = 0) -> dict:
    h = SowerHeader()
    check(lib().aerofx_sower_info(_b(geo), _b(dec), int(sub), C.byref(h)), "sower_info")
    return {k: getattr(h, k) for k, _ in SowerHeader._fields_}


def read_mesh(geo, dec, sub: int = 0) -> Mesh:
    """Subdomain `sub` as a Mesh; meta holds its global and cluster node
    numbers (the cluster numbers index the cluster's solution files)."""
    info = mesh_info(geo, dec, sub)
    N, T, F = info["num_nodes"], info["num_tets"], info["num_faces"]
    X = np.empty((N, 3))
    tets = np.empty((T, 4), dtype=np.int32)
    faces = np.empty((max(F, 1), 3), dtype=np.int32)
    codes = np.empty(max(F, 1), dtype=np.int32)
    glob = np.empty(N, dtype=np.int32)
    clus = np.empty(N, dtype=np.int32)
    check(lib().aerofx_sower_read(_b(geo), _b(dec), int(sub), X.ctypes.data_as(c_double_p),
                                  tets.ctypes.data_as(c_int_p), faces.ctypes.data_as(c_int_p),
                                  codes.ctypes.data_as(c_int_p), glob.ctypes.data_as(c_int_p),
                                  clus.ctypes.data_as(c_int_p)), "sower_read")
    return Mesh(X, tets, faces[:F].copy(), codes[:F].copy(),
                dict(source=(os.fspath(geo), os.fspath(dec), int(sub)), node_glob=glob, node_clus=clus,
                     num_clus_nodes=info["num_clus_nodes"]))


def write_mesh(mesh: Mesh, geo, dec) -> None:
    X = np.ascontiguousarray(mesh.X, dtype=np.float64)
    tets = np.ascontiguousarray(mesh.tets, dtype=np.int32)
    faces = np.ascontiguousarray(mesh.faces, dtype=np.int32)
    codes = np.ascontiguousarray(mesh.codes, dtype=np.int32)
    check(lib().aerofx_sower_write(_b(geo), _b(dec), mesh.num_nodes, mesh.num_tets, mesh.num_faces,
                                   X.ctypes.data_as(c_double_p), tets.ctypes.data_as(c_int_p),
                                   faces.ctypes.data_as(c_int_p), codes.ctypes.data_as(c_int_p)), "sower_write")


def solution_info(path) -> dict:
    n, dim, nsol = C.c_int(), C.c_int(), C.c_int()
    check(lib().aerofx_solution_info(_b(path), C.byref(n), C.byref(dim), C.byref(nsol)), "solution_info")
    return dict(num_clus_nodes=n.value, dim=dim.value, nsol=nsol.value)


def write_solution(path, step: int, tag: float, U) -> None:
    """Solution `step` (0 creates the file) of U [n, dim] with its tag."""
    if hasattr(U, "detach"):
        U = U.detach().cpu().numpy()
    U = np.ascontiguousarray(U, dtype=np.float64)
    n, dim = U.shape
    check(lib().aerofx_solution_write(_b(path), int(step), float(tag), n, dim, U.ctypes.data_as(c_double_p)),
          "solution_write")


def read_solution(path, step: int, dim: int | None = None, node_clus=None):
    """(U [n, dim], tag) of solution `step`: all cluster rows, or the rows
    node_clus (e.g. a subdomain's meta["node_clus"])."""
    info = solution_info(path)
    dim = info["dim"] if dim is None else int(dim)
    if node_clus is None:
        n, ptr = info["num_clus_nodes"], None
    else:
        node_clus = np.ascontiguousarray(node_clus, dtype=np.int32)
        n, ptr = len(node_clus), node_clus.ctypes.data_as(c_int_p)
    U = np.zeros((n, dim))
    tag = C.c_double()
    check(lib().aerofx_solution_read(_b(path), int(step), C.byref(tag), n, dim, ptr, U.ctypes.data_as(c_double_p)),
          "solution_read")
    return U, tag.value
