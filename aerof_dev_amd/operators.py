"""Python mirror of the reference's SpaceOperator<5> surface over the C ABI.

Device arrays are torch tensors (float64, contiguous, on the GPU); torch is
only the allocator and stream here.  Every call goes to the HIP kernels in
libaerofx.so -- there is no other code path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import Mesh as _CMesh
from ._abi import c_double_p, c_int_p, check, lib
from .mesh import FlowParams, Mesh


def _ptr(t) -> int:
    if not t.is_cuda:
        raise ValueError("device tensor expected")
    if not t.is_contiguous():
        raise ValueError("contiguous tensor expected")
    import torch
    if t.dtype != torch.float64:
        raise ValueError("float64 tensor expected")
    return t.data_ptr()


class SpaceOperator:
    """SpaceOperator<5> (SpaceOperator.h:171-173) bound to one mesh.

    The context owns the mesh tables (edges, dual normals, control volumes,
    LSQ factors), as the reference's SpaceOperator owns its geometry state.
    """

    def __init__(self, mesh: Mesh, params: FlowParams, stream=None):
        self.mesh = mesh
        self.params = params
        X = np.ascontiguousarray(mesh.X, dtype=np.float64)
        tets = np.ascontiguousarray(mesh.tets, dtype=np.int32)
        faces = np.ascontiguousarray(mesh.faces, dtype=np.int32)
        codes = np.ascontiguousarray(mesh.codes, dtype=np.int32)
        cm = _CMesh(mesh.num_nodes, mesh.num_tets, mesh.num_faces, X.ctypes.data_as(c_double_p),
                    tets.ctypes.data_as(c_int_p), faces.ctypes.data_as(c_int_p),
                    codes.ctypes.data_as(c_int_p))
        p = params.c()
        h = C.c_void_p()
        check(lib().aerofx_ctx_create(C.byref(cm), C.byref(p), C.byref(h)), "ctx_create")
        self._h = h
        N, E, F = C.c_int(), C.c_int(), C.c_int()
        check(lib().aerofx_ctx_sizes(h, C.byref(N), C.byref(E), C.byref(F)), "ctx_sizes")
        self.num_nodes, self.num_edges, self.num_faces = N.value, E.value, F.value
        self.set_stream(stream)

    def close(self):
        if getattr(self, "_h", None):
            lib().aerofx_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream=None):
        """Binds the torch stream (default: torch's current stream)."""
        import torch
        if stream is None and torch.cuda.is_available():
            stream = torch.cuda.current_stream()
        if stream is None:
            check(lib().aerofx_ctx_set_stream(self._h, None, 1), "set_stream")
        else:
            check(lib().aerofx_ctx_set_stream(self._h, C.c_void_p(stream.cuda_stream), 0), "set_stream")

    def profile(self, enable: bool = True):
        """Turns per-kernel event timing on/off (resets the buckets)."""
        check(lib().aerofx_ctx_profile(self._h, int(enable)), "ctx_profile")

    def timers(self) -> dict:
        """{bucket: (total ms, launches)} for buckets with launches."""
        from ._abi import TIMER_NAMES
        ms = (C.c_double * 8)()
        n = (C.c_longlong * 8)()
        check(lib().aerofx_ctx_timers(self._h, ms, n), "ctx_timers")
        return {TIMER_NAMES[k]: (ms[k], n[k]) for k in range(8) if n[k]}

    def sync(self):
        check(lib().aerofx_ctx_sync(self._h), "sync")

    def tables(self):
        """Host copies: edge table, ctrlVol, edge normals, face normals, LSQ R."""
        N, E, F = self.num_nodes, self.num_edges, self.num_faces
        ptr = np.empty((E, 2), np.int32)
        cv = np.empty(N, np.float64)
        en = np.empty((E, 3), np.float64)
        fn = np.empty((F, 3), np.float64)
        R = np.empty((N, 6), np.float64)
        check(lib().aerofx_ctx_tables(self._h, ptr.ctypes.data_as(c_int_p), cv.ctypes.data_as(c_double_p),
                                      en.ctypes.data_as(c_double_p), fn.ctypes.data_as(c_double_p),
                                      R.ctypes.data_as(c_double_p)), "ctx_tables")
        return dict(ptr=ptr, ctrl_vol=cv, edge_normals=en, face_normals=fn, lsq=R)

    def compute_residual(self, U, R, fluxes_only: bool = False) -> int:
        """computeResidual (SpaceOperator.C:524-634): R = fluxes / V.

        Returns the count of non-physical reconstructions (0 when valid)."""
        return check(lib().aerofx_residual_ex(self._h, _ptr(U), _ptr(R), 1 if fluxes_only else 0),
                     "residual")

    def compute_residual_async(self, U, R):
        check(lib().aerofx_residual_ex(self._h, _ptr(U), _ptr(R), 2), "residual")

    def compute_gradients(self, V, ddx, ddy, ddz):
        """DistNodalGrad::compute, least squares, on primitive V."""
        check(lib().aerofx_gradients(self._h, _ptr(V), _ptr(ddx), _ptr(ddy), _ptr(ddz)), "gradients")

    def conservative_to_primitive(self, U, V):
        check(lib().aerofx_primitive(self._h, _ptr(U), _ptr(V)), "primitive")
