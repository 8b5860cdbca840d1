"""aerof_dev_amd -- MI355X-native edge-based FV residual and Newton-Krylov.

The compute path is libaerofx.so (hand-written gfx950 HIP kernels behind the
C ABI in include/aerofx.h).  This package holds the ctypes binding and a
Python mirror of the reference's operator surface.
"""
from ._abi import LIB_PATH, AerofxError, lib  # noqa: F401
from .mesh import FlowParams, LowMachPrec, Mesh, ViscousParams, box_mesh, channel_mesh, init_state, kuhn_box  # noqa: F401

__all__ = ["LIB_PATH", "AerofxError", "lib", "FlowParams", "LowMachPrec", "Mesh", "ViscousParams", "box_mesh", "channel_mesh",
           "init_state", "kuhn_box"]
