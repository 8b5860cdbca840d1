"""Test-only torch.autograd.Function over the residual and its adjoint kernels.

forward(U, X) evaluates the GPU residual on the mesh with coordinates X;
backward(gR) returns ((dR/dU)^T gR, (dR/dX)^T gR) from the adjoint entry
point.  A context is built per distinct X (geometry is set up on the host
from X), so this is for small test meshes only.
"""
import dataclasses

import numpy as np


def residual_fn(mesh, params):
    import torch

    from aerof_dev_amd.operators import SpaceOperator

    cache = {}

    def op_for(X):
        key = X.detach().cpu().numpy().tobytes()
        op = cache.get(key)
        if op is None:
            Xh = np.ascontiguousarray(X.detach().cpu().numpy(), dtype=np.float64)
            op = SpaceOperator(dataclasses.replace(mesh, X=Xh), params)
            cache.clear()
            cache[key] = op
        return op

    class Residual(torch.autograd.Function):
        @staticmethod
        def forward(ctx, U, X):
            op = op_for(X)
            Uc = U.detach().contiguous()
            R = torch.empty_like(Uc)
            ierr = op.compute_residual(Uc, R)
            if ierr:
                raise RuntimeError(f"non-physical reconstruction count {ierr}")
            ctx.save_for_backward(Uc, X.detach())
            ctx.op = op
            return R

        @staticmethod
        def backward(ctx, gR):
            U, X = ctx.saved_tensors
            gU = torch.empty_like(U)
            gX = torch.empty_like(X)
            ctx.op.compute_transpose_derivative_of_residual(U, gR.contiguous(), gU, gX)
            return gU, gX

    return Residual.apply
