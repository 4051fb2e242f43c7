"""torch.autograd binding of the projector.

GaussProject.apply(I, x, r, space) returns the sinogram A(theta); its backward
is the hand-written adjoint kernel (gt_derivs in linear mode, order 1), i.e.
the reference's analytic derivative derivs_RadProj / GaussTomo.gradient
(GaussDictCode/bin/NewtonBases.py:954-1069, GaussDictCode/atomFuncs.py:130-144)
applied to the incoming gradient.  Only the atom parameters are saved: the
backward recomputes every per-(atom, angle) detector Gaussian and reduces the
pixel moments again, so no residual x Jacobian intermediates are stored.
"""
from __future__ import annotations

import torch

from . import ops


class _Atoms:
    def __init__(self, I, x, r):
        self.I, self.x, self.r = I, x, r


class GaussProject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, I, x, r, space):
        I32, x32, r32 = (t.detach().to(torch.float32).contiguous() for t in (I, x, r))
        R = ops.forward(_Atoms(I32, x32, r32), space)
        ctx.save_for_backward(I32, x32, r32)
        ctx.space = space
        ctx.dtypes = (I.dtype, x.dtype, r.dtype)
        return R

    @staticmethod
    def backward(ctx, grad_R):
        I32, x32, r32 = ctx.saved_tensors
        dim = x32.shape[1]
        _, g, _ = ops.derivs(_Atoms(I32, x32, r32), ctx.space, grad_R.contiguous(), "linear", order=1,
                             want_f=False)
        gI = g[:, 0].to(ctx.dtypes[0])
        gx = g[:, 1:1 + dim].to(ctx.dtypes[1])
        gr = g[:, 1 + dim:].to(ctx.dtypes[2])
        return gI, gx, gr, None


def project(I, x, r, space):
    """Differentiable sinogram of atoms (I, x, r) in `space`."""
    return GaussProject.apply(I, x, r, space)


def l2_loss(I, x, r, space, data):
    """1/2 ||A(theta) - data||^2 (transport_loss.l2_squared_loss, :567-585), differentiable."""
    R = project(I, x, r, space)
    d = R - data
    return 0.5 * (d.double() * d.double()).sum()
