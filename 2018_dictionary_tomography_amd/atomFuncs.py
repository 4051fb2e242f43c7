"""GaussTomo and SingleParticleGaussTomo: the reference's projector
operators over the HIP library.

Same methods, argument meaning and return shapes as
GaussDictCode/atomFuncs.py:77-172; the four module functions of
GaussTomo._getOps (atomFuncs.py:106-113) are replaced by batched device
calls, so derivs / L2_derivs cost one launch sequence for all atoms instead of
one per atom.  SingleParticleGaussTomo (:161-172) binds the pixel-integrated
(point-spread) projector instead.
"""
from __future__ import annotations

import numpy as np
import torch

from . import ops, psf
from .spaces import ProjElement, VolElement


class GaussTomo:
    def __init__(self, Atoms, Volume, Sino, device="GPU"):
        if str(device).lower() == "cpu":
            raise ValueError("CPU processing not supported: the projector runs on the GPU only")
        self.domain = Atoms
        self.embedding = Volume
        self.range = Sino
        self.linear = False

    # -- forward ---------------------------------------------------------
    def __call__(self, atoms, out=None):
        """Sinogram of `atoms` (atomFuncs.py:115-119)."""
        data = out.data if out is not None else None
        R = ops.forward(atoms, self.range, out=data)
        return out if out is not None else ProjElement(self.range, R)

    def discretise(self, atoms, out=None):
        """Volume render of `atoms` on the embedding grid (atomFuncs.py:121-125)."""
        data = out.data if out is not None else None
        u = ops.volume(atoms, self.embedding, out=data)
        return out if out is not None else VolElement(self.embedding, u)

    # -- derivatives -----------------------------------------------------
    def _derivs(self, atoms, C, order, mode):
        C = C.data if hasattr(C, "data") and not isinstance(C, torch.Tensor) else C
        f, g, H = ops.derivs(atoms, self.range, C, mode, order)
        f = f.cpu().numpy().reshape(-1, 1)
        g = g.cpu().numpy()
        if order == 1:
            return f, g
        return f, g, H.cpu().numpy()

    def derivs(self, atoms, C, order=2):
        """Per atom: R(a_i).C, its gradient and Hessian (atomFuncs.py:130-142).
        Returns f (n, 1), df (n, P), ddf (n, P, P) as host arrays."""
        return self._derivs(atoms, C, order, "linear")

    def gradient(self, x):
        """Linear map d -> derivs(x, d, 1)[1] (atomFuncs.py:144)."""
        return lambda d: self.derivs(x, d, order=1)[1]

    def L2_derivs(self, atoms, C, order=2):
        """Per atom: |R(a_i) - C|^2 / 2, gradient, Hessian (atomFuncs.py:146-158)."""
        return self._derivs(atoms, C, order, "L2")

    # -- device-resident variants (no host copies) ------------------------
    def derivs_device(self, atoms, C, order=2, mode="linear"):
        return ops.derivs(atoms, self.range, C, mode, order)


class SingleParticleGaussTomo(GaussTomo):
    """GaussTomo with the point-spread projector (atomFuncs.py:161-172):
    rule 'exact' integrates detector values over pixels and the volume over
    voxels; rule 'reference' takes the reference's own samples of them
    (psf module)."""

    def __init__(self, Atoms, Volume, Sino, device="GPU", rule="exact"):
        super().__init__(Atoms, Volume, Sino, device=device)
        psf.check_rule(rule)
        self.rule = rule

    def __call__(self, atoms, out=None):
        data = out.data if out is not None else None
        R = psf.forward(atoms, self.range, out=data, rule=self.rule)
        return out if out is not None else ProjElement(self.range, R)

    def discretise(self, atoms, out=None):
        data = out.data if out is not None else None
        u = psf.volume(atoms, self.embedding, out=data, rule=self.rule)
        return out if out is not None else VolElement(self.embedding, u)

    def _derivs(self, atoms, C, order, mode):
        C = C.data if hasattr(C, "data") and not isinstance(C, torch.Tensor) else C
        f, g, H = psf.derivs(atoms, self.range, C, mode, order, self.rule)
        f = f.cpu().numpy().reshape(-1, 1)
        g = g.cpu().numpy()
        if order == 1:
            return f, g
        return f, g, H.cpu().numpy()

    def derivs_device(self, atoms, C, order=2, mode="linear"):
        return psf.derivs(atoms, self.range, C, mode, order, self.rule)
