"""Regularisers of the atom fits (GaussDictCode/regularisation.py:10-442).

Same classes, constructor arguments and methods as the reference:

    null(dim)                          regularisation.py:22-38
    Joubert(dim, intensity, location, radius)   :41-108
    Mass(dim, m, k=2, s=1)             :111-210
    Volume(dim, r0, k=2, s=1)          :213-318
    Radius(dim, r0, s, k)              :321-386
    Iso(dim, s)                        :389-442

Each is a sum over atoms of one term per atom.  The terms are evaluated on the
device by gt_reg_eval in FP64 (the same code the fitter kernels add to every
atom update); reg(atoms) returns the total (+inf outside the domain, as the
reference), grad(atoms, axis) the gradient blocks in the reference's shapes
and hess(atoms) the per-atom Hessians.

Differences from the reference, each a repair of a defect there:
  * grad and hess are the exact derivatives of the value.  The reference's
    Volume.grad leaves out the factor k (:248-262), Volume.hess is not the
    derivative and multiplies by the undefined self.s (:286-318), Mass.hess
    skips the (r1, r0) and (r2, r1) cross terms (range(i - 1), :199-205) and
    Joubert.hess returns None (:93-94).
  * Joubert's domain test looks at the diagonal slots of r, the ones inside
    its logarithm; the reference tests every slot (:63-64), so any atom with a
    non-positive off-diagonal slot made the value +inf.
  * grad(atoms, axis=None) returns [dI, dx, dr] concatenated for every class
    (the reference's null and Joubert have no such branch).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import ops
from ._lib import (GT_REG_ISO, GT_REG_JOUBERT, GT_REG_MASS, GT_REG_NPARAM, GT_REG_NULL, GT_REG_RADIUS,
                   GT_REG_VOLUME, check, load, stream_handle)


class Regulariser:
    """Base class: a per-atom term evaluated on the device."""

    kind = GT_REG_NULL
    default_axis = None

    def __init__(self, dim):
        self.dim = dim

    def params(self):
        """The GT_REG_NPARAM doubles gt_reg_eval and the fit slots take."""
        return np.zeros(GT_REG_NPARAM)

    def terms(self, atoms, derivs=2):
        """Per-atom values (n,), and for derivs >= 1 gradients (n, P) and for
        derivs == 2 Hessians (n, P, P), FP64 device tensors."""
        n = atoms.I.numel()
        dev = atoms.I.device
        P = 1 + atoms.x.shape[1] + atoms.r.reshape(max(n, 1), -1).shape[1]
        val = torch.empty(n, dtype=torch.float64, device=dev)
        g = torch.empty((n, P), dtype=torch.float64, device=dev) if derivs >= 1 else None
        H = torch.empty((n, P, P), dtype=torch.float64, device=dev) if derivs >= 2 else None
        prm = (ctypes.c_double * GT_REG_NPARAM)(*[float(v) for v in self.params()])
        a = ops.atoms_struct(atoms)
        check(load().gt_reg_eval(ctypes.byref(a), self.kind, prm, ops._ptr(val), ops._ptr(g), ops._ptr(H),
                                 stream_handle(dev)), "gt_reg_eval")
        return val, g, H

    def __call__(self, atoms):
        if atoms.I.numel() == 0:
            return 0.0
        val = self.terms(atoms, 0)[0]
        return float(val.sum())

    def grad(self, atoms, axis="default"):
        """axis 0/'I': dI (n,); 1/'x': dx (n, dim); 2/'r': dr (n, nr);
        None: the three concatenated, as the reference's
        concatenate((dI.reshape(-1), dx.reshape(-1), dr.reshape(-1)))."""
        if axis == "default":
            axis = self.default_axis
        D = atoms.x.shape[1]
        g = self.terms(atoms, 1)[1].cpu().numpy()
        dI, dx, dr = g[:, 0], g[:, 1:1 + D], g[:, 1 + D:]
        if axis in (0, "I"):
            return dI
        if axis in (1, "x"):
            return dx
        if axis in (2, "r"):
            return dr
        return np.concatenate((dI.reshape(-1), dx.reshape(-1), dr.reshape(-1)))

    def hess(self, atoms):
        """Hessian of the term of each atom: (P, P) for one atom, else (n, P, P)."""
        H = self.terms(atoms, 2)[2].cpu().numpy()
        return H[0] if H.shape[0] == 1 else H


class null(Regulariser):
    kind = GT_REG_NULL
    default_axis = 0


class Joubert(Regulariser):
    """-a log I + b/2 |x|^2 + c v - d log v - e per atom, v = r^dim (iso) or
    the product of the diagonal slots (regularisation.py:41-108)."""

    kind = GT_REG_JOUBERT
    default_axis = 0

    def __init__(self, dim, intensity, location, radius):
        self.dim = dim
        self.param = (intensity, location, radius[1], radius[1] * radius[0],
                      radius[0] * radius[1] * (1 - math.log(radius[0])))

    def params(self):
        return np.array(self.param, dtype=np.float64)

    def __mul__(self, scal):
        if np.isscalar(scal):
            return Joubert(self.dim, self.param[0] * scal, self.param[1] * scal,
                           (self.param[3] / self.param[2], self.param[2] * scal))
        return NotImplemented

    def __rmul__(self, scal):
        return self.__mul__(scal)


class Mass(Regulariser):
    """s (|I R| - m)^k per atom, R = r^dim (iso) or 1 / product of the diagonal
    slots (regularisation.py:111-210)."""

    kind = GT_REG_MASS

    def __init__(self, dim, m, k=2, s=1):
        self.dim = dim
        self.param = (m, k, s)

    def params(self):
        return np.array(list(self.param) + [0.0, 0.0], dtype=np.float64)


class Volume(Regulariser):
    """s |v - r0^dim|^k per atom, v = r^-dim (iso) or the product of the
    diagonal slots (regularisation.py:213-318)."""

    kind = GT_REG_VOLUME

    def __init__(self, dim, r0, k=2, s=1):
        self.dim = dim
        self.param = (abs(r0) ** dim, k, s)

    def params(self):
        return np.array(list(self.param) + [0.0, 0.0], dtype=np.float64)


class Radius(Regulariser):
    """s sum (r_i - r0)^-k over the diagonal slots, iso dim (r - r0)^-k
    (regularisation.py:321-386)."""

    kind = GT_REG_RADIUS

    def __init__(self, dim, r0, s, k):
        self.dim = dim
        self.r0, self.s, self.k = r0, s, -k

    def params(self):
        return np.array([self.r0, self.s, self.k, 0.0, 0.0], dtype=np.float64)


class Iso(Regulariser):
    """s/2 times the sum of the squared off-diagonal slots (regularisation.py:389-442)."""

    kind = GT_REG_ISO

    def __init__(self, dim, s):
        self.dim = dim
        self.s = s

    def params(self):
        return np.array([self.s, 0.0, 0.0, 0.0, 0.0], dtype=np.float64)


def as_regulariser(reg):
    """None -> null; a Regulariser passes through."""
    if reg is None:
        return null(None)
    if isinstance(reg, Regulariser):
        return reg
    raise TypeError(f"not a regulariser of this package: {reg!r}")
