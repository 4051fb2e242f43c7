"""Fourier-space evaluation of Gaussian atoms.

evaluate(atoms, grid)          FTGaussian3D.evaluate on a separable 3D grid
                               (GaussDictCode/bin/FTGaussian3D.py:12-52, 95-110,
                               151-164), computed by gt_ft_eval.
analytic_gauss_fourier         (I, x, r) of f(y) = I exp(-|U(y-x)|^2 / 2) ->
analytic_gauss_ifourier        FT-space parameters and back
                               (Fourier_Transform.py:293-376); host FP64.
get_grid(coords)               GaussFT.getGrid (Fourier_Transform.py:163-174).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import ops
from ._lib import check, gt_kgrid, load, stream_handle


def get_grid(coords):
    """Frequencies matching an FFT of samples on `coords`: fftshift(fftfreq(n, dx))."""
    out = []
    for g in coords:
        g = np.asarray(g)
        out.append(np.fft.fftshift(np.fft.fftfreq(g.size, g[1] - g[0])).astype(g.dtype))
    return out


def analytic_gauss_fourier(I, x, r):
    """FT of I exp(-|U(y-x)|^2/2) is newI exp(-i newx.k) exp(-|newU^T k|^2/2) with
    newI = I (2 pi)^{d/2} / det U, newx = 2 pi x, newU = 2 pi U^{-1} (upper
    triangular, same slots); isotropic radii: newI = I (2 pi)^{d/2} r^d, newr = 2 pi / r."""
    I, x, r = (np.asarray(v, dtype=np.float64) for v in (I, x, r))
    n, D = x.shape
    tau = 2 * np.pi
    nI = np.empty(n)
    nr = np.empty_like(r)
    nx = tau * x
    if r.shape[1] == 1:
        nI[:] = I * np.sqrt(tau) ** D * r[:, 0] ** D
        nr[:, 0] = tau / r[:, 0]
        return nI, nx, nr
    nI[:] = I * np.sqrt(tau) ** D / np.prod(r[:, :D], axis=1)
    nr[:] = _inverse_slots(r, D, tau)
    return nI, nx, nr


def analytic_gauss_ifourier(I, x, r):
    """Inverse of analytic_gauss_fourier."""
    I, x, r = (np.asarray(v, dtype=np.float64) for v in (I, x, r))
    n, D = x.shape
    tau = 2 * np.pi
    nx = x / tau
    nr = np.empty_like(r)
    if r.shape[1] == 1:
        nr[:, 0] = tau / r[:, 0]
        return I / (nr[:, 0] ** D * np.sqrt(tau) ** D), nx, nr
    nr[:] = _inverse_slots(r, D, tau)
    return I * np.prod(nr[:, :D], axis=1) / np.sqrt(tau) ** D, nx, nr


def _inverse_slots(r, D, scale):
    """Slots of scale * U^{-1} for upper-triangular U given by its slots."""
    out = np.empty_like(r)
    if D == 2:
        out[:, 0] = scale / r[:, 0]
        out[:, 1] = scale / r[:, 1]
        out[:, 2] = -scale * r[:, 2] / (r[:, 0] * r[:, 1])
    else:
        out[:, 0] = scale / r[:, 0]
        out[:, 1] = scale / r[:, 1]
        out[:, 2] = scale / r[:, 2]
        out[:, 3] = -scale * r[:, 3] / (r[:, 0] * r[:, 1])
        out[:, 4] = -scale * r[:, 4] / (r[:, 1] * r[:, 2])
        out[:, 5] = scale * (r[:, 3] * r[:, 4] / r[:, 1] - r[:, 5]) / (r[:, 0] * r[:, 2])
    return out


class DeviceKGrid:
    def __init__(self, k, device):
        f32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=device)
        self.k = [f32(v) for v in k]
        self.shape = tuple(int(v.numel()) for v in self.k)
        sorted_ = int(all(ops._ascending(np.asarray(v)) for v in k))
        self.struct = gt_kgrid(3, self.shape[0], self.shape[1], self.shape[2], sorted_, self.k[0].data_ptr(),
                               self.k[1].data_ptr(), self.k[2].data_ptr())


def evaluate(atoms, k, out=None):
    """sum over FT-space atoms of I exp(-|U^T k|^2/2) exp(-i x.k) (cut at 30) on the
    separable grid k = (k0, k1, k2).  Returns a complex64 tensor (m0, m1, m2)."""
    lib = load()
    dev = atoms.I.device
    grid = k if isinstance(k, DeviceKGrid) else DeviceKGrid(k, dev)
    a = ops.atoms_struct(atoms)
    if a.dim != 3:
        raise ValueError("only 3D atoms are supported by the Fourier kernel")
    if out is None:
        out = torch.empty(grid.shape, dtype=torch.complex64, device=dev)
    need = lib.gt_ft_workspace(ctypes.byref(a), ctypes.byref(grid.struct))
    ws = ops._ws.get(need, dev)
    check(lib.gt_ft_eval(ctypes.byref(a), ctypes.byref(grid.struct), ctypes.c_void_p(out.data_ptr()),
                         ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle(dev)), "gt_ft_eval")
    return out


class FTAtoms:
    """FT-space atom set on the device (float32 I, x, r)."""

    def __init__(self, I, x, r, device):
        f = lambda v: torch.as_tensor(np.ascontiguousarray(v, dtype=np.float32), device=device)
        self.I, self.x, self.r = f(I), f(np.reshape(x, (-1, 3))), f(np.reshape(r, (len(I), -1)))
