"""Fourier-space evaluation and derivatives of Gaussian atoms.

evaluate(atoms, S)             FTGaussian3D.evaluate / FTGaussian2D.evaluate
                               (GaussDictCode/bin/FTGaussian3D.py:12-52,
                               FTGaussian2D.py:13-35) with the reference's
                               dispatch on the sample set S: a separable 3D grid
                               (gt_ft_eval), 3D or 2.5D hyperplanes, a 2D grid or
                               2D lines (gt_ft_planes_eval; float64 samples give
                               complex128 through gt_ft_planes_eval_f64).
derivs(atoms, S, C)            FTGaussian3D.derivs (:55-92): complex L2 value,
derivs_all(atoms, S, C)        gradient and Hessian over hyperplane samples, of
                               atom 0 as the reference / of every atom
                               (gt_ft_derivs).
GaussFTVolume(ASpace, PSpace)  Fourier_Transform.py:177-207: atoms -> their FT on
                               the Fourier-slice planes of a projection space.
analytic_gauss_fourier         (I, x, r) of f(y) = I exp(-|U(y-x)|^2 / 2) ->
analytic_gauss_ifourier        FT-space parameters and back
                               (Fourier_Transform.py:293-376); host FP64.
get_grid(coords)               GaussFT.getGrid (Fourier_Transform.py:163-174).
apply_weight_hyperplane        _applyweight_hyperplane (:259-290).
dft_projections(arr, space)    the projection-space DFT of GaussFT.__DFT
                               (:103-160): Fourier-slice data to compare with
                               GaussFTVolume.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib, ops
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


def evaluate_grid(atoms, k, out=None):
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
    """FT-space atom set on the device: I (n,), x (n, dim), r (n, nr) tensors
    (float32 by default)."""

    def __init__(self, I, x, r, device, dim=None, dtype=np.float32):
        f = lambda v: torch.as_tensor(np.ascontiguousarray(v, dtype=dtype), device=device)
        n = len(np.asarray(I).reshape(-1))
        dim = dim or (np.asarray(x).size // max(n, 1) if n else 3)
        r = np.asarray(r)
        # an empty set keeps the slot count of its (0, k) array (default: anisotropic)
        r = r.reshape(n, -1) if n else r.reshape(0, r.shape[-1] if r.ndim == 2 else dim * (dim + 1) // 2)
        self.I, self.x, self.r = f(np.reshape(I, -1)), f(np.reshape(x, (-1, dim))), f(r)


# ---------------------------------------------------------------------------
# Sample planes: GaussDictCode/bin/FTGaussian3D.py:12-38, FTGaussian2D.py:13-35
# ---------------------------------------------------------------------------
def _is_f64(v):
    return (isinstance(v, torch.Tensor) and v.dtype == torch.float64) or \
        (not isinstance(v, torch.Tensor) and np.asarray(v).dtype == np.float64)


def _ndim(v):
    return v.dim() if isinstance(v, torch.Tensor) else np.ndim(v)


class DevicePlanes:
    """Device copy of a plane sample set and its gt_kplanes struct."""

    def __init__(self, form, k0, k1, w0, w1, device, f64=False):
        dt = torch.float64 if f64 else torch.float32
        cv = lambda a: None if a is None else torch.as_tensor(np.ascontiguousarray(
            a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a, dtype=np.float64), device=device).to(dt)
        self.k0, self.k1, self.w0, self.w1 = cv(k0), cv(k1), cv(w0), cv(w1)
        self.form, self.f64 = form, f64
        m0 = self.k0.numel()
        if form == _lib.GT_KP_2D_LINE:
            nA, m1 = self.w0.shape[0], 1
        elif form == _lib.GT_KP_2D_GRID:
            nA, m1 = 1, self.k1.numel()
        else:
            nA, m1 = self.w0.shape[0], self.k1.numel()
        self.nA, self.m0, self.m1 = int(nA), int(m0), int(m1)
        asc = lambda t: t is None or ops._ascending(t.cpu().numpy())
        sorted_ = int(asc(self.k0) and (form == _lib.GT_KP_2D_LINE or asc(self.k1)))
        ptr = lambda t: t.data_ptr() if t is not None else None
        cls = _lib.gt_kplanes64 if f64 else _lib.gt_kplanes
        self.struct = cls(form, self.nA, self.m0, self.m1, sorted_, ptr(self.k0), ptr(self.k1), ptr(self.w0),
                          ptr(self.w1))


def _planes_of(atoms, S, f64):
    """(form, DevicePlanes, output shape) of the reference's sample sets."""
    dev = atoms.I.device
    dim = atoms.x.shape[1]
    if dim == 3:
        if len(S) == 4:
            k0, k1, w0, w1 = S
            P = DevicePlanes(_lib.GT_KP_3D, k0, k1, w0, w1, dev, f64)
        elif _ndim(S[2]) == 2:
            k0, k1, w0 = S
            P = DevicePlanes(_lib.GT_KP_2P5D, k0, k1, w0, None, dev, f64)
        else:
            return None
        return P, (P.nA, P.m0, P.m1)
    k0, k1 = S
    if _ndim(k1) == 1:
        P = DevicePlanes(_lib.GT_KP_2D_GRID, k0, k1, None, None, dev, f64)
        return P, (P.m0, P.m1)
    P = DevicePlanes(_lib.GT_KP_2D_LINE, k0, None, k1, None, dev, f64)
    return P, (P.nA, P.m0)


def evaluate(atoms, S, out=None):
    """FT of FT-space atoms on the samples S, with the reference's dispatch
    (FTGaussian3D.py:12-52, FTGaussian2D.py:13-35): 3D atoms: (k0, k1, w0, w1)
    hyperplanes -> (nA, m0, m1); (k0, k1, w0) with w0 of shape (nA, 2) the
    2.5D planes -> (nA, m0, m1); (k0, k1, k2) a separable grid -> (m0, m1, m2).
    2D atoms: (k0, k1) a grid -> (m0, m1); (k, w) with w (nA, 2) lines ->
    (nA, m).  complex64; for 2D atoms complex128 when k0 is float64 (the
    reference's 2D f8 signatures)."""
    if isinstance(S, DeviceKGrid):
        return evaluate_grid(atoms, S, out)
    # the reference's 3D evaluator is float32 only; its 2D one follows k0's dtype
    f64 = atoms.x.shape[1] == 2 and _is_f64(S[0])
    got = _planes_of(atoms, S, f64)
    if got is None:
        return evaluate_grid(atoms, S, out)
    P, shape = got
    lib = load()
    dev = atoms.I.device
    n = atoms.I.numel()
    if out is None:
        out = torch.empty(shape, dtype=torch.complex128 if f64 else torch.complex64, device=dev)
    if f64:
        I, x, r = (t.to(torch.float64).contiguous() for t in (atoms.I, atoms.x, atoms.r))
        a = _lib.gt_atoms64(atoms.x.shape[1], n, int(r.reshape(max(n, 1), -1).shape[1] == 1) if n else 0,
                            I.data_ptr(), x.data_ptr(), r.data_ptr())
        need = lib.gt_ft_planes_workspace_f64(n, ctypes.byref(P.struct))
        ws = ops._ws.get(need, dev)
        check(lib.gt_ft_planes_eval_f64(ctypes.byref(a), ctypes.byref(P.struct), ctypes.c_void_p(out.data_ptr()),
                                        ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle(dev)),
              "gt_ft_planes_eval_f64")
    else:
        a = ops.atoms_struct(atoms)
        need = lib.gt_ft_planes_workspace(n, ctypes.byref(P.struct))
        ws = ops._ws.get(need, dev)
        check(lib.gt_ft_planes_eval(ctypes.byref(a), ctypes.byref(P.struct), ctypes.c_void_p(out.data_ptr()),
                                    ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle(dev)),
              "gt_ft_planes_eval")
    return out


def derivs_all(atoms, S, C, order=2):
    """For every 3D FT-space atom against all of C (complex, (nA, m0, m1)) on
    the hyperplane samples S = (k0, k1, w0[, w1]): f (n,), g (n, P), H (n, P, P)
    FP64 device tensors (FTGaussian3D.derivs semantics per atom, cut at
    |L k|^2 > 60); order 1 leaves H out."""
    if atoms.x.shape[1] != 3:
        raise NotImplementedError("Fourier derivatives exist for 3D atoms only (FTGaussian2D.derivs raises)")
    got = _planes_of(atoms, S, False)
    if got is None:
        raise NotImplementedError("Fourier derivatives are defined on hyperplane samples only (FTGaussian3D.py:88-90)")
    P, shape = got
    lib = load()
    dev = atoms.I.device
    C = C.asarray() if hasattr(C, "asarray") else C
    C = torch.as_tensor(C, device=dev).to(torch.complex64).reshape(shape).contiguous()
    n = atoms.I.numel()
    Pn = 5 if (n and atoms.r.reshape(n, -1).shape[1] == 1) else 10
    f = torch.empty(n, dtype=torch.float64, device=dev)
    g = torch.empty((n, Pn), dtype=torch.float64, device=dev)
    H = torch.empty((n, Pn, Pn), dtype=torch.float64, device=dev) if order == 2 else None
    a = ops.atoms_struct(atoms)
    need = lib.gt_ft_derivs_workspace(ctypes.byref(a), ctypes.byref(P.struct))
    ws = ops._ws.get(need, dev)
    check(lib.gt_ft_derivs(ctypes.byref(a), ctypes.byref(P.struct), ctypes.c_void_p(C.data_ptr()), ops._ptr(f),
                           ops._ptr(g), ops._ptr(H), ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream_handle(dev)),
          "gt_ft_derivs")
    return f, g, H


def derivs(atoms, S, C):
    """FTGaussian3D.derivs (FTGaussian3D.py:55-92): (f, Df, DDf) of atom 0 as
    host arrays."""
    one = ops._AtomView(atoms.I[:1], atoms.x[:1], atoms.r[:1])
    f, g, H = derivs_all(one, S, C, 2)
    return float(f[0]), g[0].cpu().numpy(), H[0].cpu().numpy()


class GaussFTVolume:
    """Fourier_Transform.GaussFTVolume (Fourier_Transform.py:177-207): FT-space
    atoms -> their values on the Fourier-slice planes of a projection space:
    frequencies get_grid(PSpace.detector) along the detector axes, the planes'
    orientation from PSpace.ortho.  L2_derivs, which the reference leaves
    unimplemented (:201-207), returns (f (n, 1), df (n, P), ddf (n, P, P)) per
    atom like GaussTomo.L2_derivs."""

    def __init__(self, ASpace, PSpace):
        from .spaces import ProjSpace
        self.domain = ASpace
        freq = tuple(get_grid(PSpace.detector))
        self.range = ProjSpace(orientations=PSpace.orientations, ortho=PSpace.ortho, detector=freq)
        self.grid = freq + tuple(PSpace.ortho)

    def __call__(self, atoms):
        from .spaces import ProjElement
        return ProjElement(self.range, evaluate(atoms, self.grid))

    def L2_derivs(self, atoms, C, order=2):
        f, g, H = derivs_all(atoms, self.grid, C, order)
        f = f.cpu().numpy().reshape(-1, 1)
        if order == 1:
            return f, g.cpu().numpy()
        return f, g.cpu().numpy(), H.cpu().numpy()


def apply_weight_hyperplane(f, det, w, dx, x0):
    """_applyweight_hyperplane (Fourier_Transform.py:259-290): f times prod(dx)
    times exp(-2 pi i sum_j det[j] x0[j]) along the detector frequency axes
    (x0 the first detector coordinates; the reference's 2D branch uses dx[0],
    2.5D and 3D dx[0] dx[1])."""
    f = torch.as_tensor(f)
    dev = f.device
    x0 = [2 * np.pi * float(np.asarray(v).reshape(-1)[0]) for v in x0]
    if len(det) == 1:
        s = float(dx[0])
        if all(v == 0 for v in x0):
            return f * s
        ph = torch.as_tensor(np.exp(-1j * np.asarray(det[0], float) * x0[0]), device=dev)
        return f * s * ph.reshape(1, -1).to(f.dtype if f.is_complex() else torch.complex128)
    s = float(dx[0]) * float(dx[1])
    p0 = np.exp(-1j * np.asarray(det[0], float) * x0[0]).reshape(1, -1, 1)
    p1 = np.exp(-1j * np.asarray(det[1], float) * x0[1]).reshape(1, 1, -1)
    ph = torch.as_tensor(p0 * p1, device=dev)
    return f * s * ph.to(f.dtype if f.is_complex() else torch.complex128)


def dft_projections(arr, space, inverse=False):
    """GaussFT.__DFT for projection data (Fourier_Transform.py:103-160): along
    the detector axes, fftshift(fftn(arr)) weighted by apply_weight_hyperplane
    (detector spacing and the phase of the first detector coordinate), or the
    inverse.  arr: (nA, n0[, n1]) tensor; returns complex."""
    arr = torch.as_tensor(arr.asarray() if hasattr(arr, "asarray") else arr)
    det = [np.asarray(d, float) for d in space.detector]
    ax = tuple(range(1, 1 + len(det)))
    freq = get_grid(det)
    if not inverse:
        dx = [d[1] - d[0] for d in det]
        x0 = [d[0] for d in det]
        FT = torch.fft.fftshift(torch.fft.fftn(arr, dim=ax), dim=ax)
        return apply_weight_hyperplane(FT, freq, space.ortho, dx, x0)
    dx = [1.0 / (d[1] - d[0]) for d in det]
    x0 = [-d[0] for d in det]
    FT = apply_weight_hyperplane(arr, freq, space.ortho, dx, x0)
    return torch.fft.ifftn(torch.fft.ifftshift(FT, dim=ax), dim=ax)
