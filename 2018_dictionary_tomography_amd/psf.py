"""Point-spread ("single particle") projector over the HIP library.

Device API (torch tensors, current HIP stream):
    forward(atoms, space, out=None, rule="exact")          sinogram, float32
    derivs(atoms, space, C, mode, order=2, rule="exact")   per-atom (f, g, H) in FP64
    volume(atoms, vol, out=None, rule="exact")             render, float32

Module-level drop-ins for GaussDictCode/bin/point_spread_functions.py (the
module SingleParticleGaussTomo._getOps binds, GaussDictCode/atomFuncs.py:168-172):
    RadProj(atom, Rad, R)                    point_spread_functions.py:371-385
    L2derivs_RadProj(atom, Rad, C, order=2)  :1077-1115
    derivs_RadProj(atom, Rad, C, order=2)    :1314-1350
    VolProj(atom, y, u)                      :168-197
each with an optional rule= as the device API.

rule "exact": every detector value is the integral of the atom's line
integral over its pixel (the square of side |p[1] - p[0]| around the detector
coordinate), every volume value the integral of the profile over its voxel.

rule "reference": the reference's own numbers (__rf2 / __rf3 :413-527,
__drf2 / __drf3 :924-1075, __if2 / __if3 :222-291).  Per (atom, cell), with n
the distance of the (projected) centre from the cell centre and R = 1 / min
diag U: n + R < s gives the atom's total with no derivative, n - R > s gives
0, and otherwise the integrand is sampled at 11 points per axis, from one
tenth of the interval past its lower end to one tenth past its upper end, on
the cell cut to [-R, R] about the centre (the mean of the samples times the
interval); s is half the first detector axis' spacing on both detector axes
and half the second volume axis' spacing on every volume axis.  Isotropic
derivatives use the exact chain rule in both rules.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import ops
from ._lib import GT_MODE_L2, GT_MODE_LINEAR, GT_PSF_EXACT, GT_PSF_SAMPLED, check, load, stream_handle

_MODES = {"linear": GT_MODE_LINEAR, "L2": GT_MODE_L2, "l2": GT_MODE_L2, GT_MODE_LINEAR: GT_MODE_LINEAR,
          GT_MODE_L2: GT_MODE_L2}
_RULES = {"exact": GT_PSF_EXACT, "reference": GT_PSF_SAMPLED, "sampled": GT_PSF_SAMPLED}


def check_rule(rule):
    """The library's code for a cell rule name ('exact', or 'reference' /
    'sampled' for the reference's own samples)."""
    if rule not in _RULES:
        raise ValueError(f"rule must be one of {sorted(_RULES)}, got {rule!r}")
    return _RULES[rule]


def forward(atoms, space, out=None, rule="exact"):
    code = check_rule(rule)
    lib = load()
    dev = atoms.I.device
    geom = ops.device_geometry(space, dev)
    a = ops.atoms_struct(atoms)
    if out is None:
        out = torch.empty(geom.shape, dtype=torch.float32, device=dev)
    if tuple(out.shape) != tuple(geom.shape) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float32 tensor of shape {geom.shape}")
    need = lib.gt_psf_forward_workspace(ctypes.byref(a), ctypes.byref(geom.struct))
    ws = ops._ws.get(need, dev)
    check(lib.gt_psf_forward_rule(ctypes.byref(a), ctypes.byref(geom.struct), code, ops._ptr(out), ops._ptr(ws),
                                  ws.numel(), stream_handle(dev)), "gt_psf_forward")
    return out


def derivs(atoms, space, C, mode="L2", order=2, rule="exact"):
    """Per atom, against all of C: mode 'L2': f = 1/2 sum (P - C)^2, 'linear':
    f = sum P C, with gradient g (n, P) and Hessian H (n, P, P), FP64."""
    code = check_rule(rule)
    lib = load()
    mode = _MODES[mode]
    dev = atoms.I.device
    geom = ops.device_geometry(space, dev)
    a = ops.atoms_struct(atoms)
    C = C.data if hasattr(C, "data") and not isinstance(C, torch.Tensor) else C
    C = torch.as_tensor(C, device=dev).to(torch.float32).contiguous()
    if tuple(C.shape) != tuple(geom.shape):
        raise ValueError(f"C must have shape {geom.shape}, got {tuple(C.shape)}")
    P = ops.n_params(a.dim, a.iso)
    f = torch.zeros(a.n, dtype=torch.float64, device=dev)
    g = torch.empty((a.n, P), dtype=torch.float64, device=dev)
    H = torch.empty((a.n, P, P), dtype=torch.float64, device=dev) if order > 1 else None
    need = lib.gt_psf_derivs_workspace(ctypes.byref(a), ctypes.byref(geom.struct))
    ws = ops._ws.get(need, dev)
    check(lib.gt_psf_derivs_rule(ctypes.byref(a), ctypes.byref(geom.struct), ops._ptr(C), mode, order, code,
                                 ops._ptr(f), ops._ptr(g), ops._ptr(H), ops._ptr(ws), ws.numel(), stream_handle(dev)),
          "gt_psf_derivs")
    return f, g, H


def volume(atoms, vol, out=None, rule="exact"):
    code = check_rule(rule)
    lib = load()
    dev = atoms.I.device
    dv = ops.device_volume(vol, dev)
    a = ops.atoms_struct(atoms)
    if out is None:
        out = torch.empty(dv.shape, dtype=torch.float32, device=dev)
    if tuple(out.shape) != tuple(dv.shape) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float32 tensor of shape {dv.shape}")
    need = lib.gt_psf_volume_workspace(ctypes.byref(a), ctypes.byref(dv.struct))
    ws = ops._ws.get(need, dev)
    check(lib.gt_psf_volume_rule(ctypes.byref(a), ctypes.byref(dv.struct), code, ops._ptr(out), ops._ptr(ws),
                                 ws.numel(), stream_handle(dev)), "gt_psf_volume")
    return out


# ---------------------------------------------------------------------------
# Module-level drop-ins with the reference's argument conventions
# ---------------------------------------------------------------------------
def RadProj(atom, Rad, R, rule="exact"):
    """Write the point-spread sinogram of `atom` into R (the reference
    zero-fills R and accumulates, point_spread_functions.py:371-385)."""
    a = ops._atoms_on_device(atom)
    if isinstance(R, torch.Tensor) and R.is_cuda:
        out = forward(a, ops._range_space(Rad), out=R if R.is_contiguous() else None, rule=rule)
        if out is not R:
            R.copy_(out)
        return R
    out = forward(a, ops._range_space(Rad), rule=rule)
    np.copyto(R, out.cpu().numpy().reshape(np.shape(R)))
    return R


def _single(atom, Rad, C, order, mode, rule):
    a = ops._atoms_on_device(atom)
    a0 = ops._AtomView(a.I[:1].contiguous(), a.x[:1].contiguous(), a.r[:1].contiguous(), a.space)
    if hasattr(C, "asarray"):
        C = C.asarray()
    C = C if isinstance(C, torch.Tensor) else torch.as_tensor(np.asarray(C, dtype=np.float32))
    f, g, H = derivs(a0, ops._range_space(Rad), C.to(a0.I.device), mode, 2 if order > 1 else 1, rule)
    g = g[0].cpu().numpy()
    H = H[0].cpu().numpy() if H is not None else np.zeros((g.size, g.size))
    return f[0].item(), g, H


def L2derivs_RadProj(atom, Rad, C, order=2, rule="exact"):
    """(f, Df, DDf) of |P(atom[0]) - C|^2 / 2 (point_spread_functions.py:1077-1115)."""
    return _single(atom, Rad, C, order, "L2", rule)


def derivs_RadProj(atom, Rad, C, order=2, rule="exact"):
    """(f, Df, DDf) of P(atom[0]) . C (point_spread_functions.py:1314-1350)."""
    return _single(atom, Rad, C, order, "linear", rule)


def VolProj(atom, y, u, rule="exact"):
    """Write the point-spread render of `atom` on the grid y into u
    (point_spread_functions.py:168-197)."""
    a = ops._atoms_on_device(atom)
    if isinstance(u, torch.Tensor) and u.is_cuda:
        out = volume(a, y, out=u if u.is_contiguous() else None, rule=rule)
        if out is not u:
            u.copy_(out)
        return u
    out = volume(a, y, rule=rule)
    np.copyto(u, out.cpu().numpy().reshape(np.shape(u)))
    return u
