"""Device-side operators over the C ABI, and the reference's module-level
"myRad" interface.

Batched device API (torch tensors in, torch tensors out, current HIP stream):
    forward(atoms, space, out=None)              -> R (nA, n0[, n1]) float32
    derivs(atoms, space, C, mode, order=2)       -> f (n,), g (n, P), H (n, P, P) float64
    volume(atoms, vol, out=None)                 -> u float32

Module-level drop-ins for GaussDictCode/bin/NewtonBases.py and
NewtonBases2D.py (the four functions GaussTomo._getOps binds,
GaussDictCode/atomFuncs.py:106-113):
    RadProj(atom, Rad, R)                        NewtonBases.py:714-726
    derivs_RadProj(atom, Rad, C, order=2)        NewtonBases.py:954-990
    L2derivs_RadProj(atom, Rad, C, order=2)      NewtonBases.py:759-789
    VolProj(atom, y, u)                          NewtonBases.py:691-695
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import GT_MODE_L2, GT_MODE_LINEAR, GT_MODE_RESID, check, gt_atoms, gt_geom, gt_vol, load, stream_handle

MODES = {"linear": GT_MODE_LINEAR, "L2": GT_MODE_L2, "l2": GT_MODE_L2, "resid": GT_MODE_RESID,
         "residual": GT_MODE_RESID, GT_MODE_LINEAR: GT_MODE_LINEAR, GT_MODE_L2: GT_MODE_L2,
         GT_MODE_RESID: GT_MODE_RESID}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _ascending(p):
    p = np.asarray(p)
    return bool(p.size <= 1 or np.all(np.diff(p) > 0))


class DeviceGeometry:
    """Device copy of a ProjSpace (float32 arrays) and its gt_geom struct."""

    def __init__(self, space, device):
        T = np.asarray(space.orientations, dtype=np.float64)
        ortho, det = space.ortho, space.detector
        lenT = T.shape[1]
        dim = 2 if len(det) == 1 else 3
        if dim == 2 and (lenT != 2 or len(ortho) != 1):
            raise ValueError("2D geometry needs 2-vector orientations and one ortho array")
        if dim == 3 and lenT == 3 and len(ortho) != 2:
            raise ValueError("3D geometry with 3-vector orientations needs two ortho arrays")
        if dim == 3 and lenT == 2 and len(ortho) != 1:
            raise ValueError("single-axis 3D geometry needs one ortho array")
        f32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=device)
        self.dim, self.lenT = dim, lenT
        self.nA = T.shape[0]
        self.n0 = det[0].size
        self.n1 = det[1].size if dim == 3 else 1
        self.T = f32(T)
        self.w0 = f32(ortho[0])
        self.w1 = f32(ortho[1]) if (dim == 3 and lenT == 3) else None
        self.p0 = f32(det[0])
        self.p1 = f32(det[1]) if dim == 3 else None
        self.sorted = int(all(_ascending(p) for p in det))
        self.shape = (self.nA, self.n0, self.n1) if dim == 3 else (self.nA, self.n0)
        self.struct = gt_geom(dim, lenT, self.nA, self.n0, self.n1, self.sorted, self.T.data_ptr(),
                              self.w0.data_ptr(), self.w1.data_ptr() if self.w1 is not None else None,
                              self.p0.data_ptr(), self.p1.data_ptr() if self.p1 is not None else None)


def device_geometry(space, device=None):
    device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    if isinstance(space, DeviceGeometry):
        return space
    cache = getattr(space, "_dev", None)
    if cache is None:
        return DeviceGeometry(space, device)
    key = str(device)
    if key not in cache:
        cache[key] = DeviceGeometry(space, device)
    return cache[key]


def atoms_struct(atoms):
    """gt_atoms view of an AtomElement-like object (I, x, r device float32 tensors)."""
    I, x, r = atoms.I, atoms.x, atoms.r
    for t in (I, x, r):
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
            raise TypeError("atoms must be contiguous float32 CUDA tensors")
    n = I.numel()
    dim = x.shape[1] if x.dim() == 2 else x.numel() // max(n, 1)
    iso = int(r.reshape(n, -1).shape[1] == 1) if n > 0 else int(getattr(getattr(atoms, "space", None), "isotropic", 0))
    return gt_atoms(int(dim), int(n), iso, I.data_ptr(), x.data_ptr(), r.data_ptr())


class _Workspace:
    """Grow-only device scratch buffer per (device, stream)."""

    def __init__(self):
        self.buf = {}

    def get(self, nbytes, device):
        key = (str(device), torch.cuda.current_stream(device).cuda_stream)
        b = self.buf.get(key)
        if b is None or b.numel() < nbytes:
            b = torch.empty(max(nbytes, 1 << 20), dtype=torch.uint8, device=device)
            self.buf[key] = b
        return b


_ws = _Workspace()


def forward(atoms, space, out=None):
    """R = sum of atom line integrals (RadProj), written into out."""
    lib = load()
    geom = device_geometry(space, atoms.I.device)
    a = atoms_struct(atoms)
    if out is None:
        out = torch.empty(geom.shape, dtype=torch.float32, device=atoms.I.device)
    if tuple(out.shape) != tuple(geom.shape) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float32 tensor of shape {geom.shape}")
    need = lib.gt_forward_workspace(ctypes.byref(a), ctypes.byref(geom.struct))
    ws = _ws.get(need, atoms.I.device)
    check(lib.gt_forward(ctypes.byref(a), ctypes.byref(geom.struct), _ptr(out), _ptr(ws), ws.numel(),
                         stream_handle(atoms.I.device)), "gt_forward")
    return out


def n_params(dim, iso):
    return 1 + dim + (1 if iso else dim * (dim + 1) // 2)


def derivs(atoms, space, C, mode="L2", order=2, want_f=True):
    """Per-atom (f, g, H) in FP64 on the device.  mode: 'linear', 'L2' or 'resid'."""
    lib = load()
    mode = MODES[mode]
    dev = atoms.I.device
    geom = device_geometry(space, dev)
    a = atoms_struct(atoms)
    C = C.data if hasattr(C, "data") and not isinstance(C, torch.Tensor) else C
    C = torch.as_tensor(C, device=dev).to(torch.float32).contiguous()
    if tuple(C.shape) != tuple(geom.shape):
        raise ValueError(f"C must have shape {geom.shape}, got {tuple(C.shape)}")
    P = n_params(a.dim, a.iso)
    n = a.n
    f = torch.zeros(n, dtype=torch.float64, device=dev) if (want_f and mode != GT_MODE_RESID) else None
    g = torch.empty((n, P), dtype=torch.float64, device=dev)
    H = torch.empty((n, P, P), dtype=torch.float64, device=dev) if order > 1 else None
    need = lib.gt_derivs_workspace(ctypes.byref(a), ctypes.byref(geom.struct), mode)
    ws = _ws.get(need, dev)
    check(lib.gt_derivs(ctypes.byref(a), ctypes.byref(geom.struct), _ptr(C), mode, order, _ptr(f), _ptr(g), _ptr(H),
                        _ptr(ws), ws.numel(), stream_handle(dev)), "gt_derivs")
    return f, g, H


def forward_derivs(atoms, space, data, order=2, out=None, want_f=True):
    """One sweep of the l2 fit: resid = forward(atoms) - data (written into
    out), and per atom the gradient / Hessian block of 1/2||resid||^2 (FP64);
    f = 1/2||resid||^2 as a 0-d FP64 tensor.  Returns (resid, f, g, H)."""
    lib = load()
    dev = atoms.I.device
    geom = device_geometry(space, dev)
    a = atoms_struct(atoms)
    data = data.data if hasattr(data, "data") and not isinstance(data, torch.Tensor) else data
    data = torch.as_tensor(data, device=dev).to(torch.float32).contiguous()
    if tuple(data.shape) != tuple(geom.shape):
        raise ValueError(f"data must have shape {geom.shape}, got {tuple(data.shape)}")
    if out is None:
        out = torch.empty(geom.shape, dtype=torch.float32, device=dev)
    if tuple(out.shape) != tuple(geom.shape) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float32 tensor of shape {geom.shape}")
    P = n_params(a.dim, a.iso)
    f = torch.zeros((), dtype=torch.float64, device=dev) if want_f else None
    g = torch.empty((a.n, P), dtype=torch.float64, device=dev)
    H = torch.empty((a.n, P, P), dtype=torch.float64, device=dev) if order > 1 else None
    need = lib.gt_derivs_workspace(ctypes.byref(a), ctypes.byref(geom.struct), GT_MODE_RESID)
    ws = _ws.get(need, dev)
    check(lib.gt_forward_derivs(ctypes.byref(a), ctypes.byref(geom.struct), _ptr(data), order, _ptr(out), _ptr(f),
                                _ptr(g), _ptr(H), _ptr(ws), ws.numel(), stream_handle(dev)), "gt_forward_derivs")
    return out, f, g, H


def forward_derivs_noalloc(atoms, space, data, order=2, out=None, want_f=True):
    """forward_derivs through gt_forward_derivs_noalloc: the library takes
    the workspace from its per-GPU scratch array (at most
    gt_noalloc_scratch_bytes()) instead of a caller's buffer, so the library
    call allocates no device memory (this wrapper makes f, g, H, and the
    residual when `out` is omitted, as forward_derivs does).  Calls that
    overlap on one GPU race on the scratch array."""
    lib = load()
    dev = atoms.I.device
    geom = device_geometry(space, dev)
    a = atoms_struct(atoms)
    data = data.data if hasattr(data, "data") and not isinstance(data, torch.Tensor) else data
    data = torch.as_tensor(data, device=dev).to(torch.float32).contiguous()
    if tuple(data.shape) != tuple(geom.shape):
        raise ValueError(f"data must have shape {geom.shape}, got {tuple(data.shape)}")
    if out is None:
        out = torch.empty(geom.shape, dtype=torch.float32, device=dev)
    if tuple(out.shape) != tuple(geom.shape) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float32 tensor of shape {geom.shape}")
    P = n_params(a.dim, a.iso)
    f = torch.zeros((), dtype=torch.float64, device=dev) if want_f else None
    g = torch.empty((a.n, P), dtype=torch.float64, device=dev)
    H = torch.empty((a.n, P, P), dtype=torch.float64, device=dev) if order > 1 else None
    check(lib.gt_forward_derivs_noalloc(ctypes.byref(a), ctypes.byref(geom.struct), _ptr(data), order, _ptr(out),
                                        _ptr(f), _ptr(g), _ptr(H), stream_handle(dev)), "gt_forward_derivs_noalloc")
    return out, f, g, H


class DeviceVolume:
    def __init__(self, space, device):
        cv = space.coord_vectors
        f32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=device)
        self.dim = len(cv)
        self.y = [f32(c) for c in cv]
        self.shape = tuple(c.size for c in cv)
        sorted_ = int(all(_ascending(c) for c in cv))
        m2 = self.shape[2] if self.dim == 3 else 1
        self.struct = gt_vol(self.dim, self.shape[0], self.shape[1], m2, sorted_, self.y[0].data_ptr(),
                             self.y[1].data_ptr(), self.y[2].data_ptr() if self.dim == 3 else None)


def device_volume(space, device):
    if isinstance(space, DeviceVolume):
        return space
    if not hasattr(space, "coord_vectors"):
        from .spaces import VolSpace

        space = VolSpace(space)
    cache = getattr(space, "_dev", None)
    if cache is None:
        return DeviceVolume(space, device)
    key = str(device)
    if key not in cache:
        cache[key] = DeviceVolume(space, device)
    return cache[key]


def volume(atoms, vol, out=None):
    """Volume render (VolProj / GaussTomo.discretise)."""
    lib = load()
    dev = atoms.I.device
    dv = device_volume(vol, dev)
    a = atoms_struct(atoms)
    if out is None:
        out = torch.empty(dv.shape, dtype=torch.float32, device=dev)
    if tuple(out.shape) != tuple(dv.shape) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float32 tensor of shape {dv.shape}")
    need = lib.gt_volume_workspace(ctypes.byref(a), ctypes.byref(dv.struct))
    ws = _ws.get(need, dev)
    check(lib.gt_volume(ctypes.byref(a), ctypes.byref(dv.struct), _ptr(out), _ptr(ws), ws.numel(),
                        stream_handle(dev)), "gt_volume")
    return out


# ---------------------------------------------------------------------------
# Module-level drop-ins with the reference's argument conventions
# ---------------------------------------------------------------------------
class _AtomView:
    def __init__(self, I, x, r, space=None):
        self.I, self.x, self.r, self.space = I, x, r, space


def _atoms_on_device(atom, device=None):
    dev = device or torch.device("cuda", torch.cuda.current_device())
    conv = lambda v: (v.detach().to(dev, torch.float32) if isinstance(v, torch.Tensor)
                      else torch.as_tensor(np.asarray(v, dtype=np.float32), device=dev)).contiguous()
    I = conv(atom.I).reshape(-1)
    n = I.numel()
    x = conv(atom.x).reshape(n, -1)
    r = conv(atom.r).reshape(n, -1)
    return _AtomView(I, x, r, getattr(atom, "space", None))


def _range_space(Rad):
    return Rad.range if hasattr(Rad, "range") else Rad


def RadProj(atom, Rad, R):
    """Write the sinogram of `atom` into R (NewtonBases.py:714-726)."""
    a = _atoms_on_device(atom)
    if isinstance(R, torch.Tensor) and R.is_cuda:
        out = forward(a, _range_space(Rad), out=R if R.is_contiguous() else None)
        if out is not R:
            R.copy_(out)
        return R
    out = forward(a, _range_space(Rad))
    np.copyto(R, out.cpu().numpy().reshape(np.shape(R)))
    return R


def _single_atom_derivs(atom, Rad, C, order, mode):
    a = _atoms_on_device(atom)
    a0 = _AtomView(a.I[:1].contiguous(), a.x[:1].contiguous(), a.r[:1].contiguous(), a.space)
    if hasattr(C, "asarray"):
        C = C.asarray()
    C = C if isinstance(C, torch.Tensor) else torch.as_tensor(np.asarray(C, dtype=np.float32))
    f, g, H = derivs(a0, _range_space(Rad), C.to(a0.I.device), mode, 2 if order > 1 else 1)
    f = f[0].item()
    g = g[0].cpu().numpy()
    # The reference returns a (P, P) array at every order (left unfilled at
    # order 1, NewtonBases.py:393,444); GaussTomo.derivs indexes it, so order 1
    # gives zeros here.
    H = H[0].cpu().numpy() if H is not None else np.zeros((g.size, g.size))
    return f, g, H


def derivs_RadProj(atom, Rad, C, order=2):
    """(f, Df, DDf) of R(atom[0]) . C (NewtonBases.py:954-990)."""
    return _single_atom_derivs(atom, Rad, C, order, "linear")


def L2derivs_RadProj(atom, Rad, C, order=2):
    """(f, Df, DDf) of |R(atom[0]) - C|^2 / 2 (NewtonBases.py:759-789)."""
    return _single_atom_derivs(atom, Rad, C, order, "L2")


def VolProj(atom, y, u):
    """Write the volume render of `atom` on the grid y into u (NewtonBases.py:691-695)."""
    a = _atoms_on_device(atom)
    if isinstance(u, torch.Tensor) and u.is_cuda:
        out = volume(a, y, out=u if u.is_contiguous() else None)
        if out is not u:
            u.copy_(out)
        return u
    out = volume(a, y)
    np.copyto(u, out.cpu().numpy().reshape(np.shape(u)))
    return u
