"""Atom, projection and volume spaces (the data model of the projector API).

Mirrors GaussDictCode/dictionary_def.py without ODL: atoms live on the GPU as
float32 torch tensors in the reference layout (I[n], x[n, dim], r[n, k] with
r slots [r00, r11, (r22,) r01, (r12, r02)], dictionary_def.py:147-214), and
projection geometry is explicit per-angle arrays (dictionary_def.py:263-283).
"""
from __future__ import annotations

import numpy as np
import torch


def default_device():
    return torch.device("cuda", torch.cuda.current_device())


# ---------------------------------------------------------------------------
# Geometry vectors, GaussDictCode/atomFuncs.py:13-74
# ---------------------------------------------------------------------------
def theta_to_vec(theta):
    """Ray directions for a tuple of angle vectors (atomFuncs.py:13-47).

    One angle vector: T = (-sin t, cos t) (2D and single-axis 3D).
    Two angle vectors (3D Euler grid): T = (-sin t0 cos t1, cos t0 cos t1, sin t1)."""
    theta = [np.asarray(t, dtype=np.float64).reshape(-1) for t in theta]
    if len(theta) == 1:
        t = theta[0]
        return np.stack([-np.sin(t), np.cos(t)], axis=-1)
    if len(theta) == 2:
        t0, t1 = np.meshgrid(theta[0], theta[1], indexing="ij")
        return np.stack([-np.sin(t0) * np.cos(t1), np.cos(t0) * np.cos(t1), np.sin(t1)], axis=-1)
    raise ValueError("theta must hold one or two angle vectors")


def perp(theta):
    """Detector axes for a tuple of angle vectors (atomFuncs.py:50-74)."""
    theta = [np.asarray(t, dtype=np.float64).reshape(-1) for t in theta]
    if len(theta) == 1:
        t = theta[0]
        return (np.stack([np.cos(t), np.sin(t)], axis=-1),)
    if len(theta) == 2:
        t0, t1 = np.meshgrid(theta[0], theta[1], indexing="ij")
        w0 = np.stack([np.cos(t0), np.sin(t0), np.zeros_like(t0)], axis=-1)
        w1 = np.stack([np.sin(t0) * np.sin(t1), -np.cos(t0) * np.sin(t1), np.cos(t1)], axis=-1)
        return w0, w1
    raise ValueError("theta must hold one or two angle vectors")


def cell_centres(lo, hi, n):
    """Cell-centred coordinates of a uniform partition of [lo, hi] into n cells
    (what odl.uniform_partition(lo, hi, n).grid.coord_vectors holds)."""
    h = (hi - lo) / n
    return lo + h * (np.arange(n) + 0.5)


def boundary_nodes(lo, hi, n):
    """Nodes of a uniform partition with nodes_on_bdry=True (angle grids,
    GaussDictCode/__init__.py:143-144)."""
    return np.linspace(lo, hi, n)


def axis_tilt_geometry(angles, axis, ref=None):
    """Explicit (orientations, (w0, w1)) for a tilt series rotating about a unit
    axis: w1 = axis, w0 the rotated reference direction, T = w0 x w1."""
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    if ref is None:
        ref = np.array([1.0, 0.0, 0.0]) if abs(axis[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e0 = ref - axis * (ref @ axis)
    e0 /= np.linalg.norm(e0)
    e1 = np.cross(axis, e0)
    angles = np.asarray(angles, dtype=np.float64).reshape(-1)
    c, s = np.cos(angles)[:, None], np.sin(angles)[:, None]
    w0 = c * e0 + s * e1
    w1 = np.tile(axis, (angles.size, 1))
    T = np.cross(w0, w1)
    return T, (w0, w1)


# ---------------------------------------------------------------------------
# Atoms, GaussDictCode/dictionary_def.py:72-214
# ---------------------------------------------------------------------------
class AtomSpace:
    def __init__(self, dim, isotropic=True):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self.dim = dim
        self.isotropic = isotropic
        self.nr = 1 if isotropic else dim * (dim + 1) // 2

    def element(self, *args, **kwargs):
        return AtomElement(self, *args, **kwargs)

    def random(self, n, seed=None, device=None):
        """Same draw as AtomSpace.random (dictionary_def.py:130-139): r first,
        then x, then I, from numpy's legacy generator."""
        rs = np.random.RandomState(seed)
        if self.isotropic:
            r = rs.rand(n, 1) + 1
        else:
            r = rs.rand(n, self.nr) + 1
            r[:, self.dim:] /= 4
        x = 0.8 * (2 * rs.rand(n, self.dim) - 1)
        I = 2 * rs.rand(n)
        return AtomElement(self, x, r, I, device=device)

    def zero(self, device=None):
        return AtomElement(self, np.zeros((0, self.dim)), np.zeros((0, self.nr)), np.zeros(0), device=device)

    def __eq__(self, other):
        return getattr(other, "dim", None) == self.dim and getattr(other, "isotropic", None) == self.isotropic


def _as_f32(v, device):
    if isinstance(v, torch.Tensor):
        return v.detach().to(device=device, dtype=torch.float32)
    return torch.as_tensor(np.asarray(v, dtype=np.float32), device=device)


class AtomElement:
    """n atoms with device tensors I (n,), x (n, dim), r (n, nr), all float32."""

    def __init__(self, space, x=None, r=None, I=None, device=None):
        self.space = space
        device = device or default_device()
        dim, nr = space.dim, space.nr
        x = _as_f32(x, device).reshape(-1, dim)
        n = x.shape[0]
        r = _as_f32(r, device)
        if r.numel() == 1:
            val = r.reshape(())
            r = torch.zeros((n, nr), device=device, dtype=torch.float32)
            r[:, : (1 if space.isotropic else dim)] = val
        elif (not space.isotropic) and r.numel() == n and nr > 1:
            r = torch.cat([r.reshape(n, 1).repeat(1, dim), torch.zeros((n, nr - dim), device=device)], 1)
        r = r.reshape(n, -1) if n > 0 else r.reshape(0, nr)
        if r.shape[1] != nr:
            raise ValueError(f"r must have {nr} slots per atom, got {r.shape[1]}")
        I = _as_f32(I, device)
        if I.numel() == 1 and n != 1:
            I = I.reshape(()).repeat(n)
        I = I.reshape(-1)
        if I.numel() != n:
            raise ValueError("I must have one entry per atom")
        self.x, self.r, self.I = x.contiguous(), r.contiguous(), I.contiguous()

    @property
    def size(self):
        return int(self.I.numel())

    def __len__(self):
        return self.size

    @property
    def device(self):
        return self.I.device

    def copy(self):
        return AtomElement(self.space, self.x.clone(), self.r.clone(), self.I.clone(), device=self.device)

    def __getitem__(self, i):
        if isinstance(i, int):
            i = slice(i, i + 1) if i != -1 else slice(-1, None)
        return AtomElement(self.space, self.x[i].clone(), self.r[i].clone(), self.I[i].clone(), device=self.device)

    def __setitem__(self, i, atoms):
        if isinstance(i, int):
            i = slice(i, i + 1)
        self.x[i] = atoms.x
        self.r[i] = atoms.r
        self.I[i] = atoms.I

    @property
    def data(self):
        """[x | r | I] per atom (the layout dictionary_def.py:208-212 intends)."""
        return torch.cat([self.x, self.r, self.I.reshape(-1, 1)], 1)

    def asarray(self):
        return self.data


# ---------------------------------------------------------------------------
# Projection and volume spaces
# ---------------------------------------------------------------------------
class ProjSpace:
    """Projection geometry: orientations (nA, lenT), ortho (1 or 2 arrays),
    detector coordinate vectors (1 or 2).  Built from angle vectors through
    theta_to_vec/perp, or from explicit arrays (dictionary_def.py:263-283)."""

    def __init__(self, angles=None, detector=None, orientations=None, ortho=None):
        if angles is not None:
            if not isinstance(angles, (tuple, list)) or np.isscalar(angles[0]):
                angles = (angles,)
            self.orientations = theta_to_vec(angles)
            self.ortho = perp(angles)
        elif orientations is not None and ortho is not None:
            self.orientations = np.asarray(orientations, dtype=np.float64)
            self.ortho = tuple(np.asarray(o, dtype=np.float64) for o in ortho)
        else:
            raise ValueError("give either angles or orientations and ortho")
        d = self.orientations.shape[-1]
        self.orientations = self.orientations.reshape(-1, d)
        self.ortho = tuple(o.reshape(-1, o.shape[-1]) for o in self.ortho)
        if detector is None:
            raise ValueError("detector coordinate vectors are required")
        if not isinstance(detector, (tuple, list)) or np.isscalar(detector[0]):
            detector = (detector,)
        self.detector = [np.asarray(p, dtype=np.float64).reshape(-1) for p in detector]
        self.shape = [self.orientations.shape[0]] + [p.size for p in self.detector]
        self._dev = {}

    @property
    def dim(self):
        return 2 if len(self.detector) == 1 else 3

    def element(self, arr=None):
        return ProjElement(self, 0 if arr is None else arr)

    def zero(self):
        return ProjElement(self, 0)


class ProjElement:
    """Projection data (float32), or its Fourier-slice transform (complex)."""

    def __init__(self, space, u, device=None):
        self.space = space
        device = device or default_device()
        if isinstance(u, torch.Tensor):
            u = u.to(device=device, dtype=u.dtype if u.is_complex() else torch.float32)
        elif np.iscomplexobj(u):
            u = torch.as_tensor(np.asarray(u), device=device)
        else:
            u = torch.as_tensor(np.asarray(u, dtype=np.float32), device=device)
        if u.numel() == 1:
            u = u.reshape(()).expand(*space.shape).contiguous()
        if tuple(u.shape) != tuple(space.shape):
            raise ValueError("Shape of array does not match shape of space")
        self.data = u.contiguous()
        self.shape = tuple(u.shape)

    def asarray(self):
        return self.data

    def copy(self):
        return ProjElement(self.space, self.data.clone())


class VolSpace:
    """Volume grid given by its coordinate vectors (VolSpace.coord_vectors,
    dictionary_def.py:217-230).  VolSpace(box=(lo, hi), shape=n, dim=d) builds
    cell-centred vectors like odl.uniform_discr."""

    def __init__(self, coord_vectors=None, box=None, shape=None, dim=None):
        if coord_vectors is None:
            lo, hi = box
            dim = dim or 3
            shape = [shape] * dim if np.isscalar(shape) else list(shape)
            coord_vectors = [cell_centres(lo, hi, n) for n in shape]
        self.coord_vectors = tuple(np.asarray(c, dtype=np.float64).reshape(-1) for c in coord_vectors)
        self.shape = tuple(c.size for c in self.coord_vectors)
        self.dim = len(self.coord_vectors)
        self._dev = {}

    def element(self, arr):
        return VolElement(self, arr)


class VolElement:
    def __init__(self, space, u):
        self.space = space
        self.data = u if isinstance(u, torch.Tensor) else torch.as_tensor(np.asarray(u, dtype=np.float32))
        self.shape = tuple(self.data.shape)

    def asarray(self):
        return self.data
