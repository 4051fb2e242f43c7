"""Problem factory: standardGaussTomo and its helpers without ODL.

Restates GaussDictCode/__init__.py:19-256:

    standardGaussTomo(...) -> (Radon, fidelity, data, ASpace, PSpace, (reg, GD))   :19-69
    getGeom(...)           -> (vol, angles, detector)                            :125-162
    getReg(dim, reg), getSolver(solver), getFid(dim, fidelity, device)           :177-210

The reference builds its grids with ODL partitions.  Here they are the
same coordinate vectors held by UniformPartition: cell centres for the
volume and the detector (odl.uniform_discr / odl.uniform_partition,
:136,158-160), and nodes on both ends of the interval for the angles
(nodes_on_bdry=True, :143-144).  A scalar angle_num gives one angle vector
(2D, or single-axis "2.5D" 3D geometry rotating about z); a pair gives the
3D Euler grid (:141-142).

Not provided: the ODL solver path (solver='odl', TV reconstructions) and
ODL's voxel ray transform, which the reference uses to simulate data from a
voxel ground truth (:27-31).  Ground truth given as atoms (an AtomElement)
is projected with the Gaussian projector instead; with neither data nor
ground truth the data are zero, as in the reference.
"""
from __future__ import annotations

import functools

import numpy as np
import torch

from . import fit
from .atomFuncs import GaussTomo, SingleParticleGaussTomo
from .losses import l2_squared_loss
from .regularisation import Iso, Joubert, Mass, Radius, Volume, null
from .spaces import AtomElement, AtomSpace, ProjElement, ProjSpace, VolSpace


class UniformPartition:
    """Axis-aligned uniform partition of [min_pt, max_pt] into `shape` cells
    with its sample points: cell centres, or (nodes_on_bdry=True) points
    spread so that the first and last sit on the interval's ends."""

    def __init__(self, min_pt, max_pt, shape, nodes_on_bdry=False):
        self.min_pt = tuple(float(v) for v in np.atleast_1d(min_pt))
        self.max_pt = tuple(float(v) for v in np.atleast_1d(max_pt))
        self.shape = tuple(int(v) for v in np.atleast_1d(shape))
        if not (len(self.min_pt) == len(self.max_pt) == len(self.shape)):
            raise ValueError("min_pt, max_pt and shape must have the same length")
        if any(n < 1 for n in self.shape):
            raise ValueError("every axis needs at least one cell")
        self.nodes_on_bdry = bool(nodes_on_bdry)
        vecs = []
        for lo, hi, n in zip(self.min_pt, self.max_pt, self.shape):
            if self.nodes_on_bdry:
                if n < 2:
                    raise ValueError("nodes on the boundary need at least two points per axis")
                vecs.append(np.linspace(lo, hi, n))
            else:
                vecs.append(lo + (hi - lo) / n * (np.arange(n) + 0.5))
        self.coord_vectors = tuple(vecs)

    @property
    def ndim(self):
        return len(self.shape)

    @classmethod
    def from_points(cls, points):
        """The partition whose sample points are the given uniformly spaced
        vector(s) (what odl.uniform_partition_fromgrid keeps, :247-256)."""
        if np.isscalar(points[0]):
            points = (points,)
        obj = cls.__new__(cls)
        vecs = tuple(np.asarray(p, dtype=np.float64).reshape(-1) for p in points)
        for v in vecs:
            if v.size > 2 and not np.allclose(np.diff(v), v[1] - v[0], rtol=1e-6, atol=1e-12):
                raise ValueError("angles must be uniformly spaced")
        obj.coord_vectors = vecs
        obj.shape = tuple(v.size for v in vecs)
        obj.min_pt = tuple(float(v[0]) for v in vecs)
        obj.max_pt = tuple(float(v[-1]) for v in vecs)
        obj.nodes_on_bdry = True
        return obj


def _padshape(arr, shape):
    """Broadcast a value over a shape, the rule the problem factory applies to
    its box and size arguments (GaussDictCode/__init__.py:223-244):

    * shape a single length n (an int or a one-entry list): a value with one
      entry (or a scalar) becomes n copies of it; a value with n entries is
      kept as given; any other length is an error;
    * shape a list of lengths [n0, n1, ...]: a value with one entry becomes
      one tuple of n_i copies per length; a value with one entry per length
      is broadcast entry by entry; any other length is an error."""
    lengths = [shape] if np.isscalar(shape) else list(shape)
    entries = [arr] if np.isscalar(arr) else list(arr)
    one = len(entries) == 1
    if len(lengths) == 1:
        n = lengths[0]
        if one:
            return (entries[0],) * n
        if len(entries) != n:
            raise ValueError(f"cannot broadcast {len(entries)} values to length {n}: give one value or {n}")
        return arr
    if one:
        return [(entries[0],) * n for n in lengths]
    if len(entries) != len(lengths):
        raise ValueError(f"cannot broadcast {len(entries)} values over {len(lengths)} axes: give one value or one "
                         "per axis")
    return [_padshape(v, n) for v, n in zip(entries, lengths)]


def getGeom(vol_box, vol_size, gt, dim, angles, angle_num, angle_range, det_box, det_size, data):
    """(vol, angles, detector) of GaussDictCode/__init__.py:125-162: vol a
    VolSpace of cell centres on vol_box (default [-1, 1]^dim), angles and
    detector UniformPartitions (detector default [-1.3, 1.3]^(dim-1))."""
    if vol_box is None:
        vol_box = [-1, 1]
    if vol_size is None:
        if gt is None or isinstance(gt, AtomElement):
            raise ValueError("If <gt> is not provided then <vol_size> must.")
        vol_size = np.shape(gt)
    vol_box = _padshape(vol_box, [dim, dim])
    vol_size = _padshape(vol_size, dim)
    vol_part = UniformPartition(vol_box[0], vol_box[1], vol_size)
    vol = VolSpace(coord_vectors=vol_part.coord_vectors)
    vol.partition = vol_part

    if angles is None:
        if angle_range is None or angle_num is None:
            raise ValueError("give either <angles> or both <angle_range> and <angle_num>")
        if dim == 3 and not (np.isscalar(angle_num) or len(angle_num) == 1):
            angle_range = _padshape(angle_range, [dim - 1, dim - 1])
        angles = UniformPartition(angle_range[0], angle_range[1], angle_num, nodes_on_bdry=True)
    elif not isinstance(angles, UniformPartition):
        angles = UniformPartition.from_points(angles)

    if det_box is None:
        det_box = [-1.3, 1.3]
    if det_size is None:
        if data is not None:
            det_size = np.shape(data.asarray() if hasattr(data, "asarray") else data)[1:]
        elif gt is not None and not isinstance(gt, AtomElement):
            det_size = np.shape(gt)[1:]
        else:
            raise ValueError("If <det_size> is not provided then either <data> or <gt> must.")
    det_box = _padshape(det_box, [dim - 1, dim - 1])
    det_size = _padshape(det_size, dim - 1)
    detector = UniformPartition(det_box[0], det_box[1], det_size)
    return vol, angles, detector


def getReg(dim, reg):
    """None -> null; ('Joubert' | 'Mass' | 'Radius' | 'Iso', *args) -> that
    regulariser (GaussDictCode/__init__.py:175-190).  ('Volume', *args) is
    accepted too (the reference defines Volume but does not map it here)."""
    if reg is None:
        return null(None)
    table = {"Joubert": Joubert, "Mass": Mass, "Radius": Radius, "Iso": Iso, "Volume": Volume}
    if not isinstance(reg, (tuple, list)) or reg[0] not in table:
        raise ValueError("Given <reg>, %s, was not recognised." % repr(reg))
    return table[reg[0]](dim, *reg[1:])


def getSolver(solver):
    """'Newton' -> Newton linesearch; 'Newton_block' -> its block variant;
    '*_block' -> block gradient descent (GaussDictCode/__init__.py:193-202).
    Other names map to GD_lib.linesearch in the reference, which cannot run
    (it passes the un-indexed gradient to _step, GD_lib.py:50-53)."""
    if solver == "Newton":
        return fit.newton_linesearch
    if solver == "Newton_block":
        return fit.newton_linesearch_block
    if solver.endswith("_block"):
        return fit.linesearch_block
    raise ValueError(f"solver {solver!r}: GD_lib.linesearch is not provided (its reference form cannot run); "
                     "use 'Newton', 'Newton_block' or '<name>_block'")


def getFid(dim, fidelity, device=None):
    """'l2_squared' -> l2_squared_loss (GaussDictCode/__init__.py:205-210);
    a fidelity object passes through."""
    if fidelity == "l2_squared":
        return l2_squared_loss(dim)
    if fidelity == "Transport":
        raise NotImplementedError("the Transport (Wasserstein) fidelity is not part of this package")
    return fidelity


def standardGaussTomo(gt=None, data=None, noise=None, dim=3, device="GPU", isotropic=False, vol_box=None,
                      vol_size=None, angles=None, angle_range=None, angle_num=None, axis=(0, 0, 1), det_box=None,
                      det_size=None, fidelity="l2_squared", reg=None, solver="Newton", _operator=GaussTomo):
    """GaussDictCode/__init__.py:19-69 for the Gaussian solvers.  Returns
    (Radon, fidelity, data, ASpace, PSpace, (reg, GD)).

    `axis` only entered the reference's ODL geometry; the Gaussian projector
    rotates single-axis tilt series about z (atomFuncs.py:13-74)."""
    if solver == "odl":
        raise NotImplementedError("the ODL solver path (voxel reconstructions) is not part of this package")
    if data is None and det_size is None and isinstance(gt, AtomElement):
        raise ValueError("with atoms as ground truth give <det_size>")
    vol, angles, detector = getGeom(vol_box, vol_size, gt, dim, angles, angle_num, angle_range, det_box, det_size,
                                    data)
    ASpace = AtomSpace(dim, isotropic=isotropic)
    PSpace = ProjSpace(angles.coord_vectors, detector.coord_vectors)
    Radon = _operator(ASpace, vol, PSpace, device=device)

    if data is None:
        if gt is None:
            data = PSpace.zero().asarray()
        elif isinstance(gt, AtomElement):
            data = Radon(gt).asarray()
        else:
            raise NotImplementedError("data from a voxel ground truth need ODL's ray transform; "
                                      "give <data> or atoms as <gt>")
    data = data.asarray() if hasattr(data, "asarray") else data
    if noise is not None:
        if np.isscalar(noise):  # a variance (GaussDictCode/__init__.py:32-35)
            noise = np.random.randn(*np.shape(data)) * noise ** .5
        if isinstance(data, torch.Tensor):
            noise = torch.as_tensor(noise, dtype=data.dtype, device=data.device)
        data = data + noise
    shape = tuple(np.shape(data))
    data = ProjElement(PSpace, data.reshape((-1,) + shape[len(shape) - dim + 1:]))

    fidelity = getFid(dim, fidelity, device)
    reg, GD = getReg(dim, reg), getSolver(solver)
    return Radon, fidelity, data, ASpace, PSpace, (reg, GD)


def standardSingleParticleGaussTomo(gt=None, data=None, noise=None, dim=3, device="GPU", isotropic=False,
                                    vol_box=None, vol_size=None, angles=None, angle_range=None, angle_num=None,
                                    axis=(0, 0, 1), det_box=None, det_size=None, fidelity="l2_squared", reg=None,
                                    solver="Newton", rule="exact"):
    """GaussDictCode/__init__.py:72-122: standardGaussTomo with the
    point-spread projector (SingleParticleGaussTomo with cell rule `rule`).
    Data simulated from atoms as ground truth use that projector."""
    return standardGaussTomo(gt, data, noise, dim, device, isotropic, vol_box, vol_size, angles, angle_range,
                             angle_num, axis, det_box, det_size, fidelity, reg, solver,
                             _operator=functools.partial(SingleParticleGaussTomo, rule=rule))
