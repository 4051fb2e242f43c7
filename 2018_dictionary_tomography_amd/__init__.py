"""MI355X-native Gaussian-atom tomography: the projector, its derivatives,
the volume render, the Fourier evaluation and the fitters of
robtovey/2018_dictionary_tomography, computed by hand-written gfx950 HIP
kernels behind the C ABI in include/gausstomo.h.

The package name starts with a digit, so import it with
``importlib.import_module("2018_dictionary_tomography_amd")``.
"""
from . import _lib, formats, kl, ops, spaces
from ._lib import GTError
from .atomFuncs import GaussTomo, SingleParticleGaussTomo
from .kl import doKL_LagrangeStep_iso, doKL_ProjGDStep_iso
from .losses import l2_squared_loss
from .ops import L2derivs_RadProj, RadProj, VolProj, derivs_RadProj
from .regularisation import Iso, Joubert, Mass, Radius, Volume, null
from .spaces import (AtomElement, AtomSpace, ProjElement, ProjSpace, VolElement, VolSpace, axis_tilt_geometry,
                     boundary_nodes, cell_centres, perp, theta_to_vec)
from .problem import (UniformPartition, getFid, getGeom, getReg, getSolver, standardGaussTomo,
                      standardSingleParticleGaussTomo)

__all__ = [
    "GTError", "GaussTomo", "SingleParticleGaussTomo", "standardSingleParticleGaussTomo", "RadProj", "derivs_RadProj",
    "L2derivs_RadProj", "VolProj", "AtomSpace", "doKL_ProjGDStep_iso", "doKL_LagrangeStep_iso", "kl", "formats",
    "AtomElement", "ProjSpace", "ProjElement", "VolSpace", "VolElement", "theta_to_vec", "perp",
    "axis_tilt_geometry", "cell_centres", "boundary_nodes", "ops", "spaces", "l2_squared_loss", "null", "Joubert",
    "Mass", "Volume", "Radius", "Iso", "UniformPartition", "standardGaussTomo", "getGeom", "getReg", "getSolver",
    "getFid",
]
