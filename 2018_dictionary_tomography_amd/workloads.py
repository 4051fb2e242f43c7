"""Synthetic workloads: atom draws and geometries of the benchmark configurations.

Atom parametrisation: the configurations give a lower-triangular Cholesky
factor L of a precision matrix (density ~ exp(-(y-x)^T L L^T (y-x) / 2)).  The
projector's atom is I sqrt(30/pi) exp(-30 |U (y - x)|^2)
(GaussDictCode/bin/NewtonBases.py:40-46, 237-257), so U = L^T / sqrt(60) and
the r slots are [U00, U11, U22, U01, U12, U02] (NewtonBases.py:319-325).
"""
from __future__ import annotations

import numpy as np

from .spaces import axis_tilt_geometry, boundary_nodes, cell_centres, perp, theta_to_vec

SQRT60 = np.sqrt(60.0)


def cholesky_to_slots(L):
    """(n, 3, 3) lower-triangular L -> (n, 6) projector r slots (U = L^T / sqrt 60)."""
    L = np.asarray(L, dtype=np.float64)
    return np.stack([L[:, 0, 0], L[:, 1, 1], L[:, 2, 2], L[:, 1, 0], L[:, 2, 1], L[:, 2, 0]], 1) / SQRT60


def draw_aniso3(n, rng, centre=0.7, rho=(0.02, 0.1), offdiag=5.0, intensity=(0.5, 1.5)):
    """3D anisotropic atoms: centres U[-c, c]^3, L diagonal 1/rho with rho
    uniform, off-diagonal N(0, offdiag), intensities uniform.  Returns
    (I, x, r, L) in float64."""
    x = rng.uniform(-centre, centre, (n, 3))
    I = rng.uniform(intensity[0], intensity[1], n)
    rr = rng.uniform(rho[0], rho[1], (n, 3))
    L = np.zeros((n, 3, 3))
    for i in range(3):
        L[:, i, i] = 1.0 / rr[:, i]
    L[:, 1, 0] = rng.normal(0, offdiag, n)
    L[:, 2, 0] = rng.normal(0, offdiag, n)
    L[:, 2, 1] = rng.normal(0, offdiag, n)
    return I, x, cholesky_to_slots(L), L


def draw_iso2(n, rng, centre=0.8, rho=(0.02, 0.08), intensity=(0.5, 1.5)):
    """2D isotropic atoms with standard deviation rho: r = 1 / (rho sqrt 60)."""
    x = rng.uniform(-centre, centre, (n, 2))
    rr = rng.uniform(rho[0], rho[1], n)
    I = rng.uniform(intensity[0], intensity[1], n)
    return I, x, (1.0 / (rr * SQRT60)).reshape(n, 1)


def tilt_series(n_angles, max_deg=60.0, axis=(0.0, 1.0, 0.0), det=128, box=1.0):
    """Parallel-beam tilt series over [-max, max] degrees about `axis`, square
    cell-centred detector on [-box, box]^2.  Returns (T, (w0, w1), [p0, p1])."""
    ang = boundary_nodes(-np.deg2rad(max_deg), np.deg2rad(max_deg), n_angles)
    T, ortho = axis_tilt_geometry(ang, axis, ref=np.array([1.0, 0.0, 0.0]))
    p = cell_centres(-box, box, det)
    return T, ortho, [p, p.copy()]


def half_turn_series(n_angles, axis, det=256, box=1.0):
    """n_angles over [0, pi) about `axis`."""
    ang = np.arange(n_angles) * np.pi / n_angles
    T, ortho = axis_tilt_geometry(ang, axis)
    p = cell_centres(-box, box, det)
    return T, ortho, [p, p.copy()]


def fan_2d(n_angles=180, det=128, box=1.0):
    """2D parallel beam, angles k pi / n_angles, detector cell centres on [-box, box]."""
    ang = np.arange(n_angles) * np.pi / n_angles
    return theta_to_vec([ang]), perp([ang]), [cell_centres(-box, box, det)]


def random_axis(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)
