"""Pins of the oracle's point-spread projector (oracle.psf_radproj,
psf_derivs, psf_volproj: pixel / voxel integrals of the pinned point
projector and render, point_spread_functions.py:168-1350).

  * the pixel integrals of a projection add up to the atom's mass,
    I (pi/30) / det U (3D, 2.5D) and I sqrt(pi/30) / det U (2D), as the
    reference's __total (:60-65); the voxel integrals likewise;
  * 2D isotropic atoms: the pixel integral of (I/rho) exp(-30 rho^2 t^2) in
    closed form with erf;
  * derivatives: central differences of the oracle's own value.
"""
import numpy as np
import pytest
from scipy.special import erf


def _U(r):
    r = np.asarray(r, float)
    if r.size == 3:
        return np.array([[r[0], r[2]], [0, r[1]]])
    return np.array([[r[0], r[3], r[5]], [0, r[1], r[4]], [0, 0, r[2]]])


def _geoms(oracle):
    phi = 0.3
    T = np.array([[-np.sin(phi), 0, np.cos(phi)]])
    w0 = np.array([[np.cos(phi), 0, np.sin(phi)]])
    w1 = np.array([[0, 1.0, 0]])
    p = np.linspace(-1, 1, 40)
    g3 = oracle.Geometry(T, (w0, w1), [p, p])
    th = np.array([0.7])
    g25 = oracle.Geometry(oracle.theta_to_vec([th]), oracle.perp([th]), [p, p])
    g2 = oracle.Geometry(oracle.theta_to_vec([th]), oracle.perp([th]), [np.linspace(-1, 1, 50)])
    return g3, g25, g2


def test_psf_mass_identity(oracle):
    g3, g25, g2 = _geoms(oracle)
    I = np.array([1.2])
    x3, r3 = np.array([[0.1, -0.2, 0.05]]), np.array([[6.0, 5.0, 7.0, 1.0, -0.5, 0.3]])
    for g in (g3, g25):
        P = oracle.psf_radproj(I, x3, r3, g, 16)
        assert abs(P.sum() - 1.2 * (np.pi / 30) / np.linalg.det(_U(r3[0]))) < 1e-10 * P.sum()
    x2, r2 = np.array([[0.1, -0.2]]), np.array([[6.0, 5.0, 1.5]])
    P = oracle.psf_radproj(I, x2, r2, g2, 16)
    assert abs(P.sum() - 1.2 * np.sqrt(np.pi / 30) / np.linalg.det(_U(r2[0]))) < 1e-10 * P.sum()
    # voxels: int I sqrt(30/pi) exp(-30|U y|^2) = I (pi/30)^(D/2) sqrt(30/pi) / det U
    y = np.linspace(-1, 1, 24)
    u = oracle.psf_volproj(I, x3, r3, [y, y, y], 12)
    assert abs(u.sum() - 1.2 * (np.pi / 30) / np.linalg.det(_U(r3[0]))) < 1e-9 * u.sum()
    u = oracle.psf_volproj(I, x2, r2, [y, y], 12)
    assert abs(u.sum() - 1.2 * np.sqrt(np.pi / 30) / np.linalg.det(_U(r2[0]))) < 1e-9 * u.sum()


def test_psf_2d_iso_closed_form(oracle):
    _, _, g2 = _geoms(oracle)
    I, x, rho = np.array([0.9]), np.array([[0.2, -0.3]]), 8.0
    P = oracle.psf_radproj(I, x, np.array([[rho]]), g2, 24)[0]
    c = np.asarray(g2.detector[0])
    s = (c[1] - c[0]) / 2
    mu = x[0] @ np.asarray(g2.ortho[0])[0]
    ex = (I[0] / rho) * np.sqrt(np.pi / 30) / rho / 2 * (erf(np.sqrt(30) * rho * (c + s - mu)) -
                                                          erf(np.sqrt(30) * rho * (c - s - mu)))
    assert np.abs(P - ex).max() < 1e-12 * ex.max()


@pytest.mark.parametrize("which", ["3d", "2p5d", "2d"])
@pytest.mark.parametrize("mode", ["linear", "L2"])
def test_psf_derivs_finite_differences(oracle, which, mode):
    g3, g25, g2 = _geoms(oracle)
    g = {"3d": g3, "2p5d": g25, "2d": g2}[which]
    rng = np.random.default_rng(11)
    if which == "2d":
        I, x, r = np.array([1.1]), np.array([[0.05, -0.1]]), np.array([[7.0, 6.0, 1.2]])
        shape = (1, 50)
    else:
        I, x, r = np.array([1.1]), np.array([[0.05, -0.1, 0.1]]), np.array([[7.0, 6.0, 8.0, 1.0, -0.6, 0.4]])
        shape = (1, 40, 40)
    C = rng.normal(0, 1e-3, shape)
    th = np.concatenate((I, x[0], r[0]))
    D = x.shape[1]

    def val(t):
        P = oracle.psf_radproj(t[:1], t[1:1 + D][None], t[1 + D:][None], g, 12)
        return float(np.sum(P * C)) if mode == "linear" else 0.5 * float(np.sum((P - C) ** 2))

    f, Df, DDf = oracle.psf_derivs(I, x, r, g, C, mode, 2, 12)
    assert abs(f[0] - val(th)) < 1e-10 * abs(val(th))
    E = np.eye(th.size) * 1e-6
    fd = np.array([(val(th + e) - val(th - e)) / 2e-6 for e in E])
    assert np.abs(Df[0] - fd).max() < 1e-5 * np.abs(fd).max()

    def grad(t):
        return oracle.psf_derivs(t[:1], t[1:1 + D][None], t[1 + D:][None], g, C, mode, 1, 12)[1][0]

    E = np.eye(th.size) * 1e-5
    Hfd = np.array([(grad(th + e) - grad(th - e)) / 2e-5 for e in E])
    assert np.abs(DDf[0] - Hfd).max() < 1e-4 * np.abs(Hfd).max()


@pytest.mark.parametrize("block", [False, True])
def test_psf_newton_fit_restatement(oracle, block):
    """newton_fit with the pixel-integrated projector: E never increases, the
    last E is the fidelity of the final state, and a noise-free one-atom
    problem is recovered from a perturbed start."""
    _, _, g2 = _geoms(oracle)
    th = np.array([0.1, 0.9, 1.7, 2.5])
    p = np.linspace(-1, 1, 30)
    g = oracle.Geometry(oracle.theta_to_vec([th]), oracle.perp([th]), [p])
    # an atom about three pixels wide (std 0.1 on pixels of 0.069)
    I, x, r = np.array([1.0]), np.array([[0.1, -0.2]]), np.array([[1.3, 1.1, 0.2]])
    data = oracle.psf_radproj(I, x, r, g, 12)
    Io, xo, ro, E, F = oracle.newton_fit(I * 0.8, x + 0.03, r * 1.1, g, data, [30, 1], tol=1e-10, block=block,
                                         projector="psf", n_sub=12)
    assert np.all(np.diff(E) <= 0)
    assert abs(E[-1] - oracle.l2_squared_loss(oracle.psf_radproj(Io, xo, ro, g, 12), data)) < 1e-12 * E[0]
    if block:  # block coordinate steps converge linearly
        assert E[-1] < 1e-5 * E[0] and np.abs(xo - x).max() < 1e-4 and abs(Io[0] - 1.0) < 1e-2
    else:
        assert E[-1] < 1e-20 * E[0] and np.abs(xo - x).max() < 1e-8 and np.abs(ro - r).max() < 1e-8
