"""Pins of the oracle's sampled point-spread rule (oracle.psf_radproj_sampled,
psf_derivs_sampled, psf_volproj_sampled: the reference's own pixel and voxel
sampling, GaussDictCode/bin/point_spread_functions.py:25-88, 201-291,
389-527, 924-1075).

  * the sample offsets and weights of one axis against values worked out by
    hand from the reference's loop (:440-458: Y starts at the lower limit and
    steps by a tenth of the interval before each of 11 samples; the result is
    the sample mean times the interval);
  * pixels and voxels whose interval is not cut by R: the pinned point
    projector / render evaluated on each cell's 11 (x 11 (x 11)) sample
    points, weighted by (2 s / 11) per axis, with s half the first detector
    axis' spacing (projections) or the second volume axis' spacing (renders)
    on every axis -- also for cells of another size;
  * pixels and voxels that hold the atom's whole radius: the atom's total,
    which equals the sum of the pinned pixel / voxel integrals; every other
    cell of such an atom is zero, and its L2 gradient is zero;
  * derivatives: central differences of the oracle's own sampled value for
    atoms wide enough that the cut cells carry less than 1e-12 of the mass
    (there the samples do not move with the atom).
"""
import numpy as np
import pytest

from conftest import relerr


def _U(r):
    r = np.asarray(r, float)
    if r.size == 3:
        return np.array([[r[0], r[2]], [0, r[1]]])
    return np.array([[r[0], r[3], r[5]], [0, r[1], r[4]], [0, 0, r[2]]])


def _geoms(oracle, n=24, d1=None):
    phi = 0.3
    T = np.array([[-np.sin(phi), 0, np.cos(phi)], [0.0, 1.0, 0.0]])
    w0 = np.array([[np.cos(phi), 0, np.sin(phi)], [1.0, 0.0, 0.0]])
    w1 = np.array([[0, 1.0, 0], [0.0, 0.0, 1.0]])
    p = np.linspace(-0.6, 0.6, n)
    q = p if d1 is None else np.arange(-n // 2, n // 2) * d1
    g3 = oracle.Geometry(T, (w0, w1), [p, q])
    th = np.array([0.7, 2.1])
    g25 = oracle.Geometry(oracle.theta_to_vec([th]), oracle.perp([th]), [p, q])
    g2 = oracle.Geometry(oracle.theta_to_vec([th]), oracle.perp([th]), [np.linspace(-0.6, 0.6, 2 * n)])
    return {"3d": g3, "2p5d": g25, "2d": g2}


def _atom(which, wide=True):
    if which == "2d":
        r = np.array([[3.0, 2.5, 0.6]]) if wide else np.array([[150.0, 140.0, 8.0]])
        return np.array([1.3]), np.array([[0.04, -0.07]]), r
    r = np.array([[3.0, 2.5, 3.5, 0.5, -0.4, 0.3]]) if wide else np.array([[60.0, 55.0, 70.0, 4.0, -3.0, 2.0]])
    return np.array([1.3]), np.array([[0.04, -0.07, 0.05]]), r


def test_rule_axis_known_values(oracle):
    # c = <x - y, W> = 0.3, R = 0.5, s = 0.25: lo = max(-0.2, -0.25) = -0.2,
    # hi = min(0.8, 0.25) = 0.25, step 0.045, samples -0.155, -0.11, ..., 0.295
    t, w = oracle._psf_rule_axis(np.array([0.3]), 0.5, 0.25)
    assert np.allclose(t[0], -0.2 + 0.045 * np.arange(1, 12), rtol=0, atol=1e-15)
    assert abs(w[0] - 0.45 / 11) < 1e-16
    # an interval inside the cell: c = 0.01, R = 0.1, s = 0.25 -> [-0.09, 0.11]
    t, w = oracle._psf_rule_axis(np.array([0.01]), 0.1, 0.25)
    assert abs(t[0, 0] - (-0.09 + 0.02)) < 1e-15 and abs(t[0, -1] - (0.11 + 0.02)) < 1e-15
    assert abs(w[0] - 0.2 / 11) < 1e-16


def _uncut_cells(X, R, s):
    return np.all(np.abs(X) + s <= R, axis=-1)


@pytest.mark.parametrize("which", ["3d", "2p5d", "2d"])
@pytest.mark.parametrize("square", [True, False])
def test_uncut_pixels_are_point_samples(oracle, which, square):
    g = _geoms(oracle, d1=None if square else 0.037)[which]
    I, x, r = _atom(which)
    P = oracle.psf_radproj_sampled(I, x, r, g)
    D = g.dim
    R = 1 / np.min(np.diag(_U(r[0])))
    p0 = g.detector[0]
    s = (p0[1] - p0[0]) / 2          # the first axis' spacing, for both axes
    frames = oracle._psf_frames(g)
    checked = 0
    rng = np.random.default_rng(3)
    for a, (T, W0, W1) in enumerate(frames):
        if D == 2:
            ks = [(k,) for k in range(p0.size)]
        else:
            p1 = g.detector[1]
            ks = [(k0, k1) for k0 in range(p0.size) for k1 in range(p1.size)]
        for k in ks:
            y = p0[k[0]] * W0 + (g.detector[1][k[1]] * W1 if D == 3 else 0)
            X = x[0] - y
            off = [X @ W0] + ([X @ W1] if D == 3 else [])
            if not _uncut_cells(np.array(off), R, s) or rng.random() > 0.25:
                continue
            # the cell's samples: detector coordinates p - s + (2 s / 10) j, j = 1..11
            nodes = [p0[k[0]] - s + 0.2 * s * np.arange(1, 12)]
            if D == 3:
                nodes.append(g.detector[1][k[1]] - s + 0.2 * s * np.arange(1, 12))
            sub = oracle.Geometry(g.orientations[a:a + 1], tuple(o[a:a + 1] for o in g.ortho), nodes)
            v = oracle.radproj(I, x, r, sub)[0]
            ref = v.sum() * (2 * s / 11) ** len(g.detector)
            # the pinned projector is cut where m > 1 (exp(-30) of the peak), the rule is not
            assert abs(P[(a,) + k] - ref) <= 1e-12 * abs(ref) + 1e-12 * np.abs(P).max()
            checked += 1
    assert checked >= 10


def test_rule_shift_against_pixel_integral(oracle):
    # the samples' centroid is 0.6 of the interval from its lower end, so for a
    # profile nearly linear across a pixel the sampled value is the integral
    # over the pixel shifted by a tenth of its width (2D, wide iso atom)
    g = _geoms(oracle, n=40)["2d"]
    I, x, r = np.array([1.0]), np.array([[0.02, -0.01]]), np.array([[1.2]])
    P = oracle.psf_radproj_sampled(I, x, r, g)
    p0 = g.detector[0]
    s = (p0[1] - p0[0]) / 2
    sh = oracle.Geometry(g.orientations, g.ortho, [p0 + 0.2 * s])
    E = oracle.psf_radproj(I, x, r, sh, 24)
    E0 = oracle.psf_radproj(I, x, r, g, 24)
    R = 1 / 1.2
    c = np.abs(np.asarray(x[0]) @ np.asarray(g.ortho[0]).T)[:, None] - p0[None, :]
    inner = np.abs(c) + s < 0.5 * R
    assert inner.sum() > 20
    assert relerr(P[inner], E[inner]) < 2e-4
    assert relerr(P[inner], E0[inner]) > 2e-3       # not the unshifted integral


@pytest.mark.parametrize("which", ["3d", "2p5d", "2d"])
def test_total_pixel(oracle, which):
    g = _geoms(oracle)[which]
    I, _, r = _atom(which, wide=False)
    D = g.dim
    p0 = g.detector[0]
    # centre the atom on pixel (5, 7) of angle 0 (and a little off it)
    T, W0, W1 = oracle._psf_frames(g)[0]
    y = p0[5] * W0 + (g.detector[1][7] * W1 if D == 3 else 0)
    x = (y + 0.002 * W0)[None, :]
    P = oracle.psf_radproj_sampled(I, x, r, g)
    mass = oracle.psf_radproj(I, x, r, g, 64)[0].sum()   # fine nodes: the atom is narrower than a pixel
    k = (0, 5, 7) if D == 3 else (0, 5)
    assert abs(P[k] - mass) < 1e-9 * mass
    other = P[0].copy()
    other[k[1:]] = 0
    assert np.all(other == 0)
    C = np.random.default_rng(1).normal(0, 1e-3, P.shape)
    f, Df, DDf = oracle.psf_derivs_sampled(I, x, r, g, C, "L2", 2)
    # angle 1 may sample the atom; angle 0 adds (total - C)^2 / 2 and no gradient
    only1 = oracle.Geometry(g.orientations[1:], tuple(o[1:] for o in g.ortho), g.detector)
    f1, Df1, DDf1 = oracle.psf_derivs_sampled(I, x, r, only1, C[1:], "L2", 2)
    d0 = P[0] - C[0]
    assert abs(f[0] - (f1[0] + 0.5 * (d0 * d0).sum())) < 1e-12 * abs(f[0])
    assert np.allclose(Df, Df1, rtol=1e-12, atol=0) and np.allclose(DDf, DDf1, rtol=1e-12, atol=0)
    # angle 0 alone: no sampled pixel, so no gradient or Hessian at all
    only0 = oracle.Geometry(g.orientations[:1], tuple(o[:1] for o in g.ortho), g.detector)
    f0, Df0, DDf0 = oracle.psf_derivs_sampled(I, x, r, only0, C[:1], "L2", 2)
    assert abs(f0[0] - 0.5 * (d0 * d0).sum()) < 1e-12 * f0[0]
    assert np.all(Df0 == 0) and np.all(DDf0 == 0)


@pytest.mark.parametrize("which", ["3d", "2p5d", "2d"])
@pytest.mark.parametrize("mode", ["linear", "L2"])
def test_sampled_derivs_finite_differences(oracle, which, mode):
    g = _geoms(oracle, n=20)[which]
    I, x, r = _atom(which)
    P = oracle.psf_radproj_sampled(I, x, r, g)
    C = P * 0.7 + np.random.default_rng(2).normal(0, 1e-3, P.shape)
    f, Df, DDf = oracle.psf_derivs_sampled(I, x, r, g, C, mode, 2)
    D = g.dim
    th = np.concatenate((I, x[0], r[0]))

    def val(t):
        Pt = oracle.psf_radproj_sampled(t[:1], t[1:1 + D][None], t[1 + D:][None], g)
        return 0.5 * ((Pt - C) ** 2).sum() if mode == "L2" else (Pt * C).sum()

    def grad(t):
        return oracle.psf_derivs_sampled(t[:1], t[1:1 + D][None], t[1 + D:][None], g, C, mode, 1)[1][0]

    assert abs(f[0] - val(th)) < 1e-13 * abs(f[0])
    gfd = np.zeros(th.size)
    Hfd = np.zeros((th.size, th.size))
    for i in range(th.size):
        e = np.zeros(th.size)
        e[i] = 1e-5 * max(1.0, abs(th[i]))
        gfd[i] = (val(th + e) - val(th - e)) / (2 * e[i])
        Hfd[i] = (grad(th + e) - grad(th - e)) / (2 * e[i])
    assert relerr(Df[0], gfd) < 1e-6
    assert relerr(DDf[0], Hfd) < 1e-5
    assert np.allclose(DDf[0], DDf[0].T, rtol=0, atol=1e-12 * np.abs(DDf[0]).max())


def test_sampled_iso_chain_rule(oracle):
    g = _geoms(oracle, n=20)["3d"]
    I, x = np.array([0.8]), np.array([[0.03, 0.02, -0.04]])
    rho = 2.7
    C = np.random.default_rng(4).normal(0, 1e-2, g.shape)
    f, Df, DDf = oracle.psf_derivs_sampled(I, x, np.array([[rho]]), g, C, "L2", 2)
    fa, Dfa, DDfa = oracle.psf_derivs_sampled(I, x, np.array([[rho, rho, rho, 0, 0, 0]]), g, C, "L2", 2)
    E = oracle.iso_collapse(3)
    assert abs(f[0] - fa[0]) < 1e-14 * abs(f[0])
    assert relerr(Df[0], Dfa[0] @ E) < 1e-14
    assert relerr(DDf[0], E.T @ DDfa[0] @ E) < 1e-14


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("square", [True, False])
def test_uncut_voxels_are_point_samples(oracle, D, square):
    n = 14
    y = [np.linspace(-0.5, 0.5, n) for _ in range(D)]
    if not square:
        y[1] = np.arange(n) * 0.061 - 0.4     # s comes from this axis for every axis
    I = np.array([0.9])
    x = np.array([[0.03, -0.05, 0.02][:D]])
    r = np.array([[2.4, 2.0, 0.5]]) if D == 2 else np.array([[2.4, 2.0, 2.2, 0.4, -0.3, 0.2]])
    u = oracle.psf_volproj_sampled(I, x, r, y)
    R = 1 / np.min(np.diag(_U(r[0])))
    s = (y[1][1] - y[1][0]) / 2
    rng = np.random.default_rng(5)
    checked = 0
    for idx in np.ndindex(*u.shape):
        yv = np.array([y[i][idx[i]] for i in range(D)])
        if not _uncut_cells(x[0] - yv, R, s) or rng.random() > (0.3 if D == 2 else 0.05):
            continue
        nodes = [yv[i] - s + 0.2 * s * np.arange(1, 12) for i in range(D)]
        ref = oracle.volproj(I, x, r, nodes).sum() * (2 * s / 11) ** D
        assert abs(u[idx] - ref) <= 1e-12 * abs(ref) + 1e-12 * np.abs(u).max()
        checked += 1
    assert checked >= 8


@pytest.mark.parametrize("D", [2, 3])
def test_total_voxel(oracle, D):
    n = 12
    y = [np.linspace(-0.5, 0.5, n) for _ in range(D)]
    I = np.array([1.1])
    x = np.array([[y[0][4] + 0.003, y[1][6] - 0.002, y[-1][5]][:D]])
    r = np.array([[50.0, 45.0, 3.0]]) if D == 2 else np.array([[50.0, 45.0, 55.0, 3.0, -2.0, 1.0]])
    u = oracle.psf_volproj_sampled(I, x, r, y)
    # the atom's mass (the sum of its voxel integrals, pinned in test_oracle_psf.py for wide atoms)
    mass = I[0] * (np.pi / 30 if D == 3 else np.sqrt(np.pi / 30)) / np.linalg.det(_U(r[0]))
    idx = (4, 6, 5)[:D]
    assert abs(u[idx] - mass) < 1e-9 * mass
    u[idx] = 0
    assert np.all(u == 0)
