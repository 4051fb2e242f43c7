"""Pins of the FP64 oracle against mathematics that fixes its results.

The reference ships no golden vectors or tests (its __main__ self-checks are
stale, NewtonBases.py:1155-1229), so each oracle function is checked here
against something independent of its own formulas:
  * closed forms (centre-ray value, detector mass, volume peak and mass),
  * brute-force 1D quadrature of the volume profile along a ray,
  * sympy differentiation of I exp(-30 m) / |U T| (the reference's own check,
    corrected from exp(-n/2) to the live exp(-30 n)),
  * central finite differences of the oracle's own forward projection,
  * identities (Gauss-Newton split, adjoint, iso chain rule, symmetries),
  * the FT of a Gaussian sampled on a fine grid, and the parameter round trip.
"""
import numpy as np
import pytest
from scipy import integrate

from conftest import relerr


def _U(r):
    r = np.asarray(r, float)
    if r.size == 3:
        return np.array([[r[0], r[2]], [0, r[1]]])
    return np.array([[r[0], r[3], r[5]], [0, r[1], r[4]], [0, 0, r[2]]])


def _one_atom(seed=0):
    rng = np.random.default_rng(seed)
    I = np.array([rng.uniform(0.5, 1.5)])
    x = rng.uniform(-0.3, 0.3, (1, 3))
    r = np.array([[rng.uniform(3, 6), rng.uniform(3, 6), rng.uniform(3, 6), rng.normal(0, 1),
                   rng.normal(0, 1), rng.normal(0, 1)]])
    return I, x, r


def _tilt(phi):
    T = np.array([[-np.sin(phi), 0.0, np.cos(phi)]])
    w0 = np.array([[np.cos(phi), 0.0, np.sin(phi)]])
    w1 = np.array([[0.0, 1.0, 0.0]])
    return T, w0, w1


def test_geometry_orthonormal(oracle):
    th = np.linspace(0, np.pi, 7)
    T, (w0,) = oracle.theta_to_vec([th]), oracle.perp([th])
    assert np.allclose((T * w0).sum(1), 0) and np.allclose((T * T).sum(1), 1) and np.allclose((w0 * w0).sum(1), 1)
    t0, t1 = np.linspace(0, np.pi, 4), np.linspace(-1, 1, 3)
    T3 = oracle.theta_to_vec([t0, t1]).reshape(-1, 3)
    a, b = [w.reshape(-1, 3) for w in oracle.perp([t0, t1])]
    for u, v in ((T3, a), (T3, b), (a, b)):
        assert np.allclose((u * v).sum(1), 0, atol=1e-14)
    for u in (T3, a, b):
        assert np.allclose((u * u).sum(1), 1)
    # explicit trig form of atomFuncs.py:13-47
    assert np.allclose(T3[0], [-np.sin(t0[0]) * np.cos(t1[0]), np.cos(t0[0]) * np.cos(t1[0]), np.sin(t1[0])])


def test_centre_ray_closed_form(oracle):
    I, x, r = _one_atom(1)
    T, w0, w1 = _tilt(0.4)
    g = oracle.Geometry(T, (w0, w1), [np.array([w0[0] @ x[0]]), np.array([w1[0] @ x[0]])])
    R = oracle.radproj(I, x, r, g)[0, 0, 0]
    assert abs(R - I[0] / np.linalg.norm(_U(r[0]) @ T[0])) < 1e-14


def test_detector_mass_3d_and_2d(oracle):
    I, x, r = _one_atom(2)
    x[:] = 0.05
    T, w0, w1 = _tilt(-0.7)
    p = (np.arange(600) + 0.5) / 300 - 1
    R = oracle.radproj(I, x, r, oracle.Geometry(T, (w0, w1), [p, p]))
    mass = R.sum() * (2 / 600) ** 2
    assert abs(mass - I[0] * (np.pi / 30) / np.prod(r[0, :3])) < 1e-9 * mass
    # 2D: integral of R over the detector line = I sqrt(pi/30) / det U
    r2 = np.array([[4.0, 6.0, 1.5]])
    th = np.array([0.3])
    g2 = oracle.Geometry(oracle.theta_to_vec([th]), oracle.perp([th]), [(np.arange(4000) + 0.5) / 2000 - 1])
    R2 = oracle.radproj(np.array([1.3]), np.array([[0.1, -0.2]]), r2, g2)
    assert abs(R2.sum() * (2 / 4000) - 1.3 * np.sqrt(np.pi / 30) / 24.0) < 1e-10


def test_forward_equals_quadrature_of_volume_profile(oracle):
    """R(y) = int f(y + s T) ds with f = I sqrt(30/pi) exp(-30 |U(y-x)|^2) (no projection algebra)."""
    I, x, r = _one_atom(3)
    U = _U(r[0])
    T, w0, w1 = _tilt(0.9)
    rng = np.random.default_rng(0)
    p = rng.uniform(-0.25, 0.25, (6, 2)) + [w0[0] @ x[0], w1[0] @ x[0]]
    for p0, p1 in p:
        g = oracle.Geometry(T, (w0, w1), [np.array([p0]), np.array([p1])])
        R = oracle.radproj(I, x, r, g)[0, 0, 0]
        y = p0 * w0[0] + p1 * w1[0]
        f = lambda s: I[0] * np.sqrt(30 / np.pi) * np.exp(-30 * np.sum((U @ (y + s * T[0] - x[0])) ** 2))
        q, _ = integrate.quad(f, -3, 3, epsabs=1e-14, epsrel=1e-12, limit=200)
        assert abs(R - q) <= 1e-9 * max(abs(q), 1e-12) + 1e-13


def test_invariants(oracle):
    rng = np.random.default_rng(4)
    I, x, r = np.array([0.7, 1.2]), rng.uniform(-0.3, 0.3, (2, 3)), np.array([[4, 5, 6, 1, -1, .5], [6, 4, 5, 0, 1, -1.]])
    ang = np.linspace(0, np.pi, 5, endpoint=False)
    p = np.linspace(-1, 1, 23)
    g = oracle.Geometry(oracle.theta_to_vec([ang]), oracle.perp([ang]), [p, p])
    R = oracle.radproj(I, x, r, g)
    # linear in I, additive over atoms
    assert relerr(oracle.radproj(2 * I, x, r, g), 2 * R) < 1e-14
    assert relerr(oracle.radproj(I[:1], x[:1], r[:1], g) + oracle.radproj(I[1:], x[1:], r[1:], g), R) < 1e-14
    # T -> -T leaves line integrals unchanged
    g2 = oracle.Geometry(-g.orientations, g.ortho, g.detector)
    assert relerr(oracle.radproj(I, x, r, g2), R) < 1e-13
    # translation along z (the 2.5D detector axis) shifts the sinogram by whole pixels
    dz = p[1] - p[0]
    Rs = oracle.radproj(I, x + [0, 0, 3 * dz], r, g)
    assert relerr(Rs[:, :, 3:], R[:, :, :-3]) < 1e-12


def _sympy_ray_derivs(I, x, r, y, T, C):
    import sympy as sp

    sI, sx, sr = sp.Symbol("I"), sp.symbols("x0:3"), sp.symbols("r0:6")
    U = sp.Matrix([[sr[0], sr[3], sr[5]], [0, sr[1], sr[4]], [0, 0, sr[2]]])
    Y = sp.Matrix([y[i] - sx[i] for i in range(3)])
    Tv = sp.Matrix(T)
    rY, rT = U * Y, U * Tv
    nT2 = (rT.T * rT)[0]
    m = (rY.T * rY)[0] - (rY.T * rT)[0] ** 2 / nT2
    Rexpr = sI * sp.exp(-30 * m) / sp.sqrt(nT2)
    E = (Rexpr - C) ** 2 / 2
    var = (sI,) + tuple(sx) + tuple(sr)
    subs = dict(zip(var, [I] + list(x) + list(r)))
    grad = [sp.diff(E, v) for v in var]
    g = np.array([float(d.evalf(subs=subs)) for d in grad])
    H = np.array([[float(sp.diff(d, v).evalf(subs=subs)) for v in var] for d in grad])
    return float(E.evalf(subs=subs)), g, H


def test_derivs_against_sympy(oracle):
    """The reference's own symbolic check, with the live profile exp(-30 m)."""
    I, x, r = _one_atom(5)
    T, w0, w1 = _tilt(0.3)
    c = np.array([w0[0] @ x[0], w1[0] @ x[0]]) + [0.03, -0.02]
    g = oracle.Geometry(T, (w0, w1), [c[:1], c[1:]])
    C = np.array([[[0.05]]])
    f, gr, H = oracle.derivs(I, x, r, g, C, "L2", 2)
    y = c[0] * w0[0] + c[1] * w1[0]
    fs, gs, Hs = _sympy_ray_derivs(I[0], x[0], r[0], y, T[0], 0.05)
    assert abs(f[0] - fs) < 1e-12 * abs(fs)
    assert relerr(gr[0], gs) < 1e-10
    assert relerr(H[0], Hs) < 1e-10


def _fd_geometry():
    ang = np.linspace(-1.0, 1.0, 5)
    T = np.stack([-np.sin(ang), 0 * ang, np.cos(ang)], 1)
    w0 = np.stack([np.cos(ang), 0 * ang, np.sin(ang)], 1)
    w1 = np.tile([0.0, 1.0, 0.0], (5, 1))
    p = np.linspace(-0.9, 0.9, 31)
    return T, w0, w1, p


def _theta(I, x, r):
    return np.concatenate([I, x[0], r[0]])


def _split(th, D=3):
    return th[:1], th[1:1 + D][None], th[1 + D:][None]


@pytest.mark.parametrize("mode", ["L2", "linear"])
def test_derivs_finite_differences(oracle, mode):
    I, x, r = _one_atom(6)
    T, w0, w1, p = _fd_geometry()
    g = oracle.Geometry(T, (w0, w1), [p, p])
    rng = np.random.default_rng(1)
    C = oracle.radproj(I, x, r, g) * 0.6 + rng.normal(0, 0.01, (5, 31, 31))

    def energy(th):
        R = oracle.radproj(*_split(th), g)
        return 0.5 * np.sum((R - C) ** 2) if mode == "L2" else np.sum(R * C)

    def grad(th):
        return oracle.derivs(*_split(th), g, C, mode, 2)[1][0]

    th = _theta(I, x, r)
    f, gr, H = oracle.derivs(I, x, r, g, C, mode, 2)
    assert abs(f[0] - energy(th)) < 1e-12 * abs(f[0])
    h = 1e-6
    gfd = np.array([(energy(th + h * e) - energy(th - h * e)) / (2 * h) for e in np.eye(10)])
    assert relerr(gr[0], gfd) < 1e-7
    Hfd = np.array([(grad(th + h * e) - grad(th - h * e)) / (2 * h) for e in np.eye(10)])
    assert relerr(H[0], Hfd) < 1e-6
    assert np.allclose(H[0], H[0].T, rtol=0, atol=1e-12 * np.abs(H[0]).max())


def test_gauss_newton_and_adjoint_identities(oracle):
    I, x, r = _one_atom(7)
    T, w0, w1, p = _fd_geometry()
    g = oracle.Geometry(T, (w0, w1), [p, p])
    R = oracle.radproj(I, x, r, g)
    C = R * 0.3 + 0.02
    _, gL, HL = oracle.derivs(I, x, r, g, C, "L2", 2)
    _, gl, Hl = oracle.derivs(I, x, r, g, R - C, "linear", 2)
    _, _, GN = oracle.derivs(I, x, r, g, R, "L2", 2)
    assert relerr(gL, gl) < 1e-12
    assert relerr(HL, Hl + GN) < 1e-12
    # adjoint: <J d, c> = <d, J^T c>, J d by central differences of the forward
    rng = np.random.default_rng(3)
    d = rng.normal(size=10)
    c = rng.normal(size=R.shape)
    th = _theta(I, x, r)
    h = 1e-6
    Jd = (oracle.radproj(*_split(th + h * d), g) - oracle.radproj(*_split(th - h * d), g)) / (2 * h)
    Jtc = oracle.derivs(I, x, r, g, c, "linear", 1)[1][0]
    assert abs(np.sum(Jd * c) - d @ Jtc) < 1e-7 * abs(d @ Jtc)


def test_iso_chain_rule(oracle):
    """Isotropic atoms: derivative in rho equals the sum over the diagonal slots,
    and matches finite differences of the iso forward (2D and 3D)."""
    for D in (2, 3):
        if D == 2:
            th = np.array([0.2, 1.1])
            geom = oracle.Geometry(oracle.theta_to_vec([th]), oracle.perp([th]), [np.linspace(-1, 1, 41)])
            I, x, r = np.array([0.9]), np.array([[0.1, -0.1]]), np.array([[5.0]])
        else:
            T, w0, w1, p = _fd_geometry()
            geom = oracle.Geometry(T, (w0, w1), [p, p])
            I, x, r = np.array([0.9]), np.array([[0.1, -0.1, 0.05]]), np.array([[5.0]])
        Rr = oracle.radproj(I, x, r, geom)
        C = 0.5 * Rr + 0.01
        f, gi, Hi = oracle.derivs(I, x, r, geom, C, "L2", 2)
        en = lambda rho: 0.5 * np.sum((oracle.radproj(I, x, np.array([[rho]]), geom) - C) ** 2)
        h = 1e-6
        gfd = (en(5.0 + h) - en(5.0 - h)) / (2 * h)
        assert abs(gi[0, -1] - gfd) < 1e-7 * abs(gfd)
        hfd = (en(5.0 + h) - 2 * en(5.0) + en(5.0 - h)) / h ** 2
        assert abs(Hi[0, -1, -1] - hfd) < 1e-3 * abs(hfd)


def test_volume_closed_forms(oracle):
    I, x, r = _one_atom(8)
    x[:] = [[0.01, -0.02, 0.03]]
    u = oracle.volproj(I, x, r, [x[0, :1], x[0, 1:2], x[0, 2:]])
    assert abs(u[0, 0, 0] - I[0] * np.sqrt(30 / np.pi)) < 1e-14
    c = (np.arange(200) + 0.5) / 100 - 1
    u = oracle.volproj(I, x, r, [c, c, c])
    mass = u.sum() * 0.01 ** 3
    assert abs(mass - I[0] * (np.pi / 30) / np.prod(r[0, :3])) < 1e-8 * mass


def test_fourier_against_sampled_gaussian(oracle):
    """FT of I exp(-|U(y-x)|^2/2) by a fine Riemann sum vs the analytic map + evaluation."""
    I, x = np.array([1.3]), np.array([[0.1, -0.05, 0.07]])
    r = np.array([[9.0, 8.0, 10.0, 2.0, -1.0, 1.5]])
    U = _U(r[0])
    n, L = 96, 1.0
    c = (np.arange(n) + 0.5) / n * 2 * L - L
    Y = np.stack(np.meshgrid(c, c, c, indexing="ij"), -1) - x[0]
    f = I[0] * np.exp(-0.5 * np.sum((Y @ U.T) ** 2, -1))
    ks = np.array([[0, 0, 0], [1, 0, 0], [0, 2, -1], [1.5, -1, 2], [3, 1, 0]], float)
    yy = Y + x[0]
    dft = np.array([np.sum(f * np.exp(-2j * np.pi * (yy @ k))) * (2 * L / n) ** 3 for k in ks])
    nI, nx, nr = oracle.analytic_gauss_fourier(I, x, r)
    ana = oracle.ft_points(nI, nx, nr, ks)
    assert relerr(ana, dft) < 1e-9
    # round trip of the parameter maps (Fourier_Transform.py:455-469)
    bI, bx, br = oracle.analytic_gauss_ifourier(nI, nx, nr)
    assert relerr(bI, I) < 1e-14 and relerr(bx, x) < 1e-14 and relerr(br, r) < 1e-13


def test_fourier_cutoff_and_iso(oracle):
    nI, nx = np.array([2.0]), np.array([[0.3, 0.2, -0.1]])
    nr = np.array([[0.5, 0.6, 0.7, 0.1, 0.0, -0.1]])
    k = np.array([[20.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    v = oracle.ft_points(nI, nx, nr, k)
    assert v[0] == 0  # |U^T k|^2 = 100 >= 30
    assert abs(v[1] - 2.0) < 1e-15
    vi = oracle.ft_points(nI, nx, np.array([[4.0]]), np.array([[3.0, 4.0, 0.0]]), iso=True)
    assert abs(vi[0] - 2.0 * np.exp(-0.5 * 25 / 16) * np.exp(-1j * (0.9 + 0.8))) < 1e-14


def test_newton_energy_monotone_and_recovery(oracle):
    """Noise-free 2-atom 2D problem: energy never increases and the fit recovers the atoms."""
    th = np.linspace(0, np.pi, 30, endpoint=False)
    g = oracle.Geometry(oracle.theta_to_vec([th]), oracle.perp([th]), [(np.arange(96) + 0.5) / 48 - 1])
    I0, x0, r0 = np.array([1.0, 0.8]), np.array([[-0.4, 0.1], [0.35, -0.2]]), np.array([[3.0, 3.5, 0.5], [4.0, 3.0, -0.3]])
    data = oracle.radproj(I0, x0, r0, g)
    I, x, r, E = oracle.newton_linesearch(I0 * 0.8, x0 + 0.03, r0 * 1.1, g, data, [25, 1])
    assert np.all(np.diff(E) <= 1e-12 * E[0])
    assert E[-1] < 1e-10 * E[0]
    assert np.allclose(x, x0, atol=1e-5) and np.allclose(I, I0, atol=1e-4)


def test_gd_block_energy_monotone(oracle):
    th = np.linspace(0, np.pi, 20, endpoint=False)
    g = oracle.Geometry(oracle.theta_to_vec([th]), oracle.perp([th]), [(np.arange(64) + 0.5) / 32 - 1])
    I0, x0, r0 = np.array([1.0, 0.8]), np.array([[-0.4, 0.1], [0.35, -0.2]]), np.array([[3.0], [4.0]])
    data = oracle.radproj(I0, x0, r0, g)
    I, x, r, E = oracle.gd_linesearch_block(I0, x0 + 0.02, r0, g, data, [10, 1])
    assert np.all(np.diff(E) <= 0)
    assert E[-1] < 0.5 * E[0]
