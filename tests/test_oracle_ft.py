"""Pins of the oracle's Fourier derivatives on hyperplane samples
(oracle.ft_derivs, restating FTGaussian3D.py:199-652).

The value f is compared with an independent closed-form evaluation of
1/2 sum |g - C|^2, the gradient and Hessian with central differences of that
evaluation, for the 3D planes, the 2.5D planes and the isotropic form; the
sample points with the oracle's pinned point evaluator (ft_points).
"""
import numpy as np
import pytest


def _value(th, K, C, iso):
    """1/2 sum |g - C|^2 of one atom, g = I exp(-|L k|^2/2 - i x.k), cut above 60."""
    I, x = th[0], th[1:4]
    if iso:
        N = (K * K).sum(1) / th[4] ** 2
    else:
        L = np.array([[th[4], 0, 0], [th[7], th[5], 0], [th[9], th[8], th[6]]])
        N = ((K @ L.T) ** 2).sum(1)
    g = I * np.exp(-0.5 * N - 1j * (K @ x)) * (N <= 60)
    return 0.5 * np.sum(np.abs(g - C) ** 2)


def _case(rng, nA=3, m0=9, m1=7, form="3d"):
    k0, k1 = np.linspace(-4, 4, m0), np.linspace(-3.5, 3.5, m1)
    th = rng.uniform(0, np.pi, nA)
    if form == "3d":
        w0 = np.stack([np.cos(th), np.sin(th), 0 * th], 1)
        w1 = np.stack([-np.sin(th) * 0.6, np.cos(th) * 0.6, 0.8 + 0 * th], 1)
    else:
        w0, w1 = np.stack([np.cos(th), np.sin(th)], 1), None
    C = rng.normal(size=(nA, m0, m1)) + 1j * rng.normal(size=(nA, m0, m1))
    return k0, k1, w0, w1, C


@pytest.mark.parametrize("form,iso", [("3d", False), ("2p5d", False), ("3d", True), ("2p5d", True)])
def test_ft_derivs_finite_differences(oracle, form, iso):
    rng = np.random.default_rng(4)
    k0, k1, w0, w1, C = _case(rng, form=form)
    K = oracle.ft_plane_points(k0, k1, w0, w1).reshape(-1, 3)
    I = np.array([1.3])
    x = rng.normal(0, 0.3, (1, 3))
    r = np.array([[0.8]]) if iso else np.array([[0.9, 0.7, 1.1, 0.3, -0.2, 0.15]])
    f, Df, DDf = oracle.ft_derivs(I, x, r, k0, k1, w0, w1, C, iso=iso)
    th = np.concatenate((I, x[0], r[0]))
    P = th.size
    val = lambda t: _value(t, K, C.reshape(-1), iso)
    assert abs(f[0] - val(th)) < 1e-12 * val(th)
    h = 1e-5
    E = np.eye(P) * h
    g_fd = np.array([(val(th + e) - val(th - e)) / (2 * h) for e in E])
    assert np.abs(Df[0] - g_fd).max() < 1e-6 * np.abs(g_fd).max()
    H_fd = np.array([[(val(th + a + b) - val(th + a - b) - val(th - a + b) + val(th - a - b)) / (4 * h * h)
                      for b in E] for a in E])
    assert np.abs(DDf[0] - H_fd).max() < 1e-4 * np.abs(H_fd).max()
    assert np.allclose(DDf[0], DDf[0].T)


def test_plane_points_match_point_evaluator(oracle):
    """ft_plane_points gives the sample points FTGaussian3D's hyperplane kernels
    use: the oracle's pinned ft_hyperplane equals ft_points on them."""
    rng = np.random.default_rng(5)
    for form in ("3d", "2p5d"):
        k0, k1, w0, w1, _ = _case(rng, form=form)
        I, x = np.array([1.0, 0.7]), rng.normal(0, 0.3, (2, 3))
        r = np.array([[0.9, 0.7, 1.1, 0.3, -0.2, 0.15], [1.2, 0.8, 0.9, -0.1, 0.2, 0.0]])
        K = oracle.ft_plane_points(k0, k1, w0, w1)
        a = oracle.ft_points(I, x, r, K.reshape(-1, 3)).reshape(K.shape[:3])
        b = oracle.ft_hyperplane(I, x, r, k0, k1, w0, w1)
        assert np.allclose(a, b, rtol=0, atol=1e-14)


def test_newton_fit_on_fourier_planes(oracle):
    """newton_fit with the Fourier-plane projector (a GaussFTVolume fit):
    energy never increases, its last entry is the loss of the final atoms,
    and noise-free FT atoms are recovered from a perturbed start."""
    rng = np.random.default_rng(3)
    n = 2
    I = rng.uniform(0.5, 1.5, n)
    x = rng.uniform(-0.5, 0.5, (n, 3))
    r = np.concatenate((rng.uniform(3, 6, (n, 3)), rng.normal(0, 0.5, (n, 3))), 1)
    nI, nx, nr = oracle.analytic_gauss_fourier(I, x, r * np.sqrt(60))
    nI = nI * 2000  # samples of order 1, so the fidelity dominates the initial damping
    th = np.linspace(0, np.pi, 6, endpoint=False)
    w0 = np.stack([np.cos(th), np.sin(th), 0 * th], 1)
    w1 = np.tile([0, 0, 1.0], (6, 1))
    k = np.linspace(-40, 40, 25)
    g = (k, k, w0, w1)
    data = oracle.ft_hyperplane(nI, nx, nr, k, k, w0, w1)
    Io, xo, ro, E, F = oracle.newton_fit(nI * 0.8, nx + 0.02, nr * 1.1, g, data, [14, 1], tol=1e-12, projector="ft")
    assert np.all(np.diff(E) <= 0)
    assert abs(E[-1] - oracle.l2_squared_loss(oracle.ft_hyperplane(Io, xo, ro, k, k, w0, w1), data)) < 1e-12 * E[0]
    assert E[-1] < 1e-7 * E[0]
    assert np.abs(xo - nx).max() < 1e-4 and np.abs(Io / nI - 1).max() < 1e-3
