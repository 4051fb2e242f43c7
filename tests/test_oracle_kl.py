"""Pins of the oracle's KL moment initialiser (oracle.kl_sino_moments,
kl_orthog_op, kl_newton_sino, kl_step; KL_GaussRadon.py:57-394) and of the
greedy (guess) path of oracle.newton_fit (NewtonGaussian.py:38-50).

  * kl_newton_sino solves f + t log f = a + t log b (the residual of the
    defining equation), with the special cases t = 0 and b <= 1e-15;
  * kl_orthog_op returns M exactly when every angle's first moment is the
    projection of M onto its detector plane (3D, 2.5D, 2D);
  * kl_sino_moments against per-pixel sums over an explicit coordinate mesh;
  * kl_step recovers an isotropic atom (I, x, radius) from its own
    noise-free projection: each angle's projection of the atom is centred
    on the projection of x with spread 1/(60 r^2) per detector axis, which
    is what moment matching inverts;
  * the greedy Newton path: E never increases inside a stage, its last
    entry is the energy of the final atoms, and two atoms are found from a
    noise-free sinogram.
"""
import numpy as np
import pytest


def _g2(oracle, nA=12, n0=96):
    th = np.arange(nA) * np.pi / nA
    return oracle.Geometry(oracle.theta_to_vec([th]), oracle.perp([th]), [np.linspace(-1, 1, n0)])


def _g25(oracle, nA=10, n=64):
    th = np.arange(nA) * np.pi / nA
    p = np.linspace(-1, 1, n)
    return oracle.Geometry(oracle.theta_to_vec([th]), oracle.perp([th]), [p, p])


def _g3(oracle, nA=9, n=64):
    th0 = np.linspace(0, np.pi, nA, endpoint=False)
    th1 = np.linspace(-0.6, 0.6, nA)
    T = np.stack([-np.sin(th0) * np.cos(th1), np.cos(th0) * np.cos(th1), np.sin(th1)], 1)
    w0 = np.stack([np.cos(th0), np.sin(th0), 0 * th0], 1)
    w1 = np.stack([np.sin(th0) * np.sin(th1), -np.cos(th0) * np.sin(th1), np.cos(th1)], 1)
    p = np.linspace(-1, 1, n)
    return oracle.Geometry(T, (w0, w1), [p, p])


def test_newton_sino_solves_its_equation(oracle):
    rng = np.random.default_rng(0)
    a = rng.normal(0, 2, 500)
    b = np.exp(rng.normal(0, 3, 500))
    for t in (1e-3, 0.1, 1.0, 7.0):
        f = oracle.kl_newton_sino(a, b, t)
        rhs = a + t * np.log(b)
        ok = f > 1e-290
        lhs = f[ok] + t * np.log(f[ok])
        assert np.all(np.abs(lhs - rhs[ok]) < 1e-12 * np.maximum(1, np.abs(rhs[ok])))
        # the rest: roots at the bottom of the FP64 range (log f < -660)
        assert np.all(rhs[~ok] / t < -660)
        assert ok.mean() > 0.5
    assert np.array_equal(oracle.kl_newton_sino(a, b, 0.0), a)
    b[:10] = 1e-16
    assert np.all(oracle.kl_newton_sino(a, b, 0.5)[:10] == 0)


@pytest.mark.parametrize("which", ["2d", "2p5d", "3d"])
def test_orthog_op_recovers_the_centre(oracle, which):
    rng = np.random.default_rng(1)
    g = {"2d": _g2, "2p5d": _g25, "3d": _g3}[which](oracle)
    T = np.asarray(g.orientations)
    nA = T.shape[0]
    d = 2 if which == "2d" else 3
    M = rng.normal(0, 0.3, d)
    If = rng.uniform(0.5, 2.0, nA)
    Ify = []
    for a in range(nA):
        P = np.eye(d)
        P[:T.shape[1], :T.shape[1]] -= np.outer(T[a], T[a])
        Ify.append(If[a] * (P @ M))
    op, vec = oracle.kl_orthog_op(If, np.array(Ify), T)
    assert np.allclose(np.linalg.solve(op, vec), M, rtol=0, atol=1e-12)


def test_sino_moments_against_mesh_sums(oracle):
    rng = np.random.default_rng(2)
    g = _g3(oracle, 3, 20)
    f = rng.uniform(0, 1, (3, 20, 20))
    m = oracle.kl_sino_moments(f, g)
    P0, P1 = np.meshgrid(g.detector[0], g.detector[1], indexing="ij")
    for a in range(3):
        ref = [f[a].sum(), (f[a] * P0).sum(), (f[a] * P1).sum(), (f[a] * P0 ** 2).sum(), (f[a] * P1 ** 2).sum()]
        assert np.allclose(m[a], ref, rtol=1e-13, atol=0)
    g2 = _g2(oracle, 4, 30)
    f2 = rng.uniform(0, 1, (4, 30))
    m2 = oracle.kl_sino_moments(f2, g2)
    assert np.allclose(m2[:, 2], 0) and np.allclose(m2[:, 4], 0)
    assert np.allclose(m2[:, 1], f2 @ g2.detector[0], rtol=1e-13)


@pytest.mark.parametrize("which,lagrange", [("2d", False), ("2p5d", False), ("3d", False), ("2d", True),
                                            ("3d", True)])
def test_kl_step_recovers_an_atom_from_its_projection(oracle, which, lagrange):
    g = {"2d": _g2, "2p5d": _g25, "3d": _g3}[which](oracle)
    D = g.dim
    x = np.array([[0.12, -0.2, 0.05][:D]])
    rho = 1.0 / (0.08 * np.sqrt(60))  # spread 0.08: several pixels, well inside the detector
    r = np.array([[rho]])
    I = np.array([1.7])
    S0 = oracle.radproj(I, x, r, g)
    Ik, xk, rk = oracle.kl_step(S0, np.array([0.3]), np.zeros((1, D)), np.array([[3.0]]), g, 1.0, lagrange=lagrange)
    assert abs(Ik[0] - 1.7) < 1e-6 * 1.7
    assert np.abs(xk - x).max() < 1e-8
    assert abs(rk[0, 0] - rho) < 1e-6 * rho
    # anisotropic layout: the same atom, diagonal slots set, others zero
    nr = D * (D + 1) // 2
    Ik, xk, rk = oracle.kl_step(S0, np.array([0.3]), np.zeros((1, D)), np.ones((1, nr)), g, 1.0, lagrange=lagrange)
    assert np.abs(rk[0, :D] - rho).max() < 1e-6 * rho and np.all(rk[0, D:] == 0)


def test_kl_step_t_below_one_still_converges(oracle):
    """t < 1: f mixes the residual with the current projection; with a
    noise-free atom the fixed point is the atom."""
    g = _g2(oracle)
    x, r, I = np.array([[0.1, 0.3]]), np.array([[1.0 / (0.06 * np.sqrt(60))]]), np.array([0.8])
    S0 = oracle.radproj(I, x, r, g)
    for lag in (False, True):
        Ik, xk, rk = oracle.kl_step(S0, I, x, r, g, 0.5, lagrange=lag)
        assert abs(Ik[0] - 0.8) < 1e-4 and np.abs(xk - x).max() < 1e-6 and abs(rk[0, 0] / r[0, 0] - 1) < 1e-4


def test_kl_step_without_positive_mass_returns_a_narrow_atom(oracle):
    g = _g2(oracle, 4, 20)
    I, x, r = oracle.kl_step(-np.ones((4, 20)), np.array([1.0]), np.zeros((1, 2)), np.array([[5.0]]), g, 1.0)
    assert 1.0 / (r[0, 0] * np.sqrt(60)) < 1e-4 and I[0] == 1.0


def test_greedy_newton_with_kl_guess(oracle):
    g = _g2(oracle, 16, 96)
    D = 2
    It = np.array([1.0, 0.7])
    xt = np.array([[-0.4, 0.2], [0.35, -0.25]])
    rt = np.array([[1 / (0.06 * np.sqrt(60))], [1 / (0.09 * np.sqrt(60))]])
    data = oracle.radproj(It, xt, rt, g)

    def guess(res, I1, x1, r1):
        return oracle.kl_step(res + oracle.radproj(I1, x1, r1, g), I1, x1, r1, g, 1.0)

    I0, x0, r0 = np.ones(2), np.zeros((2, D)), np.full((2, 1), 5.0)
    I, x, r, E, F = oracle.newton_fit(I0, x0, r0, g, data, [10, 1, 5], tol=1e-8, guess=guess)
    assert E.size == F.size
    assert abs(E[-1] - oracle.l2_squared_loss(oracle.radproj(I, x, r, g), data)) < 1e-12 * E[0]
    assert E[-1] < 1e-6 * E[0]
    order = np.argsort(x[:, 0])
    assert np.abs(x[order] - xt).max() < 1e-3
    # without a guess the first stage's start energy is that of all atoms
    _, _, _, E2, _ = oracle.newton_fit(I0, x0, r0, g, data, [2, 1])
    assert abs(E2[0] - oracle.l2_squared_loss(oracle.radproj(I0, x0, r0, g), data)) < 1e-12 * E2[0]
