"""Pins of the oracle's regularisers and of its regularised / block Newton
restatement (GaussDictCode/regularisation.py:22-442, NewtonGaussian.py:18-237).

Values are checked at points where the reference's formulas reduce to
numbers fixed by their construction (Joubert's volume term vanishes with zero
slope at v = r0, which is what its constants d and e are chosen for; Mass
and Volume vanish at their targets; Radius at unit distance is s per slot),
derivatives against central differences of the value, and the domains
against the reference's tests.  The regularised fitter is checked against the
unregularised one (null regulariser), for energy monotonicity, and for what a
strong regulariser must do to the fit.
"""
import numpy as np
import pytest


def _th(I, x, r):
    return np.concatenate(([I], x, r)).astype(float)


def _fd(f, th, h=1e-6):
    g = np.zeros(th.size)
    for i in range(th.size):
        e = np.zeros(th.size)
        e[i] = h
        g[i] = (f(th + e) - f(th - e)) / (2 * h)
    return g


def test_joubert_volume_term_vanishes_at_r0(oracle):
    """c v - d log v - e has its minimum 0 at v = r0 (constants of
    regularisation.py:47-51): at I = 1, x = 0 the whole value is 0 there."""
    r0, r1 = 0.7, 2.5
    for D, iso in ((3, False), (2, False), (3, True), (2, True)):
        R = oracle.Reg("Joubert", D, 1.3, 2.0, (r0, r1))
        if iso:
            r = np.array([r0 ** (1.0 / D)])
        else:
            diag = [r0 / 1.2, 1.2] + ([1.0] if D == 3 else [])
            r = np.array(diag + [0.4, -0.7, 0.1][: D * (D + 1) // 2 - D])
        th = _th(1.0, np.zeros(D), r)
        assert abs(R.term(th, iso)) < 1e-12
        g, H = R.derivs(th, iso)
        # dv-slope zero: the r-gradient vanishes, the I-gradient is -a / I
        assert np.allclose(g[1 + D:], 0, atol=1e-12)
        assert np.isclose(g[0], -1.3)
    # value off the minimum: -a log 2 + b/2 |x|^2 with the volume term at r0
    R = oracle.Reg("Joubert", 3, 1.0, 2.0, (0.5, 3.0))
    th = _th(2.0, [0.1, -0.2, 0.2], [0.5, 1.0, 1.0, 0.3, -0.4, 0.2])
    assert np.isclose(R.term(th, False), -np.log(2.0) + 0.09)


def test_other_regulariser_values(oracle):
    # Mass: s (|I R| - m)^k with R = 1 / prod(diag); zero at I R = m
    M = oracle.Reg("Mass", 3, 1.5, 2, 3.0)
    assert np.isclose(M.term(_th(3.0, [0, 0, 0], [2.0, 1.0, 1.0, 0.5, 0.5, 0.5]), False), 0.0)
    assert np.isclose(M.term(_th(4.0, [0, 0, 0], [2.0, 1.0, 1.0, 0.5, 0.5, 0.5]), False), 3.0 * 0.25)
    assert np.isclose(M.term(_th(3.0, [0, 0, 0], [2.0]), True), 3.0 * (24 - 1.5) ** 2)  # iso: I r^D
    # Volume: s |prod(diag) - r0^D|^k; iso r^-D
    V = oracle.Reg("Volume", 3, 0.5, 2, 2.0)
    assert np.isclose(V.term(_th(1.0, [0, 0, 0], [0.25, 0.5, 1.0, 9, 9, 9]), False), 2.0 * (0.125 - 0.125) ** 2)
    assert np.isclose(V.term(_th(1.0, [0, 0, 0], [1.0, 1.0, 1.0, 9, 9, 9]), False), 2.0 * 0.875 ** 2)
    assert np.isclose(V.term(_th(1.0, [0, 0, 0], [2.0]), True), 0.0)
    # Radius: s sum (r_i - r0)^-k; iso D (r - r0)^-k
    Rr = oracle.Reg("Radius", 2, 0.5, 1.5, 3)
    assert np.isclose(Rr.term(_th(1.0, [0, 0], [1.5, 1.5, 7.0]), False), 1.5 * 2)
    assert np.isclose(Rr.term(_th(1.0, [0, 0], [2.5]), True), 1.5 * 2 / 8)
    # Iso: s/2 sum of squared off-diagonal slots
    S = oracle.Reg("Iso", 3, 4.0)
    assert np.isclose(S.term(_th(1.0, [0, 0, 0], [1, 1, 1, 0.3, -0.4, 0.2]), False), 2.0 * 0.29)
    assert S.term(_th(1.0, [0, 0, 0], [1.0]), True) == 0.0
    assert oracle.Reg("null", 3).term(_th(1.0, [0, 0, 0], [1.0]), True) == 0.0


def test_regulariser_domains(oracle):
    """+inf exactly where the reference returns infty (regularisation.py:60-64,
    123-124, 225-226, 334-335), Joubert on the diagonal slots only."""
    J = oracle.Reg("Joubert", 3, 1.0, 1.0, (0.5, 2.0))
    good = [1.0, 1.0, 1.0, -0.5, -0.5, -0.5]
    assert np.isfinite(J.term(_th(1.0, [0, 0, 0], good), False))  # negative off-diagonal slots are fine
    assert J.term(_th(0.0, [0, 0, 0], good), False) == np.inf
    assert J.term(_th(1.0, [0, 0, 0], [1.0, 0.0, 1.0, 0, 0, 0]), False) == np.inf
    M = oracle.Reg("Mass", 3, 1.0)
    assert M.term(_th(-1e-3, [0, 0, 0], good), False) == np.inf
    assert np.isfinite(M.term(_th(0.0, [0, 0, 0], good), False))
    V = oracle.Reg("Volume", 2, 1.0)
    assert V.term(_th(1.0, [0, 0], [1.0, -1.0, 0.0]), False) == np.inf
    Rr = oracle.Reg("Radius", 3, 0.5, 1.0, 2)
    assert Rr.term(_th(1.0, [0, 0, 0], [1.0, 0.5, 1.0, 0, 0, 0]), False) == np.inf
    assert Rr.term(_th(1.0, [0, 0, 0], [0.5]), True) == np.inf


@pytest.mark.parametrize("name,args", [("Joubert", (1.3, 0.7, (0.6, 2.0))), ("Mass", (0.8, 2, 1.5)),
                                       ("Mass", (0.3, 3, 0.5)), ("Volume", (0.9, 2, 2.0)),
                                       ("Volume", (0.7, 3, 1.0)), ("Radius", (0.2, 0.3, 2)),
                                       ("Iso", (1.7,))])
@pytest.mark.parametrize("D,iso", [(3, False), (2, False), (3, True), (2, True)])
def test_regulariser_derivatives_are_derivatives(oracle, name, args, D, iso):
    R = oracle.Reg(name, D, *args)
    rng = np.random.default_rng(3)
    nr = 1 if iso else D * (D + 1) // 2
    for _ in range(3):
        r = np.concatenate((rng.uniform(0.8, 1.6, D if not iso else 1), rng.normal(0, 0.4, nr - (D if not iso else 1))))
        th = _th(rng.uniform(0.5, 1.5), rng.uniform(-0.5, 0.5, D), r)
        g, H = R.derivs(th, iso)
        scale = 1 + np.abs(g).max()
        assert np.abs(g - _fd(lambda t: R.term(t, iso), th)).max() < 1e-6 * scale
        Hfd = np.stack([_fd(lambda t: R.derivs(t, iso)[0][i], th) for i in range(th.size)])
        assert np.abs(H - Hfd).max() < 1e-5 * (1 + np.abs(H).max())
        assert np.allclose(H, H.T, atol=1e-12 * (1 + np.abs(H).max()))


def _problem2d(oracle, na=24, nd=64):
    th = np.linspace(0, np.pi, na, endpoint=False)
    g = oracle.Geometry(oracle.theta_to_vec([th]), oracle.perp([th]), [(np.arange(nd) + 0.5) / (nd / 2) - 1])
    I0, x0, r0 = np.array([1.0, 0.8]), np.array([[-0.4, 0.1], [0.35, -0.2]]), np.array([[3.0, 3.5, 0.5], [4.0, 3.0, -0.3]])
    return g, I0, x0, r0


def test_newton_fit_null_equals_linesearch(oracle):
    """The regularised restatement with no regulariser is the plain one."""
    g, I0, x0, r0 = _problem2d(oracle)
    data = oracle.radproj(I0, x0, r0, g)
    a = oracle.newton_linesearch(I0 * 0.8, x0 + 0.03, r0 * 1.1, g, data, [6, 1])
    b = oracle.newton_fit(I0 * 0.8, x0 + 0.03, r0 * 1.1, g, data, [6, 1])
    for u, v in zip(a, b[:4]):
        assert np.array_equal(u, v)
    E, F = b[3], b[4]
    # F holds each proposal's fidelity: equal to E where accepted, above it where rejected
    assert F.size == E.size and F[0] == E[0]
    acc = np.r_[False, np.diff(E) != 0]
    assert np.array_equal(F[acc], E[acc]) and np.all(F[~acc] >= E[~acc])


def test_regularised_newton_monotone_and_pulls_towards_target(oracle):
    """E = F + reg never increases; a strong Mass regulariser moves the
    fitted masses I / det U towards its target."""
    g, I0, x0, r0 = _problem2d(oracle)
    data = oracle.radproj(I0, x0, r0, g)
    reg = oracle.Reg("Mass", 2, 0.05, 2, 1e3)
    I, x, r, E, F = oracle.newton_fit(I0, x0, r0, g, data, [8, 1], reg=reg)
    assert np.all(np.diff(E) <= 1e-12 * E[0])
    mass0 = I0 / (r0[:, 0] * r0[:, 1])
    mass = I / (r[:, 0] * r[:, 1])
    assert np.all(np.abs(mass - 0.05) < 0.5 * np.abs(mass0 - 0.05))
    # energy bookkeeping: E is the fidelity of the accepted state plus the regulariser
    assert np.isclose(E[-1], oracle.l2_squared_loss(oracle.radproj(I, x, r, g), data) + reg.total(I, x, r))


def test_block_newton_monotone_and_recovers(oracle):
    g, I0, x0, r0 = _problem2d(oracle)
    data = oracle.radproj(I0, x0, r0, g)
    I, x, r, E, F = oracle.newton_fit(I0 * 0.9, x0 + 0.02, r0 * 1.05, g, data, [30, 1], block=True)
    assert np.all(np.diff(E) <= 1e-12 * E[0])
    assert E[-1] < 1e-6 * E[0]
    assert np.allclose(x, x0, atol=1e-3)
    # three blocks per atom visit: at least 3 * n entries per outer iteration
    assert E.size >= 1 + 3 * 2
