"""Fourier evaluation on sample planes, the complex L2 derivatives over
hyperplane samples, GaussFTVolume and the projection DFT, against the FP64
oracle.

Tolerances (norm-wise relative to the largest reference entry): plane values
1e-5 (float32 samples) and 1e-12 (float64 samples); f, gradients and Hessian
blocks 1e-5; the Fourier-slice check of a projected sinogram (discretisation:
detector sampling and truncation) 1e-4.
"""
import importlib

import numpy as np
import pytest
import torch

from conftest import f32

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Fo():
    return importlib.import_module("2018_dictionary_tomography_amd.fourier")


def _atoms3(rng, n, iso=False):
    I = rng.uniform(0.5, 1.5, n)
    x = rng.uniform(-0.8, 0.8, (n, 3))
    if iso:
        r = rng.uniform(0.25, 0.6, (n, 1))
    else:
        r = np.concatenate((rng.uniform(0.12, 0.3, (n, 3)), rng.normal(0, 0.05, (n, 3))), 1)
    return I, x, r


def _frames(rng, nA):
    th = np.sort(rng.uniform(0, np.pi, nA))
    ph = rng.uniform(-0.6, 0.6, nA)
    w0 = np.stack([np.cos(th), np.sin(th), 0 * th], 1)
    w1 = np.stack([-np.sin(th) * np.sin(ph), np.cos(th) * np.sin(ph), np.cos(ph)], 1)
    return w0, w1, np.stack([np.cos(th), np.sin(th)], 1)


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


@pytest.mark.parametrize("iso", [False, True])
def test_planes_3d_and_2p5d_vs_fp64(gt, oracle, cuda, Fo, iso):
    rng = np.random.default_rng(10 + iso)
    I, x, r = _atoms3(rng, 150, iso)
    w0, w1, w2 = _frames(rng, 7)
    k0, k1 = np.linspace(-30, 30, 61), np.linspace(-24, 26, 45)
    atoms = Fo.FTAtoms(I, x, r, cuda)
    got = Fo.evaluate(atoms, (k0, k1, w0, w1)).cpu().numpy()
    assert got.shape == (7, 61, 45) and got.dtype == np.complex64
    ref = oracle.ft_hyperplane(f32(I), f32(x), f32(r), f32(k0), f32(k1), f32(w0), f32(w1), iso=iso)
    assert _rel(got, ref) < 1e-5
    got = Fo.evaluate(atoms, (k0, k1, w2)).cpu().numpy()
    ref = oracle.ft_hyperplane(f32(I), f32(x), f32(r), f32(k0), f32(k1), f32(w2), None, iso=iso)
    assert _rel(got, ref) < 1e-5
    # unsorted sample vectors: no culling, same values
    perm = rng.permutation(61)
    got_u = Fo.evaluate(atoms, (k0[perm], k1, w0, w1)).cpu().numpy()
    ref_u = oracle.ft_hyperplane(f32(I), f32(x), f32(r), f32(k0[perm]), f32(k1), f32(w0), f32(w1), iso=iso)
    assert _rel(got_u, ref_u) < 1e-5


@pytest.mark.parametrize("iso", [False, True])
def test_2d_grid_and_lines_f32_f64(gt, oracle, cuda, Fo, iso):
    """FTGaussian2D.evaluate (FTGaussian2D.py:13-35): grid and hyperplane
    samplers, float32 -> complex64 and float64 -> complex128."""
    rng = np.random.default_rng(20 + iso)
    n = 80
    I = rng.uniform(0.5, 1.5, n)
    x = rng.uniform(-0.8, 0.8, (n, 2))
    r = rng.uniform(0.1, 0.3, (n, 1)) if iso else np.concatenate(
        (rng.uniform(0.1, 0.3, (n, 2)), rng.normal(0, 0.05, (n, 1))), 1)
    k0, k1 = np.linspace(-40, 40, 81), np.linspace(-35, 30, 52)
    th = rng.uniform(0, np.pi, 9)
    w = np.stack([np.cos(th), np.sin(th)], 1)
    for dt, tol in ((np.float32, 1e-5), (np.float64, 1e-12)):
        atoms = Fo.FTAtoms(I, x, r, cuda, dim=2, dtype=dt)
        cast = (lambda v: f32(v)) if dt == np.float32 else (lambda v: np.asarray(v, float))
        g = Fo.evaluate(atoms, (k0.astype(dt), k1.astype(dt))).cpu().numpy()
        assert g.shape == (81, 52) and g.dtype == (np.complex64 if dt == np.float32 else np.complex128)
        mesh = np.stack(np.meshgrid(cast(k0), cast(k1), indexing="ij"), -1).reshape(-1, 2)
        ref = oracle.ft_points(cast(I), cast(x), cast(r), mesh, iso=iso).reshape(g.shape)
        assert _rel(g, ref) < tol
        line = Fo.evaluate(atoms, (k0.astype(dt), w.astype(dt))).cpu().numpy()
        assert line.shape == (9, 81)
        K = (cast(k0)[None, :, None] * cast(w)[:, None, :]).reshape(-1, 2)
        ref = oracle.ft_points(cast(I), cast(x), cast(r), K, iso=iso).reshape(line.shape)
        assert _rel(line, ref) < tol


@pytest.mark.parametrize("form,iso", [("3d", False), ("2p5d", False), ("3d", True), ("2p5d", True)])
def test_ft_derivs_vs_fp64(gt, oracle, cuda, Fo, form, iso):
    rng = np.random.default_rng(30 + 2 * iso + (form == "3d"))
    I, x, r = _atoms3(rng, 12, iso)
    w0, w1, w2 = _frames(rng, 5)
    k0, k1 = np.linspace(-25, 25, 51), np.linspace(-20, 22, 43)
    S = (k0, k1, w0, w1) if form == "3d" else (k0, k1, w2)
    atoms = Fo.FTAtoms(I, x, r, cuda)
    # data: the FT of a perturbed atom set plus noise, so residuals are O(1)
    C = Fo.evaluate(Fo.FTAtoms(I * 1.1, x + 0.02, r, cuda), S).cpu().numpy()
    C = (C + 0.05 * (rng.normal(size=C.shape) + 1j * rng.normal(size=C.shape))).astype(np.complex64)
    f, g, H = (t.cpu().numpy() for t in Fo.derivs_all(atoms, S, torch.as_tensor(C, device=cuda)))
    wo = (w0, w1) if form == "3d" else (w2, None)
    fo, go, Ho = oracle.ft_derivs(f32(I), f32(x), f32(r), f32(k0), f32(k1), f32(wo[0]),
                                  None if wo[1] is None else f32(wo[1]), C.astype(np.complex128), iso=iso)
    P = 5 if iso else 10
    assert g.shape == (12, P) and H.shape == (12, P, P)
    assert _rel(f, fo) < 1e-5
    for j in range(12):
        assert _rel(g[j], go[j]) < 1e-5, j
        assert _rel(H[j], Ho[j]) < 1e-5, j
    # the reference's single-atom form
    f0, g0, H0 = Fo.derivs(atoms, S, C)
    assert abs(f0 - f[0]) <= 1e-12 * abs(f[0]) and np.array_equal(g0, g[0]) and np.array_equal(H0, H[0])


def test_gauss_ft_volume_and_projection_dft(gt, oracle, cuda, Fo):
    """GaussFTVolume evaluates FT-space atoms on the Fourier-slice planes of a
    projection space; the DFT of the sinogram of the same atoms in real space
    (dft_projections) agrees with it (Fourier slice theorem)."""
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(3)
    n = 20
    # atoms wide enough for the detector sampling (no aliasing) and narrow
    # enough to stay inside the detector (no truncation): std 0.02 .. 0.05
    I, x, r, _ = W.draw_aniso3(n, rng, rho=(0.025, 0.05), offdiag=3.0)
    x *= 0.6
    T, ortho, det = W.tilt_series(9, 60.0, det=192, box=1.0)
    space = gt.ProjSpace(orientations=T, ortho=ortho, detector=det)
    real = gt.AtomElement(gt.AtomSpace(3, isotropic=False), x, r, I, device=cuda)
    sino = gt.ops.forward(real, space)
    FT = Fo.dft_projections(sino, space)
    # projector atom I sqrt(30/pi) exp(-30 |U(y - x)|^2) = I' exp(-|U'(y - x)|^2 / 2)
    nI, nx, nr = Fo.analytic_gauss_fourier(I * np.sqrt(30 / np.pi), x, r * np.sqrt(60.0))
    vol = Fo.GaussFTVolume(gt.AtomSpace(3, isotropic=False), space)
    ana = vol(Fo.FTAtoms(nI, nx, nr, cuda)).asarray()
    assert ana.shape == FT.shape and ana.is_complex()
    assert float((FT - ana).abs().max() / ana.abs().max()) < 1e-4
    # L2_derivs: per atom, the derivs of the planes
    C = ana.cpu().numpy() * 0.9
    f, df, ddf = vol.L2_derivs(Fo.FTAtoms(nI, nx, nr, cuda), C)
    fa, ga, Ha = Fo.derivs_all(Fo.FTAtoms(nI, nx, nr, cuda), vol.grid, torch.as_tensor(C, device=cuda))
    assert f.shape == (n, 1) and df.shape == (n, 10) and ddf.shape == (n, 10, 10)
    assert np.array_equal(df, ga.cpu().numpy())
    # inverse DFT returns the sinogram
    back = Fo.dft_projections(FT, space, inverse=True)
    assert float((back.real - sino).abs().max() / sino.abs().max()) < 1e-5


def _ft_fit_case(gt, Fo, cuda, seed):
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    T, ortho, det = W.tilt_series(5, 60.0, det=24)
    space = gt.ProjSpace(orientations=T, ortho=ortho, detector=det)
    vol = Fo.GaussFTVolume(gt.AtomSpace(3, isotropic=False), space)
    rng = np.random.default_rng(seed)
    n = 2
    I = rng.uniform(0.8, 1.2, n)
    x = rng.uniform(-0.4, 0.4, (n, 3))
    r = np.concatenate((rng.uniform(0.5, 0.8, (n, 3)), rng.normal(0, 0.1, (n, 3))), 1)
    data = Fo.evaluate(Fo.FTAtoms(I, x, r, cuda), vol.grid)
    noise = rng.normal(size=data.shape) + 1j * rng.normal(size=data.shape)
    data = data + torch.as_tensor(0.01 * float(data.abs().max()) * noise, dtype=torch.complex64, device=cuda)
    start = (f32(I * 0.85), f32(x + 0.03), f32(r * 1.05))
    g = tuple(f32(v) for v in vol.grid)
    return vol, data, start, g


@pytest.mark.parametrize("solver", ["Newton", "GD_block"])
def test_fit_with_gauss_ft_volume_matches_oracle(gt, oracle, cuda, Fo, solver):
    """The reference's fitters call only Radon(.) and Radon.L2_derivs, so
    with a GaussFTVolume they fit FT-space atoms to Fourier-plane data (the
    complex l2 loss): same energy histories as the FP64 restatement."""
    vol, data, (I0, x0, r0), g = _ft_fit_case(gt, Fo, cuda, 61 + (solver == "Newton"))
    recon = Fo.FTAtoms(I0, x0, r0, cuda)
    recon, E, F = gt.getSolver(solver)(recon, data, [4, 1], None, None, vol, tol=1e-6)
    d = data.cpu().numpy().astype(np.complex128)
    if solver == "Newton":
        Io, xo, ro, Eo, Fo_ = oracle.newton_fit(I0, x0, r0, g, d, [4, 1], tol=1e-6, projector="ft")
        assert np.all(np.abs(F - Fo_) < 1e-5 * np.maximum(Fo_, Fo_[0]))
    else:
        Io, xo, ro, Eo = oracle.gd_linesearch_block(I0, x0, r0, g, d, [4, 1], tol=1e-6, projector="ft")
    assert E.size == Eo.size and Eo[-1] < 0.5 * Eo[0]
    assert np.abs(E - Eo).max() < 1e-5 * Eo[0]
    assert np.abs(recon.x.cpu().numpy() - xo).max() < 1e-4
    with pytest.raises(NotImplementedError):
        gt.getSolver(solver)(recon, data, [1, 1], None, None, vol, guess=lambda res, a: a)
