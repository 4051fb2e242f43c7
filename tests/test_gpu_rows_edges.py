"""Edge cases of the point-spread, Fourier-plane and KL-moment entry points:
empty atom sets, a single angle, a non-square detector, and the 16-byte
load path of the sinogram moments (rows of a multiple of 4 pixels), each
against the FP64 oracle or the exact empty result."""
import importlib

import numpy as np
import pytest
import torch

from conftest import f32

pytestmark = pytest.mark.gpu


def _mods():
    return (importlib.import_module("2018_dictionary_tomography_amd.psf"),
            importlib.import_module("2018_dictionary_tomography_amd.fourier"),
            importlib.import_module("2018_dictionary_tomography_amd.kl"))


def _og(oracle, space):
    return oracle.Geometry(f32(space.orientations), tuple(f32(o) for o in space.ortho),
                           [f32(q) for q in space.detector])


def test_empty_atom_sets(gt, cuda):
    psf, Fo, _ = _mods()
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    T, ortho, det = W.tilt_series(3, 60.0, det=16)
    space = gt.ProjSpace(orientations=T, ortho=ortho, detector=det)
    none = gt.AtomElement(gt.AtomSpace(3, isotropic=False), np.zeros((0, 3)), np.zeros((0, 6)), np.zeros(0),
                          device=cuda)
    P = psf.forward(none, space)
    assert P.shape == (3, 16, 16) and torch.count_nonzero(P) == 0
    f, g, H = psf.derivs(none, space, torch.ones(space.shape, device=cuda), "L2", 2)
    assert f.numel() == 0 and g.shape == (0, 10) and H.shape == (0, 10, 10)
    u = psf.volume(none, gt.VolSpace(box=(-1, 1), shape=8, dim=3))
    assert u.shape == (8, 8, 8) and torch.count_nonzero(u) == 0
    ft = Fo.FTAtoms(np.zeros(0), np.zeros((0, 3)), np.zeros((0, 6)), cuda)
    k = np.linspace(-5, 5, 11)
    planes = Fo.evaluate(ft, (k, k, np.asarray(ortho[0]), np.asarray(ortho[1])))
    assert planes.shape == (3, 11, 11) and torch.count_nonzero(planes) == 0
    f, g, H = Fo.derivs_all(ft, (k, k, np.asarray(ortho[0]), np.asarray(ortho[1])), planes)
    assert f.numel() == 0 and g.shape[0] == 0


def test_psf_single_angle_non_square_detector(gt, oracle, cuda):
    psf, _, _ = _mods()
    rng = np.random.default_rng(12)
    th = np.array([0.4])
    space = gt.ProjSpace(orientations=gt.theta_to_vec([th]), ortho=gt.perp([th]),
                         detector=[np.linspace(-1, 1, 17), np.linspace(-0.8, 0.9, 29)])
    n = 4
    I = rng.uniform(0.5, 1.5, n)
    x = rng.uniform(-0.4, 0.4, (n, 3))
    r = np.concatenate((rng.uniform(3, 8, (n, 3)), rng.normal(0, 0.5, (n, 3))), 1)
    atoms = gt.AtomElement(gt.AtomSpace(3, isotropic=False), x, r, I, device=cuda)
    P = psf.forward(atoms, space).cpu().numpy()
    ref = oracle.psf_radproj(f32(I), f32(x), f32(r), _og(oracle, space), 24)
    assert P.shape == (1, 17, 29)
    assert np.abs(P - ref).max() < 1e-5 * np.abs(ref).max()


def test_sino_moments_vector_path_vs_fp64(gt, oracle, cuda):
    _, _, KL = _mods()
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    T, ortho, det = W.tilt_series(5, 60.0, det=128)   # rows of 128 pixels: 16-byte loads
    space = gt.ProjSpace(orientations=T, ortho=ortho, detector=det)
    rng = np.random.default_rng(4)
    S = torch.as_tensor(rng.normal(0, 1, space.shape), dtype=torch.float32, device=cuda)
    Q = torch.as_tensor(rng.uniform(0, 2, space.shape), dtype=torch.float32, device=cuda)
    g = _og(oracle, space)
    Sh, Qh = S.double().cpu().numpy(), Q.double().cpu().numpy()
    for mode, a, b, f in ((0, 0.3, 0.6, np.maximum(0, 0.3 * Sh + 0.6 * Qh)),
                          (1, 0.7, 1.1, oracle.kl_newton_sino(Sh, 1.1 * Qh, 0.7))):
        m = KL.sino_moments(S, space, Q, mode, a, b)
        ref = oracle.kl_sino_moments(f, g)
        assert np.abs(m - ref).max() < 1e-6 * np.abs(ref).max(), mode
    # an odd offset breaks 16-byte alignment: the scalar path gives the same sums
    big = torch.zeros(S.numel() + 1, device=cuda)
    big[1:] = S.reshape(-1)
    m_al = KL.sino_moments(S, space)
    m_un = KL.sino_moments(big[1:].view(space.shape), space)
    assert np.abs(m_al - m_un).max() <= 1e-12 * np.abs(m_al).max()
