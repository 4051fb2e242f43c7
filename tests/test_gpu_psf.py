"""The point-spread projector (pixel / voxel integrals) against the FP64
oracle: values, L2 and linear derivatives, the voxel render, and the
SingleParticleGaussTomo / factory API.

Tolerances (norm-wise relative to the largest reference entry): 1e-5 for
values, gradients and Hessian blocks.  Atoms range from a fraction of a pixel
to several pixels wide, so single pixels, pixels cutting an atom and atoms
inside one pixel all occur.
"""
import importlib

import numpy as np
import pytest
import torch

from conftest import f32

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def psf():
    return importlib.import_module("2018_dictionary_tomography_amd.psf")


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def _case(gt, which, iso, rng, n=6, det=24, nA=4):
    th = np.sort(rng.uniform(0, np.pi, nA))
    p = np.linspace(-1, 1, det)
    if which == "2d":
        space = gt.ProjSpace(orientations=gt.theta_to_vec([th]), ortho=gt.perp([th]), detector=[p])
        D = 2
    elif which == "2p5d":
        space = gt.ProjSpace(orientations=gt.theta_to_vec([th]), ortho=gt.perp([th]), detector=[p, p * 0.9])
        D = 3
    else:
        W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
        T, ortho, _ = W.tilt_series(nA, 60.0, axis=(0.2, 1.0, 0.1), det=det)
        space = gt.ProjSpace(orientations=T, ortho=ortho, detector=[p, p * 0.9])
        D = 3
    I = rng.uniform(0.5, 1.5, n)
    x = rng.uniform(-0.5, 0.5, (n, D))
    # widths (1/U) from about a sixth of a pixel to several pixels
    pix = p[1] - p[0]
    scale = np.exp(rng.uniform(np.log(1 / (3 * pix)), np.log(1 / (0.25 * pix)), n)) / np.sqrt(60)
    if iso:
        r = scale[:, None]
    else:
        nr = D * (D + 1) // 2
        r = np.concatenate((scale[:, None] * rng.uniform(0.8, 1.25, (n, D)),
                            scale[:, None] * rng.normal(0, 0.2, (n, nr - D))), 1)
    aspace = gt.AtomSpace(D, isotropic=iso)
    return I, x, r, space, aspace


def _og(oracle, space):
    return oracle.Geometry(f32(space.orientations), tuple(f32(o) for o in space.ortho),
                           [f32(q) for q in space.detector])


@pytest.mark.parametrize("which", ["3d", "2p5d", "2d"])
@pytest.mark.parametrize("iso", [False, True])
def test_psf_forward_vs_fp64(gt, oracle, cuda, psf, which, iso):
    rng = np.random.default_rng(7 + 2 * iso + {"3d": 0, "2p5d": 10, "2d": 20}[which])
    I, x, r, space, aspace = _case(gt, which, iso, rng)
    atoms = gt.AtomElement(aspace, x, r, I, device=cuda)
    P = psf.forward(atoms, space).cpu().numpy()
    ref = oracle.psf_radproj(f32(I), f32(x), f32(r), _og(oracle, space), 24)
    assert P.shape == ref.shape
    assert _rel(P, ref) < 1e-5


@pytest.mark.parametrize("which", ["3d", "2p5d", "2d"])
@pytest.mark.parametrize("iso", [False, True])
@pytest.mark.parametrize("mode", ["L2", "linear"])
def test_psf_derivs_vs_fp64(gt, oracle, cuda, psf, which, iso, mode):
    rng = np.random.default_rng(40 + 2 * iso + {"3d": 0, "2p5d": 10, "2d": 20}[which] + 100 * (mode == "L2"))
    I, x, r, space, aspace = _case(gt, which, iso, rng, n=3, det=20, nA=3)
    atoms = gt.AtomElement(aspace, x, r, I, device=cuda)
    data = psf.forward(gt.AtomElement(aspace, x + 0.01, r, I * 1.1, device=cuda), space)
    C = data + torch.as_tensor(rng.normal(0, 0.02 * float(data.abs().max()), data.shape), dtype=torch.float32,
                               device=cuda)
    f, g, H = (t.cpu().numpy() for t in psf.derivs(atoms, space, C, mode, 2))
    fo, go, Ho = oracle.psf_derivs(f32(I), f32(x), f32(r), _og(oracle, space), C.double().cpu().numpy(), mode, 2, 16)
    assert _rel(f, fo) < 1e-5
    for j in range(I.size):
        assert _rel(g[j], go[j]) < 1e-5, (j, g[j], go[j])
        assert _rel(H[j], Ho[j]) < 1e-5, j
    # order 1 gives the same gradient
    _, g1, H1 = psf.derivs(atoms, space, C, mode, 1)
    assert H1 is None and np.array_equal(g1.cpu().numpy(), g)


@pytest.mark.parametrize("D", [3, 2])
def test_psf_volume_vs_fp64(gt, oracle, cuda, psf, D):
    rng = np.random.default_rng(60 + D)
    n = 5
    y = np.linspace(-1, 1, 20 if D == 3 else 40)
    pix = y[1] - y[0]
    scale = np.exp(rng.uniform(np.log(1 / (2 * pix)), np.log(1 / (0.3 * pix)), n)) / np.sqrt(60)
    nr = D * (D + 1) // 2
    r = np.concatenate((scale[:, None] * rng.uniform(0.8, 1.25, (n, D)),
                        scale[:, None] * rng.normal(0, 0.2, (n, nr - D))), 1)
    I = rng.uniform(0.5, 1.5, n)
    x = rng.uniform(-0.5, 0.5, (n, D))
    atoms = gt.AtomElement(gt.AtomSpace(D, isotropic=False), x, r, I, device=cuda)
    vol = gt.VolSpace([y] * D)
    u = psf.volume(atoms, vol).cpu().numpy()
    ref = oracle.psf_volproj(f32(I), f32(x), f32(r), [f32(y)] * D, 12)
    assert u.shape == ref.shape
    assert _rel(u, ref) < 1e-5


def test_psf_volume_many_atoms_vs_fp64(gt, oracle, cuda, psf):
    """300 atoms: more than one 256-record staging chunk per brick, each
    brick's list shared by the block's thread groups; repeated calls give the
    same bits (the groups' sums are added in a fixed order)."""
    rng = np.random.default_rng(71)
    n = 300
    y = np.linspace(-1, 1, 12)
    pix = y[1] - y[0]
    scale = np.exp(rng.uniform(np.log(1 / (2 * pix)), np.log(1 / (0.5 * pix)), n)) / np.sqrt(60)
    r = np.concatenate((scale[:, None] * rng.uniform(0.8, 1.25, (n, 3)), scale[:, None] * rng.normal(0, 0.2, (n, 3))),
                       1)
    I = rng.uniform(0.5, 1.5, n)
    x = rng.uniform(-0.8, 0.8, (n, 3))
    atoms = gt.AtomElement(gt.AtomSpace(3, isotropic=False), x, r, I, device=cuda)
    vol = gt.VolSpace([y] * 3)
    first = psf.volume(atoms, vol)
    ref = oracle.psf_volproj(f32(I), f32(x), f32(r), [f32(y)] * 3, 10)
    assert _rel(first.cpu().numpy(), ref) < 1e-5
    for _ in range(2):
        assert torch.equal(psf.volume(atoms, vol), first)


def test_single_particle_operator_and_factory(gt, oracle, cuda, psf):
    """SingleParticleGaussTomo has GaussTomo's methods over the point-spread
    kernels; the module drop-ins write into host arrays; the factory builds
    the operator and simulates its data with it."""
    rng = np.random.default_rng(5)
    I, x, r, space, aspace = _case(gt, "3d", False, rng, n=4, det=20, nA=3)
    vol = gt.VolSpace(box=(-1, 1), shape=12, dim=3)
    op = gt.SingleParticleGaussTomo(aspace, vol, space)
    atoms = gt.AtomElement(aspace, x, r, I, device=cuda)
    R = op(atoms)
    assert torch.equal(R.asarray(), psf.forward(atoms, space))
    assert op.discretise(atoms).asarray().shape == (12, 12, 12)
    f, df, ddf = op.L2_derivs(atoms, R.asarray() * 0.9)
    assert f.shape == (4, 1) and df.shape == (4, 10) and ddf.shape == (4, 10, 10)
    Rh = np.zeros(space.shape, dtype=np.float32)
    psf.RadProj(atoms, op, Rh)
    assert np.array_equal(Rh, R.asarray().cpu().numpy())
    f0, g0, H0 = psf.L2derivs_RadProj(atoms, op, R.asarray() * 0.9)
    assert np.allclose(g0, df[0], rtol=0, atol=0) and abs(f0 - f[0, 0]) <= 1e-12 * abs(f[0, 0])
    # factory: data simulated by the point-spread operator
    Radon, fid, data, ASpace, PSpace, (reg, GD) = gt.standardSingleParticleGaussTomo(
        gt=atoms, dim=3, vol_size=12, angle_range=(-1, 1), angle_num=5, det_size=[16, 16])
    assert isinstance(Radon, gt.SingleParticleGaussTomo)
    assert torch.equal(data.asarray(), psf.forward(atoms, PSpace))


def _fit_case(gt, psf, rng, which, cuda):
    """A small fit: atoms one to three pixels wide, data = the point-spread
    sinogram of the true atoms plus 1% noise, start perturbed (intensity x0.85,
    radius x1.05, centres by 0.03)."""
    p = np.linspace(-1, 1, 20)
    pix = p[1] - p[0]
    if which == "2d":
        th = np.sort(rng.uniform(0, np.pi, 5))
        space = gt.ProjSpace(orientations=gt.theta_to_vec([th]), ortho=gt.perp([th]), detector=[p])
        D, iso, n = 2, True, 3
    else:
        W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
        T, ortho, _ = W.tilt_series(4, 60.0, axis=(0.2, 1.0, 0.1), det=20)
        space = gt.ProjSpace(orientations=T, ortho=ortho, detector=[p, p * 0.9])
        D, iso, n = 3, False, 3
    I = rng.uniform(0.8, 1.2, n)
    x = rng.uniform(-0.5, 0.5, (n, D))
    scale = rng.uniform(1 / (3 * pix), 1 / pix, n) / np.sqrt(60)
    if iso:
        r = scale[:, None]
    else:
        nr = D * (D + 1) // 2
        r = np.concatenate((scale[:, None] * rng.uniform(0.9, 1.1, (n, D)),
                            scale[:, None] * rng.normal(0, 0.1, (n, nr - D))), 1)
    aspace = gt.AtomSpace(D, isotropic=iso)
    # intensities scaled so that the sinogram peaks at 1: the fidelity's
    # Hessian then dominates the initial damping 1/h = 1
    I = I / float(psf.forward(gt.AtomElement(aspace, x, r, I, device=cuda), space).abs().max())
    clean = psf.forward(gt.AtomElement(aspace, x, r, I, device=cuda), space)
    data = clean + 0.01 * float(clean.abs().max()) * torch.as_tensor(rng.normal(size=clean.shape),
                                                                      dtype=torch.float32, device=cuda)
    x0 = x + rng.normal(0, 0.03, x.shape)
    return f32(I * 0.85), f32(x0), f32(r * 1.05), space, aspace, data


@pytest.mark.parametrize("which,block", [("3d", False), ("3d", True), ("2d", False), ("2d", True)])
def test_newton_with_single_particle_operator_matches_oracle(gt, oracle, cuda, psf, which, block):
    """NewtonGaussian.linesearch / linesearch_block only call Radon(.) and
    Radon.L2_derivs, so with SingleParticleGaussTomo they fit the
    pixel-integrated model: energy histories as the FP64 restatement with the
    pixel-integrated projector."""
    rng = np.random.default_rng(90 + 2 * (which == "2d") + block)
    I0, x0, r0, space, aspace, data = _fit_case(gt, psf, rng, which, cuda)
    vol = gt.VolSpace(box=(-1, 1), shape=8, dim=aspace.dim)
    op = gt.SingleParticleGaussTomo(aspace, vol, space)
    recon = gt.AtomElement(aspace, x0, r0, I0, device=cuda)
    solver = gt.getSolver("Newton_block" if block else "Newton")
    recon, E, Fh = solver(recon, data, [3, 1], gt.l2_squared_loss(aspace.dim), None, op, tol=1e-6)
    Io, xo, ro, Eo, Fo = oracle.newton_fit(I0, x0, r0, _og(oracle, space), data.double().cpu().numpy(), [3, 1],
                                           tol=1e-6, block=block, projector="psf", n_sub=12)
    assert E.size == Eo.size and Eo[-1] < 0.5 * Eo[0]
    assert np.all(np.diff(E) <= 0)
    assert np.abs(E - Eo).max() < 1e-5 * Eo[0]
    # proposals (also rejected ones, whose fidelity can exceed F[0] many times)
    assert np.all(np.abs(Fh - Fo) < 1e-5 * np.maximum(Fo, Fo[0]))
    assert np.abs(recon.x.cpu().numpy() - xo).max() < 1e-4
    # the final energy is the fidelity of the final state under the same operator
    R = op(recon).asarray()
    assert abs(0.5 * float(((R - data).double() ** 2).sum()) - E[-1]) < 1e-5 * E[-1]


def test_gd_block_with_single_particle_operator_matches_oracle(gt, oracle, cuda, psf):
    rng = np.random.default_rng(97)
    I0, x0, r0, space, aspace, data = _fit_case(gt, psf, rng, "3d", cuda)
    op = gt.SingleParticleGaussTomo(aspace, gt.VolSpace(box=(-1, 1), shape=8, dim=3), space)
    recon = gt.AtomElement(aspace, x0, r0, I0, device=cuda)
    F = importlib.import_module("2018_dictionary_tomography_amd.fit")
    recon, E, _ = F.linesearch_block(recon, data, [3, 1], gt.l2_squared_loss(3), None, op, tol=1e-6)
    Io, xo, ro, Eo = oracle.gd_linesearch_block(I0, x0, r0, _og(oracle, space), data.double().cpu().numpy(), [3, 1],
                                                tol=1e-6, projector="psf", n_sub=12)
    assert E.size == Eo.size and Eo[-1] < Eo[0]
    assert np.abs(E - Eo).max() < 1e-5 * Eo[0]


def test_host_newton_with_point_kernels_equals_device_fitter(gt, cuda, psf):
    """The host-driven loop over the point projector's kernels and the
    device-resident fitter run the same algorithm."""
    F = importlib.import_module("2018_dictionary_tomography_amd.fit")
    rng = np.random.default_rng(99)
    I0, x0, r0, space, aspace, _ = _fit_case(gt, psf, rng, "3d", cuda)
    truth = gt.AtomElement(aspace, x0 + 0.02, r0, I0 * 1.1, device=cuda)
    data = gt.ops.forward(truth, space)
    a = gt.AtomElement(aspace, x0, r0, I0, device=cuda)
    b = gt.AtomElement(aspace, x0, r0, I0, device=cuda)
    _, Eh, Fh = F.newton_host(a, space, data, [3, 1], F.POINT, tol=1e-6)
    _, Ed, Fd = F.newton_linesearch(b, data, [3, 1], None, None, None, tol=1e-6, space=space)
    assert Eh.size == Ed.size
    assert np.abs(Eh - Ed).max() < 1e-5 * Ed[0] and np.abs(Fh - Fd).max() < 1e-5 * Fd[0]
    assert np.abs(a.x.cpu().numpy() - b.x.cpu().numpy()).max() < 1e-4


@pytest.mark.parametrize("mode", ["L2", "linear"])
def test_psf_wide_atoms_vs_fp64(gt, oracle, cuda, psf, mode):
    """Atoms 3-10 pixels wide, the spread of the 3D workload's atoms on its
    detector: every pixel takes the unrolled 3 x 3 / 4 x 4 rules."""
    rng = np.random.default_rng(77 + (mode == "L2"))
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    T, ortho, _ = W.tilt_series(3, 60.0, axis=(0.2, 1.0, 0.1), det=40)
    p = np.linspace(-1, 1, 40)
    space = gt.ProjSpace(orientations=T, ortho=ortho, detector=[p, p])
    pix = p[1] - p[0]
    n = 3
    sig = rng.uniform(3 * pix, 10 * pix, n)
    r = np.concatenate(((1.0 / (sig * np.sqrt(60)))[:, None] * rng.uniform(0.9, 1.1, (n, 3)),
                        (1.0 / (sig * np.sqrt(60)))[:, None] * rng.normal(0, 0.15, (n, 3))), 1)
    I = rng.uniform(0.5, 1.5, n)
    x = rng.uniform(-0.3, 0.3, (n, 3))
    atoms = gt.AtomElement(gt.AtomSpace(3, isotropic=False), x, r, I, device=cuda)
    P = psf.forward(atoms, space).cpu().numpy()
    ref = oracle.psf_radproj(f32(I), f32(x), f32(r), _og(oracle, space), 12)
    assert _rel(P, ref) < 1e-5
    C = torch.as_tensor(ref * 0.9 + rng.normal(0, 0.01 * ref.max(), ref.shape), dtype=torch.float32, device=cuda)
    f, g, H = (t.cpu().numpy() for t in psf.derivs(atoms, space, C, mode, 2))
    fo, go, Ho = oracle.psf_derivs(f32(I), f32(x), f32(r), _og(oracle, space), C.double().cpu().numpy(), mode, 2, 8)
    assert _rel(f, fo) < 1e-5
    for j in range(n):
        assert _rel(g[j], go[j]) < 1e-5 and _rel(H[j], Ho[j]) < 1e-5, j
