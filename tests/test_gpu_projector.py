"""GPU parity of the projector, derivatives, volume render and Fourier
evaluation against the FP64 oracle (same float32 inputs on both sides).

Tolerances (norm-wise relative, ||y - y64||_inf / ||y64||_inf):
  forward / volume / Fourier values       1e-5
  gradients, Hessian blocks, f            1e-5
"""
import numpy as np
import pytest
import torch

from conftest import f32, relerr

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _atoms(gt, I, x, r, dim, dev):
    r = np.asarray(r)
    nr = r.shape[1] if r.ndim == 2 else (r.size // max(len(I), 1))
    space = gt.AtomSpace(dim, isotropic=(nr == 1))
    return gt.AtomElement(space, x, r, I, device=dev)


def _space3(gt, T, ortho, det):
    return gt.ProjSpace(orientations=T, ortho=ortho, detector=det)


def _ogeom(oracle, space):
    return oracle.Geometry(f32(space.orientations), tuple(f32(o) for o in space.ortho),
                           [f32(p) for p in space.detector])


def _c2_case(gt, N, nA, det, seed):
    from importlib import import_module

    W = import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(seed)
    I, x, r, _ = W.draw_aniso3(N, rng)
    T, ortho, dv = W.tilt_series(nA, 60.0, det=det)
    return I, x, r, _space3(gt, T, ortho, dv), rng


@pytest.mark.parametrize("N,nA,det", [(1, 3, 32), (37, 9, 48), (200, 61, 128)])
def test_forward_3d(gt, oracle, cuda, N, nA, det):
    I, x, r, space, _ = _c2_case(gt, N, nA, det, seed=N)
    atoms = _atoms(gt, I, x, r, 3, cuda)
    R = gt.ops.forward(atoms, space).cpu().numpy()
    R64 = oracle.radproj(f32(I), f32(x), f32(r), _ogeom(oracle, space))
    assert R.shape == R64.shape
    assert relerr(R, R64) < TOL


def test_forward_2p5d_and_nonsquare(gt, oracle, cuda):
    rng = np.random.default_rng(5)
    W = __import__("importlib").import_module("2018_dictionary_tomography_amd.workloads")
    I, x, r, _ = W.draw_aniso3(40, rng)
    ang = np.linspace(0, np.pi, 13, endpoint=False)
    space = gt.ProjSpace(angles=ang, detector=(gt.cell_centres(-1.2, 1.2, 70), gt.cell_centres(-1, 1, 45)))
    atoms = _atoms(gt, I, x, r, 3, cuda)
    R = gt.ops.forward(atoms, space).cpu().numpy()
    R64 = oracle.radproj(f32(I), f32(x), f32(r), _ogeom(oracle, space))
    assert R.shape == (13, 70, 45)
    assert relerr(R, R64) < TOL


@pytest.mark.parametrize("N,shape", [(1100, (37, 53)), (300, (5, 100)), (300, (130, 7)), (65, (1, 1)),
                                     (513, (64, 200))])
def test_forward_3d_tile_order_shapes(gt, oracle, cuda, N, shape):
    """Detectors that are not multiples of the 8 x 64 pixel tiles or their
    8 x 16 wave rectangles, tile grids that are not powers of two (empty
    blocks in the centre-first block order), several 256-record staging chunks
    per tile, a one-pixel detector; repeated calls give the same bits."""
    W = __import__("importlib").import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(N + shape[0])
    I, x, r, _ = W.draw_aniso3(N, rng)
    T, ortho, _ = W.tilt_series(7, 60.0, det=8)
    det = [gt.cell_centres(-1, 1, shape[0]), gt.cell_centres(-0.9, 1.1, shape[1])]
    space = _space3(gt, T, ortho, det)
    atoms = _atoms(gt, I, x, r, 3, cuda)
    R = gt.ops.forward(atoms, space)
    R64 = oracle.radproj(f32(I), f32(x), f32(r), _ogeom(oracle, space))
    assert R.shape == R64.shape
    assert relerr(R.cpu().numpy(), R64) < TOL
    for _ in range(3):
        assert torch.equal(gt.ops.forward(atoms, space), R)


def test_forward_2d_iso_c1(gt, oracle, cuda):
    W = __import__("importlib").import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(11)
    I, x, r = W.draw_iso2(50, rng)
    T, ortho, det = W.fan_2d(180, 128)
    space = gt.ProjSpace(orientations=T, ortho=ortho, detector=det)
    atoms = _atoms(gt, I, x, r, 2, cuda)
    R = gt.ops.forward(atoms, space).cpu().numpy()
    R64 = oracle.radproj(f32(I), f32(x), f32(r), _ogeom(oracle, space))
    assert R.shape == (180, 128)
    assert relerr(R, R64) < TOL


def test_forward_unsorted_detector_disables_culling(gt, oracle, cuda):
    I, x, r, space, _ = _c2_case(gt, 12, 5, 32, seed=3)
    det = [space.detector[0][::-1].copy(), space.detector[1]]
    space2 = _space3(gt, space.orientations, space.ortho, det)
    atoms = _atoms(gt, I, x, r, 3, cuda)
    R = gt.ops.forward(atoms, space2).cpu().numpy()
    R64 = oracle.radproj(f32(I), f32(x), f32(r), _ogeom(oracle, space2))
    assert relerr(R, R64) < TOL


def test_forward_empty_and_offdetector(gt, oracle, cuda):
    I, x, r, space, _ = _c2_case(gt, 4, 5, 32, seed=4)
    empty = _atoms(gt, np.zeros(0), np.zeros((0, 3)), np.zeros((0, 6)), 3, cuda)
    assert np.all(gt.ops.forward(empty, space).cpu().numpy() == 0)
    far = x + 10.0  # every footprint lies off the detector
    atoms = _atoms(gt, I, far, r, 3, cuda)
    assert np.all(gt.ops.forward(atoms, space).cpu().numpy() == 0)


def _derivs_case(gt, oracle, cuda, space, I, x, r, mode, dim):
    atoms = _atoms(gt, I, x, r, dim, cuda)
    og = _ogeom(oracle, space)
    R64 = oracle.radproj(f32(I), f32(x), f32(r), og)
    rng = np.random.default_rng(7)
    C = f32(R64 * 0.7 + rng.normal(0, 0.01 * np.abs(R64).max(), R64.shape))
    f, g, H = gt.ops.derivs(atoms, space, torch.as_tensor(C, dtype=torch.float32, device=cuda), mode, 2)
    omode = "linear" if mode == "linear" else "L2"
    if mode == "resid":
        # residual mode == L2 with C_j = data - R_all + R_j for every atom j; here C is the residual.
        Rj = [oracle.radproj(f32(I[j:j + 1]), f32(x[j:j + 1]), f32(r[j:j + 1]), og) for j in range(len(I))]
        go = np.zeros_like(g.cpu().numpy())
        Ho = np.zeros_like(H.cpu().numpy())
        for j in range(len(I)):
            _, gj, Hj = oracle.derivs(f32(I[j:j + 1]), f32(x[j:j + 1]), f32(r[j:j + 1]), og, Rj[j] - C, "L2", 2)
            go[j], Ho[j] = gj[0], Hj[0]
        fo = None
    else:
        fo, go, Ho = oracle.derivs(f32(I), f32(x), f32(r), og, C, omode, 2)
    g, H = g.cpu().numpy(), H.cpu().numpy()
    print(f"derivs {mode} {dim}D: worst per-atom relative error gradient "
          f"{max(relerr(g[j], go[j]) for j in range(len(I))):.3e}, Hessian "
          f"{max(relerr(H[j], Ho[j]) for j in range(len(I))):.3e}")
    for j in range(len(I)):
        assert relerr(g[j], go[j]) < TOL, (j, relerr(g[j], go[j]))
        assert relerr(H[j], Ho[j]) < TOL, (j, relerr(H[j], Ho[j]))
        assert np.allclose(H[j], H[j].T, rtol=0, atol=1e-12 * np.abs(H[j]).max())
    if fo is not None:
        assert relerr(f.cpu().numpy(), fo) < TOL


@pytest.mark.parametrize("mode", ["L2", "linear", "resid"])
def test_derivs_3d(gt, oracle, cuda, mode):
    I, x, r, space, _ = _c2_case(gt, 24, 17, 64, seed=21)
    _derivs_case(gt, oracle, cuda, space, I, x, r, mode, 3)


@pytest.mark.parametrize("dim", [3, 2])
def test_forward_derivs_fused(gt, oracle, cuda, dim):
    """gt_forward_derivs = forward, minus data, residual-mode derivatives: bit for
    bit the separate calls, f = 1/2||resid||^2, and the gradient equals the
    oracle's L2 derivative against C = data - R + R_j (atom by atom)."""
    W = __import__("importlib").import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(31)
    if dim == 3:
        I, x, r, space, _ = _c2_case(gt, 12, 9, 40, seed=31)
    else:
        I, x, r = W.draw_iso2(7, rng)
        T, ortho, det = W.fan_2d(40, 64)
        space = gt.ProjSpace(orientations=T, ortho=ortho, detector=det)
    atoms = _atoms(gt, I, x, r, dim, cuda)
    data = gt.ops.forward(_atoms(gt, I * 1.1, x + 0.01, r, dim, cuda), space)
    resid, f, g, H = gt.ops.forward_derivs(atoms, space, data)
    R = gt.ops.forward(atoms, space)
    assert torch.equal(resid, R - data)
    _, g2, H2 = gt.ops.derivs(atoms, space, R - data, "resid", 2)
    assert torch.equal(g, g2) and torch.equal(H, H2)
    r64 = resid.double()
    assert abs(float(f) - 0.5 * float((r64 * r64).sum())) <= 1e-12 * float(f)
    # full-loss derivative of atom j = L2 derivative against data - R + R_j (NewtonGaussian.py:59-60)
    og = _ogeom(oracle, space)
    R64 = R.double().cpu().numpy()
    d64 = data.double().cpu().numpy()
    for j in (0, len(I) - 1):
        Rj = oracle.radproj(f32(I[j:j + 1]), f32(x[j:j + 1]), f32(np.asarray(r)[j:j + 1]), og)
        _, go, Ho = oracle.derivs(f32(I[j:j + 1]), f32(x[j:j + 1]), f32(np.asarray(r)[j:j + 1]), og,
                                  d64 - R64 + Rj, "L2", 2)
        assert relerr(g[j].cpu().numpy(), go[0]) < 1e-4 and relerr(H[j].cpu().numpy(), Ho[0]) < 1e-4
    # empty atom set: resid = -data, nothing else
    empty = _atoms(gt, I[:0], x[:0], np.asarray(r)[:0], dim, cuda)
    resid0, f0, g0, _ = gt.ops.forward_derivs(empty, space, data)
    assert torch.equal(resid0, -data) and g0.numel() == 0


@pytest.mark.parametrize("mode", ["L2", "linear"])
def test_derivs_2p5d(gt, oracle, cuda, mode):
    W = __import__("importlib").import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(8)
    I, x, r, _ = W.draw_aniso3(10, rng)
    ang = np.linspace(0, np.pi, 11, endpoint=False)
    space = gt.ProjSpace(angles=ang, detector=(gt.cell_centres(-1, 1, 48), gt.cell_centres(-1, 1, 40)))
    _derivs_case(gt, oracle, cuda, space, I, x, r, mode, 3)


@pytest.mark.parametrize("mode", ["L2", "linear"])
@pytest.mark.parametrize("iso", [True, False])
def test_derivs_2d(gt, oracle, cuda, mode, iso):
    W = __import__("importlib").import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(9)
    I, x, r = W.draw_iso2(12, rng)
    if not iso:
        r = np.concatenate([r, 1.3 * r, rng.normal(0, 1, (12, 1))], 1)
    T, ortho, det = W.fan_2d(60, 96)
    space = gt.ProjSpace(orientations=T, ortho=ortho, detector=det)
    _derivs_case(gt, oracle, cuda, space, I, x, r, mode, 2)


def test_derivs_3d_iso(gt, oracle, cuda):
    rng = np.random.default_rng(12)
    I = rng.uniform(0.5, 1.5, 8)
    x = rng.uniform(-0.5, 0.5, (8, 3))
    r = rng.uniform(2, 6, (8, 1))
    _, _, _, space, _ = _c2_case(gt, 2, 9, 48, seed=2)
    _derivs_case(gt, oracle, cuda, space, I, x, r, "L2", 3)


def test_volume_3d_2d(gt, oracle, cuda):
    I, x, r, space, _ = _c2_case(gt, 30, 3, 16, seed=30)
    vol = gt.VolSpace(box=(-1, 1), shape=(40, 33, 50), dim=3)
    u = gt.ops.volume(_atoms(gt, I, x, r, 3, cuda), vol).cpu().numpy()
    u64 = oracle.volproj(f32(I), f32(x), f32(r), [f32(c) for c in vol.coord_vectors])
    assert relerr(u, u64) < TOL
    W = __import__("importlib").import_module("2018_dictionary_tomography_amd.workloads")
    I2, x2, r2 = W.draw_iso2(20, np.random.default_rng(1))
    vol2 = gt.VolSpace(box=(-1, 1), shape=(70, 90), dim=2)
    u2 = gt.ops.volume(_atoms(gt, I2, x2, r2, 2, cuda), vol2).cpu().numpy()
    u264 = oracle.volproj(f32(I2), f32(x2), f32(r2), [f32(c) for c in vol2.coord_vectors])
    assert relerr(u2, u264) < TOL


def test_volume_3d_many_atoms(gt, oracle, cuda):
    """300 atoms (several 128-record staging chunks per brick) on a volume
    whose sizes are off the 4 x 4 x 64 voxel bricks; repeated calls give the
    same bits."""
    I, x, r, _, _ = _c2_case(gt, 300, 3, 16, seed=33)
    vol = gt.VolSpace(box=(-1, 1), shape=(31, 21, 70), dim=3)
    atoms = _atoms(gt, I, x, r, 3, cuda)
    first = gt.ops.volume(atoms, vol)
    u64 = oracle.volproj(f32(I), f32(x), f32(r), [f32(c) for c in vol.coord_vectors])
    assert relerr(first.cpu().numpy(), u64) < TOL
    for _ in range(2):
        assert torch.equal(gt.ops.volume(atoms, vol), first)


def test_dropin_module_functions(gt, oracle, cuda):
    """RadProj / L2derivs_RadProj / derivs_RadProj / VolProj with the reference conventions."""
    I, x, r, space, _ = _c2_case(gt, 5, 7, 40, seed=50)

    class Obj:
        pass

    atom = Obj()
    atom.I, atom.x, atom.r = I.astype(np.float32), x.astype(np.float32), r.astype(np.float32)
    Rad = Obj()
    Rad.range = space
    R = np.zeros(space.shape, dtype=np.float32)
    gt.RadProj(atom, Rad, R)
    og = _ogeom(oracle, space)
    R64 = oracle.radproj(f32(I), f32(x), f32(r), og)
    assert relerr(R, R64) < TOL
    C = f32(R64 * 0.5)
    f, Df, DDf = gt.L2derivs_RadProj(atom, Rad, C)
    fo, go, Ho = oracle.derivs(f32(I[:1]), f32(x[:1]), f32(r[:1]), og, C, "L2", 2)
    assert Df.shape == (10,) and DDf.shape == (10, 10)
    assert relerr(Df, go[0]) < TOL and relerr(DDf, Ho[0]) < TOL and abs(f - fo[0]) <= TOL * abs(fo[0])
    f, Df, DDf = gt.derivs_RadProj(atom, Rad, C)
    fo, go, Ho = oracle.derivs(f32(I[:1]), f32(x[:1]), f32(r[:1]), og, C, "linear", 2)
    assert relerr(Df, go[0]) < TOL and relerr(DDf, Ho[0]) < TOL
    # order 1 still returns a (P, P) third value, as GaussTomo.derivs indexes it
    f1, Df1, DDf1 = gt.derivs_RadProj(atom, Rad, C, order=1)
    assert DDf1.shape == (10, 10) and relerr(Df1, go[0]) < TOL
    y = [f32(c) for c in gt.VolSpace(box=(-1, 1), shape=20, dim=3).coord_vectors]
    u = np.zeros((20, 20, 20), dtype=np.float32)
    gt.VolProj(atom, gt.VolSpace(y), u)
    assert relerr(u, oracle.volproj(f32(I), f32(x), f32(r), y)) < TOL


def test_gausstomo_operator(gt, oracle, cuda):
    I, x, r, space, _ = _c2_case(gt, 9, 6, 32, seed=60)
    vol = gt.VolSpace(box=(-1, 1), shape=24, dim=3)
    Radon = gt.GaussTomo(gt.AtomSpace(3, isotropic=False), vol, space)
    atoms = _atoms(gt, I, x, r, 3, cuda)
    sino = Radon(atoms)
    og = _ogeom(oracle, space)
    R64 = oracle.radproj(f32(I), f32(x), f32(r), og)
    assert relerr(sino.asarray().cpu().numpy(), R64) < TOL
    f, df, ddf = Radon.L2_derivs(atoms, sino)
    assert f.shape == (9, 1) and df.shape == (9, 10) and ddf.shape == (9, 10, 10)
    f1, df1 = Radon.derivs(atoms, sino, order=1)
    assert f1.shape == (9, 1) and df1.shape == (9, 10)
    grad = Radon.gradient(atoms)(sino)
    assert np.allclose(grad, df1)
    u = Radon.discretise(atoms).asarray().cpu().numpy()
    assert relerr(u, oracle.volproj(f32(I), f32(x), f32(r), [f32(c) for c in vol.coord_vectors])) < TOL
