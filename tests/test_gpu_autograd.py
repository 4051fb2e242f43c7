"""Differentiation (5000 anisotropic atoms, 90 angles over [0, pi) about a seeded
random axis, 256x256 detector): the autograd backward (hand-written adjoint
kernel) against
  * the reference analytic derivative (oracle derivs_RadProj, FP64) on 200
    random parameters, tolerance 1e-5 relative to the largest of them, and
  * central finite differences of the loss in FP64 with h = 1e-3, computed
    footprint-locally (exact: a parameter of atom j only changes R_j):
    L(t+h) - L(t-h) = sum (R_j+ - R_j-) (r - R_j + (R_j+ + R_j-)/2).
    The plain h = 1e-3 difference carries O(h^2) truncation error (up to a
    few 1e-4 relative for centre coordinates of the narrowest atoms), so it
    is checked at 1e-2; the Richardson combination of h and h/2 (O(h^4)) is
    checked at 2e-5 of the largest gradient.
"""
import importlib

import numpy as np
import pytest
import torch

from conftest import f32

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3(gt, cuda):
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(3)
    N = 5000
    I, x, r, _ = W.draw_aniso3(N, rng)
    axis = W.random_axis(rng)
    T, ortho, det = W.half_turn_series(90, axis, det=256)
    space = gt.ProjSpace(orientations=T, ortho=ortho, detector=det)
    atoms_gt = gt.AtomElement(gt.AtomSpace(3, isotropic=False), x, r, I, device=cuda)
    clean = gt.ops.forward(atoms_gt, space)
    b = clean + torch.as_tensor(rng.normal(0, 0.01 * float(clean.max()), clean.shape), dtype=torch.float32,
                                device=cuda)
    # evaluation point: perturbed ground truth
    x1 = x + rng.normal(0, 0.005, x.shape)
    I1 = I * rng.uniform(0.9, 1.1, N)
    return dict(I=I1, x=x1, r=r, space=space, b=b, T=T, ortho=ortho, det=det, rng=rng)


def test_autograd_matches_analytic_and_fd(gt, oracle, cuda, c3):
    AG = importlib.import_module("2018_dictionary_tomography_amd.autograd")
    I = torch.tensor(c3["I"], dtype=torch.float32, device=cuda, requires_grad=True)
    x = torch.tensor(c3["x"], dtype=torch.float32, device=cuda, requires_grad=True)
    r = torch.tensor(c3["r"], dtype=torch.float32, device=cuda, requires_grad=True)
    loss = AG.l2_loss(I, x, r, c3["space"], c3["b"])
    loss.backward()
    G = torch.cat([I.grad[:, None], x.grad, r.grad], 1).double().cpu().numpy()
    with torch.no_grad():
        resid = (AG.project(I, x, r, c3["space"]) - c3["b"]).double().cpu().numpy()

    rng = np.random.default_rng(17)
    picks = rng.choice(5000 * 10, size=200, replace=False)
    atoms = sorted(set(int(p) // 10 for p in picks))
    g = oracle.Geometry(f32(c3["T"]), tuple(f32(o) for o in c3["ortho"]), [f32(p) for p in c3["det"]])
    Iv, xv, rv = f32(c3["I"]), f32(c3["x"]), f32(c3["r"])
    # analytic derivative of the reference (linear functional with C = residual)
    _, go, _ = oracle.derivs(Iv[atoms], xv[atoms], rv[atoms], g, resid, "linear", 1)
    ref = {a: go[k] for k, a in enumerate(atoms)}
    vals = np.array([G[p // 10, p % 10] for p in picks])
    refs = np.array([ref[p // 10][p % 10] for p in picks])
    scale = np.abs(refs).max()
    print(f"autograd at 5000 atoms x 90 x 256^2: max |gpu - fp64| / max |fp64| = "
          f"{np.abs(vals - refs).max() / scale:.3e} over {len(picks)} parameters")
    assert np.abs(vals - refs).max() < 1e-5 * scale

    # footprint-local central differences in FP64
    def dL(j, k, h):
        th = np.concatenate([Iv[j:j + 1], xv[j], rv[j]])
        def Rj(t):
            return oracle.radproj(t[:1], t[1:4][None], t[4:][None], g)
        e = np.zeros(10)
        e[k] = h
        Rp, Rm, R0 = Rj(th + e), Rj(th - e), Rj(th)
        return np.sum((Rp - Rm) * (resid - R0 + 0.5 * (Rp + Rm))) / (2 * h)

    sub = picks[:40]
    fd1 = np.array([dL(p // 10, p % 10, 1e-3) for p in sub])
    fd2 = np.array([dL(p // 10, p % 10, 5e-4) for p in sub])
    rich = (4 * fd2 - fd1) / 3
    gv = np.array([G[p // 10, p % 10] for p in sub])
    sc = np.abs(rich).max()
    assert np.abs(fd1 - gv).max() < 1e-2 * sc
    assert np.abs(rich - gv).max() < 2e-5 * sc


def test_autograd_gradcheck_small(gt, cuda):
    """torch.autograd.gradcheck-style check of the Function's plumbing on a small case
    (FP32 forward, so compare directional derivatives with a loose tolerance).  The
    weights and directions come from a seeded generator, and the error is measured
    against sum |g_i d_i|, which a random direction cannot cancel to near zero."""
    AG = importlib.import_module("2018_dictionary_tomography_amd.autograd")
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(0)
    I, x, r, _ = W.draw_aniso3(6, rng, rho=(0.06, 0.1))
    T, ortho, det = W.tilt_series(7, 60.0, det=40)
    space = gt.ProjSpace(orientations=T, ortho=ortho, detector=det)
    t = [torch.tensor(v, dtype=torch.float32, device=cuda, requires_grad=True) for v in (I, x, r)]
    gen = torch.Generator(device=cuda)
    gen.manual_seed(11)
    w = torch.randn(space.shape, device=cuda, generator=gen)
    (AG.project(*t, space) * w).sum().backward()
    worst = 0.0
    for _ in range(4):
        d = [torch.randn(v.shape, device=cuda, generator=gen) * 1e-3 for v in t]
        with torch.no_grad():
            fp = (AG.project(*[a + b for a, b in zip(t, d)], space) * w).double().sum()
            fm = (AG.project(*[a - b for a, b in zip(t, d)], space) * w).double().sum()
        lhs = float((fp - fm) / 2)
        rhs = float(sum((a.grad.double() * b.double()).sum() for a, b in zip(t, d)))
        scale = float(sum((a.grad.double() * b.double()).abs().sum() for a, b in zip(t, d)))
        worst = max(worst, abs(lhs - rhs) / scale)
    print(f"directional derivative vs central difference: {worst:.2e} of sum |g d|")
    assert worst < 2e-3
