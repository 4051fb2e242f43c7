"""Device Newton fitter and block GD vs the FP64 oracle restatements of
NewtonGaussian.linesearch / GD_lib.linesearch_block (same inputs, same
RandomState(1) permutations).

The GPU computes derivatives in FP32 (1e-7 relative) and solves / compares
energies in FP64; the oracle is FP64 throughout.  Accept/reject decisions and
hence trajectories agree as long as no decision is within ~1e-7 of a tie, so
the energy histories are compared entry-wise at 1e-5 relative to E[0].
"""
import importlib

import numpy as np
import pytest
import torch

from conftest import f32

pytestmark = pytest.mark.gpu


def _setup3(gt, n, nA, det, seed):
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(seed)
    I, x, r, _ = W.draw_aniso3(n, rng, rho=(0.05, 0.1))
    T, ortho, dv = W.tilt_series(nA, 60.0, det=det)
    space = gt.ProjSpace(orientations=T, ortho=ortho, detector=dv)
    return I, x, r, space, rng


def _oracle_geom(oracle, space):
    return oracle.Geometry(f32(space.orientations), tuple(f32(o) for o in space.ortho),
                           [f32(p) for p in space.detector])


def final_steps(trace, n, blocks):
    """Step sizes after an oracle fit, from its trace: h[t][j] of the last
    update of block t of atom j (1 for an atom never updated)."""
    h = np.ones((len(blocks), n))
    for t in trace:
        if not t.get("skip"):
            h[blocks.index(t.get("t", t.get("block"))), t["j"]] = t["h"]
    return h


def test_newton_3d_matches_oracle(gt, oracle, cuda):
    F = importlib.import_module("2018_dictionary_tomography_amd.fit")
    I, x, r, space, rng = _setup3(gt, 6, 11, 40, seed=5)
    aspace = gt.AtomSpace(3, isotropic=False)
    truth = gt.AtomElement(aspace, x, r, I, device=cuda)
    data = gt.ops.forward(truth, space)
    data = data + torch.as_tensor(rng.normal(0, 0.01 * float(data.max()), data.shape), dtype=torch.float32,
                                  device=cuda)
    x0 = x + rng.normal(0, 0.02, x.shape)
    I0 = I * 0.8
    start = gt.AtomElement(aspace, x0, r, I0, device=cuda)
    req, s = F.newton_fit(start, space, data, [4, 1], tol=1e-6)
    assert s.done == 1 and s.jj == 4 * 6
    Eg = req.Ehist[: s.jj + 1].cpu().numpy()
    og = _oracle_geom(oracle, space)
    Io, xo, ro, Eo = oracle.newton_linesearch(f32(I0), f32(x0), f32(r), og, data.double().cpu().numpy(), [4, 1],
                                              tol=1e-6)
    assert abs(Eg[0] - Eo[0]) < 1e-6 * Eo[0]
    assert np.all(np.diff(Eg) <= 0)
    assert np.abs(Eg - Eo).max() < 1e-5 * Eo[0]
    assert np.abs(req.x.cpu().numpy() - xo).max() < 1e-4
    # the same accept/reject decisions: equal update counts and step sizes
    tr = []
    oracle.newton_fit(f32(I0), f32(x0), f32(r), og, data.double().cpu().numpy(), [4, 1], tol=1e-6, trace=tr)
    assert len(tr) == s.jj
    assert np.array_equal(req.h.cpu().numpy(), final_steps(tr, 6, [0])[0])
    # energy bookkeeping: the tracked E equals 1/2||resid||^2 of the stored residual
    r64 = req.resid.double()
    assert abs(float(0.5 * (r64 * r64).sum()) - s.E) < 1e-9 * s.E
    # the stored residual equals a fresh forward of the fitted atoms minus the data
    fresh = gt.ops.forward(req, space) - req.data
    assert float((fresh - req.resid).abs().max()) < 1e-5 * float(req.data.abs().max())


def test_newton_2d_iso_matches_oracle(gt, oracle, cuda):
    F = importlib.import_module("2018_dictionary_tomography_amd.fit")
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(9)
    I, x, r = W.draw_iso2(5, rng, rho=(0.04, 0.08))
    T, ortho, det = W.fan_2d(60, 96)
    space = gt.ProjSpace(orientations=T, ortho=ortho, detector=det)
    aspace = gt.AtomSpace(2, isotropic=True)
    data = gt.ops.forward(gt.AtomElement(aspace, x, r, I, device=cuda), space)
    x0, I0 = x + rng.normal(0, 0.01, x.shape), I * 1.2
    req, s = F.newton_fit(gt.AtomElement(aspace, x0, r * 1.1, I0, device=cuda), space, data, [6, 1], tol=1e-7)
    Eg = req.Ehist[: s.jj + 1].cpu().numpy()
    Io, xo, ro, Eo = oracle.newton_linesearch(f32(I0), f32(x0), f32(r * 1.1), _oracle_geom(oracle, space),
                                              data.double().cpu().numpy(), [6, 1], tol=1e-7)
    n = min(len(Eg), len(Eo))
    assert np.all(np.diff(Eg) <= 0)
    assert np.abs(Eg[:n] - Eo[:n]).max() < 1e-5 * Eo[0]
    assert Eg[-1] < 1e-2 * Eg[0]


def test_newton_multi_slot_equals_single(gt, cuda):
    """Several fits in one launch give bit-identical results to each fit alone."""
    F = importlib.import_module("2018_dictionary_tomography_amd.fit")
    reqs = []
    for seed in (1, 2, 3):
        I, x, r, space, rng = _setup3(gt, 4 + seed, 7 + seed, 24 + 8 * seed, seed)
        aspace = gt.AtomSpace(3, isotropic=False)
        data = gt.ops.forward(gt.AtomElement(aspace, x, r, I, device=cuda), space)
        start = gt.AtomElement(aspace, x + 0.01, r, I * 0.9, device=cuda)
        reqs.append((start, space, data))
    alone = [F.newton_fit(a, s, d, [3, 1])[0] for a, s, d in reqs]
    mixed = [F.FitRequest(a, s, d, [3, 1]) for a, s, d in reqs]
    table = torch.as_tensor(np.frombuffer(b"".join(m.slot_bytes() for m in mixed), dtype=np.uint8).copy(), device=cuda)
    for _ in range(100):
        F.run_slots(table, 3, 3, 5)
        if all(s.done for s in F.read_slots(table, 3)):
            break
    for a, m in zip(alone, mixed):
        assert torch.equal(a.x, m.x) and torch.equal(a.r, m.r) and torch.equal(a.I, m.I)
        assert torch.equal(a.resid, m.resid)


def test_newton_group_size_does_not_change_result(gt, oracle, cuda):
    """A fit advanced by G = 1, 2, 3, 5 workgroups (angles shared between
    them, per-angle reductions in a fixed order), with the speculative commit
    or without it, is bit-identical to G = 1 and follows the oracle."""
    F = importlib.import_module("2018_dictionary_tomography_amd.fit")
    I, x, r, space, rng = _setup3(gt, 7, 23, 36, seed=12)
    aspace = gt.AtomSpace(3, isotropic=False)
    data = gt.ops.forward(gt.AtomElement(aspace, x, r, I, device=cuda), space)
    data = data + torch.as_tensor(rng.normal(0, 0.01 * float(data.max()), data.shape), dtype=torch.float32,
                                  device=cuda)
    x0 = x + rng.normal(0, 0.02, x.shape)
    start = gt.AtomElement(aspace, x0, r, I * 0.9, device=cuda)
    runs = []
    for G, staged in ((1, True), (2, True), (3, True), (5, True), (1, False), (3, False)):
        # without a staging slab: commit after the accept decision instead of restoring on reject
        req = F.FitRequest(start, space, data, [3, 1], tol=1e-6, stage=staged)
        assert (req.stage is None) == (not staged)
        table = torch.as_tensor(np.frombuffer(req.slot_bytes(), dtype=np.uint8).copy(), device=cuda)
        blocks = torch.as_tensor(F.block_table([G], [4]), device=cuda)
        for _ in range(100):
            F.run_blocks(table, 1, blocks, G, 3, 1 << 20)
            s = F.read_slots(table, 1)[0]
            if s.done:
                break
        assert s.done == 1 and s.updates == 3 * 7
        runs.append((req, s))
    (r1, s1) = runs[0]
    for req, s in runs[1:]:
        assert torch.equal(req.x, r1.x) and torch.equal(req.r, r1.r) and torch.equal(req.I, r1.I)
        assert torch.equal(req.resid, r1.resid) and torch.equal(req.Ehist, r1.Ehist)
        assert s.E == s1.E and s.accepted == s1.accepted
    Eg = r1.Ehist[: s1.jj + 1].cpu().numpy()
    Io, xo, ro, Eo = oracle.newton_linesearch(f32(I * 0.9), f32(x0), f32(r), _oracle_geom(oracle, space),
                                              data.double().cpu().numpy(), [3, 1], tol=1e-6)
    assert np.abs(Eg - Eo).max() < 1e-5 * Eo[0]


def test_grouped_launch_rejects_oversubscription(gt, cuda):
    F = importlib.import_module("2018_dictionary_tomography_amd.fit")
    cap = F.max_blocks()
    assert cap >= 1
    table = torch.zeros(F.SLOT_BYTES, dtype=torch.uint8, device=cuda)  # one inactive slot
    blocks = torch.full((cap + 1, 4), -1, dtype=torch.int32, device=cuda)
    with pytest.raises(gt.GTError, match="holds"):
        F.run_blocks(table, 1, blocks, cap + 1, 3, 1)
    F.run_blocks(table, 1, blocks, cap, 3, 1)  # idle blocks only: returns at once
    torch.cuda.synchronize()


def test_gd_block_2d_matches_oracle(gt, oracle, cuda):
    """GD_lib.linesearch_block at the 2D workload's size: 50 isotropic atoms
    (centres U[-0.8, 0.8]^2, radii U[0.02, 0.08], intensities U[0.5, 1.5]),
    180 angles over [0, pi), 128-pixel detector, data = forward + 1% noise,
    start = centres + N(0, 0.02); 3 outer iterations (450 block steps).
    Energies entry-wise within 1e-5 of E[0], the same step sizes."""
    F = importlib.import_module("2018_dictionary_tomography_amd.fit")
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(21)
    I, x, r = W.draw_iso2(50, rng)
    T, ortho, det = W.fan_2d(180, 128)
    space = gt.ProjSpace(orientations=T, ortho=ortho, detector=det)
    aspace = gt.AtomSpace(2, isotropic=True)
    data = gt.ops.forward(gt.AtomElement(aspace, x, r, I, device=cuda), space)
    data = data + torch.as_tensor(rng.normal(0, 0.01 * float(data.max()), data.shape), dtype=torch.float32,
                                  device=cuda)
    x0 = x + rng.normal(0, 0.02, x.shape)
    start = gt.AtomElement(aspace, x0, r, I, device=cuda)
    h = np.full((3, 50), 1e-2)
    _, Eg = F.gd_linesearch_block(start, data, [3, 1], space, h=h)
    tr = []
    _, _, _, Eo = oracle.gd_linesearch_block(f32(I), f32(x0), f32(r), _oracle_geom(oracle, space),
                                             data.double().cpu().numpy(), [3, 1], trace=tr)
    assert np.all(np.diff(Eg) <= 0)
    assert len(Eg) == len(Eo) == 1 + 3 * 50 * 3
    assert np.abs(Eg - Eo).max() < 1e-5 * Eo[0]
    ho = final_steps(tr, 50, list("xrI"))
    assert np.allclose(h, ho, rtol=1e-12)
