"""Bookkeeping of the greedy (guess) fitting path, fit.greedy, on the CPU
with a stand-in projector and stage: the order and arguments of the guess
calls, the stage sizes and iteration counts (including the last pass of a
three-entry max_iter), the replaced energy entries (NewtonGaussian.py:38-50),
and the write-back of each stage's atoms."""
import importlib

import numpy as np
import torch

F = importlib.import_module("2018_dictionary_tomography_amd.fit")
S = importlib.import_module("2018_dictionary_tomography_amd.spaces")


def _setup(N=3):
    space = S.ProjSpace(orientations=np.array([[0.0, 1.0], [1.0, 0.0]]), ortho=(np.array([[1.0, 0.0], [0.0, 1.0]]),),
                        detector=[np.linspace(-1, 1, 4)])
    aspace = S.AtomSpace(2, isotropic=True)
    recon = S.AtomElement(aspace, np.zeros((N, 2)), np.ones((N, 1)), np.zeros(N), device="cpu")
    data = torch.arange(8, dtype=torch.float32).reshape(2, 4) / 8
    return space, recon, data


def _forward(atoms, space):
    # a linear stand-in projector: every pixel holds the sum of the intensities
    return torch.full(space.shape, float(atoms.I.sum()), dtype=torch.float32)


class _Stage:
    kernels = F.Kernels(_forward, None, None)

    def __init__(self):
        self.calls = []

    def __call__(self, sub, mi):
        n = len(sub)
        self.calls.append((n, list(mi)))
        sub.I.mul_(0.5)  # the stage moves the atoms; the greedy loop must keep them
        e0 = float(10 * len(self.calls))
        return np.array([e0, e0 - 1, e0 - 2]), np.array([e0, e0 - 0.5, e0 - 1.5])


def test_guess_calls_stage_sizes_and_energy_entries():
    space, recon, data = _setup()
    seen = []

    def guess(res, atom):
        n = len(seen) + 1
        # the residual handed over is data - R(recon[:n]) with the current atoms
        seen.append(res.asarray().clone())
        atom.I.fill_(float(n))
        return atom

    stage = _Stage()
    out, E, Fh = F.greedy(recon, data, [2, 1], None, space, guess, stage)
    assert out is recon
    assert [c[0] for c in stage.calls] == [1, 2, 3] and all(c[1] == [2, 1] for c in stage.calls)
    # guess n saw the residual of atoms 0..n-1 after the earlier stages (I halved each stage)
    assert torch.allclose(seen[0], data - 0.0)
    assert torch.allclose(seen[1], data - 0.5)              # atom 0: 1 * 0.5
    assert torch.allclose(seen[2], data - (0.25 + 1.0))     # atom 0 halved twice, atom 1: 2 * 0.5
    # energies: each placement replaces the last entry; stage histories follow without their start
    e = lambda tot: 0.5 * float(((torch.full((2, 4), tot) - data).double() ** 2).sum())
    assert np.allclose(E, [e(1.0), 9.0, e(0.5 + 2.0), 19.0, e(0.25 + 1.0 + 3.0), 29.0, 28.0])
    assert np.allclose(Fh, [e(1.0), 9.5, e(0.5 + 2.0), 19.5, e(0.25 + 1.0 + 3.0), 29.5, 28.5])
    assert np.allclose(recon.I.numpy(), [0.125, 0.5, 1.5])


def test_three_entry_max_iter_adds_a_last_pass():
    space, recon, data = _setup(2)
    stage = _Stage()
    _, E, _ = F.greedy(recon, data, [3, 1, 5], None, space, lambda res, a: a, stage)
    assert stage.calls == [(1, [3, 1]), (2, [3, 1]), (2, [5, 1])]
    # the last pass starts from the previous stage's final entry (no replacement)
    assert E.size == 1 + 2 + 2 + 2 and np.isclose(E[4], 18.0) and np.isclose(E[5], 29.0)


def test_scalar_max_iter():
    space, recon, data = _setup(2)
    stage = _Stage()
    F.greedy(recon, data, 4, None, space, lambda res, a: a, stage)
    assert stage.calls == [(1, [4, 1]), (2, [4, 1])]
