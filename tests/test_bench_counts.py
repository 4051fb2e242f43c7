"""The in-support pair counts behind the benchmark's roofline figures
(bench.support_pairs_columns, bench.ellipsoid_lattice_count: roots of one
quadratic per detector column / grid column) against brute-force evaluation
of every (atom, pixel) or (atom, grid point) pair, on the CPU."""
import importlib

import numpy as np
import torch

import bench


def _U(r):
    U = np.zeros((r.shape[0], 3, 3))
    for k, (i, j) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]):
        U[:, i, j] = r[:, k]
    return U


def test_projector_support_count_matches_brute_force():
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(4)
    I, x, r, _ = W.draw_aniso3(12, rng)
    T, (w0, w1), (p0, p1) = W.tilt_series(7, 60.0, det=40)
    U = _U(r)
    P = U.transpose(0, 2, 1) @ U
    brute = 0
    G0, G1 = np.meshgrid(p0, p1, indexing="ij")
    for a in range(T.shape[0]):
        for j in range(12):
            z = P[j] @ T[a]
            Q = P[j] - np.outer(z, z) / (T[a] @ z)
            Y = G0[..., None] * w0[a] + G1[..., None] * w1[a] - x[j]
            m = np.einsum("...i,ij,...j->...", Y, Q, Y)
            brute += int((m <= 1.0).sum())
    assert brute > 0
    assert bench.support_pairs_columns(torch, r, x, T, w0, w1, p0, p1) == brute


def test_lattice_counts_match_brute_force():
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(5)
    _, x, r, _ = W.draw_aniso3(6, rng)
    U = _U(r * np.sqrt(60.0) / 12.0)  # ellipsoids a few grid steps across
    k = np.arange(-16, 16, dtype=np.float64)
    K = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3)
    # Fourier form: |U^T k|^2 < 30 about the origin
    M = U @ U.transpose(0, 2, 1)
    brute = sum(int((np.einsum("pi,ij,pj->p", K, Mj, K) < 30.0).sum()) for Mj in M)
    assert brute > 0
    assert bench.ellipsoid_lattice_count(torch, M, np.zeros((6, 3)), k, k, k, 30.0, strict=True) == brute
    # volume form: |U (y - x)|^2 <= 1 about the atom centres, on cell centres
    c = (np.arange(24) + 0.5) / 12 - 1.0
    Y = np.stack(np.meshgrid(c, c, c, indexing="ij"), -1).reshape(-1, 3)
    U2 = _U(r)
    M2 = U2.transpose(0, 2, 1) @ U2
    brute = sum(int((np.einsum("pi,ij,pj->p", Y - x[j], M2[j], Y - x[j]) <= 1.0).sum()) for j in range(6))
    assert brute > 0
    assert bench.ellipsoid_lattice_count(torch, M2, x, c, c, c, 1.0, strict=False) == brute
