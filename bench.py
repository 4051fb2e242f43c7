#!/usr/bin/env python3
"""Benchmark of the Gaussian-atom projector hot path on one MI355X.

Workload (the 3D anisotropic configuration): 1000 atoms, centres U[-0.7,0.7]^3,
intensities U[0.5,1.5], Cholesky precision factors with diagonal 1/rho,
rho ~ U[0.02,0.1], off-diagonal N(0,5); 61-angle tilt series over +-60 deg
about y onto a 128x128 detector on [-1,1]^2; data = forward + 1% noise.

One step = forward projection of the current atoms + residual + per-atom
gradient and 10x10 Hessian block of 1/2||A(theta)-b||^2 (the derivative pass
of one Newton sweep).  value = dense-equivalent atom.ray evaluations per second
= 2 * N * nA * n0 * n1 / step time (one forward and one derivative sweep).

Prints one JSON line.  Flags: --gpus N (only 1 is supported), --steps K,
--warmup W, --cpu-atoms M (atoms in the CPU-baseline sample, 0 disables it).
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Frozen algorithmic work per in-support (atom, pixel) pair (FP32 flops, FMA = 2).
FLOPS_FWD = 11
FLOPS_L2 = 85
PEAK_FP32_TFLOPS = 157.3      # MI355X published FP32 (vector, packed / matrix)
PEAK_FP32_NONPACKED = 78.65   # 256 CU x 4 SIMD x 16 lanes x 2 flop x 2.4 GHz
PEAK_TRANS_PER_S = 256 * 4 * 16 * 2.4e9 / 4  # quarter-rate transcendental
PEAK_HBM_GBS = 8000.0
# moments kernel, 1000 atoms x 61 angles x 128^2: 38.0 MiB read + 7.4 MiB written per launch
MOMENTS_HBM_BYTES = int((38.0 + 7.4) * 2**20)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-atoms", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--atoms", type=int, default=1000)
    ap.add_argument("--fit-iters", type=int, default=20, help="Newton outer iterations of the fit leg (0 skips it)")
    ap.add_argument("--serve-requests", type=int, default=256, help="requests of the serving leg (0 skips it)")
    ap.add_argument("--serve-nmax", type=int, default=500, help="upper end of the log-uniform atom count")
    ap.add_argument("--plain-requests", type=int, default=6, help="requests timed one at a time")
    ap.add_argument("--autograd-reps", type=int, default=3,
                    help="timed forward+backward passes of the autograd leg (0 skips it)")
    ap.add_argument("--fourier-reps", type=int, default=5, help="timed evaluations of the Fourier leg (0 skips it)")
    ap.add_argument("--extras-reps", type=int, default=5,
                    help="timed calls of the point-spread / Fourier-plane / KL-moment leg (0 skips it)")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving leg and print its JSON")
    ap.add_argument("--skip-checks", action="store_true",
                    help="skip the FP64 parity spot check and the in-support pair count (counter runs)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="before the steps, copy a buffer of this many MiB on the device once: a plain copy kernel "
                         "of known size against which HBM byte counters are scaled")
    return ap.parse_args()


def fit_leg(torch, gtp, space, data, start, iters):
    """20-iteration damped Newton fit of the 1000-atom workload on the device
    (FP64 block solves and energy deltas), one request alone."""
    F = importlib.import_module("2018_dictionary_tomography_amd.fit")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    group = F.default_group(space.shape[0])
    req, s = F.newton_fit(start, space, data, [iters, 1], record=True, budget=4000, group=group)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    E = req.Ehist[: s.jj + 1].cpu().numpy()
    fresh = gtp.ops.forward(req, space)
    drift = float((req.resid - (fresh - req.data)).abs().max() / fresh.abs().max())
    return {
        "outer_iterations": int(s.it), "atom_updates": int(s.updates), "accepted": int(s.accepted),
        "seconds": t, "updates_per_s": s.updates / t, "E0": float(E[0]), "E_final": float(s.E),
        "monotone": bool(np.all(np.diff(E) <= 0)), "residual_drift": drift, "workgroups": group,
        "note": "device-resident fit advanced by a group of workgroups sharing the angles; FP64 solves and "
                "energy deltas",
    }


def serving_leg(torch, args, dev):
    """Poisson stream at 200 req/s with the serving workload's distributions
    (angles U{30..120} over +-60 deg, detector side from {64,96,128,192,256},
    30-iteration Newton), atom count log-uniform in [50, serve_nmax].  Served
    = continuous batching into the slot table with graph-replayed grouped
    launches; plain = the same fits one request at a time, each with the whole
    device (its default workgroup group), on a seeded random subset.  Request
    payloads (start atoms, data) are generated before either is timed."""
    Sv = importlib.import_module("2018_dictionary_tomography_amd.serving")
    specs = Sv.c5_stream(args.serve_requests, rate=200.0, seed=args.seed, n_range=(50, args.serve_nmax))
    log = (lambda m: print(m, file=sys.stderr, flush=True)) if args.serve_only else None
    t0 = time.perf_counter()
    payloads = Sv.prepare_payloads(specs, dev)
    torch.cuda.synchronize()
    prep_s = time.perf_counter() - t0
    if log:
        log(f"payloads of {len(specs)} requests built in {prep_s:.1f}s")
    server = Sv.FitServer(dev, keep_results=True)
    st = server.run(specs, log=log, payloads=payloads)
    served = st.summary()
    pick = sorted(np.random.default_rng(args.seed + 1).choice(len(specs), min(args.plain_requests, len(specs)),
                                                              replace=False).tolist())
    times, alone = [], {}
    for i in pick:
        t, a = Sv.plain_fits([specs[i]], dev, keep_results=True, payloads=payloads)
        times += t
        alone.update(a)
        if log:
            log(f"plain rid={i} N={specs[i].N} nA={specs[i].nA} S={specs[i].S} service={t[0]:.2f}s")
    same = all(torch.equal(alone[k][j], st.results[k][j]) for k in alone for j in range(3)) and \
        all(alone[k][3] == st.results[k][3] for k in alone)
    plain_fps = 1.0 / float(np.mean(times))
    return {
        "requests": len(specs), "rate_per_s": 200.0, "atoms_range": [50, args.serve_nmax],
        "iterations": 30, "served": served, "client_payload_s": prep_s,
        "plain": {"requests_timed": pick, "service_s": times, "fits_per_s": plain_fps,
                  "note": "same device fitter, one request at a time, whole device (default group)"},
        "speedup_fits_per_s": served["fits_per_s"] / plain_fps,
        "alone_equals_mixed": bool(same),
        "max_residual_drift": max(a[6] for a in alone.values()),
    }


def autograd_leg(torch, gtp, dev, reps):
    """Differentiation workload: 5000 atoms drawn as the 3D workload, 90 angles
    over [0, pi) about a seeded random unit axis, 256x256 detector; loss
    1/2||A(theta) - b||^2 with b = ground truth + 1% noise, evaluated at a
    perturbed ground truth; one timed pass = forward + hand-written backward
    (the adjoint kernel), no stored residual x Jacobian intermediates."""
    AG = importlib.import_module("2018_dictionary_tomography_amd.autograd")
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(3)
    N, nA, S = 5000, 90, 256
    I, x, r, _ = W.draw_aniso3(N, rng)
    T, ortho, det = W.half_turn_series(nA, W.random_axis(rng), det=S)
    space = gtp.ProjSpace(orientations=T, ortho=ortho, detector=det)
    clean = gtp.ops.forward(gtp.AtomElement(gtp.AtomSpace(3, isotropic=False), x, r, I, device=dev), space)
    b = clean + torch.as_tensor(rng.normal(0, 0.01 * float(clean.max()), clean.shape), dtype=torch.float32,
                                device=dev)
    t = [torch.tensor(v, dtype=torch.float32, device=dev, requires_grad=True)
         for v in (I * rng.uniform(0.9, 1.1, N), x + rng.normal(0, 0.005, x.shape), r)]

    def one():
        for v in t:
            v.grad = None
        loss = AG.l2_loss(*t, space, b)
        loss.backward()
        return loss

    one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        loss = one()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    return {"workload": "5000 atoms, 90 angles over [0, pi) about a random axis, 256x256 detector",
            "ms_per_forward_backward": ms, "evals_per_s": 2.0 * N * nA * S * S / (ms * 1e-3),
            "loss": float(loss.detach()), "reps": reps}


def fourier_leg(torch, gtp, dev, reps):
    """Fourier workload: analytic FT of 2000 atoms drawn as the 3D workload on
    the integer grid [-64, 64)^3 in complex64; error against an FP64
    restatement of the same closed form (torch, independent of the library) on
    4096 sampled grid points."""
    Fo = importlib.import_module("2018_dictionary_tomography_amd.fourier")
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(44)
    I, x, r, _ = W.draw_aniso3(2000, rng)
    nI, nx, nr = Fo.analytic_gauss_fourier(I, x, r * np.sqrt(60.0))
    k = [np.arange(-64, 64, dtype=np.float64)] * 3
    atoms = Fo.FTAtoms(nI, nx, nr, dev)
    grid = Fo.DeviceKGrid(k, dev)
    out = Fo.evaluate(atoms, grid)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        Fo.evaluate(atoms, grid, out=out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    # FP64 restatement: sum I exp(-|U^T k|^2 / 2) [|U^T k|^2 < 30] exp(-i x.k), U upper triangular
    idx = np.random.default_rng(0).integers(0, 128, (4096, 3))
    K = torch.as_tensor(np.stack([k[0][idx[:, 0]], k[1][idx[:, 1]], k[2][idx[:, 2]]], 1), device=dev)
    d = lambda a: torch.as_tensor(np.float32(a).astype(np.float64), device=dev)
    I64, x64, r64 = d(nI), d(nx), d(nr)
    U = torch.zeros((len(nI), 3, 3), dtype=torch.float64, device=dev)
    for s_, (i, j) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]):
        U[:, i, j] = r64[:, s_]
    ref = torch.zeros(4096, dtype=torch.complex128, device=dev)
    for c in range(0, len(nI), 250):
        Ut = torch.einsum("nli,pl->npi", U[c:c + 250], K)
        norm = (Ut * Ut).sum(-1)
        amp = I64[c:c + 250, None] * torch.exp(-0.5 * norm) * (norm < 30.0)
        ph = x64[c:c + 250] @ K.T
        ref += (amp * torch.exp(-1j * ph)).sum(0)
    got = out[idx[:, 0], idx[:, 1], idx[:, 2]].to(torch.complex128)
    err = float((got - ref).abs().max() / ref.abs().max())
    pairs = 2000.0 * 128 ** 3
    return {"workload": "2000 atoms on the integer grid [-64, 64)^3, complex64", "ms_per_evaluation": ms,
            "atom_k_pairs_per_s": pairs / (ms * 1e-3), "rel_err_vs_fp64_sampled": err, "reps": reps}


def _timed(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def _kernel_ms(torch, kernel_id, fn, reps):
    """Average duration (ms) of the launches of one library kernel id over
    reps calls of fn, from the library's HIP-event hooks."""
    L = importlib.import_module("2018_dictionary_tomography_amd._lib")
    torch.cuda.synchronize()
    L.profile_enable(True)
    L.profile_read(kernel_id)
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    L.profile_enable(False)
    tot, cnt = L.profile_read(kernel_id)
    return tot / max(cnt, 1)


def extras_leg(torch, gtp, dev, reps):
    """Rows the reference times nowhere, on the 3D workload's geometry (61
    angles +-60 deg about y, 128x128): the volume render of 1000 atoms drawn
    as the 3D workload on 128^3, the point-spread projector (forward
    and L2 gradient + Hessian blocks of 1000 atoms drawn as the 3D
    workload), the Fourier-slice planes of 2000 FT atoms and their complex
    L2 derivatives, and the KL initialiser's sinogram moments on a 90 x
    256^2 sinogram (HBM-bound: 8 bytes read per pixel)."""
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    psf = importlib.import_module("2018_dictionary_tomography_amd.psf")
    Fo = importlib.import_module("2018_dictionary_tomography_amd.fourier")
    KL = importlib.import_module("2018_dictionary_tomography_amd.kl")
    rng = np.random.default_rng(45)
    T, ortho, det = W.tilt_series(61, 60.0, det=128)
    space = gtp.ProjSpace(orientations=T, ortho=ortho, detector=det)
    rays = 61 * 128 * 128
    out = {}
    I, x, r, _ = W.draw_aniso3(1000, rng)
    atoms = gtp.AtomElement(gtp.AtomSpace(3, isotropic=False), x, r, I, device=dev)
    P = psf.forward(atoms, space)
    C = P * 0.9
    ms_f = _timed(torch, lambda: psf.forward(atoms, space, out=P), reps)
    ms_d = _timed(torch, lambda: psf.derivs(atoms, space, C, "L2", 2), reps)
    vol = gtp.VolSpace(box=(-1, 1), shape=64, dim=3)
    u = psf.volume(atoms, vol)
    ms_v = _timed(torch, lambda: psf.volume(atoms, vol, out=u), reps)
    vol128 = gtp.VolSpace(box=(-1, 1), shape=128, dim=3)
    u128 = gtp.ops.volume(atoms, vol128)
    ms_pv = _timed(torch, lambda: gtp.ops.volume(atoms, vol128, out=u128), reps)
    L = importlib.import_module("2018_dictionary_tomography_amd._lib")
    k_pv = _kernel_ms(torch, L.GT_K_VOLUME, lambda: gtp.ops.volume(atoms, vol128, out=u128), reps)
    out["volume"] = {"workload": "VolProj of the same 1000 atoms on 128^3 voxel centres",
                     "ms": ms_pv, "kernel_ms": k_pv, "note": "ms: wall clock per call (host-bound); kernel_ms: "
                     "the render kernel alone (HIP events)"}
    out["psf"] = {"workload": "1000 atoms, 61 x 128^2, pixel-integrated; voxel-integrated render on 64^3",
                  "forward_ms": ms_f, "derivs_L2_ms": ms_d, "volume_ms": ms_v,
                  "forward_atom_pixel_per_s": 1000 * rays / (ms_f * 1e-3),
                  "derivs_atom_pixel_per_s": 1000 * rays / (ms_d * 1e-3)}
    I, x, r, _ = W.draw_aniso3(2000, rng)
    nI, nx, nr = Fo.analytic_gauss_fourier(I, x, r * np.sqrt(60.0))
    ft = Fo.FTAtoms(nI, nx, nr, dev)
    vol = Fo.GaussFTVolume(gtp.AtomSpace(3, isotropic=False), space)
    planes = Fo.evaluate(ft, vol.grid)
    ms_p = _timed(torch, lambda: Fo.evaluate(ft, vol.grid, out=planes), reps)
    ms_k = _timed(torch, lambda: Fo.derivs_all(ft, vol.grid, planes * 0.9), reps)
    out["fourier_planes"] = {"workload": "2000 FT atoms on the 61 Fourier-slice planes of 128^2 samples",
                             "evaluate_ms": ms_p, "derivs_ms": ms_k,
                             "atom_k_pairs_per_s": 2000 * rays / (ms_p * 1e-3)}
    T3, o3, d3 = W.tilt_series(90, 90.0, det=256)
    sp3 = gtp.ProjSpace(orientations=T3, ortho=o3, detector=d3)
    S = torch.randn(sp3.shape, device=dev)
    Q = torch.rand(sp3.shape, device=dev)
    m = KL.sino_moments_device(S, sp3, Q, 0, 0.5, 0.5)
    ms_m = _timed(torch, lambda: KL.sino_moments_device(S, sp3, Q, 0, 0.5, 0.5, out=m), max(reps, 20))
    nbytes = 8.0 * S.numel()
    out["kl_moments"] = {"workload": "90 x 256^2 sinogram, f = max(0, a S + b Q), 5 FP64 moments per angle",
                         "ms": ms_m, "GB_per_s": nbytes / (ms_m * 1e-3) / 1e9,
                         "hbm_frac": nbytes / (ms_m * 1e-3) / 1e9 / PEAK_HBM_GBS,
                         "bytes_per_launch": nbytes}
    return out


def support_pairs(torch, I, x, r, T, w0, w1, p0, p1):
    """Exact count of (atom, pixel) pairs with m <= 1 (in-support) and the FP64
    sinogram, both from the closed form of the line integral, on the GPU in
    FP64 (elementwise torch, independent of the library):
    m(p) = (Gp - x)^T Q (Gp - x), Q = U^T (Id - t t^T/|t|^2) U, t = U T,
    R(p) = sum over atoms of I/|t| exp(-30 m) [m <= 1]."""
    dev = torch.device("cuda")
    d = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
    r, x, I = d(r), d(x), d(I)
    N = r.shape[0]
    U = torch.zeros((N, 3, 3), dtype=torch.float64, device=dev)
    for k, (i, j) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]):
        U[:, i, j] = r[:, k]
    P0, P1 = torch.meshgrid(d(p0), d(p1), indexing="ij")
    total = 0
    R64 = torch.zeros((T.shape[0],) + P0.shape, dtype=torch.float64, device=dev)
    P = U.transpose(1, 2) @ U
    for a in range(T.shape[0]):
        Ta, g0, g1 = d(T[a]), d(w0[a]), d(w1[a])
        t = U @ Ta
        s = (t * t).sum(1)
        z = P @ Ta
        Q = P - z[:, :, None] * z[:, None, :] / s[:, None, None]
        alpha = I / s.sqrt()
        for c in range(0, N, 64):
            Y = [P0[None] * g0[i] + P1[None] * g1[i] - x[c:c + 64, i, None, None] for i in range(3)]
            q = lambda i, j: Q[c:c + 64, i, j, None, None]
            m = (q(0, 0) * Y[0] * Y[0] + q(1, 1) * Y[1] * Y[1] + q(2, 2) * Y[2] * Y[2]
                 + 2 * (q(0, 1) * Y[0] * Y[1] + q(1, 2) * Y[1] * Y[2] + q(0, 2) * Y[0] * Y[2]))
            inside = m <= 1.0
            total += int(inside.sum().item())
            R64[a] += (alpha[c:c + 64, None, None] * torch.exp(-30.0 * m) * inside).sum(0)
    return total, R64


def cpu_baseline(I, x, r, T, ortho, det, data, n_atoms):
    """Oracle (FP64 C restatement of the reference's dense loops, OpenMP) on a
    bounded sample: forward + L2 derivatives of the first n_atoms atoms over the
    full geometry.  Rate in dense atom.ray evaluations per second."""
    from oracle import oracle as O

    O.build()
    g = O.Geometry(T, ortho, det)
    rays = T.shape[0] * det[0].size * det[1].size
    sl = slice(0, n_atoms)
    t0 = time.perf_counter()
    R = O.radproj(I[sl], x[sl], r[sl], g)
    t1 = time.perf_counter()
    O.derivs(I[sl], x[sl], r[sl], g, data - R, "L2", 2)
    t2 = time.perf_counter()
    return {
        "value": 2.0 * n_atoms * rays / (t2 - t0),
        "unit": "atom-ray evals/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": f"{n_atoms} of the atoms, full {T.shape[0]}x{det[0].size}x{det[1].size} geometry: "
                  f"dense FP64 forward ({t1 - t0:.2f} s) + L2 gradient/Hessian ({t2 - t1:.2f} s)",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "multi-GPU sharding is described in DESIGN.md but not run", file=sys.stderr)
        return 2
    import torch

    gtp = importlib.import_module("2018_dictionary_tomography_amd")
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    from importlib import import_module

    L = import_module("2018_dictionary_tomography_amd._lib")
    dev = torch.device("cuda", 0)
    if args.serve_only:
        print(json.dumps({"serving": serving_leg(torch, args, dev)}))
        return 0
    rng = np.random.default_rng(args.seed)
    N = args.atoms
    I, x, r, _ = W.draw_aniso3(N, rng)
    T, ortho, det = W.tilt_series(61, 60.0, det=128)
    space = gtp.ProjSpace(orientations=T, ortho=ortho, detector=det)
    aspace = gtp.AtomSpace(3, isotropic=False)
    truth = gtp.AtomElement(aspace, x, r, I, device=dev)
    clean = gtp.ops.forward(truth, space)
    noise = torch.as_tensor(rng.normal(0, 0.01 * float(clean.max()), clean.shape), dtype=torch.float32, device=dev)
    data = clean + noise
    x_cur = x + rng.normal(0, 0.01, x.shape)
    cur = gtp.AtomElement(aspace, x_cur, r, I, device=dev)
    R = torch.empty_like(data)

    def step():
        # forward, residual and every atom's gradient / Hessian block in one call
        return gtp.ops.forward_derivs(cur, space, data, 2, out=R, want_f=False)

    if args.calib_copy_mib > 0:
        src = torch.ones(args.calib_copy_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    L.profile_enable(True)
    for k in range(7):
        L.profile_read(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.profile_enable(False)
    kt = {name: L.profile_read(kid) for name, kid in
          (("setup", L.GT_K_SETUP), ("forward", L.GT_K_FORWARD), ("moments", L.GT_K_MOMENTS),
           ("finish", L.GT_K_FINISH))}
    ms = (t1 - t0) * 1e3 / args.steps
    rays = 61 * 128 * 128
    value = 2.0 * N * rays / (ms * 1e-3)

    # in-support pair count and forward error against an FP64 restatement of the sinogram
    if args.skip_checks:
        P_in, err_fwd = 0, None
    else:
        f64 = lambda a: np.float32(a).astype(np.float64)
        P_in, R64 = support_pairs(torch, f64(I), f64(x_cur), f64(r), f64(T), f64(ortho[0]), f64(ortho[1]),
                                  f64(det[0]), f64(det[1]))
        Rg = gtp.ops.forward(cur, space).double()
        err_fwd = float((Rg - R64).abs().max() / R64.abs().max())
    mom_ms = kt["moments"][0] / max(kt["moments"][1], 1)
    fwd_ms = kt["forward"][0] / max(kt["forward"][1], 1)
    achieved = P_in * FLOPS_L2 / (mom_ms * 1e-3) / 1e12
    roof = {
        "bound": "mfma",
        "engine": "fp32-valu (no MFMA: exp-bound per-pixel work; 'mfma' = compute-bound class)",
        "kernel": "moments_kernel<2,RESID> (per-atom gradient/Hessian pixel moments)",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
        "frac_nonpacked": round(achieved / PEAK_FP32_NONPACKED, 4),
        "trans_frac": round(P_in / (mom_ms * 1e-3) / PEAK_TRANS_PER_S, 4),
        "avg_kernel_ms": round(mom_ms, 5),
        "in_support_pairs": P_in,
        "flops_per_pair": FLOPS_L2,
        # HBM bytes per launch of the moments kernel from the two rocprofv3 counter
        # passes (FETCH_SIZE x 2.0 and WRITE_SIZE x 1.0, both scaled against a
        # 256 MiB copy), recorded with the command in the profiles directory
        "traffic": MOMENTS_HBM_BYTES,
        "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, profiles/step_1000atoms_61x128x128_traffic.md",
        "forward_kernel": {"avg_ms": round(fwd_ms, 5), "achieved_tflops": round(P_in * FLOPS_FWD / (fwd_ms * 1e-3) / 1e12, 3),
                           "trans_frac": round(P_in / (fwd_ms * 1e-3) / PEAK_TRANS_PER_S, 4)},
        "kernel_ms": {k: round(v[0] / max(v[1], 1), 5) for k, v in kt.items()},
    }
    fit = fit_leg(torch, gtp, space, data, cur, args.fit_iters) if args.fit_iters > 0 else None
    serving = serving_leg(torch, args, dev) if args.serve_requests > 0 else None
    autograd = autograd_leg(torch, gtp, dev, args.autograd_reps) if args.autograd_reps > 0 else None
    fourier = fourier_leg(torch, gtp, dev, args.fourier_reps) if args.fourier_reps > 0 else None
    extras = extras_leg(torch, gtp, dev, args.extras_reps) if args.extras_reps > 0 else None
    cpu = None
    if args.cpu_atoms > 0:
        cpu = cpu_baseline(np.float32(I).astype(float), np.float32(x_cur).astype(float),
                           np.float32(r).astype(float), np.float32(T).astype(float),
                           tuple(np.float32(o).astype(float) for o in ortho),
                           [np.float32(p).astype(float) for p in det],
                           (data.cpu().numpy()).astype(float), args.cpu_atoms)
    out = {
        "metric": "fwd+grad atom-ray evals/s",
        "value": value,
        "unit": "atom-ray evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "3D anisotropic, 1000 atoms, 61 angles +-60 deg about y, 128x128 detector, "
                               "forward + per-atom gradient and 10x10 Hessian blocks",
                   "atoms": N, "angles": 61, "detector": [128, 128], "seed": args.seed},
        "rel_err_fwd_vs_fp64": err_fwd,
        "roofline": roof,
        "cpu_baseline": cpu,
        "newton_fit": fit,
        "serving": serving,
        "autograd": autograd,
        "fourier": fourier,
        "extras": extras,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
