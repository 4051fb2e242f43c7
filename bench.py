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
FLOPS_FWD = 11   # forward: quadratic (Horner) 9, amplitude 1, accumulate 1; + 1 exp
FLOPS_ADJ = 25   # adjoint (linear functional): value 11 + 7 residual-weighted moment FMAs; + 1 exp
FLOPS_L2 = 85    # L2 gradient + Hessian: value 11, 7 first moments, 15 + 15 degree-4 moments; + 1 exp
FLOPS_FT = 15    # Fourier: U~^T k 6, norm 3 (+2), phase 3, complex multiply-add 4; + exp + sincos
FLOPS_VOL = 20   # volume: y - x 3, U (y - x) 9, squared norm 5, scale 1, amplitude 1, accumulate 1; + 1 exp
PEAK_FP32_TFLOPS = 157.3      # MI355X published FP32 (vector, packed / matrix)
PEAK_FP32_NONPACKED = 78.65   # 256 CU x 4 SIMD x 16 lanes x 2 flop x 2.4 GHz
PEAK_TRANS_PER_S = 256 * 4 * 16 * 2.4e9 / 4  # quarter-rate transcendental
PEAK_HBM_GBS = 8000.0
# moments kernel, 1000 atoms x 61 angles x 128^2, per launch: FETCH_SIZE 19481.5 KiB x 1.9998 (the
# counter's reading of a 256 MiB copy) read + WRITE_SIZE 7625.0 KiB written (38.0 + 7.4 MiB)
MOMENTS_HBM_BYTES = int((19481.5 * 1.9998 + 7625.0) * 1024)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # a step takes about 0.3 ms: 500 timed steps after 50 warm-up steps are
    # 0.2 s of GPU time, long enough for the clocks to settle
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-atoms", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--atoms", type=int, default=1000)
    ap.add_argument("--kernel-sample-every", type=int, default=8,
                    help="bracket every k-th moments launch of the timed steps with HIP events")
    ap.add_argument("--fit-iters", type=int, default=20, help="Newton outer iterations of the fit leg (0 skips it)")
    ap.add_argument("--serve-requests", type=int, default=2000, help="requests of the serving leg (0 skips it)")
    ap.add_argument("--serve-nmax", type=int, default=5000, help="upper end of the log-uniform atom count")
    ap.add_argument("--plain-cap", type=float, default=3.0,
                    help="seconds a sampled plain fit may run before it is cut to a timed prefix")
    ap.add_argument("--autograd-reps", type=int, default=3,
                    help="timed forward+backward passes of the autograd leg (0 skips it)")
    ap.add_argument("--fourier-reps", type=int, default=5, help="timed evaluations of the Fourier leg (0 skips it)")
    ap.add_argument("--extras-reps", type=int, default=5,
                    help="timed calls of the point-spread / Fourier-plane / KL-moment leg (0 skips it)")
    ap.add_argument("--gd-iters", type=int, default=100, help="outer iterations of the 2D block-GD leg (0 skips it)")
    ap.add_argument("--gd-cpu-iters", type=int, default=5,
                    help="outer iterations of the 2D block-GD leg timed on the CPU oracle")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving leg and print its JSON")
    ap.add_argument("--steady-fits", type=int, default=0,
                    help="run only the fit kernel's steady state with this many fits at once and print its JSON")
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


def stratified_sample(specs, n_edges, s_bins, seed):
    """One request per (atom-count bin, detector-size bin) cell, drawn with a
    seeded shuffle: n_edges are the atom-count bin edges, s_bins tuples of
    detector sides."""
    order = np.random.default_rng(seed).permutation(len(specs))
    pick = []
    for a, b in zip(n_edges[:-1], n_edges[1:]):
        for sides in s_bins:
            for i in order:
                sp = specs[i]
                if a <= sp.N < b and sp.S in sides:
                    pick.append(int(i))
                    break
    return sorted(pick)


def _cost_features(nA, S):
    nA, S = np.asarray(nA, float), np.asarray(S, float)
    return np.stack([np.ones_like(nA), nA, S * S, nA * S * S], -1)


def steady_leg(torch, args, dev, launches=4, budget=100):
    """Steady state of the served fitter: args.steady_fits requests of the
    serving stream (a seed-5 pick of its first 2000), one slot and one
    workgroup each, all advanced together by `launches` grouped launches of
    `budget` updates per fit after one warm-up launch of 10.  Reports atom
    updates/s and, for the first fit of each (detector side, more than 75
    angles) class, the fit kernel's wall-clock ticks (100 MHz) per phase and
    update: setup, moments + per-angle contraction, angle sums, solve,
    proposal set-up, energy pass, decision, commit."""
    Sv = importlib.import_module("2018_dictionary_tomography_amd.serving")
    F = importlib.import_module("2018_dictionary_tomography_amd.fit")
    specs = Sv.c5_stream(2000, rate=200.0, seed=2024, n_range=(50, 5000))
    rng = np.random.default_rng(5)
    sub = [specs[i] for i in sorted(rng.choice(len(specs), args.steady_fits, replace=False))]
    first = {}
    for k, spec in enumerate(sub):
        first.setdefault((spec.S, spec.nA > 75), k)
    watched = sorted(first.values())
    ticks = torch.zeros((len(watched), 16), dtype=torch.int64, device=dev)
    reqs = []
    for k, spec in enumerate(sub):
        start, space, data = Sv.build_request(spec, dev)
        req = F.FitRequest(start, space, data, [spec.max_iter, 1], record=False)
        if k in watched:
            req.slot.phase_ticks = ticks[watched.index(k)].data_ptr()
        reqs.append(req)
    torch.cuda.synchronize()
    raw = b"".join(r.slot_bytes() for r in reqs)
    table = torch.as_tensor(np.frombuffer(raw, dtype=np.uint8).copy(), device=dev)
    n = len(reqs)
    F.run_slots(table, n, 3, 10)
    torch.cuda.synchronize()
    u0 = sum(sl.updates for sl in F.read_slots(table, n))
    ticks.zero_()
    t0 = time.perf_counter()
    for _ in range(launches):
        F.run_slots(table, n, 3, budget)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    slots = F.read_slots(table, n)
    u1 = sum(sl.updates for sl in slots)
    per = ticks.cpu().numpy() / 100.0 / (launches * budget)
    names = ["setup", "moments_contraction", "angle_sums", "solve", "proposal", "energy", "decision", "commit"]
    return {"fits": n, "updates": int(u1 - u0), "seconds": dt, "updates_per_s": (u1 - u0) / dt,
            "accepted_frac": sum(sl.accepted for sl in slots) / max(u1, 1),
            "us_per_update": [{"N": sub[k].N, "angles": sub[k].nA, "side": sub[k].S,
                               **{nm: round(float(per[i][j]), 1) for j, nm in enumerate(names)}}
                              for i, k in enumerate(watched)]}


def serving_leg(torch, args, dev):
    """Poisson stream at 200 req/s with the serving workload's distributions
    (N log-uniform in [50, serve_nmax], angles U{30..120} over +-60 deg,
    detector side from {64,96,128,192,256}, 30-iteration Newton).

    Served = continuous batching into the slot table with graph-replayed
    grouped launches.  Plain = the same fits one request at a time with the
    whole device, each at the group size that runs it fastest (best_group).
    Plain is measured on a stratified sample (atom-count quartiles x three
    detector-size bins = 12 requests): a sampled fit runs in full when it
    takes at most plain_cap seconds, else for its first atom updates within
    that time, and its time per update is kept.  A per-update cost model
    u(nA, S) = b0 + b1 nA + b2 S^2 + b3 nA S^2 is fitted to the sample (its
    residual is reported) and applied to every request with its update count
    from the served run (a fit makes the same updates alone and served, the
    results being equal): plain fits/s = requests / sum of modelled times.
    Request payloads (start atoms, data) are generated before either is
    timed."""
    Sv = importlib.import_module("2018_dictionary_tomography_amd.serving")
    specs = Sv.c5_stream(args.serve_requests, rate=200.0, seed=args.seed, n_range=(50, args.serve_nmax))
    log = (lambda m: print(m, file=sys.stderr, flush=True))
    t0 = time.perf_counter()
    payloads = Sv.prepare_payloads(specs, dev)
    torch.cuda.synchronize()
    prep_s = time.perf_counter() - t0
    log(f"serving: payloads of {len(specs)} requests built in {prep_s:.1f}s")
    server = Sv.FitServer(dev, keep_results=True)
    st = server.run(specs, log=None, payloads=payloads, progress=log, progress_every=20.0)
    served = st.summary()
    log(f"serving: served {served['succeeded']} fits in {served['makespan_s']:.1f}s "
        f"({served['fits_per_s']:.2f} fits/s)")
    updates = {rid: int(v[5]) for rid, v in st.results.items()}
    # plain: stratified sample
    edges = np.exp(np.linspace(np.log(50), np.log(args.serve_nmax + 1), 5))
    pick = stratified_sample(specs, edges, ((64, 96), (128,), (192, 256)), args.seed + 1)
    rows, same = [], True
    for i in pick:
        sp = specs[i]
        g, per = Sv.best_group(sp, payloads[sp.rid], dev)
        total = updates[sp.rid]
        cap = max(sp.N, int(args.plain_cap / per[g]))
        dt, req, s = Sv.plain_fit_timed(sp, payloads[sp.rid], dev, g, max_updates=min(total, cap))
        full = bool(s.done)
        if full:
            Im, xm, rm, Em, donem, updm = st.results[sp.rid]
            same &= (torch.equal(req.I, Im) and torch.equal(req.x, xm) and torch.equal(req.r, rm) and s.E == Em
                     and s.updates == updm)
        rows.append({"rid": int(sp.rid), "N": sp.N, "nA": sp.nA, "S": sp.S, "group": int(g),
                     "updates_timed": int(s.updates), "updates_total": total, "seconds_timed": dt,
                     "us_per_update": dt / max(s.updates, 1) * 1e6, "full_fit": full,
                     "group_us_per_update": {str(k): round(v * 1e6, 1) for k, v in per.items()}})
        log(f"plain rid={sp.rid} N={sp.N} nA={sp.nA} S={sp.S} G={g} {s.updates}/{total} updates in {dt:.2f}s")
    u = np.array([r["us_per_update"] for r in rows]) * 1e-6
    X = _cost_features([r["nA"] for r in rows], [r["S"] for r in rows])
    beta = np.linalg.lstsq(X / u[:, None], np.ones_like(u), rcond=None)[0]   # relative least squares
    resid = X @ beta / u - 1.0
    allX = _cost_features([sp.nA for sp in specs], [sp.S for sp in specs])
    per_fit = np.array([updates[sp.rid] for sp in specs]) * (allX @ beta)
    plain_fps = len(specs) / float(per_fit.sum())
    # the plain server's latency on the same arrival stream: one request at a
    # time, first come first served, with the modelled service times
    free, lat = 0.0, []
    for sp, t in sorted(zip(specs, per_fit), key=lambda z: z[0].arrival):
        free = max(free, sp.arrival) + float(t)
        lat.append(free - sp.arrival)
    for r, x in zip(rows, X):
        r["model_seconds_full"] = float(r["updates_total"] * (x @ beta))
    return {
        "requests": len(specs), "rate_per_s": 200.0, "atoms_range": [50, args.serve_nmax],
        "angles_range": [30, 120], "detector_sides": [64, 96, 128, 192, 256], "iterations": 30,
        "served": served, "client_payload_s": prep_s,
        "plain": {
            "fits_per_s": plain_fps,
            "p50_service_s": float(np.percentile(per_fit, 50)), "p99_service_s": float(np.percentile(per_fit, 99)),
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "latency_note": "first come first served on the same arrivals with the modelled service times",
            "sample": rows,
            "cost_model": {"form": "us/update = b0 + b1 nA + b2 S^2 + b3 nA S^2", "coef_s": beta.tolist(),
                           "rms_rel_residual": float(np.sqrt(np.mean(resid ** 2))),
                           "max_rel_residual": float(np.abs(resid).max())},
            "note": "same device fitter, one request at a time with the whole device at its fastest group size; "
                    "service times = served update counts x modelled time per update"},
        "speedup_fits_per_s": served["fits_per_s"] / plain_fps,
        "p99_latency_ratio_plain_over_served": float(np.percentile(lat, 99)) / served["p99_latency_s"],
        "alone_equals_mixed": bool(same),
        "alone_equals_mixed_checked": int(sum(r["full_fit"] for r in rows)),
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
    L = importlib.import_module("2018_dictionary_tomography_amd._lib")
    kt = _kernel_table(torch, one, reps, {"setup": L.GT_K_SETUP, "forward": L.GT_K_FORWARD,
                                          "moments": L.GT_K_MOMENTS, "finish": L.GT_K_FINISH})
    P_in = support_pairs_columns(torch, t[2].detach().cpu().numpy(), t[1].detach().cpu().numpy(), T, ortho[0],
                                 ortho[1], det[0], det[1])
    return {"workload": "5000 atoms, 90 angles over [0, pi) about a random axis, 256x256 detector",
            "ms_per_forward_backward": ms, "evals_per_s": 2.0 * N * nA * S * S / (ms * 1e-3),
            "loss": float(loss.detach()), "reps": reps, "kernel_ms": kt,
            "roofline": {"forward": roof_entry("forward3_kernel", P_in, kt["forward"], FLOPS_FWD, 1),
                         "adjoint": roof_entry("moments_kernel<2,LINEAR> (hand-written backward)", P_in,
                                               kt["moments"], FLOPS_ADJ, 1)}}


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
    L = importlib.import_module("2018_dictionary_tomography_amd._lib")
    kms = _kernel_table(torch, lambda: Fo.evaluate(atoms, grid, out=out), reps, {"ft": L.GT_K_FT})["ft"]
    Uc = U.cpu().numpy()
    P_in = ellipsoid_lattice_count(torch, Uc @ Uc.transpose(0, 2, 1), np.zeros((len(nI), 3)), k[0], k[1], k[2],
                                   30.0, strict=True)
    return {"workload": "2000 atoms on the integer grid [-64, 64)^3, complex64", "ms_per_evaluation": ms,
            "atom_k_pairs_per_s": pairs / (ms * 1e-3), "rel_err_vs_fp64_sampled": err, "reps": reps,
            "roofline": roof_entry("ft_patch_kernel", P_in, kms, FLOPS_FT, 2,
                                   "in-support = (atom, k) pairs with |U~^T k|^2 < 30")}


def _timed(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def _kernel_table(torch, fn, reps, ids):
    """Average launch duration (ms) of several library kernels over reps
    calls of fn, from the library's HIP-event hooks (on each kernel's stream)."""
    L = importlib.import_module("2018_dictionary_tomography_amd._lib")
    torch.cuda.synchronize()
    L.profile_enable(True)
    for k in ids.values():
        L.profile_read(k)
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    L.profile_enable(False)
    out = {}
    for name, k in ids.items():
        tot, cnt = L.profile_read(k)
        out[name] = tot / max(cnt, 1)
    return out


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
    # the reference's own sampled cell rule on the same workload
    ms_fs = _timed(torch, lambda: psf.forward(atoms, space, out=P, rule="reference"), reps)
    ms_ds = _timed(torch, lambda: psf.derivs(atoms, space, C, "L2", 2, rule="reference"), reps)
    ms_vs = _timed(torch, lambda: psf.volume(atoms, vol, out=u, rule="reference"), reps)
    vol128 = gtp.VolSpace(box=(-1, 1), shape=128, dim=3)
    u128 = gtp.ops.volume(atoms, vol128)
    ms_pv = _timed(torch, lambda: gtp.ops.volume(atoms, vol128, out=u128), reps)
    L = importlib.import_module("2018_dictionary_tomography_amd._lib")
    k_pv = _kernel_ms(torch, L.GT_K_VOLUME, lambda: gtp.ops.volume(atoms, vol128, out=u128), reps)
    rc = np.asarray(r, dtype=np.float32).astype(np.float64)
    Uv = np.zeros((rc.shape[0], 3, 3))
    for k_, (i_, j_) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]):
        Uv[:, i_, j_] = rc[:, k_]
    cv = vol128.coord_vectors
    P_vol = ellipsoid_lattice_count(torch, Uv.transpose(0, 2, 1) @ Uv, np.float32(x).astype(np.float64), cv[0],
                                    cv[1], cv[2], 1.0, strict=False)
    out["volume"] = {"workload": "VolProj of the same 1000 atoms on 128^3 voxel centres",
                     "ms": ms_pv, "kernel_ms": k_pv, "note": "ms: wall clock per call (host-bound); kernel_ms: "
                     "the render kernel alone (HIP events)",
                     "roofline": roof_entry("volume3_kernel", P_vol, k_pv, FLOPS_VOL, 1,
                                            "in-support = (atom, voxel) pairs with |U(y - x)|^2 <= 1")}
    out["psf"] = {"workload": "1000 atoms, 61 x 128^2, pixel-integrated; voxel-integrated render on 64^3",
                  "forward_ms": ms_f, "derivs_L2_ms": ms_d, "volume_ms": ms_v,
                  "forward_atom_pixel_per_s": 1000 * rays / (ms_f * 1e-3),
                  "derivs_atom_pixel_per_s": 1000 * rays / (ms_d * 1e-3),
                  "reference_rule": {"forward_ms": ms_fs, "derivs_L2_ms": ms_ds, "volume_ms": ms_vs,
                                     "note": "11 (x 11 (x 11)) samples per cell, total / zero cells as "
                                             "point_spread_functions.py"}}
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


def _count_in(lo, hi, grid):
    """Number of ascending grid values in [lo, hi] (elementwise)."""
    import torch
    return (torch.searchsorted(grid, hi, right=True) - torch.searchsorted(grid, lo, right=False)).clamp_min(0)


def support_pairs_columns(torch, r, x, T, w0, w1, p0, p1, chunk=1024):
    """Exact number of (atom, pixel) pairs with m <= 1 on a 2D detector,
    counted per detector column (FP64 torch, independent of the library):
    along a column p1 the ray offset is G p - x = p0 w0 + d, d = p1 w1 - x,
    so m = a p0^2 + b p0 + c with a = w0^T Q w0, b = 2 w0^T Q d, c = d^T Q d
    (Q = U^T (Id - t t^T/|t|^2) U, t = U T), and the rows inside are the
    detector coordinates between the roots of m = 1."""
    dev = torch.device("cuda" if torch.cuda.is_available() else "cpu")
    d64 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
    r, x = d64(r), d64(x)
    N = r.shape[0]
    U = torch.zeros((N, 3, 3), dtype=torch.float64, device=dev)
    for k, (i, j) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]):
        U[:, i, j] = r[:, k]
    P = U.transpose(1, 2) @ U
    g0, g1 = d64(p0), d64(p1)
    total = 0
    for a in range(T.shape[0]):
        Ta, v0, v1 = d64(T[a]), d64(w0[a]), d64(w1[a])
        z = P @ Ta
        Q = P - z[:, :, None] * z[:, None, :] / (z @ Ta)[:, None, None]
        Qw0 = Q @ v0
        A = Qw0 @ v0
        for c in range(0, N, chunk):
            D = g1[None, :, None] * v1[None, None, :] - x[c:c + chunk, None, :]
            QD = torch.matmul(D, Q[c:c + chunk].transpose(1, 2).contiguous())
            B = 2.0 * (D * Qw0[c:c + chunk, None, :]).sum(-1)
            C = (D * QD).sum(-1) - 1.0
            Ac = A[c:c + chunk, None]
            disc = B * B - 4.0 * Ac * C
            ok = (disc >= 0) & (Ac > 0)
            sq = disc.clamp_min(0).sqrt()
            lo, hi = (-B - sq) / (2 * Ac), (-B + sq) / (2 * Ac)
            total += int((_count_in(lo.contiguous(), hi.contiguous(), g0) * ok).sum())
    return total


def ellipsoid_lattice_count(torch, M, centre, k0, k1, k2, bound, strict, chunk=256):
    """Number of grid points y = (k0[i], k1[j], k2[l]) with (y - c)^T M (y - c)
    below `bound` (strictly if `strict`), per atom M (n, 3, 3) and centre c
    (n, 3), summed over atoms: per (k0, k1) column the inequality is a
    quadratic in the third coordinate, whose roots bound the points inside."""
    dev = torch.device("cuda" if torch.cuda.is_available() else "cpu")
    d64 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
    M, centre = d64(M), d64(centre)
    a0, a1, a2 = d64(k0), d64(k1), d64(k2)
    Y0, Y1 = torch.meshgrid(a0, a1, indexing="ij")
    total = 0
    for c in range(0, M.shape[0], chunk):
        m = M[c:c + chunk]
        e0 = Y0[None] - centre[c:c + chunk, 0, None, None]
        e1 = Y1[None] - centre[c:c + chunk, 1, None, None]
        g = lambda i, j: m[:, i, j, None, None]
        A = g(2, 2).expand_as(e0)
        B = 2.0 * (g(0, 2) * e0 + g(1, 2) * e1)
        C = g(0, 0) * e0 * e0 + 2.0 * g(0, 1) * e0 * e1 + g(1, 1) * e1 * e1 - bound
        disc = B * B - 4.0 * A * C
        ok = disc > 0 if strict else disc >= 0
        sq = disc.clamp_min(0).sqrt()
        lo = (-B - sq) / (2 * A) + centre[c:c + chunk, 2, None, None]
        hi = (-B + sq) / (2 * A) + centre[c:c + chunk, 2, None, None]
        if strict:
            n = (torch.searchsorted(a2, hi.reshape(-1).contiguous(), right=False)
                 - torch.searchsorted(a2, lo.reshape(-1).contiguous(), right=True)).clamp_min(0)
        else:
            n = _count_in(lo.reshape(-1).contiguous(), hi.reshape(-1).contiguous(), a2)
        total += int((n * ok.reshape(-1)).sum())
    return total


def roof_entry(kernel, pairs, ms, flops, trans, note=None):
    """Roofline block of one kernel: algorithmic FP32 flops and transcendentals
    per in-support pair times the pairs of one launch, over its average
    launch time (HIP events on its stream)."""
    ach = pairs * flops / (ms * 1e-3) / 1e12
    out = {"bound": "mfma", "engine": "fp32-valu", "kernel": kernel, "in_support_pairs": int(pairs),
           "flops_per_pair": flops, "transcendentals_per_pair": trans, "avg_kernel_ms": round(ms, 5),
           "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
           "frac": round(ach / PEAK_FP32_TFLOPS, 4), "frac_nonpacked": round(ach / PEAK_FP32_NONPACKED, 4),
           "trans_frac": round(pairs * trans / (ms * 1e-3) / PEAK_TRANS_PER_S, 4)}
    if note:
        out["note"] = note
    return out


def cpu_baseline(I, x, r, T, ortho, det, data, n_atoms):
    """The reference's dense loops on the host cores (one ray x all atoms for
    the forward, one atom x all rays for the L2 gradient and Hessian,
    NewtonBases.py:729-870), restated in C (oracle/gt_oracle.c) and compiled
    here with gcc -O3 -march=native -fopenmp in FP32 (the reference's float32)
    and FP64, on a bounded sample: the first n_atoms atoms over the full
    geometry.  value = the FP32 rate in dense atom.ray evaluations per second."""
    import tempfile
    from oracle import oracle as O

    g = O.Geometry(T, ortho, det)
    rays = T.shape[0] * det[0].size * det[1].size
    sl = slice(0, n_atoms)
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for bits in (32, 64):
            cb = O.CpuBaseline(tmp, bits)
            t0 = time.perf_counter()
            R = cb.radproj(I[sl], x[sl], r[sl], g)
            t1 = time.perf_counter()
            cb.derivs(I[sl], x[sl], r[sl], g, data - R, "L2", 2)
            t2 = time.perf_counter()
            out[f"fp{bits}"] = {"value": 2.0 * n_atoms * rays / (t2 - t0), "forward_s": t1 - t0,
                                "derivs_s": t2 - t1, "threads": cb.threads(), "flags": cb.flags}
    return {
        "value": out["fp32"]["value"],
        "unit": "atom-ray evals/s",
        "cores": out["fp32"]["threads"],
        "kind": "port",
        "sample": f"{n_atoms} of the atoms, full {T.shape[0]}x{det[0].size}x{det[1].size} geometry: dense forward "
                  f"+ L2 gradient/Hessian (the reference's loops in C, gcc -O3 -march=native, OpenMP); value is "
                  f"FP32, fp64 beside it",
        "fp32": out["fp32"], "fp64": out["fp64"],
        "cpu": _cpu_model(),
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def gd2d_leg(torch, gtp, dev, iters, cpu_iters):
    """The 2D configuration: 50 isotropic atoms (centres U[-0.8, 0.8]^2,
    radii U[0.02, 0.08], intensities U[0.5, 1.5]), 180 angles over [0, pi),
    128-pixel detector on [-1, 1], data = forward + 1% noise, block gradient
    descent (GD_lib.linesearch_block) from centres + N(0, 0.02).  GPU: the
    drop-in fitter for `iters` outer iterations.  CPU: the FP64 oracle's
    restatement (Python loop over the C projector, OpenMP) for cpu_iters
    outer iterations, a bounded sample of the same run."""
    from oracle import oracle as O
    F = importlib.import_module("2018_dictionary_tomography_amd.fit")
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(7)
    I, x, r = W.draw_iso2(50, rng)
    T, ortho, det = W.fan_2d(180, 128)
    space = gtp.ProjSpace(orientations=T, ortho=ortho, detector=det)
    aspace = gtp.AtomSpace(2, isotropic=True)
    data = gtp.ops.forward(gtp.AtomElement(aspace, x, r, I, device=dev), space)
    data = data + torch.as_tensor(rng.normal(0, 0.01 * float(data.max()), data.shape), dtype=torch.float32,
                                  device=dev)
    x0 = x + rng.normal(0, 0.02, x.shape)
    start = gtp.AtomElement(aspace, x0, r, I, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, Eg = F.gd_linesearch_block(start, data, [iters, 1], space)
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    out = {"workload": "50 isotropic 2D atoms, 180 angles over [0, pi), 128-pixel detector, block GD",
           "gpu": {"outer_iterations": iters, "block_steps": len(Eg) - 1, "seconds": tg,
                   "s_per_iteration": tg / iters, "E0": float(Eg[0]), "E_final": float(Eg[-1])}}
    if cpu_iters > 0:
        f32 = lambda a: np.float32(a).astype(np.float64)
        g = O.Geometry(f32(T), tuple(f32(o) for o in ortho), [f32(p) for p in det])
        t0 = time.perf_counter()
        _, _, _, Eo = O.gd_linesearch_block(f32(I), f32(x0), f32(r), g, data.double().cpu().numpy(), [cpu_iters, 1])
        tc = time.perf_counter() - t0
        n = min(len(Eg), len(Eo))
        out["cpu_fp64_oracle"] = {"outer_iterations": cpu_iters, "seconds": tc, "s_per_iteration": tc / cpu_iters,
                                  "threads": O.num_threads()}
        out["energy_match_rel_E0"] = float(np.abs(Eg[:n] - Eo[:n]).max() / Eo[0])
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (--gpus {args.gpus} requested); the path's sharding over "
              "several GPUs is a design, not run here", file=sys.stderr)
        return 2
    import torch

    gtp = importlib.import_module("2018_dictionary_tomography_amd")
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    from importlib import import_module

    L = import_module("2018_dictionary_tomography_amd._lib")
    dev = torch.device("cuda", 0)
    if args.calib_copy_mib > 0:
        src = torch.ones(args.calib_copy_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.serve_only:
        print(json.dumps({"serving": serving_leg(torch, args, dev)}))
        return 0
    if args.steady_fits > 0:
        print(json.dumps({"steady": steady_leg(torch, args, dev)}))
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

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    for k in range(7):
        L.profile_read(k)
    # The timed steps bracket only the moments kernel (the roofline's) with HIP
    # events on its stream, and only every k-th launch of it: a bracketed
    # launch leaves a gap of about 9 us before the next kernel, so bracketing
    # every step would add that to every step.  The other kernels are timed in
    # a second, untimed pass.
    every = max(1, args.kernel_sample_every)
    L.profile_select([L.GT_K_MOMENTS])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        if i % every == 0:
            L.profile_enable(True)
            step()
            L.profile_enable(False)
        else:
            step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kt_mom = L.profile_read(L.GT_K_MOMENTS)
    L.profile_select(None)
    L.profile_enable(True)
    for _ in range(min(args.steps, 20)):
        step()
    torch.cuda.synchronize()
    L.profile_enable(False)
    kt = {name: L.profile_read(kid) for name, kid in
          (("setup", L.GT_K_SETUP), ("forward", L.GT_K_FORWARD), ("moments", L.GT_K_MOMENTS),
           ("finish", L.GT_K_FINISH))}
    kt["moments"] = kt_mom
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
        "kernel_ms_note": "moments: HIP events around every %d-th launch of it inside the timed steps (only that "
                          "kernel bracketed, %d launches); setup, forward, finish: a second, untimed pass with every "
                          "kernel bracketed" % (every, kt_mom[1]),
    }
    fit = fit_leg(torch, gtp, space, data, cur, args.fit_iters) if args.fit_iters > 0 else None
    serving = serving_leg(torch, args, dev) if args.serve_requests > 0 else None
    autograd = autograd_leg(torch, gtp, dev, args.autograd_reps) if args.autograd_reps > 0 else None
    fourier = fourier_leg(torch, gtp, dev, args.fourier_reps) if args.fourier_reps > 0 else None
    extras = extras_leg(torch, gtp, dev, args.extras_reps) if args.extras_reps > 0 else None
    gd2d = gd2d_leg(torch, gtp, dev, args.gd_iters, args.gd_cpu_iters) if args.gd_iters > 0 else None
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
        "gd_2d": gd2d,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
