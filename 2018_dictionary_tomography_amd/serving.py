"""Continuous-batching server of Newton fits.

Requests (atoms, geometry, data) arrive over time; each one becomes a slot of
a device-resident slot table (gt_fit_slot) as soon as a slot is free and its
HBM footprint fits the budget.  One grouped launch (gt_fit_run_blocks) over
every fit workgroup the device holds advances every active slot by a work
quota: the block table gives each active fit one workgroup, and spare
workgroups go to the fits with the most remaining work (a group of G
workgroups shares a fit's angles).  The launch is captured once in a HIP
graph and replayed; between replays the host admits arrivals into free
slots, retires finished fits and rewrites the block table.  State (atoms,
residual, step sizes) never leaves the device while a fit runs.  A fit's
gradients, Hessians and energy changes are reduced per angle in a fixed
order, so its result depends neither on its group size nor on what else is
in the table (tests/test_gpu_serving.py checks this bit for bit).

`plain_fits` / `plain_fit_timed` run the same fits one request at a time,
each with the whole device (`best_group` measures the group size at which a
fit alone runs fastest), the baseline the served throughput is compared with.
"""
from __future__ import annotations

import ctypes
import importlib
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import fit as F
from . import ops
from .spaces import AtomElement, AtomSpace, ProjSpace

W = importlib.import_module(__package__ + ".workloads")


@dataclass
class RequestSpec:
    """One fit request of the serving workload."""
    rid: int
    arrival: float  # seconds after the stream starts
    N: int
    nA: int
    S: int
    seed: int
    max_iter: int = 30

    def footprint_bytes(self):
        """HBM held while the request is in flight: data, residual and staging
        sinograms, atoms (twice), permutations, step sizes, the slot scratch
        (per-angle gradient/Hessian records and energy changes), geometry."""
        sino = self.nA * self.S * self.S * 4
        scratch = self.nA * (65 * 8 + 8) + 1024
        return 3 * sino + self.N * (10 * 4 * 2 + 8) + self.max_iter * self.N * 4 + scratch + self.nA * 12 * 4 + 8 * self.S


def c5_stream(n_requests=2000, rate=200.0, seed=0, n_range=(50, 5000), angle_range=(30, 120),
              sides=(64, 96, 128, 192, 256), max_iter=30):
    """Poisson arrivals at `rate`; N log-uniform in n_range, angle count uniform,
    detector side uniform over `sides`."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n_requests)
    t = np.cumsum(gaps) - gaps[0]
    lo, hi = np.log(n_range[0]), np.log(n_range[1])
    out = []
    for i in range(n_requests):
        N = int(round(np.exp(rng.uniform(lo, hi))))
        nA = int(rng.integers(angle_range[0], angle_range[1] + 1))
        S = int(rng.choice(sides))
        out.append(RequestSpec(i, float(t[i]), N, nA, S, int(rng.integers(1 << 30)), max_iter))
    return out


def build_request(spec: RequestSpec, device):
    """Synthetic request: ground-truth atoms drawn as the 3D workload, tilt series
    over +-60 deg about y, data = forward + 1% noise, start = truth with centres
    perturbed by N(0, 0.01)."""
    rng = np.random.default_rng(spec.seed)
    I, x, r, _ = W.draw_aniso3(spec.N, rng)
    T, ortho, det = W.tilt_series(spec.nA, 60.0, det=spec.S)
    space = ProjSpace(orientations=T, ortho=ortho, detector=det)
    aspace = AtomSpace(3, isotropic=False)
    truth = AtomElement(aspace, x, r, I, device=device)
    data = ops.forward(truth, space)
    gen = torch.Generator(device=device)
    gen.manual_seed(spec.seed)
    data += torch.randn(data.shape, generator=gen, device=device) * (0.01 * data.max())
    start = AtomElement(aspace, x + rng.normal(0, 0.01, x.shape), r, I, device=device)
    return start, space, data



def prepare_payloads(specs, device):
    """Client side: every request's (start atoms, geometry, data), built before
    the stream is served so that neither the served nor the plain timing
    includes generating the synthetic data."""
    return {spec.rid: build_request(spec, device) for spec in specs}


def work_per_update(spec):
    """Relative cost of one atom update on one workgroup, in detector-pixel
    units: the pixels touched grow with angles x detector area, plus a fixed
    part (setup, FP64 contraction and solve) worth about 5.3e5 pixels."""
    return spec.nA * spec.S * spec.S + 5.3e5


# Fixed cost of one atom update advanced by a group (setup, contraction,
# barriers, the solve repeated in every member), pixel units.
GROUP_UPDATE_OVERHEAD = 4.1e5


def update_cost(spec, group):
    """Relative cost (time) of one atom update advanced by `group` workgroups:
    the pixel work is shared, the fixed part is not."""
    if group <= 1:
        return work_per_update(spec)
    return spec.nA * spec.S * spec.S / group + GROUP_UPDATE_OVERHEAD


def useful_group(nA):
    """Largest group worth giving one fit: one angle per workgroup (a fit
    alone runs fastest there or within a few percent of it, measured over
    the serving workload's shapes)."""
    return max(1, int(nA))


def allocate_groups(specs, remaining, capacity):
    """Share `capacity` fit workgroups among the active requests: one each,
    then one more at a time to the request whose remaining work per workgroup
    is largest, up to its useful group size.  With as many active fits as
    workgroups (the steady state of an overloaded stream) every fit has one;
    spare workgroups appear as the stream drains."""
    import heapq

    n = len(specs)
    groups = [1] * n
    spare = capacity - n
    if spare <= 0:
        return groups
    heap = [(-remaining[i] * update_cost(specs[i], 1), i) for i in range(n) if useful_group(specs[i].nA) > 1]
    heapq.heapify(heap)
    while spare > 0 and heap:
        _, i = heapq.heappop(heap)
        groups[i] += 1
        spare -= 1
        if groups[i] < useful_group(specs[i].nA):
            heapq.heappush(heap, (-remaining[i] * update_cost(specs[i], groups[i]), i))
    return groups


@dataclass
class ServeStats:
    completed: int = 0
    makespan: float = 0.0
    latencies: list = field(default_factory=list)
    service: list = field(default_factory=list)
    launches: int = 0
    admitted_bytes_peak: int = 0
    queued_peak: int = 0
    refused: int = 0
    failed: int = 0        # fits whose workgroups timed out at a barrier (done = 3)
    singular: int = 0      # fits stopped by a singular Newton system (done = 2)
    admit_s: float = 0.0   # host + device time spent admitting requests
    launch_s: float = 0.0  # time from each launch to its completion
    trace: list = field(default_factory=list)  # per launch: (start s, active fits, workgroups used, s, updates)
    results: dict = field(default_factory=dict)

    def summary(self):
        """fits_per_s counts fits that ended normally (done = 1) only; latency
        is arrival to completion (queueing included), service is admission to
        completion."""
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        svc = np.array(self.service) if self.service else np.zeros(1)
        ok = self.completed - self.failed - self.singular
        return {
            "completed": self.completed,
            "succeeded": ok,
            "fits_per_s": ok / self.makespan if self.makespan > 0 else 0.0,
            "makespan_s": self.makespan,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "p50_service_s": float(np.percentile(svc, 50)),
            "p99_service_s": float(np.percentile(svc, 99)),
            "launches": self.launches,
            "peak_admitted_bytes": self.admitted_bytes_peak,
            "peak_queue": self.queued_peak,
            "refused": self.refused,
            "failed": self.failed,
            "singular": self.singular,
            "admit_s": self.admit_s,
            "launch_s": self.launch_s,
        }


class FitServer:
    """Slot table + block table on the device; one graph-replayed grouped
    launch advances every active fit by its quota.  max_slots defaults to the
    number of fit workgroups the device holds at once, so every active fit
    always has at least one workgroup."""

    def __init__(self, device=None, max_slots=None, hbm_budget=64 << 30, work_per_launch=4.0e7, use_graph=True,
                 keep_results=False, grouped=True, stage_max_atoms=0):
        self.dev = device or torch.device("cuda", torch.cuda.current_device())
        self.capacity = F.max_blocks()
        self.max_slots = min(max_slots or self.capacity, self.capacity)
        self.hbm_budget = hbm_budget
        self.work_per_launch = work_per_launch
        self.use_graph = use_graph
        self.keep_results = keep_results
        # fits of at most this many atoms get a staging slab (see FitRequest;
        # none by default: measured slower at every size of the serving mix)
        self.stage_max_atoms = stage_max_atoms
        self.grouped = grouped
        self.table = torch.zeros(self.max_slots * F.SLOT_BYTES, dtype=torch.uint8, device=self.dev)
        self.host = torch.zeros(self.max_slots * F.SLOT_BYTES, dtype=torch.uint8).pin_memory()
        self.blocks = torch.full((self.capacity, 4), -1, dtype=torch.int32, device=self.dev)
        self.blocks_host = torch.full((self.capacity, 4), -1, dtype=torch.int32).pin_memory()
        self.slots = [None] * self.max_slots  # (spec, FitRequest, t_admit)
        self.graph = None

    def _raw_launch(self):
        F.run_blocks(self.table, self.max_slots, self.blocks, self.capacity, 3, 1 << 30)

    def _launch(self):
        if not self.use_graph:
            self._raw_launch()
            return
        if self.graph is None:
            s = torch.cuda.Stream(self.dev)
            s.wait_stream(torch.cuda.current_stream(self.dev))
            with torch.cuda.stream(s):
                self._raw_launch()  # warm-up on the side stream
            torch.cuda.current_stream(self.dev).wait_stream(s)
            torch.cuda.synchronize(self.dev)
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self._raw_launch()
        self.graph.replay()

    def _write_slot(self, i, slot):
        b = torch.frombuffer(bytearray(bytes(slot)), dtype=torch.uint8)
        self.table[i * F.SLOT_BYTES:(i + 1) * F.SLOT_BYTES].copy_(b.to(self.dev))

    def _read_active(self):
        """(index, gt_fit_slot) of every occupied slot, read back after a launch."""
        self.host.copy_(self.table)
        raw = self.host.numpy()
        return [(i, F.gt_fit_slot.from_buffer_copy(raw, i * F.SLOT_BYTES))
                for i in range(self.max_slots) if self.slots[i] is not None]

    def _plan(self, progress):
        """Rewrite the block table: groups and quotas of the active slots."""
        act = [i for i in range(self.max_slots) if self.slots[i] is not None]
        specs = [self.slots[i][0] for i in act]
        remaining = []
        for i, spec in zip(act, specs):
            done = progress.get(i, 0)
            remaining.append(max(1, spec.max_iter * spec.N - done))
        groups = allocate_groups(specs, remaining, self.capacity) if self.grouped else [1] * len(act)
        rows = np.full((self.capacity, 4), -1, dtype=np.int32)
        b = 0
        for i, spec, g in zip(act, specs, groups):
            q = max(1, int(self.work_per_launch / update_cost(spec, g)))
            for m in range(g):
                rows[b] = (i, m, g, q)
                b += 1
        self.blocks_host.copy_(torch.from_numpy(rows))
        self.blocks.copy_(self.blocks_host)
        self.used_blocks = b
        return dict(zip(act, groups))

    def run(self, specs, time_scale=1.0, log=None, payloads=None, progress=None, progress_every=20.0):
        """Serve `specs` (sorted by arrival).  time_scale stretches the arrival
        clock (1.0 = real time).  payloads: prepare_payloads() output (else each
        request's data is generated at admission).  log: called with a line per
        completed fit; progress: called with a status line every
        progress_every seconds.  Returns ServeStats."""
        st = ServeStats()
        pending = sorted(specs, key=lambda s: s.arrival)
        queue = []
        reserved = 0
        t0 = time.perf_counter()
        nxt = 0
        active = 0
        progress_map = {}
        # the empty table must not be left with stale slots
        self.table.zero_()
        torch.cuda.synchronize(self.dev)
        dirty = True
        next_note = progress_every
        while nxt < len(pending) or queue or active:
            now = time.perf_counter() - t0
            while nxt < len(pending) and pending[nxt].arrival * time_scale <= now:
                spec = pending[nxt]
                if spec.footprint_bytes() > self.hbm_budget:
                    st.refused += 1
                else:
                    queue.append(spec)
                nxt += 1
            st.queued_peak = max(st.queued_peak, len(queue))
            # admit FIFO while a slot and memory are free
            free = [i for i in range(self.max_slots) if self.slots[i] is None]
            ta = time.perf_counter()
            while queue and free and reserved + queue[0].footprint_bytes() <= self.hbm_budget:
                spec = queue.pop(0)
                i = free.pop(0)
                start, space, data = payloads[spec.rid] if payloads else build_request(spec, self.dev)
                req = F.FitRequest(start, space, data, [spec.max_iter, 1], record=False,
                                   stage=spec.N <= self.stage_max_atoms)
                self._write_slot(i, req.slot)
                self.slots[i] = (spec, req, time.perf_counter() - t0)
                progress_map[i] = 0
                reserved += spec.footprint_bytes()
                active += 1
                dirty = True
            st.admit_s += time.perf_counter() - ta
            st.admitted_bytes_peak = max(st.admitted_bytes_peak, reserved)
            if active == 0:
                if nxt < len(pending):
                    time.sleep(max(0.0, min(0.01, pending[nxt].arrival * time_scale - now)))
                continue
            if progress is not None and now >= next_note:
                progress(f"serving: t={now:.0f}s completed={st.completed} active={active} queued={len(queue)} "
                         f"launches={st.launches}")
                next_note = now + progress_every
            if dirty:
                self._plan(progress_map)
                dirty = False
            tl = time.perf_counter()
            self._launch()
            st.launches += 1
            torch.cuda.synchronize(self.dev)
            dt = time.perf_counter() - tl
            st.launch_s += dt
            before = sum(progress_map.values())
            rows = self._read_active()
            for i, s in rows:
                progress_map[i] = s.updates
            st.trace.append((tl - t0, active, self.used_blocks, dt, sum(progress_map.values()) - before))
            for i, s in rows:
                if not (s.done or not s.active):
                    continue
                spec, req, t_admit = self.slots[i]
                t_done = time.perf_counter() - t0
                st.latencies.append(t_done - spec.arrival * time_scale)
                st.service.append(t_done - t_admit)
                st.completed += 1
                st.failed += int(s.done == 3)
                st.singular += int(s.done == 2)
                if self.keep_results:
                    st.results[spec.rid] = (req.I.clone(), req.x.clone(), req.r.clone(), s.E, s.done, s.updates)
                reserved -= spec.footprint_bytes()
                active -= 1
                self.slots[i] = None
                progress_map.pop(i, None)
                self._write_slot(i, F.gt_fit_slot())  # inactive
                dirty = True
                if log:
                    log(f"done rid={spec.rid} N={spec.N} nA={spec.nA} S={spec.S} updates={s.updates} "
                        f"latency={st.latencies[-1]:.2f}s")
        st.makespan = time.perf_counter() - t0
        return st


def group_candidates(nA, capacity):
    """Group sizes tried for a plain fit of nA angles: the default (about one
    angle per wave), 16, 32, 64 and one angle per workgroup."""
    c = {F.default_group(nA), 16, 32, 64, int(nA)}
    return sorted(g for g in c if 1 <= g <= min(int(nA), capacity))


def best_group(spec, payload, device, updates=None, candidates=None):
    """Fastest group size for a plain fit of this request: each candidate
    advances a fresh copy of the fit by the same number of atom updates (after
    a short warm-up) and the one with the least time per update wins.
    Returns (group, {group: seconds per update})."""
    start, space, data = payload
    cands = candidates or group_candidates(spec.nA, F.max_blocks())
    n = updates or max(16, min(spec.N, 96))
    per = {}
    for g in cands:
        req = F.FitRequest(start, space, data, [spec.max_iter, 1], record=False)
        table = torch.as_tensor(np.frombuffer(req.slot_bytes(), dtype=np.uint8).copy(), device=device)
        blocks = torch.as_tensor(F.block_table([g]), device=device)
        F.run_blocks(table, 1, blocks, g, req.dim, 4)
        torch.cuda.synchronize(device)
        u0 = F.read_slots(table, 1)[0].updates
        t0 = time.perf_counter()
        F.run_blocks(table, 1, blocks, g, req.dim, n)
        torch.cuda.synchronize(device)
        dt = time.perf_counter() - t0
        done = F.read_slots(table, 1)[0].updates - u0
        per[g] = dt / max(done, 1)
    return min(per, key=per.get), per


def plain_fit_timed(spec, payload, device, group, max_updates=None, work_per_launch=4.0e7):
    """One plain fit (the request alone on the device, `group` workgroups),
    stopped after max_updates atom updates if given.  Returns (seconds,
    FitRequest, final slot)."""
    start, space, data = payload
    req = F.FitRequest(start, space, data, [spec.max_iter, 1], record=False)
    table = torch.as_tensor(np.frombuffer(req.slot_bytes(), dtype=np.uint8).copy(), device=device)
    blocks = torch.as_tensor(F.block_table([group]), device=device) if group > 1 else None
    budget = max(1, int(work_per_launch / update_cost(spec, group)))
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    while True:
        left = budget if max_updates is None else min(budget, max_updates - F.read_slots(table, 1)[0].updates)
        if left <= 0:
            break
        if blocks is None:
            F.run_slots(table, 1, req.dim, left)
        else:
            F.run_blocks(table, 1, blocks, group, req.dim, left)
        s = F.read_slots(table, 1)[0]
        if s.done or not s.active:
            break
    torch.cuda.synchronize(device)
    dt = time.perf_counter() - t0
    s = F.read_slots(table, 1)[0]
    req.final = s
    return dt, req, s


def plain_fits(specs, device=None, keep_results=False, work_per_launch=4.0e7, group=None, payloads=None):
    """The same fits one request at a time, each with the whole device to
    itself: fit.newton_fit with its default group (group=1: one workgroup).
    Returns (service times in s, results).  With keep_results, each result
    also carries the drift of the incrementally updated residual:
    max|resid - (forward(final atoms) - data)| / max|forward(final atoms)|."""
    dev = device or torch.device("cuda", torch.cuda.current_device())
    times, results = [], {}
    for spec in specs:
        start, space, data = payloads[spec.rid] if payloads else build_request(spec, dev)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        g = F.default_group(spec.nA) if group is None else group
        req, s = F.newton_fit(start, space, data, [spec.max_iter, 1], record=False, group=g,
                              budget=max(1, int(work_per_launch / update_cost(spec, g))))
        torch.cuda.synchronize(dev)
        times.append(time.perf_counter() - t0)
        if keep_results:
            fresh = ops.forward(req, space)
            drift = float((req.resid - (fresh - req.data)).abs().max() / fresh.abs().max())
            results[spec.rid] = (req.I.clone(), req.x.clone(), req.r.clone(), s.E, s.done, s.updates, drift)
    return times, results
