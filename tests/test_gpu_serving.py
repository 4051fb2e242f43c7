"""Continuous-batching server: more requests than slots, admission by memory,
and every request's result identical (bit for bit) alone and in the mix."""
import importlib

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _small_stream(n, seed):
    Sv = importlib.import_module("2018_dictionary_tomography_amd.serving")
    specs = Sv.c5_stream(n, rate=500.0, seed=seed, n_range=(4, 40), angle_range=(6, 14), sides=(24, 32, 48),
                         max_iter=3)
    return Sv, specs


def test_server_alone_equals_mixed(gt, cuda):
    Sv, specs = _small_stream(14, seed=3)
    srv = Sv.FitServer(cuda, max_slots=5, work_per_launch=4.0e4, keep_results=True)
    st = srv.run(specs)
    assert st.completed == len(specs) and st.refused == 0
    assert st.launches > 3
    s = st.summary()
    assert s["p99_latency_s"] >= s["p50_latency_s"] > 0
    _, alone = Sv.plain_fits(specs, cuda, keep_results=True)
    for rid, (I, x, r, E, done, upd, drift) in alone.items():
        Im, xm, rm, Em, donem, updm = st.results[rid]
        assert done == donem == 1 and upd == updm
        assert torch.equal(I, Im) and torch.equal(x, xm) and torch.equal(r, rm)
        assert E == Em
        # the residual kept by incremental updates stays that of the final atoms
        assert drift < 1e-5


def test_server_groups_equal_single_workgroup(gt, cuda):
    """With angle counts that make the server give fits several workgroups,
    every result still equals the one-workgroup fit bit for bit."""
    Sv = importlib.import_module("2018_dictionary_tomography_amd.serving")
    specs = Sv.c5_stream(5, rate=500.0, seed=8, n_range=(4, 12), angle_range=(33, 48), sides=(24, 32), max_iter=2)
    srv = Sv.FitServer(cuda, max_slots=8, work_per_launch=2.0e5, keep_results=True)
    st = srv.run(specs)
    assert st.completed == len(specs) and st.failed == 0
    _, alone = Sv.plain_fits(specs, cuda, keep_results=True, group=1)
    for rid, (I, x, r, E, done, upd, drift) in alone.items():
        Im, xm, rm, Em, donem, updm = st.results[rid]
        assert done == donem == 1 and upd == updm
        assert torch.equal(I, Im) and torch.equal(x, xm) and torch.equal(r, rm) and E == Em


def test_server_grouped_mix_of_large_fits(gt, cuda):
    """Several fits of 60-100 angles in one launch, each advanced by 4-7
    workgroups whose timing drifts apart: every result still equals the
    one-workgroup fit bit for bit (members must not read state that another
    member of the same update may already have rewritten)."""
    Sv = importlib.import_module("2018_dictionary_tomography_amd.serving")
    specs = Sv.c5_stream(6, rate=1000.0, seed=21, n_range=(40, 90), angle_range=(60, 100), sides=(64, 96, 128),
                         max_iter=3)
    pay = Sv.prepare_payloads(specs, cuda)
    srv = Sv.FitServer(cuda, keep_results=True)
    st = srv.run(specs, payloads=pay)
    assert st.completed == len(specs) and st.failed == 0
    _, alone = Sv.plain_fits(specs, cuda, keep_results=True, group=1, payloads=pay)
    for rid, (I, x, r, E, done, upd, drift) in alone.items():
        Im, xm, rm, Em, donem, updm = st.results[rid]
        assert done == donem == 1 and upd == updm
        assert torch.equal(I, Im) and torch.equal(x, xm) and torch.equal(r, rm) and E == Em


def test_server_memory_admission(gt, cuda):
    Sv, specs = _small_stream(6, seed=4)
    big = max(s.footprint_bytes() for s in specs)
    # budget for at most one request in flight: requests are admitted one by one
    srv = Sv.FitServer(cuda, max_slots=4, hbm_budget=big, work_per_launch=4.0e4, use_graph=False)
    st = srv.run(specs)
    assert st.completed == len(specs)
    assert st.admitted_bytes_peak <= big
    # a request larger than the whole budget is refused, not queued forever
    srv2 = Sv.FitServer(cuda, max_slots=4, hbm_budget=big // 1000, work_per_launch=4.0e4, use_graph=False)
    st2 = srv2.run(specs[:2])
    assert st2.refused == 2 and st2.completed == 0
