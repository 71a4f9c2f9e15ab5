"""Serving loop (mcb_serve_run): requests arriving over time share the device
as one continuously re-formed batch.

Each request's result must be the same alone and in the mix. Cases never
exchange data and every kernel is deterministic, so the stated tolerance is
zero: phi and the FP64 diagnostics are compared bit for bit.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_trace(mcb, n, rate=400.0, step_scale=0.04, seed=3):
    """A short serving trace: the serving workload's geometry with fewer steps."""
    return mcb.serve_trace(n, rate, step_scale, seed)


def with_outputs(reqs):
    bufs = []
    for q in reqs:
        p = q.c.params
        b = np.zeros(p.nx * p.ny * p.nz, dtype=np.float32)
        q.phi_out = b.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        bufs.append(b)
    return bufs


def one(mcb, q):
    arr = (mcb.Request * 1)()
    ctypes.memmove(ctypes.byref(arr[0]), ctypes.byref(q), ctypes.sizeof(q))
    arr[0].arrival_s = 0.0
    return arr


@pytest.mark.parametrize("storage", ["fp16", "fp32"])
def test_each_result_is_the_same_alone_and_in_the_mix(gpu_lib, storage):
    mcb = gpu_lib
    st = mcb.STORE_FP16 if storage == "fp16" else mcb.STORE_FP32
    reqs = small_trace(mcb, 10)
    bufs = with_outputs(reqs)
    res, stats = mcb.serve(reqs, storage=st, tick_steps=16)
    assert stats.completed == 10 and stats.rejected == 0
    assert stats.peak_cases >= 2, "the trace should overlap requests on the device"
    for i in (0, 4, 9):
        alone = one(mcb, reqs[i])
        b = np.zeros_like(bufs[i])
        alone[0].phi_out = b.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        r1, s1 = mcb.serve(alone, storage=st, tick_steps=7)
        assert r1[0].status == mcb.OK and res[i].status == mcb.OK
        nbad = int(np.count_nonzero(b != bufs[i]))
        assert nbad == 0, f"request {i}: phi differs in {nbad} cells between alone and in the mix"
        assert r1[0].sum_phi == res[i].sum_phi and r1[0].kinetic == res[i].kinetic
        assert r1[0].lattice_updates == res[i].lattice_updates


def test_admission_under_a_small_budget_keeps_results(gpu_lib):
    mcb = gpu_lib
    reqs = small_trace(mcb, 8, seed=5)
    bufs = with_outputs(reqs)
    res, stats = mcb.serve(reqs, storage=mcb.STORE_FP16)
    sizes = [mcb.case_bytes(mcb_params(mcb, q)) for q in reqs]
    budget = max(sizes) + min(sizes)      # the largest case fits, with little room beside it
    reqs2 = small_trace(mcb, 8, seed=5)
    bufs2 = with_outputs(reqs2)
    res2, stats2 = mcb.serve(reqs2, storage=mcb.STORE_FP16, budget=budget)
    assert stats2.completed == 8
    assert stats2.peak_bytes <= budget
    assert stats2.peak_cases < stats.peak_cases or stats.peak_cases <= 2
    for i in range(8):
        assert np.array_equal(bufs[i], bufs2[i]), i
        assert res2[i].admit_s >= res2[i].arrival_s
        assert res2[i].finish_s >= res2[i].admit_s


def mcb_params(mcb, q):
    p = mcb.Params()
    ctypes.memmove(ctypes.byref(p), ctypes.byref(q.c.params), ctypes.sizeof(p))
    p.storage = mcb.STORE_FP16
    return p


def test_oversized_request_is_rejected_and_the_rest_served(gpu_lib):
    mcb = gpu_lib
    reqs = small_trace(mcb, 6, seed=7)
    sizes = [mcb.case_bytes(mcb_params(mcb, q)) for q in reqs]
    big = int(np.argmax(sizes))
    budget = sorted(sizes)[-2]            # everything but the largest fits
    if sizes[big] == budget:
        pytest.skip("trace has two largest requests of the same size")
    res, stats = mcb.serve(reqs, storage=mcb.STORE_FP16, budget=budget)
    assert res[big].status == mcb.ERR_CAPACITY
    assert stats.rejected == 1 and stats.completed == 5
    for i, r in enumerate(res):
        if i != big:
            assert r.status == mcb.OK


def test_sequential_plain_matches_batched_fp32(gpu_lib):
    """The baseline (one request at a time on the single-case fused path) and
    the batched loop compute the same FP32 numbers."""
    mcb = gpu_lib
    reqs = small_trace(mcb, 5, seed=11)
    bufs = with_outputs(reqs)
    res, stats = mcb.serve(reqs, storage=mcb.STORE_FP32)
    reqs2 = small_trace(mcb, 5, seed=11)
    bufs2 = with_outputs(reqs2)
    res2, stats2 = mcb.serve(reqs2, storage=mcb.STORE_FP32, sequential=True)
    assert stats2.completed == 5 and stats2.peak_cases == 1
    for i in range(5):
        assert np.array_equal(bufs[i], bufs2[i]), i
        assert res[i].sum_phi == res2[i].sum_phi


def test_serving_statistics(gpu_lib):
    mcb = gpu_lib
    reqs = small_trace(mcb, 12, seed=13)
    res, stats = mcb.serve(reqs, storage=mcb.STORE_FP16)
    ttr = sorted(r.finish_s - r.arrival_s for r in res)
    assert stats.completed == 12
    assert stats.p50_s == ttr[5] and stats.p99_s == ttr[11] and stats.max_s == ttr[11]
    lu = sum(r.lattice_updates for r in res)
    assert stats.lattice_updates == lu
    assert abs(stats.mlups - lu / stats.makespan_s / 1e6) < 1e-6 * stats.mlups
    for r in res:
        assert np.isfinite(r.sum_phi) and r.sum_phi > 0
