"""Host side of batching and serving: the workload generators and the
argument checks of mcb_serve_run (no GPU needed)."""
import ctypes
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def mcb():
    import mcb as m
    return m


def test_batch_preset_draws(mcb):
    sides, radii = [], []
    for i in range(400):
        c = mcb.case_preset("batch", 1000 + i)
        p = c.params
        assert p.nx == p.ny == p.nz and p.nx in (32, 48, 64, 96, 128)
        assert p.bc_mode == mcb.BC_PERIODIC and p.storage == mcb.STORE_FP32
        assert 0.01 <= p.sigma <= 0.1 and 0.52 <= p.tau_f <= 0.8 and p.tau_g == 1.0
        assert c.nsteps == 2000 and c.n_bubbles == 1
        b = c.bubbles[0]
        assert 0.15 * p.nx <= b.r <= 0.3 * p.nx
        assert 0 <= b.cx < p.nx and 0 <= b.cy < p.ny and 0 <= b.cz < p.nz
        sides.append(p.nx)
    assert set(sides) == {32, 48, 64, 96, 128}
    # Same seed, same case.
    a, b = mcb.case_preset("batch", 7), mcb.case_preset("batch", 7)
    assert bytes(a) == bytes(b)


def test_serve_trace_draws(mcb):
    n, rate = 500, 4.0
    reqs = mcb.serve_trace(n, rate, 1.0, 0)
    sides = np.array([q.c.params.nx for q in reqs])
    steps = np.array([q.c.nsteps for q in reqs])
    arr = np.array([q.arrival_s for q in reqs])
    assert set(sides) <= set(range(48, 193, 16)) and len(set(sides)) == 10
    assert steps.min() >= 500 and steps.max() <= 5000
    # log-uniform: the median is near sqrt(500 * 5000)
    assert abs(math.log(np.median(steps) / math.sqrt(500 * 5000))) < 0.2
    for q in reqs:
        p = q.c.params
        assert p.storage == mcb.STORE_FP16 and p.nx == p.ny == p.nz
        assert 0.1 * p.nx <= q.c.bubbles[0].r <= 0.3 * p.nx
    assert np.all(np.diff(arr) > 0) and arr[0] > 0
    # Poisson arrivals: mean gap 1/rate, coefficient of variation near 1.
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert abs(gaps.mean() * rate - 1) < 0.15
    assert abs(gaps.std() / gaps.mean() - 1) < 0.15
    # Deterministic, and the step scale only scales steps.
    again = mcb.serve_trace(n, rate, 0.1, 0)
    assert [q.arrival_s for q in again] == list(arr)
    # (C lround: halves away from zero)
    assert all(b.c.nsteps == max(1, math.floor(a.c.nsteps * 0.1 + 0.5)) for a, b in zip(reqs, again))


def test_serve_run_rejects_bad_arguments(mcb):
    reqs = mcb.serve_trace(3, 10.0, 0.01, 1)
    cfg = mcb.ServeConfig()
    res = (mcb.RequestResult * 3)()
    st = mcb.ServeStats()
    lib = mcb.lib()
    assert lib.mcb_serve_run(None, reqs, 3, res, ctypes.byref(st)) == mcb.ERR_INVALID_ARG
    cfg.storage = 7
    assert lib.mcb_serve_run(ctypes.byref(cfg), reqs, 3, res, ctypes.byref(st)) == mcb.ERR_INVALID_ARG
    cfg.storage = mcb.STORE_FP16
    reqs[2].arrival_s = 0.0   # not sorted by arrival
    assert lib.mcb_serve_run(ctypes.byref(cfg), reqs, 3, res, ctypes.byref(st)) == mcb.ERR_INVALID_ARG
    assert lib.mcb_serve_trace(ctypes.c_uint64(0), 3, 0.0, 1.0, reqs) == mcb.ERR_INVALID_ARG
    assert lib.mcb_serve_trace(ctypes.c_uint64(0), 3, 1.0, -1.0, reqs) == mcb.ERR_INVALID_ARG


def test_case_bytes(mcb):
    # Every plane and field of a case is n cells rounded up to 256, plus one
    # 256-cell block of skew between planes.
    n = 64 ** 3 + 256
    # tau_g != 1: f and g populations (A/B); 16 bits add a second phi (g anchor)
    p = mcb.make_params(64, 64, 64, tau_g=0.8)
    assert mcb.case_bytes(p) == 4 * 19 * n * 4 + n * 4
    p.storage = mcb.STORE_FP16
    assert mcb.case_bytes(p) == 4 * 19 * n * 2 + 2 * n * 4
    # tau_g = 1: f populations and g in moment form: two phi, and u, n (FP32)
    # or B = 3 phi u + phi_n n (FP16, 16-bit forms)
    p = mcb.make_params(64, 64, 64, tau_g=1.0)
    assert mcb.case_bytes(p) == 2 * 19 * n * 4 + 8 * n * 4
    p.storage = mcb.STORE_FP16
    assert mcb.case_bytes(p) == 2 * 19 * n * 2 + 2 * n * 4 + 3 * n * 2
    p.tau_f = 0.4
    assert mcb.case_bytes(p) == 0   # invalid parameters
