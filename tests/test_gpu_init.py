"""Initial bubble field computed on the device against the host and the oracle.

The serving loop and mcb_batch_add compute the initial phi on the GPU
(k_bubble_phi), and the operator face exposes initPhase (src/init.cu:27-49)
as a kernel; the host initialisation (mcb_host_init_phi) and the oracle
compute it on the CPU. All of them evaluate the same profile with the same
IEEE-only tanh, so they must agree bit for bit; one ulp of phi grows past
the parity tolerance within a few hundred steps (the model is chaotic at the
rounding level). Served FP32 requests are compared with the oracle after
their steps, sequentially and batched.
"""
import ctypes

import numpy as np
import pytest

from conftest import f32_bubbles, oracle_for

pytestmark = pytest.mark.gpu

OBS = ("phi", "rho", "ux", "uy", "uz")
SERVE_SIDES = list(range(48, 193, 16))


def serve_case_of_side(mcb, side):
    """The first 'serve' preset (seeds 0, 1, ...) with cube side `side`."""
    for seed in range(2000):
        c = mcb.case_preset("serve", seed)
        if c.params.nx == side:
            return c
    raise AssertionError(f"no serve preset of side {side}")


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_device_bubble_phi_equals_host_init_for_every_serve_side(gpu_lib, storage):
    mcb = gpu_lib
    st = mcb.STORE_FP32 if storage == "fp32" else mcb.STORE_FP16
    with mcb.Batch(st) as b:
        for side in SERVE_SIDES:
            c = serve_case_of_side(mcb, side)
            p = c.params
            p.storage = st
            cid = b.add(p, mcb.case_bubbles(c), c.width)
            got = b.read(cid, "phi")
            want = mcb.host_init_phi(p, mcb.case_bubbles(c), c.width)
            nbad = int(np.count_nonzero(got != want))
            assert nbad == 0, f"side {side}: device phi differs from the host's in {nbad} cells"
            b.remove(cid)


@pytest.mark.parametrize("bc", ["periodic", "reference"])
def test_initPhase_op_equals_host_init_and_oracle(gpu_lib, bc):
    mcb = gpu_lib
    nx, ny, nz = 37, 30, 26
    bcm = mcb.BC_PERIODIC if bc == "periodic" else mcb.BC_REFERENCE
    p = mcb.make_params(nx, ny, nz, bc_mode=bcm, weights=mcb.WEIGHTS_LITERAL)
    bubbles = [(5.3, 14.2, 12.9, 7.5), (30.1, 3.7, 20.2, 6.25)]   # near faces: wraps in periodic mode
    consts = mcb.op_consts(p)
    consts.n_bubbles = len(bubbles)
    for i, bb in enumerate(bubbles):
        for j in range(4):
            consts.bubbles[4 * i + j] = bb[j]
    consts.width = 3.0
    phi = mcb.DeviceArray.from_numpy(np.full((nz, ny, nx), -1.0, np.float32))
    if bcm == mcb.BC_REFERENCE:   # faces keep the buffer's contents (the reference's memset 0)
        z = np.zeros((nz, ny, nx), np.float32)
        mcb.check(mcb.lib().mcb_memcpy_htod(phi.ptr, z.ctypes.data_as(ctypes.c_void_p), z.nbytes))
    mcb.op("initPhase", phi, nx, ny, nz, consts)
    got = phi.to_numpy((nz, ny, nx))
    phi.free()
    want = mcb.host_init_phi(p, bubbles, 3.0)
    assert np.array_equal(got, want), f"{int(np.count_nonzero(got != want))} cells differ from the host init"
    o = oracle_for(p)
    o.init(f32_bubbles(bubbles), 3.0)
    assert np.array_equal(got, o.field("phi"))


def requests_of(mcb, cases, steps):
    reqs = (mcb.Request * len(cases))()
    bufs = []
    for i, c in enumerate(cases):
        ctypes.memmove(ctypes.byref(reqs[i].c), ctypes.byref(c), ctypes.sizeof(c))
        reqs[i].c.nsteps = steps
        reqs[i].arrival_s = 0.0
        p = c.params
        b = np.zeros((p.nz, p.ny, p.nx), dtype=np.float32)
        reqs[i].phi_out = b.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        bufs.append(b)
    return reqs, bufs


def oracle_phi(case, steps):
    o = oracle_for(case.params)
    o.init(f32_bubbles([(b.cx, b.cy, b.cz, b.r) for b in case.bubbles[:case.n_bubbles]]), case.width)
    o.step(steps)
    o.macros()
    return o.field("phi").copy(), o


@pytest.mark.parametrize("sequential", [True, False], ids=["sequential", "batched"])
def test_served_fp32_requests_match_the_oracle(gpu_lib, sequential):
    """Three FP32 requests (two of the same side, with different physics: the
    sequential loop re-targets one context at both) served on the device;
    each final phi and sum of phi equals the oracle's after its steps."""
    mcb = gpu_lib
    c48 = serve_case_of_side(mcb, 48)
    c64 = serve_case_of_side(mcb, 64)
    other48 = None
    for seed in range(2000):
        c = mcb.case_preset("serve", seed)
        if c.params.nx == 48 and (c.params.tau_f, c.params.sigma) != (c48.params.tau_f, c48.params.sigma):
            other48 = c
            break
    assert other48 is not None
    cases = [c48, c64, other48]
    steps = 40
    reqs, bufs = requests_of(mcb, cases, steps)
    res, stats = mcb.serve(reqs, storage=mcb.STORE_FP32, sequential=sequential, tick_steps=16)
    assert stats.completed == len(cases)
    for i, c in enumerate(cases):
        assert res[i].status == mcb.OK
        want, o = oracle_phi(c, steps)
        nbad = int(np.count_nonzero(bufs[i] != want))
        assert nbad == 0, f"request {i} (side {c.params.nx}): phi differs from the oracle in {nbad} cells"
        s_or = float(np.sum(want.astype(np.float64)))
        assert abs(res[i].sum_phi - s_or) <= 1e-9 * abs(s_or)
