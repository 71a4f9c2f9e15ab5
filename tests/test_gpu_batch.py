"""Batched execution (mcb_batch_*): many independent cases in one launch per pass.

Every case of a batch must evolve exactly as it does on its own: the batched
kernels run the same tile bodies as the single-case kernels, and cases share
no data. FP32 cases are therefore compared bit for bit with single-case
fused runs (which are themselves bit-exact against the oracle), and one case
is compared with the oracle directly. Two batched steps are replayed from a
hipGraph; the replays are checked against eager launches after every replay.
"""
import numpy as np
import pytest

from conftest import f32_bubbles, oracle_for

pytestmark = pytest.mark.gpu

OBS = ("phi", "rho", "ux", "uy", "uz")

# Uneven lattices with their own physics; (17, 33, 9) exercises partial tiles.
CASES = [
    ((24, 20, 16), dict(tau_f=0.55, tau_g=0.8, sigma=0.1)),
    ((40, 36, 20), dict(tau_f=0.6, tau_g=1.0, sigma=0.05)),
    ((17, 33, 9), dict(tau_f=0.52, tau_g=1.0, sigma=0.01)),
    ((64, 16, 12), dict(tau_f=0.7, tau_g=0.9, sigma=0.08)),
]


def case_params(mcb, shape, phys, storage):
    nx, ny, nz = shape
    return mcb.make_params(nx, ny, nz, storage=storage, path=mcb.PATH_FUSED, **phys)


def case_bubbles(shape):
    nx, ny, nz = shape
    return [(nx * 0.5 + 0.3, ny * 0.5 - 0.6, nz * 0.5 + 0.2, min(shape) * 0.3)]


def single_run(mcb, shape, phys, steps, storage):
    lat = mcb.Lattice(case_params(mcb, shape, phys, storage))
    lat.init_bubbles(case_bubbles(shape), 3.0)
    lat.step(steps)
    out = lat.observables()
    lat.close()
    return out


def assert_same(got, want, label):
    for k in OBS:
        nbad = int(np.count_nonzero(got[k] != want[k]))
        assert nbad == 0, f"{label}: {k} differs in {nbad} cells"


def test_batch_fp32_matches_single_cases(gpu_lib):
    mcb = gpu_lib
    steps = 13  # six graph replays of two steps, then one eager step
    with mcb.Batch(mcb.STORE_FP32) as b:
        ids = [b.add(case_params(mcb, s, ph, mcb.STORE_FP32), case_bubbles(s)) for s, ph in CASES]
        assert len(b) == len(CASES)
        b.step(steps)
        for cid, (shape, phys) in zip(ids, CASES):
            assert b.case_steps(cid) == steps
            assert_same(b.observables(cid), single_run(mcb, shape, phys, steps, mcb.STORE_FP32), f"case {shape}")


def test_batch_case_matches_oracle(gpu_lib):
    mcb = gpu_lib
    shape, phys = CASES[2]
    p = case_params(mcb, shape, phys, mcb.STORE_FP32)
    o = oracle_for(p)
    o.init(f32_bubbles(case_bubbles(shape)), 3.0)
    o.step(9)
    o.macros()
    with mcb.Batch(mcb.STORE_FP32) as b:
        b.add(case_params(mcb, *CASES[0], mcb.STORE_FP32), case_bubbles(CASES[0][0]))
        cid = b.add(p, case_bubbles(shape))
        b.step(9)
        got = b.observables(cid)
    assert_same(got, {k: o.field(k) for k in OBS}, "batched case vs oracle")


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_case_alone_equals_in_the_mix(gpu_lib, storage):
    """A case that joins a running batch, sees others leave, and finishes gives
    the same numbers as the same case alone."""
    mcb = gpu_lib
    st = mcb.STORE_FP32 if storage == "fp32" else mcb.STORE_FP16
    shape, phys = CASES[0]
    with mcb.Batch(st) as alone:
        a = alone.add(case_params(mcb, shape, phys, st), case_bubbles(shape))
        alone.step(13)
        want = alone.observables(a)
    with mcb.Batch(st) as mix:
        y = mix.add(case_params(mcb, *CASES[1], st), case_bubbles(CASES[1][0]))
        z = mix.add(case_params(mcb, *CASES[3], st), case_bubbles(CASES[3][0]))
        mix.step(5)  # odd: the new case joins at the other A/B parity
        x = mix.add(case_params(mcb, shape, phys, st), case_bubbles(shape))
        mix.step(7)
        mix.remove(y)
        mix.step(6)
        assert mix.case_steps(x) == 13 and mix.case_steps(z) == 18
        got = mix.observables(x)
    assert_same(got, want, f"{storage} alone vs in the mix")


def test_graph_replays_match_eager_steps(gpu_lib):
    mcb = gpu_lib
    with mcb.Batch(mcb.STORE_FP32) as g, mcb.Batch(mcb.STORE_FP32) as e:
        gid = [g.add(case_params(mcb, s, ph, mcb.STORE_FP32), case_bubbles(s)) for s, ph in CASES]
        eid = [e.add(case_params(mcb, s, ph, mcb.STORE_FP32), case_bubbles(s)) for s, ph in CASES]
        for replay in range(4):
            g.step(2)          # one graph replay
            e.step(1)          # eager launches
            e.step(1)
            for a, c in zip(gid, eid):
                pa, pc = g.read(a, "phi"), e.read(c, "phi")
                nbad = int(np.count_nonzero(pa != pc))
                assert nbad == 0, f"replay {replay}: phi differs in {nbad} cells"
        for a, c in zip(gid, eid):
            assert_same(g.observables(a), e.observables(c), "after replays")


def test_batch_reduce_and_populations(gpu_lib):
    mcb = gpu_lib
    shape, phys = CASES[1]
    lat = mcb.Lattice(case_params(mcb, shape, phys, mcb.STORE_FP32))
    lat.init_bubbles(case_bubbles(shape), 3.0)
    lat.step(4)
    lat.observables()
    s_ref, k_ref, _ = lat.reduce()
    f_ref = lat.read("f")
    with mcb.Batch(mcb.STORE_FP32) as b:
        cid = b.add(case_params(mcb, shape, phys, mcb.STORE_FP32), case_bubbles(shape))
        b.step(4)
        b.compute_macros(cid)
        s, k, ok = b.reduce(cid)
        assert ok and s == s_ref and k == k_ref
        f = b.read(cid, "f")
        assert np.array_equal(f, f_ref)
        # Populations written back reproduce the state exactly.
        g = b.read(cid, "g")
        b.write(cid, "f", f)
        b.write(cid, "g", g)
        assert np.array_equal(b.read(cid, "f"), f) and np.array_equal(b.read(cid, "g"), g)
    lat.close()


def test_batch_rejections(gpu_lib):
    mcb = gpu_lib
    shape, phys = CASES[0]
    with mcb.Batch(mcb.STORE_FP32) as b:
        with pytest.raises(mcb.McbError) as ei:
            b.add(case_params(mcb, shape, phys, mcb.STORE_FP16), case_bubbles(shape))
        assert ei.value.status == mcb.ERR_UNSUPPORTED
        p = case_params(mcb, shape, phys, mcb.STORE_FP32)
        p.bc_mode = mcb.BC_REFERENCE
        with pytest.raises(mcb.McbError) as ei:
            b.add(p, case_bubbles(shape))
        assert ei.value.status == mcb.ERR_UNSUPPORTED
        cid = b.add(case_params(mcb, shape, phys, mcb.STORE_FP32), case_bubbles(shape))
        with pytest.raises(mcb.McbError):
            b.read(cid, "rho")  # macros not computed yet
        b.compute_macros(cid)
        b.read(cid, "rho")
        b.step(1)
        with pytest.raises(mcb.McbError):
            b.read(cid, "rho")  # stale after a step
        with pytest.raises(mcb.McbError) as ei:
            b.remove(cid + 100)
        assert ei.value.status == mcb.ERR_INVALID_ARG
        with pytest.raises(mcb.McbError):
            b.set_tiling(cid, 48, 8, 4)
        b.set_tiling(cid, 16, 16, 4)
        b.step(3)
        assert b.case_steps(cid) == 4
        b.remove(cid)
        assert len(b) == 0
        b.step(2)  # an empty batch steps nothing


def test_tilings_and_autotune_agree(gpu_lib):
    mcb = gpu_lib
    shape, phys = CASES[1]
    outs = []
    for t in ((64, 8, 16), (32, 8, 4), (16, 16, 7), (64, 4, 20)):
        with mcb.Batch(mcb.STORE_FP32) as b:
            cid = b.add(case_params(mcb, shape, phys, mcb.STORE_FP32), case_bubbles(shape))
            b.set_tiling(cid, *t)
            b.step(5)
            outs.append(b.observables(cid))
    for o in outs[1:]:
        assert_same(o, outs[0], "tilings")
    tx, ty, zc, ms = mcb.autotune(case_params(mcb, (48, 48, 48), phys, mcb.STORE_FP32))
    assert (tx, ty) in mcb.tilings() and 1 <= zc <= 48 and ms > 0
