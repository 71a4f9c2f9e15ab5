"""The benchmark configurations at their own sizes.

* 'batch' (64 uneven cases, sides 32-128): all 64 presets advance in one FP32
  batch; every case must be bit-identical to its own single-case run (host
  initialisation, single-case kernels), and two of them to the oracle.
* cfg1 (128^3, tau_g = 0.8, g stored as populations): 1000 steps of the
  fused path bit-identical to the FP32 transcription.
* cfg2 (256^3, two bubbles): for sigma in {0.01, 0.05, 0.1} the Laplace
  jump of each bubble in FP16 and BF16 storage agrees with the FP32 path to
  1 %; BF16 conserves the phase over 200 steps; both 16-bit forms commute
  with a one-cell shift at this size.
"""
import numpy as np
import pytest

from conftest import f32_bubbles, oracle_for, rel_l2

pytestmark = pytest.mark.gpu

OBS = ("phi", "rho", "ux", "uy", "uz")
CS2 = 1.0 / 3.0


def assert_same(got, want, label):
    for k in OBS:
        nbad = int(np.count_nonzero(got[k] != want[k]))
        assert nbad == 0, f"{label}: {k} differs in {nbad} cells"


def test_batch_presets_all_64_in_one_batch(gpu_lib):
    mcb = gpu_lib
    steps = 20
    cases = [mcb.case_preset("batch", 1000 + i) for i in range(64)]
    sides = sorted({c.params.nx for c in cases})
    assert set(sides) <= {32, 48, 64, 96, 128} and len(sides) >= 4
    with mcb.Batch(mcb.STORE_FP32) as b:
        ids = [b.add_case(c) for c in cases]
        b.step(steps)
        for i, (cid, c) in enumerate(zip(ids, cases)):
            assert b.case_steps(cid) == steps
            got = b.observables(cid)
            with mcb.Lattice.from_case(c) as lat:
                lat.step(steps)
                want = lat.observables()
            assert_same(got, want, f"case {i} (side {c.params.nx})")
            if i in (int(np.argmin([q.params.nx for q in cases])), int(np.argmax([q.params.nx for q in cases]))):
                o = oracle_for(c.params)
                o.init(f32_bubbles(mcb.case_bubbles(c)), c.width)
                o.step(steps)
                o.macros()
                assert_same(got, {k: o.field(k) for k in OBS}, f"case {i} (side {c.params.nx}) vs oracle")


def test_cfg1_1000_steps_match_the_oracle(gpu_lib):
    """128^3, seeded off-centre bubble R in [12, 24], tau_f 0.55, tau_g 0.8
    (g as populations), sigma 0.1: 1000 fused steps bit-identical to the
    transcription (so within the 1e-5 relative-L2 bar on phi, rho, u)."""
    mcb = gpu_lib
    c = mcb.case_preset("cfg1", 0)
    p = c.params
    assert (p.nx, p.ny, p.nz) == (128, 128, 128) and abs(p.tau_g - 0.8) < 1e-7
    steps = 1000
    with mcb.Lattice.from_case(c) as lat:
        lat.compute_macros()
        s0, _, _ = lat.reduce()
        lat.step(steps)
        got = lat.observables()
        s1, _, ok = lat.reduce()
    assert ok
    o = oracle_for(p)
    o.init(f32_bubbles(mcb.case_bubbles(c)), c.width)
    o.step(steps)
    o.macros()
    want = {k: o.field(k) for k in OBS}
    for k in OBS:
        assert rel_l2(got[k], want[k]) <= 1e-5, k
    assert_same(got, want, "cfg1 after 1000 steps")
    print(f"cfg1: sum phi relative change over {steps} steps {(s1 - s0) / s0:.2e}")


def cfg2_case(mcb, sigma):
    c = mcb.case_preset("cfg2", 0)
    c.params.sigma = sigma
    return c


def delta_p_per_bubble(rho, bubbles, width):
    """cs^2 (mean rho inside bubble i at r < R_i - 2W  -  mean rho of the
    liquid at r_j > R_j + 2W for every bubble j), minimum-image distances."""
    rho = rho.astype(np.float64)
    nz, ny, nx = rho.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    mi = lambda d, n: (d + n / 2) % n - n / 2
    rs = [np.sqrt(mi(x - b[0], nx) ** 2 + mi(y - b[1], ny) ** 2 + mi(z - b[2], nz) ** 2) for b in bubbles]
    outside = np.ones(rho.shape, bool)
    for r, b in zip(rs, bubbles):
        outside &= r > b[3] + 2 * width
    out = rho[outside].mean()
    return [CS2 * (rho[r < b[3] - 2 * width].mean() - out) for r, b in zip(rs, bubbles)]


@pytest.mark.parametrize("sigma", [0.01, 0.05, 0.1])
def test_cfg2_laplace_jump_of_16bit_storage_within_1pct_of_fp32(gpu_lib, sigma):
    mcb = gpu_lib
    steps = 1000
    c = cfg2_case(mcb, sigma)
    bubbles = mcb.case_bubbles(c)
    assert len(bubbles) == 2
    dps = {}
    for name, st in (("fp32", mcb.STORE_FP32), ("fp16", mcb.STORE_FP16), ("bf16", mcb.STORE_BF16)):
        with mcb.Lattice.from_case(c, storage=st) as lat:
            lat.step(steps)
            lat.compute_macros()
            dps[name] = delta_p_per_bubble(lat.read("rho"), bubbles, c.width)
    print(f"sigma {sigma}: dp per bubble fp32 {dps['fp32']}, fp16 {dps['fp16']}, bf16 {dps['bf16']}, "
          f"2 sigma / R {[2 * sigma / b[3] for b in bubbles]}")
    for name in ("fp16", "bf16"):
        for i, (a, ref) in enumerate(zip(dps[name], dps["fp32"])):
            assert ref > 0, "the jump should be positive"
            assert abs(a / ref - 1) <= 0.01, f"{name}, bubble {i}: dp {a:.6e} vs FP32 {ref:.6e}"


def test_cfg2_bf16_conserves_phase_over_200_steps(gpu_lib):
    mcb = gpu_lib
    c = cfg2_case(mcb, 0.1)
    with mcb.Lattice.from_case(c, storage=mcb.STORE_BF16) as lat:
        lat.compute_macros()
        s0, _, ok0 = lat.reduce()
        lat.step(200)
        lat.compute_macros()
        s1, _, ok1 = lat.reduce()
    assert ok0 and ok1
    drift = abs(s1 - s0) / abs(s0)
    print(f"bf16: relative drift of sum phi over 200 steps {drift:.2e}")
    assert drift <= 1e-7


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_cfg2_16bit_commutes_with_a_shift(gpu_lib, storage):
    mcb = gpu_lib
    st = mcb.STORE_FP16 if storage == "fp16" else mcb.STORE_BF16
    c = cfg2_case(mcb, 0.1)
    p = c.params
    p.storage = st
    phi0 = mcb.host_init_phi(p, mcb.case_bubbles(c), c.width)
    out = []
    for shifted in (False, True):
        with mcb.Lattice(p) as lat:
            lat.init_phi(np.roll(phi0, (1, 1, 1), axis=(2, 1, 0)) if shifted else phi0)
            lat.step(3)
            out.append(lat.observables())
    for k in OBS:
        want = np.roll(out[0][k], (1, 1, 1), axis=(2, 1, 0))
        nbad = int(np.count_nonzero(out[1][k] != want))
        assert nbad == 0, f"{storage}: {k} after a shift along x, y and z differs in {nbad} cells"
