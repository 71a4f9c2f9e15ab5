"""The benchmark case at its full size (256^3, two seeded bubbles, sigma 0.1).

At this size the oracle still finishes two steps in seconds on the host's
cores, so the fused FP32 path is compared with it bit for bit; longer runs
are checked through properties that do not depend on the size: the total
phase is conserved (the phase pass only moves phi between cells; FP64 sum),
every field stays finite, and the step commutes with a one-cell shift of the
periodic lattice.
"""
import numpy as np
import pytest

from conftest import f32_bubbles, oracle_for

pytestmark = pytest.mark.gpu

OBS = ("phi", "rho", "ux", "uy", "uz")


def cfg2(mcb):
    return mcb.case_preset("cfg2", 0)


def test_fp32_two_steps_match_the_oracle(gpu_lib):
    mcb = gpu_lib
    c = cfg2(mcb)
    assert (c.params.nx, c.params.ny, c.params.nz) == (256, 256, 256)
    with mcb.Lattice.from_case(c) as lat:
        o = oracle_for(c.params)
        o.init(f32_bubbles(mcb.case_bubbles(c)), c.width)
        lat.step(2)
        o.step(2)
        got = lat.observables()
        o.macros()
        for k in OBS:
            nbad = int(np.count_nonzero(got[k] != o.field(k)))
            assert nbad == 0, f"{k} differs from the oracle in {nbad} cells"


# Relative drift of the FP64 sum of phi after 200 steps, from phi as the
# reference's phiCalc gives it on the initial state. FP32: the transcription's
# own rounding (measured 6.5e-8; 9e-8 after 800 steps). 16-bit storage at
# tau_g = 1 (the (phi, B) moment form) carries a per-step mass-drift
# correction on an exact fixed-point sum (measured 1.8e-8; within 4e-8 over
# 800 steps; -3.9e-7 without the correction).
@pytest.mark.parametrize("storage,bound", [("fp32", 1e-7), ("fp16", 1e-7)])
def test_phase_is_conserved_over_200_steps(gpu_lib, storage, bound):
    mcb = gpu_lib
    st = {"fp32": mcb.STORE_FP32, "fp16": mcb.STORE_FP16}[storage]
    c = cfg2(mcb)
    with mcb.Lattice.from_case(c, storage=st) as lat:
        lat.compute_macros()
        s0, _, ok0 = lat.reduce()
        lat.step(200)
        lat.compute_macros()
        s1, ke, ok1 = lat.reduce()
        assert ok0 and ok1, "non-finite field"
        drift = abs(s1 - s0) / abs(s0)
        print(f"{storage}: sum phi {s0:.10g} -> {s1:.10g}, relative drift {drift:.2e}, KE {ke:.4g}")
        assert drift <= bound, f"{storage}: relative drift of sum phi {drift:.2e} over 200 steps"
        assert ke > 0


def test_fp32_commutes_with_a_shift(gpu_lib):
    mcb = gpu_lib
    c = cfg2(mcb)
    p = c.params
    phi0 = mcb.host_init_phi(p, mcb.case_bubbles(c), c.width)
    out = []
    for shifted in (False, True):
        with mcb.Lattice(p) as lat:
            lat.init_phi(np.roll(phi0, (1, 1), axis=(2, 1)) if shifted else phi0)
            lat.step(3)
            out.append(lat.observables())
    for k in OBS:
        want = np.roll(out[0][k], (1, 1), axis=(2, 1))
        nbad = int(np.count_nonzero(out[1][k] != want))
        assert nbad == 0, f"{k} after a shift along x and y differs in {nbad} cells"
