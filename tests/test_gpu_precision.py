"""16-bit population storage (FP16 / BF16 offsets, FP32 arithmetic) against
an FP64 run of the oracle.

The model is chaotic at the rounding level (a one-ulp change grows to
percent-level differences in u over 1000 steps), so the per-field error of a
reduced-precision run is measured against the FP32 transcription's own distance
from FP64 on the same case: the 16-bit forms may be at most RATIO times
further from FP64 than FP32 is (plus a small floor). The physical observable,
the Laplace pressure jump, must agree with FP64 to 1 %, and the phase must be
conserved.
"""
import numpy as np
import pytest

from conftest import f32_bubbles, oracle_for, rel_l2

pytestmark = pytest.mark.gpu

CS2 = 1.0 / 3.0
DP_TOL = 0.01          # |dp / dp_fp64 - 1|
# Per-field bound: RATIO x (FP32's distance from FP64) + FLOOR[field]. BF16
# keeps 8 significand bits to FP16's 11 and f is stored as offsets from the
# lattice weights, so in smooth, quiet flows (tau_f = 1, spurious currents of
# ~1e-4) its u carries several percent of error where FP32 is far closer to
# FP64; its bounds state that. Both forms must meet the Laplace-jump bar.
RATIO = {"fp16": 4.0, "bf16": 16.0}
FLOOR = {"fp16": {"phi": 2e-3, "rho": 2e-3, "u": 2e-3},
         "bf16": {"phi": 1e-2, "rho": 1e-2, "u": 0.15}}
SUM_PHI_TOL = 1e-6     # relative change of sum(phi) over the run

# (n, R, sigma, tau_f, tau_g, steps): Laplace cases (near equilibrium) and the
# 64^3 plain-method case, with g stored as 16-bit populations (tau_g != 1)
# and in FP32 moment form (tau_g = 1).
CASES = [
    (48, 10.0, 0.1, 1.0, 0.8, 800),
    (48, 10.0, 0.01, 1.0, 0.8, 800),
    (64, 10.0, 0.1, 0.55, 0.8, 1000),
    (48, 10.0, 0.1, 1.0, 1.0, 800),
    (48, 10.0, 0.01, 0.7, 1.0, 800),
    (64, 10.0, 0.1, 0.52, 1.0, 1000),
]


def delta_p(rho, c, R, width=3.0):
    rho = rho.astype(np.float64)
    nz, ny, nx = rho.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    mi = lambda d, n: (d + n / 2) % n - n / 2
    r = np.sqrt(mi(x - c[0], nx) ** 2 + mi(y - c[1], ny) ** 2 + mi(z - c[2], nz) ** 2)
    return CS2 * (rho[r < R - 2 * width].mean() - rho[r > R + 2 * width].mean())


def errors(got, ref):
    u_ref = np.sqrt(sum(ref[k] ** 2 for k in ("ux", "uy", "uz")))
    du = np.sqrt(sum(np.sum((got[k].astype(np.float64) - ref[k]) ** 2) for k in ("ux", "uy", "uz")))
    return {"phi": rel_l2(got["phi"], ref["phi"]),
            "rho": rel_l2(got["rho"].astype(np.float64) - 1.0, ref["rho"] - 1.0),
            "u": float(du / np.linalg.norm(u_ref))}


_cache = {}


def fp64_and_fp32(mcb, case):
    if case not in _cache:
        n, R, sigma, tf, tg, steps = case
        c = (n / 2, n / 2, n / 2)
        p = mcb.make_params(n, n, n, tau_f=tf, tau_g=tg, sigma=sigma, path=mcb.PATH_FUSED)
        o = oracle_for(p, prec=64)
        o.init(f32_bubbles([(*c, R)]), 3.0)
        o.step(steps)
        o.macros()
        ref = {k: o.field(k).astype(np.float64) for k in ("phi", "rho", "ux", "uy", "uz")}
        lat = mcb.Lattice(p)
        lat.init_bubbles([(*c, R)], 3.0)
        lat.step(steps)
        e32 = errors(lat.observables(), ref)
        lat.close()
        _cache[case] = (ref, delta_p(ref["rho"], c, R), e32)
    return _cache[case]


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c[0]}-sigma{c[2]}-taug{c[4]}")
def test_16bit_storage_against_fp64(gpu_lib, storage, case):
    mcb = gpu_lib
    n, R, sigma, tf, tg, steps = case
    c = (n / 2, n / 2, n / 2)
    ref, dp64, e32 = fp64_and_fp32(mcb, case)
    st = mcb.STORE_FP16 if storage == "fp16" else mcb.STORE_BF16
    p = mcb.make_params(n, n, n, tau_f=tf, tau_g=tg, sigma=sigma, storage=st)
    lat = mcb.Lattice(p)
    lat.init_bubbles([(*c, R)], 3.0)
    lat.observables()
    s0, _, ok0 = lat.reduce()
    lat.step(steps)
    got = lat.observables()
    s1, _, ok1 = lat.reduce()
    lat.close()
    assert ok0 and ok1
    e = errors(got, ref)
    dp = delta_p(got["rho"], c, R)
    print(f"{storage} {case}: errors {e}, fp32 errors {e32}, dp {dp:.6e} fp64 {dp64:.6e}, "
          f"sum phi rel change {(s1 - s0) / s0:.2e}")
    assert abs(dp / dp64 - 1) <= DP_TOL, f"Laplace jump {dp} vs FP64 {dp64}"
    assert abs(s1 - s0) / s0 <= SUM_PHI_TOL
    for k in e:
        assert e[k] <= RATIO[storage] * e32[k] + FLOOR[storage][k], f"{k}: {e[k]:.3e} vs FP32's {e32[k]:.3e}"


CX = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0]
CY = [0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1]
CZ = [0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1]
# Significand bits of the 16-bit forms (FP16 11, BF16 8).
SIG_BITS = {"fp16": 11, "bf16": 8}


@pytest.mark.parametrize("tau_g", [0.8, 1.0])
@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_16bit_one_step_matches_fp64_on_a_strong_random_state(gpu_lib, storage, tau_g):
    """One step from a random state far from equilibrium: velocities up to
    0.15 per component, a random non-equilibrium part, and a strong
    surface-tension force (sigma = 1 on a sharp bubble of radius 4.5, |F| up
    to ~0.1). The 16-bit form's post-step populations equal the FP64 oracle's,
    run from the same decoded state, to the 16-bit rounding of the stored
    offsets (plus, on the rest link, the folded residuals of the other 18).
    Here the cubic He-forcing terms (27/4 (c.u)^2 (c.F) w_l ~ 1e-3), the
    velocity-force products and the regularised Pi terms are each one to two
    orders above that bound, so a wrong coefficient or sign in the 16-bit
    arithmetic fails this test, which multi-step comparisons near equilibrium
    would not show."""
    mcb = gpu_lib
    st = mcb.STORE_FP16 if storage == "fp16" else mcb.STORE_BF16
    shape = (20, 18, 16)                      # nx, ny, nz
    nz, ny, nx = shape[::-1]
    p = mcb.make_params(*shape, tau_f=0.7, tau_g=tau_g, sigma=1.0, storage=st)
    lat = mcb.Lattice(p)
    lat.init_bubbles([(9.6, 8.3, 7.7, 4.5)], 1.5)
    o = oracle_for(p, prec=64)
    w = o.weights().reshape(19, 1, 1, 1)
    rng = np.random.default_rng(20261202)
    u = rng.uniform(-0.15, 0.15, size=(3, nz, ny, nx))
    rho = 1.0 + rng.uniform(-0.05, 0.05, size=(nz, ny, nx))
    cu = sum(np.array(cc, dtype=np.float64).reshape(19, 1, 1, 1) * u[a] for a, cc in enumerate((CX, CY, CZ)))
    feq = w * rho * (1.0 + 3.0 * cu + 4.5 * cu ** 2 - 1.5 * (u ** 2).sum(axis=0))
    lat.write("f", feq + w * rng.normal(0.0, 0.02, size=feq.shape))
    f0, g0 = lat.read("f"), lat.read("g")
    o.init([(9.6, 8.3, 7.7, 4.5)], 1.5)
    o.field("f")[:] = f0
    o.field("g")[:] = g0
    lat.step(1)
    o.step(1)
    got, ref = lat.read("f").astype(np.float64), o.field("f").astype(np.float64)

    q = SIG_BITS[storage]
    off = np.abs(ref - w)
    bound = off * 2.0 ** (1 - q) + 2e-6       # one rounding of the offset, with margin
    bound[0] += (off[1:] * 2.0 ** -q).sum(axis=0)
    err = np.abs(got - ref)
    worst = np.unravel_index(np.argmax(err / bound), err.shape)
    print(f"{storage} tau_g {tau_g}: max |df| {err.max():.3e}, max err/bound {np.max(err / bound):.3f} "
          f"at {worst}, offsets up to {off.max():.3e}")
    assert np.all(err <= bound), f"link {worst[0]}: {err[worst]:.3e} > {bound[worst]:.3e}"

    # The macros of the new state (the collide pass's macro mode).
    obs = lat.observables()
    o.macros()
    tol = 4 * 2.0 ** -q * np.abs(off).max() * 19
    for k in ("rho", "ux", "uy", "uz"):
        assert np.max(np.abs(obs[k] - o.field(k))) <= tol, k
    lat.close()
    o.close()


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_16bit_moment_form_g_readout(gpu_lib, storage):
    """At tau_g = 1 the 16-bit forms keep g as (phi, B): its readout is
    initDist's w_l phi0 at init and tracks the FP32 moment form after steps;
    writing g switches the case to stored populations."""
    mcb = gpu_lib
    st = mcb.STORE_FP16 if storage == "fp16" else mcb.STORE_BF16
    shape = (24, 20, 16)
    bub = [(12.2, 9.7, 8.1, 5.0)]
    a = mcb.Lattice(mcb.make_params(*shape, tau_f=0.6, tau_g=1.0))
    b = mcb.Lattice(mcb.make_params(*shape, tau_f=0.6, tau_g=1.0, storage=st))
    a.init_bubbles(bub, 3.0)
    b.init_bubbles(bub, 3.0)
    assert np.array_equal(a.read("g"), b.read("g")), "initial g = w_l phi0 in both"
    a.step(4)
    b.step(4)
    ga, gb = a.read("g"), b.read("g")
    assert np.max(np.abs(ga - gb)) <= (2e-5 if storage == "fp16" else 2e-4)
    b.write("g", gb)
    b.step(2)
    a.step(2)
    assert rel_l2(b.observables()["phi"], a.observables()["phi"]) < 1e-4
    a.close()
    b.close()


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_16bit_populations_decode_to_the_fp32_state(gpu_lib, storage):
    """Right after init the decoded populations are the reference's initial
    populations (w_l and w_l phi) to 16-bit offset precision, and writing
    them back and reading again changes nothing beyond that precision."""
    mcb = gpu_lib
    st = mcb.STORE_FP16 if storage == "fp16" else mcb.STORE_BF16
    shape = (24, 20, 16)
    p32 = mcb.make_params(*shape, tau_f=0.6, tau_g=0.8)
    p16 = mcb.make_params(*shape, tau_f=0.6, tau_g=0.8, storage=st)
    bub = [(12.2, 9.7, 8.1, 5.0)]
    a, b = mcb.Lattice(p32), mcb.Lattice(p16)
    a.init_bubbles(bub, 3.0)
    b.init_bubbles(bub, 3.0)
    # Links 1..18 within the 16-bit rounding of their offsets; the rest link
    # also carries the cell's residual of the other 18 links.
    tol = (2e-4 if storage == "fp16" else 2e-3) / 18.0
    tol0 = 18 * tol

    def close(x, y, what):
        assert np.max(np.abs(x[1:] - y[1:])) <= tol, what
        assert np.max(np.abs(x[0] - y[0])) <= tol0, what

    for name in ("f", "g"):
        close(a.read(name), b.read(name), name)
    for name in ("f", "g"):
        y = b.read(name)
        b.write(name, y)
        close(b.read(name), y, f"{name} written back")
    a.step(6)
    b.step(6)
    ga, gb = a.observables(), b.observables()
    assert rel_l2(gb["phi"], ga["phi"]) < 1e-3
    a.close()
    b.close()
