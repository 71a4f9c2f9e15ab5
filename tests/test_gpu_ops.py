"""Operator face (include/mcb_ops.h) against the oracle, kernel by kernel.

Each reference kernel of src/kernels.cuh has a libmcb entry point of the same
name. Every test fills the reference's buffers with the same seeded fields on
both sides, runs one kernel on the GPU and the matching oracle stage on the
CPU, and requires every output buffer to be bit-identical.
"""
import zlib

import numpy as np
import pytest

from conftest import oracle_for

pytestmark = pytest.mark.gpu

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)

SCALARS = ["phi", "normx", "normy", "normz", "indicator", "curvature", "ffx", "ffy", "ffz",
           "ux", "uy", "uz", "rho", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz"]

# kernel -> buffer names in argument order (src/kernels.cuh)
SIGNATURES = {
    "initTensor": ["pxx", "pyy", "pzz", "pxy", "pxz", "pyz", "rho"],
    "initDist": ["rho", "phi", "f", "g"],
    "phiCalc": ["phi", "g"],
    "gradCalc": ["phi", "normx", "normy", "normz", "indicator"],
    "curvatureCalc": ["curvature", "indicator", "normx", "normy", "normz", "ffx", "ffy", "ffz"],
    "momentiCalc": ["ux", "uy", "uz", "rho", "ffx", "ffy", "ffz", "f", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz"],
    "collisionFluid": ["f", "ux", "uy", "uz", "ffx", "ffy", "ffz", "rho", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz"],
    "collisionPhase": ["g", "ux", "uy", "uz", "phi", "normx", "normy", "normz"],
    "streamingCalc": ["g", "g_out"],
    "fgBoundary": ["f", "rho", "g", "phi"],
    "boundaryConditions": ["phi"],
    "movingWallBC": ["f", "rho"],
}


def random_state(shape, seed):
    """Plausible LBM buffers: rho ~ 1, small u and forces, unit normals, phi in [0, 1]."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    s = {}
    s["phi"] = rng.uniform(0.0, 1.0, shape)
    n = rng.normal(size=(3,) + shape)
    n /= np.linalg.norm(n, axis=0)
    s["normx"], s["normy"], s["normz"] = n
    s["indicator"] = rng.uniform(0.0, 0.6, shape)
    s["curvature"] = rng.normal(0.0, 0.1, shape)
    for k in ("ffx", "ffy", "ffz"):
        s[k] = rng.normal(0.0, 1e-3, shape)
    for k in ("ux", "uy", "uz"):
        s[k] = rng.normal(0.0, 1e-2, shape)
    s["rho"] = 1.0 + rng.normal(0.0, 1e-2, shape)
    for k in ("pxx", "pyy", "pzz", "pxy", "pxz", "pyz"):
        s[k] = rng.normal(0.0, 1e-3, shape)
    w = W.reshape(19, 1, 1, 1)
    s["f"] = w * (1.0 + rng.normal(0.0, 1e-2, (19,) + shape))
    s["g"] = w * rng.uniform(0.0, 1.0, (19,) + shape)
    s["g_out"] = rng.normal(size=(19,) + shape)
    return {k: v.astype(np.float32) for k, v in s.items()}


def run_op(mcb, name, params, state):
    shape = (params.nz, params.ny, params.nx)
    o = oracle_for(params)
    for k, v in state.items():
        o.field(k)[...] = v
    dev = {k: mcb.DeviceArray.from_numpy(v) for k, v in state.items()}
    consts = mcb.op_consts(params)
    args = [dev[k] for k in SIGNATURES[name]]
    mcb.op(name, *args, params.nx, params.ny, params.nz, consts)
    o.stage(name)
    # streamingCalc on the oracle swaps g and g_out; compare the physical buffers.
    pairs = {k: k for k in state}
    if name == "streamingCalc":
        pairs = {"g": "g_out", "g_out": "g"}
    bad = []
    for k, ok in pairs.items():
        got = dev[k].to_numpy(state[k].shape)
        want = o.field(ok)
        if not np.array_equal(got, want):
            nb = int(np.count_nonzero(got != want))
            bad.append(f"{k}: {nb} cells differ (max abs {np.max(np.abs(got.astype(np.float64) - want)):.3e})")
    for d in dev.values():
        d.free()
    return bad


CASES = [
    ("periodic", (24, 20, 16), 0),
    ("periodic", (17, 9, 11), 0),
    ("reference", (16, 12, 20), 1),
]


@pytest.mark.parametrize("mode,shape,bc", CASES)
@pytest.mark.parametrize("name", [k for k in SIGNATURES if k not in ("fgBoundary", "boundaryConditions", "movingWallBC")])
def test_op_matches_oracle(gpu_lib, mode, shape, bc, name):
    mcb = gpu_lib
    nx, ny, nz = shape
    for tau_g in ((1.0, 0.8) if name == "collisionPhase" else (1.0,)):
        p = mcb.make_params(nx, ny, nz, bc_mode=bc, weights=mcb.WEIGHTS_LITERAL, tau_f=0.55, tau_g=tau_g, sigma=0.1)
        bad = run_op(mcb, name, p, random_state((nz, ny, nx), seed=zlib.crc32(f"{name}/{mode}".encode())))
        assert not bad, f"{name} ({mode}, tau_g={tau_g}): " + "; ".join(bad)


@pytest.mark.parametrize("name", ["fgBoundary", "boundaryConditions", "movingWallBC"])
def test_boundary_ops_match_oracle(gpu_lib, name):
    mcb = gpu_lib
    p = mcb.make_params(16, 12, 20, bc_mode=mcb.BC_REFERENCE, weights=mcb.WEIGHTS_LITERAL, tau_f=0.505)
    bad = run_op(mcb, name, p, random_state((20, 12, 16), seed=7))
    assert not bad, f"{name}: " + "; ".join(bad)


def test_op_rejects_bad_arguments(gpu_lib):
    mcb = gpu_lib
    p = mcb.make_params(8, 8, 8)
    c = mcb.op_consts(p)
    with pytest.raises(mcb.McbError):
        mcb.op("phiCalc", None, None, 8, 8, 8, c)
    d = mcb.DeviceArray(19 * 8 * 8 * 8 * 4)
    with pytest.raises(mcb.McbError):
        mcb.op("phiCalc", d, d, 2, 8, 8, c)          # side below 3
    with pytest.raises(mcb.McbError):
        mcb.op("streamingCalc", d, d, 8, 8, 8, c)    # in place streaming
    d.free()
