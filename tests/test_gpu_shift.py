"""Translation equivariance of the fused step, bit for bit.

Every pass computes a cell from its own neighbourhood with the same
operations wherever the cell falls in a workgroup's tile (interior, tile
edge, halo ring, partial tile at the lattice edge, 16-bit staging pair).
The periodic lattice therefore commutes with shifts: starting from the
initial phi rolled by one cell along an axis, the state after k steps is the
unshifted run's state rolled the same way, in every bit. At tau_g = 1 no
global quantity enters the step (the 16-bit mass-drift correction belongs to
stored 16-bit g), so this holds for the FP32, FP16 and BF16 forms alike.
A one-cell shift along x moves every cell across the tile's halo ring and
flips the parity of the 16-bit phase pass's even-aligned staging pairs; the
odd row length exercises its single-cell staging. For FP32 the oracle tests
pin the values themselves; for the 16-bit forms this pins the data movement
of their kernels, whose values have no bit-exact oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OBS = ("phi", "rho", "ux", "uy", "uz")


def run(mcb, params, phi0, steps):
    with mcb.Lattice(params) as lat:
        lat.init_phi(phi0)
        lat.step(steps)
        return lat.observables()


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("shape", [(48, 40, 36), (47, 33, 21), (66, 24, 20), (80, 20, 12)])
def test_step_commutes_with_shifts(gpu_lib, storage, shape):
    mcb = gpu_lib
    st = {"fp32": mcb.STORE_FP32, "fp16": mcb.STORE_FP16, "bf16": mcb.STORE_BF16}[storage]
    nx, ny, nz = shape
    p = mcb.make_params(nx, ny, nz, storage=st, tau_f=0.55, tau_g=1.0, sigma=0.1, path=mcb.PATH_FUSED)
    bubbles = [(nx * 0.45 + 0.3, ny * 0.55 - 0.7, nz * 0.5 + 0.2, min(shape) * 0.25)]
    phi0 = mcb.host_init_phi(p, bubbles)
    steps = 6
    base = run(mcb, p, phi0, steps)
    for axis, name in ((2, "x"), (1, "y"), (0, "z")):
        got = run(mcb, p, np.roll(phi0, 1, axis=axis), steps)
        for k in OBS:
            want = np.roll(base[k], 1, axis=axis)
            nbad = int(np.count_nonzero(got[k] != want))
            assert nbad == 0, f"{storage} {shape}: {k} after a shift along {name} differs in {nbad} cells"


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
@pytest.mark.parametrize("shape", [(80, 20, 12), (47, 18, 10)])
def test_16bit_tilings_agree(gpu_lib, storage, shape):
    """Every tile shape and z-chunk of the 16-bit kernels gives the same bits
    (the full-row TX x 4 tiles included)."""
    mcb = gpu_lib
    st = {"fp16": mcb.STORE_FP16, "bf16": mcb.STORE_BF16}[storage]
    nx, ny, nz = shape
    p = mcb.make_params(nx, ny, nz, storage=st, tau_f=0.6, tau_g=1.0, sigma=0.05, path=mcb.PATH_FUSED)
    phi0 = mcb.host_init_phi(p, [(nx * 0.5 + 0.3, ny * 0.5 - 0.4, nz * 0.5, min(shape) * 0.3)])
    ref, used = None, []
    for tx, ty in mcb.tilings():
        for zc in (3, nz):
            with mcb.Lattice(p) as lat:
                lat.init_phi(phi0)
                try:
                    lat.set_tiling(tx, ty, zc)
                except mcb.McbError as e:   # a shape the 16-bit kernels are not built for
                    assert e.status == mcb.ERR_INVALID_ARG, e
                    continue
                lat.step(5)
                obs = lat.observables()
            used.append((tx, ty))
            if ref is None:
                ref = obs
            for k in OBS:
                nbad = int(np.count_nonzero(obs[k] != ref[k]))
                assert nbad == 0, f"{storage} {shape} tiling {tx}x{ty} zc={zc}: {k} differs in {nbad} cells"
    assert (80, 4) in used and (64, 8) in used
