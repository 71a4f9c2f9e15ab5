"""The mcb driver end to end, and failure handling of the context API.

The driver keeps the reference's surface (src/main.cu): `mcb <VS> <ID>`
writes <root>bin/<VS>/<ID>/<ID>_info.txt and <ID>_phi<STEP:06d>.bin after
iteration STEP for STEP % save == 0. The reference preset runs the fused
path with ghost-layer boundaries (the staged kernels on request); its files
must equal the oracle's d_phi bit for bit and parse under the
post-processing rules.
"""
import os
import subprocess

import numpy as np
import pytest

import postfmt
from conftest import f32_bubbles, oracle_for

pytestmark = pytest.mark.gpu


def test_driver_reference_preset_matches_oracle(gpu_lib, tmp_path):
    mcb = gpu_lib
    root = str(tmp_path) + "/"
    r = subprocess.run([mcb.DRIVER_PATH, "D3Q19", "000", "1", "--preset", "reference", "--steps", "200",
                        "--root", root], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    sim = root + "bin/D3Q19/000/"
    info = postfmt.sim_info(sim)
    assert (info["NX"], info["NY"], info["NZ_TOTAL"], info["Nsteps"]) == (64, 64, 256, 200)
    assert postfmt.macr_steps(sim) == [0, 100, 200]
    c = mcb.case_preset("reference")
    o = oracle_for(c.params)
    o.init(f32_bubbles(mcb.case_bubbles(c)), c.width)
    done = 0
    for step in (0, 100, 200):
        o.step(step + 1 - done)        # the file of STEP holds d_phi after iteration STEP
        done = step + 1
        got = postfmt.read_macr(postfmt.macr_file(sim, "000", "phi", step), sim)
        want = np.swapaxes(o.field("phi"), 0, 2)
        assert np.array_equal(got, want), f"step {step}"


def test_driver_periodic_moment_form_matches_oracle(gpu_lib, tmp_path):
    """cfg0 with tau_g = 1 runs the fused path with g in moment form; its phi
    files equal the oracle's d_phi bit for bit."""
    mcb = gpu_lib
    root = str(tmp_path) + "/"
    r = subprocess.run([mcb.DRIVER_PATH, "D3Q19", "gm", "--preset", "cfg0", "--tau-g", "1", "--steps", "40",
                        "--save-every", "20", "--root", root], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    sim = root + "bin/D3Q19/gm/"
    assert postfmt.macr_steps(sim) == [0, 20, 40]
    c = mcb.case_preset("cfg0")
    c.params.tau_g = 1.0
    o = oracle_for(c.params)
    o.init(f32_bubbles(mcb.case_bubbles(c)), c.width)
    done = 0
    for step in (0, 20, 40):
        o.step(step + 1 - done)
        done = step + 1
        got = postfmt.read_macr(postfmt.macr_file(sim, "gm", "phi", step), sim)
        want = np.swapaxes(o.field("phi"), 0, 2)
        assert np.array_equal(got, want), f"step {step}"


def test_driver_periodic_preset_with_extra_fields(gpu_lib, tmp_path):
    mcb = gpu_lib
    root = str(tmp_path) + "/"
    r = subprocess.run([mcb.DRIVER_PATH, "D3Q19", "cfg0", "--preset", "cfg0", "--steps", "50", "--save-every",
                        "25", "--extra-fields", "--root", root], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    sim = root + "bin/D3Q19/cfg0/"
    assert postfmt.macr_steps(sim) == [0, 25, 50]     # the extra names do not match post's *phi* glob
    # After iteration STEP the driver writes d_phi and the observables of the
    # state (mcb_compute_macros: phiCalc..momentiCalc of iteration STEP + 1).
    c = mcb.case_preset("cfg0")
    o = oracle_for(c.params)
    o.init(f32_bubbles(mcb.case_bubbles(c)), c.width)
    done = 0
    for step in (0, 25, 50):
        o.step(step + 1 - done)
        done = step + 1
        phi = postfmt.read_macr(postfmt.macr_file(sim, "cfg0", "phi", step), sim)
        assert np.array_equal(phi, np.swapaxes(o.field("phi"), 0, 2)), f"phi {step}"
        o.macros()
        for name in ("rho", "ux", "uy", "uz"):
            got = postfmt.read_macr(postfmt.macr_file(sim, "cfg0", name, step), sim)
            assert np.array_equal(got, np.swapaxes(o.field(name), 0, 2)), f"{name} {step}"


def test_state_errors_are_statuses(gpu_lib):
    mcb = gpu_lib
    lat = mcb.Lattice(mcb.make_params(16, 16, 16))
    with pytest.raises(mcb.McbError) as e:
        lat.step(1)                       # no initial state yet
    assert e.value.status == mcb.ERR_INVALID_ARG
    lat.init_bubbles([(8, 8, 8, 4)])
    with pytest.raises(mcb.McbError):
        lat.write("phi", np.zeros(10, np.float32))
    with pytest.raises(mcb.McbError):
        lat.init_phi(np.zeros(7, np.float32))
    with pytest.raises(mcb.McbError):
        lat.set_tiling(48, 3, 4)
    lat.close()


@pytest.mark.parametrize("path", ["fused", "staged"])
def test_nonfinite_populations_are_detected(gpu_lib, path):
    mcb = gpu_lib
    p = mcb.make_params(16, 16, 16, tau_f=0.6, path=mcb.PATH_FUSED if path == "fused" else mcb.PATH_STAGED)
    lat = mcb.Lattice(p)
    lat.init_bubbles([(8, 8, 8, 4)])
    lat.step(2)
    lat.compute_macros()
    s, ke, ok = lat.reduce()
    assert ok and np.isfinite(s)
    g = lat.read("g")
    g[3, 4, 5, 6] = np.nan               # fault injection through the ABI
    lat.write("g", g)
    lat.step(2)
    lat.compute_macros()
    s, ke, ok = lat.reduce()
    assert not ok and not np.isfinite(s)
    lat.close()


@pytest.mark.parametrize("tau_g", ["0.8", "1"])
def test_driver_resumes_from_a_state_image(gpu_lib, tmp_path, tau_g):
    """--save-state after 20 iterations, then --load-state in a second run:
    the phi files of the resumed run are the bytes of one run straight
    through."""
    mcb = gpu_lib
    common = ["--preset", "cfg0", "--tau-g", tau_g, "--save-every", "20"]
    straight, resumed = str(tmp_path / "a") + "/", str(tmp_path / "b") + "/"
    image = str(tmp_path / "state.img")
    runs = [
        [mcb.DRIVER_PATH, "D3Q19", "s", *common, "--steps", "40", "--root", straight],
        [mcb.DRIVER_PATH, "D3Q19", "s", *common, "--steps", "19", "--save-state", image, "--root", resumed],
        [mcb.DRIVER_PATH, "D3Q19", "s", *common, "--steps", "40", "--load-state", image, "--root", resumed],
    ]
    for cmd in runs:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
    assert "Resumed from" in r.stdout
    for step in (20, 40):
        a = open(postfmt.macr_file(straight + "bin/D3Q19/s/", "s", "phi", step), "rb").read()
        b = open(postfmt.macr_file(resumed + "bin/D3Q19/s/", "s", "phi", step), "rb").read()
        assert a == b, f"step {step}"
    # An image of another case is refused with a status, not a crash.
    r = subprocess.run([mcb.DRIVER_PATH, "D3Q19", "t", "--preset", "cfg1", "--steps", "2", "--load-state", image,
                        "--root", resumed], capture_output=True, text=True, timeout=300)
    assert r.returncode == 2 and "load state" in r.stderr
