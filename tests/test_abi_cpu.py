"""libmcb's C ABI and host logic, without a GPU.

Covers: every symbol declared in include/*.h is exported; parameter checks
reject bad input with a status (never abort); the context API reports a
missing device as a status; presets are deterministic; the host initial
bubble equals the oracle's bit for bit; the info and field files follow the
reference's format and parse under its post-processing rules.
"""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import mcb
import oracle as O
import postfmt
from conftest import ROOT, f32_bubbles, oracle_for

HEADERS = [os.path.join(ROOT, "include", h) for h in ("mcb.h", "mcb_ops.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(mcb_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(mcb.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_prototypes_cover_the_headers():
    assert set(declared_functions()) == set(mcb.PROTOTYPES)


def test_abi_version():
    assert mcb.lib().mcb_abi_version() == 1


@pytest.mark.parametrize("field,value", [
    ("nx", 2), ("ny", 0), ("nz", -4), ("tau_f", 0.5), ("tau_f", float("nan")), ("tau_g", 0.4),
    ("sigma", -1.0), ("sigma", float("inf")), ("bc_mode", 7), ("weight_set", 3), ("storage", 9),
    ("path", 5), ("device", -1), ("normals", 3), ("normals", -1), ("f_form", 2), ("f_form", -1),
])
def test_validate_rejects(field, value):
    p = mcb.make_params(16, 16, 16)
    setattr(p, field, value)
    assert mcb.lib().mcb_validate_params(ctypes.byref(p)) == mcb.ERR_INVALID_ARG
    h = ctypes.c_void_p()
    assert mcb.lib().mcb_create(ctypes.byref(p), None, ctypes.byref(h)) == mcb.ERR_INVALID_ARG
    assert not h.value


def test_validate_rejects_huge_lattice():
    p = mcb.make_params(1024, 1024, 1024)
    assert mcb.lib().mcb_validate_params(ctypes.byref(p)) == mcb.ERR_INVALID_ARG


def test_reference_params_are_var_cu():
    p = mcb.reference_params()
    assert (p.nx, p.ny, p.nz) == (64, 64, 256)
    assert np.float32(p.tau_f) == np.float32(0.505) and np.float32(p.sigma) == np.float32(0.1)
    assert p.bc_mode == mcb.BC_REFERENCE and p.weight_set == mcb.WEIGHTS_LITERAL and p.tau_g == 1.0
    assert mcb.lib().mcb_validate_params(ctypes.byref(p)) == mcb.OK


def test_null_arguments_are_rejected():
    L = mcb.lib()
    assert L.mcb_validate_params(None) == mcb.ERR_INVALID_ARG
    assert L.mcb_create(None, None, None) == mcb.ERR_INVALID_ARG
    assert L.mcb_step(None, 1) == mcb.ERR_INVALID_ARG
    assert L.mcb_destroy(None) == mcb.ERR_INVALID_ARG
    assert L.mcb_case_preset(b"nope", 0, None) == mcb.ERR_INVALID_ARG
    assert L.mcb_footprint_bytes(None) == 0


@pytest.mark.skipif(mcb.device_count() > 0, reason="a GPU is visible")
def test_create_without_device_reports_status():
    h = ctypes.c_void_p()
    p = mcb.make_params(16, 16, 16)
    assert mcb.lib().mcb_create(ctypes.byref(p), None, ctypes.byref(h)) == mcb.ERR_NO_DEVICE
    assert not h.value
    with pytest.raises(mcb.McbError):
        mcb.Lattice(p)


def test_unsupported_combinations():
    """The fused kernels take ghost-layer boundaries only in FP32 with g in
    moment form (tau_g = 1) and without the dead movingWallBC; these
    combinations are refused before any device is touched."""
    L = mcb.lib()
    h = ctypes.c_void_p()
    for kw in (dict(tau_g=0.8), dict(moving_wall=True), dict(storage=mcb.STORE_FP16)):
        p = mcb.make_params(16, 16, 16, bc_mode=mcb.BC_REFERENCE, path=mcb.PATH_FUSED, **kw)
        assert L.mcb_create(ctypes.byref(p), None, ctypes.byref(h)) == mcb.ERR_UNSUPPORTED, kw
    # f in moment form: fused FP32, g in moment form, periodic only.
    for kw in (dict(tau_g=0.8), dict(storage=mcb.STORE_FP16), dict(path=mcb.PATH_STAGED),
               dict(bc_mode=mcb.BC_REFERENCE)):
        p = mcb.make_params(16, 16, 16, f_form=mcb.F_MOMENTS, **kw)
        assert L.mcb_create(ctypes.byref(p), None, ctypes.byref(h)) == mcb.ERR_UNSUPPORTED, kw


def test_largest_fused_lattice():
    """The fused kernels address a buffer's 19 population planes through one
    32-bit buffer resource: FP32 lattices beyond 4 GiB of populations
    (400^3: 64M cells x 19 x 4 B) are refused on the fused path before any
    device is touched, and an unset path falls back to the staged kernels."""
    L = mcb.lib()
    h = ctypes.c_void_p()
    p = mcb.make_params(400, 400, 400, path=mcb.PATH_FUSED)
    assert L.mcb_create(ctypes.byref(p), None, ctypes.byref(h)) == mcb.ERR_UNSUPPORTED
    assert not h.value
    if mcb.device_count() == 0:
        # Accepted shapes get as far as the device lookup: 352^3 (43.6M cells,
        # 3.3 GB of f) on the fused path, 400^3 with the path left to the
        # library (staged), and 400^3 with 16-bit populations (half the bytes).
        for q in (mcb.make_params(352, 352, 352, path=mcb.PATH_FUSED), mcb.make_params(400, 400, 400),
                  mcb.make_params(400, 400, 400, storage=mcb.STORE_FP16, path=mcb.PATH_FUSED)):
            assert L.mcb_create(ctypes.byref(q), None, ctypes.byref(h)) == mcb.ERR_NO_DEVICE


def test_footprint():
    n = 64 ** 3
    p = mcb.make_params(64, 64, 64, tau_g=0.8)
    fused = mcb.lib().mcb_footprint_bytes(ctypes.byref(p))
    assert fused >= 4 * 19 * n * 4                      # f and g, two buffers each
    p.tau_g = 1.0
    moments = mcb.lib().mcb_footprint_bytes(ctypes.byref(p))
    assert 2 * 19 * n * 4 + 15 * n * 4 <= moments < fused   # g as phi (two buffers), u (and n)
    # normals: AUTO stores n at 64^3 (short phase chunks) and recomputes it at
    # 256^3; RECOMPUTED and STORED force either.
    p.normals = mcb.NORMALS_RECOMPUTED
    assert mcb.lib().mcb_footprint_bytes(ctypes.byref(p)) == moments - 3 * (n + 256) * 4
    p.normals = mcb.NORMALS_STORED
    assert mcb.lib().mcb_footprint_bytes(ctypes.byref(p)) == moments
    # f in moment form: 13 planes A/B instead of 19, u among them.
    p.f_form = mcb.F_MOMENTS
    assert mcb.lib().mcb_footprint_bytes(ctypes.byref(p)) == moments - (2 * 6 + 3) * (n + 256) * 4
    p.f_form = mcb.F_POPULATIONS
    n2 = 256 ** 3
    p256 = mcb.make_params(256, 256, 256)
    auto = mcb.lib().mcb_footprint_bytes(ctypes.byref(p256))
    p256.normals = mcb.NORMALS_STORED
    assert auto == mcb.lib().mcb_footprint_bytes(ctypes.byref(p256)) - 3 * (n2 + 256) * 4
    p.path = mcb.PATH_STAGED
    assert mcb.lib().mcb_footprint_bytes(ctypes.byref(p)) >= 3 * 19 * 64 ** 3 * 4


def test_presets_are_deterministic_and_separated():
    for name in ("reference", "cfg0", "cfg1", "cfg2"):
        a, b = mcb.case_preset(name, 5), mcb.case_preset(name, 5)
        assert bytes(a) == bytes(b)
    assert bytes(mcb.case_preset("cfg1", 1)) != bytes(mcb.case_preset("cfg1", 2))
    for seed in range(20):
        c = mcb.case_preset("cfg2", seed)
        (x0, y0, z0, r0), (x1, y1, z1, r1) = mcb.case_bubbles(c)
        assert 16 <= r0 <= 32 and 16 <= r1 <= 32
        d = np.array([x0 - x1, y0 - y1, z0 - z1])
        d = (d + 128) % 256 - 128
        assert np.linalg.norm(d) >= r0 + r1 + 12 - 1e-3
    c = mcb.case_preset("cfg1", 3)
    assert 12 <= c.bubbles[0].r <= 24 and c.nsteps == 10000
    ref = mcb.case_preset("reference")
    assert mcb.case_bubbles(ref) == [(32.0, 32.0, 128.0, 20.0)] and ref.nsteps == 25000 and ref.save_every == 100


@pytest.mark.parametrize("name", ["reference", "cfg0", "cfg1"])
def test_host_initial_phi_equals_oracle(name):
    c = mcb.case_preset(name, 11)
    phi = mcb.host_init_phi(c.params, mcb.case_bubbles(c), c.width)
    o = oracle_for(c.params)
    o.init(f32_bubbles(mcb.case_bubbles(c)), c.width)
    assert np.array_equal(phi.view(np.uint32), o.field("phi").view(np.uint32))


def test_host_initial_phi_two_bubbles_min_image():
    p = mcb.make_params(40, 36, 32)
    bubbles = [(2.5, 3.0, 30.0, 6.0), (25.0, 20.0, 12.0, 8.0)]
    phi = mcb.host_init_phi(p, bubbles, 3.0)
    o = oracle_for(p)
    o.init(f32_bubbles(bubbles), 3.0)
    assert np.array_equal(phi, o.field("phi"))
    assert phi[30, 3, 2] > 0.99 and phi[0, 0, 0] > 0.5    # wraps around the corner


def test_info_file_matches_reference_bytes(tmp_path):
    path = str(tmp_path / "000_info.txt")
    st = mcb.lib().mcb_write_info_file(path.encode(), b"000", b"D3Q19", 64, 64, 256, 100, 25000, 0.505, -1.0)
    assert st == mcb.OK
    want = open(os.path.join(ROOT, "tests", "golden", "reference_info_000.txt"), "rb").read()
    assert open(path, "rb").read() == want


def test_info_and_field_files_parse_with_post_rules(tmp_path):
    d = str(tmp_path) + "/"
    nx, ny, nz = 5, 4, 3
    st = mcb.lib().mcb_write_info_file((d + "007_info.txt").encode(), b"007", b"D3Q19", nx, ny, nz, 100, 300,
                                       0.55, 1234.5)
    assert st == mcb.OK
    info = postfmt.sim_info(d)
    assert info["ID"] == "007" and info["Prc"] == "float" and (info["NX"], info["NY"], info["NZ"]) == (5, 4, 3)
    assert info["NZ_TOTAL"] == 3 and info["Nsteps"] == 300 and abs(info["Tau"] - 0.55) < 1e-6
    assert "1.234500e+03" in open(d + "007_info.txt").read()
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    field = (x + 10 * y + 100 * z).astype(np.float32)
    for step in (0, 100, 300):
        a = np.ascontiguousarray(field + step)
        st = mcb.lib().mcb_write_field_file(d.encode(), b"007", b"phi", step, a.ctypes.data_as(mcb._fp), a.size)
        assert st == mcb.OK
    b = np.ascontiguousarray(field)
    mcb.lib().mcb_write_field_file(d.encode(), b"007", b"rho", 100, b.ctypes.data_as(mcb._fp), b.size)
    assert postfmt.macr_steps(d) == [0, 100, 300]           # rho files are not picked up
    arr = postfmt.read_macr(postfmt.macr_file(d, "007", "phi", 100), d)
    assert arr.shape == (nx, ny, nz)
    assert arr[3, 2, 1] == 3 + 20 + 100 + 100
    assert os.path.getsize(postfmt.macr_file(d, "007", "phi", 0)) == nx * ny * nz * 4


def test_write_errors_are_statuses(tmp_path):
    bad = str(tmp_path / "no" / "such" / "dir" / "x_info.txt")
    assert mcb.lib().mcb_write_info_file(bad.encode(), b"1", b"D3Q19", 3, 3, 3, 1, 1, 0.6, -1) == mcb.ERR_IO


def test_bench_refuses_multiple_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_driver_usage_without_arguments():
    drv = mcb.DRIVER_PATH
    if not os.path.exists(drv):
        pytest.skip("driver not built")
    r = subprocess.run([drv], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([drv, "D3Q27", "001"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "D3Q19" in r.stderr


@pytest.mark.parametrize("blob", [b"", b"MCBSTATE", b"x" * 256])
def test_state_info_rejects_what_is_not_a_state_image(blob):
    """mcb_state_info reads an image's header without a device; anything that
    is not a whole image of this version is refused with a status."""
    with pytest.raises(mcb.McbError) as e:
        mcb.state_info(blob)
    assert e.value.status == mcb.ERR_INVALID_ARG


def test_state_calls_without_a_device_report_status():
    n = ctypes.c_size_t()
    assert mcb.lib().mcb_state_size(None, ctypes.byref(n)) == mcb.ERR_INVALID_ARG
    assert mcb.lib().mcb_load_state(None, None, 0) == mcb.ERR_INVALID_ARG
    assert mcb.lib().mcb_batch_add_state(None, None, 0, None) == mcb.ERR_INVALID_ARG
