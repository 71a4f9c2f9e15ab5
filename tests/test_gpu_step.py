"""GPU parity of the full time step against the oracle (FP32 transcription).

Both the fused path (k_phase + k_collide) and the staged path (the
reference's kernel sequence, mcb_ops.h) must reproduce the transcription
bit for bit: the kernels round every operation as the transcription does
(-ffp-contract=off, IEEE division/sqrt, denormals kept). The stated
tolerance of the path (1e-5 relative L2 on phi, rho, u after 1000 steps)
is asserted too, so a lost bit-exactness shows up as a named failure with
its size, not as a silent pass.
"""
import numpy as np
import pytest

from conftest import f32_bubbles, oracle_for, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5
OBS = ("phi", "rho", "ux", "uy", "uz")


def run_pair(mcb, params, bubbles, steps, width=3.0):
    lat = mcb.Lattice(params)
    lat.init_bubbles(bubbles, width)
    o = oracle_for(params)
    o.init(f32_bubbles(bubbles), width)
    lat.step(steps)
    o.step(steps)
    got = lat.observables()
    o.macros()
    want = {k: o.field(k).copy() for k in OBS}
    return lat, o, got, want


def assert_parity(got, want, label, exact=True):
    for k in OBS:
        err = rel_l2(got[k], want[k])
        assert err <= TOL, f"{label}: {k} rel L2 {err:.3e} > {TOL}"
        if exact:
            nbad = int(np.count_nonzero(got[k] != want[k]))
            assert nbad == 0, f"{label}: {k} differs from the oracle in {nbad} cells (rel L2 {err:.3e})"


# Fused FP32 at tau_g = 1 keeps g in moment form with the interface normal
# either stored by the collide pass or recomputed by the phase pass.
NORMALS = {"n-stored": 1, "n-recomputed": 2}
# Tile shapes of the collide pass with f in moment form.
FM_TILES = {(64, 8), (32, 8), (32, 16), (16, 16), (16, 8)}


@pytest.mark.parametrize("path", ["fused", "fused-n-recomputed", "fused-f-moments", "staged"])
@pytest.mark.parametrize("shape,tau_g", [((32, 32, 32), 1.0), ((48, 40, 36), 0.8), ((20, 24, 17), 1.0)])
def test_periodic_short_bitexact(gpu_lib, path, shape, tau_g):
    mcb = gpu_lib
    if path == "fused-f-moments" and tau_g != 1.0:
        pytest.skip("f in moment form needs g in moment form (tau_g = 1)")
    nx, ny, nz = shape
    p = mcb.make_params(nx, ny, nz, tau_f=0.55, tau_g=tau_g, sigma=0.1,
                        path=mcb.PATH_STAGED if path == "staged" else mcb.PATH_FUSED,
                        normals=mcb.NORMALS_RECOMPUTED if path == "fused-n-recomputed" else mcb.NORMALS_STORED,
                        f_form=mcb.F_MOMENTS if path == "fused-f-moments" else mcb.F_POPULATIONS)
    bubbles = [(nx * 0.5 + 0.3, ny * 0.5 - 1.2, nz * 0.5 + 0.7, min(shape) * 0.25)]
    lat, o, got, want = run_pair(mcb, p, bubbles, 20)
    assert_parity(got, want, f"{path} {shape} tau_g={tau_g}")
    lat.close()


@pytest.mark.parametrize("path", ["fused", "staged"])
def test_cfg0_1000_steps(gpu_lib, path):
    """64^3, R=10 centred, tau_f 0.55, tau_g 0.8, sigma 0.1, 1000 steps."""
    mcb = gpu_lib
    c = mcb.case_preset("cfg0")
    c.params.path = mcb.PATH_FUSED if path == "fused" else mcb.PATH_STAGED
    lat, o, got, want = run_pair(mcb, c.params, mcb.case_bubbles(c), 1000)
    assert_parity(got, want, f"cfg0 {path}")
    # Phase conservation: sum(phi) after 1000 steps vs the initial bubble.
    # The path carries the transcription's bits, so this is the FP32
    # transcription's own rounding drift with the tau_g = 0.8 extension on
    # this small lattice: 1.0e-7 measured (the oracle's FP64 run: 1e-14; the
    # 128^3 case of the same physics drifts 2.4e-8 over 1000 steps).
    phi0 = mcb.host_init_phi(c.params, mcb.case_bubbles(c), c.width).astype(np.float64).sum()
    s, ke, finite = lat.reduce()
    assert finite
    assert abs(s - o.sum_phi()) <= 1e-9 * phi0
    assert abs(s - phi0) / phi0 <= 2e-7, f"sum(phi) drift {abs(s - phi0) / phi0:.3e}"
    assert abs(ke - o.kinetic()) <= 1e-9 * max(1.0, abs(ke))
    lat.close()


@pytest.mark.parametrize("f_form", ["f-populations", "f-moments"])
@pytest.mark.parametrize("normals", list(NORMALS))
def test_moment_form_1000_steps(gpu_lib, normals, f_form):
    """tau_g = 1 keeps g as (phi, u, n) and recomputes g* in the phase pass,
    with n stored or itself recomputed from phi, and f as populations or as
    the 13 moments collisionFluid reads: 1000 steps stay bit-identical to the
    transcription."""
    mcb = gpu_lib
    p = mcb.make_params(48, 48, 48, tau_f=0.55, tau_g=1.0, sigma=0.1, path=mcb.PATH_FUSED, normals=NORMALS[normals],
                        f_form=mcb.F_MOMENTS if f_form == "f-moments" else mcb.F_POPULATIONS)
    bubbles = [(24.3, 23.6, 24.1, 10.0)]
    lat, o, got, want = run_pair(mcb, p, bubbles, 1000)
    assert_parity(got, want, f"moment form ({normals}, {f_form}) 48^3 x 1000")
    lat.close()


def test_f_moments_read_populations_and_refuse_writes(gpu_lib):
    """With f in moment form, f and g read in the reference layout equal the
    transcription's populations before and after steps (f rebuilt from the
    moments with collisionFluid's operations); populations cannot be
    written in this form."""
    mcb = gpu_lib
    p = mcb.make_params(30, 22, 18, tau_f=0.6, tau_g=1.0, sigma=0.08, path=mcb.PATH_FUSED, f_form=mcb.F_MOMENTS)
    bubbles = [(14.5, 10.0, 9.5, 5.5)]
    o = oracle_for(p)
    o.init(f32_bubbles(bubbles))
    lat = mcb.Lattice(p)
    lat.init_bubbles(bubbles)
    for steps in (0, 1, 6):
        lat.step(steps)
        o.step(steps)
        for name in ("f", "g"):
            got = lat.read(name).reshape(19, -1)
            assert np.array_equal(got, o.field(name).reshape(19, -1)), f"{name} after {steps} more steps"
    with pytest.raises(mcb.McbError) as e:
        lat.write("f", lat.read("f"))
    assert e.value.status == mcb.ERR_UNSUPPORTED
    with pytest.raises(mcb.McbError):
        lat.write("g", lat.read("g"))
    got = lat.observables()
    o.macros()
    for k in OBS:
        assert np.array_equal(got[k], o.field(k)), k
    lat.close()


# FP32 contexts keep populations in planes at the first shape (nx*ny = 480)
# and in 64-cell blocks at the second (nx*ny a multiple of 64).
LAYOUT_SHAPES = [(24, 20, 16), (32, 24, 16)]


@pytest.mark.parametrize("shape", LAYOUT_SHAPES)
@pytest.mark.parametrize("normals", list(NORMALS))
def test_moment_form_populations(gpu_lib, normals, shape):
    """g read from the moment form is the reference layout, before and after
    steps; writing g switches the context to stored populations, bit for bit."""
    mcb = gpu_lib
    p = mcb.make_params(*shape, tau_f=0.6, tau_g=1.0, sigma=0.08, path=mcb.PATH_FUSED, normals=NORMALS[normals])
    bubbles = [(12.0, 10.5, 8.0, 5.0)]
    o = oracle_for(p)
    o.init(f32_bubbles(bubbles))
    lat = mcb.Lattice(p)
    lat.init_bubbles(bubbles)
    assert np.array_equal(lat.read("g").reshape(19, -1), o.field("g").reshape(19, -1)), "initial g"
    assert np.array_equal(lat.read("phi"), o.field("phi"))
    lat.step(5)
    o.step(5)
    g = lat.read("g")
    assert np.array_equal(g.reshape(19, -1), o.field("g").reshape(19, -1)), "g after 5 steps"
    assert np.array_equal(lat.read("f").reshape(19, -1), o.field("f").reshape(19, -1))
    lat.write("g", g)             # now stored as populations
    lat.step(3)
    o.step(3)
    got = lat.observables()
    o.macros()
    for k in OBS:
        assert np.array_equal(got[k], o.field(k)), k
    lat.close()


@pytest.mark.parametrize("tau_g", [0.8, 1.0])
def test_fused_tilings_agree(gpu_lib, tau_g):
    """Every tile shape and z-chunk gives the same bits; at tau_g = 1 with the
    normals stored and recomputed alike."""
    mcb = gpu_lib
    bubbles = [(20.5, 17.0, 14.2, 7.5)]
    ref = None
    forms = [(n, mcb.F_POPULATIONS) for n in NORMALS.values()] + [(1, mcb.F_MOMENTS), (2, mcb.F_MOMENTS)]
    for normals, f_form in (forms if tau_g == 1.0 else [(0, mcb.F_POPULATIONS)]):
        p = mcb.make_params(40, 36, 28, tau_f=0.6, tau_g=tau_g, sigma=0.05, path=mcb.PATH_FUSED, normals=normals,
                            f_form=f_form)
        for tx, ty in mcb.tilings():
            if f_form == mcb.F_MOMENTS and (tx, ty) not in FM_TILES:
                continue
            for zc in (1, 5, 28):
                lat = mcb.Lattice(p)
                lat.init_bubbles(bubbles)
                lat.set_tiling(tx, ty, zc)
                lat.step(7)
                obs = lat.observables()
                lat.close()
                if ref is None:
                    ref = obs
                for k in OBS:
                    assert np.array_equal(obs[k], ref[k]), \
                        f"normals {normals}, f_form {f_form}, tiling {tx}x{ty} zc={zc}: {k} differs"


@pytest.mark.parametrize("shape", LAYOUT_SHAPES)
def test_populations_roundtrip_and_layout(gpu_lib, shape):
    """F/G readout is the reference layout (post-stream, pre-collision) for both paths."""
    mcb = gpu_lib
    bubbles = [(12.0, 10.0, 8.0, 5.0)]
    outs = {}
    for path in ("fused", "staged"):
        p = mcb.make_params(*shape, tau_f=0.55, tau_g=0.8, path=mcb.PATH_FUSED if path == "fused" else mcb.PATH_STAGED)
        lat = mcb.Lattice(p)
        lat.init_bubbles(bubbles)
        lat.step(5)
        outs[path] = (lat.read("f"), lat.read("g"))
        # write back what was read: the next step must be unchanged
        lat.write("f", outs[path][0])
        lat.write("g", outs[path][1])
        lat.step(1)
        outs[path + "+1"] = lat.observables()
        lat.close()
    assert np.array_equal(outs["fused"][0], outs["staged"][0])
    assert np.array_equal(outs["fused"][1], outs["staged"][1])
    o = oracle_for(mcb.make_params(*shape, tau_f=0.55, tau_g=0.8))
    o.init(f32_bubbles(bubbles))
    o.step(5)
    assert np.array_equal(outs["fused"][0].reshape(19, -1), o.field("f").reshape(19, -1))
    assert np.array_equal(outs["fused"][1].reshape(19, -1), o.field("g").reshape(19, -1))
    o.step(1)
    o.macros()
    for k in OBS:
        assert np.array_equal(outs["fused+1"][k], o.field(k))
        assert np.array_equal(outs["staged+1"][k], o.field(k))


# The second has partial tiles in x and y; the third keeps FP32 populations
# in 64-cell blocks.
REF_SHAPES = [(24, 20, 40), (37, 29, 33), (32, 24, 20)]


@pytest.mark.parametrize("path", ["fused", "fused-n-recomputed", "staged"])
@pytest.mark.parametrize("shape", REF_SHAPES)
def test_reference_mode(gpu_lib, path, shape):
    """Ghost-layer boundaries of the reference against the oracle: the fused
    path (faces as fixed pull-form sources) and the staged kernels. phi is
    compared after every 10 iterations as the reference's d_phi is saved
    (boundaryConditions applied), then every macroscopic field and the f
    populations in the reference layout."""
    mcb = gpu_lib
    nx, ny, nz = shape
    p = mcb.make_params(nx, ny, nz, bc_mode=mcb.BC_REFERENCE, weights=mcb.WEIGHTS_LITERAL, tau_f=0.505,
                        tau_g=1.0, sigma=0.1, path=mcb.PATH_STAGED if path == "staged" else mcb.PATH_FUSED,
                        normals=mcb.NORMALS_RECOMPUTED if path == "fused-n-recomputed" else mcb.NORMALS_STORED)
    bubbles = [(nx * 0.5, ny * 0.5, nz * 0.5, min(shape) * 0.3)]
    lat = mcb.Lattice(p)
    lat.init_bubbles(bubbles)
    o = oracle_for(p)
    o.init(f32_bubbles(bubbles))
    assert np.array_equal(lat.read("phi"), o.field("phi"))
    for _ in range(3):
        lat.step(10)
        o.step(10)
        assert np.array_equal(lat.read("phi"), o.field("phi"))
    f = lat.read("f")
    assert np.array_equal(f.reshape(19, -1), o.field("f").reshape(19, -1))
    lat.compute_macros()
    o.macros()
    for k, ok in (("phi", "phi"), ("rho", "rho"), ("ux", "ux"), ("uy", "uy"), ("uz", "uz"), ("fx", "ffx"),
                  ("fy", "ffy"), ("fz", "ffz"), ("normx", "normx"), ("normy", "normy"), ("normz", "normz")):
        got = lat.read(k)
        nbad = int(np.count_nonzero(got != o.field(ok)))
        assert nbad == 0, f"{path}: {k} differs from the oracle in {nbad} cells"
    lat.close()


@pytest.mark.parametrize("shape", [(26, 22, 30), (32, 22, 30)])
@pytest.mark.parametrize("normals", list(NORMALS))
def test_reference_mode_fused_write_f_and_resume(gpu_lib, normals, shape):
    """Fused ghost-layer path: F written through the ABI feeds the next
    iteration as the reference's d_f does (face populations included once),
    and a state image resumes the run bit for bit, also in a context that
    keeps the normals the other way; compared with the oracle given the
    same F."""
    mcb = gpu_lib
    nx, ny, nz = shape
    p = mcb.make_params(nx, ny, nz, bc_mode=mcb.BC_REFERENCE, weights=mcb.WEIGHTS_LITERAL, tau_f=0.55,
                        tau_g=1.0, sigma=0.1, path=mcb.PATH_FUSED, normals=NORMALS[normals])
    bubbles = [(13.0, 11.0, 15.0, 6.5)]
    lat = mcb.Lattice(p)
    lat.init_bubbles(bubbles)
    o = oracle_for(p)
    o.init(f32_bubbles(bubbles))
    lat.step(6)
    o.step(6)
    rng = np.random.default_rng(5)
    f = lat.read("f")
    f = (f * (1.0 + 0.01 * rng.standard_normal(f.shape))).astype(np.float32)
    lat.write("f", f)
    o.field("f")[...] = f.reshape(o.field("f").shape)
    lat.step(4)
    o.step(4)
    image = lat.save_state()
    lat.step(5)
    o.step(5)
    got = lat.observables()
    o.macros()
    for k in OBS:
        assert np.array_equal(got[k], o.field(k)), k
    for other in NORMALS.values():
        p.normals = other
        with mcb.Lattice(p) as lat2:
            lat2.load_state(image)
            lat2.step(5)
            got2 = lat2.observables()
        for k in OBS:
            assert np.array_equal(got2[k], got[k]), f"resumed (normals {other}) {k}"
    with pytest.raises(mcb.McbError):
        lat.read("g")           # the reference layout of g is kept by the staged path only
    lat.close()


@pytest.mark.parametrize("shape", [(30, 26, 22), (32, 26, 22)])
@pytest.mark.parametrize("kind", ["moments", "moments-n-recomputed", "f-moments", "populations", "reference"])
def test_graph_replayed_steps_match_eager_steps(gpu_lib, kind, shape):
    """mcb_step replays pairs of fused steps from a captured hipGraph. Each
    replay, with eager work (observables, a field readout) and host
    synchronisation between replays, must give the bits of the same steps
    launched one at a time."""
    mcb = gpu_lib
    kw = dict(tau_f=0.6, sigma=0.08, path=mcb.PATH_FUSED)
    if kind == "populations":
        kw["tau_g"] = 0.8
    if kind == "reference":
        kw.update(bc_mode=mcb.BC_REFERENCE, weights=mcb.WEIGHTS_LITERAL)
    if kind == "moments-n-recomputed":
        kw["normals"] = mcb.NORMALS_RECOMPUTED
    if kind == "f-moments":
        kw["f_form"] = mcb.F_MOMENTS
    p = mcb.make_params(*shape, **kw)
    bubbles = [(15.2, 13.1, 11.3, 6.0)]
    g, e = mcb.Lattice(p), mcb.Lattice(p)
    g.init_bubbles(bubbles)
    e.init_bubbles(bubbles)
    g.step(1)                      # the first step (initial state) is eager on both
    e.step(1)
    for replay in range(6):
        g.step(2)                  # one graph replay
        e.step(1)                  # eager launches
        e.step(1)
        a, b = g.read("phi"), e.read("phi")
        assert np.array_equal(a, b), f"{kind}, replay {replay}: phi differs in {int(np.count_nonzero(a != b))} cells"
        if replay % 2 == 0:
            ga, eb = g.observables(), e.observables()
            for k in OBS:
                assert np.array_equal(ga[k], eb[k]), f"{kind}, replay {replay}: {k}"
    g.close()
    e.close()


def test_steps_inside_a_caller_graph(gpu_lib):
    """A caller capturing the context's stream into its own hipGraph records
    mcb_step's eager launches (the library captures nothing of its own while
    the stream is capturing); replaying the caller's graph advances the
    lattice as the same steps launched directly."""
    import ctypes
    mcb = gpu_lib
    hip = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p
    stream, graph, exe = vp(), vp(), vp()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    p = mcb.make_params(30, 26, 22, tau_f=0.6, sigma=0.08, path=mcb.PATH_FUSED)
    bubbles = [(15.2, 13.1, 11.3, 6.0)]
    g = mcb.Lattice(p, stream=stream)
    e = mcb.Lattice(p)
    g.init_bubbles(bubbles)
    e.init_bubbles(bubbles)
    g.step(1)
    e.step(1)
    assert hip.hipStreamBeginCapture(stream, 1) == 0          # thread-local capture mode
    g.step(2)                                                 # recorded, not run
    assert hip.hipStreamEndCapture(stream, ctypes.byref(graph)) == 0
    assert hip.hipGraphInstantiate(ctypes.byref(exe), graph, None, None, ctypes.c_size_t(0)) == 0
    for replay in range(3):
        assert hip.hipGraphLaunch(exe, stream) == 0
        assert hip.hipStreamSynchronize(stream) == 0
        e.step(2)
        a, b = g.read("phi"), e.read("phi")
        assert np.array_equal(a, b), f"replay {replay}: phi differs in {int(np.count_nonzero(a != b))} cells"
    hip.hipGraphExecDestroy(exe)
    hip.hipGraphDestroy(graph)
    g.close()
    e.close()
    hip.hipStreamDestroy(stream)
