"""State images (mcb_save_state / mcb_load_state, mcb_batch_*_state): a run
saved after k steps and loaded into a fresh context or batch continues bit
for bit like the run that never stopped, in every storage form and g
representation; images of other lattices or parameters are refused.

The reference has no restart (it writes phi snapshots only, src/main.cu:184-189);
the property pinned here is the definition of a checkpoint: save, load,
continue == continue.
"""
import numpy as np
import pytest

from conftest import oracle_for, f32_bubbles

pytestmark = pytest.mark.gpu

BUB = [(11.3, 9.6, 8.2, 5.0), (3.1, 14.2, 12.5, 3.5)]
SHAPE = (24, 20, 18)            # nx, ny, nz
# FP32 contexts keep populations in planes at SHAPE (nx*ny = 480) and in
# 64-cell blocks where nx*ny is a multiple of 64 (the second shape); images
# hold planes either way.
SHAPES = [SHAPE, (32, 24, 18)]


def params(mcb, storage, tau_g, path=None, bc=None, shape=SHAPE):
    kw = dict(tau_f=0.6, tau_g=tau_g, sigma=0.1, storage=storage)
    if path is not None:
        kw["path"] = path
    if bc is not None:
        kw["bc_mode"] = bc
    return mcb.make_params(*shape, **kw)


def snapshot(lat):
    obs = lat.observables()
    obs["f"] = lat.read("f")
    obs["g"] = lat.read("g")
    return obs


def assert_same(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs in {np.count_nonzero(a[k] != b[k])} values"


FORMS = [
    ("fp32", 0.8), ("fp32", 1.0), ("fp16", 0.8), ("fp16", 1.0), ("bf16", 1.0), ("bf16", 0.8),
]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("k", [0, 7])
@pytest.mark.parametrize("form", FORMS, ids=lambda f: f"{f[0]}-taug{f[1]}")
def test_save_load_continues_bit_for_bit(gpu_lib, form, k, shape):
    mcb = gpu_lib
    st = {"fp32": mcb.STORE_FP32, "fp16": mcb.STORE_FP16, "bf16": mcb.STORE_BF16}[form[0]]
    p = params(mcb, st, form[1], shape=shape)
    a = mcb.Lattice(p)
    a.init_bubbles(BUB, 3.0)
    a.step(k)
    image = a.save_state()
    phi_saved = a.read("phi")
    pi, steps = mcb.state_info(image)
    assert steps == k and (pi.nx, pi.ny, pi.nz) == shape
    a.step(9)
    ref = snapshot(a)
    a.close()
    # A fresh context (no initial state of its own) resumes from the image;
    # phi read before any step is the phi the saved context reported.
    b = mcb.Lattice(p)
    b.load_state(image)
    assert np.array_equal(b.read("phi"), phi_saved), f"{form} k={k}: phi read after loading"
    b.step(9)
    assert_same(ref, snapshot(b), f"{form} k={k}")
    # Loading into a context that has stepped an odd number of times (the
    # other A/B buffer) gives the same run.
    c = mcb.Lattice(p)
    c.init_bubbles([(5.0, 5.0, 5.0, 2.0)], 3.0)
    c.step(3)
    c.load_state(image)
    c.step(9)
    assert_same(ref, snapshot(c), f"{form} k={k}, loaded at odd parity")
    b.close()
    c.close()


@pytest.mark.parametrize("shape", SHAPES)
def test_fp32_state_matches_the_oracle_after_restart(gpu_lib, shape):
    """The resumed FP32 run equals the FP32 transcription run straight through."""
    mcb = gpu_lib
    p = params(mcb, mcb.STORE_FP32, 1.0, shape=shape)
    a = mcb.Lattice(p)
    a.init_bubbles(BUB, 3.0)
    a.step(5)
    image = a.save_state()
    a.close()
    b = mcb.Lattice(p)
    b.load_state(image)
    b.step(6)
    got = b.observables()
    b.close()
    o = oracle_for(p)
    o.init(f32_bubbles(BUB), 3.0)
    o.step(11)
    o.macros()
    for name in ("phi", "rho", "ux", "uy", "uz"):
        assert np.array_equal(got[name], o.field(name)), name
    o.close()


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_populations_image_into_moment_form_context(gpu_lib, storage, shape):
    """At tau_g = 1 writing g switches a case to stored g populations; its
    image loads into a fresh (moment-form) context, which switches too."""
    mcb = gpu_lib
    st = mcb.STORE_FP32 if storage == "fp32" else mcb.STORE_FP16
    p = params(mcb, st, 1.0, shape=shape)
    a = mcb.Lattice(p)
    a.init_bubbles(BUB, 3.0)
    a.step(4)
    a.write("g", a.read("g"))
    a.step(2)
    image = a.save_state()
    a.step(5)
    ref = snapshot(a)
    a.close()
    b = mcb.Lattice(p)
    b.init_bubbles(BUB, 3.0)
    b.load_state(image)
    b.step(5)
    assert_same(ref, snapshot(b), storage)
    b.close()


def test_staged_reference_mode_restart(gpu_lib):
    mcb = gpu_lib
    p = params(mcb, mcb.STORE_FP32, 1.0, path=mcb.PATH_STAGED, bc=mcb.BC_REFERENCE)
    a = mcb.Lattice(p)
    a.init_bubbles(BUB, 3.0)
    a.step(6)
    image = a.save_state()
    a.step(5)
    ref = snapshot(a)
    a.close()
    b = mcb.Lattice(p)
    b.load_state(image)
    b.step(5)
    assert_same(ref, snapshot(b), "staged")
    b.close()


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_batch_case_resumes_in_another_batch(gpu_lib, storage):
    """A case saved from one batch resumes, through mcb_batch_add_state, in
    another batch with other cases and at the other batch parity, bit for
    bit like the case that stayed."""
    mcb = gpu_lib
    st = mcb.STORE_FP32 if storage == "fp32" else mcb.STORE_FP16
    cases = [mcb.case_preset("batch", 3000 + i) for i in range(3)]
    with mcb.Batch(st) as b1, mcb.Batch(st) as b2:
        ids = [b1.add_case(c, storage=st) for c in cases]
        b1.step(5)
        image = b1.save_state(ids[1])
        assert mcb.state_info(image)[1] == 5
        other = b2.add_case(mcb.case_preset("batch", 3100), storage=st)
        b2.step(3)                       # odd: the resumed case joins at the other parity
        moved = b2.add_state(image)
        b1.step(8)
        b2.step(8)
        assert b2.case_steps(moved) == 13
        for name in ("f", "g"):
            assert np.array_equal(b1.read(ids[1], name), b2.read(moved, name)), name
        b1.compute_macros(ids[1])
        r1 = {k: b1.read(ids[1], k) for k in ("phi", "rho", "ux")}
        b2.compute_macros(moved)
        r2 = {k: b2.read(moved, k) for k in ("phi", "rho", "ux")}
        assert_same(r1, r2, "batch resume")
        # load_state into an existing case of the same parameters
        b2.load_state(other, b2.save_state(other))
        b2.step(2)


@pytest.mark.parametrize("shape", SHAPES)
def test_moment_form_images_cross_normal_modes(gpu_lib, shape):
    """An FP32 moment-form image holds phi and u, not the normal n (gradCalc
    of that phi): a context that stores n rebuilds it on loading, one that
    recomputes n needs nothing, and a batch case (n stored) resumes from a
    single context's image; every continuation is bit for bit the run that
    never stopped."""
    mcb = gpu_lib
    modes = (mcb.NORMALS_STORED, mcb.NORMALS_RECOMPUTED)
    ref = None
    for src in modes:
        p = mcb.make_params(*shape, tau_f=0.6, tau_g=1.0, sigma=0.1, path=mcb.PATH_FUSED, normals=src)
        with mcb.Lattice(p) as a:
            a.init_bubbles(BUB, 3.0)
            a.step(7)
            image = a.save_state()
            a.step(9)
            got = snapshot(a)
        if ref is None:
            ref = got
        assert_same(dict(ref), got, f"normals {src} straight through")
        for dst in modes:
            p.normals = dst
            b = mcb.Lattice(p)
            b.load_state(image)
            b.step(9)
            assert_same(dict(ref), snapshot(b), f"image of normals {src} into normals {dst}")
            b.close()
        with mcb.Batch(mcb.STORE_FP32) as bt:
            cid = bt.add_state(image)
            bt.step(9)
            for name in ("f", "g"):
                assert np.array_equal(bt.read(cid, name), ref[name]), f"batch from normals {src}: {name}"


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("tau_g", [0.8, 1.0])
def test_context_and_batch_images_interchange(gpu_lib, tau_g, shape):
    """FP32 images are one format whatever the population layout: a single
    context's image resumes as a batch case (populations in planes) and a
    batch case's image resumes in a single context, each bit for bit like
    the run that never stopped."""
    mcb = gpu_lib
    p = params(mcb, mcb.STORE_FP32, tau_g, path=mcb.PATH_FUSED, shape=shape)
    with mcb.Lattice(p) as a:
        a.init_bubbles(BUB, 3.0)
        a.step(6)
        image = a.save_state()
        a.step(7)
        ref = snapshot(a)
    with mcb.Batch(mcb.STORE_FP32) as bt:
        other = bt.add_case(mcb.case_preset("batch", 3200), storage=mcb.STORE_FP32)
        bt.step(1)                       # the resumed case joins at the other parity
        cid = bt.add_state(image)
        bt.step(7)
        for name in ("f", "g"):
            assert np.array_equal(bt.read(cid, name), ref[name]), f"context image in a batch: {name}"
        back = bt.save_state(cid)
        bt.remove(other)
    with mcb.Lattice(p) as b:
        b.load_state(back)
        got = snapshot(b)
    assert_same(dict(ref), got, "batch image in a context")


def test_f_moments_image_continues_bit_for_bit(gpu_lib):
    """f in moment form: an image (the 13 moment planes and phi) resumes the
    run bit for bit in a fresh context, in both normal modes; a context that
    keeps f as populations refuses it."""
    mcb = gpu_lib
    p = mcb.make_params(*SHAPE, tau_f=0.6, tau_g=1.0, sigma=0.1, path=mcb.PATH_FUSED, f_form=mcb.F_MOMENTS)
    with mcb.Lattice(p) as a:
        a.init_bubbles(BUB, 3.0)
        a.step(7)
        image = a.save_state()
        a.step(9)
        ref = snapshot(a)
    for normals in (mcb.NORMALS_STORED, mcb.NORMALS_RECOMPUTED):
        p.normals = normals
        b = mcb.Lattice(p)
        b.load_state(image)
        b.step(9)
        assert_same(dict(ref), snapshot(b), f"f-moment image, normals {normals}")
        b.close()
    p.f_form = mcb.F_POPULATIONS
    with mcb.Lattice(p) as c:
        c.init_bubbles(BUB, 3.0)
        with pytest.raises(mcb.McbError):
            c.load_state(image)


def test_state_images_of_other_cases_are_refused(gpu_lib):
    mcb = gpu_lib
    a = mcb.Lattice(params(mcb, mcb.STORE_FP32, 1.0))
    a.init_bubbles(BUB, 3.0)
    image = a.save_state()
    for other in (mcb.make_params(24, 20, 16, tau_f=0.6, tau_g=1.0, sigma=0.1),
                  params(mcb, mcb.STORE_FP32, 0.8),
                  params(mcb, mcb.STORE_FP16, 1.0)):
        b = mcb.Lattice(other)
        with pytest.raises(mcb.McbError) as e:
            b.load_state(image)
        assert e.value.status == mcb.ERR_INVALID_ARG
        b.close()
    with pytest.raises(mcb.McbError):
        a.load_state(image[:-4])                  # truncated
    bad = bytearray(image)
    bad[0] ^= 0xFF
    with pytest.raises(mcb.McbError):
        a.load_state(bytes(bad))                  # not an image
    with mcb.Batch(mcb.STORE_FP16) as b:
        with pytest.raises(mcb.McbError):
            b.add_state(image)                    # FP32 image into an FP16 batch
    a.close()
