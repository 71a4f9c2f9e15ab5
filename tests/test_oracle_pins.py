"""Direct CPU pins of oracle stages that conservation checks alone leave open.

momentiCalc's explicit Pi link sums (src/lbm.cu:213-218), collisionFluid's
regularised non-equilibrium term (src/lbm.cu:278-286) and fgBoundary
(src/lbm.cu:365-401) are compared with closed forms derived from the D3Q19
moment identities, or with a brute-force restatement over target cells, on
random states far from equilibrium. A wrong link in a Pi sum, a sign slip in
the 2 c_a c_b Pi_ab term, or omega in place of (1 - omega) keeps mass and
momentum (which the checks in test_oracle_kat.py cover) but fails these.
The initial profile's FP32 tanh is compared with the C library's tanh.
"""
import numpy as np

import oracle as O

C = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0],
              [0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1],
              [0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1]])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
CS2 = 1.0 / 3.0
PI_KEYS = {(0, 0): "pxx", (1, 1): "pyy", (2, 2): "pzz", (0, 1): "pxy", (0, 2): "pxz", (1, 2): "pyz"}


def random_nonequilibrium(o, seed, shape):
    """f far from equilibrium (+-40% per link, uneven across links) and a
    random force field; returns (f, F)."""
    rng = np.random.default_rng(seed)
    f = W.reshape(19, 1, 1, 1) * (1 + 0.4 * rng.uniform(-1, 1, (19,) + shape))
    F = 0.05 * rng.standard_normal((3,) + shape)
    o.field("f")[...] = f
    for a, k in enumerate(("ffx", "ffy", "ffz")):
        o.field(k)[...] = F[a]
    return f, F


def macroscopic(f, F):
    rho = f.sum(0)
    u = np.einsum("al,l...->a...", C, f) / rho + 0.5 * F / rho
    m2 = np.einsum("al,bl,l...->ab...", C, C, f)
    return rho, u, m2


def pi_closed_form(rho, u, F, m2):
    """Pi_ab = sum_l c_a c_b (f_l - feq_l + HeF_l). With the D3Q19 identities
    sum w c_a c_b = cs2 delta_ab and sum w c_a c_b c_c c_d = cs2^2 (delta delta
    + delta delta + delta delta), the second moment of feq is
    rho (cs2 delta_ab + u_a u_b) and that of He's forcing term is
    0.5 (F_a u_b + F_b u_a) - 1.5 (u.F) u_a u_b."""
    uF = np.einsum("a...,a...->...", u, F)
    eye = np.eye(3).reshape(3, 3, *([1] * rho.ndim))
    uu = np.einsum("a...,b...->ab...", u, u)
    Fu = np.einsum("a...,b...->ab...", F, u)
    return m2 - rho * (CS2 * eye + uu) + 0.5 * (Fu + Fu.transpose(1, 0, *range(2, Fu.ndim))) - 1.5 * uF * uu


def test_pi_link_sums_match_the_moment_closed_form():
    shape = (4, 5, 6)
    o = O.Oracle(6, 5, 4, prec=64, tau_f=0.6)
    f, F = random_nonequilibrium(o, 11, shape)
    o.stage("momentiCalc")
    rho, u, m2 = macroscopic(f, F)
    assert np.allclose(o.field("rho"), rho, rtol=0, atol=1e-14)
    for a, k in enumerate(("ux", "uy", "uz")):
        assert np.allclose(o.field(k), u[a], rtol=0, atol=1e-14), k
    pi = pi_closed_form(rho, u, F, m2)
    # Pi is of order 0.05 here: a link missing from a sum moves it by ~w f.
    assert np.abs(pi).max() > 1e-2
    for (a, b), k in PI_KEYS.items():
        assert np.allclose(o.field(k), pi[a, b], rtol=0, atol=1e-13), k


def test_fluid_collision_second_moment_matches_the_closed_form():
    """sum_l c_a c_b f*_l = rho (cs2 delta_ab + u_a u_b) + (1 - omega) Pi_ab
    + 0.5 (F_a u_b + F_b u_a) - 1.5 (u.F) u_a u_b: the regularised term
    (w_l / 2cs2^2) (c_a c_b - cs2 delta_ab) Pi_ab has second moment Pi_ab and
    zero mass and momentum. tau_f = 0.6 makes omega and 1 - omega differ."""
    shape = (5, 6, 7)
    tau = 0.6
    o = O.Oracle(7, 6, 5, prec=64, tau_f=tau)
    f, F = random_nonequilibrium(o, 12, shape)
    o.stage("momentiCalc")
    rho, u, m2 = macroscopic(f, F)
    pi = np.stack([np.stack([o.field(PI_KEYS[tuple(sorted((a, b)))]) for b in range(3)]) for a in range(3)])
    o.stage("collisionFluid")          # pushes f*_l(x) to x + c_l
    fs = o.field("f")
    post = np.stack([np.roll(fs[l], shift=(-C[2, l], -C[1, l], -C[0, l]), axis=(0, 1, 2)) for l in range(19)])
    omega = 1.0 / tau
    uF = np.einsum("a...,a...->...", u, F)
    uu = np.einsum("a...,b...->ab...", u, u)
    Fu = np.einsum("a...,b...->ab...", F, u)
    eye = np.eye(3).reshape(3, 3, 1, 1, 1)
    want = rho * (CS2 * eye + uu) + (1 - omega) * pi + 0.5 * (Fu + Fu.transpose(1, 0, 2, 3, 4)) - 1.5 * uF * uu
    got = np.einsum("al,bl,l...->ab...", C, C, post)
    assert np.allclose(got, want, rtol=0, atol=1e-13)
    # The regularised part matters at this tolerance: omega in place of
    # (1 - omega) would move the second moment by |2 omega - 1| |Pi|.
    assert np.abs((omega - (1 - omega)) * pi).max() > 1e-3
    # Mass and momentum as before.
    assert np.allclose(post.sum(0), rho, rtol=0, atol=1e-14)


def test_fg_boundary_matches_a_target_cell_restatement():
    """fgBoundary (lbm.cu:365-401): for every face node e and link l with
    e + c_l inside the lattice, f_l(e + c_l) = rho(e) w_l and
    g_l(e + c_l) = phi(e) w_l; every other population is untouched. Restated
    over target cells x: x - c_l is a face node and x lies inside."""
    nx, ny, nz = 7, 6, 5
    o = O.Oracle(nx, ny, nz, prec=32, bc_mode=O.REFERENCE, weights=O.W_LITERAL)
    rng = np.random.default_rng(13)
    shape = (nz, ny, nx)
    f0 = rng.uniform(0, 1, (19,) + shape).astype(np.float32)
    g0 = rng.uniform(0, 1, (19,) + shape).astype(np.float32)
    rho = rng.uniform(0.5, 1.5, shape).astype(np.float32)
    phi = rng.uniform(0, 1, shape).astype(np.float32)
    o.field("f")[...] = f0
    o.field("g")[...] = g0
    o.field("rho")[...] = rho
    o.field("phi")[...] = phi
    o.stage("fgBoundary")
    w = o.weights().astype(np.float32)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f_want, g_want = f0.copy(), g0.copy()
    for l in range(19):
        sx, sy, sz = x - C[0, l], y - C[1, l], z - C[2, l]
        inside = (sx >= 0) & (sx < nx) & (sy >= 0) & (sy < ny) & (sz >= 0) & (sz < nz)
        face = (sx == 0) | (sx == nx - 1) | (sy == 0) | (sy == ny - 1) | (sz == 0) | (sz == nz - 1)
        hit = inside & face
        src_rho = rho[np.clip(sz, 0, nz - 1), np.clip(sy, 0, ny - 1), np.clip(sx, 0, nx - 1)]
        src_phi = phi[np.clip(sz, 0, nz - 1), np.clip(sy, 0, ny - 1), np.clip(sx, 0, nx - 1)]
        f_want[l][hit] = (src_rho * w[l])[hit]
        g_want[l][hit] = (src_phi * w[l])[hit]
        if l == 0:
            assert hit.sum() == 2 * (nx * ny + ny * nz + nx * nz) - 4 * (nx + ny + nz) + 8
    assert np.array_equal(o.field("f"), f_want)
    assert np.array_equal(o.field("g"), g_want)
    # Interior targets of interior sources keep their populations.
    assert np.array_equal(o.field("f")[0, 2:-2, 2:-2, 2:-2], f0[0, 2:-2, 2:-2, 2:-2])


def test_fp32_tanh_of_the_initial_profile_is_the_rounded_libm_tanh():
    """Every float in [0, 20): tanh_f32 equals the C library's double tanh
    rounded to float (tanh is odd and tanh_f32 is built symmetric; beyond 20
    both are exactly 1)."""
    n, diff, far = O.tanh_scan(0, np.float32(20.0).view(np.uint32) - 1)
    assert n == int(np.float32(20.0).view(np.uint32))
    assert far == 0
    assert diff == 0
