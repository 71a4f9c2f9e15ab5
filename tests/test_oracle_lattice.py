"""D3Q19 tables (src/var.cu:29-31, 40-45) against what the mathematics fixes.

Pins the oracle's link table and both weight sets: exact moment isotropy in
rationals, opposite-link pairing, the float sums of the two weight sets, and
that libmcb's constants equal the oracle's.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

CX = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0]
CY = [0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1]
CZ = [0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1]
C = np.array([CX, CY, CZ]).T
W_EXACT = [Fraction(1, 3)] + [Fraction(1, 18)] * 6 + [Fraction(1, 36)] * 12


def test_velocity_set_is_d3q19():
    vecs = {tuple(c) for c in C}
    assert len(vecs) == 19
    assert (0, 0, 0) in vecs
    norms = sorted(int((np.array(c) ** 2).sum()) for c in C)
    assert norms == [0] + [1] * 6 + [2] * 12


def test_opposite_pairs():
    # links (2k-1, 2k) are opposite for k = 1..9
    for k in range(1, 10):
        assert tuple(-C[2 * k - 1]) == tuple(C[2 * k])


def test_exact_isotropy_in_rationals():
    d = lambda a, b: 1 if a == b else 0
    assert sum(W_EXACT) == 1
    for a in range(3):
        assert sum(w * c[a] for w, c in zip(W_EXACT, C)) == 0
        for b in range(3):
            assert sum(w * c[a] * c[b] for w, c in zip(W_EXACT, C)) == Fraction(d(a, b), 3)
            for g in range(3):
                assert sum(w * c[a] * c[b] * c[g] for w, c in zip(W_EXACT, C)) == 0
                for e in range(3):
                    want = Fraction(d(a, b) * d(g, e) + d(a, g) * d(b, e) + d(a, e) * d(b, g), 9)
                    got = sum(w * c[a] * c[b] * c[g] * c[e] for w, c in zip(W_EXACT, C))
                    assert got == want


def _float_weight_sum(w):
    return sum(Fraction(float(np.float32(x))) for x in w)


def test_literal_weights_sum_exceeds_one_by_2_pow_minus_26():
    o = O.Oracle(4, 4, 4, prec=32, weights=O.W_LITERAL)
    w = o.weights()
    assert np.array_equal(np.float32(w), np.float32([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12))
    assert _float_weight_sum(w) == 1 + Fraction(1, 2 ** 26)


def test_unit_sum_weights_are_exact_and_within_one_ulp():
    o = O.Oracle(4, 4, 4, prec=32, weights=O.W_UNIT_SUM)
    w = np.float32(o.weights())
    assert _float_weight_sum(w) == 1
    lit = np.float32([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    ulps = np.abs(w.view(np.int32).astype(np.int64) - lit.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1
    assert w[7] * 2 == w[1]


@pytest.mark.parametrize("weights", [0, 1])
def test_libmcb_constants_equal_oracle(weights):
    import mcb
    p = mcb.make_params(8, 8, 8, weights=weights, tau_f=0.505, tau_g=0.8, sigma=0.1)
    c = mcb.op_consts(p)
    o = O.Oracle(8, 8, 8, prec=32, weights=weights)
    w = np.float32(o.weights())
    assert np.float32(c.w[0]) == w[0] and np.float32(c.w[1]) == w[1] and np.float32(c.w[2]) == w[7]
    assert np.float32(c.omega) == np.float32(1.0) / np.float32(0.505)     # var.cu:23
    assert np.float32(c.sharp_c) == np.float32(0.15) * np.float32(3.0)    # var.cu:24
    assert np.float32(c.omega_g) == np.float32(1.0) / np.float32(0.8)
