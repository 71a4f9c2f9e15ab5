import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "multic-bubble-cuda_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libmcb.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def oracle_for(params, prec=32):
    """Oracle state with the same physics as a libmcb parameter set."""
    import oracle as O
    return O.Oracle(params.nx, params.ny, params.nz, prec=prec, bc_mode=params.bc_mode,
                    weights=params.weight_set, tau_f=float(params.tau_f), tau_g=float(params.tau_g),
                    sigma=float(params.sigma), sharp_c=float(params.sharp_c),
                    moving_wall=bool(params.moving_wall))


def f32_bubbles(bubbles):
    return [tuple(float(np.float32(v)) for v in b) for b in bubbles]


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return num / den if den > 0 else num


@pytest.fixture(scope="session")
def gpu_lib():
    import mcb
    if mcb.device_count() < 1:
        pytest.fail("no GPU visible to libmcb")
    return mcb
