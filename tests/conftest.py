import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libostr_hip.so)")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def small_case(oracle_lib):
    """Config-1 workload (200x100x20, 9 views): phantom, noisy counts, blank."""
    from newrecon_amd.geometry import scan_small
    from newrecon_amd.phantom import breast_phantom, poisson_counts
    s = scan_small()
    x = breast_phantom(s.geom, s.phantom_seed, s.n_fg, s.n_calc)
    l = oracle_lib.fproj(s.geom, s.angles, x, 0, s.n_views)
    y = poisson_counts(l, s.i0, s.noise_seed)
    b = np.full_like(y, s.i0)
    return s, x, y, b


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
