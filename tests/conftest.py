import os
import sys

import pytest

# torch's wheel carries its own libamdhip64 under the same soname as the one the
# HIP library links; whichever loads first serves the process. Importing torch
# before any test loads the library keeps one HIP runtime in the process, so
# tests that hand torch device tensors to the library can initialise torch's
# device in any order or subset.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def small_tables(cfg, n=1000):
    """Reduced eikonal discretisation for fast CPU/GPU parity tests."""
    cfg.n_bt = n
    cfg.n_kt2 = n
    cfg.n_fb = n
    return cfg
