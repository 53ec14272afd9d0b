import os
import sys

import pytest

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
