"""GPU test of gra_hip_eval_points_noalloc, the entry point that allocates no
device memory: through the C ABI, its weights, |M|^2 and flags equal
gra_hip_eval_points_device's bit for bit on two calls in a row, the second with
a smaller input (the scratch array keeps the first call's contents), and a size
beyond the scratch fails with GRA_E_ARG. Device buffers are the caller's (torch
tensors here: plumbing only)."""
import ctypes as C
import sys

import numpy as np
import pytest

from graniitti_amd import abi
from graniitti_amd.process import MContinuum
from conftest import small_tables

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _device_eval(lib, h, fn, u_dev, n, torch):
    W = torch.full((n,), -1.0, dtype=torch.float64, device="cuda")
    a2 = torch.full((n,), -1.0, dtype=torch.float64, device="cuda")
    fl = torch.full((n,), 255, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    rc = fn(h, C.c_void_p(u_dev.data_ptr()), n, C.c_void_p(W.data_ptr()),
            C.c_void_p(a2.data_ptr()), C.c_void_p(fl.data_ptr()))
    if rc != abi.GRA_OK:
        raise abi.GraError(f"status {rc}: {lib.gra_hip_last_error(h).decode()}")
    torch.cuda.synchronize()
    return W.cpu().numpy(), a2.cpu().numpy(), fl.cpu().numpy()


@pytest.mark.parametrize("preset", [2, 1, 3, 7])
def test_noalloc_matches_device_entry_point(preset):
    torch = pytest.importorskip("torch")
    cfg = abi.cfg_default(preset)
    if cfg.screening:
        small_tables(cfg, 1000)
    lib = abi.load()
    sizes = (50000, 7001) if preset != 3 else (6000, 901)
    say(f"case: noalloc vs device entry point, preset {preset}, n = {sizes}")
    rng = np.random.default_rng(100 + preset)
    with MContinuum(cfg) as P:
        for n in sizes:
            u = torch.tensor(rng.random((n, cfg.dim)), dtype=torch.float64, device="cuda")
            Wd, ad, fd = _device_eval(lib, P.h, lib.gra_hip_eval_points_device, u, n, torch)
            Wn, an, fn = _device_eval(lib, P.h, lib.gra_hip_eval_points_noalloc, u, n, torch)
            say(f"preset {preset} n={n}: nonzero {int((Wd != 0).sum())}")
            assert (Wd != 0).sum() > n // 20
            np.testing.assert_array_equal(fn, fd)
            np.testing.assert_array_equal(Wn, Wd)
            np.testing.assert_array_equal(an, ad)


def test_noalloc_rejects_sizes_beyond_the_scratch():
    torch = pytest.importorskip("torch")
    cfg = small_tables(abi.cfg_default(2), 1000)
    lib = abi.load()
    with MContinuum(cfg) as P:
        n = 2_000_000  # 2e6 screened candidate records need ~0.45 GB
        u = torch.zeros((16, cfg.dim), dtype=torch.float64, device="cuda")
        W = torch.zeros(16, dtype=torch.float64, device="cuda")
        fl = torch.zeros(16, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        # the size check comes before any device access, so the short buffers are not read
        rc = lib.gra_hip_eval_points_noalloc(P.h, C.c_void_p(u.data_ptr()), n,
                                             C.c_void_p(W.data_ptr()), None,
                                             C.c_void_p(fl.data_ptr()))
        assert rc == abi.GRA_E_ARG
        assert "scratch" in lib.gra_hip_last_error(P.h).decode()
