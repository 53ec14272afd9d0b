"""GPU (HIP, gfx950) against the FP64 oracle for the tensor-pomeron continuum of
a pion pair, "PP[CONTENSOR]<C> -> pi+ pi-" (MTensorPomeron::ME4 2xS,
src/MTensorPomeron.cc:726-763), on identical phase-space points.

The GPU evaluates each (proton vertex)(propagator)(meson vertex) chain as the
scalar it reduces to; the oracle keeps the reference's tensor network. As for the
2xV amplitude, |M|^2 depends on t and on pt = pa - p1 - p3 formed from 6.5 TeV
components, so the pointwise bound is 1e-9 with a round-off median."""
import sys

import numpy as np
import pytest

from graniitti_amd import abi
from graniitti_amd.process import MContinuum, MGraniitti, vegas_cfg
import oracle_lib as ol
from conftest import small_tables
from test_gpu_parity import compare, points

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_2xs_unscreened_pointwise():
    cfg = abi.cfg_default(9)
    say("case: preset 9 (tensor pomeron pi+ pi-) FP64 unscreened, 20000 points, seed 21")
    u = points(20000, cfg.dim, 21)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    rel = compare(Wg, Wo, fg, fo, RTOL, "2xS fp64")
    assert np.median(rel) < 1e-13
    assert (Wo > 0).sum() > 5000


def test_2xs_screened_pointwise():
    cfg = small_tables(abi.cfg_default(10), 1000)
    say("case: preset 10 (tensor pomeron pi+ pi-) FP64 screened, 1000-node tables, 1500 points, "
        "seed 22")
    u = points(1500, cfg.dim, 22)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    rel = compare(Wg, Wo, fg, fo, RTOL, "2xS fp64 screened")
    assert np.median(rel) < 1e-13


def test_2xs_factorized_pointwise():
    cfg = abi.cfg_default(9)
    f = abi.cfg_default(4)
    cfg.phase_space = abi.GRA_PS_F
    cfg.rap_min, cfg.rap_max = f.rap_min, f.rap_max
    cfg.gen_M_min, cfg.gen_M_max = f.gen_M_min, f.gen_M_max
    say("case: PP[CONTENSOR]<F> -> pi+ pi- FP64, 10000 points, seed 23")
    u = points(10000, cfg.dim, 23)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    compare(Wg, Wo, fg, fo, RTOL, "2xS <F> fp64")


def test_2xs_vegas_sigma_matches_oracle_iteration():
    # one VEGAS iteration of 2e5 calls: the accumulated sums agree with the oracle's
    cfg = abi.cfg_default(9)
    say("case: preset 9 one VEGAS iteration, 200000 calls")
    import ctypes as C
    g = abi.GraGrid()
    abi.check(abi.load().gra_hip_grid_init(C.byref(g), cfg.dim, 128))
    with MContinuum(cfg) as P:
        sg = P.vegas_iterate(g, 200000, 99, 0)
    so = ol.OracleRun(cfg).vegas_iterate(g, 200000, 99, 0)
    say(f"2xS sum f GPU {sg.fsum:.12e} oracle {so.fsum:.12e}")
    assert abs(sg.fsum - so.fsum) <= 1e-9 * so.fsum
    assert abs(sg.n_valid - so.n_valid) <= 1
