"""PP[RES+CON]<C> -> pi+ pi- (tests/processes/ALICECUTS.json: continuum plus eight
resonances with spin correlations in the Collins-Soper frame) on the GPU against the
FP64 oracle point by point, with and without the eikonal screening loop, which
evaluates the full amplitude at every loop node (PROC_17 sets lts.hamp = {A})."""
import sys

import numpy as np
import pytest

from graniitti_amd import abi
from graniitti_amd.process import MContinuum, MGraniitti, vegas_cfg
import oracle_lib as ol
from conftest import small_tables
from test_gpu_parity import compare, points

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_res_con_fp64_pointwise():
    cfg = abi.cfg_default(6)
    say("case: PP[RES+CON] FP64, 20000 points, seed 41")
    u = points(20000, cfg.dim, 41)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    assert ((fo & 7) == 7).sum() > 15000
    compare(Wg, Wo, fg, fo, 1e-10, "RES+CON fp64")


def test_res_con_screened_fp64_pointwise():
    cfg = small_tables(abi.cfg_default(7), 1000)
    say("case: PP[RES+CON] screened FP64, 1500 points, seed 42")
    u = points(1500, cfg.dim, 42)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    compare(Wg, Wo, fg, fo, 1e-10, "RES+CON screened fp64")


def test_res_con_without_resonances_is_the_continuum():
    cfg = abi.cfg_default(6)
    cfg.n_res = 0
    say("case: PP[RES+CON] with no resonances vs PP[CON], 20000 points, seed 43")
    u = points(20000, cfg.dim, 43)
    with MContinuum(cfg) as P:
        Wr, _, fr = P.eval(u)
    with MContinuum(abi.cfg_default(1)) as P:
        Wc, _, fc = P.eval(u)
    assert np.array_equal(fr, fc)
    assert np.array_equal(Wr, Wc)


def test_res_con_events_and_f2_peak():
    cfg = abi.cfg_default(6)
    say("case: 4000 unweighted PP[RES+CON] events")
    with MContinuum(cfg) as P:
        G = MGraniitti(P)
        r = G.Integrate(vegas_cfg(1 << 18, 5, burnin=3, seed=11))
        p, f, tot = G.Generate(4000, calls=1 << 20, seed=5)
    say(f"sigma {r.sigma:.5e} +- {r.sigma_err:.1e} barn")
    assert p.shape == (4000, 4, 4)
    E = cfg.beam_E1
    tot4 = p[:, 0] + p[:, 1] + p[:, 2] + p[:, 3]
    assert np.abs(tot4 - np.array([2 * E, 0, 0, 0])).max() < 1e-6 * E
    X = p[:, 2] + p[:, 3]
    MX = np.sqrt(X[:, 0] ** 2 - (X[:, 1:] ** 2).sum(-1))
    # the f2(1270) band holds more events per GeV than its continuum-only neighbours
    band = ((MX > 1.2) & (MX < 1.35)).sum() / 0.15
    side = ((MX > 1.6) & (MX < 1.9)).sum() / 0.3
    say(f"events per GeV: f2 band {band:.0f}, 1.6-1.9 GeV {side:.0f}")
    assert band > 2 * side
