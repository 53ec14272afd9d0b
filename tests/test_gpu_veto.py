"""Veto domains on the GPU (VETOCUTS; MProcess::VetoCuts and FindVetoCuts,
src/MProcess.cc:1069-1082, 1157-1175) against the FP64 oracle on identical
points: flags exactly, weights to the tolerance of each path (FP64
reference-order path 1e-10 relative per point, FP32 screening loop 1e-4), and
the unweighted events of a generation pass, none of which may carry a central
final state inside a domain."""
import sys

import numpy as np
import pytest

from graniitti_amd import abi
from graniitti_amd.process import MContinuum, MGraniitti, vegas_cfg
import oracle_lib as ol
from conftest import small_tables
from test_gpu_parity import compare, points
from test_oracle_veto import DOMAINS, in_domain, with_veto

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_veto_flags(fg, Wg):
    vetoed = (fg & abi.FLAG_VETO_OK) == 0
    assert np.all(Wg[vetoed] == 0.0)
    return int(vetoed.sum())


@pytest.mark.parametrize("preset", [1, 3])
def test_veto_fp64_pointwise(preset):
    base = abi.cfg_default(preset)
    base.screening = 0
    base.amp_precision = abi.GRA_AMP_FP64
    cfg = with_veto(base, DOMAINS)
    say(f"case: preset {preset} unscreened FP64 with two veto domains, 20000 points, seed 21")
    u = points(20000, cfg.dim, 21)
    with MContinuum(cfg) as P:
        Wg, _, fg = P.eval(u)
    Wo, _, fo = ol.OracleRun(cfg).eval(u)
    compare(Wg, Wo, fg, fo, 1e-9 if preset == 3 else 1e-10, f"preset {preset} veto fp64")
    n = check_veto_flags(fg, Wg)
    say(f"vetoed {n} of {len(u)}")
    assert n > 1000


def test_veto_screened_fp32():
    cfg = with_veto(small_tables(abi.cfg_default(2), 1000), DOMAINS)
    say("case: preset 2 screened FP32 with two veto domains, 3000 points, seed 22")
    u = points(3000, cfg.dim, 22)
    with MContinuum(cfg) as P:
        Wg, _, fg = P.eval(u)
    ocfg = abi.GraCfg.from_buffer_copy(cfg)
    ocfg.amp_precision = abi.GRA_AMP_FP64
    Wo, _, fo = ol.OracleRun(ocfg).eval(u)
    rel = compare(Wg, Wo, fg, fo, 1e-4, "preset 2 veto fp32 screened")
    assert np.median(rel) < 4e-6
    assert check_veto_flags(fg, Wg) > 200


def test_generated_events_avoid_the_veto_domains():
    cfg = with_veto(abi.cfg_default(1), DOMAINS)
    say("case: 3000 unweighted events, preset 1 with two veto domains")
    with MContinuum(cfg) as P:
        G = MGraniitti(P)
        G.Integrate(vegas_cfg(1 << 18, 4, burnin=3, seed=17))
        p, f, tot = G.Generate(3000, calls=1 << 20, seed=9)
    say(f"events {len(p)} counters {tot}")
    assert len(p) == 3000
    for ev in p:
        assert not any(in_domain(q, DOMAINS) for q in ev[2:4])
