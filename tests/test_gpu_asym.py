"""Unequal beam energies on the GPU (MContinuum::BNBuildKin's boost to the lab,
src/MContinuum.cc:386-395) against the FP64 oracle on identical points, with the
fiducial cuts applied in the lab: the FP64 reference-order path to 1e-10
relative per point (1e-9 for the rho rho tensor process), the FP32 tree
amplitude to its usual 2e-5, and the events of a generation pass carrying the
lab beams' four-momentum."""
import sys

import numpy as np
import pytest

from graniitti_amd import abi
from graniitti_amd.process import MContinuum, MGraniitti, vegas_cfg
import oracle_lib as ol
from test_gpu_parity import compare, points
from test_oracle_asym import asym_cfg, beams

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("preset", [1, 3])
def test_unequal_beams_fp64_pointwise(preset):
    _, cfg = asym_cfg(preset)
    cfg.screening = 0
    cfg.amp_precision = abi.GRA_AMP_FP64
    say(f"case: preset {preset} unscreened FP64, beams {cfg.beam_E1:.1f} + {cfg.beam_E2:.1f} GeV, "
        "20000 points, seed 31")
    u = points(20000, cfg.dim, 31)
    with MContinuum(cfg) as P:
        Wg, _, fg = P.eval(u)
    Wo, _, fo = ol.OracleRun(cfg).eval(u)
    compare(Wg, Wo, fg, fo, 1e-9 if preset == 3 else 1e-10, f"preset {preset} unequal beams fp64")


def test_unequal_beams_fp32_tree_amplitude():
    _, cfg = asym_cfg(1)
    cfg.amp_precision = abi.GRA_AMP_FP32
    say("case: preset 1 FP32 amplitude, unequal beams, 20000 points, seed 32")
    u = points(20000, cfg.dim, 32)
    with MContinuum(cfg) as P:
        Wg, _, fg = P.eval(u)
    ocfg = abi.GraCfg.from_buffer_copy(cfg)
    ocfg.amp_precision = abi.GRA_AMP_FP64
    Wo, _, fo = ol.OracleRun(ocfg).eval(u)
    rel = compare(Wg, Wo, fg, fo, 2e-5, "preset 1 unequal beams fp32")
    assert np.median(rel) < 2e-6


def test_unequal_beams_events_carry_the_lab_beams():
    _, cfg = asym_cfg(1)
    say("case: 2000 unweighted events, preset 1, unequal beams")
    with MContinuum(cfg) as P:
        G = MGraniitti(P)
        G.Integrate(vegas_cfg(1 << 18, 4, burnin=3, seed=19))
        p, f, tot = G.Generate(2000, calls=1 << 20, seed=8)
    say(f"events {len(p)} counters {tot}")
    assert len(p) == 2000
    b1, b2 = beams(cfg)
    total = p[:, 0] + p[:, 1] + p[:, 2] + p[:, 3]
    assert np.abs(total - (b1 + b2)).max() < 1e-6 * cfg.beam_E1
