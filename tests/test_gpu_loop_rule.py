"""The screened pi+ pi- workload in its stated loop discretisation (preset 11: 32
Gauss-Legendre kt nodes x 16 azimuths, A_eik on 64 log-spaced kt nodes) on the
GPU: pointwise against the oracle, and its cross section against the reference's
discretisation (preset 2: Simpson 3/8 on 16 x 31 nodes, A_eik from the 10001-row
table) and against refined grids, with VEGAS."""
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


def test_config_mode_fp64_pointwise():
    cfg = small_tables(abi.cfg_default(11), 1000)
    cfg.amp_precision = abi.GRA_AMP_FP64
    say("case: preset 11 in FP64 (Gauss loop, log-spaced A_eik), 1000-row tables, 2000 points")
    u = points(2000, cfg.dim, 51)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    compare(Wg, Wo, fg, fo, 1e-10, "preset 11 fp64")


def test_config_mode_fp32_error():
    cfg = small_tables(abi.cfg_default(11), 1000)
    say("case: preset 11 FP32 amplitude, 1000-row tables, 4000 points")
    u = points(4000, cfg.dim, 52)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    ref = small_tables(abi.cfg_default(11), 1000)
    ref.amp_precision = abi.GRA_AMP_FP64
    Wo, ao, fo = ol.OracleRun(ref).eval(u)
    rel = compare(Wg, Wo, fg, fo, 1e-4, "preset 11 fp32")
    assert np.median(rel) < 4e-6


def sigma(cfg, seed=99):
    with MContinuum(cfg) as P:
        r = MGraniitti(P).Integrate(vegas_cfg(1 << 20, 1, burnin=5, seed=seed, precision=3e-4,
                                              max_iter=60))
    return r.sigma, r.sigma_err


def test_cross_section_against_loop_discretisations():
    # FP32 amplitude, full tables; 3e-4 relative VEGAS errors. Refining the
    # reference's kt grid moves sigma towards the Gauss rule, whose own
    # refinements in kt and phi move it much less.
    runs = {}
    for name, rule, nphi, nkt, nlog in [
            ("Simpson 16x31 (preset 2)", abi.GRA_LOOP_SIMPSON38, 15, 30, 0),
            ("Simpson 16x61", abi.GRA_LOOP_SIMPSON38, 15, 60, 0),
            ("Simpson 31x31", abi.GRA_LOOP_SIMPSON38, 30, 30, 0),
            ("Gauss 16x32, table", abi.GRA_LOOP_GAUSS, 15, 31, 0),
            ("Gauss 16x64, table", abi.GRA_LOOP_GAUSS, 15, 63, 0),
            ("Gauss 32x32, table", abi.GRA_LOOP_GAUSS, 31, 31, 0),
            ("Gauss 16x32, 64 log nodes (preset 11)", abi.GRA_LOOP_GAUSS, 15, 31, 64)]:
        cfg = abi.cfg_default(2)
        cfg.loop_rule, cfg.n_loop_phi, cfg.n_loop_kt, cfg.eik_log_nodes = rule, nphi, nkt, nlog
        say(f"case: VEGAS sigma with the loop {name}")
        runs[name] = sigma(cfg)
    base = runs["Simpson 16x31 (preset 2)"]
    for k, (s, e) in runs.items():
        say(f"{k:40s} sigma {s:.6e} +- {e:.1e}  ({s / base[0] - 1:+.4f} vs preset 2)")
    g64 = runs["Gauss 16x64, table"][0]
    d_ref = abs(base[0] - g64)
    assert abs(runs["Simpson 16x61"][0] - g64) < 0.5 * d_ref
    assert abs(runs["Gauss 16x32, table"][0] - g64) < 0.3 * d_ref
    assert abs(runs["Gauss 32x32, table"][0] - runs["Gauss 16x32, table"][0]) < 0.3 * d_ref
    assert abs(runs["Gauss 16x32, 64 log nodes (preset 11)"][0] / base[0] - 1) < 0.1
