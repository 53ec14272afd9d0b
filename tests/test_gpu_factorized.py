"""The <F> phase-space class (MFactorized, B51 kinematics) on the GPU:
PP[CON]<F> -> pi+ pi- against the FP64 oracle point by point, and against the
<C> class through the cross section, as the reference's
tests/run_compare_PS/run.sh compares them ("With these cuts, <F> and <C> should
match one to one")."""
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


def f_cfg(preset, amp=None, tables=None):
    cfg = abi.cfg_default(preset)
    if tables:
        small_tables(cfg, tables)
    if amp is not None:
        cfg.amp_precision = amp
    return cfg


def test_factorized_fp64_pointwise():
    cfg = f_cfg(4)
    say("case: PP[CON]<F> FP64, 20000 points, seed 31")
    u = points(20000, cfg.dim, 31)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    assert ((fo & 7) == 7).sum() > 5000
    compare(Wg, Wo, fg, fo, 1e-10, "F fp64")


def test_factorized_fp64_pointwise_screened():
    cfg = f_cfg(5, abi.GRA_AMP_FP64, 1000)
    say("case: PP[CON]<F> screened FP64, 1500 points, seed 32")
    u = points(1500, cfg.dim, 32)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    compare(Wg, Wo, fg, fo, 1e-10, "F fp64 screened")


def test_factorized_fp32_screened_error():
    cfg = f_cfg(5, None, 1000)
    say("case: PP[CON]<F> screened FP32 amplitude, 3000 points, seed 33")
    u = points(3000, cfg.dim, 33)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(f_cfg(5, abi.GRA_AMP_FP64, 1000)).eval(u)
    rel = compare(Wg, Wo, fg, fo, 1e-4, "F fp32 screened")
    say(f"F fp32 |M|^2 rel err: median {np.median(rel):.2e} p99 {np.quantile(rel, 0.99):.2e}")
    assert np.median(rel) < 4e-6


def test_factorized_and_continuum_cross_sections_agree():
    # the fiducial region (|eta| < 0.9, pT > 0.15 on both pions, M < 4 GeV) lies
    # inside both generation regions: <F> |Y| < 0.9, M < 4; <C> |y_i| < 0.9 and
    # kt < 10 (|p3T - p4T| reaches 5.2 GeV there, beyond the default kt < 5)
    cF = f_cfg(4)
    cF.sys_M_max = 4.0
    cC = abi.cfg_default(1)
    cC.sys_M_max = 4.0
    cC.kt_max = 10.0
    res = {}
    for name, cfg in (("F", cF), ("C", cC)):
        say(f"case: VEGAS sigma, <{name}> class, 12 x 2^20 calls")
        with MContinuum(cfg) as P:
            r = MGraniitti(P).Integrate(vegas_cfg(1 << 20, 12, burnin=4, seed=77))
        res[name] = (r.sigma, r.sigma_err)
        say(f"<{name}>: sigma {r.sigma:.6e} +- {r.sigma_err:.2e} barn")
    (sF, eF), (sC, eC) = res["F"], res["C"]
    assert eF < 2e-3 * sF and eC < 2e-3 * sC
    assert abs(sF - sC) < 4.0 * np.hypot(eF, eC), (sF, sC)


def test_factorized_events():
    cfg = f_cfg(4)
    say("case: 2000 unweighted <F> events")
    with MContinuum(cfg) as P:
        G = MGraniitti(P)
        G.Integrate(vegas_cfg(1 << 18, 4, burnin=3, seed=9))
        p, f, tot = G.Generate(2000, calls=1 << 20, seed=3)
    assert p.shape == (2000, 4, 4)
    E = cfg.beam_E1
    tot4 = p[:, 0] + p[:, 1] + p[:, 2] + p[:, 3]
    assert np.abs(tot4 - np.array([2 * E, 0, 0, 0])).max() < 1e-6 * E
    m2 = p[:, 2:, 0] ** 2 - (p[:, 2:, 1:] ** 2).sum(-1)
    assert np.allclose(m2, cfg.m_central ** 2, atol=1e-8)
    X = p[:, 2] + p[:, 3]
    MX = np.sqrt(X[:, 0] ** 2 - (X[:, 1:] ** 2).sum(-1))
    YX = 0.5 * np.log((X[:, 0] + X[:, 3]) / (X[:, 0] - X[:, 3]))
    assert MX.max() <= cfg.gen_M_max * (1 + 1e-12)
    assert np.abs(YX).max() <= cfg.rap_max * (1 + 1e-9)
