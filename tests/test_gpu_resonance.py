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


def test_res_con_fp32_amplitude():
    # the FP32-amplitude mode of the resonance processes: tree amplitudes in FP64
    # (unscreened weights equal the FP64 mode's), screening-loop nodes with FP64
    # kinematics and an FP32 amplitude, against the FP64 oracle
    say("case: PP[RES+CON] FP32 amplitude unscreened, 5000 points, seed 44")
    u = points(5000, 8, 44)
    c32 = abi.cfg_default(6)
    c32.amp_precision = abi.GRA_AMP_FP32
    with MContinuum(c32) as P:
        W32, _, f32 = P.eval(u)
    with MContinuum(abi.cfg_default(6)) as P:
        W64, _, f64 = P.eval(u)
    assert np.array_equal(f32, f64) and np.array_equal(W32, W64)
    # Preset 7 (continuum + eight resonances) point by point. Preset 8 (the lone
    # f2 with |J_z| = 2, no cuts, no continuum) reaches weights 1e-150 below its
    # largest far from the pole and at the zeros of sin^4(theta_CS) sin^4(theta_q),
    # below FP32's range: the pointwise bound holds for weights above 1e-8 of the
    # largest, and the summed deviation is bounded against the summed weight.
    for preset, n, seed in ((7, 1500, 45), (8, 1500, 47)):
        cfg = small_tables(abi.cfg_default(preset), 1000)
        if not cfg.screening:
            cfg.screening = 1
        ref = small_tables(abi.cfg_default(preset), 1000)
        ref.screening = cfg.screening
        cfg.amp_precision = abi.GRA_AMP_FP32
        say(f"case: preset {preset} screened, FP32 loop amplitude, {n} points, seed {seed}")
        u = points(n, cfg.dim, seed)
        with MContinuum(cfg) as P:
            Wg, ag, fg = P.eval(u)
        Wo, ao, fo = ol.OracleRun(ref).eval(u)
        if preset == 7:
            rel = compare(Wg, Wo, fg, fo, 1e-4, f"preset {preset} fp32 screened")
        else:
            assert np.array_equal(fg, fo)
            big = Wo > 1e-8 * Wo.max()
            rel = np.abs(Wg[big] - Wo[big]) / Wo[big]
            dev = np.abs(Wg - Wo).sum() / Wo.sum()
            say(f"preset 8 fp32 screened: {big.sum()} of {n} points above 1e-8 of the largest "
                f"weight: max rel {rel.max():.2e}; sum |dW| / sum W {dev:.2e}")
            assert rel.max() < 1e-4 and dev < 1e-5
        say(f"preset {preset} fp32 rel err: median {np.median(rel):.2e} "
            f"p99 {np.quantile(rel, 0.99):.2e}")
        assert np.median(rel) < 4e-6


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


def test_res_f_pointwise():
    cfg = abi.cfg_default(8)
    say("case: PP[RES]<F> f2(1270) FP64, 20000 points, seed 44")
    u = points(20000, cfg.dim, 44)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    assert ((fo & 7) == 7).sum() > 10000
    # a lone f2 has |A|^2 ~ sin^4 theta_CS sin^4 theta_q, which vanishes on the axes; there
    # the spin trace is a small difference of O(1) terms, formed as D^+ rho D literally by
    # the oracle and contracted as sum rho_ab g_a g_b* on the GPU, so the two roundings
    # differ by ~eps / trace relative: 1e-9 bound, and the bulk at rounding level
    rel = compare(Wg, Wo, fg, fo, 1e-9, "RES<F> fp64")
    assert np.median(rel) < 1e-14


def test_res_f_screened_pointwise():
    cfg = small_tables(abi.cfg_default(8), 1000)
    cfg.screening = 1
    say("case: PP[RES]<F> f2(1270) screened FP64, 1500 points, seed 45")
    u = points(1500, cfg.dim, 45)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    compare(Wg, Wo, fg, fo, 1e-10, "RES<F> screened fp64")


def frame_cos(p, frame, E, mb):
    """cos of the pi+ polar angle in the system rest frame about the frame's axis"""
    from test_oracle_resonance import boost_rest, cos_to_cs_axis
    out = np.empty(len(p))
    for i in range(len(p)):
        X, pip = p[i, 2] + p[i, 3], p[i, 2]
        if frame == abi.GRA_FRAME_CS:
            out[i] = cos_to_cs_axis(X, pip, E, mb)
            continue
        q = boost_rest(X, pip)[1:]
        axis = X[1:] / np.linalg.norm(X[1:]) if frame == abi.GRA_FRAME_HX else np.array([0, 0, 1.0])
        out[i] = q @ axis / np.linalg.norm(q)
    return out


@pytest.mark.parametrize("frame", [abi.GRA_FRAME_CS, abi.GRA_FRAME_HX, abi.GRA_FRAME_CM])
def test_jw_frames_decay_angle(frame):
    # tests/run_JW_frames/run.sh: f2(1270) with |J_z| = 2 in the chosen frame decays as
    # |d^2_{20}|^2 ~ (1 - cos^2)^2 about that frame's axis; without fiducial cuts the
    # decay direction is isotropic in phase space, so unweighted events follow it
    cfg = abi.cfg_default(8)
    cfg.spin_frame = frame
    say(f"case: 20000 unweighted PP[RES]<F> f2(1270) events, frame {frame}")
    with MContinuum(cfg) as P:
        G = MGraniitti(P)
        G.Integrate(vegas_cfg(1 << 18, 4, burnin=3, seed=21 + frame))
        p, f, tot = G.Generate(20000, calls=1 << 20, seed=7 + frame)
    edges = np.linspace(-1, 1, 21)
    F = lambda c: (15.0 / 16.0) * (c - 2 * c ** 3 / 3 + c ** 5 / 5)
    exp_pol = len(p) * np.diff(F(edges))
    res = {}
    for fr in (abi.GRA_FRAME_CS, abi.GRA_FRAME_HX, abi.GRA_FRAME_CM):
        h, _ = np.histogram(frame_cos(p, fr, cfg.beam_E1, cfg.m_beam), edges)
        res[fr] = ((h - exp_pol) ** 2 / exp_pol).sum()
    say(f"frame {frame}: chi2/19 against (1 - c^2)^2 about CS {res[abi.GRA_FRAME_CS]:.1f}, "
        f"HX {res[abi.GRA_FRAME_HX]:.1f}, CM {res[abi.GRA_FRAME_CM]:.1f}")
    assert res[frame] < 19 + 5 * np.sqrt(2 * 19)
    # the helicity axis (the system's direction) differs from the near-beam CS and CM axes
    assert max(v for fr, v in res.items() if fr != frame) > 4 * res[frame]


@pytest.mark.parametrize("variant", ["bw_fixed_gj", "bw_barrier_hx", "bw_spin_cm", "no_spin_corr"])
def test_res_con_variants_pointwise(variant):
    # the Breit-Wigner forms, spin frames and spin switches the ALICECUTS preset leaves unused
    cfg = abi.cfg_default(6)
    bw, frame = {"bw_fixed_gj": (abi.GRA_BW_FIXED, abi.GRA_FRAME_GJ),
                 "bw_barrier_hx": (abi.GRA_BW_BARRIER, abi.GRA_FRAME_HX),
                 "bw_spin_cm": (abi.GRA_BW_SPIN, abi.GRA_FRAME_CM),
                 "no_spin_corr": (abi.GRA_BW_RUNNING, abi.GRA_FRAME_CS)}[variant]
    for r in range(cfg.n_res):
        cfg.res[r].BW = bw
    cfg.spin_frame = frame
    if frame == abi.GRA_FRAME_GJ:
        # the GJ axis is q1 itself, so an f2 with |J_z| = 2 has a production trace of
        # exactly 0 and ProdAmp = sqrt(rounding noise), different in every evaluation
        # order: give the tensors the unpolarised density instead
        for r in range(cfg.n_res):
            if cfg.res[r].J == 2:
                for k in range(25):
                    cfg.res[r].rho_re[k] = 0.2 if k % 6 == 0 else 0.0
    if variant == "no_spin_corr":
        cfg.spin_gen = cfg.spin_dec = 0
    say(f"case: PP[RES+CON] {variant}, 10000 points, seed 46")
    u = points(10000, cfg.dim, 46)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    compare(Wg, Wo, fg, fo, 1e-10, f"RES+CON {variant}")
