"""The flow sampler in the reference's NEURO integrator slot (MGraniitti::SampleNeuro,
src/MGraniitti.cc:1527-1648) on the GPU:

- per-point output of gra_hip_flow_sample: the weights w = f vw / q and points x,
  whose integrand the process evaluates identically;
- unweighted events from the flow (gra_hip_flow_generate, SaveEvent's hit-or-miss,
  src/MGraniitti.cc:1651-1726): the accepted set equals a host restatement of the
  test U < w / maxw on the same weights with the oracle's Philox, and the events
  conserve four-momentum;
- the VEGAS fallback when the effective sample size is too small;
- the flow of the screened pi+ pi- process at the stated size: 6 layers, K = 16,
  2 x 128 BF16, batches of 2^18, 500 Adam steps at lr 1e-3 on the full eikonal
  tables, against plain VEGAS (sigma within 3 combined standard errors, CV^2 and
  unweighting efficiency)."""
import ctypes as C
import sys
import time

import numpy as np
import pytest

from graniitti_amd import abi
from graniitti_amd.process import MContinuum, MGraniitti, vegas_cfg
import oracle_lib as ol
from conftest import small_tables

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def flow_cfg(D, L, K, H, depth, batch, seed=7, lr=1e-3, ess=0.1):
    fc = abi.GraFlowCfg()
    fc.dim, fc.layers, fc.bins, fc.hidden, fc.depth = D, L, K, H, depth
    fc.steps, fc.batch, fc.lr = 0, batch, lr
    fc.beta1, fc.beta2, fc.eps = 0.9, 0.999, 1e-8
    fc.seed, fc.ess_fallback = seed, ess
    return fc


@pytest.fixture(scope="module")
def trained():
    """Screened pi+ pi- (1000-row tables): VEGAS grid, then a flow trained 30 steps"""
    cfg = small_tables(abi.cfg_default(2), 1000)
    P = MContinuum(cfg)
    G = MGraniitti(P)
    res = G.Integrate(vegas_cfg(1 << 18, 3, burnin=4, seed=41))
    fc = flow_cfg(8, 6, 16, 128, 2, 16384)
    abi.check(P.lib.gra_hip_flow_init(P.h, C.byref(fc)), P.h)
    abi.check(P.lib.gra_hip_flow_set_grid(P.h, C.byref(G.grid)), P.h)
    st = abi.GraFlowStats()
    abi.check(P.lib.gra_hip_flow_train(P.h, 30, C.byref(st)), P.h)
    yield P, G, res
    P.close()


def flow_sample(P, n, seed):
    w = np.empty(n)
    x = np.empty((n, P.cfg.dim))
    st = abi.GraFlowStats()
    abi.check(P.lib.gra_hip_flow_sample(P.h, n, seed, abi.dptr(w), abi.dptr(x), C.byref(st)), P.h)
    return w, x, st


def test_flow_sample_per_point_output(trained):
    P, G, res = trained
    n = 70000  # several chunks, the last one partial
    say(f"case: flow per-point output, n = {n}")
    w, x, st = flow_sample(P, n, 5)
    assert np.all((x >= 0) & (x < 1)) and np.all(w >= 0)
    assert st.n == n
    assert abs(w.mean() - st.sigma) <= 1e-12 * st.sigma
    assert abs(w.max() - st.maxw) == 0.0
    # the integrand at the returned points: w > 0 exactly where W > 0
    W, a2, fl = P.eval(x)
    assert np.array_equal(w > 0, W > 0)
    # same seed, same points and weights
    w2, x2, _ = flow_sample(P, n, 5)
    assert np.array_equal(w, w2) and np.array_equal(x, x2)


def flow_generate(P, n, seed, pas, maxw, cap):
    buf = (abi.GraEvent * cap)()
    st = abi.GraGenStats()
    abi.check(P.lib.gra_hip_flow_generate(P.h, n, seed, pas, maxw, buf, cap, C.byref(st)), P.h)
    a = np.ctypeslib.as_array(buf)[: st.stored]
    return a, st


def test_flow_generation_matches_host_restatement(trained):
    P, G, res = trained
    n, seed = 300000, 17
    say(f"case: flow hit-or-miss, n = {n}, seed {seed}, pass 0")
    w, x, st = flow_sample(P, n, seed)
    maxw = float(np.quantile(w, 0.995))  # a low maxw, so some points overflow
    ev, gs = flow_generate(P, n, seed, 0, maxw, n)
    W, a2, fl = P.eval(x)
    valid = (fl & 7) == 7
    # the flow's acceptance stream: pass 0 of the flow generation (0xD0000000)
    U = np.array([ol.point_uniforms(seed, 0xD0000000, i, 1)[0] for i in range(n)])
    over = w > maxw
    acc = valid & (U < w / maxw) & ~over
    say(f"flow generation: valid {valid.sum()} accepted {acc.sum()} overflow {over.sum()} "
        f"(GPU {gs.valid} {gs.accepted} {gs.overflow})")
    assert gs.trials == n and gs.valid == valid.sum()
    assert gs.accepted == acc.sum() and gs.overflow == over.sum()
    assert gs.stored == gs.accepted
    idx = np.array(ev["index"], dtype=np.int64)
    np.testing.assert_array_equal(idx, np.nonzero(acc)[0])
    np.testing.assert_array_equal(np.array(ev["f"]), w[idx])
    # kinematics of the events: the oracle's at the same points, conserved
    R = ol.OracleRun(P.cfg)
    p = np.array(ev["p"])[:, :4, :]
    beam = np.array([P.cfg.beam_E1 + P.cfg.beam_E2, 0, 0, 0])
    for e in range(0, len(idx), max(1, len(idx) // 200)):
        ok, pf, legs, sc = R.kinematics(x[idx[e]])
        assert ok
        np.testing.assert_allclose(p[e], pf[1:5], rtol=1e-10, atol=1e-12)
        assert np.abs(p[e].sum(axis=0) - beam).max() < 1e-6


def test_flow_integrate_falls_back_to_vegas(trained):
    P, G, res = trained
    say("case: flow_integrate with ESS threshold 0.99 (fallback) and 0 (no fallback)")
    rc = abi.GraFlowRunCfg()
    rc.train_steps, rc.n_max, rc.seed, rc.precision = 0, 1 << 20, 3, 0.0
    rc.vegas = vegas_cfg(1 << 17, 3, burnin=2, seed=55)
    rc.ess_min = 0.99
    g = abi.GraGrid()
    out = abi.GraFlowResult()
    abi.check(P.lib.gra_hip_flow_integrate(P.h, C.byref(rc), C.byref(g), C.byref(out)), P.h)
    assert out.fallback == 1 and out.ess < 0.99
    assert out.n < (1 << 20)  # the first chunk's ESS already triggers it
    # the result is that of the same VEGAS run made directly
    g2 = abi.GraGrid()
    v = abi.GraVegasResult()
    abi.check(P.lib.gra_hip_vegas_run(P.h, C.byref(rc.vegas), C.byref(g2), C.byref(v)), P.h)
    assert out.sigma == v.sigma and out.sigma_err == v.sigma_err
    assert out.vegas.sigma == v.sigma
    rc.ess_min = 0.0
    out2 = abi.GraFlowResult()
    abi.check(P.lib.gra_hip_flow_integrate(P.h, C.byref(rc), C.byref(g), C.byref(out2)), P.h)
    assert out2.fallback == 0 and out2.n == (1 << 20)
    comb = np.hypot(out2.sigma_err, v.sigma_err)
    say(f"flow sigma {out2.sigma:.6e} +- {out2.sigma_err:.2e}, VEGAS {v.sigma:.6e} "
        f"+- {v.sigma_err:.2e}, pull {(out2.sigma - v.sigma) / comb:+.2f}")
    assert abs(out2.sigma - v.sigma) < 3 * comb


def test_flow_integrate_stops_training_on_a_plateau(trained):
    # windows of 5 steps: with a 90% improvement required per window the plateau
    # test stops after the second window; with the test off every step is taken
    P, G, res = trained
    say("case: flow_integrate training with the CV^2 plateau test")
    rc = abi.GraFlowRunCfg()
    rc.train_steps, rc.n_max, rc.seed, rc.precision, rc.ess_min = 40, 1 << 18, 3, 0.0, 0.0
    rc.check_every, rc.plateau = 5, 0.9
    rc.vegas = vegas_cfg(1 << 17, 3, burnin=2, seed=55)
    g = abi.GraGrid()
    out = abi.GraFlowResult()
    abi.check(P.lib.gra_hip_flow_integrate(P.h, C.byref(rc), C.byref(g), C.byref(out)), P.h)
    say(f"plateau 0.9: {out.train_steps_run} steps, window CV^2 {out.train_cv2:.3f}")
    assert out.train_steps_run == 10 and out.train_cv2 > 0
    rc.plateau = 0.0
    out2 = abi.GraFlowResult()
    abi.check(P.lib.gra_hip_flow_integrate(P.h, C.byref(rc), C.byref(g), C.byref(out2)), P.h)
    assert out2.train_steps_run == 40 and out2.fallback == 0


def test_flow_on_the_screened_process_at_the_stated_size():
    # flow 6 layers, K = 16, MLP 2 x 128 (BF16 MFMA, FP32 accumulation), Adam at
    # lr 1e-3 for 500 steps on batches of 2^18 points weighted by the screened
    # pi+ pi- integrand (FP64 event weight), full eikonal tables
    cfg = abi.cfg_default(2)
    say("case: flow at the stated size on preset 2: 500 steps x 2^18, then 2^21 points")
    with MContinuum(cfg) as P:
        G = MGraniitti(P)
        v = G.Integrate(vegas_cfg(1 << 20, 10, burnin=5, seed=12345678))
        fc = flow_cfg(8, 6, 16, 128, 2, 1 << 18, seed=2024, lr=1e-3)
        abi.check(P.lib.gra_hip_flow_init(P.h, C.byref(fc)), P.h)
        abi.check(P.lib.gra_hip_flow_set_grid(P.h, C.byref(G.grid)), P.h)
        s0 = abi.GraFlowStats()
        abi.check(P.lib.gra_hip_flow_sample(P.h, 1 << 21, 11, None, None, C.byref(s0)), P.h)
        t0 = time.perf_counter()
        st = abi.GraFlowStats()
        abi.check(P.lib.gra_hip_flow_train(P.h, 500, C.byref(st)), P.h)
        t_train = time.perf_counter() - t0
        w, x, s1 = flow_sample(P, 1 << 21, 12)
    comb = np.hypot(v.sigma_err, s1.sigma_err)
    pull = (s1.sigma - v.sigma) / comb
    say(f"VEGAS sigma {v.sigma:.6e} +- {v.sigma_err:.2e}; flow sigma {s1.sigma:.6e} +- "
        f"{s1.sigma_err:.2e} (pull {pull:+.2f}); CV^2 {s0.cv2:.3f} -> {s1.cv2:.3f}; ESS "
        f"{s0.ess:.3f} -> {s1.ess:.3f}; training {t_train:.2f} s")
    assert s1.cv2 < 0.6 * s0.cv2
    assert abs(pull) < 3
    assert s1.fallback == 0
