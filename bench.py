"""Benchmark of the screened 2->4 continuum inner loop on one MI355X.

Workload (default): pp -> p pi+ pi- p, sqrt(s) = 13 TeV, |eta| < 0.9, pT > 0.15 GeV,
single-channel eikonal screening (16 phi x 31 kt loop nodes, MaxLoopKT 3 GeV,
10001-row eikonal tables), FP64 kinematics and Jacobian, FP32 amplitude with FP32
row sums of the loop nodes added in FP64, 2^20 points per pass.

The run follows the reference's SampleVegas (src/MGraniitti.cc:1074-1135):
1. integration: VEGAS with 5 burn-in iterations (grid adaptation) and 4
   iterations of 2^20 points that fix max f, the largest f = W vw seen while
   integrating (the maxf that SaveEvent's hit-or-miss uses);
2. W untimed generation passes (other iterations' points), which also bring the
   GPU to its steady clock: the first few tens of milliseconds of work run slower;
3. generation, the timed steps: K passes of 2^20 points through the adapted grid,
   each drawing the points (Philox), weighting them (tree + screening kernels),
   accepting U < f / max f (src/MGraniitti.cc:1651-1663) and writing the accepted
   events' four-momenta to host memory.
value = unweighted events produced per second by the timed passes. Because max f
comes from a fixed integration sample, the value does not depend on K.

Extra fields: points/s of the generation passes, time to 0.1% relative sigma error
of plain VEGAS (a separate run, grid adaptation included), the flow leg (the
learned sampler on the same process, with the same max-weight convention: the
largest weight of a 4 x 2^20 integration sample, then produced events/s from its
own generation passes, and its time to 0.1% with training included), the flow on
the screened rho rho tensor-pomeron process, PP[RES+CON] throughput, the k_screen
roofline and the CPU baseline (the FP64 oracle on the host's cores).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from graniitti_amd import abi  # noqa: E402
from graniitti_amd.process import MContinuum, MGraniitti, vegas_cfg  # noqa: E402

# Algorithmic work of the screening loop: counted in the run by the op-counting
# instance of the FP32 node (gra_hip_screen_opcount) over OPCOUNT_POINTS screened
# points of the last timed pass: every operation the node formulation executes,
# rejected nodes included (fma = 2 FLOP; add, sub, mul, log2, exp2, rcp, sqrt,
# sin, cos = 1; compares 0). A gfx950 SIMD issues up to two FP32 VALU
# instructions per quad-cycle (SQ_ACTIVE_INST_VALU2 counts the quad-cycles that
# do), so the bound is the 157.3 TFLOP/s FP32 vector peak, even without packed FP32.
OPCOUNT_POINTS = 65536
VALU_PEAK_TFLOPS = 157.3
# HBM bytes of one screening launch at the default workload, from rocprofv3
# FETCH_SIZE / WRITE_SIZE runs of this command (the counters cannot be read inside
# the run): profiles/k_screen_hbm_traffic_screened_pipi.md
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_screen_hbm_traffic_screened_pipi.md")
MAXF_SAMPLE = 4  # integration iterations (of the pass size) that fix max f / max w
BURNIN = 5  # VEGAS burn-in iterations of every integration
EVENT_CAP = 1 << 17


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    """Host threads of this process's CPU share (the cores it may run on, capped
    by OMP_NUM_THREADS where the host sets it)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(cfg, grid, eff, n_sample):
    """The oracle's FP64 restatement (built with the reference's optimisation
    flags) with the reference's threading model on a bounded sample of the same
    workload and grid: fork/join threads per iteration with one global mutex per
    point (src/MGraniitti.cc:1389-1419) at 1.0x and at the reference's default
    1.5x the hardware threads (include/Graniitti/MGraniitti.h:135-147), and with
    per-thread accumulators. value = the fastest mode's points/s x the GPU run's
    unweighting efficiency."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol

    cfg64 = abi.GraCfg.from_buffer_copy(cfg)
    cfg64.amp_precision = abi.GRA_AMP_FP64
    hw = cpu_threads()
    modes = [("mutex", hw, 1), ("mutex", int(round(1.5 * hw)), 1), ("per_thread", hw, 0)]
    t0 = time.time()
    R = ol.OracleRun(cfg64, hw)
    t_tables = time.time() - t0
    runs = []
    for name, th, mtx in modes:
        t0 = time.time()
        st = R.vegas_iterate(grid, n_sample, 987654321, 0, nthreads=th, use_mutex=mtx)
        dt = time.time() - t0
        runs.append({"mode": name, "threads": th, "points_per_s": n_sample / dt, "seconds": dt,
                     "n_valid": int(st.n_valid)})
        log(f"cpu baseline {name} x{th}: {n_sample / dt:.4g} points/s ({dt:.1f} s)")
    R.close()
    best = max(runs, key=lambda r: r["points_per_s"])
    return {
        "value": best["points_per_s"] * eff, "unit": "unweighted events/s",
        "cores": best["threads"], "kind": "port", "host_cpus": os.cpu_count(),
        "cpu_share": hw, "points_per_s": best["points_per_s"], "modes": runs,
        "sample": f"{n_sample} points of the same workload on the same adapted grid per mode, "
                  f"FP64 oracle built with the reference's flags (-O3 -ftree-vectorize), "
                  f"eikonal tables ({t_tables:.1f} s) not included; value = the fastest mode's "
                  f"points/s x the GPU run's unweighting efficiency",
    }


def vegas_generate(P, G, n, seed, first_iter, passes, maxf):
    """Generation passes on the adapted grid, events into pinned host memory:
    (counters summed over the passes, seconds)"""
    tot = {"trials": 0, "valid": 0, "accepted": 0, "overflow": 0, "stored": 0}
    with abi.PinnedEvents(EVENT_CAP) as ev:
        t0 = time.perf_counter()
        for k in range(passes):
            st = abi.GraGenStats()
            abi.check(P.lib.gra_hip_generate(P.h, C.byref(G.grid), n, seed, first_iter + k, maxf,
                                             ev.buf, EVENT_CAP, C.byref(st)), P.h)
            for key in tot:
                tot[key] += getattr(st, key)
        dt = time.perf_counter() - t0
    return tot, dt


def flow_cfg(dim, layers, bins, hidden, depth, batch, seed=2024):
    fc = abi.GraFlowCfg()
    fc.dim, fc.layers, fc.bins, fc.hidden, fc.depth = dim, layers, bins, hidden, depth
    fc.steps, fc.batch, fc.lr = 0, batch, 1e-3
    fc.beta1, fc.beta2, fc.eps = 0.9, 0.999, 1e-8
    fc.seed, fc.ess_fallback = seed, 0.1
    return fc


def flow_leg(P, grid, fc, steps, n_pass, vegas, gen_passes, plateau_steps=0):
    """The learned sampler on the process of P: an RQS flow composed with the
    adapted VEGAS grid (identity at initialisation), trained by weight-variance
    minimisation (Adam, lr 1e-3) on batches of its own weighted points, then
    - gra_hip_flow_integrate to 0.1% relative sigma error (training included),
      with the VEGAS fallback below an effective sample size of 10%;
    - max w of a 4 x n_pass sample (the same convention as VEGAS's max f);
    - produced unweighted events/s of gen_passes flow generation passes."""
    lib = P.lib
    abi.check(lib.gra_hip_flow_init(P.h, C.byref(fc)), P.h)
    abi.check(lib.gra_hip_flow_set_grid(P.h, C.byref(grid)), P.h)
    s0 = abi.GraFlowStats()
    abi.check(lib.gra_hip_flow_sample(P.h, n_pass, 11, None, None, C.byref(s0)), P.h)
    rc = abi.GraFlowRunCfg()
    rc.train_steps, rc.n_max, rc.seed, rc.precision, rc.ess_min = steps, 64 * n_pass, 1000, 1e-3, 0.1
    rc.vegas = vegas_cfg(n_pass, 1, burnin=5, seed=31337, precision=1e-3, max_iter=60)
    g_fb = abi.GraGrid()
    res = abi.GraFlowResult()
    abi.check(lib.gra_hip_flow_integrate(P.h, C.byref(rc), C.byref(g_fb), C.byref(res)), P.h)
    log(f"flow: trained {steps} steps x {fc.batch} in {res.train_seconds:.3f} s; sampled "
        f"{res.n} points in {res.sample_seconds:.3f} s; sigma {res.sigma:.6e} +- "
        f"{res.sigma_err:.2e}, fallback {res.fallback}")
    sm = abi.GraFlowStats()
    abi.check(lib.gra_hip_flow_sample(P.h, MAXF_SAMPLE * n_pass, 2000, None, None, C.byref(sm)),
              P.h)
    maxw = sm.maxw
    tot = {"trials": 0, "valid": 0, "accepted": 0, "overflow": 0, "stored": 0}
    with abi.PinnedEvents(EVENT_CAP) as ev:
        t0 = time.perf_counter()
        for k in range(gen_passes):
            st = abi.GraGenStats()
            abi.check(lib.gra_hip_flow_generate(P.h, n_pass, 3000, k, maxw, ev.buf, EVENT_CAP,
                                                C.byref(st)), P.h)
            for key in tot:
                tot[key] += getattr(st, key)
        t_gen = time.perf_counter() - t0
    comb = np.hypot(res.sigma_err, vegas["sigma_err"])
    out = {
        "layers": fc.layers, "bins": fc.bins, "hidden": f"{fc.depth}x{fc.hidden}",
        "mlp": "bf16 mfma, fp32 accumulate", "train_steps": steps, "train_batch": int(fc.batch),
        "train_s": res.train_seconds, "sample_s_to_0p1pct": res.sample_seconds,
        "time_to_0p1pct_s": res.seconds_to_precision if res.reached else None,
        "reached_0p1pct": bool(res.reached), "points_to_0p1pct": int(res.n),
        "vegas_fallback": bool(res.fallback),
        "sigma_barn": res.sigma, "sigma_err_barn": res.sigma_err,
        "sigma_pull_vs_vegas": (res.sigma - vegas["sigma"]) / comb,
        "consistent_3sigma": bool(abs(res.sigma - vegas["sigma"]) < 3 * comb),
        "cv2": res.cv2, "cv2_vegas_grid": s0.cv2,
        "variance_reduction": s0.cv2 / res.cv2 if res.cv2 > 0 else None,
        "ess": res.ess, "ess_vegas_grid": s0.ess,
        "maxw_sample_points": MAXF_SAMPLE * n_pass,
        "generation": {"passes": gen_passes, "points_per_pass": n_pass, **tot,
                       "seconds": t_gen,
                       "unweighting_efficiency": tot["accepted"] / tot["trials"],
                       "points_per_s": tot["trials"] / t_gen,
                       "unweighted_events_per_s": tot["stored"] / t_gen},
        "time_ratio_vs_vegas": (res.seconds_to_precision / vegas["time_to_0p1pct_s"]
                                if res.reached and vegas["time_to_0p1pct_s"] else None),
    }
    if plateau_steps:
        # the same flow from a fresh initialisation, its training stopped by the CV^2
        # plateau test (windows of 50 steps, less than 5% improvement per window)
        abi.check(lib.gra_hip_flow_init(P.h, C.byref(fc)), P.h)
        abi.check(lib.gra_hip_flow_set_grid(P.h, C.byref(grid)), P.h)
        rc.train_steps, rc.check_every, rc.plateau = plateau_steps, 50, 0.05
        rp = abi.GraFlowResult()
        abi.check(lib.gra_hip_flow_integrate(P.h, C.byref(rc), C.byref(g_fb), C.byref(rp)), P.h)
        out["plateau_run"] = {
            "max_steps": plateau_steps, "window": 50, "min_improvement": 0.05,
            "train_steps_run": int(rp.train_steps_run), "train_s": rp.train_seconds,
            "last_window_cv2": rp.train_cv2, "cv2": rp.cv2, "ess": rp.ess,
            "time_to_0p1pct_s": rp.seconds_to_precision if rp.reached else None,
            "vegas_fallback": bool(rp.fallback),
            "time_ratio_vs_vegas": (rp.seconds_to_precision / vegas["time_to_0p1pct_s"]
                                    if rp.reached and vegas["time_to_0p1pct_s"] else None),
        }
    log(f"flow: {out}")
    return out


def vegas_precision_run(P, n, burnin, seed):
    """Plain VEGAS to 0.1% relative sigma error, grid adaptation included"""
    G = MGraniitti(P)
    vc = vegas_cfg(n, 1, burnin=burnin, seed=seed, precision=1e-3, max_iter=60)
    res = G.Integrate(vc)
    t = res.seconds_to_precision if res.seconds_to_precision > 0 else None
    out = {"sigma": res.sigma, "sigma_err": res.sigma_err, "iterations": res.iterations,
           "restarts": res.restarts, "seconds_total": res.seconds_total,
           "calls_total": int(res.calls_total), "time_to_0p1pct_s": t,
           "integration_points_per_s": res.calls_total / res.seconds_total}
    log(f"VEGAS precision run: {out}")
    return out


WORKLOADS = {
    1: "pp->p pi+ pi- p 13 TeV, Regge continuum, unscreened, FP64",
    2: "pp->p pi+ pi- p 13 TeV, Regge continuum, screened (16x31 loop nodes), FP32 amplitude",
    3: "pp->p rho0 rho0 p -> p 2(pi+ pi-) p 13 TeV, tensor pomeron 2xV, screened (16x31 loop "
       "nodes), FP32 amplitude",
    6: "pp->p pi+ pi- p 13 TeV, Regge continuum + 8 resonances (PP[RES+CON], ALICECUTS), "
       "unscreened, FP64",
    7: "pp->p pi+ pi- p 13 TeV, Regge continuum + 8 resonances (PP[RES+CON], ALICECUTS), "
       "screened (16x31 loop nodes), FP64",
}

COPY_BYTES = 256 << 20


def res_leg():
    """PP[RES+CON] -> pi+ pi- (continuum + the eight ALICECUTS resonances): VEGAS
    integration throughput unscreened (2^20 points/iteration, FP64) and screened
    (2^18), with the loop nodes in FP64 and in the FP32-amplitude mode."""
    out = {}
    for name, preset, n, warm, steps, prec in (
            ("unscreened", 6, 1 << 20, 2, 5, abi.GRA_AMP_FP64),
            ("screened", 7, 1 << 18, 1, 3, abi.GRA_AMP_FP64),
            ("screened_fp32_nodes", 7, 1 << 18, 1, 3, abi.GRA_AMP_FP32)):
        cfg = abi.cfg_default(preset)
        cfg.amp_precision = prec
        with MContinuum(cfg) as P:
            grid = abi.GraGrid()
            abi.check(P.lib.gra_hip_grid_init(C.byref(grid), cfg.dim, 128))
            it, acc = 0, [0.0, 0.0]
            for k in range(warm + steps):
                if k == warm:
                    t0 = time.perf_counter()
                st = P.vegas_iterate(grid, n, 4242, it)
                it += 1
                abi.check(P.lib.gra_hip_grid_optimize(C.byref(grid), C.byref(st), 1.5))
                if k >= warm:
                    I = st.fsum / n
                    v = max((st.f2sum / n - I * I) / n, 1e-300)
                    acc[0] += I / v
                    acc[1] += 1.0 / v
            dt = time.perf_counter() - t0
        pps = n * steps / dt
        out[name] = {
            "preset": preset, "amplitude": "fp32 loop nodes" if prec == abi.GRA_AMP_FP32 else "fp64",
            "points_per_step": n, "steps": steps, "points_per_s": pps,
            "sigma_barn": acc[0] / acc[1], "sigma_err_barn": acc[1] ** -0.5}
        log(f"PP[RES+CON] {name}: {pps:.3e} points/s, sigma {acc[0] / acc[1]:.5e} barn")
    return out


def rhorho_leg(args):
    """The screened rho rho tensor-pomeron process (preset 3, FP32 amplitude):
    VEGAS to 0.1% at 2^21 points per iteration, and the flow of 10 layers, K = 32,
    3 x 256 against it."""
    cfg = abi.cfg_default(3)
    n = 1 << 21
    with MContinuum(cfg) as P:
        v = vegas_precision_run(P, n, 5, 1717)
        G = MGraniitti(P)
        G.Integrate(vegas_cfg(n, MAXF_SAMPLE, burnin=5, seed=1818))
        fc = flow_cfg(cfg.dim, 10, 32, 256, 3, 1 << 18)
        fl = flow_leg(P, G.grid, fc, args.flow5_steps, 1 << 20, v, 2, args.flow5_steps)
    return {"preset": 3, "vegas": v, "flow": fl}


def config_mode_leg(N, burnin, base):
    """The workload in its stated loop discretisation (preset 11: 32 Gauss-Legendre
    kt nodes x 16 azimuths, A_eik on 64 log-spaced kt nodes) against the reference's
    (Simpson 3/8 on 16 x 31 nodes, 10001-row A_eik table): VEGAS to 0.1%."""
    with MContinuum(abi.cfg_default(11)) as P:
        v = vegas_precision_run(P, N, burnin, 424242)
    comb = np.hypot(v["sigma_err"], base["sigma_err"])
    return {"preset": 11, "sigma_barn": v["sigma"], "sigma_err_barn": v["sigma_err"],
            "time_to_0p1pct_s": v["time_to_0p1pct_s"],
            "delta_sigma_rel": v["sigma"] / base["sigma"] - 1.0,
            "delta_sigma_pull": (v["sigma"] - base["sigma"]) / comb}


def calibrate_copy(reps=3):
    """Device-to-device copies of COPY_BYTES (the runtime's blit kernel): the
    reference dispatches that scale the FETCH_SIZE / WRITE_SIZE counters of a
    profiled run to bytes."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(COPY_BYTES)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(COPY_BYTES), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def screen_traffic():
    """HBM bytes per screening launch measured with FETCH_SIZE / WRITE_SIZE, or None."""
    if not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=30,
                    help="untimed generation passes before the timed ones")
    ap.add_argument("--points", type=int, default=1 << 20)
    ap.add_argument("--preset", type=int, default=2)
    ap.add_argument("--no-res", action="store_true",
                    help="skip the PP[RES+CON] resonance throughput leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision-run", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 19)
    ap.add_argument("--no-flow", action="store_true")
    ap.add_argument("--no-flow5", action="store_true",
                    help="skip the flow on the screened rho rho process")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run 256 MiB device copies (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--flow-steps", type=int, default=500)
    ap.add_argument("--flow5-steps", type=int, default=200)
    ap.add_argument("--flow-layers", type=int, default=6)
    ap.add_argument("--flow-bins", type=int, default=16)
    ap.add_argument("--flow-hidden", type=int, default=128)
    ap.add_argument("--flow-depth", type=int, default=2)
    ap.add_argument("--flow-batch", type=int, default=1 << 18)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: every box here has one MI355X; "
                          "the multi-GPU path is a design note only (bench.py --gpus 1)"}))
        sys.exit(2)

    if args.calibrate_copy:
        calibrate_copy()
    cfg = abi.cfg_default(args.preset)
    N = args.points
    P = MContinuum(cfg)
    G = MGraniitti(P)
    seed = 12345678
    log(f"workload preset {args.preset}: {N} points/pass, burn-in {BURNIN}, "
        f"generation passes {args.warmup} untimed + {args.steps} timed")
    # integration: burn-in adapts the grid, MAXF_SAMPLE iterations fix max f
    integ = G.Integrate(vegas_cfg(N, MAXF_SAMPLE, burnin=BURNIN, seed=seed))
    maxf = integ.maxf
    # untimed warm-up passes (iterations 900.. ; the timed passes draw 1000..)
    if args.warmup > 0:
        vegas_generate(P, G, N, seed, 900, args.warmup, maxf)
    for s in range(4):
        P.kernel_time(s, reset=True)
    # timed steps: generation passes
    tot, dt = vegas_generate(P, G, N, seed, 1000, args.steps, maxf)
    ms_tree, n_tree = P.kernel_time(0)
    ms_scr, n_scr = P.kernel_time(1)
    pts_per_s = tot["trials"] / dt
    eff = tot["accepted"] / tot["trials"]
    value = tot["stored"] / dt
    if tot["stored"] != tot["accepted"]:
        log(f"warning: {tot['accepted'] - tot['stored']} accepted events beyond the buffer")
    log(f"timed: {dt:.3f} s, {pts_per_s:.3e} points/s, efficiency {eff:.5f}, "
        f"{value:.4e} events/s; integration sigma {integ.sigma:.5e} +- {integ.sigma_err:.2e} b")

    roofline = None
    # per-node operation counts are those of the FP32 Regge continuum node
    if (cfg.screening and n_scr > 0 and cfg.process == abi.GRA_PROC_REGGE_CON
            and cfg.amp_precision == abi.GRA_AMP_FP32):
        nodes = (cfg.n_loop_phi + 1) * (cfg.n_loop_kt + 1)
        cand_per_launch = tot["valid"] / n_scr
        avg_ms = ms_scr / n_scr
        oc = (C.c_double * 12)()
        abi.check(P.lib.gra_hip_screen_opcount(P.h, OPCOUNT_POINTS, oc), P.h)
        flop_pt = (oc[3] + oc[4] + 2 * oc[5] + oc[6]) / oc[0]
        achieved = flop_pt * cand_per_launch / (avg_ms * 1e-3) / 1e12
        roofline = {
            "bound": "valu", "kernel": "k_screen (screening loop)",
            "achieved": achieved, "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / VALU_PEAK_TFLOPS, "traffic": screen_traffic(),
            "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT),
            "avg_launch_ms": avg_ms, "launches": n_scr,
            "flop_per_point": flop_pt, "flop_per_node": flop_pt / nodes,
            "flop_source": f"op-counting run of the node code over {int(oc[0])} screened "
                           f"points of the last pass (gra_hip_screen_opcount)",
            "accepted_node_fraction": oc[2] / oc[1],
            "flop_per_accepted_node": oc[8] + oc[9] + 2 * oc[10] + oc[11],
            "accepted_node_ops": {"add": oc[8], "mul": oc[9], "fma": oc[10], "trans": oc[11]},
            "node_dtype": "f32", "nodes_per_point": nodes,
            "screened_points_per_launch": cand_per_launch,
        }
        # the FP32 rate this GPU sustains without packed FP32 (which this build does not
        # use): independent FMA chains, measured now on the same device
        ceil = C.c_double()
        abi.check(P.lib.gra_hip_fp32_fma_rate(C.byref(ceil)))
        roofline["attainable_fp32_no_packed_tflops"] = ceil.value
        roofline["frac_of_attainable"] = achieved / ceil.value
        log(f"k_screen {achieved:.1f} TFLOP/s: {achieved / VALU_PEAK_TFLOPS:.3f} of the published "
            f"peak, {achieved / ceil.value:.3f} of the {ceil.value:.1f} TFLOP/s FMA-chain rate "
            f"without packed FP32")
        log(f"op count: {int(oc[0])} points, accepted nodes {oc[2] / oc[1]:.4f}, "
            f"{flop_pt / nodes:.2f} FLOP per node, one accepted node {oc[8]:.0f} add "
            f"{oc[9]:.0f} mul {oc[10]:.0f} fma {oc[11]:.0f} trans")
    vegas = {"sigma": integ.sigma, "sigma_err": integ.sigma_err, "time_to_0p1pct_s": None}
    prec_res = None
    if not args.no_precision_run:
        prec_res = vegas_precision_run(P, N, BURNIN, 424242)
        vegas = prec_res
    cmode = None
    if prec_res is not None and args.preset == 2:
        cmode = config_mode_leg(N, BURNIN, prec_res)
        log(f"config mode: {cmode}")
    res = None
    if not args.no_res and args.preset == 2:
        try:
            res = res_leg()
        except Exception as e:  # reported, never fatal to the headline line
            res = {"error": str(e)}
    flow = None
    if not args.no_flow:
        fc = flow_cfg(cfg.dim, args.flow_layers, args.flow_bins, args.flow_hidden,
                      args.flow_depth, args.flow_batch)
        flow = flow_leg(P, G.grid, fc, args.flow_steps, N, vegas, 4, args.flow_steps)
    flow5 = None
    if not args.no_flow and not args.no_flow5 and args.preset == 2:
        try:
            flow5 = rhorho_leg(args)
        except Exception as e:
            flow5 = {"error": str(e)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, G.grid, eff, args.cpu_sample)
        log(f"cpu baseline: {cpu}")
    P.close()
    line = {
        "metric": "time to 0.1% sigma rel. error & unweighted events/s, flow vs VEGAS, 1 MI355X",
        "value": value, "unit": "unweighted events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": WORKLOADS.get(args.preset, f"preset {args.preset}") +
                   f", VEGAS generation passes of {N} points",
                   "preset": args.preset, "points_per_step": N, "integrator": "VEGAS"},
        "points_per_s": pts_per_s, "unweighting_efficiency": eff,
        "generation": {**tot, "maxf": maxf,
                       "maxf_sample": f"{MAXF_SAMPLE} x {N} integration points after "
                                      f"{BURNIN} burn-in iterations"},
        "time_to_0p1pct_s": vegas.get("time_to_0p1pct_s"), "precision_run": prec_res,
        "sigma_barn": integ.sigma, "sigma_err_barn": integ.sigma_err,
        "kernel_ms": {"tree": ms_tree, "screen": ms_scr},
        "config_mode": cmode, "flow": flow, "flow_rhorho": flow5, "resonances": res,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
