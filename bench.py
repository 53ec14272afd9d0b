"""Benchmark of the screened 2->4 continuum inner loop on one MI355X.

Workload (default): pp -> p pi+ pi- p, sqrt(s) = 13 TeV, |eta| < 0.9, pT > 0.15 GeV,
single-channel eikonal screening (16 phi x 31 kt loop nodes, MaxLoopKT 3 GeV,
10001-node eikonal tables), FP64 kinematics and Jacobian, FP32 amplitude with
compensated FP32 loop sums, VEGAS with 2^20 points per iteration.

A step is one VEGAS iteration: 2^20 points drawn (Philox), mapped through the
grid, weighted (tree kernel + screening kernel), accumulated, and the grid
updated. W untimed warm-up iterations adapt the grid; K iterations are timed.

value = unweighted events/s = points/s x unweighting efficiency <f>/max f of
the timed iterations (MGraniitti::SaveEvent hit-or-miss, src/MGraniitti.cc:1651-1663).
Extra fields: points/s, time to 0.1% relative sigma error (a separate VEGAS run
to that precision, grid adaptation included), sigma.

"flow": the learned sampler on the same process (6 RQS coupling layers, 16 bins,
2 x 128 hidden in BF16 MFMA), composed with the adapted grid, trained 500 Adam
steps on batches of 2^18 of its own weighted points and then sampled to 0.1%
relative sigma error; its time to precision is given with and without training,
with sigma consistency against VEGAS, variance reduction and unweighting
efficiency against the grid it starts from.
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

# Algorithmic work of one screening-loop node of the FP32 amplitude kernel
# (kinematics as changes from the tree point, invariants, Regge amplitude, row
# sum), counted on the source: add / mul = 1, fma = 2, each transcendental or
# reciprocal = 1. A gfx950 SIMD issues up to two FP32 VALU instructions per
# quad-cycle (the SQ_ACTIVE_INST_VALU2 counter counts the quad-cycles that do),
# so the bound is the 157.3 TFLOP/s FP32 vector peak, even without packed FP32.
FLOP_PER_NODE = 147
VALU_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(cfg, grid, eff, n_sample, threads):
    """The oracle's multithreaded FP64 restatement on a bounded sample of the
    same workload (same grid), per-thread accumulators."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol

    cfg64 = abi.GraCfg.from_buffer_copy(cfg)
    cfg64.amp_precision = abi.GRA_AMP_FP64
    t0 = time.time()
    R = ol.OracleRun(cfg64, threads)
    t_tables = time.time() - t0
    t0 = time.time()
    st = R.vegas_iterate(grid, n_sample, 987654321, 0, nthreads=threads)
    dt = time.time() - t0
    R.close()
    pts = n_sample / dt
    return {
        "value": pts * eff, "unit": "unweighted events/s", "cores": threads, "kind": "port",
        "points_per_s": pts,
        "sample": f"{n_sample} points of the same workload on the same adapted grid, "
                  f"FP64 oracle, {threads} threads ({dt:.1f} s; eikonal tables {t_tables:.1f} s "
                  f"not included); value = points/s x the GPU run's unweighting efficiency",
        "n_valid": int(st.n_valid),
    }


def flow_stats_sums(st):
    """(sum w, sum w^2, max w, n) of one flow_sample call, from its moments."""
    n = float(st.n)
    var = st.sigma_err ** 2 * n
    return st.sigma * n, (var + st.sigma ** 2) * n, st.maxw, n


def flow_leg(P, grid, args, vegas_sigma, vegas_err):
    """Learned sampler on the same process: RQS flow (by default 6 coupling
    layers, 16 bins, 2 x 128 hidden, BF16 MFMA) composed with the adapted VEGAS
    grid, trained by weight-variance minimisation (Adam, lr 1e-3) on batches of
    its own weighted points (2^18 by default), then sampled in chunks of 2^20
    until the relative sigma error reaches 0.1%."""
    lib = P.lib
    fc = abi.GraFlowCfg()
    fc.dim, fc.layers, fc.bins = P.cfg.dim, args.flow_layers, args.flow_bins
    fc.hidden, fc.depth = args.flow_hidden, args.flow_depth
    fc.steps, fc.batch, fc.lr = args.flow_steps, args.flow_batch, 1e-3
    fc.beta1, fc.beta2, fc.eps = 0.9, 0.999, 1e-8
    fc.seed, fc.ess_fallback = 2024, 0.1
    abi.check(lib.gra_hip_flow_init(P.h, C.byref(fc)), P.h)
    abi.check(lib.gra_hip_flow_set_grid(P.h, C.byref(grid)), P.h)
    chunk = 1 << 20
    # the untrained flow is the identity: it samples exactly the frozen VEGAS grid
    s0 = abi.GraFlowStats()
    abi.check(lib.gra_hip_flow_sample(P.h, chunk, 11, C.byref(s0)), P.h)
    st = abi.GraFlowStats()
    t0 = time.perf_counter()
    abi.check(lib.gra_hip_flow_train(P.h, args.flow_steps, C.byref(st)), P.h)
    t_train = time.perf_counter() - t0
    log(f"flow: trained {args.flow_steps} steps x {args.flow_batch} in {t_train:.3f} s; "
        f"last batch cv2 {st.cv2:.3f}")
    s1 = s2 = mx = n = 0.0
    t0 = time.perf_counter()
    calls = 0
    while True:
        one = abi.GraFlowStats()
        abi.check(lib.gra_hip_flow_sample(P.h, chunk, 1000 + calls, C.byref(one)), P.h)
        calls += 1
        a, b, m, k = flow_stats_sums(one)
        s1, s2, mx, n = s1 + a, s2 + b, max(mx, m), n + k
        mean = s1 / n
        err = np.sqrt(max(s2 / n - mean * mean, 0.0) / n)
        if err <= 1e-3 * mean or calls >= 64:
            break
    t_sample = time.perf_counter() - t0
    var = max(s2 / n - mean * mean, 0.0)
    cv2 = var / mean ** 2
    eff = mean / mx
    pts = n / t_sample
    comb = np.hypot(err, vegas_err)
    res = {
        "layers": args.flow_layers, "bins": args.flow_bins,
        "hidden": f"{args.flow_depth}x{args.flow_hidden}", "mlp": "bf16 mfma, fp32 accumulate",
        "train_steps": args.flow_steps, "train_batch": args.flow_batch, "train_s": t_train,
        "sample_s_to_0p1pct": t_sample, "time_to_0p1pct_s": t_train + t_sample,
        "points_to_0p1pct": int(n), "reached_0p1pct": bool(err <= 1e-3 * mean),
        "sigma_barn": mean, "sigma_err_barn": err,
        "sigma_pull_vs_vegas": (mean - vegas_sigma) / comb,
        "consistent_3sigma": bool(abs(mean - vegas_sigma) < 3 * comb),
        "cv2": cv2, "cv2_vegas_grid": s0.cv2, "variance_reduction": s0.cv2 / cv2 if cv2 > 0 else None,
        "ess": 1.0 / (1.0 + cv2), "ess_vegas_grid": s0.ess,
        "unweighting_efficiency": eff, "unweighting_efficiency_vegas_grid": s0.sigma / s0.maxw,
        "points_per_s": pts, "unweighted_events_per_s": pts * eff,
        "vegas_fallback": bool(1.0 / (1.0 + cv2) < fc.ess_fallback),
    }
    log(f"flow: {res}")
    return res


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
    """PP[RES+CON] -> pi+ pi- (continuum + the eight ALICECUTS resonances, FP64 amplitude):
    VEGAS throughput unscreened (2^20 points/iteration) and screened (2^18)."""
    out = {}
    for preset, n, warm, steps in ((6, 1 << 20, 2, 5), (7, 1 << 18, 1, 3)):
        cfg = abi.cfg_default(preset)
        with MContinuum(cfg) as P:
            grid = abi.GraGrid()
            abi.check(P.lib.gra_hip_grid_init(C.byref(grid), cfg.dim, 128))
            it, fs, mx, acc = 0, 0.0, 0.0, [0.0, 0.0]
            for k in range(warm + steps):
                if k == warm:
                    t0 = time.perf_counter()
                st = P.vegas_iterate(grid, n, 4242, it)
                it += 1
                abi.check(P.lib.gra_hip_grid_optimize(C.byref(grid), C.byref(st), 1.5))
                if k >= warm:
                    fs += st.fsum
                    mx = max(mx, st.maxf)
                    I = st.fsum / n
                    v = max((st.f2sum / n - I * I) / n, 1e-300)
                    acc[0] += I / v
                    acc[1] += 1.0 / v
            dt = time.perf_counter() - t0
        pps = n * steps / dt
        eff = fs / (n * steps) / mx if mx > 0 else 0.0
        out["screened" if preset == 7 else "unscreened"] = {
            "preset": preset, "points_per_step": n, "steps": steps, "points_per_s": pps,
            "unweighting_efficiency": eff, "unweighted_events_per_s": pps * eff,
            "sigma_barn": acc[0] / acc[1], "sigma_err_barn": acc[1] ** -0.5}
        log(f"PP[RES+CON] preset {preset}: {pps:.3e} points/s, eff {eff:.4f}, "
            f"sigma {acc[0] / acc[1]:.5e} barn")
    return out


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
    """HBM bytes per screening launch measured with FETCH_SIZE / WRITE_SIZE
    (profiles/k_screen_hbm_traffic_screened_pipi.md), or None."""
    path = os.path.join(ROOT, "profiles", "k_screen_hbm_traffic_screened_pipi.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--points", type=int, default=1 << 20)
    ap.add_argument("--preset", type=int, default=2)
    ap.add_argument("--no-res", action="store_true",
                    help="skip the PP[RES+CON] resonance throughput leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision-run", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 20)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-flow", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run 256 MiB device copies (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--flow-steps", type=int, default=500)
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
    lib = P.lib
    grid = abi.GraGrid()
    abi.check(lib.gra_hip_grid_init(C.byref(grid), cfg.dim, 128))
    seed = 12345678
    it = 0
    log(f"workload preset {args.preset}: {N} points/iteration, warmup {args.warmup}, "
        f"steps {args.steps}")
    for _ in range(args.warmup):
        st = P.vegas_iterate(grid, N, seed, it)
        it += 1
        abi.check(lib.gra_hip_grid_optimize(C.byref(grid), C.byref(st), 1.5))
    for s in range(4):
        P.kernel_time(s, reset=True)
    fsum = f2sum = 0.0
    maxf = 0.0
    n_valid = 0
    state = [0.0, 0.0, 0.0]
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = P.vegas_iterate(grid, N, seed, it)
        it += 1
        abi.check(lib.gra_hip_grid_optimize(C.byref(grid), C.byref(st), 1.5))
        fsum += st.fsum
        f2sum += st.f2sum
        maxf = max(maxf, st.maxf)
        n_valid += st.n_valid
        # iteration combination (src/MGraniitti.cc:1222-1246) for the sigma report
        I = st.fsum / N
        v = max((st.f2sum / N - I * I) / N, 1e-300)
        state[0] += I / v
        state[1] += 1.0 / v
    t1 = time.perf_counter()
    dt = t1 - t0
    ms_tree, n_tree = P.kernel_time(0)
    ms_scr, n_scr = P.kernel_time(1)
    ms_acc, n_acc = P.kernel_time(2)
    pts_per_s = N * args.steps / dt
    mean_f = fsum / (N * args.steps)
    eff = mean_f / maxf if maxf > 0 else 0.0
    value = pts_per_s * eff
    sigma = state[0] / state[1]
    sigma_err = 1.0 / np.sqrt(state[1])
    log(f"timed: {dt:.3f} s, {pts_per_s:.3e} points/s, eff {eff:.4f}, sigma {sigma:.5e} +- "
        f"{sigma_err:.2e} barn")

    roofline = None
    # per-node operation counts are those of the FP32 Regge continuum node
    if (cfg.screening and n_scr > 0 and cfg.process == abi.GRA_PROC_REGGE_CON
            and cfg.amp_precision == abi.GRA_AMP_FP32):
        nodes = (cfg.n_loop_phi + 1) * (cfg.n_loop_kt + 1)
        cand_per_launch = n_valid / n_scr
        avg_ms = ms_scr / n_scr
        flop = FLOP_PER_NODE * nodes * cand_per_launch
        achieved = flop / (avg_ms * 1e-3) / 1e12
        roofline = {
            "bound": "valu", "kernel": "k_screen (screening loop)",
            "achieved": achieved, "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / VALU_PEAK_TFLOPS, "traffic": screen_traffic(),
            "avg_launch_ms": avg_ms, "launches": n_scr,
            "flop_per_node": FLOP_PER_NODE, "node_dtype": "f32",
            "nodes_per_point": nodes, "screened_points_per_launch": cand_per_launch,
        }
    # time to 0.1% relative sigma error with plain VEGAS (GPU)
    t_prec = None
    prec_res = None
    if not args.no_precision_run:
        G = MGraniitti(P)
        vc = vegas_cfg(N, 1, burnin=args.warmup, seed=424242, precision=1e-3, max_iter=60)
        res = G.Integrate(vc)
        t_prec = res.seconds_to_precision if res.seconds_to_precision > 0 else None
        prec_res = {"sigma": res.sigma, "sigma_err": res.sigma_err, "iterations": res.iterations,
                    "seconds_total": res.seconds_total, "calls_total": int(res.calls_total)}
        log(f"precision run: {prec_res}, t(0.1%) = {t_prec}")
    flow = None
    res = None
    if not args.no_res and args.preset == 2:
        try:
            res = res_leg()
        except Exception as e:  # reported, never fatal to the headline line
            res = {"error": str(e)}
    if not args.no_flow:
        flow = flow_leg(P, grid, args, prec_res["sigma"] if prec_res else sigma,
                        prec_res["sigma_err"] if prec_res else sigma_err)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, grid, eff, args.cpu_sample, args.cpu_threads)
        log(f"cpu baseline: {cpu}")
    P.close()
    line = {
        "metric": "time to 0.1% sigma rel. error & unweighted events/s, flow vs VEGAS, 1 MI355X",
        "value": value, "unit": "unweighted events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": WORKLOADS.get(args.preset, f"preset {args.preset}") +
                   f", VEGAS {N} points/iteration",
                   "preset": args.preset, "points_per_step": N, "integrator": "VEGAS"},
        "points_per_s": pts_per_s, "unweighting_efficiency": eff,
        "time_to_0p1pct_s": t_prec, "precision_run": prec_res,
        "sigma_barn": sigma, "sigma_err_barn": sigma_err,
        "kernel_ms": {"tree": ms_tree, "screen": ms_scr, "accumulate": ms_acc},
        "flow": flow, "resonances": res, "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
