"""GPU flow sampler against the PyTorch float64 reference (tests/flow_ref.py) and
finite differences; training reduces the weight variance and keeps sigma."""
import ctypes as C
import sys

import numpy as np
import pytest

from graniitti_amd import abi
from graniitti_amd.process import MContinuum, MGraniitti, vegas_cfg
from flow_ref import FlowRef, layout

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def flow_cfg(D, L, K, H, depth, fp32, batch=4096, seed=7, lr=1e-3):
    fc = abi.GraFlowCfg()
    fc.dim, fc.layers, fc.bins, fc.hidden = D, L, K, H
    fc.depth = -depth if fp32 else depth
    fc.steps, fc.batch, fc.lr = 0, batch, lr
    fc.beta1, fc.beta2, fc.eps = 0.9, 0.999, 1e-8
    fc.seed, fc.ess_fallback = seed, 0.1
    return fc


def init_random(P, fc, scale, seed):
    lib = P.lib
    abi.check(lib.gra_hip_flow_init(P.h, C.byref(fc)), P.h)
    n = C.c_uint64()
    abi.check(lib.gra_hip_flow_param_count(P.h, C.byref(n)), P.h)
    th = np.empty(n.value, dtype=np.float32)
    abi.check(lib.gra_hip_flow_get_params(P.h, abi.fptr(th)), P.h)
    lay, nref, ldp = layout(fc.dim, fc.layers, fc.bins, fc.hidden, abs(fc.depth))
    assert nref == n.value
    rng = np.random.default_rng(seed)
    for d in lay:  # perturb every block except the padding columns of the output layer
        for key in ("W1", "b1"):
            off, shp = d[key]
            th[off:off + int(np.prod(shp))] += rng.normal(0, scale, int(np.prod(shp)))
        for (Wo, Ws), (bo, bs) in d["Wh"]:
            th[Wo:Wo + Ws[0] * Ws[1]] += rng.normal(0, scale, Ws[0] * Ws[1])
        off, (H, ld) = d["Wo"]
        valid = len(d["B"]) * (3 * fc.bins - 1)
        W = th[off:off + H * ld].reshape(H, ld)
        W[:, :valid] += rng.normal(0, scale, (H, valid))
        off, _ = d["bo"]
        th[off:off + valid] += rng.normal(0, scale, valid)
    abi.check(lib.gra_hip_flow_set_params(P.h, abi.fptr(th)), P.h)
    return th


def forward(P, z):
    n, D = z.shape
    u = np.empty((n, D), dtype=np.float32)
    lq = np.empty(n, dtype=np.float32)
    z = np.ascontiguousarray(z, dtype=np.float32)
    abi.check(P.lib.gra_hip_flow_forward(P.h, abi.fptr(z), n, abi.fptr(u), abi.fptr(lq)), P.h)
    return u, lq


def grad(P, u, c, npar):
    n = u.shape[0]
    g = np.empty(npar, dtype=np.float32)
    lq = np.empty(n, dtype=np.float32)
    u = np.ascontiguousarray(u, dtype=np.float32)
    c = np.ascontiguousarray(c, dtype=np.float32)
    abi.check(P.lib.gra_hip_flow_grad(P.h, abi.fptr(u), abi.fptr(c), n, abi.fptr(g),
                                      abi.fptr(lq)), P.h)
    return g, lq


@pytest.fixture(scope="module")
def proc():
    cfg = abi.cfg_default(1)
    P = MContinuum(cfg)
    yield P
    P.close()


@pytest.mark.parametrize("fp32,tol", [(True, 2e-4), (False, 6e-2)])
def test_flow_forward_matches_reference(proc, fp32, tol):
    D, L, K, H, depth = 8, 4, 8, 64, 2
    say(f"case: flow forward fp32={fp32} D={D} L={L} K={K} H={H}")
    fc = flow_cfg(D, L, K, H, depth, fp32)
    th = init_random(proc, fc, 0.3, 11)
    z = np.random.default_rng(1).random((256, D)).astype(np.float32)
    u, lq = forward(proc, z)
    ref = FlowRef(th, D, L, K, H, depth)
    import torch
    with torch.no_grad():
        lq_ref, z_ref = ref.logq(torch.tensor(u, dtype=torch.float64))
    err_z = np.abs(z_ref.numpy() - z).max()
    err_lq = np.abs(lq_ref.numpy() - lq).max()
    say(f"flow forward fp32={fp32}: max |z - F(u)| {err_z:.2e}, max |dlogq| {err_lq:.2e}")
    assert err_z < tol and err_lq < 5 * tol
    assert np.all((u >= 0) & (u < 1))


def test_flow_gradient_matches_reference(proc):
    D, L, K, H, depth = 8, 4, 8, 64, 2
    say("case: flow gradient fp32 vs torch autograd")
    fc = flow_cfg(D, L, K, H, depth, True)
    th = init_random(proc, fc, 0.3, 12)
    rng = np.random.default_rng(2)
    u = rng.random((256, D)).astype(np.float32)
    c = rng.random(256).astype(np.float32)
    g, lq = grad(proc, u, c, len(th))
    ref = FlowRef(th, D, L, K, H, depth)
    gr, lqr = ref.grad(u.astype(np.float64), c.astype(np.float64))
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    say(f"flow grad fp32: relative L2 error {rel:.2e}, logq err {np.abs(lq - lqr).max():.2e}")
    assert rel < 1e-4
    big = np.abs(gr) > 1e-2 * np.abs(gr).max()
    assert np.all(np.abs(g[big] - gr[big]) <= 1e-3 * np.abs(gr[big]) + 1e-6 * np.abs(gr).max())


def test_flow_gradient_bf16_direction(proc):
    D, L, K, H, depth = 8, 4, 8, 64, 2
    say("case: flow gradient bf16 MFMA vs torch autograd")
    fc = flow_cfg(D, L, K, H, depth, False)
    th = init_random(proc, fc, 0.3, 13)
    rng = np.random.default_rng(3)
    u = rng.random((1024, D)).astype(np.float32)
    c = rng.random(1024).astype(np.float32)
    g, lq = grad(proc, u, c, len(th))
    gr, _ = FlowRef(th, D, L, K, H, depth).grad(u.astype(np.float64), c.astype(np.float64))
    cos = float(g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr)))
    say(f"flow grad bf16: cosine {cos:.5f}")
    assert cos > 0.995


def test_flow_backward_finite_differences_large():
    # config-5 sized flow (D = 14, 10 layers, K = 32, 3 x 256) in FP32: central
    # differences of L = -sum c log q against the backward kernel, rel. tol 1e-3
    cfg = abi.cfg_default(1)
    cfg.dim = 14
    cfg.w_central = 0.149
    cfg.m_central = 0.775
    cfg.m_daughter = 0.13957061
    say("case: flow FD check D=14 L=10 K=32 H=256 depth=3")
    with MContinuum(cfg) as P:
        fc = flow_cfg(14, 10, 32, 256, 3, True)
        th = init_random(P, fc, 0.05, 14)
        rng = np.random.default_rng(4)
        u = rng.random((256, 14)).astype(np.float32)
        c = rng.random(256).astype(np.float32) / 256
        g, _ = grad(P, u, c, len(th))
        order = np.argsort(-np.abs(g))
        pick = list(order[:12]) + list(rng.choice(order[:2000], 12, replace=False))
        worst = 0.0
        for k in pick:
            hstep = 2e-3 * max(1.0, abs(float(th[k])))
            vals = []
            for sgn in (+1, -1):
                t2 = th.copy()
                t2[k] += sgn * hstep
                abi.check(P.lib.gra_hip_flow_set_params(P.h, abi.fptr(t2)), P.h)
                _, lq = grad(P, u, c, len(th))
                vals.append(-(c.astype(np.float64) * lq.astype(np.float64)).sum())
            fd = (vals[0] - vals[1]) / (2 * hstep)
            rel = abs(fd - g[k]) / max(abs(g[k]), 1e-3 * abs(g[order[0]]))
            worst = max(worst, rel)
        abi.check(P.lib.gra_hip_flow_set_params(P.h, abi.fptr(th)), P.h)
        say(f"flow FD check: worst relative deviation {worst:.2e} over {len(pick)} parameters")
        assert worst < 1e-3


def test_flow_training_reduces_variance_and_keeps_sigma():
    cfg = abi.cfg_default(2)
    cfg.n_bt = cfg.n_kt2 = cfg.n_fb = 2000
    say("case: flow training on the screened process (small tables), 40 steps x 16384")
    with MContinuum(cfg) as P:
        G = MGraniitti(P)
        res = G.Integrate(vegas_cfg(1 << 18, 4, burnin=4, seed=99))
        fc = flow_cfg(8, 6, 16, 128, 2, False, batch=16384, lr=1e-3)
        abi.check(P.lib.gra_hip_flow_init(P.h, C.byref(fc)), P.h)
        abi.check(P.lib.gra_hip_flow_set_grid(P.h, C.byref(G.grid)), P.h)
        s0 = abi.GraFlowStats()
        abi.check(P.lib.gra_hip_flow_sample(P.h, 1 << 18, 5, C.byref(s0)), P.h)
        st = abi.GraFlowStats()
        abi.check(P.lib.gra_hip_flow_train(P.h, 40, C.byref(st)), P.h)
        s1 = abi.GraFlowStats()
        abi.check(P.lib.gra_hip_flow_sample(P.h, 1 << 18, 6, C.byref(s1)), P.h)
    say(f"VEGAS sigma {res.sigma:.5e} +- {res.sigma_err:.2e}; flow before: sigma {s0.sigma:.5e} "
        f"cv2 {s0.cv2:.3f} ess {s0.ess:.3f}; after: sigma {s1.sigma:.5e} +- {s1.sigma_err:.2e} "
        f"cv2 {s1.cv2:.3f} ess {s1.ess:.3f}")
    assert s1.cv2 < s0.cv2
    comb = np.sqrt(res.sigma_err ** 2 + s1.sigma_err ** 2)
    assert abs(s1.sigma - res.sigma) < 3 * comb
