"""GPU flow sampler against the PyTorch float64 reference (tests/flow_ref.py) and
finite differences; training reduces the weight variance and keeps sigma."""
import ctypes as C
import sys

import numpy as np
import pytest

from graniitti_amd import abi
from graniitti_amd.process import MContinuum, MGraniitti, vegas_cfg
from flow_ref import FlowRef, layout, prm_stride

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
        # the 3K - 1 parameter columns of each transformed dimension (padded blocks)
        cols = np.concatenate([j * prm_stride(fc.bins) + np.arange(3 * fc.bins - 1)
                               for j in range(len(d["B"]))])
        W = th[off:off + H * ld].reshape(H, ld)
        W[:, cols] += rng.normal(0, scale, (H, len(cols)))
        off, _ = d["bo"]
        th[off + cols] += rng.normal(0, scale, len(cols))
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


# K = 8 runs the unrolled spline kernels, K = 10 the runtime-K loops; the BF16
# cases run the fused sampling layer, at H = 64 and at the two flow shapes the
# samplers use: 6 layers, K = 16, 2 x 128 (D = 8) and 10 layers, K = 32,
# 3 x 256 (D = 14). The BF16 path is compared with the float64 reference that
# rounds the same operands to bfloat16 (weights of the MFMA products, hidden
# activations), which leaves only the FP32 accumulation and the splines' FP32
# arithmetic as differences (tolerances tol for z, lqtol for log q); against the
# unrounded reference the 8-bit mantissas of the operands move log q by up to
# ~0.4 at these random parameters, which is reported.
# The 10-layer flow is taken at weight perturbations of 0.01 (log q in [-2.5,
# 4.4], the range of a trained sampler): at 0.05 its log q spans 5 to 72 and
# the composed map is so ill-conditioned that the FP32 path itself differs from
# float64 by up to 8 in log q (MI355X: FP32 max dz 7e-2, BF16 1.2e-1), so no
# tolerance there separates a kernel error from amplified rounding. The trained
# 10-layer flow is compared in test_flow_training_on_tensor_process.
@pytest.mark.parametrize("fp32,tol,lqtol,D,L,K,H,depth,scale", [
    (True, 2e-4, 1e-3, 8, 4, 8, 64, 2, 0.05),
    (False, 2e-3, 2e-2, 8, 4, 8, 64, 2, 0.05),
    (False, 2e-3, 2e-2, 8, 4, 10, 64, 2, 0.05),
    (False, 2e-3, 2e-2, 8, 6, 16, 128, 2, 0.05),
    (True, 2e-5, 1e-3, 14, 10, 32, 256, 3, 0.01),
    (False, 2e-3, 2e-2, 14, 10, 32, 256, 3, 0.01),
])
def test_flow_forward_matches_reference(proc, fp32, tol, lqtol, D, L, K, H, depth, scale):
    say(f"case: flow forward fp32={fp32} D={D} L={L} K={K} H={H} depth={depth} scale={scale}")
    fc = flow_cfg(D, L, K, H, depth, fp32)
    P = proc if D == proc.cfg.dim else MContinuum(rho_cfg())
    try:
        th = init_random(P, fc, scale, 11)
        z = np.random.default_rng(1).random((256, D)).astype(np.float32)
        u, lq = forward(P, z)
    finally:
        if P is not proc:
            P.close()
    import torch
    with torch.no_grad():
        lq_ref, z_ref = FlowRef(th, D, L, K, H, depth, bf16=not fp32).logq(
            torch.tensor(u, dtype=torch.float64))
        lq_64, z_64 = FlowRef(th, D, L, K, H, depth).logq(torch.tensor(u, dtype=torch.float64))
    err_z = np.abs(z_ref.numpy() - z).max()
    err_lq = np.abs(lq_ref.numpy() - lq).max()
    say(f"flow forward fp32={fp32}: max |z - F(u)| {err_z:.2e}, max |dlogq| {err_lq:.2e} "
        f"(median {np.median(np.abs(lq_ref.numpy() - lq)):.1e}); against the unrounded "
        f"reference: {np.abs(z_64.numpy() - z).max():.2e}, {np.abs(lq_64.numpy() - lq).max():.2e}")
    assert err_z < tol and err_lq < lqtol
    assert np.all((u >= 0) & (u < 1))


@pytest.mark.parametrize("D,L,K,H,depth,scale,n", [
    (8, 6, 16, 128, 2, 0.05, 8192),
    (14, 4, 32, 256, 3, 0.01, 5056),
])
def test_flow_weight_gradient_tiles(proc, D, L, K, H, depth, scale, n):
    # BF16 weight gradients of the 128 x 128-tile kernel (hidden and output
    # products, bias rows from an all-ones operand) and of its 16 x 128 input-layer
    # form, against the split 64 x 64-tile kernel on the same BF16 operands: only
    # the FP32 summation order differs. n spans several row chunks of 512.
    say(f"case: BF16 weight-gradient tiles vs split tiles, D={D} L={L} K={K} H={H} n={n}")
    rng = np.random.default_rng(5)
    u = rng.random((n, D)).astype(np.float32)
    c = rng.random(n).astype(np.float32)
    gs = []
    P = MContinuum(rho_cfg()) if D == 14 else proc
    for split in (0, 1):
        fc = flow_cfg(D, L, K, H, depth, False)
        fc.wgrad_split = split
        th = init_random(P, fc, scale, 21)
        g, _ = grad(P, u, c, len(th))
        gs.append(g.astype(np.float64))
    if D == 14:
        P.close()
    rel = np.linalg.norm(gs[0] - gs[1]) / np.linalg.norm(gs[1])
    dmax = np.abs(gs[0] - gs[1]).max() / np.abs(gs[1]).max()
    say(f"tiles vs split: relative L2 {rel:.2e}, max |dg| / max |g| {dmax:.2e}")
    assert rel < 2e-5 and dmax < 2e-5
    if H == 128:
        # the whole BF16 gradient against float64 autograd: six layers of BF16
        # operands and BF16 backward products (both kernels: cosine 0.9849)
        fc.wgrad_split = 0
        init_random(proc, fc, scale, 21)
        g, _ = grad(proc, u[:2048], c[:2048], len(th))
        for rounded in (False, True):
            gr, _ = FlowRef(th, D, L, K, H, depth, bf16=rounded).grad(
                u[:2048].astype(np.float64), c[:2048].astype(np.float64))
            cos = float(g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr)))
            say(f"tile-kernel gradient, 2048 rows, vs float64 (operands rounded to bf16: "
                f"{rounded}): cosine {cos:.5f}")
            if not rounded:
                assert cos > 0.98


@pytest.mark.parametrize("K", [8, 10, 16])
def test_flow_gradient_matches_reference(proc, K):
    D, L, H, depth = 8, 4, 64, 2
    say(f"case: flow gradient fp32 vs torch autograd, K={K}")
    fc = flow_cfg(D, L, K, H, depth, True)
    th = init_random(proc, fc, 0.05, 12)
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
    th = init_random(proc, fc, 0.05, 13)
    rng = np.random.default_rng(3)
    u = rng.random((1024, D)).astype(np.float32)
    c = rng.random(1024).astype(np.float32)
    g, lq = grad(proc, u, c, len(th))
    gr, _ = FlowRef(th, D, L, K, H, depth).grad(u.astype(np.float64), c.astype(np.float64))
    cos = float(g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr)))
    say(f"flow grad bf16: cosine {cos:.5f}")
    assert cos > 0.995


@pytest.mark.parametrize("D,L,K,H,depth", [(8, 6, 16, 128, 2), (14, 4, 32, 256, 3)])
def test_unrolled_splines_match_runtime_loops(proc, D, L, K, H, depth):
    """The unrolled spline kernels (K = 16, 32) against the runtime-K loops, which
    the K = 10 cases check against the float64 reference: fused BF16 sampling
    layer (u, log q) and spline backward (FP32 gradient) on the same parameters."""
    say(f"case: unrolled vs runtime-K splines D={D} L={L} K={K} H={H} depth={depth}")
    P = proc if D == proc.cfg.dim else MContinuum(rho_cfg())
    rng = np.random.default_rng(5)
    z = rng.random((1024, D)).astype(np.float32)
    uu = rng.random((1024, D)).astype(np.float32)
    c = rng.random(1024).astype(np.float32)
    out = []
    try:
        for loops in (0, 1):
            fc = flow_cfg(D, L, K, H, depth, False)
            fc.spline_loops = loops
            init_random(P, fc, 0.05, 21)
            u, lq = forward(P, z)
            fc = flow_cfg(D, L, K, H, depth, True)
            fc.spline_loops = loops
            th = init_random(P, fc, 0.05, 22)
            g, lqg = grad(P, uu, c, len(th))
            out.append((u, lq, g, lqg))
    finally:
        if P is not proc:
            P.close()
    (u0, lq0, g0, lg0), (u1, lq1, g1, lg1) = out
    du, dl = np.abs(u0 - u1).max(), np.abs(lq0 - lq1).max()
    dg = np.abs(g0 - g1).max() / np.abs(g1).max()
    exact = np.array_equal(u0, u1) and np.array_equal(lq0, lq1) and np.array_equal(g0, g1)
    say(f"unrolled vs loops: max |du| {du:.2e}, |dlogq| {dl:.2e}, grad {dg:.2e}, bit-exact {exact}")
    assert du <= 1e-6 and dl <= 1e-5 and dg <= 1e-6
    assert np.abs(lg0 - lg1).max() <= 1e-5


def test_fused_layer_output_groups_agree():
    """The fused sampling layer computes the output product and the splines in
    groups of transformed dimensions (out_groups; 0 picks the largest group whose
    parameter tile fits the block's LDS share, 4 of 7 here): every grouping gives
    the same u and log q bit for bit, D = 14, 10 layers, K = 32, 3 x 256. A group
    of all 7 dimensions needs more LDS than a CU has and is refused."""
    D, L, K, H, depth = 14, 10, 32, 256, 3
    say("case: fused layer output groups 1, 2, 3, auto at D=14 L=10 K=32 H=256 depth=3")
    z = np.random.default_rng(8).random((2048, D)).astype(np.float32)
    out = {}
    with MContinuum(rho_cfg()) as P:
        for g in (1, 2, 3, 0):
            fc = flow_cfg(D, L, K, H, depth, False)
            fc.out_groups = g
            init_random(P, fc, 0.05, 31)
            out[g] = forward(P, z)
        fc = flow_cfg(D, L, K, H, depth, False)
        fc.out_groups = 7
        assert P.lib.gra_hip_flow_init(P.h, C.byref(fc)) == abi.GRA_E_ARG
    for g in (1, 2, 3):
        assert np.array_equal(out[g][0], out[0][0]), g
        assert np.array_equal(out[g][1], out[0][1]), g
    say("output groups 1, 2, 3, auto: identical u and log q")


def test_fused_layer_groups_wider_than_the_block():
    """At H = 128 a block has 256 threads for 64 rows: a group of all 7
    transformed dimensions (D = 14, out_groups = 7) has 448 (row, dimension)
    splines, which the block's threads take in turns. Same u and log q as one
    dimension per group."""
    D, L, K, H, depth = 14, 4, 16, 128, 2
    say("case: fused layer groups 1 and 7 at D=14 L=4 K=16 H=128 depth=2")
    z = np.random.default_rng(9).random((1024, D)).astype(np.float32)
    out = {}
    with MContinuum(rho_cfg()) as P:
        for g in (1, 7):
            fc = flow_cfg(D, L, K, H, depth, False)
            fc.out_groups = g
            init_random(P, fc, 0.05, 33)
            out[g] = forward(P, z)
    assert np.array_equal(out[7][0], out[1][0]) and np.array_equal(out[7][1], out[1][1])
    say("output groups 1 and 7 at H=128: identical u and log q")


def rho_cfg():
    cfg = abi.cfg_default(1)
    cfg.dim = 14
    cfg.w_central = 0.149
    cfg.m_central = 0.775
    cfg.m_daughter = 0.13957061
    return cfg


def test_flow_gradient_matches_reference_large():
    # flow of the 2->6 sampler (D = 14, 10 layers, K = 32, 3 x 256) in FP32 at a
    # strongly non-identity point (log q spans tens of units): the backward kernels
    # against float64 autograd of the same function
    say("case: flow gradient fp32 vs torch autograd, D=14 L=10 K=32 H=256 depth=3")
    with MContinuum(rho_cfg()) as P:
        fc = flow_cfg(14, 10, 32, 256, 3, True)
        th = init_random(P, fc, 0.05, 14)
        rng = np.random.default_rng(4)
        u = rng.random((256, 14)).astype(np.float32)
        c = rng.random(256).astype(np.float32) / 256
        g, lq = grad(P, u, c, len(th))
    gr, lqr = FlowRef(th, 14, 10, 32, 256, 3).grad(u.astype(np.float64), c.astype(np.float64))
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    say(f"flow grad fp32 (large): relative L2 error {rel:.2e}, logq err {np.abs(lq - lqr).max():.2e}")
    assert rel < 1e-3


def test_flow_backward_finite_differences_large():
    # Central differences in FP32 of L = -sum c log q against the backward kernel,
    # rel. tol 1e-3, for the 2->6 sampler's flow size. log q is piecewise smooth
    # (ReLU units, spline bins): a step that moves pre-activations across zero
    # biases the difference, and FP32 rounding of log q bounds the step from
    # below. A near-identity point keeps log q small (so its rounding is small in
    # absolute terms), and directions along the gradient maximise the signal.
    # The full gradient direction and each layer's output block (whose gradient
    # runs through the spline backward and every later layer's MLP and input
    # derivative) are held to 1e-3; the directions that move hidden weights,
    # whose steps cross ReLU kinks, to 1e-2 (measured 1.5e-3 to 3.4e-3). A hidden
    # direction restricted to units clear of their kink needs a step so small
    # that the FP32 rounding of log q dominates the difference (4e-3). The exact
    # comparison at this size is against float64 autograd above (1.2e-4).
    say("case: flow FD check D=14 L=10 K=32 H=256 depth=3")
    with MContinuum(rho_cfg()) as P:
        fc = flow_cfg(14, 10, 32, 256, 3, True)
        th = init_random(P, fc, 0.001, 14).astype(np.float64)
        lay, _, _ = layout(14, 10, 32, 256, 3)
        rng = np.random.default_rng(4)
        u = rng.random((256, 14)).astype(np.float32)
        c = rng.random(256).astype(np.float32) / 256
        g, lq0 = grad(P, u, c, len(th))
        g = g.astype(np.float64)
        L0 = -(c.astype(np.float64) * lq0.astype(np.float64)).sum()

        def loss_at(t2):
            abi.check(P.lib.gra_hip_flow_set_params(P.h, abi.fptr(t2.astype(np.float32))), P.h)
            _, lq = grad(P, u, c, len(th))
            return -(c.astype(np.float64) * lq.astype(np.float64)).sum()

        def central(v, h):
            tp = (th + h * v).astype(np.float32).astype(np.float64)
            tm = (th - h * v).astype(np.float32).astype(np.float64)
            return (loss_at(tp) - loss_at(tm)) / float((tp - tm) @ v)  # the step as rounded

        def fd_along(v):
            # Richardson extrapolation of steps h and 2h removes the O(h^2) curvature term
            gv = float(g @ v)
            h = 1e-2 * abs(L0) / abs(gv)
            fd = (4 * central(v, h) - central(v, 2 * h)) / 3
            return abs(fd - gv) / abs(gv)

        def block_dir(blocks):
            v = np.zeros_like(g)
            for off, shp in blocks:
                sz = int(np.prod(shp))
                v[off:off + sz] = g[off:off + sz]
            return v / np.linalg.norm(v)

        rel_full = fd_along(g / np.linalg.norm(g))
        rel_out, rel_hid = [], []
        for l in (0, 3, 6, 9):
            d = lay[l]
            rel_out.append(fd_along(block_dir([d["Wo"], d["bo"]])))
            rel_hid.append(fd_along(block_dir([d["W1"], d["b1"]] + [x for p in d["Wh"] for x in p])))
        abi.check(P.lib.gra_hip_flow_set_params(P.h, abi.fptr(th.astype(np.float32))), P.h)
    say(f"flow FD check: full gradient {rel_full:.2e}; output blocks "
        f"{' '.join('%.1e' % r for r in rel_out)}; hidden blocks "
        f"{' '.join('%.1e' % r for r in rel_hid)}")
    assert rel_full < 1e-3
    assert max(rel_out) < 1e-3
    assert max(rel_hid) < 1e-2


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
        abi.check(P.lib.gra_hip_flow_sample(P.h, 1 << 18, 5, None, None, C.byref(s0)), P.h)
        st = abi.GraFlowStats()
        abi.check(P.lib.gra_hip_flow_train(P.h, 40, C.byref(st)), P.h)
        s1 = abi.GraFlowStats()
        abi.check(P.lib.gra_hip_flow_sample(P.h, 1 << 18, 6, None, None, C.byref(s1)), P.h)
    say(f"VEGAS sigma {res.sigma:.5e} +- {res.sigma_err:.2e}; flow before: sigma {s0.sigma:.5e} "
        f"cv2 {s0.cv2:.3f} ess {s0.ess:.3f}; after: sigma {s1.sigma:.5e} +- {s1.sigma_err:.2e} "
        f"cv2 {s1.cv2:.3f} ess {s1.ess:.3f}")
    assert s1.cv2 < s0.cv2
    comb = np.sqrt(res.sigma_err ** 2 + s1.sigma_err ** 2)
    assert abs(s1.sigma - res.sigma) < 3 * comb


def test_flow_training_on_tensor_process():
    # the 2->6 sampler's flow shape (D = 14, 10 layers, K = 32, 3 x 256, BF16)
    # on the screened tensor-pomeron rho rho process with small eikonal tables
    cfg = abi.cfg_default(3)
    cfg.amp_precision = abi.GRA_AMP_FP64
    cfg.n_bt = cfg.n_kt2 = cfg.n_fb = 1000
    say("case: flow D=14 L=10 K=32 H=256 depth=3 on the screened rho rho process, 30 steps x 16384")
    with MContinuum(cfg) as P:
        G = MGraniitti(P)
        res = G.Integrate(vegas_cfg(1 << 18, 4, burnin=4, seed=77))
        fc = flow_cfg(14, 10, 32, 256, 3, False, batch=16384, lr=1e-3)
        abi.check(P.lib.gra_hip_flow_init(P.h, C.byref(fc)), P.h)
        abi.check(P.lib.gra_hip_flow_set_grid(P.h, C.byref(G.grid)), P.h)
        s0 = abi.GraFlowStats()
        abi.check(P.lib.gra_hip_flow_sample(P.h, 1 << 18, 5, None, None, C.byref(s0)), P.h)
        st = abi.GraFlowStats()
        abi.check(P.lib.gra_hip_flow_train(P.h, 30, C.byref(st)), P.h)
        s1 = abi.GraFlowStats()
        abi.check(P.lib.gra_hip_flow_sample(P.h, 1 << 18, 6, None, None, C.byref(s1)), P.h)
        n = C.c_uint64()
        abi.check(P.lib.gra_hip_flow_param_count(P.h, C.byref(n)), P.h)
        th = np.empty(n.value, dtype=np.float32)
        abi.check(P.lib.gra_hip_flow_get_params(P.h, abi.fptr(th)), P.h)
        z = np.random.default_rng(2).random((256, 14)).astype(np.float32)
        u, lq = forward(P, z)
    say(f"VEGAS sigma {res.sigma:.5e} +- {res.sigma_err:.2e}; flow before: cv2 {s0.cv2:.3f} "
        f"ess {s0.ess:.3f}; after: sigma {s1.sigma:.5e} +- {s1.sigma_err:.2e} cv2 {s1.cv2:.3f} "
        f"ess {s1.ess:.3f} fallback {s1.fallback}")
    assert s1.cv2 < s0.cv2
    assert s1.fallback == (1 if s1.ess < fc.ess_fallback else 0)
    # the trained BF16 flow's fused sampling layers against the float64
    # reference with the same bfloat16 roundings
    import torch
    with torch.no_grad():
        lq_ref, z_ref = FlowRef(th, 14, 10, 32, 256, 3, bf16=True).logq(
            torch.tensor(u, dtype=torch.float64))
    err_z = np.abs(z_ref.numpy() - z).max()
    err_lq = np.abs(lq_ref.numpy() - lq).max()
    say(f"trained flow: log q in [{lq.min():.2f}, {lq.max():.2f}], max |z - F(u)| {err_z:.2e}, "
        f"max |dlogq| {err_lq:.2e}")
    assert err_z < 2e-3 and err_lq < 2e-2
    comb = np.sqrt(res.sigma_err ** 2 + s1.sigma_err ** 2)
    assert abs(s1.sigma - res.sigma) < 3 * comb
