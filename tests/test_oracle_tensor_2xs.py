"""Checks that pin the oracle's tensor-pomeron ME4 for a pseudoscalar pair
(pp -> p pi+ pi- p, "2xS", src/MTensorPomeron.cc:726-763), and the spin-mode
dispatch of the C ABI (src/MTensorPomeron.cc:643-654).

The oracle contracts the reference's rank-4 tensors with plain index sums. What
the mathematics fixes, and a plausible index, metric or sign mistake breaks:
- the scalar form each (proton vertex)(Pomeron propagator)(meson vertex) chain
  must reduce to, 2 (P.K)^2 - P^2 K^2 / 2 times its factors (the propagator's
  trace term vanishes against the traceless meson vertex);
- Lorentz invariance of the amplitude under common boosts and rotations;
- Bose symmetry: exchanging the two pions maps the t- onto the u-channel;
- four identical proton-helicity replicas with |M|^2 = sum / 4.
The couplings and the overall phase convention are not fixed by these and stay
unpinned."""
import ctypes as C

import numpy as np
import pytest

from graniitti_amd import abi
import oracle_lib as ol

G = np.diag([1.0, -1.0, -1.0, -1.0])


def mdot(a, b):
    return a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3]


@pytest.fixture(scope="module")
def run9():
    cfg = abi.cfg_default(9)
    R = ol.OracleRun(cfg, 2)
    yield R
    R.close()


def _points(R, n, seed):
    E = R.cfg.beam_E1
    pz = np.sqrt(E * E - R.cfg.m_beam ** 2)
    pa, pb = np.array([E, 0, 0, pz]), np.array([E, 0, 0, -pz])
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        ok, pf, legs, sc = R.kinematics(rng.random(R.cfg.dim))
        if ok:
            out.append(np.array([pa, pb, pf[1], pf[2], pf[3], pf[4]]))
    return out


def _boost(p, beta):
    b = np.asarray(beta, dtype=float)
    b2 = b @ b
    g = 1.0 / np.sqrt(1.0 - b2)
    L = np.eye(4)
    L[0, 0] = g
    L[0, 1:] = L[1:, 0] = g * b
    L[1:, 1:] += (g - 1.0) * np.outer(b, b) / b2
    return p @ L.T


def _kelly_f1(Q2, mp):
    tau = abs(Q2) / (2 * mp) ** 2
    GE = (1 - 0.24 * tau) / (1 + 10.98 * tau + 12.82 * tau ** 2 + 21.97 * tau ** 3)
    GM = 2.792847337 * (1 + 0.12 * tau) / (1 + 10.97 * tau + 18.86 * tau ** 2 + 6.55 * tau ** 3)
    return (GE + tau * GM) / (1 + tau)


def _reduced(cfg, p):
    """The amplitude from the scalar chains, with the factors written out."""
    pa, pb, p1, p2, p3, p4 = p
    t1, t2 = mdot(pa - p1, pa - p1), mdot(pb - p2, pb - p2)
    s = lambda a, b: mdot(a + b, a + b)
    P1, P2 = p1 + pa, p2 + pb
    pt, pu = pa - p1 - p3, p4 - pa + p1
    gP = cfg.tp_gPpipi

    def chain(Pv, kp, k, ss, t):
        K = kp + k
        q = kp - k
        F = (complex(0, -ss * cfg.tp_ap_P)) ** (cfg.tp_delta_P + cfg.tp_ap_P * t) / (4 * ss)
        FM = 1 / (1 - mdot(q, q) / 0.5)
        c = -3j * cfg.tp_gPNN * _kelly_f1(t, cfg.tp_mp_ff)
        d = -2j * gP * FM
        return c * F * d * (2 * mdot(Pv, K) ** 2 - 0.5 * mdot(Pv, Pv) * mdot(K, K))

    M2 = cfg.m_prop ** 2
    ff = lambda v: 1 / (1 - (mdot(v, v) - M2) / cfg.b_POW)
    Mt = (chain(P1, pt, -p3, s(p1, p3), t1) * chain(P2, p4, pt, s(p2, p4), t2) *
          (1j / (mdot(pt, pt) - M2)) * ff(pt) ** 2)
    Mu = (chain(P1, p4, pu, s(p1, p4), t1) * chain(P2, pu, -p3, s(p2, p3), t2) *
          (1j / (mdot(pu, pu) - M2)) * ff(pu) ** 2)
    return -1j * (Mt + Mu)


def test_literal_network_equals_the_scalar_chains(run9):
    worst = 0.0
    for p in _points(run9, 200, 1):
        a = run9.tensor_2xs_amplitude(p)
        b = _reduced(run9.cfg, p)
        worst = max(worst, abs(a - b) / abs(b))
    # FP64 rounding of the 6.5 TeV components enters q^2 and t at ~1e-12
    assert worst < 1e-10, worst


def test_amplitude_is_lorentz_invariant(run9):
    # the FP64 boosts of 6.5 TeV beams move t = (pa - p1)^2 by ~1e-8 relative
    for p in _points(run9, 20, 2):
        a = run9.tensor_2xs_amplitude(p)
        for beta in ([0.3, 0.0, 0.0], [0.0, -0.5, 0.2], [0.1, 0.1, 0.6]):
            b = run9.tensor_2xs_amplitude(_boost(p, beta))
            assert abs(b - a) <= 1e-7 * abs(a)
        c, s = np.cos(0.7), np.sin(0.7)
        Rz = np.array([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1.0]])
        Rx = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, c, -s], [0, 0, s, c]])
        for Rm in (Rz, Rx):
            b = run9.tensor_2xs_amplitude(p @ Rm.T)
            assert abs(b - a) <= 1e-7 * abs(a)


def test_bose_symmetry_of_the_pion_pair(run9):
    for p in _points(run9, 40, 3):
        q = p.copy()
        q[[4, 5]] = q[[5, 4]]
        a, b = run9.tensor_2xs_amplitude(p), run9.tensor_2xs_amplitude(q)
        assert abs(a - b) <= 1e-9 * abs(a)


def test_helicity_replicas_and_average(run9):
    rng = np.random.default_rng(5)
    seen = 0
    for _ in range(300):
        x = rng.random(run9.cfg.dim)
        h = run9.bare_amplitudes(x)
        if h is None:
            continue
        assert h.shape == (4,)
        assert np.all(h == h[0])
        W, a2, fl = run9.eval(x[None, :])
        if fl[0] & abi.FLAG_FID_OK:
            seen += 1
            assert abs(a2[0] - (np.abs(h) ** 2).sum() / 4) <= 1e-12 * a2[0]
    assert seen > 20


def _create(cfg):
    lib = abi.load()
    h = C.c_void_p()
    rc = lib.gra_hip_create(C.byref(cfg), C.byref(h))
    msg = lib.gra_hip_last_error(None).decode()
    if rc == abi.GRA_OK:
        lib.gra_hip_destroy(h)
    return rc, msg


@pytest.mark.parametrize("case", ["spin_half", "spin_three", "2xv_stable", "2xs_wide", "2xs_fp32"])
def test_spin_mode_dispatch_rejects_unbuilt_cases(case):
    # MTensorPomeron::ME4 picks its mode from the central spins and throws for
    # other pairs; the stable-pion card must never reach the 2xV formula
    cfg = abi.cfg_default(9)
    if case == "spin_half":
        cfg.central_spin2 = 1
    elif case == "spin_three":
        cfg.central_spin2 = 3
    elif case == "2xv_stable":
        cfg.central_spin2 = 2
    elif case == "2xs_wide":
        cfg = abi.cfg_default(3)
        cfg.central_spin2 = 0
    else:
        cfg.amp_precision = abi.GRA_AMP_FP32
    rc, msg = _create(cfg)
    assert rc == abi.GRA_E_ARG, (case, rc, msg)
    assert "CONTENSOR" in msg or "spin" in msg
