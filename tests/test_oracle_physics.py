"""Pins of the oracle's physics: SolvePz, 2->4 kinematics, the phase-space
weight, the Regge amplitude and the eikonal chain. Each check is one a plausible
slip (dropped term, wrong sign or index, swapped operand) would fail."""
import ctypes as C
import math

import mpmath
import numpy as np
import pytest

from graniitti_amd import abi
import oracle_lib as ol
from conftest import small_tables

L = ol.oracle()
MP = 0.938272081
S = 13000.0 ** 2


def test_solvepz_residuals_and_mpmath_root():
    rng = np.random.default_rng(1)
    for _ in range(200):
        m3 = MP
        m4 = MP if rng.random() < 0.5 else MP * (1 + rng.random())
        pt3, pt4 = rng.uniform(0, 2.5, 2)
        pz5, E5 = rng.uniform(-50, 50), 0.0
        E5 = math.sqrt(pz5 ** 2 + rng.uniform(0.1, 30) ** 2 + rng.uniform(0, 4))
        p3z = L.ora_solvepz(m3, m4, pt3, pt4, pz5, E5, S)
        p4z = -(pz5 + p3z)
        E3 = math.sqrt(m3 ** 2 + pt3 ** 2 + p3z ** 2)
        E4 = math.sqrt(m4 ** 2 + pt4 ** 2 + p4z ** 2)
        assert abs(E3 + E4 + E5 - math.sqrt(S)) < 1e-9 * math.sqrt(S)
        assert p3z > 0 and p4z < 0
        # independent root of the same equations at 40 digits
        mpmath.mp.dps = 40
        f = lambda z: (mpmath.sqrt(m3 ** 2 + pt3 ** 2 + z ** 2) +
                       mpmath.sqrt(m4 ** 2 + pt4 ** 2 + (pz5 + z) ** 2) + E5 - mpmath.sqrt(S))
        z = mpmath.findroot(f, p3z)
        assert abs(float(z) - p3z) < 1e-10 * abs(float(z))


@pytest.fixture(scope="module")
def run1():
    R = ol.OracleRun(abi.cfg_default(1))
    yield R
    R.close()


def m2(v):
    return v[0] ** 2 - v[1] ** 2 - v[2] ** 2 - v[3] ** 2


def test_kinematics_conservation_and_shell(run1):
    rng = np.random.default_rng(2)
    n_ok = 0
    for _ in range(500):
        x = rng.random(8)
        ok, pf, legs, sc = run1.kinematics(x)
        if not ok:
            continue
        n_ok += 1
        beams = np.array([13000.0, 0, 0, 0])
        tot = pf[1] + pf[2] + pf[3] + pf[4]
        np.testing.assert_allclose(tot, beams, rtol=0, atol=2e-8)
        np.testing.assert_allclose(pf[0], pf[3] + pf[4], rtol=0, atol=1e-12)
        assert abs(math.sqrt(m2(pf[1])) - MP) < 1e-6
        for k in (3, 4):
            assert abs(m2(pf[k]) - 0.13957061 ** 2) < 1e-10
        # sampled variables come back: rapidities and central pT from (p1T, p2T, kt)
        y = [0.5 * math.log((pf[k][0] + pf[k][3]) / (pf[k][0] - pf[k][3])) for k in (3, 4)]
        np.testing.assert_allclose(y, [-0.9 + 1.8 * x[6], -0.9 + 1.8 * x[7]], atol=1e-12)
        kt = 5.0 * x[4]
        q = np.array([kt * math.cos(2 * math.pi * x[5]), kt * math.sin(2 * math.pi * x[5])])
        np.testing.assert_allclose(pf[3][1:3] - pf[4][1:3], q, atol=1e-12)
        # Lorentz scalars: t1 from the components, s13 boost invariant
        pa = np.array([6500.0, 0, 0, math.sqrt(6500.0 ** 2 - MP ** 2)])
        assert abs(sc[0] - m2(pa - pf[1])) < 1e-9
        assert sc[0] < 0 and sc[1] < 0
        b = 0.6
        g = 1 / math.sqrt(1 - b * b)
        boost = lambda v: np.array([g * (v[0] - b * v[3]), v[1], v[2], g * (v[3] - b * v[0])])
        assert abs(m2(boost(pf[1] + pf[3])) - sc[4]) < 1e-7 * sc[4]
    assert n_ok > 450


def _jacobian_weight(run, x, h=1e-7):
    """dPhi_4/d^8x from finite differences of (p1T, p2T, p3T, y3, y4)(x)."""
    def vars_of(xx):
        ok, pf, _, _ = run.kinematics(xx)
        assert ok
        y3 = 0.5 * math.log((pf[3][0] + pf[3][3]) / (pf[3][0] - pf[3][3]))
        y4 = 0.5 * math.log((pf[4][0] + pf[4][3]) / (pf[4][0] - pf[4][3]))
        return np.array([pf[1][1], pf[1][2], pf[2][1], pf[2][2], pf[3][1], pf[3][2], y3, y4]), pf
    v0, pf = vars_of(x)
    J = np.zeros((8, 8))
    for k in range(8):
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        J[:, k] = (vars_of(xp)[0] - vars_of(xm)[0]) / (2 * h)
    E1, E2 = pf[1][0], pf[2][0]
    jz = abs(pf[1][3] / E1 - pf[2][3] / E2)
    # dPhi_4 = (2pi)^-8 2^-4 d2p1T d2p2T d2p3T dy3 dy4 / (E1 E2 |p1z/E1 - p2z/E2|)
    return abs(np.linalg.det(J)) * (2 * math.pi) ** -8 / 16 / (E1 * E2 * jz)


def test_phase_space_weight_is_the_4body_jacobian(run1):
    rng = np.random.default_rng(3)
    checked = 0
    while checked < 20:
        x = rng.uniform(0.05, 0.95, 8)
        w = L.ora_phase_space_weight(run1.h, abi.dptr(x))
        if w == 0:
            continue
        ref = _jacobian_weight(run1, x)
        assert abs(w - ref) < 2e-5 * ref, (w, ref)
        checked += 1


def test_regge_eta_linear_close_to_exact():
    # reference testbench3 ReggeEtaLinear vs ReggeEta within 5% near t = 0
    a0, ap = 1.088, 0.20
    for t in np.linspace(-0.5, -0.01, 20):
        e = np.zeros(2)
        el = np.zeros(2)
        L.ora_regge_eta(a0 + ap * t, 1.0, abi.dptr(e))
        L.ora_regge_eta_linear(t, a0, ap, 1.0, abi.dptr(el))
        assert abs(complex(*e) - complex(*el)) < 0.05 * abs(complex(*e))
    # closed form of the signature factor at alpha = 1 + delta: -(1 + e^{-i pi a}) / sin(pi a)
    e = np.zeros(2)
    L.ora_regge_eta(1.25, 1.0, abi.dptr(e))
    want = -(1 + np.exp(-1j * math.pi * 1.25)) / math.sin(math.pi * 1.25)
    assert abs(complex(*e) - want) < 1e-14


def test_regge_propagator_modulus_and_scaling():
    cfg = abi.cfg_default(1)
    out = np.zeros(2)
    for s, t in [(1e3, -0.1), (5e4, -0.5), (2e6, -1.2)]:
        L.ora_regge_prop(C.byref(cfg), s, t, abi.dptr(out))
        e = np.zeros(2)
        L.ora_regge_eta(1.088, 1.0, abi.dptr(e))
        a = 1.088 + 0.2 * t
        assert abs(abs(complex(*out)) - abs(complex(*e)) * s ** a) < 1e-12 * abs(complex(*out))
        # s-scaling: P(2s)/P(s) = 2^alpha(t)
        o2 = np.zeros(2)
        L.ora_regge_prop(C.byref(cfg), 2 * s, t, abi.dptr(o2))
        assert abs(complex(*o2) / complex(*out) - 2 ** a) < 1e-12


def test_amplitude_t_u_symmetry(run1):
    # swapping the two central pions (x6 <-> x7 and kt -> -kt) swaps t <-> u and
    # s13 <-> s14, s24 <-> s23; the pi+ pi- amplitude A_t + A_u is symmetric
    rng = np.random.default_rng(4)
    n = 0
    for _ in range(300):
        x = rng.random(8)
        xs = x.copy()
        xs[6], xs[7] = x[7], x[6]
        xs[5] = (x[5] + 0.5) % 1.0
        h1 = np.zeros(2)
        h2 = np.zeros(2)
        if not (L.ora_bare_amplitudes(run1.h, abi.dptr(x), abi.dptr(h1)) and
                L.ora_bare_amplitudes(run1.h, abi.dptr(xs), abi.dptr(h2))):
            continue
        a1, a2 = complex(*h1), complex(*h2)
        assert abs(a1 - a2) < 1e-7 * abs(a1)
        n += 1
    assert n > 250


def test_pomeron_trajectory_limits():
    cfg = abi.cfg_default(2)
    a0 = L.ora_pom_alpha(C.byref(cfg), -1e-9)
    assert abs(a0 - (1 + cfg.delta_P)) < 1e-6
    # pion-loop insert (KMR, hep-ph/0007359): alpha(t) = 1 + Delta + alpha' t - h(t) with
    # h(t) -> |t| beta_pi^2 / (32 pi^3) (ln(1/m_pi^2) - 8/3) as t -> 0, beta_pi = 2/3 gN_P
    beta = 2.0 / 3.0 * cfg.gN_P
    slope = beta ** 2 / (32 * math.pi ** 3) * (math.log(1 / cfg.mpi_const ** 2) - 8.0 / 3.0)
    for t in (-1e-6, -1e-5):
        gap = 1 + cfg.delta_P + cfg.alpha_P * t - L.ora_pom_alpha(C.byref(cfg), t)
        assert abs(gap / abs(t) - slope) < 2e-3 * slope


@pytest.fixture(scope="module")
def run2():
    R = ol.OracleRun(abi.cfg_default(2), 8)
    yield R
    R.close()


def test_eikonal_cross_sections_ballpark(run2):
    # sigma_tot, sigma_el from the density table as MEikonal::S3CalcXS computes
    # them (src/MEikonal.cc:254-376); TOTEM 13 TeV: 110.6 +- 3.4 mb, 31.0 +- 1.7 mb
    mbt, _ = run2.tables()
    N = 6000
    bmin, bmax = 1e-6, 10.0 / (6.582119514e-25 * 2.99792458e8 * 1e15)
    bt = bmin + (bmax - bmin) / N * np.arange(N + 1)
    step = (bmax - bmin) / 10000
    idx = np.clip(np.floor((bt - bmin) / step).astype(int), 0, 9999)
    x0, x1 = mbt[idx, 0], mbt[idx + 1, 0]
    y0 = mbt[idx, 1] + 1j * mbt[idx, 2]
    y1 = mbt[idx + 1, 1] + 1j * mbt[idx + 1, 2]
    Om = y0 + (bt - x0) * (y1 - y0) / (x1 - x0)
    A = 1j * (1 - np.exp(1j * Om / 2))
    w = np.ones(N + 1)
    w[1:-1:2] = 4
    w[2:-1:2] = 2
    simpson = lambda f: (bmax - bmin) / N / 3 * (w * f).sum()
    gev2mb = (6.582119514e-25 * 2.99792458e8) ** 2 / 1e-28 * 1e3
    stot = 2 * math.pi * simpson(2 * A.imag * bt) * gev2mb
    sel = 2 * math.pi * simpson(np.abs(A) ** 2 * bt) * gev2mb
    assert 90 < stot < 130, stot
    assert 22 < sel < 40, sel


def test_screening_suppresses_and_survival_factor(run2):
    # S^2 = <|M_screened|^2> / <|M_bare|^2> is a suppression in (0.05, 0.5)
    cfg1 = abi.cfg_default(1)
    R1 = ol.OracleRun(cfg1)
    u = np.random.default_rng(5).random((3000, 8))
    W2, _, _ = run2.eval(u)
    W1, _, _ = R1.eval(u)
    S2 = W2.sum() / W1.sum()
    assert 0.05 < S2 < 0.5, S2
    R1.close()


def test_screening_node_amplitudes_come_from_the_table(run2):
    # A_eik at the loop nodes is the linear interpolation of the MSA table, with
    # kt^2 below the first node clamped to MinKT2 = 1e-6 (IArray1D::Interpolate1D)
    _, msa = run2.tables()
    for kt2 in [1e-8, 1e-6, 0.37, 3.0, 8.99]:
        got = run2.msa(kt2)
        a = max(kt2, 1e-6)
        step = (25.0 - 1e-6) / 10000
        i = min(int(math.floor((a - 1e-6) / step)), 9999)
        y0 = complex(msa[i, 1], msa[i, 2])
        y1 = complex(msa[i + 1, 1], msa[i + 1, 2])
        want = y0 + (y1 - y0) / (msa[i + 1, 0] - msa[i, 0]) * (a - msa[i, 0])
        assert abs(got - want) <= 1e-14 * abs(want)


def test_small_table_screening_close_to_full(run2):
    # table and Fourier-Bessel resolution converge: 5000 intervals agree with 10000 at 1e-4
    cfg = small_tables(abi.cfg_default(2), 5000)
    cfg.amp_precision = abi.GRA_AMP_FP64
    R = ol.OracleRun(cfg)
    u = np.random.default_rng(6).random((300, 8))
    Wf, _, _ = run2.eval(u)
    Ws, _, _ = R.eval(u)
    nz = Wf > 0
    rel = np.abs(Ws[nz] / Wf[nz] - 1)
    # pointwise the tree and loop terms can nearly cancel; the sum is what converges
    assert np.median(rel) < 1e-4 and abs(Ws.sum() / Wf.sum() - 1) < 1e-4
    R.close()
