"""Pins of the oracle's form factors and Breit-Wigner forms by what the reference's
own tests and the mathematics fix:

- the Kelly proton form factor F1 (form::F1 with G_E, G_M of Kelly's fit,
  src/MForm.cc:542-645): F1(0) = 1, and the reference's own check that it stays
  within 40% of the dipole Dirac form factor MTensorPomeron::F1_ for t in [-5, 0]
  and equals it at t = 0 to 1e-5 (tests/testbench3.cc:90-124);
- form::CBW (src/MForm.cc:743-828): the fixed-width, running-width and barrier
  forms coincide at the pole, where CBW = i / (M0 Gamma); the fixed form is the
  Lorentzian 1 / ((m^2 - M0^2)^2 + M0^2 Gamma^2) in |CBW|^2 with phase
  arg(-1 / (m^2 - M0^2 + i M0 Gamma)); the barrier form's width scales as
  (q / q0)^(2J+1) M0 / m; the spin-dependent numerators are -(1 - m^2/M0^2) for
  J = 1 (a zero at the pole) and -(7/6 - 4/3 m^2/M0^2 + 2/3 m^4/M0^4) for J = 2;
- PARAM_REGGE::ResonanceFormFactor: a Gaussian in s_hat of width S0^2;
- form::CohFlux: its limits at Q^2 -> 0 for pt >> xi m_p and pt << xi m_p,
  against CODATA alpha and mu_p;
- MRegge::PhotoProp: the signature factor's phase, and the slope
  B = B0 + 2 alpha' log s of |A|^2 with the reference's stated B0 ~ 11 (rho)
  and 4.8 (J/psi), frozen below m = 0.7 GeV."""
import ctypes as C

import numpy as np
import pytest

from graniitti_amd import abi
import oracle_lib as ol

MP = 0.938272081
MU_P = 2.792847337


def kelly(t):
    return ol.oracle().ora_kelly_f1(t, MP)


def dipole_dirac_f1(t):
    # MTensorPomeron::F1_ (src/MTensorPomeron.cc:2470-2473) with GD = 1 / (1 - t / 0.71)^2
    r = t / (4 * MP * MP)
    return (1 - r * MU_P) / (1 - r) / (1 - t / 0.71) ** 2


def test_kelly_f1_at_zero():
    assert kelly(0.0) == 1.0
    assert kelly(0.0) == pytest.approx(dipole_dirac_f1(0.0), rel=1e-5)


def test_kelly_f1_against_the_dipole_form_as_the_reference_tests():
    # testbench3: 100 uniform t in [-5, 0], EPS_KELLY = 0.4 (Catch's Approx epsilon is relative)
    rng = np.random.default_rng(3)
    for t in rng.uniform(-5.0, 0.0, 100):
        k, d = kelly(t), dipole_dirac_f1(t)
        assert abs(k - d) <= 0.4 * abs(d) + 1e-12, (t, k, d)


def test_kelly_f1_is_even_in_q2_sign_and_falls():
    # form::F1 takes |Q^2|; it decreases monotonically from 1 on t in [-5, 0]
    ts = np.linspace(0, -5, 200)
    v = np.array([kelly(t) for t in ts])
    assert np.all(np.diff(v) < 0)
    assert all(kelly(t) == kelly(-t) for t in (-0.1, -1.0, -3.0))


def _res(M, W, J, bw):
    r = abi.GraRes()
    r.mass, r.width, r.J, r.P, r.BW = M, W, J, (-1 if J % 2 else 1), bw
    return r


def cbw(r, m2, mA=0.13957061, mB=0.13957061):
    o = np.zeros(2)
    ol.oracle().ora_cbw(C.byref(r), m2, mA, mB, abi.dptr(o))
    return complex(o[0], o[1])


@pytest.mark.parametrize("M,W,J", [(1.2755, 0.1867, 2), (0.77526, 0.1491, 1), (0.98, 0.08, 0)])
def test_breit_wigner_forms_agree_at_the_pole(M, W, J):
    want = 1j / (M * W)
    for bw in (abi.GRA_BW_FIXED, abi.GRA_BW_RUNNING, abi.GRA_BW_BARRIER):
        z = cbw(_res(M, W, J, bw), M * M)
        assert abs(z - want) <= 1e-12 * abs(want), (bw, z, want)


def test_fixed_width_breit_wigner_is_the_lorentzian():
    M, W = 1.2755, 0.1867
    r = _res(M, W, 2, abi.GRA_BW_FIXED)
    for m in np.linspace(0.4, 3.0, 50):
        m2 = m * m
        z = cbw(r, m2)
        lor = 1.0 / ((m2 - M * M) ** 2 + (M * W) ** 2)
        assert abs(abs(z) ** 2 - lor) <= 1e-12 * lor
        assert abs(z - (-1.0 / complex(m2 - M * M, M * W))) <= 1e-12 * abs(z)


def test_running_and_barrier_widths():
    M, W = 1.2755, 0.1867
    mA = mB = 0.13957061
    q = lambda m2: np.sqrt((m2 - (mA + mB) ** 2) * (m2 - (mA - mB) ** 2)) / (2 * np.sqrt(m2))
    for m in (0.6, 1.0, 1.6, 2.4):
        m2 = m * m
        run = cbw(_res(M, W, 2, abi.GRA_BW_RUNNING), m2)
        assert abs(run - (-1.0 / complex(m2 - M * M, m * W))) <= 1e-12 * abs(run)
        for J in (0, 1, 2):
            bar = cbw(_res(M, W, J, abi.GRA_BW_BARRIER), m2)
            width = W * M * M / m * (q(m2) / q(M * M)) ** (2 * J + 1)
            assert abs(bar - (-1.0 / complex(m2 - M * M, width))) <= 1e-11 * abs(bar)


def test_spin_dependent_numerators():
    M, W = 1.2755, 0.1867
    pole = cbw(_res(M, W, 1, abi.GRA_BW_SPIN), M * M)
    assert pole == 0
    for m in (0.7, 1.1, 2.0):
        m2 = m * m
        den = complex(m2 - M * M, M * W)
        z1 = cbw(_res(M, W, 1, abi.GRA_BW_SPIN), m2)
        assert abs(z1 - (-(1 - m2 / M ** 2)) / den) <= 1e-12 * abs(z1)
        z2 = cbw(_res(M, W, 2, abi.GRA_BW_SPIN), m2)
        n2 = -(7 / 6 - 4 / 3 * m2 / M ** 2 + 2 / 3 * m2 ** 2 / M ** 4)
        assert abs(z2 - n2 / den) <= 1e-12 * abs(z2)
        z0 = cbw(_res(M, W, 0, abi.GRA_BW_SPIN), m2)
        assert abs(z0 - (-1.0) / den) <= 1e-12 * abs(z0)


# ---------------------------------------------------------------------------
# PARAM_REGGE::ResonanceFormFactor (src/MRegge.cc:1233-1235): a Gaussian in
# s_hat around M^2 with width S0^2
# ---------------------------------------------------------------------------
def test_resonance_form_factor_is_a_gaussian_in_s_hat():
    f = ol.oracle().ora_res_ff
    M2, S0 = 1.2755 ** 2, 0.7
    assert f(M2, M2, S0) == 1.0
    assert f(M2 + S0 ** 2, M2, S0) == pytest.approx(np.exp(-1.0), rel=1e-14)
    assert f(M2 - 2 * S0 ** 2, M2, S0) == pytest.approx(np.exp(-4.0), rel=1e-14)
    for d in (0.05, 0.3, 0.9):
        assert f(M2 + d, M2, S0) == pytest.approx(f(M2 - d, M2, S0), rel=1e-14)


# ---------------------------------------------------------------------------
# form::CohFlux (src/MForm.cc:655-690), the coherent photon flux of the proton
# with the reference's 2 -> N sampling factors (1/xi, 1/pt^2, 16 pi^2). Its two
# limits at Q^2 -> 0, where Kelly's G_E = 1 and G_M = mu_p, fix the
# normalisation with CODATA constants that the oracle does not share:
#   pt >> xi m_p: xi pt^2 f / (16 pi^2) -> (alpha/pi) [(1 - xi) + xi^2 mu_p^2 / 4]
#   pt << xi m_p: the electric part vanishes as pt^2, the magnetic part leaves
#                 f -> 4 pi alpha mu_p^2 / (m_p^2 xi)
# ---------------------------------------------------------------------------
ALPHA = 1.0 / 137.035999084


@pytest.mark.parametrize("xi", [1e-3, 0.05, 0.3])
def test_photon_flux_at_large_transverse_momentum(xi):
    pt = 1e4  # (xi m_p / pt)^2 below 1e-9
    f = ol.oracle().ora_coh_flux(xi, -1e-14, pt, MP)
    want = ALPHA / np.pi * ((1 - xi) + xi * xi * MU_P ** 2 / 4)
    assert xi * pt * pt * f / (16 * np.pi ** 2) == pytest.approx(want, rel=1e-8)


@pytest.mark.parametrize("xi", [0.01, 0.2])
def test_photon_flux_at_small_transverse_momentum(xi):
    f = ol.oracle().ora_coh_flux
    want = 4 * np.pi * ALPHA * MU_P ** 2 / (MP * MP * xi)
    assert f(xi, -1e-14, 1e-7 * xi, MP) == pytest.approx(want, rel=1e-8)
    # the approach to the limit is quadratic in pt (the electric part and the
    # expansion of the magnetic part)
    d1 = f(xi, -1e-14, 1e-3 * xi, MP) - want
    d2 = f(xi, -1e-14, 2e-3 * xi, MP) - want
    assert d2 / d1 == pytest.approx(4.0, rel=1e-3)


# ---------------------------------------------------------------------------
# MRegge::PhotoProp, elastic leg (src/MRegge.cc:849-880): the pomeron exchange
# of vector-meson photoproduction. Its phase is the signature factor's, and
# |A|^2 falls as exp(B t) with B = B0(m^2) + 2 alpha' log s (the reference's
# comment: B = B0 + 2 alpha' log(W^2 / W0^2), W0 = 1 GeV), B0 ~ 11 GeV^-2 for the
# rho and 4.8 for the J/psi, frozen below m = 0.7 GeV.
# ---------------------------------------------------------------------------
def photo_prop(cfg, s, t, m2):
    o = np.zeros(2)
    ol.oracle().ora_photo_prop(C.byref(cfg), s, t, m2, abi.dptr(o))
    return complex(o[0], o[1])


def eta(cfg, t):
    o = np.zeros(2)
    ol.oracle().ora_regge_eta_linear(t, cfg.regge_a0[0], cfg.regge_ap[0], cfg.regge_sgn[0],
                                     abi.dptr(o))
    return complex(o[0], o[1])


def slope_b(cfg, s, m2, t=-0.2, h=1e-4):
    def lg(tt):
        return np.log(abs(photo_prop(cfg, s, tt, m2)) ** 2 / abs(eta(cfg, tt)) ** 2)
    return (lg(t + h) - lg(t - h)) / (2 * h)


def test_photo_prop_phase_is_the_signature_factor():
    cfg = abi.cfg_default(6)
    for s, t in [(100.0, -0.1), (5e3, -0.5), (1e5, -1.2)]:
        a, e = photo_prop(cfg, s, t, 0.775 ** 2), eta(cfg, t)
        assert np.angle(a) == pytest.approx(np.angle(e), abs=1e-12)


def test_photo_prop_slope():
    cfg = abi.cfg_default(6)
    ap = cfg.regge_ap[0]
    s1, s2 = 1e2, 1e4
    b_rho = slope_b(cfg, s1, 0.775 ** 2) - 2 * ap * np.log(s1)
    b_jpsi = slope_b(cfg, s1, 3.0969 ** 2) - 2 * ap * np.log(s1)
    assert 10.5 < b_rho < 12.0
    assert 4.6 < b_jpsi < 5.0
    # the energy dependence: B grows by 2 alpha' log(s2 / s1)
    assert slope_b(cfg, s2, 0.775 ** 2) - slope_b(cfg, s1, 0.775 ** 2) == pytest.approx(
        2 * ap * np.log(s2 / s1), rel=1e-5)
    # frozen below m = 0.7 GeV
    assert slope_b(cfg, s1, 0.1) == pytest.approx(slope_b(cfg, s1, 0.49), rel=1e-12)
