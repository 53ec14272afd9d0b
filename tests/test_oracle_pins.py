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
  J = 1 (a zero at the pole) and -(7/6 - 4/3 m^2/M0^2 + 2/3 m^4/M0^4) for J = 2.
The EPA photon flux normalisation and PhotoProp's slope stay unpinned."""
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
