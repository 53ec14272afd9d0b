"""Unequal beam energies (MContinuum::BNBuildKin, src/MContinuum.cc:386-395): the
<C> kinematics are built in the centre-of-mass frame and boosted to the lab.
Pinned on the CPU by what the boost must preserve:

- the event weight is Lorentz invariant: with the fiducial cuts off, a point of
  the asymmetric run has the weight of the same point of a symmetric run at the
  same sqrt(s) (|M|^2 depends on invariants only, and the phase-space weight's
  pt1 pt2 / (E1 E2 |p1z/E1 - p2z/E2|) is invariant under boosts along z);
- the final state carries the lab beams' four-momentum;
- the lab momenta are the boosted centre-of-mass momenta (pt unchanged,
  rapidities shifted by the beams' rapidity);
and the validation keeps unequal beams to the unscreened <C> class."""
import ctypes as C

import numpy as np
import pytest
from scipy.optimize import brentq

from graniitti_amd import abi
import oracle_lib as ol


def asym_cfg(preset, E1=8000.0):
    sym = abi.cfg_default(preset)
    m = sym.m_beam
    s = 2 * m * m + 2 * (sym.beam_E1 * sym.beam_E2 + np.sqrt(sym.beam_E1 ** 2 - m * m) *
                         np.sqrt(sym.beam_E2 ** 2 - m * m))
    p1 = np.sqrt(E1 * E1 - m * m)
    E2 = brentq(lambda e: 2 * m * m + 2 * (E1 * e + p1 * np.sqrt(e * e - m * m)) - s, m * 1.0001,
                E1)
    cfg = abi.GraCfg.from_buffer_copy(sym)
    cfg.beam_E1, cfg.beam_E2 = E1, E2
    return sym, cfg


def beams(cfg):
    m = cfg.m_beam
    b1 = np.array([cfg.beam_E1, 0, 0, np.sqrt(cfg.beam_E1 ** 2 - m * m)])
    b2 = np.array([cfg.beam_E2, 0, 0, -np.sqrt(cfg.beam_E2 ** 2 - m * m)])
    return b1, b2


@pytest.mark.parametrize("preset", [1, 3])
def test_weight_is_invariant_under_the_boost(preset):
    sym, asym = asym_cfg(preset)
    for c in (sym, asym):
        c.fid_active = 0
        c.screening = 0
        c.amp_precision = abi.GRA_AMP_FP64
    u = np.random.default_rng(7).random((2000, sym.dim))
    W0, _, f0 = ol.OracleRun(sym).eval(u)
    W1, _, f1 = ol.OracleRun(asym).eval(u)
    assert np.array_equal(f0, f1)
    ok = W0 > 0
    assert ok.sum() > 500
    rel = np.abs(W1[ok] - W0[ok]) / W0[ok]
    print(f"preset {preset}: {ok.sum()} points, max rel {rel.max():.2e}")
    # the invariants are formed from TeV-scale components in either frame (t from
    # p_beam - p at 8 TeV against 6.5 TeV): rounding of those subtractions, not a
    # frame dependence, leaves 1e-10 (pi pi) to 1.4e-9 (rho rho, four decay legs)
    assert rel.max() < (1e-9 if preset == 1 else 1e-8)


def test_lab_momenta_are_the_boosted_centre_of_mass_momenta():
    sym, asym = asym_cfg(1)
    b1, b2 = beams(asym)
    y_beams = 0.5 * np.log((b1[0] + b2[0] + b1[3] + b2[3]) / (b1[0] + b2[0] - b1[3] - b2[3]))
    O0, O1 = ol.OracleRun(sym), ol.OracleRun(asym)
    u = np.random.default_rng(8).random((300, sym.dim))
    n = 0
    for x in u:
        ok0, pf0, _, _ = O0.kinematics(x)
        ok1, pf1, _, _ = O1.kinematics(x)
        assert ok0 == ok1
        if not ok0:
            continue
        n += 1
        tot = pf1[1] + pf1[2] + pf1[3] + pf1[4]
        assert np.abs(tot - (b1 + b2)).max() < 1e-9 * asym.beam_E1
        for k in (1, 2, 3, 4):
            assert np.hypot(pf1[k, 1], pf1[k, 2]) == pytest.approx(
                np.hypot(pf0[k, 1], pf0[k, 2]), rel=1e-12)
            rap = lambda p: 0.5 * np.log((p[0] + p[3]) / (p[0] - p[3]))
            if k >= 3:  # the forward protons' rapidities lose digits at 13 TeV
                assert rap(pf1[k]) - rap(pf0[k]) == pytest.approx(y_beams, abs=1e-9)
    assert n > 100


def _validate(cfg):
    """validate() through an entry point that needs no GPU"""
    lib = abi.load()
    buf = C.create_string_buffer(256)
    rc = lib.gra_hip_eikonal_cache_name(C.byref(cfg), 0, buf, len(buf))
    return rc, lib.gra_hip_last_error(None).decode()


def test_unequal_beams_are_kept_to_the_unscreened_c_class():
    _, cfg = asym_cfg(1)
    assert _validate(cfg)[0] == abi.GRA_OK
    _, cfg = asym_cfg(2)  # screened
    rc, msg = _validate(cfg)
    assert rc == abi.GRA_E_ARG and "asymmetric" in msg, msg
    _, cfg = asym_cfg(4)  # <F>
    rc, msg = _validate(cfg)
    assert rc == abi.GRA_E_ARG and "asymmetric" in msg, msg
