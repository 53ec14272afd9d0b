"""Veto domains (VETOCUTS; MProcess::VetoCuts and FindVetoCuts,
src/MProcess.cc:1069-1082, 1157-1175) in the oracle and in the C ABI's
validation, on the CPU.

The oracle's veto is checked against a restatement written here from the
oracle's own final-state four-vectors (ora_kinematics): a point is vetoed iff
one of its central final states (the two central particles, or the four decay
products of a vector-meson pair) has pt and pseudorapidity inside a domain, and
a vetoed point's weight is zero while every other point keeps the weight of the
same configuration without vetoes, bit for bit. The reference holds no vectors
for this; the domains below are chosen so that both outcomes occur often."""
import ctypes as C

import numpy as np
import pytest

from graniitti_amd import abi
import oracle_lib as ol

# (eta_min, eta_max, pt_min, pt_max): a central window above 0.25 GeV and a
# forward strip below 0.4 GeV, inside the fiducial |eta| < 0.9 of the presets
DOMAINS = [(-0.3, 0.2, 0.25, 1e5), (0.5, 0.9, 0.0, 0.4)]


def with_veto(base, domains):
    cfg = abi.GraCfg.from_buffer_copy(base)
    cfg.veto_active = 1
    cfg.n_veto = len(domains)
    for d, (e0, e1, p0, p1) in enumerate(domains):
        cfg.veto_eta_min[d] = e0
        cfg.veto_eta_max[d] = e1
        cfg.veto_pt_min[d] = p0
        cfg.veto_pt_max[d] = p1
    return cfg


def final_states(cfg, pf, legs):
    return legs.reshape(4, 4) if cfg.m_daughter > 0.0 else pf[3:5]


def in_domain(p, domains):
    pt = np.hypot(p[1], p[2])
    pm = np.sqrt(p[1] ** 2 + p[2] ** 2 + p[3] ** 2)
    eta = 0.5 * np.log((pm + p[3]) / (pm - p[3]))
    return any(p0 <= pt <= p1 and e0 <= eta <= e1 for e0, e1, p0, p1 in domains)


@pytest.mark.parametrize("preset", [1, 3])
def test_oracle_veto_matches_final_state_restatement(preset):
    base = abi.cfg_default(preset)
    base.screening = 0
    base.amp_precision = abi.GRA_AMP_FP64
    vcfg = with_veto(base, DOMAINS)
    rng = np.random.default_rng(41 + preset)
    u = rng.random((3000, base.dim))
    W0, _, f0 = ol.OracleRun(base).eval(u)
    W1, _, f1 = ol.OracleRun(vcfg).eval(u)
    O = ol.OracleRun(base)
    hit = np.zeros(len(u), dtype=bool)
    kin = np.zeros(len(u), dtype=bool)
    for i in range(len(u)):
        ok, pf, legs, _ = O.kinematics(u[i])
        kin[i] = bool(ok)
        if ok:
            hit[i] = any(in_domain(p, DOMAINS) for p in final_states(base, pf, legs))
    assert np.array_equal(kin, (f0 & abi.FLAG_KIN_OK) != 0)
    veto_ok = (f1 & abi.FLAG_VETO_OK) != 0
    # VETO_OK is clear exactly at the points with a final state in a domain
    assert np.array_equal(veto_ok, ~hit)
    assert np.all((f0 & abi.FLAG_VETO_OK) != 0)
    # the other flags do not change
    other = abi.FLAG_KIN_OK | abi.FLAG_FID_OK | abi.FLAG_AMP_OK | abi.FLAG_INF
    assert np.array_equal(f0 & other, f1 & other)
    # vetoed points weigh nothing; the rest are untouched
    assert np.all(W1[hit] == 0.0)
    assert np.array_equal(W1[~hit], W0[~hit])
    fid = (f0 & abi.FLAG_FID_OK) != 0
    n_hit_valid = int(np.sum(hit & fid & (W0 > 0)))
    n_kept_valid = int(np.sum(~hit & fid & (W0 > 0)))
    print(f"preset {preset}: {n_hit_valid} valid points vetoed, {n_kept_valid} kept")
    assert n_hit_valid > 30 and n_kept_valid > 30


def test_inactive_veto_changes_nothing():
    base = abi.cfg_default(1)
    vcfg = with_veto(base, DOMAINS)
    vcfg.veto_active = 0
    u = np.random.default_rng(3).random((500, base.dim))
    W0, _, f0 = ol.OracleRun(base).eval(u)
    W1, _, f1 = ol.OracleRun(vcfg).eval(u)
    assert np.array_equal(W0, W1) and np.array_equal(f0, f1)


def test_domain_covering_everything_vetoes_every_valid_point():
    base = abi.cfg_default(1)
    vcfg = with_veto(base, [(-30.0, 30.0, 0.0, 1e6)])
    u = np.random.default_rng(4).random((500, base.dim))
    W1, _, f1 = ol.OracleRun(vcfg).eval(u)
    kin = (f1 & abi.FLAG_KIN_OK) != 0
    assert kin.sum() > 100
    assert np.all((f1[kin] & abi.FLAG_VETO_OK) == 0)
    assert np.all(W1 == 0.0)


def _validate(cfg):
    """validate() through an entry point that needs no GPU"""
    lib = abi.load()
    buf = C.create_string_buffer(256)
    rc = lib.gra_hip_eikonal_cache_name(C.byref(cfg), 0, buf, len(buf))
    return rc, lib.gra_hip_last_error(None).decode()


@pytest.mark.parametrize("field,value", [("veto_eta_max", -0.3), ("veto_pt_max", 0.1)])
def test_empty_veto_domain_is_rejected(field, value):
    # aux::AssertCut(.., dothrow = true) on each domain, src/MGraniitti.cc:950-956
    cfg = with_veto(abi.cfg_default(2), DOMAINS)
    assert _validate(cfg)[0] == abi.GRA_OK
    getattr(cfg, field)[0] = value
    rc, msg = _validate(cfg)
    assert rc == abi.GRA_E_ARG and "VETOCUTS::0" in msg, msg
    # the same domain is not checked while the veto is off
    cfg.veto_active = 0
    assert _validate(cfg)[0] == abi.GRA_OK


def test_too_many_veto_domains_are_rejected():
    cfg = with_veto(abi.cfg_default(2), DOMAINS)
    cfg.n_veto = abi.GRA_MAX_VETO + 1
    rc, msg = _validate(cfg)
    assert rc == abi.GRA_E_ARG and "n_veto" in msg, msg
