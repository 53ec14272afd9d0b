"""Unweighted event generation on the GPU (MGraniitti::SaveEvent hit-or-miss,
src/MGraniitti.cc:1651-1726): the accepted set against a host restatement of
the hit-or-miss on the same points, and the event records' kinematics."""
import ctypes as C
import sys

import numpy as np
import pytest

from graniitti_amd import abi
from graniitti_amd.process import HepMC3Writer, MContinuum, MGraniitti, vegas_cfg
import oracle_lib as ol
from conftest import small_tables
from test_hepmc3 import check_file

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def minkowski(p):
    return p[..., 0] ** 2 - (p[..., 1:] ** 2).sum(-1)


def test_hit_or_miss_matches_host_restatement():
    cfg = abi.cfg_default(1)
    say("case: generate 2^16 calls, preset 1, accepted set vs host hit-or-miss")
    n, seed, it = 1 << 16, 4242, 3
    with MContinuum(cfg) as P:
        G = MGraniitti(P)
        G.Integrate(vegas_cfg(1 << 17, 3, burnin=3, seed=7))
        # the same points on the host: Philox uniforms, the grid map, the GPU integrand
        u = np.array([ol.point_uniforms(seed, it, i, cfg.dim) for i in range(n)])
        x, vw, bins = P.vegas_map(G.grid, u)
        W, a2, fl = P.eval(x)
        f = W * vw
        valid = (fl & 7) == 7
        maxf = float(np.quantile(f[valid], 0.999))  # some points overflow
        buf = (abi.GraEvent * n)()
        st = abi.GraGenStats()
        P._chk(P.lib.gra_hip_generate(P.h, C.byref(G.grid), n, seed, it, maxf, buf, n,
                                      C.byref(st)))
        ev = np.ctypeslib.as_array(buf)[: st.stored]
    U = np.array([ol.point_uniforms(seed, 0xC0000000 | it, i, 1)[0] for i in range(n)])
    want = valid & (U < f / maxf) & ~(f > maxf)
    say(f"trials {st.trials} valid {st.valid} accepted {st.accepted} overflow {st.overflow}")
    assert st.trials == n and st.valid == valid.sum()
    assert st.overflow == (f > maxf).sum() and st.overflow > 0
    assert st.accepted == want.sum() == st.stored
    np.testing.assert_array_equal(ev["index"], np.nonzero(want)[0])
    np.testing.assert_allclose(ev["f"], f[want], rtol=1e-13)


def test_generation_into_pinned_and_pageable_buffers_agree():
    # the same pass into pageable memory (through the handle's pinned staging)
    # and into pinned memory from gra_hip_host_alloc (written by the kernel), with a
    # max_events cut that keeps the accepted points of smallest index; a pass
    # of 2^20 + 37 points covers a partial last group of 64 and several tiles
    cfg = abi.cfg_default(1)
    n, seed, it = (1 << 20) + 37, 99, 5
    say(f"case: generate {n} calls, preset 1, pinned vs pageable event buffers")
    with MContinuum(cfg) as P:
        G = MGraniitti(P)
        G.Integrate(vegas_cfg(1 << 17, 3, burnin=3, seed=7))
        maxf = G.result.maxf
        out = []
        for cap in (1 << 16, 500):
            page = (abi.GraEvent * cap)()
            with abi.PinnedEvents(cap) as pin:
                for buf in (page, pin.buf):
                    arr = np.ctypeslib.as_array(buf)
                    arr["index"] = np.uint64(0xDEADBEEF)  # entries past stored stay as they were
                    st = abi.GraGenStats()
                    P._chk(P.lib.gra_hip_generate(P.h, C.byref(G.grid), n, seed, it, maxf, buf,
                                                  cap, C.byref(st)))
                    assert np.all(arr["index"][st.stored:] == 0xDEADBEEF)
                    out.append((cap, st.trials, st.accepted, st.stored, arr[: st.stored].copy()))
    say("accepted " + ", ".join(f"{a} (stored {s}, cap {c})" for c, _, a, s, _ in out))
    for cap, trials, acc, stored, ev in out:
        assert trials == n and stored == min(acc, cap) and acc > 500
        assert np.all(np.diff(ev["index"].astype(np.int64)) > 0)
    np.testing.assert_array_equal(out[0][4], out[1][4])
    np.testing.assert_array_equal(out[2][4], out[3][4])
    np.testing.assert_array_equal(out[2][4], out[0][4][:500])


def check_records(p, cfg):
    E = cfg.beam_E1
    pz = np.sqrt(E * E - cfg.m_beam ** 2)
    beams = np.array([2 * E, 0.0, 0.0, 0.0])
    tot = p[:, 0] + p[:, 1] + p[:, 2] + p[:, 3]
    assert np.abs(tot - beams).max() < 1e-6 * E
    assert np.allclose(minkowski(p[:, 0]), cfg.m_beam ** 2, atol=1e-6 * E)
    assert np.allclose(minkowski(p[:, 1]), cfg.m_beam ** 2, atol=1e-6 * E)
    assert np.all(p[:, 0, 3] > 0) and np.all(p[:, 1, 3] < 0)
    del pz


def test_events_of_the_pion_pair_process():
    cfg = abi.cfg_default(1)
    say("case: 4000 unweighted events, preset 1")
    with MContinuum(cfg) as P:
        G = MGraniitti(P)
        G.Integrate(vegas_cfg(1 << 18, 4, burnin=3, seed=11))
        p, f, tot = G.Generate(4000, calls=1 << 20, seed=5)
    say(f"events {len(p)} counters {tot}")
    assert p.shape == (4000, 4, 4)
    check_records(p, cfg)
    assert np.allclose(minkowski(p[:, 2]), cfg.m_central ** 2, rtol=1e-9)
    # fiducial cuts hold for every event: |eta| < 0.9, pT > 0.15
    for k in (2, 3):
        pt = np.hypot(p[:, k, 1], p[:, k, 2])
        eta = np.arcsinh(p[:, k, 3] / pt)
        assert np.all(pt > cfg.fid_pt_min) and np.all(np.abs(eta) < cfg.fid_eta_max)


def test_events_of_the_tensor_pomeron_process(tmp_path):
    cfg = small_tables(abi.cfg_default(3), 1000)
    say("case: 300 unweighted events, preset 3 (rho rho -> 4 pi, FP32 amplitude, screened)")
    with MContinuum(cfg) as P:
        G = MGraniitti(P)
        G.Integrate(vegas_cfg(1 << 17, 3, burnin=3, seed=13))
        p, f, tot = G.Generate(300, calls=1 << 18, seed=6)
        sigma, sigma_err = G.GetXS()
    say(f"events {len(p)} counters {tot}")
    assert p.shape == (300, 8, 4)
    check_records(p, cfg)
    # rho -> pi+ pi-: daughters sum to their parent, pion mass shell, cuts on pions
    assert np.abs(p[:, 4] + p[:, 5] - p[:, 2]).max() < 1e-9 * cfg.beam_E1
    assert np.abs(p[:, 6] + p[:, 7] - p[:, 3]).max() < 1e-9 * cfg.beam_E1
    assert np.allclose(minkowski(p[:, 4:]), cfg.m_daughter ** 2, atol=1e-8)
    pt = np.hypot(p[:, 4:, 1], p[:, 4:, 2])
    eta = np.arcsinh(p[:, 4:, 3] / pt)
    assert np.all(pt > cfg.fid_pt_min) and np.all(np.abs(eta) < cfg.fid_eta_max)
    # the same events through the HepMC3 writer: exact round trip, conservation per vertex
    path = tmp_path / "rhorho.hepmc3"
    with HepMC3Writer(path, cfg) as w:
        w.write(p, f, sigma, sigma_err)
    check_file(path, cfg, p, f, sigma, sigma_err, weighted=False, tol=1e-6)
