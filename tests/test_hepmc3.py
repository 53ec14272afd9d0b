"""HepMC3 ASCII output (host only): events built from the oracle's FP64
kinematics are written through the C ABI and read back.

HepMC3 itself is absent here, so the file is checked against the published
Asciiv3 layout and against what the event record must satisfy: the vertex
graph of MProcess::CommonRecord (src/MProcess.cc:1588-1706), four-momentum and
charge conservation at every vertex, exact FP64 round trip of the final
state, E-line counts equal to the particles and vertices present, and the
cross section in picobarn (src/MGraniitti.cc:1696-1701)."""
import ctypes as C
import math

import numpy as np
import pytest

from graniitti_amd import abi
from graniitti_amd.process import HepMC3Writer, record_pdg

import hepmc3_read
import oracle_lib as ol

CHARGE = {2212: 1, 211: 1, -211: -1, 113: 0, 99: 0, 90: 0}


def oracle_events(preset, n, seed):
    cfg = abi.cfg_default(preset)
    cfg.screening = 0
    run = ol.OracleRun(cfg, nthreads=1)
    rng = np.random.default_rng(seed)
    n_p = 8 if cfg.m_daughter > 0 else 4
    out = []
    while len(out) < n:
        ok, pf, legs, _ = run.kinematics(rng.random(cfg.dim))
        if not ok:
            continue
        p = [pf[1], pf[2], pf[3], pf[4]]
        if n_p == 8:
            p += [legs[0][0], legs[0][1], legs[1][0], legs[1][1]]
        out.append(np.array(p))
    run.close()
    return cfg, np.array(out)


def vec(q):
    """(E, px, py, pz) of a parsed particle."""
    return np.array([q.p[3], q.p[0], q.p[1], q.p[2]])


def check_file(path, cfg, p, f, sigma, sigma_err, weighted, first_number=0, tol=1e-9):
    h = hepmc3_read.read(path)
    assert h.footer and h.version.startswith("3.")
    assert h.tools[0] == ["GRANIITTI (TUNE0)", "1.0", "Generator"]
    assert len(h.events) == len(p)
    n_p = p.shape[1]
    ids = record_pdg(cfg)
    pz1 = math.sqrt(cfg.beam_E1**2 - cfg.m_beam**2)
    for k, ev in enumerate(h.events):
        assert ev.number == first_number + k
        assert ev.units == ("GEV", "MM")
        assert ev.weights == [f[k] if weighted else 1.0]
        xs, xs_err, acc, att = ev.attributes["GenCrossSection"].split()
        assert float(xs) == pytest.approx(sigma * 1e12, rel=1e-8)
        assert float(xs_err) == pytest.approx(sigma_err * 1e12, rel=1e-8)
        assert (acc, att) == ("-1", "-1")
        # counts on the E line are those of the record
        g = ev.vertex_graph()
        assert ev.n_particles == len(ev.particles) == (13 if n_p == 8 else 9)
        assert ev.n_vertices == len(g) == (6 if n_p == 8 else 4)
        assert [q.id for q in ev.particles] == list(range(1, ev.n_particles + 1))
        P = {q.id: q for q in ev.particles}
        # beams: status 4 along +z / -z
        assert (P[1].status, P[4].status) == (4, 4)
        assert (P[1].pdg, P[4].pdg) == (ids.beam[0], ids.beam[1])
        assert vec(P[1]) == pytest.approx([cfg.beam_E1, 0, 0, pz1], abs=0)
        # the CommonRecord topology
        assert g[-3] == ([3, 6], [7])
        assert sorted(g[("end", 1)][1]) == [2, 3] and sorted(g[("end", 4)][1]) == [5, 6]
        assert (P[3].pdg, P[6].pdg, P[7].pdg) == (99, 99, 90)
        assert {P[i].status for i in (3, 6, 7)} == {81}
        # final state: exact round trip of the input momenta, in gra_event order
        finals = [2, 5] + ([9, 10, 12, 13] if n_p == 8 else [8, 9])
        assert all(P[i].status == 1 for i in finals)
        assert sum(1 for q in ev.particles if q.status == 1) == len(finals)
        order = [2, 5, 8, 11, 9, 10, 12, 13] if n_p == 8 else [2, 5, 8, 9]
        for slot, pid in enumerate(order):
            assert np.array_equal(vec(P[pid]), p[k, slot])
        if n_p == 8:
            assert (P[8].status, P[11].status) == (2, 2)
            assert sorted(g[("end", 8)][1]) == [9, 10] and sorted(g[("end", 11)][1]) == [12, 13]
        # conservation of four-momentum and charge at every vertex
        for key, (inc, outg) in g.items():
            d = sum(vec(P[i]) for i in inc) - sum(vec(P[i]) for i in outg)
            assert np.max(np.abs(d)) < tol * cfg.beam_E1, (k, key, d)
            assert sum(CHARGE[P[i].pdg] for i in inc) == sum(CHARGE[P[i].pdg] for i in outg)
        # mass column: the 4-momentum's signed mass
        for q in ev.particles:
            v = vec(q)
            m2 = v[0] ** 2 - v[1] ** 2 - v[2] ** 2 - v[3] ** 2
            assert q.mass == pytest.approx(math.copysign(math.sqrt(abs(m2)), m2),
                                           rel=1e-6, abs=1e-6 * v[0])
        assert P[3].mass < 0 and P[6].mass < 0  # space-like propagators (t < 0)
    return h


@pytest.mark.parametrize("preset", [1, 3])
def test_written_events_read_back(tmp_path, preset):
    cfg, p = oracle_events(preset, 40, seed=7 + preset)
    f = np.linspace(1e-9, 2e-9, len(p))
    path = tmp_path / "ev.hepmc3"
    with HepMC3Writer(path, cfg) as w:
        w.write(p[:25], f[:25], 1.5e-6, 2e-9)
        w.write(p[25:], f[25:], 1.5e-6, 2e-9)
    check_file(path, cfg, p, f, 1.5e-6, 2e-9, weighted=False)


def test_weighted_events_carry_f(tmp_path):
    cfg, p = oracle_events(1, 10, seed=3)
    f = np.geomspace(1e-12, 1e-6, len(p))
    path = tmp_path / "w.hepmc3"
    with HepMC3Writer(path, cfg, card="card \\ name") as w:
        w.write(p, f, 3e-7, 1e-9, weighted=True)
    h = check_file(path, cfg, p, f, 3e-7, 1e-9, weighted=True)
    assert h.tools[1] == ["card \\ name", "1.0", "Steering card"]


def test_writer_rejects_bad_events_without_writing(tmp_path):
    lib = abi.load()
    cfg, p = oracle_events(1, 3, seed=5)
    path = tmp_path / "bad.hepmc3"
    with HepMC3Writer(path, cfg) as w:
        w.write(p[:1], [1.0], 1e-6, 1e-9)
        buf = (abi.GraEvent * 2)()
        a = np.ctypeslib.as_array(buf)
        a["p"][:, :4, :] = p[1:3]
        a["n_p"][0], a["n_p"][1] = 4, 5
        rc = lib.gra_hepmc3_write(w.w, C.byref(cfg), C.byref(w.pdg), buf, 2, 1e-6, 1e-9, 0)
        assert rc == abi.GRA_E_ARG
        assert b"n_p = 5" in lib.gra_hip_last_error(None)
        bad = abi.cfg_default(1)
        bad.abi_size = 8
        assert lib.gra_hepmc3_write(w.w, C.byref(bad), C.byref(w.pdg), buf, 1, 1e-6, 1e-9,
                                    0) == abi.GRA_E_ARG
        w.write(p[2:3], [1.0], 1e-6, 1e-9)
    h = hepmc3_read.read(path)
    assert [e.number for e in h.events] == [0, 1]  # nothing of the rejected calls was written
    out = C.c_void_p()
    assert lib.gra_hepmc3_open(str(tmp_path / "no" / "such" / "dir").encode(), b"g", b"1", None,
                               C.byref(out)) == abi.GRA_E_ARG
    assert not out.value
