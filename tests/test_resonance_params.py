"""Presets 6-8 against the resonance inputs of the reference's data files
(tests/golden/resonance_params.json, written by tests/golden/make_resonance_params.py
from ALICECUTS.json, RES_<name>.json, BRANCHING.json and GENERAL.json)."""
import json
import os

import numpy as np

from graniitti_amd import abi

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "resonance_params.json")))


def check_res(q, want):
    assert q.pdg == want["pdg"]
    assert q.mass == want["M"] and q.width == want["W"]
    assert q.J == want["J"] and q.P == want["P"] and q.BW == want["BW"]
    assert q.g_A == want["g_A"] and q.g_phi == want["g_phi"] and q.g_FF == want["g_FF"]
    assert q.BR == want["BR"]
    (l, s, are, aim), = want["alpha_ls"]
    assert (l, s) == (q.J, 0) and (q.alpha_re, q.alpha_im) == (are, aim)
    n = 2 * q.J + 1
    if q.J > 0:
        assert np.array_equal(np.array(q.rho_re[:n * n]).reshape(n, n), np.array(want["rho_real"]))
        assert np.array_equal(np.array(q.rho_im[:n * n]).reshape(n, n), np.array(want["rho_imag"]))


def test_res_con_preset_matches_the_reference_inputs():
    cfg = abi.cfg_default(6)
    assert FIX["process"].startswith("PP[RES+CON]<C>")
    assert cfg.process == abi.GRA_PROC_RES_CON and cfg.phase_space == abi.GRA_PS_C
    spin = FIX["spin"]
    assert cfg.spin_gen == int(spin["SPINGEN"]) and cfg.spin_dec == int(spin["SPINDEC"])
    assert cfg.spin_frame == {"CM": 0, "HX": 1, "CS": 2, "GJ": 3}[spin["FRAME"]]
    assert cfg.regge_omega == FIX["omega"]
    # PROC_17 iterates lts.RESONANCES, a std::map keyed by name: sorted order
    names = sorted(FIX["resonances"])
    assert cfg.n_res == len(names)
    for r, name in enumerate(names):
        check_res(cfg.res[r], FIX["resonances"][name])


def test_jw_frames_preset_is_the_f2():
    cfg = abi.cfg_default(8)
    assert cfg.process == abi.GRA_PROC_RES and cfg.phase_space == abi.GRA_PS_F and cfg.n_res == 1
    check_res(cfg.res[0], FIX["resonances"]["f2_1270"])
    assert cfg.fid_active == 0
