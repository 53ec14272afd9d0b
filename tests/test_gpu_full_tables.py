"""GPU (HIP, gfx950) against the FP64 oracle at the presets' full eikonal
discretisation: MBT and MSA tables of 10001 rows from 10001-node Fourier-Bessel
quadratures (modeldata/NUMERICS.json; MEikonal::S3Constructor,
src/MEikonal.cc:46-132), as the benchmark runs them. The other GPU parity tests
use 1000-row tables to keep the oracle fast; these check the same bounds at the
full size. Also the CheckInfNan run guard (include/Graniitti/MProcess.h:317-340)
on a configuration whose weights are NaN."""
import ctypes as C
import sys

import numpy as np
import pytest

from graniitti_amd import abi
from graniitti_amd.process import MContinuum, MGraniitti, vegas_cfg
import oracle_lib as ol
from test_gpu_parity import compare, points

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_full_eikonal_tables_match_oracle():
    cfg = abi.cfg_default(2)
    say("case: eikonal tables at 10001 rows (preset 2)")
    with MContinuum(cfg) as P:
        node, mbt, msa = P.eikonal_tables()
    R = ol.OracleRun(cfg)
    ombt, omsa = R.tables()
    assert mbt.shape == (10001, 3) and msa.shape == (10001, 3)
    np.testing.assert_array_equal(mbt[:, 0], ombt[:, 0])
    np.testing.assert_array_equal(msa[:, 0], omsa[:, 0])
    for name, got, want in [("MBT", mbt, ombt), ("MSA", msa, omsa)]:
        scale = np.abs(want[:, 1] + 1j * want[:, 2]).max()
        err = np.abs((got[:, 1] + 1j * got[:, 2]) - (want[:, 1] + 1j * want[:, 2])).max()
        say(f"{name}: max |GPU - oracle| / max |value| = {err / scale:.2e}")
        assert err <= 1e-12 * scale


def test_full_tables_screened_continuum_fp64():
    cfg = abi.cfg_default(2)
    cfg.amp_precision = abi.GRA_AMP_FP64
    say("case: preset 2 in FP64, full tables, 2000 points, seed 31")
    u = points(2000, cfg.dim, 31)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    compare(Wg, Wo, fg, fo, 1e-10, "preset 2 fp64 full tables")


def test_full_tables_screened_continuum_fp32():
    cfg = abi.cfg_default(2)
    say("case: preset 2 (FP32 amplitude), full tables, 4000 points, seed 32")
    u = points(4000, cfg.dim, 32)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    ref = abi.cfg_default(2)
    ref.amp_precision = abi.GRA_AMP_FP64
    Wo, ao, fo = ol.OracleRun(ref).eval(u)
    rel = compare(Wg, Wo, fg, fo, 1e-4, "preset 2 fp32 full tables")
    say(f"preset 2 fp32 rel err: median {np.median(rel):.2e} p99 {np.quantile(rel, 0.99):.2e}")
    assert np.median(rel) < 4e-6


def test_full_tables_screened_res_con():
    cfg = abi.cfg_default(7)
    say("case: preset 7 (screened PP[RES+CON]), full tables, 500 points, seed 33")
    u = points(500, cfg.dim, 33)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    compare(Wg, Wo, fg, fo, 1e-10, "preset 7 full tables")


def test_full_tables_screened_tensor_2xs():
    cfg = abi.cfg_default(10)
    say("case: preset 10 (screened tensor pomeron pi+ pi-), full tables, 500 points, seed 34")
    u = points(500, cfg.dim, 34)
    with MContinuum(cfg) as P:
        Wg, ag, fg = P.eval(u)
    Wo, ao, fo = ol.OracleRun(cfg).eval(u)
    compare(Wg, Wo, fg, fo, 1e-9, "preset 10 full tables")


def test_nan_weights_stop_the_run():
    # a NaN coupling makes every valid point's weight NaN: the kernels set those
    # weights to 0 and flag them, and the run fails as CheckInfNan throws
    cfg = abi.cfg_default(1)
    cfg.gN_P = float("nan")
    say("case: preset 1 with gN_P = NaN, VEGAS run of 20000 calls")
    with MContinuum(cfg) as P:
        u = points(40, cfg.dim, 35)
        W, a2, fl = P.eval(u)  # at most 40 NaN weights: within the limit
        valid = ((fl & abi.FLAG_KIN_OK) != 0) & ((fl & abi.FLAG_FID_OK) != 0)
        assert valid.sum() > 5
        assert np.all(W == 0)
        assert np.all((fl[valid] & abi.FLAG_AMP_OK) == 0)
        assert np.all((fl & abi.FLAG_INF) == 0)
        G = MGraniitti(P)
        with pytest.raises(abi.GraError) as e:
            G.Integrate(vegas_cfg(20000, 2, burnin=1))
        assert "NaN weights" in str(e.value)
        assert f"status {abi.GRA_E_NUMERIC}" in str(e.value)
