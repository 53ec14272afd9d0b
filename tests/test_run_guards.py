"""CPU tests of the run guards of the C ABI (no GPU):

- the VEGAS schedule of MGraniitti::SampleVegas + VEGAS (src/MGraniitti.cc:1074-1338)
  driven through gra_vegas_run_host by passes written here: the chi2 > 50 restarts
  (10 x calls in burn-in, twice the burn-in iterations while integrating), the
  NaN/Inf and < 1e-60 integral failures, and CheckInfNan's limit of 50 NaN or 50
  Inf weights (include/Graniitti/MProcess.h:317-340);
- the validation of resonance spin densities (spin::Positivity, src/MSpin.cc:988ff).

The passes report what a VEGAS iteration over an integrand with a chosen mean and
variance would: sum f = N m, sum f^2 = N (m^2 + v), uniform bin histograms."""
import ctypes as C

import numpy as np
import pytest

from graniitti_amd import abi

DIM = 3


def _cfg(ncall=1000, burnin=3, it=4, max_iter=4):
    v = abi.GraVegasCfg()
    v.ncall = ncall
    v.seed = 7
    v.burnin_iter = burnin
    v.iter = it
    v.max_iter = max_iter
    v.bins = 16
    v.lambda_ = 1.5
    v.precision = 0.0
    v.chi2max = 0.0
    return v


class Pass:
    """A pass whose per-iteration mean and variance come from a rule
    rule(k, n_calls) -> (mean, var, n_nan, n_inf), k = pass count so far."""

    def __init__(self, rule):
        self.rule = rule
        self.calls = []
        self.fn = abi.VEGAS_PASS(self._pass)

    def _pass(self, ctx, grid, n, seed, it, out):
        k = len(self.calls)
        self.calls.append((int(n), int(it)))
        m, v, nn, ni = self.rule(k, int(n))
        st = out.contents
        C.memset(out, 0, C.sizeof(abi.GraIterStats))
        st.fsum = n * m
        st.f2sum = n * (m * m + v)
        st.maxW = st.maxf = 2 * m
        st.n_calls = st.n_valid = n
        st.n_nan = nn
        st.n_inf = ni
        for j in range(DIM):
            for b in range(16):
                st.fmat[j * abi.GRA_VEGAS_BINS + b] = n * m / 16
                st.f2mat[j * abi.GRA_VEGAS_BINS + b] = n * (m * m + v) / 16
        return 0


def _run(rule, **kw):
    lib = abi.load()
    p = Pass(rule)
    g = abi.GraGrid()
    r = abi.GraVegasResult()
    rc = lib.gra_vegas_run_host(C.byref(_cfg(**kw)), C.byref(g), DIM, p.fn, None, C.byref(r))
    msg = lib.gra_hip_last_error(None).decode()
    return rc, msg, r, p


def test_consistent_run_has_no_restart():
    rc, msg, r, p = _run(lambda k, n: (1.0, 1.0, 0, 0))
    assert rc == abi.GRA_OK, msg
    assert r.restarts == 0 and r.iterations == 4
    assert len(p.calls) == 3 + 4 and all(n == 1000 for n, _ in p.calls)
    assert r.sigma == pytest.approx(1.0) and r.calls_total == 7000
    # the error is the inverse-variance combination of 4 equal iterations
    assert r.sigma_err == pytest.approx(np.sqrt(1.0 / 1000 / 4))


def test_burnin_chi2_restarts_with_ten_times_the_calls():
    # burn-in passes alternate between means 1 and 2 (chi2 >> 50) until the calls
    # are raised; then they agree
    def rule(k, n):
        if n < 10000:
            return (1.0 if k % 2 == 0 else 2.0), 1e-4, 0, 0
        return 1.0, 1.0, 0, 0

    rc, msg, r, p = _run(rule)
    assert rc == abi.GRA_OK, msg
    assert r.restarts == 1
    # two burn-in passes at 1000 calls (the second one trips chi2), then a fresh
    # burn-in and the integration at 10000
    assert [n for n, _ in p.calls] == [1000, 1000] + [10000] * 7
    assert r.calls_total == 2 * 1000 + 7 * 10000


def test_integration_chi2_restarts_with_twice_the_burnin():
    # consistent in burn-in; the first integration's second pass disagrees
    seen = {"bad": False}

    def rule(k, n):
        if k == 4 and not seen["bad"]:
            seen["bad"] = True
            return 5.0, 1e-4, 0, 0
        return 1.0, 1e-2, 0, 0

    rc, msg, r, p = _run(rule)
    assert rc == abi.GRA_OK, msg
    assert r.restarts == 1
    # 3 burn-in + 2 integration passes, then 6 burn-in + 4 integration
    assert len(p.calls) == 3 + 2 + 6 + 4
    assert all(n == 1000 for n, _ in p.calls)
    assert r.iterations == 4


def test_restarts_are_bounded():
    rc, msg, r, p = _run(lambda k, n: ((1.0 if k % 2 == 0 else 3.0), 1e-6, 0, 0), ncall=10)
    assert rc == abi.GRA_E_NUMERIC
    assert "chi2" in msg


@pytest.mark.parametrize("mean,word", [(float("nan"), "inf/nan"), (1e-70, "1e-60")])
def test_bad_integral_fails(mean, word):
    rc, msg, r, p = _run(lambda k, n: (mean, 0.0, 0, 0))
    assert rc == abi.GRA_E_NUMERIC
    assert word in msg
    assert len(p.calls) == 2  # the guard skips the first pass of a sequence


@pytest.mark.parametrize("kind", ["nan", "inf"])
def test_too_many_nonfinite_weights_fail(kind):
    # 30 per pass: 60 > 50 after the second pass
    rule = (lambda k, n: (1.0, 1.0, 30, 0)) if kind == "nan" else (lambda k, n: (1.0, 1.0, 0, 30))
    rc, msg, r, p = _run(rule)
    assert rc == abi.GRA_E_NUMERIC
    assert f"too many {kind.capitalize() if kind == 'inf' else 'NaN'} weights" in msg
    assert len(p.calls) == 2


def test_fifty_nonfinite_weights_are_tolerated():
    rc, msg, r, p = _run(lambda k, n: (1.0, 1.0, 50 if k == 0 else 0, 50 if k == 1 else 0))
    assert rc == abi.GRA_OK, msg


# ---------------------------------------------------------------------------
# Resonance spin densities
# ---------------------------------------------------------------------------
def _validate(cfg):
    """validate() through an entry point that needs no GPU"""
    lib = abi.load()
    buf = C.create_string_buffer(256)
    rc = lib.gra_hip_eikonal_cache_name(C.byref(cfg), 0, buf, len(buf))
    return rc, lib.gra_hip_last_error(None).decode()


def _f2(cfg):
    for k in range(cfg.n_res):
        if cfg.res[k].pdg == 225:
            return cfg.res[k]
    raise AssertionError


def _set_rho(r, M):
    n = 2 * r.J + 1
    for a in range(n):
        for b in range(n):
            r.rho_re[a * n + b] = M[a, b].real
            r.rho_im[a * n + b] = M[a, b].imag


def test_presets_pass_spin_density_checks():
    for p in (6, 7, 8):
        rc, msg = _validate(abi.cfg_default(p))
        assert rc == abi.GRA_OK, (p, msg)


@pytest.mark.parametrize("case", ["trace", "hermitian", "negative"])
def test_bad_spin_density_is_rejected(case):
    cfg = abi.cfg_default(6)
    r = _f2(cfg)
    M = np.diag([0.5, 0, 0, 0, 0.5]).astype(complex)
    if case == "trace":
        M *= 0.5
    elif case == "hermitian":
        M[0, 1] = M[1, 0] = 0.1j
    else:
        # eigenvalues 0.5 +- 0.71: a negative weight for some projection
        M = np.diag([0.6, 0, 0, 0, 0.4]).astype(complex)
        M[0, 4] = M[4, 0] = 0.7
    _set_rho(r, M)
    rc, msg = _validate(cfg)
    assert rc == abi.GRA_E_ARG, msg
    assert {"trace": "normalised", "hermitian": "Hermitian", "negative": "negative"}[case] in msg


def test_valid_mixed_spin_density_is_accepted():
    # a random positive density matrix of trace 1 (complex off-diagonal terms)
    rng = np.random.default_rng(3)
    A = rng.normal(size=(5, 5)) + 1j * rng.normal(size=(5, 5))
    M = A @ A.conj().T
    M /= np.trace(M).real
    cfg = abi.cfg_default(6)
    _set_rho(_f2(cfg), M)
    rc, msg = _validate(cfg)
    assert rc == abi.GRA_OK, msg
