"""Checks that pin the oracle's tensor-pomeron ME4 (pp -> p V V p, 2xV,
src/MTensorPomeron.cc:572-898) by what the mathematics fixes: gauge
invariance of the Pomeron-vector-vector vertices, completeness of the massive
polarisation vectors, Bose symmetry of the two identical vector mesons, and
Lorentz invariance of the helicity sum. Overall normalisation conventions
(couplings, form factors, phases) are not fixed by these and stay unpinned."""
import numpy as np
import pytest

from graniitti_amd import abi
import oracle_lib as ol

G = np.diag([1.0, -1.0, -1.0, -1.0])


@pytest.fixture(scope="module")
def run3():
    cfg = abi.cfg_default(3)
    cfg.screening = 0
    cfg.amp_precision = abi.GRA_AMP_FP64
    R = ol.OracleRun(cfg, 2)
    yield R
    R.close()


def _points(R, n, seed):
    """Momenta (pa, pb, p1, p2, p3, p4) of n valid config-3 phase-space points."""
    E = R.cfg.beam_E1
    pz = np.sqrt(E * E - R.cfg.m_beam ** 2)
    pa, pb = np.array([E, 0, 0, pz]), np.array([E, 0, 0, -pz])
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        ok, pf, legs, sc = R.kinematics(rng.random(R.cfg.dim))
        if ok:
            out.append(np.array([pa, pb, pf[1], pf[2], pf[3], pf[4]]))
    return out


def _boost(p, beta):
    """Pure boost of 4-vectors (E, px, py, pz) by velocity vector beta."""
    b = np.asarray(beta, dtype=float)
    b2 = b @ b
    g = 1.0 / np.sqrt(1.0 - b2)
    L = np.eye(4)
    L[0, 0] = g
    L[0, 1:] = L[1:, 0] = g * b
    L[1:, 1:] += (g - 1.0) * np.outer(b, b) / b2
    return p @ L.T


def test_amplitude_tensor_is_transverse(run3):
    # k2^nu Gamma_{mu nu ..}(k1, k2) = 0 and k1^mu Gamma_{mu nu ..} = 0 make
    # M(r3, r4) p3^{r3} = 0 and M(r3, r4) p4^{r4} = 0
    worst = 0.0
    for p in _points(run3, 40, 1):
        M = run3.tensor_mtensor(p)
        p3, p4 = p[4], p[5]
        scale = np.abs(M).max() * max(np.abs(p3).max(), np.abs(p4).max())
        worst = max(worst, np.abs(p3 @ M).max() / scale, np.abs(M @ p4).max() / scale)
    assert worst < 1e-10


def test_polarisation_sum_is_complete(run3):
    # sum_lambda eps^mu eps*^ka = -g^{mu ka} + p^mu p^ka / m^2 for the three massive
    # helicities; with M transverse (previous test) the helicity sum must equal
    # M_{mu nu} M*_{ka la} g^{mu ka} g^{nu la}
    for p in _points(run3, 40, 2):
        h = run3.tensor_amplitudes(p)
        M = run3.tensor_mtensor(p)
        full = np.einsum("ab,cd,ac,bd->", M, np.conj(M), G, G).real
        assert abs((np.abs(h) ** 2).sum() - full) <= 1e-10 * full


def test_bose_symmetry_of_the_meson_pair(run3):
    # exchanging the identical mesons maps the t-channel onto the u-channel:
    # h(h3, h4; p3, p4) = h(h4, h3; p4, p3)
    for p in _points(run3, 40, 3):
        h = run3.tensor_amplitudes(p).reshape(3, 3)
        q = p.copy()
        q[[4, 5]] = q[[5, 4]]
        hs = run3.tensor_amplitudes(q).reshape(3, 3)
        assert np.allclose(hs.T, h, rtol=1e-10, atol=1e-10 * np.abs(h).max())


def test_helicity_sum_is_lorentz_invariant(run3):
    # FP64 boosts of the 6.5 TeV beams leave t = (pa - p1)^2 with ~1e-8 relative
    # rounding; a metric sign or index error breaks invariance at O(1)
    for p in _points(run3, 20, 4):
        a2 = (np.abs(run3.tensor_amplitudes(p)) ** 2).sum()
        for beta in ([0.3, 0.0, 0.0], [0.0, -0.5, 0.2], [0.1, 0.1, 0.6]):
            b2 = (np.abs(run3.tensor_amplitudes(_boost(p, beta))) ** 2).sum()
            assert abs(b2 - a2) <= 1e-7 * a2
        # rotation about the beam axis and about x
        c, s = np.cos(0.7), np.sin(0.7)
        Rz = np.array([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1.0]])
        Rx = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, c, -s], [0, 0, s, c]])
        for Rm in (Rz, Rx):
            b2 = (np.abs(run3.tensor_amplitudes(p @ Rm.T)) ** 2).sum()
            assert abs(b2 - a2) <= 1e-7 * a2


def test_proton_helicity_replicas_and_average(run3):
    # the high-energy proton vertex conserves helicity: the bare process carries
    # 4 identical replicas of the 9 meson-helicity amplitudes and |M|^2 = sum / 4
    rng = np.random.default_rng(5)
    for _ in range(200):
        x = rng.random(run3.cfg.dim)
        h = run3.bare_amplitudes(x)
        if h is None:
            continue
        assert h.shape == (36,)
        assert np.array_equal(h[:9], h[9:18]) and np.array_equal(h[:9], h[27:])
        W, a2, fl = run3.eval(x[None, :])
        if fl[0] & abi.FLAG_FID_OK:
            assert abs(a2[0] - (np.abs(h) ** 2).sum() / 4) <= 1e-12 * a2[0]
