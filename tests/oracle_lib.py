"""ctypes access to the oracle (oracle/liboracle.so) and, where built, to the
reference header-only pieces (oracle/_ref/libgra_ref.so). Test helper only."""
import ctypes as C
import os

import numpy as np

from graniitti_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libgra_ref.so")

D = C.c_double
P = C.POINTER
VP = C.c_void_p
I32 = C.c_int32
U64 = C.c_uint64


def _load_oracle():
    if not os.path.exists(ORACLE):
        raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
    L = C.CDLL(ORACLE)
    sig = {
        "ora_philox4x32_10": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "ora_point_uniforms": (None, [U64, C.c_uint32, U64, C.c_int, P(D)]),
        "ora_grid_init": (C.c_int, [P(abi.GraGrid), C.c_int, C.c_int]),
        "ora_grid_optimize": (None, [P(abi.GraGrid), P(D), D]),
        "ora_vegas_map": (None, [P(abi.GraGrid), P(D), P(D), P(D), P(I32)]),
        "ora_check_emc": (C.c_int, [P(D)]),
        "ora_solvepz": (D, [D, D, D, D, D, D, D]),
        "ora_regge_eta": (None, [D, D, P(D)]),
        "ora_regge_eta_linear": (None, [D, D, D, D, P(D)]),
        "ora_pom_alpha": (D, [P(abi.GraCfg), D]),
        "ora_regge_prop": (None, [P(abi.GraCfg), D, D, P(D)]),
        "ora_bessj0": (D, [D]),
        "ora_cs_integral": (None, [P(D), C.c_int, D, P(D)]),
        "ora_fb_transform": (D, [P(D), C.c_int, D, D, D]),
        "ora_simpson38_weight": (D, [C.c_int, C.c_int]),
        "ora_create": (VP, [P(abi.GraCfg), C.c_int]),
        "ora_destroy": (None, [VP]),
        "ora_tables": (C.c_int, [VP, P(D), P(D)]),
        "ora_msa": (None, [VP, D, P(D)]),
        "ora_mbt": (None, [VP, D, P(D)]),
        "ora_eval_points": (None, [VP, P(D), C.c_int64, P(D), P(D), P(C.c_uint8)]),
        "ora_kinematics": (C.c_int, [VP, P(D), P(D), P(D), P(D)]),
        "ora_bare_amplitudes": (C.c_int, [VP, P(D), P(D)]),
        "ora_tensor_amplitudes": (C.c_int, [VP, P(D), P(D)]),
        "ora_tensor_mtensor": (None, [VP, P(D), P(D)]),
        "ora_tensor_2xs_amplitude": (None, [VP, P(D), P(D)]),
        "ora_kelly_f1": (D, [D, D]),
        "ora_gauss_legendre": (None, [C.c_int, P(D), P(D)]),
        "ora_log_node_amp": (None, [VP, D, P(D)]),
        "ora_cbw": (None, [P(abi.GraRes), D, D, D, P(D)]),
        "ora_res_ff": (D, [D, D, D]),
        "ora_coh_flux": (D, [D, D, D, D]),
        "ora_photo_prop": (None, [P(abi.GraCfg), D, D, D, P(D)]),
        "ora_vegas_iterate": (None, [VP, P(abi.GraGrid), U64, U64, U64, C.c_uint32, C.c_int,
                                     C.c_int, P(abi.GraIterStats)]),
        "ora_vegas_combine": (None, [D, D, D, C.c_int, P(D), P(D)]),
        "ora_phase_space_weight": (D, [VP, P(D)]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype = r
        f.argtypes = a
    return L


_ora = None


def oracle():
    global _ora
    if _ora is None:
        _ora = _load_oracle()
    return _ora


def ref_available():
    return os.path.exists(REF)


_ref = None


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF)
        sig = {
            "ref_grid_init": (C.c_int, [C.c_int, C.c_int, P(D)]),
            "ref_grid_optimize": (C.c_int, [C.c_int, C.c_int, D, P(D), P(D)]),
            "ref_simpson38_weight2d": (C.c_int, [C.c_int, C.c_int, P(D)]),
            "ref_simpson38_integral2d": (C.c_int, [C.c_int, C.c_int, P(D), D, D, P(D)]),
            "ref_csintegral": (C.c_int, [C.c_int, P(D), D, P(D)]),
            "ref_checkemc": (C.c_int, [D, D, D, D]),
            "ref_bessj0": (D, [D]),
            "ref_m4vec": (None, [D, D, D, D, P(D)]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _ref = L
    return _ref


class OracleRun:
    """FP64 CPU restatement of one process configuration."""

    def __init__(self, cfg, nthreads=8):
        self.cfg = cfg
        self.L = oracle()
        self.h = self.L.ora_create(C.byref(cfg), nthreads)
        if not self.h:
            raise RuntimeError("ora_create failed (abi_size mismatch?)")

    def close(self):
        if self.h:
            self.L.ora_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def eval(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        n = u.shape[0]
        W = np.zeros(n)
        a2 = np.zeros(n)
        fl = np.zeros(n, dtype=np.uint8)
        self.L.ora_eval_points(self.h, abi.dptr(u), n, abi.dptr(W), abi.dptr(a2), abi.u8ptr(fl))
        return W, a2, fl

    def tables(self):
        c = self.cfg
        mbt = np.zeros((c.n_bt + 1, 3))
        msa = np.zeros((c.n_kt2 + 1, 3))
        assert self.L.ora_tables(self.h, abi.dptr(mbt), abi.dptr(msa)) == 0
        return mbt, msa

    def msa(self, kt2):
        o = np.zeros(2)
        self.L.ora_msa(self.h, kt2, abi.dptr(o))
        return complex(o[0], o[1])

    def kinematics(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        pf = np.zeros((5, 4))
        legs = np.zeros((2, 2, 4))
        sc = np.zeros(8)
        ok = self.L.ora_kinematics(self.h, abi.dptr(x), abi.dptr(pf), abi.dptr(legs), abi.dptr(sc))
        return ok, pf, legs, sc

    def bare_amplitudes(self, x):
        """Helicity amplitudes of the bare process at hypercube point x (None if the
        kinematics fail)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        h = np.zeros(2 * 64)
        nh = self.L.ora_bare_amplitudes(self.h, abi.dptr(x), abi.dptr(h))
        if nh == 0:
            return None
        return h[0:2 * nh:2] + 1j * h[1:2 * nh:2]

    def tensor_amplitudes(self, p):
        """9 helicity amplitudes (h3, h4 in -1, 0, +1) of the tensor-pomeron 2xV
        ME4 at momenta p = (pa, pb, p1, p2, p3, p4), each (E, px, py, pz)."""
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(6, 4)
        h = np.zeros(18)
        self.L.ora_tensor_amplitudes(self.h, abi.dptr(p), abi.dptr(h))
        return h[0::2] + 1j * h[1::2]

    def tensor_mtensor(self, p):
        """Amplitude tensor M(r3, r4) (indices down, (t, x, y, z)) of the 2xV ME4."""
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(6, 4)
        o = np.zeros(32)
        self.L.ora_tensor_mtensor(self.h, abi.dptr(p), abi.dptr(o))
        return (o[0::2] + 1j * o[1::2]).reshape(4, 4)

    def tensor_2xs_amplitude(self, p):
        """The tensor-pomeron 2xS amplitude at momenta p = (pa, pb, p1, p2, p3, p4)."""
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(6, 4)
        o = np.zeros(2)
        self.L.ora_tensor_2xs_amplitude(self.h, abi.dptr(p), abi.dptr(o))
        return complex(o[0], o[1])

    def vegas_iterate(self, grid, n, seed, it, nthreads=8, first=0, use_mutex=0):
        st = abi.GraIterStats()
        self.L.ora_vegas_iterate(self.h, C.byref(grid), first, n, seed, it, nthreads, use_mutex,
                                 C.byref(st))
        return st


def point_uniforms(seed, it, idx, dim):
    u = np.zeros(dim)
    oracle().ora_point_uniforms(seed, it, idx, dim, abi.dptr(u))
    return u
