"""Write tests/golden/ref_vectors.json from the reference's own header-only code.

Runs only where the reference sources exist (this container): oracle/Makefile
builds oracle/_ref/libgra_ref.so from include/Graniitti/MVEGAS.h, MMath.h,
M4Vec.h and libs/extmath.hpp of the reference, and this script records its
outputs on fixed inputs. The JSON holds numbers only (inputs and outputs).
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle_lib as ol  # noqa: E402

D = C.c_double


def arr(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def main():
    R = ol.ref()
    rng = np.random.default_rng(20240601)
    out = {}
    # VEGAS: uniform grid and three optimisation steps on random histograms
    dim, bins = 3, 128
    x = np.zeros(dim * bins)
    R.ref_grid_init(dim, bins, x.ctypes.data_as(C.POINTER(D)))
    out["grid_init"] = {"dim": dim, "bins": bins, "xmat": x.tolist()}
    steps = []
    for k in range(3):
        f2 = arr(rng.lognormal(0.0, 2.0, size=dim * bins) * (rng.random(dim * bins) < 0.9))
        if k == 2:
            f2[:bins] = 0.0  # all-empty dimension: EPS floor and NaN guard
        xin = x.copy()
        f2in = f2.copy()
        R.ref_grid_optimize(dim, bins, 1.5, x.ctypes.data_as(C.POINTER(D)),
                            f2.ctypes.data_as(C.POINTER(D)))
        steps.append({"xmat_in": xin.tolist(), "f2mat_in": f2in.tolist(), "xmat_out": x.tolist(),
                      "f2mat_out": f2.tolist()})
    out["grid_optimize"] = {"dim": dim, "bins": bins, "lambda": 1.5, "steps": steps}
    # Simpson 3/8 weights and integral
    for M, N in [(12, 15), (15, 30), (3, 3)]:
        W = np.zeros((M + 1) * (N + 1))
        R.ref_simpson38_weight2d(M, N, W.ctypes.data_as(C.POINTER(D)))
        f = arr(rng.normal(size=2 * (M + 1) * (N + 1)))
        o = np.zeros(2)
        R.ref_simpson38_integral2d(M, N, f.ctypes.data_as(C.POINTER(D)), 0.3, 0.07,
                                   o.ctypes.data_as(C.POINTER(D)))
        out[f"simpson38_{M}_{N}"] = {"W": W.tolist(), "f": f.tolist(), "hM": 0.3, "hN": 0.07,
                                     "I": o.tolist()}
    W = np.zeros(16 * 16)
    out["simpson38_bad"] = R.ref_simpson38_weight2d(12, 16, W.ctypes.data_as(C.POINTER(D)))
    # Composite Simpson (complex)
    for N in [2, 10, 1000]:
        f = arr(rng.normal(size=2 * (N + 1)))
        o = np.zeros(2)
        R.ref_csintegral(N, f.ctypes.data_as(C.POINTER(D)), 0.013, o.ctypes.data_as(C.POINTER(D)))
        out[f"csintegral_{N}"] = {"f": f.tolist(), "h": 0.013, "I": o.tolist()}
    # BESSJ0 over both branches and the switch point
    xs = np.concatenate([[0.0, 1e-12, 7.999999999, 8.0, 8.000000001, -3.0, 30.0, 1500.0],
                         rng.uniform(0, 60, 200), rng.uniform(0, 2000, 50)])
    out["bessj0"] = {"x": xs.tolist(), "y": [R.ref_bessj0(float(v)) for v in xs]}
    # CheckEMC cases
    cases = [[0, 0, 0, 0], [1e-7, 0, 0, 0], [2e-6, 0, 0, 0], [-2e-6, 0, 0, 0], [0, 2e-6, 0, 0],
             [0, 0, -2e-6, 0], [0, 0, 0, 5e-7], [0, 0, 0, -5e-7], [float("nan"), 0, 0, 0]]
    out["checkemc"] = {"cases": [[c if c == c else "nan" for c in cs] for cs in cases],
                       "ok": [R.ref_checkemc(*cs) for cs in cases]}
    # M4Vec observables (E, x, y, z): M2, M, Pt, Mt, Eta, Rap
    vecs = [[6500.0, 0.3, -0.2, 6499.99], [1.5, 0.4, 0.3, -0.9], [0.2, 0.05, 0.01, 0.1],
            [10.0, 3.0, 4.0, 5.0], [1.0, 0.9, 0.1, 0.0]]
    obs = []
    for v in vecs:
        o = np.zeros(6)
        R.ref_m4vec(*v, o.ctypes.data_as(C.POINTER(D)))
        obs.append(o.tolist())
    out["m4vec"] = {"v": vecs, "obs": obs}
    path = os.path.join(HERE, "ref_vectors.json")
    with open(path, "w") as fh:
        json.dump(out, fh)
    print("wrote", path)


if __name__ == "__main__":
    main()
