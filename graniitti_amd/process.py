"""Host-side process and integrator classes over the C ABI, shaped like the
reference's gra::MContinuum (EventWeight, GetdLIPSDim; src/MContinuum.cc:164-193,
include/Graniitti/MProcess.h:80-160) and gra::MGraniitti (SetIntegrator,
Initialize, GetXS; include/Graniitti/MGraniitti.h:105-202).

All evaluation runs in libgra_hip.so; a missing library or GPU raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi


class AuxIntData:
    """Per-point assertions (include/Graniitti/MProcess.h:37-52)."""

    def __init__(self, flags=0xF, vegasweight=1.0):
        self.vegasweight = vegasweight
        self.kinematics_ok = bool(flags & abi.FLAG_KIN_OK)
        self.fidcuts_ok = bool(flags & abi.FLAG_FID_OK)
        self.vetocuts_ok = bool(flags & abi.FLAG_VETO_OK)
        self.amplitude_ok = bool(flags & abi.FLAG_AMP_OK)

    def Valid(self):
        return self.kinematics_ok and self.fidcuts_ok and self.vetocuts_ok


def eikonal_cache_name(cfg, table):
    """File name under which table 0 (MBT) or 1 (MSA) of cfg is cached."""
    lib = abi.load()
    buf = C.create_string_buffer(256)
    abi.check(lib.gra_hip_eikonal_cache_name(C.byref(cfg), table, buf, len(buf)))
    return buf.value.decode()


class MContinuum:
    """The <C> continuum process on the GPU."""

    def __init__(self, cfg, eikonal_dir=None):
        """eikonal_dir: a directory that caches the eikonal tables in the
        reference's file format (MEikonal::S3Constructor, src/MEikonal.cc:46-132);
        tables found there are read, missing ones computed and written."""
        self.lib = abi.load()
        self.cfg = cfg
        h = C.c_void_p()
        if eikonal_dir is None:
            abi.check(self.lib.gra_hip_create(C.byref(cfg), C.byref(h)))
        else:
            abi.check(self.lib.gra_hip_create_cached(C.byref(cfg), os.fsencode(eikonal_dir),
                                                     C.byref(h)))
        self.h = h

    @classmethod
    def preset(cls, n, **overrides):
        cfg = abi.cfg_default(n)
        for k, v in overrides.items():
            setattr(cfg, k, v)
        return cls(cfg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.gra_hip_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def GetdLIPSDim(self):
        return self.cfg.dim

    def _chk(self, rc):
        abi.check(rc, self.h)

    def eval(self, u):
        """Batched EventWeight: u (n, dim) -> W (barn), |M|^2, flags."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.ndim != 2 or u.shape[1] != self.cfg.dim:
            raise ValueError(f"u must have shape (n, {self.cfg.dim})")
        n = u.shape[0]
        W = np.empty(n)
        a2 = np.empty(n)
        fl = np.empty(n, dtype=np.uint8)
        self._chk(self.lib.gra_hip_eval_points(self.h, abi.dptr(u), n, abi.dptr(W), abi.dptr(a2),
                                               abi.u8ptr(fl)))
        return W, a2, fl

    def EventWeight(self, randvec, aux=None):
        """Single-point form of the reference's EventWeight(randvec, aux)."""
        W, _, fl = self.eval(np.asarray(randvec, dtype=np.float64)[None, :])
        a = AuxIntData(int(fl[0]), aux.vegasweight if aux is not None else 1.0)
        if aux is not None:
            aux.__dict__.update(a.__dict__)
        return float(W[0])

    def vegas_iterate(self, grid, n, seed, it):
        st = abi.GraIterStats()
        self._chk(self.lib.gra_hip_vegas_iterate(self.h, C.byref(grid), n, seed, it, C.byref(st)))
        return st

    def vegas_map(self, grid, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        n, D = u.shape
        x = np.empty((n, D))
        vw = np.empty(n)
        b = np.empty((n, D), dtype=np.int32)
        self._chk(self.lib.gra_hip_vegas_map(self.h, C.byref(grid), abi.dptr(u), n, abi.dptr(x),
                                             abi.dptr(vw), b.ctypes.data_as(C.POINTER(C.c_int32))))
        return x, vw, b

    def eikonal_tables(self):
        c = self.cfg
        node = np.empty(2 * (c.n_loop_kt + 1))
        mbt = np.empty((c.n_bt + 1, 3))
        msa = np.empty((c.n_kt2 + 1, 3))
        self._chk(self.lib.gra_hip_eikonal_tables(self.h, abi.dptr(node), abi.dptr(mbt),
                                                  abi.dptr(msa)))
        return node[0::2] + 1j * node[1::2], mbt, msa

    def unweight(self, w, wmax, seed):
        w = np.ascontiguousarray(w, dtype=np.float64)
        acc = np.empty(len(w), dtype=np.uint8)
        na, no = C.c_uint64(), C.c_uint64()
        self._chk(self.lib.gra_hip_unweight(self.h, abi.dptr(w), len(w), wmax, seed,
                                            abi.u8ptr(acc), C.byref(na), C.byref(no)))
        return acc, na.value, no.value

    def kernel_time(self, which, reset=False):
        ms, n = C.c_double(), C.c_uint64()
        self._chk(self.lib.gra_hip_kernel_time(self.h, which, C.byref(ms), C.byref(n), int(reset)))
        return ms.value, n.value


def vegas_cfg(ncall, iters, burnin=3, seed=12345678, precision=0.0, max_iter=None, bins=128,
              lam=1.5, chi2max=0.0):
    vc = abi.GraVegasCfg()
    vc.ncall, vc.seed, vc.burnin_iter, vc.iter = ncall, seed, burnin, iters
    vc.max_iter = max_iter if max_iter is not None else iters
    vc.bins, vc.lambda_, vc.precision, vc.chi2max = bins, lam, precision, chi2max
    return vc


class MGraniitti:
    """Integrator driver: VEGAS on the GPU (and the flow sampler in the NEURO slot)."""

    def __init__(self, proc):
        self.proc = proc
        self.integrator = "VEGAS"
        self.grid = abi.GraGrid()
        self.result = None

    def SetIntegrator(self, name):
        if name not in ("VEGAS", "FLOW"):
            raise ValueError(f"unknown integrator {name}")
        self.integrator = name

    def Integrate(self, vc):
        res = abi.GraVegasResult()
        self.proc._chk(self.proc.lib.gra_hip_vegas_run(self.proc.h, C.byref(vc),
                                                       C.byref(self.grid), C.byref(res)))
        self.result = res
        return res

    def GetXS(self):
        return self.result.sigma, self.result.sigma_err

    def Generate(self, n_events, calls=1 << 20, seed=12345678, maxf=None, max_passes=64):
        """Unweighted events on the integrated grid (MGraniitti::Generate with the
        VEGAS generation mode, src/MGraniitti.cc:1424-1436, 1651-1726). Passes of
        `calls` points are drawn until n_events are accepted; maxf defaults to the
        largest f seen during integration. Returns (momenta (n, n_p, 4) as
        E, px, py, pz in GeV, f (n,), counters summed over the passes)."""
        if self.result is None:
            raise RuntimeError("Integrate() before Generate()")
        wmax = float(maxf if maxf is not None else self.result.maxf)
        lib, h = self.proc.lib, self.proc.h
        evs, fs = [], []
        tot = {"trials": 0, "valid": 0, "accepted": 0, "overflow": 0}
        for it in range(max_passes):
            need = n_events - sum(len(e) for e in evs)
            if need <= 0:
                break
            buf = (abi.GraEvent * need)()
            st = abi.GraGenStats()
            self.proc._chk(lib.gra_hip_generate(h, C.byref(self.grid), calls, seed, it, wmax, buf,
                                                need, C.byref(st)))
            for k in tot:
                tot[k] += getattr(st, k)
            if st.stored:
                a = np.ctypeslib.as_array(buf)[: st.stored]
                n_p = int(a["n_p"][0])
                evs.append(np.array(a["p"])[:, :n_p, :])
                fs.append(np.array(a["f"]))
        p = np.concatenate(evs) if evs else np.zeros((0, 4, 4))
        f = np.concatenate(fs) if fs else np.zeros(0)
        return p, f, tot


def record_pdg(cfg):
    """PDG ids of the event record of the built processes: p p beams; pi+ pi- for
    PP[CON], rho0 rho0 -> (pi+ pi-)(pi+ pi-) for PP[CONTENSOR]."""
    ids = abi.GraRecordPdg()
    ids.beam[0] = ids.beam[1] = 2212
    if cfg.m_daughter > 0.0:
        ids.central[0] = ids.central[1] = 113
        for i in range(2):
            ids.daughter[i][0], ids.daughter[i][1] = 211, -211
    else:
        ids.central[0], ids.central[1] = 211, -211
    return ids


class HepMC3Writer:
    """HepMC3 ASCII output of generated events (the WriterAscii the reference
    opens in MGraniitti::InitFileOutput, src/MGraniitti.cc:190-207, written by
    SaveEvent, 1665-1712, with the record of MProcess::CommonRecord,
    src/MProcess.cc:1588-1706). Host-only calls into libgra_hip.so."""

    def __init__(self, path, cfg, generator="GRANIITTI (TUNE0)", version="1.0", card=None,
                 pdg=None):
        self.lib = abi.load()
        self.cfg = cfg
        self.pdg = pdg if pdg is not None else record_pdg(cfg)
        w = C.c_void_p()
        abi.check(self.lib.gra_hepmc3_open(str(path).encode(), generator.encode(),
                                           version.encode(),
                                           card.encode() if card else None, C.byref(w)))
        self.w = w

    def write(self, p, f, sigma, sigma_err, weighted=False):
        """Append events: p (n, n_p, 4) as (E, px, py, pz) in the gra_event order,
        f (n,) the VEGAS-weighted integrand; sigma and sigma_err in barn."""
        p = np.asarray(p, dtype=np.float64)
        n, n_p = p.shape[0], p.shape[1]
        buf = (abi.GraEvent * max(n, 1))()
        a = np.ctypeslib.as_array(buf)
        a["p"][:n, :n_p, :] = p
        a["f"][:n] = np.asarray(f, dtype=np.float64)
        a["n_p"][:n] = n_p
        abi.check(self.lib.gra_hepmc3_write(self.w, C.byref(self.cfg), C.byref(self.pdg), buf, n,
                                            float(sigma), float(sigma_err), int(weighted)))

    def close(self):
        if getattr(self, "w", None):
            w, self.w = self.w, None
            abi.check(self.lib.gra_hepmc3_close(w))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
