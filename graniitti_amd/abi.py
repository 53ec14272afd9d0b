"""ctypes mirror of include/gra_hip.h and the loader of the in-tree HIP library.

The structures below must match the C declarations field for field; the
library reports sizeof(gra_cfg) in gra_cfg.abi_size, which the tests compare
with ctypes.sizeof.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgra_hip.so")

GRA_OK = 0
GRA_E_ARG, GRA_E_HIP, GRA_E_NOMEM, GRA_E_STATE, GRA_E_NUMERIC, GRA_E_IO = -1, -2, -3, -4, -5, -6
GRA_MAX_DIM = 16
GRA_VEGAS_BINS = 128
GRA_PROC_REGGE_CON = 15
GRA_PROC_TENSOR_CON = 12
GRA_PROC_RES_CON = 17
GRA_PROC_RES = 9
GRA_MAX_RES = 8
GRA_MAX_VETO = 8
GRA_BW_FIXED, GRA_BW_RUNNING, GRA_BW_BARRIER, GRA_BW_SPIN = 1, 2, 3, 4
GRA_FRAME_CM, GRA_FRAME_HX, GRA_FRAME_CS, GRA_FRAME_GJ = 0, 1, 2, 3
GRA_AMP_FP64 = 0
GRA_AMP_FP32 = 1
GRA_PS_C, GRA_PS_F = 0, 1
GRA_LOOP_SIMPSON38, GRA_LOOP_GAUSS = 0, 1
FLAG_KIN_OK, FLAG_FID_OK, FLAG_VETO_OK, FLAG_AMP_OK, FLAG_INF = 1, 2, 4, 8, 16
GRA_MAX_NONFINITE = 50

D = C.c_double
I32 = C.c_int32
U64 = C.c_uint64


class GraRes(C.Structure):
    _fields_ = [
        ("mass", D), ("width", D), ("g_A", D), ("g_phi", D), ("g_FF", D), ("BR", D),
        ("alpha_re", D), ("alpha_im", D), ("rho_re", D * 25), ("rho_im", D * 25),
        ("J", I32), ("P", I32), ("BW", I32), ("pdg", I32),
    ]


class GraCfg(C.Structure):
    _fields_ = [
        ("abi_size", I32), ("process", I32), ("amp_precision", I32), ("dim", I32),
        ("beam_E1", D), ("beam_E2", D), ("m_beam", D),
        ("m_central", D), ("w_central", D), ("offshell", D), ("m_daughter", D), ("S_factor", D),
        ("fwd_pt_min", D), ("fwd_pt_max", D), ("kt_min", D), ("kt_max", D),
        ("rap_min", D), ("rap_max", D),
        ("fid_active", I32), ("pad0", I32),
        ("fid_eta_min", D), ("fid_eta_max", D), ("fid_rap_min", D), ("fid_rap_max", D),
        ("fid_pt_min", D), ("fid_pt_max", D), ("fid_et_min", D), ("fid_et_max", D),
        ("sys_M_min", D), ("sys_M_max", D), ("sys_Y_min", D), ("sys_Y_max", D),
        ("sys_Pt_min", D), ("sys_Pt_max", D),
        ("fwd_t_min", D), ("fwd_t_max", D), ("fwd_M_min", D), ("fwd_M_max", D),
        ("delta_P", D), ("alpha_P", D), ("gN_P", D), ("fc1", D), ("fc2", D), ("fc3", D),
        ("mpi_const", D),
        ("regge_a0", D * 3), ("regge_ap", D * 3), ("regge_sgn", D * 3), ("regge_c", D * 3),
        ("regge_on", I32 * 3), ("offshellFF", I32),
        ("regge_s0", D), ("b_EXP", D), ("a_OREAR", D), ("b_OREAR", D), ("b_POW", D),
        ("m_prop", D),
        ("tp_delta_P", D), ("tp_ap_P", D), ("tp_gPNN", D), ("tp_gPvv", D * 2), ("tp_mp_ff", D),
        ("tp_gPpipi", D), ("tp_gPKK", D), ("central_spin2", I32), ("pad_tp", I32),
        ("screening", I32), ("n_loop_phi", I32), ("n_loop_kt", I32), ("n_bt", I32),
        ("n_kt2", I32), ("n_fb", I32),
        ("min_loop_kt", D), ("max_loop_kt", D), ("loop_rule", I32), ("eik_log_nodes", I32),
        ("phase_space", I32), ("pad_ps", I32), ("gen_M_min", D), ("gen_M_max", D),
        ("n_res", I32), ("spin_gen", I32), ("spin_dec", I32), ("spin_frame", I32),
        ("regge_omega", D), ("res", GraRes * GRA_MAX_RES),
        ("veto_active", I32), ("n_veto", I32),
        ("veto_eta_min", D * GRA_MAX_VETO), ("veto_eta_max", D * GRA_MAX_VETO),
        ("veto_pt_min", D * GRA_MAX_VETO), ("veto_pt_max", D * GRA_MAX_VETO),
    ]


class GraGrid(C.Structure):
    _fields_ = [("dim", I32), ("bins", I32), ("xmat", D * (GRA_MAX_DIM * GRA_VEGAS_BINS))]


class GraIterStats(C.Structure):
    _fields_ = [
        ("fsum", D), ("fsum_c", D), ("f2sum", D), ("f2sum_c", D), ("maxW", D), ("maxf", D),
        ("n_calls", U64), ("n_kin_ok", U64), ("n_fid_ok", U64), ("n_valid", U64),
        ("n_nan", U64), ("n_inf", U64),
        ("fmat", D * (GRA_MAX_DIM * GRA_VEGAS_BINS)), ("f2mat", D * (GRA_MAX_DIM * GRA_VEGAS_BINS)),
    ]


class GraVegasCfg(C.Structure):
    _fields_ = [
        ("ncall", U64), ("seed", U64), ("burnin_iter", I32), ("iter", I32), ("max_iter", I32),
        ("bins", I32), ("lambda_", D), ("precision", D), ("chi2max", D),
    ]


class GraVegasResult(C.Structure):
    _fields_ = [
        ("sigma", D), ("sigma_err", D), ("chi2", D), ("maxW", D), ("maxf", D),
        ("seconds", D), ("seconds_total", D), ("seconds_to_precision", D),
        ("calls_total", U64), ("iterations", I32), ("restarts", I32), ("kernel_ms", D),
    ]


class GraFlowCfg(C.Structure):
    _fields_ = [
        ("dim", I32), ("layers", I32), ("bins", I32), ("hidden", I32), ("depth", I32),
        ("steps", I32), ("batch", U64), ("lr", D), ("beta1", D), ("beta2", D), ("eps", D),
        ("seed", U64), ("ess_fallback", D), ("spline_loops", I32), ("out_groups", I32),
        ("wgrad_split", I32), ("pad1", I32),
    ]


class GraFlowStats(C.Structure):
    _fields_ = [
        ("sigma", D), ("sigma_err", D), ("cv2", D), ("ess", D), ("maxw", D), ("loss", D),
        ("seconds", D), ("n", U64), ("fallback", I32), ("pad0", I32),
    ]


class GraFlowRunCfg(C.Structure):
    _fields_ = [
        ("train_steps", I32), ("check_every", I32), ("n_max", U64), ("seed", U64),
        ("precision", D), ("ess_min", D), ("vegas", GraVegasCfg), ("plateau", D),
    ]


class GraFlowResult(C.Structure):
    _fields_ = [
        ("sigma", D), ("sigma_err", D), ("cv2", D), ("ess", D), ("maxw", D),
        ("train_seconds", D), ("sample_seconds", D), ("seconds_to_precision", D),
        ("n", U64), ("fallback", I32), ("reached", I32), ("vegas", GraVegasResult),
        ("train_steps_run", U64), ("train_cv2", D),
    ]


class GraEvent(C.Structure):
    _fields_ = [("p", D * 4 * 8), ("f", D), ("index", U64), ("n_p", I32), ("pad0", I32)]


class GraRecordPdg(C.Structure):
    _fields_ = [("beam", I32 * 2), ("central", I32 * 2), ("daughter", I32 * 2 * 2)]


class GraGenStats(C.Structure):
    _fields_ = [("trials", U64), ("valid", U64), ("accepted", U64), ("overflow", U64),
                ("stored", U64)]


# Every function the header declares, with its ctypes signature
P = C.POINTER
VP = C.c_void_p
VEGAS_PASS = C.CFUNCTYPE(C.c_int, VP, C.POINTER(GraGrid), U64, U64, C.c_uint32,
                         C.POINTER(GraIterStats))
SIGNATURES = {
    "gra_hip_cfg_default": (C.c_int, [P(GraCfg), I32]),
    "gra_hip_create": (C.c_int, [P(GraCfg), P(VP)]),
    "gra_hip_create_cached": (C.c_int, [P(GraCfg), C.c_char_p, P(VP)]),
    "gra_hip_eikonal_cache_name": (C.c_int, [P(GraCfg), I32, C.c_char_p, C.c_size_t]),
    "gra_eikonal_table_write": (C.c_int, [C.c_char_p, P(D), I32]),
    "gra_eikonal_table_read": (C.c_int, [C.c_char_p, P(D), I32]),
    "gra_djb2": (U64, [C.c_char_p]),
    "gra_hip_destroy": (None, [VP]),
    "gra_hip_last_error": (C.c_char_p, [VP]),
    "gra_hip_eval_points": (C.c_int, [VP, P(D), C.c_size_t, P(D), P(D), P(C.c_uint8)]),
    "gra_hip_eval_points_device": (C.c_int, [VP, VP, C.c_size_t, VP, VP, VP]),
    "gra_hip_eval_points_noalloc": (C.c_int, [VP, VP, C.c_size_t, VP, VP, VP]),
    "gra_hip_grid_init": (C.c_int, [P(GraGrid), I32, I32]),
    "gra_hip_grid_optimize": (C.c_int, [P(GraGrid), P(GraIterStats), D]),
    "gra_hip_vegas_iterate": (C.c_int, [VP, P(GraGrid), U64, U64, C.c_uint32, P(GraIterStats)]),
    "gra_hip_vegas_run": (C.c_int, [VP, P(GraVegasCfg), P(GraGrid), P(GraVegasResult)]),
    "gra_vegas_run_host": (C.c_int, [P(GraVegasCfg), P(GraGrid), I32, VEGAS_PASS, VP,
                                     P(GraVegasResult)]),
    "gra_hip_vegas_map": (C.c_int, [VP, P(GraGrid), P(D), C.c_size_t, P(D), P(D), P(I32)]),
    "gra_hip_generate": (C.c_int, [VP, P(GraGrid), U64, U64, C.c_uint32, D, P(GraEvent), U64,
                                   P(GraGenStats)]),
    "gra_hip_unweight": (C.c_int, [VP, P(D), C.c_size_t, D, U64, P(C.c_uint8), P(U64), P(U64)]),
    "gra_hip_host_alloc": (C.c_int, [C.c_size_t, P(VP)]),
    "gra_hip_host_free": (None, [VP]),
    "gra_hip_eikonal_tables": (C.c_int, [VP, P(D), P(D), P(D)]),
    "gra_hip_flow_init": (C.c_int, [VP, P(GraFlowCfg)]),
    "gra_hip_flow_set_grid": (C.c_int, [VP, P(GraGrid)]),
    "gra_hip_flow_train": (C.c_int, [VP, I32, P(GraFlowStats)]),
    "gra_hip_flow_sample": (C.c_int, [VP, U64, U64, P(D), P(D), P(GraFlowStats)]),
    "gra_hip_flow_integrate": (C.c_int, [VP, P(GraFlowRunCfg), P(GraGrid), P(GraFlowResult)]),
    "gra_hip_flow_generate": (C.c_int, [VP, U64, U64, C.c_uint32, D, P(GraEvent), U64,
                                        P(GraGenStats)]),
    "gra_hip_flow_param_count": (C.c_int, [VP, P(U64)]),
    "gra_hip_flow_get_params": (C.c_int, [VP, P(C.c_float)]),
    "gra_hip_flow_set_params": (C.c_int, [VP, P(C.c_float)]),
    "gra_hip_flow_forward": (C.c_int, [VP, P(C.c_float), C.c_size_t, P(C.c_float), P(C.c_float)]),
    "gra_hip_flow_grad": (C.c_int, [VP, P(C.c_float), P(C.c_float), C.c_size_t, P(C.c_float),
                                    P(C.c_float)]),
    "gra_hip_kernel_time": (C.c_int, [VP, I32, P(D), P(U64), I32]),
    "gra_hip_screen_opcount": (C.c_int, [VP, U64, P(D)]),
    "gra_hip_fp32_fma_rate": (C.c_int, [P(D)]),
    "gra_hepmc3_open": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, P(VP)]),
    "gra_hepmc3_write": (C.c_int, [VP, P(GraCfg), P(GraRecordPdg), P(GraEvent), U64, D, D, I32]),
    "gra_hepmc3_close": (C.c_int, [VP]),
}

_lib = None


def load(path=LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class GraError(RuntimeError):
    pass


def check(rc, handle=None):
    if rc != GRA_OK:
        lib = load()
        msg = lib.gra_hip_last_error(handle)
        raise GraError(f"status {rc}: {msg.decode() if msg else ''}")


def dptr(a):
    return a.ctypes.data_as(P(D))


def u8ptr(a):
    return a.ctypes.data_as(P(C.c_uint8))


def fptr(a):
    return a.ctypes.data_as(P(C.c_float))


def cfg_default(preset):
    lib = load()
    c = GraCfg()
    check(lib.gra_hip_cfg_default(C.byref(c), preset))
    return c


def grid_xmat(grid):
    """(dim, bins) view of a GraGrid's edges."""
    a = np.ctypeslib.as_array(grid.xmat).reshape(GRA_MAX_DIM, GRA_VEGAS_BINS)
    return a[: grid.dim, : grid.bins]


class PinnedEvents:
    """An array of n gra_event records in pinned host memory (gra_hip_host_alloc),
    the buffer the generation calls copy into at DMA rate. Use as a context
    manager or call close(); .buf is the ctypes array."""

    def __init__(self, n):
        lib = load()
        p = C.c_void_p()
        check(lib.gra_hip_host_alloc(C.sizeof(GraEvent) * n, C.byref(p)))
        self._lib, self._p, self.n = lib, p, n
        self.buf = (GraEvent * n).from_address(p.value)

    def close(self):
        if self._p:
            self.buf = None
            self._lib.gra_hip_host_free(self._p)
            self._p = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()
