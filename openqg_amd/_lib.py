"""ctypes binding of liboqg.so (include/oqg.h).

The library is the only compute path: importing this module on a machine with
a GPU but without the built library raises, and no operation falls back to
CPU code.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboqg.so")
MAX_LAYERS = 8

OQG_HOST, OQG_DEVICE = 0, 1
FIELD_Q, FIELD_QM, FIELD_P, FIELD_PM = 0, 1, 2, 3
FLAG_YSOLVE_THOMAS = 1
ENS_ONE_YSOLVE = 1
ENS_LARGEST_FIRST = 2

_dp = ctypes.POINTER(ctypes.c_double)


class OqgDesc(ctypes.Structure):
    """Mirror of `oqg_desc` (include/oqg.h)."""

    _fields_ = [
        ("nxp", ctypes.c_int), ("nyp", ctypes.c_int), ("nl", ctypes.c_int), ("cyclic", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("fnot", ctypes.c_double), ("beta", ctypes.c_double),
        ("h", ctypes.c_double * MAX_LAYERS), ("gp", ctypes.c_double * MAX_LAYERS),
        ("ah2", ctypes.c_double * MAX_LAYERS), ("ah4", ctypes.c_double * MAX_LAYERS),
        ("bcco", ctypes.c_double), ("delek", ctypes.c_double),
        ("dz_sign", ctypes.c_int), ("tau_sign", ctypes.c_int),
        ("k_topo", ctypes.c_int), ("flags", ctypes.c_int),
        ("ddyn", _dp),
        ("amat", _dp), ("ctl2m", _dp), ("ctm2l", _dp), ("rdm2", _dp),
        ("hom_sol_bc", _dp), ("hom_bt", _dp), ("hom_bc1", _dp), ("hom_bc2", _dp), ("hom_coef", _dp),
    ]


_DIAG_LAYER_FIELDS = ("ekei", "pavg", "qavg", "pmin", "pmax", "qmin", "qmax", "sfmin", "sfmax", "circ",
                      "umax", "vmax", "cflu", "cflv", "etam", "et2m")


class OqgDiag(ctypes.Structure):
    """Mirror of `oqg_diag` (include/oqg.h)."""

    _fields_ = [(n, ctypes.c_double * MAX_LAYERS) for n in _DIAG_LAYER_FIELDS] + \
               [("ctot", ctypes.c_double), ("nonfinite", ctypes.c_double), ("pint", ctypes.c_double * MAX_LAYERS)]


class OqgEnergy(ctypes.Structure):
    """Mirror of `oqg_energy` (include/oqg.h)."""

    _fields_ = [(n, ctypes.c_double * MAX_LAYERS) for n in ("ah2d", "ah4d", "val", "pos")] + \
               [(n, ctypes.c_double) for n in ("btdg", "utau", "wepm", "wapm", "entm", "enam", "pken")]


class OqgOmlDesc(ctypes.Structure):
    """Mirror of `oqg_oml_desc` (include/oqg.h)."""

    _fields_ = [(n, ctypes.c_double) for n in ("hm", "T1_rel", "dT_12", "rho_cp", "st2d", "st4d", "ycexp", "nval",
                                                "sval")] + [("fluxn", ctypes.c_int), ("fluxs", ctypes.c_int)]


_CPL_D = ("D0up", "Dmdown", "Dup0", "xlamda", "fspco", "Aup00", "Adown10", "Adown11", "hmamin", "hmadmp", "xcexp",
          "T1_rel", "dT_12", "rho_cp", "at2d", "at4d", "ahmd")
OQG_CPL_OCEAN_ENT, OQG_CPL_ATMOS_ENT = 1, 2


class OqgCplDesc(ctypes.Structure):
    """Mirror of `oqg_cpl_desc` (include/oqg.h)."""

    _fields_ = [(n, ctypes.c_double) for n in ("xo0", "yo0", "xa0", "ya0", "hmo", "hma", "rho_o", "rho_a", "cdat")] + \
               [("tau_udiff", ctypes.c_int)] + [(n, ctypes.c_double) for n in _CPL_D] + \
               [("aface", ctypes.c_double * MAX_LAYERS), ("bface", ctypes.c_double), ("dface", ctypes.c_double),
                ("xc1ast", _dp), ("dtop", _dp)]


# (name, restype, argtypes) for every entry point of include/oqg.h
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
SIGNATURES = [
    ("oqg_create", _i, [ctypes.POINTER(OqgDesc), ctypes.POINTER(_vp)]),
    ("oqg_destroy", None, [_vp]),
    ("oqg_last_error", ctypes.c_char_p, [_vp]),
    ("oqg_set_state", _i, [_vp, _vp, _vp, _vp, _vp, _i]),
    ("oqg_get_state", _i, [_vp, _vp, _vp, _vp, _vp, _i]),
    ("oqg_get_layer", _i, [_vp, _i, _i, _vp, _i]),
    ("oqg_constr_len", _i, [_vp]),
    ("oqg_set_constr", _i, [_vp, _dp]),
    ("oqg_get_constr", _i, [_vp, _dp]),
    ("oqg_init_state", _i, [_vp]),
    ("oqg_set_forcing", _i, [_vp, _vp, _i, _vp, _d, _d, _i]),
    ("oqg_step", _i, [_vp, _d, _vp, _i, _vp, _d, _d]),
    ("oqg_suppress", _i, [_vp]),
    ("oqg_run", _i, [_vp, _d, _i, _i, _i]),
    ("oqg_plan_run", _i, [_vp, _d, _i, _i, _i]),
    ("oqg_sync", _i, [_vp]),
    ("oqg_set_tile_rows", _i, [_vp, _i]),
    ("oqg_get_tile_rows", _i, [_vp]),
    ("oqg_tune", _i, [_vp, _d, _i, _i, _dp]),
    ("oqg_solve_inhomog", _i, [_vp, _i, _vp, _vp, _i]),
    ("oqg_get_modes", _i, [_vp, _dp, _dp, _dp, _dp]),
    ("oqg_set_stream", _i, [_vp, _vp]),
    ("oqg_get_stream", _vp, [_vp]),
    ("oqg_state_ptrs", _i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                            ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_size_t)]),
    ("oqg_kernel_times", _i, [_vp, _d, _i, _dp, _i, ctypes.c_char_p, _i]),
    ("oqg_adj_init", _i, [_vp]),
    ("oqg_adj_set", _i, [_vp, _vp, _vp, _vp, _vp, _i]),
    ("oqg_adj_get", _i, [_vp, _vp, _vp, _vp, _vp, _i]),
    ("oqg_adj_set_constr", _i, [_vp, _dp]),
    ("oqg_adj_get_constr", _i, [_vp, _dp]),
    ("oqg_adj_get_txi", _i, [_vp, _dp]),
    ("oqg_adj_step", _i, [_vp, _d, _i]),
    ("oqg_adj_get_wekp", _i, [_vp, _vp, _i]),
    ("oqg_adj_zero_wekp", _i, [_vp]),
    ("oqg_set_wind", _i, [_vp, _vp, _vp, _i]),
    ("oqg_adj_wind", _i, [_vp, _vp, _vp, _i]),
    ("oqg_get_forcing", _i, [_vp, _vp, _dp, _i]),
    ("oqg_get_ent", _i, [_vp, _vp, _i]),
    ("oqg_get_wind", _i, [_vp, _vp, _vp, _i]),
    ("oqg_diagnose", _i, [_vp, _d, _d, ctypes.POINTER(OqgDiag)]),
    ("oqg_diagnose_noalloc", _i, [_vp, _d, _d, ctypes.POINTER(OqgDiag)]),
    ("oqg_diagnose_energy", _i, [_vp, _d, ctypes.POINTER(OqgEnergy)]),
    ("oqg_adj_seed_ke", _i, [_vp, _d, _dp]),
    ("oqg_invert", _i, [_vp]),
    ("oqg_adj_invert", _i, [_vp]),
    ("oqg_snap_alloc", _i, [_vp, _i]),
    ("oqg_snap_save", _i, [_vp, _i]),
    ("oqg_snap_load", _i, [_vp, _i]),
    ("oqg_adj_record", _i, [_vp, _d, _i, _i, _i, _i]),
    ("oqg_adj_sweep", _i, [_vp, _d, _i, _i, _i, _i]),
    ("oqg_adj_revolve_forward", _i, [_vp, _d, _i, _i, _i]),
    ("oqg_adj_revolve_reverse", _i, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    ("oqg_revolve_cost", ctypes.c_longlong, [_i, _i]),
    ("oqg_revolve_split", _i, [_i, _i]),
    ("oqg_ens_create", _i, [ctypes.POINTER(_vp), _i, ctypes.POINTER(_vp)]),
    ("oqg_ens_destroy", None, [_vp]),
    ("oqg_ens_last_error", ctypes.c_char_p, [_vp]),
    ("oqg_ens_run", _i, [_vp, _dp, ctypes.POINTER(_i), _i, _i]),
    ("oqg_ens_sync", _i, [_vp]),
    ("oqg_ens_set_schedule", _i, [_vp, _i]),
    ("oqg_ens_get_schedule", _i, [_vp]),
    ("oqg_ens_tune", _i, [_vp, _dp, _i, _i, _dp]),
    ("oqg_oml_create", _i, [_vp, ctypes.POINTER(OqgOmlDesc), ctypes.POINTER(_vp)]),
    ("oqg_oml_destroy", None, [_vp]),
    ("oqg_oml_last_error", ctypes.c_char_p, [_vp]),
    ("oqg_oml_set_state", _i, [_vp, _vp, _vp, _i]),
    ("oqg_oml_get_state", _i, [_vp, _vp, _vp, _i]),
    ("oqg_oml_set_fnet", _i, [_vp, _vp, _i]),
    ("oqg_oml_step", _i, [_vp, _d, _i]),
    ("oqg_oml_suppress", _i, [_vp]),
    ("oqg_run_ocean", _i, [_vp, _d, _i, _i, _i, _i]),
    ("oqg_cpl_create", _i, [_vp, _vp, ctypes.POINTER(OqgCplDesc), ctypes.POINTER(_vp)]),
    ("oqg_cpl_destroy", None, [_vp]),
    ("oqg_cpl_last_error", ctypes.c_char_p, [_vp]),
    ("oqg_cpl_set_aml_state", _i, [_vp, _vp, _vp, _vp, _vp, _i]),
    ("oqg_cpl_get_aml_state", _i, [_vp, _vp, _vp, _vp, _vp, _i]),
    ("oqg_cpl_get_fnet", _i, [_vp, _vp, _vp, _i]),
    ("oqg_cpl_xforc", _i, [_vp]),
    ("oqg_cpl_forcing", _i, [_vp]),
    ("oqg_cpl_aml_step", _i, [_vp, _d, _i]),
    ("oqg_cpl_run", _i, [_vp, _i, _i, _d, _d, _i, _i, _i, _i]),
]

_lib = None


class OqgError(RuntimeError):
    pass


def load():
    """Load liboqg.so (built by `python -m openqg_amd.build`); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OqgError(f"{LIB_PATH} is not built; run `python -m openqg_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, ctx=None, what: str = ""):
    if rc != 0:
        msg = load().oqg_last_error(ctx)
        msg = msg.decode() if msg else ""
        raise OqgError(f"{what} failed (code {rc}): {msg}")
