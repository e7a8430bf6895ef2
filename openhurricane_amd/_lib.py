"""ctypes bindings of libohc.so (the C-ABI declared in include/ohc.h).

The library is loaded from the package directory only; a missing library is
an error (there is no fallback implementation).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIBPATH = PKG / "libohc.so"

OHC_OK = 0
OHC_ERR_ARG = 1
OHC_ERR_PARSE = 2
OHC_ERR_DEVICE = 3
OHC_ERR_UNSUPPORTED = 4
OHC_ERR_NUMERIC = 5
OHC_MAXP = 6

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class MechTables(C.Structure):
    _fields_ = [
        ("ns", C.c_int32), ("nr", C.c_int32), ("ne", C.c_int32),
        ("W", _dp), ("Tlow", _dp), ("Thigh", _dp), ("Tcommon", _dp),
        ("nasa_lo", _dp), ("nasa_hi", _dp), ("elem", _dp), ("ekb", _dp), ("sigma", _dp),
        ("nf", _ip), ("nb", _ip), ("f_idx", _ip), ("b_idx", _ip),
        ("f_nu", _dp), ("b_nu", _dp), ("f_ord", _dp), ("b_ord", _dp),
        ("rev", _ip), ("kind", _ip), ("arr", _dp), ("arr_lo", _dp), ("arr_rev", _dp),
        ("fo_type", _ip), ("fo", _dp), ("eff", _dp), ("aux_off", _ip), ("aux", _dp),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("flow_type", C.c_int32), ("max_steps", C.c_int32),
        ("atol", C.c_double), ("rtol", C.c_double),
        ("atol_inc", C.c_double), ("rtol_inc", C.c_double),
        ("safe_scale", C.c_double), ("min_scale", C.c_double), ("max_scale", C.c_double),
        ("alpha_inc", C.c_double), ("T_threshold", C.c_double), ("T_alpha", C.c_double),
        ("T_beta", C.c_double), ("dt_max", C.c_double), ("T_low", C.c_double),
        ("T_high", C.c_double), ("precision", C.c_int32), ("device_pointers", C.c_int32),
        ("kernel", C.c_int32), ("waves_per_cu", C.c_int32), ("schedule", C.c_int32),
        ("jacobian", C.c_int32), ("integrator", C.c_int32), ("max_slice_cells", C.c_int64),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("cells", C.c_int64), ("steps", C.c_int64), ("attempts", C.c_int64),
        ("newton_iters", C.c_int64), ("rhs_evals", C.c_int64), ("max_cell_steps", C.c_int64),
        ("dt_min", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes)
_V = C.c_void_p
_SIGS = [
    ("ohc_opts_default", C.c_int, [C.POINTER(Opts)]),
    ("ohc_mech_load", C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(_V)]),
    ("ohc_mech_from_tables", C.c_int, [C.POINTER(MechTables), C.POINTER(_V)]),
    ("ohc_mech_get_tables", C.c_int, [_V, C.POINTER(MechTables)]),
    ("ohc_mech_species_name", C.c_int, [_V, C.c_int32, C.POINTER(C.c_char_p)]),
    ("ohc_mech_synthetic", C.c_int, [C.c_int32, C.c_int32, C.c_uint64, C.c_double,
                                     C.c_double, C.POINTER(_V)]),
    ("ohc_mech_hash", C.c_int, [_V, C.POINTER(C.c_uint64)]),
    ("ohc_mech_free", C.c_int, [_V]),
    ("ohc_ctx_create", C.c_int, [_V, C.POINTER(Opts), C.c_int32, C.POINTER(_V)]),
    ("ohc_ctx_free", C.c_int, [_V]),
    ("ohc_ctx_set_stream", C.c_int, [_V, _V]),
    ("ohc_ctx_autotune", C.c_int, [_V, C.c_int64, _V, _V, _V, _V, C.c_int64, _V, C.c_double,
                                   C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    ("ohc_sync", C.c_int, [_V]),
    ("ohc_last_error", C.c_char_p, [_V]),
    ("ohc_ctx_kernel_name", C.c_char_p, [_V]),
    ("ohc_chem_solve", C.c_int, [_V, C.c_int64, _V, _V, _V, _V, C.c_int64, _V, C.c_double,
                                 _V, _V, C.c_int64, C.POINTER(Stats)]),
    ("ohc_chem_solve_update", C.c_int, [_V, C.c_int64, _V, _V, _V, _V, C.c_int64, _V,
                                        C.c_double, _V, C.POINTER(Stats)]),
    ("ohc_chem_source", C.c_int, [_V, C.c_int64, _V, _V, _V, _V, C.c_int64, _V, C.c_int64,
                                  _V, C.c_int64, C.c_int32]),
    ("ohc_eval_rhs", C.c_int, [_V, C.c_int64, _V, C.c_int64, _V, _V, C.c_int64]),
    ("ohc_eval_jacobian", C.c_int, [_V, C.c_int64, _V, C.c_int64, _V, _V, C.c_int64, _V,
                                    C.c_int64]),
    ("ohc_lu_solve", C.c_int, [_V, C.c_int64, C.c_int32, _V, C.c_int64, _V, C.c_int64, _V,
                               C.c_int64]),
    ("ohc_thermo_T_from_h", C.c_int, [_V, C.c_int64, _V, _V, _V, C.c_int64, _V, _V]),
    ("ohc_mix_props", C.c_int, [_V, C.c_int64, _V, _V, _V, C.c_int64, _V, _V, _V, C.c_int64,
                                _V, _V, _V]),
    ("ohc_thermo_T_from_e", C.c_int, [_V, C.c_int64, _V, _V, _V, C.c_int64, _V, _V]),
    ("ohc_mix_thermo", C.c_int, [_V, C.c_int64, _V, _V, _V, C.c_int64, _V, _V, _V, _V]),
    ("ohc_chem_source_kim", C.c_int, [_V, C.c_int64, _V, _V, _V, _V, C.c_int64, _V, C.c_int64,
                                      _V, C.c_int64, C.c_int32]),
    ("ohc_chem_solve_update_noalloc", C.c_int, [_V, C.c_int64, _V, _V, _V, _V, C.c_int64, _V,
                                                C.c_double, _V, C.POINTER(Stats)]),
]
OHC_NOALLOC_MAX_CELLS = 4194304

_FP = C.POINTER(C.c_float)
_SIGS += [
    ("ohc_surrogate_create", C.c_int, [_V, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(_V)]),
    ("ohc_surrogate_free", C.c_int, [_V]),
    ("ohc_surrogate_num_params", C.c_int, [_V, C.POINTER(C.c_int64)]),
    ("ohc_surrogate_get_params", C.c_int, [_V, _V]),
    ("ohc_surrogate_set_params", C.c_int, [_V, _V]),
    ("ohc_surrogate_set_norm", C.c_int, [_V, _V, _V, _V, _V, _V, _V, _V, C.c_double]),
    ("ohc_surrogate_features", C.c_int, [_V, C.c_int64, _V, _V, _V, C.c_int64, _V, _V, _V, _V]),
    ("ohc_surrogate_train_step", C.c_int, [_V, C.c_int64, _V, _V, C.c_float, _FP]),
    ("ohc_surrogate_gradient", C.c_int, [_V, C.c_int64, _V, _V, _V, _FP]),
    ("ohc_surrogate_predict", C.c_int, [_V, C.c_int64, _V, _V, _V, C.c_int64, _V]),
    ("ohc_chem_solve_update_surrogate", C.c_int, [_V, _V, C.c_int64, _V, _V, _V, _V, C.c_int64,
                                                  _V, C.c_double, _V, C.c_double,
                                                  C.POINTER(C.c_int64), C.POINTER(Stats)]),
    ("ohc_surrogate_set_target_scale", C.c_int, [_V, _V]),
    ("ohc_surrogate_set_precision", C.c_int, [_V, C.c_int32]),
    ("ohc_surrogate_set_peer", C.c_int, [_V, _V, C.c_double]),
    ("ohc_surrogate_save", C.c_int, [_V, C.c_char_p]),
    ("ohc_surrogate_load", C.c_int, [_V, C.c_char_p]),
]

EXPORTS = [s[0] for s in _SIGS]

_lib = None


class OhcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ohc error {code}: {msg}")
        self.code = code


def lib():
    """Return the loaded libohc.so, raising if it has not been built."""
    global _lib
    if _lib is None:
        if not LIBPATH.exists():
            raise OhcError(OHC_ERR_DEVICE, f"{LIBPATH} not found: run "
                           "`python -m openhurricane_amd.build` first")
        L = C.CDLL(str(LIBPATH))
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, ctx=None):
    if rc != OHC_OK:
        msg = lib().ohc_last_error(ctx)
        raise OhcError(rc, msg.decode() if msg else "")
    return rc


def default_opts(**kw) -> Opts:
    o = Opts()
    check(lib().ohc_opts_default(C.byref(o)))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o
