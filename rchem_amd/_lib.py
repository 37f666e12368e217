"""ctypes binding of librchem_gpu.so (the C ABI in include/rchem_gpu.h).

The shared library is built in-tree (``make -C rchem_amd/csrc``).  Loading it
never falls back to anything: if it is missing, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# RCHEM_LIB selects another build of the same library (an experimental build
# kept out of the tree); the default is the in-tree one.
LIB_PATH = os.environ.get("RCHEM_LIB") or os.path.join(HERE, "librchem_gpu.so")

RCHEM_OK = 0
STATUS_NAMES = {0: "OK", 1: "ERR_INVALID", 2: "ERR_HIP", 3: "ERR_NO_DEVICE",
                4: "ERR_UNSUPPORTED", 5: "ERR_NO_BASIS", 6: "ERR_NOT_CONVERGED",
                7: "ERR_NONFINITE", 8: "ERR_ALLOC"}
BOYS_EXACT = 0
BOYS_COMPAT = 1
PREC_FP64 = 0
PREC_MIXED = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Opts(ctypes.Structure):
    _fields_ = [("boys_mode", ctypes.c_int), ("precision", ctypes.c_int),
                ("schwarz_thresh", ctypes.c_double), ("mixed_thresh", ctypes.c_double)]


class JKStats(ctypes.Structure):
    _fields_ = [("quartets", ctypes.c_double), ("primitive_quartets", ctypes.c_double),
                ("integrals", ctypes.c_double), ("flops", ctypes.c_double),
                ("quartets_fp32", ctypes.c_double), ("tiles", ctypes.c_double)]


class ClassTiming(ctypes.Structure):
    _fields_ = [("la", ctypes.c_int), ("lb", ctypes.c_int), ("lc", ctypes.c_int),
                ("ld", ctypes.c_int), ("fp32", ctypes.c_int), ("launches", ctypes.c_int),
                ("ms", ctypes.c_double),
                ("quartets", ctypes.c_double), ("primitive_quartets", ctypes.c_double),
                ("flops", ctypes.c_double)]


class SCFOpts(ctypes.Structure):
    _fields_ = [("nocc", ctypes.c_int), ("max_iter", ctypes.c_int),
                ("thresh_e", ctypes.c_double), ("jk", Opts), ("verbose", ctypes.c_int),
                ("diis", ctypes.c_int), ("graph", ctypes.c_int),
                ("purify", ctypes.c_int), ("damping", ctypes.c_double),
                ("damping_until", ctypes.c_double)]


class SCFResult(ctypes.Structure):
    _fields_ = [("e_elec", ctypes.c_double), ("e_nuc", ctypes.c_double),
                ("e_total", ctypes.c_double), ("iterations", ctypes.c_int),
                ("converged", ctypes.c_int), ("fp64_from", ctypes.c_int),
                ("graph_iterations", ctypes.c_int)]


# name -> (restype, argtypes); every symbol include/rchem_gpu.h declares.
SIGNATURES = {
    "rchem_opts_default": (ctypes.c_int, [ctypes.POINTER(Opts)]),
    "rchem_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "rchem_ctx_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "rchem_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "rchem_basis_from_atoms": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip, _dp,
                                              ctypes.c_char_p]),
    "rchem_basis_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _ip, _ip, _dp,
                                          _dp]),
    "rchem_basis_nbf": (ctypes.c_int, [ctypes.c_void_p, _ip]),
    "rchem_basis_dims": (ctypes.c_int, [ctypes.c_void_p, _ip, _ip]),
    "rchem_basis_shells": (ctypes.c_int, [ctypes.c_void_p, _dp, _ip, _ip, _dp, _dp]),
    "rchem_one_electron": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp,
                                          _dp]),
    "rchem_jk": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, ctypes.POINTER(Opts)]),
    "rchem_jk_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.POINTER(Opts)]),
    "rchem_jk_device_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.POINTER(Opts)]),
    "rchem_eri_tensor": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(Opts)]),
    "rchem_schwarz": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(Opts)]),
    "rchem_eri_shell_quartets": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip, _dp,
                                                ctypes.POINTER(Opts)]),
    "rchem_jk_last_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(JKStats)]),
    "rchem_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "rchem_reset_class_timings": (ctypes.c_int, [ctypes.c_void_p]),
    "rchem_jk_class_timings": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ClassTiming),
                                              ctypes.c_int, _ip]),
    "rchem_scf_opts_default": (ctypes.c_int, [ctypes.POINTER(SCFOpts)]),
    "rchem_lds_variant_info": (ctypes.c_int, [ctypes.c_int, _ip, _ip, _ip, _ip, _ip]),
    "rchem_set_class_tuning": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 8),
    "rchem_set_class_ket_order": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 6),
    "rchem_get_class_ket_order": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 5 +
                                  [ctypes.POINTER(ctypes.c_int)]),
    "rchem_load_batch_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p,
                                                ctypes.POINTER(ctypes.c_int)]),
    "rchem_get_class_tuning": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 5 +
                               [_ip, _ip, _ip]),
    "rchem_load_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _ip]),
    # cints.h compatibility exports (include/rchem_cints_compat.h)
    "coulomb_repulsion": (ctypes.c_double, ([ctypes.c_double] * 4 + [ctypes.c_int] * 3 +
                                            [ctypes.c_double]) * 4),
    "contr_coulomb": (ctypes.c_double, ([ctypes.c_int, _dp, _dp, _dp] + [ctypes.c_double] * 3 +
                                        [ctypes.c_int] * 3) * 4),
    "overlap": (ctypes.c_double, ([ctypes.c_double] + [ctypes.c_int] * 3 +
                                  [ctypes.c_double] * 3) * 2),
    "kinetic": (ctypes.c_double, ([ctypes.c_double] + [ctypes.c_int] * 3 +
                                  [ctypes.c_double] * 3) * 2),
    "nuclear_attraction": (ctypes.c_double, ([ctypes.c_double] * 4 + [ctypes.c_int] * 3 +
                                             [ctypes.c_double]) * 2 + [ctypes.c_double] * 3),
    "Fgamma": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
    "rchem_batch_scf": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip, _ip, _dp,
                                       ctypes.c_char_p, ctypes.POINTER(SCFOpts),
                                       ctypes.POINTER(SCFResult)]),
    "rchem_rhf_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp,
                                          ctypes.POINTER(SCFOpts), ctypes.POINTER(SCFResult),
                                          _dp]),
    "rchem_scf": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp,
                                 ctypes.POINTER(SCFOpts), ctypes.POINTER(SCFResult), _dp]),
}

_lib = None


class RchemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def lib():
    """Loads librchem_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`make -C rchem_amd/csrc` (there is no fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().rchem_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o
