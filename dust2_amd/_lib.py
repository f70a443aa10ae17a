"""Loader for the in-tree C-ABI library ``libdust2_amd.so``.

The library is the product: there is no CPU fallback.  Importing this module
fails loudly if the shared object has not been built (``python -c "import
__graft_entry__ as g; g.build()"`` or ``make -C dust2_amd``).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdust2_amd.so")

_lib = None

D = C.POINTER(C.c_double)
U64 = C.POINTER(C.c_uint64)
U32 = C.POINTER(C.c_uint32)
SZ = C.POINTER(C.c_size_t)
VP = C.c_void_p
ERR = C.c_char_p
ERRLEN = C.c_size_t

_PROTOS = {
    "dust2_amd_init": (C.c_int, [ERR, ERRLEN]),
    "dust2_amd_version": (C.c_char_p, []),
    "dust2_filter_sir_alloc": (C.c_int, [D, C.c_size_t, C.c_size_t, C.c_double, D, C.c_size_t,
                                         C.c_double, D, C.c_size_t, U64, C.c_size_t, C.c_uint64,
                                         C.c_int, C.POINTER(VP), ERR, ERRLEN]),
    "dust2_filter_seir16_alloc": (C.c_int, [D, C.c_size_t, C.c_size_t, C.c_double, D,
                                            C.c_size_t, C.c_double, D, C.c_size_t, U64,
                                            C.c_size_t, C.c_uint64, C.c_int, C.POINTER(VP), ERR,
                                            ERRLEN]),
    "dust2_filter_discrete_alloc": (C.c_int, [C.c_char_p, D, C.c_size_t, C.c_size_t,
                                              C.c_double, D, C.c_size_t, C.c_double, D,
                                              C.c_size_t, U64, C.c_size_t, C.c_uint64, C.c_int,
                                              C.POINTER(VP), ERR, ERRLEN]),
    "dust2_filter_run": (C.c_int, [VP, D, C.c_int, D, C.c_size_t, SZ, C.c_size_t, SZ,
                                   C.c_size_t, D, ERR, ERRLEN]),
    "dust2_filter_update_pars": (C.c_int, [VP, D, SZ, C.c_size_t, ERR, ERRLEN]),
    "dust2_filter_last_state": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_filter_last_trajectories": (C.c_int, [VP, C.c_int, D, ERR, ERRLEN]),
    "dust2_filter_last_trajectories_raw": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_filter_last_snapshots": (C.c_int, [VP, C.c_int, D, ERR, ERRLEN]),
    "dust2_filter_last_state_selected": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_filter_last_ancestry": (C.c_int, [VP, U32, ERR, ERRLEN]),
    "dust2_filter_rng_state": (C.c_int, [VP, U64, ERR, ERRLEN]),
    "dust2_filter_set_rng_state": (C.c_int, [VP, U64, ERR, ERRLEN]),
    "dust2_filter_set_launch": (C.c_int, [VP, C.c_int, C.c_int, ERR, ERRLEN]),
    "dust2_filter_last_stats": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_filter_set_engine": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, ERR, ERRLEN]),
    "dust2_filter_set_pipeline_chains": (C.c_int, [VP, C.c_int, ERR, ERRLEN]),
    "dust2_filter_set_pipeline_resample": (C.c_int, [VP, C.c_int, ERR, ERRLEN]),
    "dust2_filter_set_beta_schedule": (C.c_int, [VP, C.c_int, D, C.c_size_t, D, C.c_size_t,
                                                 ERR, ERRLEN]),
    "dust2_filter_autotune": (C.c_int, [VP, SZ, C.c_size_t, D, C.c_size_t, SZ, ERR, ERRLEN]),
    "dust2_filter_free": (None, [VP]),
    "dust2_system_sir_alloc": (C.c_int, [D, C.c_size_t, C.c_size_t, C.c_double, C.c_double, U64,
                                         C.c_size_t, C.c_uint64, C.c_int, C.c_int,
                                         C.POINTER(VP), ERR, ERRLEN]),
    "dust2_system_discrete_alloc": (C.c_int, [C.c_char_p, D, C.c_size_t, C.c_size_t,
                                              C.c_double, C.c_double, U64, C.c_size_t,
                                              C.c_uint64, C.c_int, C.c_int, C.POINTER(VP), ERR,
                                              ERRLEN]),
    "dust2_system_discrete_run_to_time": (C.c_int, [VP, C.c_double, ERR, ERRLEN]),
    "dust2_system_discrete_simulate": (C.c_int, [VP, D, C.c_size_t, D, ERR, ERRLEN]),
    "dust2_system_discrete_state": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_system_discrete_set_state": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_system_discrete_set_state_initial": (C.c_int, [VP, ERR, ERRLEN]),
    "dust2_system_discrete_time": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_system_discrete_set_time": (C.c_int, [VP, C.c_double, ERR, ERRLEN]),
    "dust2_system_discrete_update_pars": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_system_discrete_compare_data": (C.c_int, [VP, D, C.c_size_t, D, ERR, ERRLEN]),
    "dust2_system_discrete_reorder": (C.c_int, [VP, U32, ERR, ERRLEN]),
    "dust2_system_discrete_rng_state": (C.c_int, [VP, U64, ERR, ERRLEN]),
    "dust2_system_discrete_set_rng_state": (C.c_int, [VP, U64, ERR, ERRLEN]),
    "dust2_system_discrete_free": (None, [VP]),
    "dust2_unfilter_sirode_alloc": (C.c_int, [D, C.c_size_t, C.c_size_t, C.c_double, D,
                                              C.c_size_t, D, C.c_size_t, D, C.c_int,
                                              C.POINTER(VP), ERR, ERRLEN]),
    "dust2_unfilter_sirode_run": (C.c_int, [VP, C.c_int, D, C.c_size_t, SZ, C.c_size_t, SZ,
                                            C.c_size_t, D, ERR, ERRLEN]),
    "dust2_unfilter_sirode_last_trajectories": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_unfilter_sirode_last_snapshots": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_unfilter_sirode_update_pars": (C.c_int, [VP, D, SZ, C.c_size_t, ERR, ERRLEN]),
    "dust2_unfilter_sirode_last_state": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_unfilter_sirode_last_stats": (C.c_int, [VP, D, D, ERR, ERRLEN]),
    "dust2_unfilter_sirode_run_gradient": (C.c_int, [VP, C.c_int, D, C.c_size_t, SZ,
                                                     C.c_size_t, SZ, C.c_size_t, D, ERR,
                                                     ERRLEN]),
    "dust2_unfilter_sirode_last_gradient": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_unfilter_sirode_free": (None, [VP]),
    "dust2_unfilter_sir_alloc": (C.c_int, [D, C.c_size_t, C.c_size_t, C.c_double, D,
                                           C.c_size_t, C.c_double, D, C.c_size_t,
                                           C.POINTER(VP), ERR, ERRLEN]),
    "dust2_unfilter_sir_run": (C.c_int, [VP, D, C.c_int, D, C.c_size_t, C.c_int, SZ,
                                         C.c_size_t, SZ, C.c_size_t, D, ERR, ERRLEN]),
    "dust2_unfilter_sir_last_trajectories": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_unfilter_sir_last_snapshots": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_unfilter_sir_last_gradient": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_unfilter_sir_update_pars": (C.c_int, [VP, D, SZ, C.c_size_t, ERR, ERRLEN]),
    "dust2_unfilter_sir_last_state": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_unfilter_sir_set_adjoint_form": (C.c_int, [VP, C.c_int, ERR, ERRLEN]),
    "dust2_unfilter_sir_free": (None, [VP]),
    "dust2_server_create": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(VP), ERR, ERRLEN]),
    "dust2_server_submit": (C.c_int, [VP, C.c_int, D, C.c_size_t, C.c_double, D, C.c_size_t,
                                      C.c_double, D, C.c_uint64, C.c_int64, U64, ERR, ERRLEN]),
    "dust2_amd_rng_jump": (C.c_int, [U64, C.c_size_t, C.c_int, ERR, ERRLEN]),
    "dust2_server_step": (C.c_int, [VP, SZ, SZ, ERR, ERRLEN]),
    "dust2_server_poll": (C.c_int, [VP, U64, D, C.POINTER(C.c_int), C.c_size_t, SZ, ERR,
                                    ERRLEN]),
    "dust2_server_memory": (C.c_int, [VP, SZ, SZ, ERR, ERRLEN]),
    "dust2_server_free": (None, [VP]),
    "dust2_amd_resample": (C.c_int, [D, C.c_size_t, C.c_double, U32, D, ERR, ERRLEN]),
    "dust2_amd_resample_noalloc": (C.c_int, [VP, C.c_size_t, C.c_double, VP, VP, VP, ERR,
                                             ERRLEN]),
    "dust2_amd_trajectory_lineage": (C.c_int, [D, U32, C.c_size_t, C.c_size_t, C.c_size_t,
                                               C.c_size_t, SZ, D, ERR, ERRLEN]),
    "dust2_unfilter_sirode_set_critical_times": (C.c_int, [VP, D, C.c_size_t, ERR, ERRLEN]),
    "dust2_system_sirode_alloc": (C.c_int, [D, C.c_size_t, C.c_size_t, C.c_double, D, D,
                                            C.c_size_t, C.c_size_t, C.POINTER(VP), ERR,
                                            ERRLEN]),
    "dust2_system_sirode_run_to_time": (C.c_int, [VP, C.c_double, ERR, ERRLEN]),
    "dust2_system_sirode_simulate": (C.c_int, [VP, D, C.c_size_t, D, ERR, ERRLEN]),
    "dust2_system_sirode_state": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_system_sirode_set_state": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_system_sirode_set_state_initial": (C.c_int, [VP, ERR, ERRLEN]),
    "dust2_system_sirode_time": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_system_sirode_set_time": (C.c_int, [VP, C.c_double, ERR, ERRLEN]),
    "dust2_system_sirode_update_pars": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_system_sirode_compare_data": (C.c_int, [VP, D, C.c_size_t, D, ERR, ERRLEN]),
    "dust2_system_sirode_reorder": (C.c_int, [VP, U32, ERR, ERRLEN]),
    "dust2_system_sirode_internals": (C.c_int, [VP, D, D, SZ, ERR, ERRLEN]),
    "dust2_system_sirode_free": (None, [VP]),
    "dust2_system_ode_model": (C.c_int, [C.c_char_p, SZ, ERR, ERRLEN]),
    "dust2_system_ode_alloc": (C.c_int, [C.c_char_p, D, C.c_size_t, C.c_size_t, C.c_double,
                                         C.c_double, D, D, C.c_size_t, C.c_size_t, C.c_size_t,
                                         C.c_size_t, C.c_size_t, U64, C.c_size_t, C.c_uint64,
                                         C.c_int, C.POINTER(VP), ERR, ERRLEN]),
    "dust2_system_ode_rng_state": (C.c_int, [VP, U64, ERR, ERRLEN]),
    "dust2_system_ode_set_rng_state": (C.c_int, [VP, U64, ERR, ERRLEN]),
    "dust2_system_ode_compare_data": (C.c_int, [VP, D, C.c_size_t, D, ERR, ERRLEN]),
    "dust2_system_ode_run_to_time": (C.c_int, [VP, C.c_double, ERR, ERRLEN]),
    "dust2_system_ode_simulate": (C.c_int, [VP, D, C.c_size_t, D, ERR, ERRLEN]),
    "dust2_system_ode_state": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_system_ode_set_state": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_system_ode_set_state_initial": (C.c_int, [VP, ERR, ERRLEN]),
    "dust2_system_ode_time": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_system_ode_set_time": (C.c_int, [VP, C.c_double, ERR, ERRLEN]),
    "dust2_system_ode_update_pars": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_system_ode_reorder": (C.c_int, [VP, U32, ERR, ERRLEN]),
    "dust2_system_ode_internals": (C.c_int, [VP, D, D, SZ, ERR, ERRLEN]),
    "dust2_system_ode_events": (C.c_int, [VP, SZ, D, ERR, ERRLEN]),
    "dust2_system_ode_history": (C.c_int, [VP, SZ, D, ERR, ERRLEN]),
    "dust2_system_ode_free": (None, [VP]),
    "dust2_filter_ode_alloc": (C.c_int, [C.c_char_p, D, C.c_size_t, C.c_size_t, C.c_double, D,
                                         C.c_size_t, D, C.c_size_t, C.c_double, D, U64,
                                         C.c_size_t, C.c_uint64, C.POINTER(VP), ERR, ERRLEN]),
    "dust2_filter_ode_run": (C.c_int, [VP, D, C.c_int, D, C.c_size_t, SZ, C.c_size_t, SZ,
                                       C.c_size_t, D, ERR, ERRLEN]),
    "dust2_filter_ode_update_pars": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_filter_ode_last_state": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_filter_ode_last_trajectories": (C.c_int, [VP, C.c_int, D, ERR, ERRLEN]),
    "dust2_filter_ode_last_snapshots": (C.c_int, [VP, C.c_int, D, ERR, ERRLEN]),
    "dust2_filter_ode_last_state_selected": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_amd_valu_probe": (C.c_int, [C.c_int, C.c_int, C.c_int, D, D, C.POINTER(C.c_int),
                                       ERR, ERRLEN]),
    "dust2_filter_ode_rng_state": (C.c_int, [VP, U64, ERR, ERRLEN]),
    "dust2_filter_ode_set_rng_state": (C.c_int, [VP, U64, ERR, ERRLEN]),
    "dust2_filter_ode_free": (None, [VP]),
    "dust2_unfilter_ode_alloc": (C.c_int, [C.c_char_p, D, C.c_size_t, C.c_size_t, C.c_double,
                                           D, C.c_size_t, D, C.c_size_t, C.c_double, D,
                                           C.POINTER(VP), ERR, ERRLEN]),
    "dust2_unfilter_ode_run": (C.c_int, [VP, D, C.c_int, D, C.c_size_t, SZ, C.c_size_t, SZ,
                                         C.c_size_t, D, ERR, ERRLEN]),
    "dust2_unfilter_ode_update_pars": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_unfilter_ode_last_state": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_unfilter_ode_last_trajectories": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_unfilter_ode_last_snapshots": (C.c_int, [VP, D, ERR, ERRLEN]),
    "dust2_unfilter_ode_free": (None, [VP]),
    "dust2_amd_interpolate": (C.c_int, [C.c_int, D, C.c_size_t, D, C.c_size_t, C.c_size_t, D,
                                        C.c_size_t, D, ERR, ERRLEN]),
    "dust2_amd_rng_streams": (C.c_int, [U64, C.c_size_t, C.c_uint64, C.c_size_t, U64, ERR,
                                        ERRLEN]),
    "dust2_amd_sample": (C.c_int, [C.c_int, C.c_int, U64, C.c_size_t, C.c_size_t, D, D, C.c_int,
                                   D, U32, ERR, ERRLEN]),
}


class DustError(RuntimeError):
    """An error reported by the engine (message text follows the reference)."""


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `make -C {_HERE}` (hipcc, gfx950)")
        _lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def declared_symbols():
    return sorted(_PROTOS)


def call(name, *args):
    """Call an entry point that takes (..., err, errlen); raise on failure."""
    buf = C.create_string_buffer(4096)
    rc = getattr(lib(), name)(*args, buf, C.c_size_t(len(buf)))
    if rc != 0:
        raise DustError(buf.value.decode())


def dptr(a):
    return None if a is None else a.ctypes.data_as(D)


def u64ptr(a):
    return None if a is None else a.ctypes.data_as(U64)


def u32ptr(a):
    return None if a is None else a.ctypes.data_as(U32)


def szptr(a):
    return None if a is None else a.ctypes.data_as(SZ)


def f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)
