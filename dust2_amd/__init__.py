"""dust2_amd: MI355X-native batched particle engine for dust2 models.

The compute path is the C-ABI library ``libdust2_amd.so`` (HIP, gfx950);
this package is the host-side mirror of the reference's R interface.
"""
from ._lib import DustError, LIB_PATH, declared_symbols, lib  # noqa: F401
from .interface import (  # noqa: F401
    FP32,
    FP64,
    Generator,
    dust_filter_create,
    dust_likelihood_copy,
    dust_likelihood_ensure_initialised,
    dust_likelihood_last_ancestry,
    dust_ode_control,
    dust_unfilter_create,
    dust_likelihood_last_gradient,
    dust_likelihood_last_state,
    dust_likelihood_last_stats,
    dust_likelihood_last_trajectories,
    dust_likelihood_last_trajectories_raw,
    dust_likelihood_last_snapshots,
    dust_likelihood_rng_state,
    dust_likelihood_run,
    dust_likelihood_set_rng_state,
    dust_system_compare_data,
    dust_system_create,
    dust_system_internals,
    dust_system_reorder,
    dust_system_rng_state,
    dust_system_run_to_time,
    dust_system_set_rng_state,
    dust_system_set_state,
    dust_system_set_state_initial,
    dust_system_set_time,
    dust_system_simulate,
    dust_system_state,
    dust_system_time,
    dust_system_update_pars,
    gpu_interpolate,
    gpu_resample,
    gpu_trajectory_lineage,
    gpu_rng_streams,
    gpu_sample,
    seir16,
    sir,
    sir_time_varying,
    sirode,
)

__version__ = "0.1.0"
