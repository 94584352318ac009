"""ctypes binding of libgravity_hip.so (include/gravity_hip.h).

The library is the product path: it must be built (``python -m
gravity_amd.build``) and loading fails loudly when it is missing. There is
no CPU fallback.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIBRARY = Path(__file__).resolve().parent / "libgravity_hip.so"

GS_OK, GS_EINVAL, GS_ESTATE = 0, -1, -2
GS_BOUNDED, GS_MUTUAL_GRAVITY, GS_LIQUID = 1, 2, 3
GS_BODIES_XYVV, GS_PRESSURES, GS_ACCEL, GS_MATERIALS = 0, 1, 2, 3
GS_CONTACT_COUNTS, GS_SORT_DEST, GS_FORCE_MESH, GS_CONTACTS = 4, 5, 6, 7
GS_RIGID_GROUPS, GS_RIGID_CONTACT_MEMORY, GS_RIGID_MEMBERS = 0, 1, 2
GS_RIGID_REST, GS_RIGID_SHAPES, GS_RIGID_TILE_GROUPS = 3, 4, 5
CONTACT_SLOTS = 16


class GsConfig(ctypes.Structure):
    """Mirror of gs_config (include/gravity_hip.h); field order matters."""

    _fields_ = [
        ("viewport_width", ctypes.c_int),
        ("viewport_height", ctypes.c_int),
        ("bounded", ctypes.c_int),
        ("mutual_gravity", ctypes.c_int),
        ("liquid", ctypes.c_int),
        ("capacity", ctypes.c_int),
        ("spatial_sort_interval", ctypes.c_int),
        ("grid_width", ctypes.c_int),
        ("grid_height", ctypes.c_int),
        ("gravitational_constant", ctypes.c_float),
        ("softening_cells", ctypes.c_float),
        ("local_search_sigma_multiplier", ctypes.c_float),
        ("minimum_neighbor_radius_cells", ctypes.c_int),
        ("maximum_neighbor_radius_cells", ctypes.c_int),
        ("exact_neighbor_body_limit", ctypes.c_int),
        ("maximum_home_body_interactions", ctypes.c_int),
        ("minimum_distance_cell_fraction", ctypes.c_float),
        ("minimum_assignment_window_power", ctypes.c_float),
        ("time_step", ctypes.c_float),
        ("maximum_substeps", ctypes.c_int),
        ("contact_diameter_pixels", ctypes.c_float),
        ("neighbor_cell_radius", ctypes.c_int),
        ("restitution_coefficient", ctypes.c_float),
        ("friction", ctypes.c_float),
        ("coincident_distance_squared_epsilon", ctypes.c_float),
        ("neighbor_skin_pixels", ctypes.c_float),
        ("sleep", ctypes.c_int),
        ("sleep_speed_pixels", ctypes.c_float),
        ("kernel_radius_diameters", ctypes.c_float),
        ("liquid_iterations", ctypes.c_int),
        ("liquid_viscosity", ctypes.c_float),
        ("field_acceleration_x", ctypes.c_float),
        ("field_acceleration_y", ctypes.c_float),
        ("collisions_enabled", ctypes.c_int),
        ("adaptive_collisions", ctypes.c_int),
        ("allowed_displacement_fraction", ctypes.c_float),
        ("allowed_penetration_fraction", ctypes.c_float),
        ("dense_contact_threshold", ctypes.c_int),
        ("dense_body_fraction", ctypes.c_float),
        ("minimum_dense_bodies", ctypes.c_int),
        ("trajectory_steps", ctypes.c_int),
        ("trajectory_maximum_time_step", ctypes.c_float),
        ("trajectory_target_segment_pixels", ctypes.c_float),
        ("trajectory_maximum_path_diagonals", ctypes.c_float),
        ("trajectory_minimum_dash_pixels", ctypes.c_float),
        ("block_friction", ctypes.c_float),
    ]


class GsClusterStatistics(ctypes.Structure):
    """Mirror of gs_cluster_statistics (ClusterStatistics, K/GravityPipeline.kt:47-57)."""
    _fields_ = [("center_x", ctypes.c_double), ("center_y", ctypes.c_double),
                ("covariance_area_pixels", ctypes.c_double), ("equivalent_diameter_pixels", ctypes.c_double),
                ("core90_area_pixels", ctypes.c_double), ("core99_area_pixels", ctypes.c_double),
                ("max_speed_pixels_per_second", ctypes.c_double),
                ("average_speed_pixels_per_second", ctypes.c_double), ("non_finite_bodies", ctypes.c_int)]


class GsAdaptiveStatistics(ctypes.Structure):
    """Mirror of gs_adaptive_statistics (AdaptiveCollisionStatistics, K/GravityPipeline.kt:65-71)."""
    _fields_ = [("maximum_speed_pixels_per_second", ctypes.c_float),
                ("maximum_penetration_pixels", ctypes.c_float), ("substeps", ctypes.c_int),
                ("maximum_contact_count", ctypes.c_int), ("dense_contact_bodies", ctypes.c_int)]


class GsTuning(ctypes.Structure):
    """Mirror of gs_tuning: the step's launch shapes (no reference counterpart)."""
    _fields_ = [("big_segment_limit", ctypes.c_int), ("forward_item_bodies", ctypes.c_int),
                ("deposit_runs", ctypes.c_int), ("scan_direct_blocks", ctypes.c_int)]


class GsKernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 40), ("launches", ctypes.c_int), ("total_ms", ctypes.c_double)]


# Every symbol include/gravity_hip.h declares, with its signature.
_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
SIGNATURES = {
    "gs_config_defaults": (None, [ctypes.POINTER(GsConfig)]),
    "gs_create": (_I, [ctypes.POINTER(GsConfig), _P, _I, ctypes.POINTER(_P)]),
    "gs_destroy": (None, [_P]),
    "gs_set_bool": (_I, [_P, _I, _I]),
    "gs_set_field": (_I, [_P, _F, _F]),
    "gs_set_substeps": (_I, [_P, _I]),
    "gs_initialize_acceleration": (_I, [_P, _I]),
    "gs_step": (_I, [_P, _I]),
    "gs_upload": (_I, [_P, _I, _I, _P, _P]),
    "gs_upload_accel": (_I, [_P, _I, _I, _P]),
    "gs_download": (_I, [_P, _I, _I, _I, _P]),
    "gs_stop_bodies": (_I, [_P, _I]),
    "gs_wake": (_I, [_P]),
    "gs_debug_resolve_collisions": (_I, [_P, _I, _F, _F, _I, _I]),
    "gs_debug_sort": (_I, [_P, _I]),
    "gs_debug_resolve_liquid": (_I, [_P, _I]),
    "gs_debug_density": (_I, [_P, _I, _P, _P, _P]),
    "gs_debug_density_noalloc": (_I, [_P, _I, _P, _P, _P]),
    "gs_validate_fft": (_I, [_P, _I, ctypes.POINTER(ctypes.c_double)]),
    "gs_projectile_body_count": (_I, [_P, _F, ctypes.POINTER(_I)]),
    "gs_spawn_projectile": (_I, [_P, _I, _F, _F, _F, _F, _F, ctypes.POINTER(_I)]),
    "gs_projectile_offsets": (_I, [_F, _F, _P, _I, ctypes.POINTER(_I)]),
    "gs_projectile_diameter_for_body_count": (_I, [_F, _I, ctypes.POINTER(_F)]),
    "gs_measure_cluster": (_I, [_P, _I, ctypes.POINTER(GsClusterStatistics)]),
    "gs_supports_walls": (_I, [_I, _I]),
    "gs_measure_adaptive_collisions": (_I, [_P, _I, ctypes.POINTER(GsAdaptiveStatistics)]),
    "gs_update_trajectory": (_I, [_P, _I, _F, _F, _F, _F]),
    "gs_trajectory_view": (_I, [_P, ctypes.POINTER(_P), ctypes.POINTER(_I)]),
    "gs_download_trajectory": (_I, [_P, _I, _I, _P]),
    "gs_render_view": (_I, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    "gs_render_copy": (_I, [_P, _I, _P, _P]),
    "gs_rigid_glue": (_I, [_P, _I, _I, _P, _F, _F, ctypes.POINTER(_I)]),
    "gs_spawn_block": (_I, [_P, _I, _F, _F, _F, _F, _F, _F, _F, _F, ctypes.POINTER(_I)]),
    "gs_block_offsets": (_I, [_F, _F, _F, _P, _I, ctypes.POINTER(_I)]),
    "gs_rigid_project": (_I, [_P, _I]),
    "gs_rigid_size": (_I, [_P, _P]),
    "gs_rigid_read": (_I, [_P, _I, _I, _I, _P]),
    "gs_rigid_write": (_I, [_P, _I, _I, _I, _P]),
    "gs_synchronize": (_I, [_P]),
    "gs_query": (_I, [_P, ctypes.POINTER(_I)]),
    "gs_profile_steps": (_I, [_P, _I, ctypes.POINTER(GsKernelStat), _I, ctypes.POINTER(_I)]),
    "gs_set_graphs": (_I, [_P, _I]),
    "gs_get_tuning": (_I, [_P, ctypes.POINTER(GsTuning)]),
    "gs_set_tuning": (_I, [_P, ctypes.POINTER(GsTuning)]),
    "gs_autotune": (_I, [_P, _I, ctypes.POINTER(GsTuning), ctypes.POINTER(ctypes.c_double)]),
    "gs_batch_create": (_I, [ctypes.POINTER(GsConfig), _I, ctypes.POINTER(_P), ctypes.POINTER(_I),
                             ctypes.POINTER(_P)]),
    "gs_batch_size": (_I, [_P]),
    "gs_batch_set_active": (_I, [_P, _I, _I]),
    "gs_batch_upload": (_I, [_P, _I, _I, _I, _P, _P]),
    "gs_batch_download": (_I, [_P, _I, _I, _I, _I, _P]),
    "gs_batch_initialize_acceleration": (_I, [_P]),
    "gs_batch_step": (_I, [_P]),
    "gs_adjoint_record": (_I, [_P, _I, _I, _I]),
    "gs_adjoint_vjp": (_I, [_P, _I, _P, _P, _P]),
    "gs_adjoint_order": (_I, [_P, _I, _P]),
    "gs_vjp_initial_velocity": (_I, [_P, _I, _I, _I, _P, ctypes.POINTER(_F), _P]),
    "gs_debug_acceleration_vjp": (_I, [_P, _I, _P, _P]),
    "gs_last_error": (ctypes.c_char_p, []),
}

_lib = None


def load() -> ctypes.CDLL:
    """Loads the HIP library; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY.exists():
        raise RuntimeError(
            f"{LIBRARY} is missing: build it with `python -m gravity_amd.build` "
            "(the MI355X path has no CPU fallback)")
    lib = ctypes.CDLL(str(LIBRARY))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class GravityError(RuntimeError):
    """A failed call; ``code`` is GS_EINVAL or GS_ESTATE."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def check(code: int) -> None:
    if code != GS_OK:
        msg = load().gs_last_error().decode(errors="replace")
        if code == GS_EINVAL:
            raise ValueError(msg)
        raise GravityError(code, msg)


def defaults() -> GsConfig:
    cfg = GsConfig()
    load().gs_config_defaults(ctypes.byref(cfg))
    return cfg


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)
