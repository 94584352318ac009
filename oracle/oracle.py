"""ctypes wrapper of the CPU oracle (oracle/liboracle.so): TEST INFRASTRUCTURE
ONLY. Loaded by tests/, __graft_entry__.smoke() and the CPU-baseline leg of
bench.py, as the checker; the product path never touches it."""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "liboracle.so"


class OrcTrajectoryParams(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("maximum_dt", ctypes.c_float), ("segment_px", ctypes.c_float),
                ("path_diagonals", ctypes.c_float), ("minimum_dash_px", ctypes.c_float)]


class OrcRigidParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_float), ("height", ctypes.c_float), ("bounded", ctypes.c_int),
                ("contacts", ctypes.c_int), ("group_count", ctypes.c_int), ("tile_count", ctypes.c_int),
                ("substep_dt", ctypes.c_float), ("body_radius", ctypes.c_float),
                ("block_friction", ctypes.c_float)]


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("height", ctypes.c_int), ("bounded", ctypes.c_int),
        ("mutual_gravity", ctypes.c_int), ("liquid", ctypes.c_int), ("collide", ctypes.c_int),
        ("substeps", ctypes.c_int), ("dt", ctypes.c_float), ("diameter", ctypes.c_float),
        ("restitution", ctypes.c_float), ("coincident_eps", ctypes.c_float),
        ("grid_w", ctypes.c_int), ("grid_h", ctypes.c_int), ("gravity", ctypes.c_float),
        ("softening_cells", ctypes.c_float), ("near_sigma", ctypes.c_float),
        ("near_min", ctypes.c_int), ("near_max", ctypes.c_int), ("exact_limit", ctypes.c_int),
        ("home_limit", ctypes.c_int), ("min_dist_frac", ctypes.c_float),
        ("window_floor", ctypes.c_float), ("field_x", ctypes.c_float), ("field_y", ctypes.c_float),
        ("sort_interval", ctypes.c_int), ("kernel_radius_diameters", ctypes.c_float),
        ("liquid_iterations", ctypes.c_int), ("viscosity", ctypes.c_float),
        ("sleep", ctypes.c_int), ("sleep_speed", ctypes.c_float),
    ]


def build() -> Path:
    res = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return LIBRARY


_lib = None
_P = ctypes.c_void_p


def load():
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            build()
        lib = ctypes.CDLL(str(LIBRARY))
        P, I, F, U = _P, ctypes.c_int, ctypes.c_float, ctypes.c_uint32
        sig = {
            "orc_spread_bits": (U, [U]),
            "orc_sort_key": (U, [F, F, I, I]),
            "orc_sort_axis_bits": (I, [I]),
            "orc_spatial_sort": (None, [I, I, I, P, P, P, P]),
            "orc_collision_grid_cells": (I, [I, F]),
            "orc_pair_distance_squared": (F, [U, P, U, P, F, F, I]),
            "orc_contacts": (None, [P, I, P, P, P, P, I]),
            "orc_drift": (None, [P, I, P, P, F, F]),
            "orc_solve": (None, [P, I, P, P, P, P, P, I, P, I]),
            "orc_collide_pass": (None, [P, I, P, P, P, P, F, F, I, P]),
            "orc_mass_scale": (I, [F]),
            "orc_twiddles": (None, [I, P]),
            "orc_twiddles_f64": (None, [I, P]),
            "orc_fft2d": (None, [P, I, I, P, I]),
            "orc_fft2d_f64": (None, [P, I, I, P, I]),
            "orc_kernel_init": (None, [I, I, I, I, F, F, F, F, P, P]),
            "orc_convolve": (None, [I, I, F, P, P, P, P, P]),
            "orc_deposit": (None, [I, P, P, I, I, I, I, I, P, P, P]),
            "orc_neighbor_radius": (None, [P, P, P]),
            "orc_sim_create": (P, [P, I, P, P]),
            "orc_sim_destroy": (None, [P]),
            "orc_sim_set_bodies": (None, [P, P]),
            "orc_sim_get": (None, [P, P, P, P, P]),
            "orc_sim_set_accel": (None, [P, P]),
            "orc_sim_build_force": (None, [P]),
            "orc_sim_forces": (None, [P, P, P]),
            "orc_sim_density_spectrum": (None, [P, P]),
            "orc_sim_kick": (None, [P, I]),
            "orc_sim_initialize_acceleration": (None, [P]),
            "orc_sim_step": (None, [P]),
            "orc_sim_wake": (None, [P]),
            "orc_rigid_hull": (I, [P, P, I, P]),
            "orc_rigid_prepare": (None, [P, I, P, P, ctypes.POINTER(ctypes.c_double), P]),
            "orc_rigid_project": (None, [ctypes.POINTER(OrcRigidParams), P, P, P, P, P, P, P]),
            "orc_grav64_loss": (ctypes.c_double, [P, I, P, P, I, P, P, P]),
            "orc_sim_sleep_stamp": (U, [P]),
            "orc_sim_body_awake": (None, [P, P]),
            "orc_sim_steps": (ctypes.c_long, [P]),
            "orc_sim_last_sort_dest": (None, [P, P]),
            "orc_liquid_constants": (None, [F, F, P, P, P, P, P]),
            "orc_sim_resolve_liquid": (None, [P]),
            "orc_sim_trajectory": (None, [P, F, F, F, F, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def params(width, height, *, bounded=False, mutual_gravity=True, liquid=False, collide=True, substeps=16,
           dt=1.0 / 120.0, diameter=1.12838, restitution=0.0, coincident_eps=1e-12, grid=(512, 256),
           gravity=0.035, softening=0.85, near_sigma=3.0, near_min=1, near_max=6, exact_limit=2,
           home_limit=8, min_dist_frac=0.05, window_floor=0.08, field=(0.0, 0.0), sort_interval=2,
           kernel_radius=2.0, liquid_iterations=12, viscosity=0.1, sleep=False,
           sleep_speed=6.0) -> OrcParams:
    p = OrcParams()
    p.width, p.height, p.bounded = width, height, int(bounded)
    p.mutual_gravity, p.liquid, p.collide = int(mutual_gravity), int(liquid), int(collide)
    p.substeps, p.dt, p.diameter = substeps, dt, diameter
    p.restitution, p.coincident_eps = restitution, coincident_eps
    p.grid_w, p.grid_h = grid
    p.gravity, p.softening_cells, p.near_sigma = gravity, softening, near_sigma
    p.near_min, p.near_max, p.exact_limit, p.home_limit = near_min, near_max, exact_limit, home_limit
    p.min_dist_frac, p.window_floor = min_dist_frac, window_floor
    p.field_x, p.field_y = field
    p.sort_interval = sort_interval
    p.kernel_radius_diameters, p.liquid_iterations, p.viscosity = kernel_radius, liquid_iterations, viscosity
    p.sleep, p.sleep_speed = int(sleep), sleep_speed
    return p


def grav64_loss(p: OrcParams, bodies, mats, steps, targets, dv=None, final=False):
    """FP64 gravity-only leapfrog loss (and final positions if final=True)."""
    b = np.ascontiguousarray(bodies, dtype=np.float32)
    m = np.ascontiguousarray(mats, dtype=np.float32)
    t = np.ascontiguousarray(targets, dtype=np.float32)
    d = None if dv is None else np.ascontiguousarray(dv, dtype=np.float64)
    out = np.empty((len(b), 2), np.float64) if final else None
    loss = load().orc_grav64_loss(ctypes.byref(p), len(b), _p(b), _p(m), int(steps), _p(t),
                                  _p(d) if d is not None else None, _p(out) if final else None)
    return (loss, out) if final else loss


def sort_key(x, y, wb, hb) -> int:
    return load().orc_sort_key(float(x), float(y), int(wb), int(hb))


def sort_axis_bits(extent) -> int:
    return load().orc_sort_axis_bits(int(extent))


def spatial_sort(bodies, accel, mats, width, height):
    b = np.ascontiguousarray(bodies, dtype=np.float32).copy()
    a = np.ascontiguousarray(accel, dtype=np.float32).copy()
    m = np.ascontiguousarray(mats, dtype=np.float32).copy()
    n = len(b)
    dest = np.empty(n, dtype=np.uint32)
    load().orc_spatial_sort(n, width, height, _p(b), _p(a), _p(m), _p(dest))
    return b, a, m, dest


def contacts(p: OrcParams, bodies, mats=None, cap=64):
    b = np.ascontiguousarray(bodies, dtype=np.float32)
    n = len(b)
    m = np.ascontiguousarray(mats if mats is not None else default_materials(n), dtype=np.float32)
    counts = np.empty(n, dtype=np.uint32)
    lists = np.zeros((n, cap), dtype=np.uint32)
    load().orc_contacts(ctypes.byref(p), n, _p(b), _p(m), _p(counts), _p(lists), cap)
    return counts, lists


def collide_pass(p: OrcParams, bodies, accel, mats=None, kick=0.0, drift=0.0, write_pressure=False,
                 pressures=None):
    b = np.ascontiguousarray(bodies, dtype=np.float32).copy()
    n = len(b)
    a = np.ascontiguousarray(accel, dtype=np.float32)
    m = np.ascontiguousarray(mats if mats is not None else default_materials(n), dtype=np.float32)
    pr = np.zeros(n, dtype=np.float32) if pressures is None else np.ascontiguousarray(pressures, np.float32).copy()
    counts = np.empty(n, dtype=np.uint32)
    load().orc_collide_pass(ctypes.byref(p), n, _p(b), _p(a), _p(m), _p(pr), float(kick), float(drift),
                            int(bool(write_pressure)), _p(counts))
    return b, pr, counts


def default_materials(n, friction=0.0):
    m = np.zeros((n, 4), dtype=np.float32)
    m[:, 0] = 1.0
    m[:, 1] = friction
    return m


def twiddles(max_axis, f64=False):
    t = np.empty(4 * max(max_axis - 1, 1), dtype=np.float64 if f64 else np.float32)
    (load().orc_twiddles_f64 if f64 else load().orc_twiddles)(max_axis, _p(t))
    return t


def fft2d(values, inverse=False, f64=False):
    """values: complex array (h, w); the reference's radix-2 layout."""
    h, w = values.shape
    dt = np.float64 if f64 else np.float32
    buf = np.empty((h, w, 2), dtype=dt)
    buf[..., 0], buf[..., 1] = values.real, values.imag
    tw = twiddles(max(w, h), f64)
    (load().orc_fft2d_f64 if f64 else load().orc_fft2d)(_p(buf), w, h, _p(tw), int(inverse))
    return buf[..., 0] + 1j * buf[..., 1]


def mass_scale(total_mass) -> int:
    return load().orc_mass_scale(float(total_mass))


def deposit(bodies, mats, mw, mh, pw, ph, scale):
    b = np.ascontiguousarray(bodies, dtype=np.float32)
    m = np.ascontiguousarray(mats, dtype=np.float32)
    dens = np.empty(mw * mh, dtype=np.uint32)
    cnt = np.empty(pw * ph, dtype=np.int32)
    cm = np.empty(pw * ph, dtype=np.uint32)
    load().orc_deposit(len(b), _p(b), _p(m), mw, mh, pw, ph, int(scale), _p(dens), _p(cnt), _p(cm))
    return dens.reshape(mh, mw), cnt.reshape(ph, pw), cm.reshape(ph, pw)


def liquid_constants(diameter, radius_diameters):
    out = [ctypes.c_float() for _ in range(5)]
    load().orc_liquid_constants(float(diameter), float(radius_diameters), *[ctypes.byref(o) for o in out])
    return tuple(o.value for o in out)


class Sim:
    """orc_sim: the whole GravityPipeline step on the CPU."""

    def __init__(self, p: OrcParams, bodies, mats=None):
        self.p = p
        b = np.ascontiguousarray(bodies, dtype=np.float32)
        self.n = len(b)
        m = None if mats is None else np.ascontiguousarray(mats, dtype=np.float32)
        self._h = load().orc_sim_create(ctypes.byref(p), self.n, _p(b), _p(m))
        mw = p.grid_w * (2 if p.bounded else 1)
        mh = p.grid_h * (2 if p.bounded else 1)
        self.mesh = (mh, mw)

    def close(self):
        if self._h:
            load().orc_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_bodies(self, bodies):
        b = np.ascontiguousarray(bodies, dtype=np.float32)
        load().orc_sim_set_bodies(self._h, _p(b))

    def set_accel(self, accel):
        a = np.ascontiguousarray(accel, dtype=np.float32)
        load().orc_sim_set_accel(self._h, _p(a))

    def state(self):
        b = np.empty((self.n, 4), np.float32)
        a = np.empty((self.n, 2), np.float32)
        pr = np.empty(self.n, np.float32)
        m = np.empty((self.n, 4), np.float32)
        load().orc_sim_get(self._h, _p(b), _p(a), _p(pr), _p(m))
        return b, a, pr, m

    def build_force(self):
        load().orc_sim_build_force(self._h)

    def forces(self):
        mh, mw = self.mesh
        fx = np.empty(mh * mw, np.float32)
        fy = np.empty(mh * mw, np.float32)
        load().orc_sim_forces(self._h, _p(fx), _p(fy))
        return fx.reshape(mh, mw), fy.reshape(mh, mw)

    def density_spectrum(self):
        mh, mw = self.mesh
        s = np.empty((mh, mw, 2), np.float32)
        load().orc_sim_density_spectrum(self._h, _p(s))
        return s[..., 0] + 1j * s[..., 1]

    def kick(self, initial=False):
        load().orc_sim_kick(self._h, int(initial))

    def initialize_acceleration(self):
        load().orc_sim_initialize_acceleration(self._h)

    def step(self):
        load().orc_sim_step(self._h)

    def last_sort_dest(self):
        d = np.empty(self.n, np.uint32)
        load().orc_sim_last_sort_dest(self._h, _p(d))
        return d

    def wake(self):
        load().orc_sim_wake(self._h)

    def sleep_stamp(self) -> int:
        return int(load().orc_sim_sleep_stamp(self._h))

    def body_awake(self):
        """Whether each body was awake in the last step."""
        a = np.empty(self.n, np.uint32)
        load().orc_sim_body_awake(self._h, _p(a))
        return a == np.uint32(self.sleep_stamp() - 1)

    def resolve_liquid(self):
        load().orc_sim_resolve_liquid(self._h)

    def trajectory(self, origin, velocity, steps=1024, maximum_dt=1.0, segment_px=6.0,
                   path_diagonals=1.0, minimum_dash_px=3.0):
        """The preview path (S/trajectory.comp): (steps * 6, 4) vertices."""
        t = OrcTrajectoryParams(int(steps), float(maximum_dt), float(segment_px),
                                float(path_diagonals), float(minimum_dash_px))
        out = np.empty((steps * 6, 4), np.float32)
        load().orc_sim_trajectory(self._h, float(origin[0]), float(origin[1]), float(velocity[0]),
                                  float(velocity[1]), ctypes.byref(t), _p(out))
        return out


# ---- glued blocks: the rigid projection ------------------------------------
RIGID_TILE = 256
RIGID_MAXIMUM_GROUPS = 256
RIGID_CONTACT_PARTNERS = RIGID_MAXIMUM_GROUPS + 5
RIGID_CONTACT_KEYS = RIGID_MAXIMUM_GROUPS * RIGID_CONTACT_PARTNERS


def rigid_hull(xs, ys):
    """GlueShape.hullIndices: corner indices, counter-clockwise."""
    xs = np.ascontiguousarray(xs, np.float32)
    ys = np.ascontiguousarray(ys, np.float32)
    out = np.zeros(max(len(xs), 1), np.int32)
    n = load().orc_rigid_hull(_p(xs), _p(ys), len(xs), _p(out))
    return out[:n].copy()


def rigid_prepare(offsets):
    """RigidBlocks.prepare: (rest offsets (n, 2), centroid (2,), inertia, bounds (4,))."""
    offsets = np.ascontiguousarray(offsets, np.float32).reshape(-1, 2)
    rest = np.zeros_like(offsets)
    centroid = np.zeros(2, np.float32)
    bounds = np.zeros(4, np.float32)
    inertia = ctypes.c_double()
    load().orc_rigid_prepare(_p(offsets), len(offsets), _p(rest), _p(centroid), ctypes.byref(inertia),
                             _p(bounds))
    return rest, centroid, inertia.value, bounds


def rigid_project(width, height, bounded, contacts, group_count, tile_count, groups, memory, shapes,
                  members, rest, tile_groups, bodies, substep_dt=1.0 / 120.0 / 16, body_radius=0.56419,
                  block_friction=0.5):
    """RigidBlocks.project for blocks nothing breaks. groups (uint32/int32/float32 words, 40 per group),
    memory (int32) and bodies (float32, 4 per body) are updated in place."""
    p = OrcRigidParams(float(width), float(height), int(bounded), int(contacts), int(group_count),
                       int(tile_count), float(substep_dt), float(body_radius), float(block_friction))
    for a, t in ((groups, None), (memory, np.int32), (shapes, np.float32), (members, np.int32),
                 (rest, np.float32), (tile_groups, np.int32), (bodies, np.float32)):
        assert a.flags.c_contiguous and (t is None or a.dtype == t)
    assert groups.dtype.itemsize == 4 and groups.size >= 40 * group_count
    assert memory.size >= 5 * RIGID_CONTACT_KEYS and members.size >= RIGID_TILE * tile_count
    assert rest.size >= 2 * RIGID_TILE * tile_count and tile_groups.size >= tile_count
    load().orc_rigid_project(ctypes.byref(p), _p(groups), _p(memory), _p(shapes), _p(members), _p(rest),
                             _p(tile_groups), _p(bodies))
