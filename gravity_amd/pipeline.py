"""GravityPipeline: the reference's simulation class (K/GravityPipeline.kt)
over the MI355X C ABI. Member names, argument meaning and errors follow the
Kotlin class so callers and tests read the same; ``ValueError`` stands for
Kotlin's IllegalArgumentException and ``native.GravityError`` for
IllegalStateException.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import native
from .config import GravityConfig

MAXIMUM_UNIT = np.float32(0.99999994)
MAXIMUM_FFT_AXIS = 2048


@dataclass
class ClusterStatistics:
    """K/GravityPipeline.kt:47-57"""
    centerX: float
    centerY: float
    covarianceAreaPixels: float
    equivalentDiameterPixels: float
    core90AreaPixels: float
    core99AreaPixels: float
    maxSpeedPixelsPerSecond: float
    averageSpeedPixelsPerSecond: float
    nonFiniteBodies: int


def projectile_offsets(contact_diameter: float, projectile_diameter: float) -> np.ndarray:
    """CircularProjectile.offsets (K/CircularProjectile.kt:10-37), FP32 as Kotlin."""
    f = np.float32
    d = f(contact_diameter)
    if not (d > 0 and f(projectile_diameter) >= d):
        raise ValueError("projectile must fit one body")
    radius = f(0.5) * d
    available = f(0.5) * f(projectile_diameter) - radius
    row_step = d * f(math.sqrt(3.0) * 0.5)
    out = []
    row = 0
    y = -available
    while y <= available:
        x = -available + (f(0.0) if row % 2 == 0 else f(0.5) * d)
        while x <= available:
            if x * x + y * y <= available * available:
                out.append((x, y))
            x = f(x + d)
        y = f(y + row_step)
        row += 1
    return np.array(out, dtype=np.float32).reshape(-1, 2)


class GravityPipeline:
    """GravityPipeline(config, viewportWidth, viewportHeight, profiler=null,
    scenarioState=null, bounded=false) (K/GravityPipeline.kt:73-81)."""

    def __init__(self, config: GravityConfig, viewportWidth: int, viewportHeight: int,
                 profiler=None, scenarioState: np.ndarray | None = None, bounded: bool = False):
        self.config = config
        self.viewportWidth = int(viewportWidth)
        self.viewportHeight = int(viewportHeight)
        self._lib = native.load()
        cfg = config.native(viewportWidth, viewportHeight, bounded)
        state = None
        n = 0
        if scenarioState is not None:
            state = np.ascontiguousarray(scenarioState, dtype=np.float32).reshape(-1)
            if state.size > config.particles.capacity * 4:
                raise ValueError("scenario exceeds capacity")
            n = state.size // 4
        handle = ctypes.c_void_p()
        native.check(self._lib.gs_create(ctypes.byref(cfg), native.ptr(state), n, ctypes.byref(handle)))
        self._h = handle
        self._bounded = bool(bounded)
        self._mutual = config.start.mutualGravity
        self._liquid = config.start.liquid
        self._field = (0.0, 0.0)
        self._substeps = config.collisions.maximumSubsteps
        self.boundedSpaceSupported = 2 * max(config.gravity.gridWidth, config.gravity.gridHeight) <= MAXIMUM_FFT_AXIS

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.gs_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- properties (K/GravityPipeline.kt:97-151) --------------------------
    @property
    def bounded(self) -> bool:
        return self._bounded

    @bounded.setter
    def bounded(self, value: bool) -> None:
        native.check(self._lib.gs_set_bool(self._h, native.GS_BOUNDED, int(bool(value))))
        self._bounded = bool(value)

    @property
    def mutualGravity(self) -> bool:
        return self._mutual

    @mutualGravity.setter
    def mutualGravity(self, value: bool) -> None:
        native.check(self._lib.gs_set_bool(self._h, native.GS_MUTUAL_GRAVITY, int(bool(value))))
        self._mutual = bool(value)

    @property
    def liquid(self) -> bool:
        return self._liquid

    @liquid.setter
    def liquid(self, value: bool) -> None:
        native.check(self._lib.gs_set_bool(self._h, native.GS_LIQUID, int(bool(value))))
        self._liquid = bool(value)

    @property
    def fieldAccelerationX(self) -> float:
        return self._field[0]

    @fieldAccelerationX.setter
    def fieldAccelerationX(self, value: float) -> None:
        native.check(self._lib.gs_set_field(self._h, float(value), float(self._field[1])))
        self._field = (float(value), self._field[1])

    @property
    def fieldAccelerationY(self) -> float:
        return self._field[1]

    @fieldAccelerationY.setter
    def fieldAccelerationY(self, value: float) -> None:
        native.check(self._lib.gs_set_field(self._h, float(self._field[0]), float(value)))
        self._field = (self._field[0], float(value))

    @property
    def collisionSubsteps(self) -> int:
        return self._substeps

    @collisionSubsteps.setter
    def collisionSubsteps(self, value: int) -> None:
        native.check(self._lib.gs_set_substeps(self._h, int(value)))
        self._substeps = min(max(int(value), 1), self.config.collisions.maximumSubsteps)

    # -- the step (K/GravityPipeline.kt:366-494) ----------------------------
    def initializeAcceleration(self, activeBodies: int) -> None:
        native.check(self._lib.gs_initialize_acceleration(self._h, int(activeBodies)))

    def step(self, activeBodies: int) -> None:
        native.check(self._lib.gs_step(self._h, int(activeBodies)))

    def stopBodies(self, activeBodies: int) -> None:
        native.check(self._lib.gs_stop_bodies(self._h, int(activeBodies)))

    def wake(self) -> None:
        native.check(self._lib.gs_wake(self._h))

    def synchronize(self) -> None:
        native.check(self._lib.gs_synchronize(self._h))

    def projectileBodyCount(self, diameterPixels: float) -> int:
        return len(projectile_offsets(self.config.collisions.contactDiameterPixels, diameterPixels))

    def spawnProjectile(self, firstBody: int, centerX: float, centerY: float, velocityX: float,
                        velocityY: float, diameterPixels: float) -> int:
        """K/GravityPipeline.kt:621-637: packed disc of loose unit-mass bodies."""
        offsets = projectile_offsets(self.config.collisions.contactDiameterPixels, diameterPixels)
        cap = self.config.particles.capacity
        if not (0 <= firstBody <= cap and len(offsets) <= cap - firstBody):
            return 0
        f = np.float32
        W, H = f(self.viewportWidth), f(self.viewportHeight)
        radius = f(np.hypot(offsets[:, 0], offsets[:, 1]).max()) + f(0.5) * f(self.config.collisions.contactDiameterPixels)
        cx, cy = f(centerX), f(centerY)
        if self._bounded:
            cx = min(max(cx, radius), W - radius)
            cy = min(max(cy, radius), H - radius)
        px = cx + offsets[:, 0]
        py = cy + offsets[:, 1]
        if self._bounded:
            px = np.clip(px, f(0), W * MAXIMUM_UNIT)
            py = np.clip(py, f(0), H * MAXIMUM_UNIT)
        else:
            px = px - np.floor(px / W) * W
            py = py - np.floor(py / H) * H
        bodies = np.empty((len(offsets), 4), dtype=np.float32)
        bodies[:, 0] = px / W
        bodies[:, 1] = py / H
        bodies[:, 2] = f(velocityX) / W
        bodies[:, 3] = f(velocityY) / H
        mats = np.zeros((len(offsets), 4), dtype=np.float32)
        mats[:, 0] = 1.0
        mats[:, 1] = self.config.collisions.friction
        self.upload(bodies, mats, first=firstBody)
        return len(offsets)

    # -- test hooks (T/CollisionGpuSupport.kt:43-156) -----------------------
    def upload(self, bodies: np.ndarray, materials: np.ndarray | None = None, first: int = 0) -> None:
        b = np.ascontiguousarray(bodies, dtype=np.float32).reshape(-1, 4)
        m = None if materials is None else np.ascontiguousarray(materials, dtype=np.float32).reshape(-1, 4)
        native.check(self._lib.gs_upload(self._h, int(first), len(b), native.ptr(b), native.ptr(m)))

    def uploadAcceleration(self, accel: np.ndarray, first: int = 0) -> None:
        a = np.ascontiguousarray(accel, dtype=np.float32).reshape(-1, 2)
        native.check(self._lib.gs_upload_accel(self._h, int(first), len(a), native.ptr(a)))

    def _download(self, what: int, count: int, per: int, dtype, first: int = 0) -> np.ndarray:
        out = np.empty(count * per, dtype=dtype)
        native.check(self._lib.gs_download(self._h, what, int(first), int(count), native.ptr(out)))
        return out

    def bodies(self, count: int) -> np.ndarray:
        return self._download(native.GS_BODIES_XYVV, count, 4, np.float32).reshape(count, 4)

    def pressures(self, count: int) -> np.ndarray:
        return self._download(native.GS_PRESSURES, count, 1, np.float32)

    def acceleration(self, count: int) -> np.ndarray:
        return self._download(native.GS_ACCEL, count, 2, np.float32).reshape(count, 2)

    def materials(self, count: int) -> np.ndarray:
        return self._download(native.GS_MATERIALS, count, 4, np.float32).reshape(count, 4)

    def contactCounts(self, count: int) -> np.ndarray:
        return self._download(native.GS_CONTACT_COUNTS, count, 1, np.uint32)

    def contacts(self, count: int) -> np.ndarray:
        return self._download(native.GS_CONTACTS, count, native.CONTACT_SLOTS, np.uint32).reshape(
            count, native.CONTACT_SLOTS)

    def sortDestinations(self, count: int) -> np.ndarray:
        return self._download(native.GS_SORT_DEST, count, 1, np.uint32)

    def query(self) -> dict:
        out = (ctypes.c_int * 6)()
        native.check(self._lib.gs_query(self._h, out))
        return {"capacity": out[0], "meshWidth": out[1], "meshHeight": out[2], "gridWidth": out[3],
                "gridHeight": out[4], "contactSlots": out[5]}

    def forceMesh(self) -> np.ndarray:
        q = self.query()
        mw, mh = q["meshWidth"], q["meshHeight"]
        out = np.empty(mw * mh * 2, dtype=np.float32)
        native.check(self._lib.gs_download(self._h, native.GS_FORCE_MESH, 0, mw * mh, native.ptr(out)))
        return out.reshape(mh, mw, 2)

    def densityMesh(self, activeBodies: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(fixed-point density (mh, mw) uint32, its FP32 spectrum (mh, mw)
        complex, cell table (gh, gw, 2) uint32 of bodies and mass / 256) of
        the current positions."""
        q = self.query()
        mw, mh, pw, ph = q["meshWidth"], q["meshHeight"], q["gridWidth"], q["gridHeight"]
        dens = np.empty(mw * mh, np.uint32)
        spec = np.empty(2 * mw * mh, np.float32)
        cells = np.empty(2 * pw * ph, np.uint32)
        native.check(self._lib.gs_debug_density(self._h, int(activeBodies), native.ptr(dens),
                                                native.ptr(spec), native.ptr(cells)))
        spec = spec.reshape(mh, mw, 2)
        return dens.reshape(mh, mw), spec[..., 0] + 1j * spec[..., 1], cells.reshape(ph, pw, 2)

    def validateFft(self, activeBodies: int) -> dict:
        """K/GravityPipeline.kt:378-419 (FftValidationResult)."""
        out = (ctypes.c_double * 3)()
        native.check(self._lib.gs_validate_fft(self._h, int(activeBodies), out))
        return {"maximumAbsoluteError": out[0], "maximumRelativeError": out[1],
                "rootMeanSquareError": out[2]}

    def resolveCollisions(self, activeBodies: int, kickDt: float, driftDt: float, writePressure: bool,
                          substepIndex: int | None) -> None:
        native.check(self._lib.gs_debug_resolve_collisions(
            self._h, int(activeBodies), float(kickDt), float(driftDt), int(bool(writePressure)),
            -1 if substepIndex is None else int(substepIndex)))

    def sortBodies(self, activeBodies: int) -> None:
        native.check(self._lib.gs_debug_sort(self._h, int(activeBodies)))

    # -- adjoint (no reference counterpart) ---------------------------------
    def adjointRecord(self, activeBodies: int, steps: int, checkpointEvery: int = 8) -> None:
        """`steps` steps like step(), keeping a checkpoint every checkpointEvery."""
        native.check(self._lib.gs_adjoint_record(self._h, int(activeBodies), int(steps),
                                                 int(checkpointEvery)))

    def adjointVjp(self, dLdx: np.ndarray, count: int) -> tuple[np.ndarray, np.ndarray]:
        """(dL/dx0, dL/dv0) of the recorded steps for dL/dx at their end;
        count is the recording's active body count."""
        g = self._floats(dLdx, 2 * count, "dLdx")
        x0 = np.empty(2 * count, np.float32)
        v0 = np.empty(2 * count, np.float32)
        native.check(self._lib.gs_adjoint_vjp(self._h, int(count), native.ptr(g), native.ptr(x0),
                                              native.ptr(v0)))
        return x0.reshape(count, 2), v0.reshape(count, 2)

    def adjointOrder(self, count: int) -> np.ndarray:
        out = np.empty(max(int(count), 0), np.uint32)
        native.check(self._lib.gs_adjoint_order(self._h, int(count), native.ptr(out)))
        return out

    def vjpInitialVelocity(self, activeBodies: int, steps: int, checkpointEvery: int,
                           targets: np.ndarray) -> tuple[float, np.ndarray]:
        """Loss mean |x_T - target|^2 (px^2, minimum image) and its gradient
        with respect to the initial normalised velocities."""
        self._check_active(activeBodies)
        t = self._floats(targets, 2 * activeBodies, "targets")
        grad = np.empty(2 * activeBodies, np.float32)
        loss = ctypes.c_float()
        native.check(self._lib.gs_vjp_initial_velocity(self._h, int(activeBodies), int(steps),
                                                       int(checkpointEvery), native.ptr(t),
                                                       ctypes.byref(loss), native.ptr(grad)))
        return float(loss.value), grad.reshape(activeBodies, 2)

    def accelerationVjp(self, activeBodies: int, w: np.ndarray) -> np.ndarray:
        """J_A(x)^T w at the current positions (test hook of the adjoint)."""
        self._check_active(activeBodies)
        wv = self._floats(w, 2 * activeBodies, "w")
        out = np.empty(2 * activeBodies, np.float32)
        native.check(self._lib.gs_debug_acceleration_vjp(self._h, int(activeBodies), native.ptr(wv),
                                                         native.ptr(out)))
        return out.reshape(activeBodies, 2)

    def _check_active(self, activeBodies: int) -> None:
        if not 0 <= int(activeBodies) <= self.config.particles.capacity:
            raise ValueError("active body count out of range")

    @staticmethod
    def _floats(a: np.ndarray, need: int, name: str) -> np.ndarray:
        v = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
        if v.size < need:
            raise ValueError(f"{name} holds {v.size} floats, {need} needed")
        return v

    def resolveLiquid(self, activeBodies: int) -> None:
        """The liquid solve of a step on its own (the reflection hook
        resolveLiquid, T/CollisionGpuSupport.kt:68-90)."""
        native.check(self._lib.gs_debug_resolve_liquid(self._h, int(activeBodies)))

    def setGraphs(self, enabled: bool) -> None:
        native.check(self._lib.gs_set_graphs(self._h, int(bool(enabled))))

    def profile(self, steps: int) -> list[tuple[str, int, float]]:
        """Per-stage (name, launches, total ms) over `steps` eager steps, each
        launch stamped by its own dispatch."""
        cap = 64
        stats = (native.GsKernelStat * cap)()
        written = ctypes.c_int(0)
        native.check(self._lib.gs_profile_steps(self._h, int(steps), stats, cap, ctypes.byref(written)))
        return [(stats[i].name.decode(), stats[i].launches, stats[i].total_ms) for i in range(written.value)]

    # -- diagnostics (K/GravityPipeline.kt:793-905) -------------------------
    def measureCluster(self, activeBodies: int) -> ClusterStatistics:
        b = self.bodies(activeBodies).astype(np.float64)
        W, H = self.viewportWidth, self.viewportHeight
        finite = np.isfinite(b).all(axis=1)
        nonfinite = int((~finite).sum())
        x = np.where(finite, b[:, 0], 0.0)
        y = np.where(finite, b[:, 1], 0.0)
        speed = np.where(finite, np.hypot(b[:, 2] * W, b[:, 3] * H), 0.0)
        ax, ay = 2 * np.pi * x, 2 * np.pi * y
        cx = math.atan2(np.sin(ax).sum(), np.cos(ax).sum()) / (2 * np.pi)
        cy = math.atan2(np.sin(ay).sum(), np.cos(ay).sum()) / (2 * np.pi)
        cx -= math.floor(cx)
        cy -= math.floor(cy)
        ux = (cx + ((x - cx) - np.floor(x - cx + 0.5))) * W
        uy = (cy + ((y - cy) - np.floor(y - cy + 0.5))) * H
        mx, my = ux.mean(), uy.mean()
        xx = ((ux - mx) ** 2).mean()
        yy = ((uy - my) ** 2).mean()
        xy = ((ux - mx) * (uy - my)).mean()
        area = 4 * math.pi * math.sqrt(max(xx * yy - xy * xy, 0.0))
        medx = np.sort(ux)[activeBodies // 2]
        medy = np.sort(uy)[activeBodies // 2]
        r2 = np.sort((ux - medx) ** 2 + (uy - medy) ** 2)

        def core(fraction):
            idx = min(max(int(fraction * activeBodies), 1), activeBodies) - 1
            return math.pi * r2[idx] / fraction

        return ClusterStatistics(
            centerX=float(medx - math.floor(medx / W) * W), centerY=float(medy - math.floor(medy / H) * H),
            covarianceAreaPixels=area, equivalentDiameterPixels=2 * math.sqrt(area / math.pi),
            core90AreaPixels=core(0.90), core99AreaPixels=core(0.99),
            maxSpeedPixelsPerSecond=float(speed.max()), averageSpeedPixelsPerSecond=float(speed.mean()),
            nonFiniteBodies=nonfinite)

    @staticmethod
    def supportsWalls(gravity) -> bool:
        return 2 * max(gravity.gridWidth, gravity.gridHeight) <= MAXIMUM_FFT_AXIS
