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

from . import native, tuning
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


def block_offsets(contact_diameter: float, width: float, height: float) -> np.ndarray:
    """RectangularBlock.offsets (K/RectangularBlock.kt:10-77): (n, 2) px, row by row."""
    lib = native.load()
    count = ctypes.c_int(0)
    native.check(lib.gs_block_offsets(float(contact_diameter), float(width), float(height), None, 0,
                                      ctypes.byref(count)))
    out = np.zeros((count.value, 2), np.float32)
    native.check(lib.gs_block_offsets(float(contact_diameter), float(width), float(height), native.ptr(out),
                                      count.value, ctypes.byref(count)))
    return out


def projectile_offsets(contact_diameter: float, projectile_diameter: float) -> np.ndarray:
    """CircularProjectile.offsets (K/CircularProjectile.kt:10-37): the packing's
    centre offsets in px, shape (count, 2), from the library (host only)."""
    lib = native.load()
    count = ctypes.c_int(0)
    native.check(lib.gs_projectile_offsets(float(contact_diameter), float(projectile_diameter), None, 0,
                                           ctypes.byref(count)))
    out = np.empty(2 * count.value, np.float32)
    native.check(lib.gs_projectile_offsets(float(contact_diameter), float(projectile_diameter),
                                           native.ptr(out), count.value, ctypes.byref(count)))
    return out.reshape(-1, 2)


def projectile_diameter_for_body_count(contact_diameter: float, bodies: int) -> float:
    """CircularProjectile.diameterForBodyCount (K/CircularProjectile.kt:39-60)."""
    d = ctypes.c_float(0.0)
    native.check(native.load().gs_projectile_diameter_for_body_count(float(contact_diameter), int(bodies),
                                                                     ctypes.byref(d)))
    return float(d.value)


@dataclass(frozen=True)
class AdaptiveCollisionStatistics:
    """K/GravityPipeline.kt:65-71"""
    maximumSpeedPixelsPerSecond: float
    maximumPenetrationPixels: float
    substeps: int
    maximumContactCount: int
    denseContactBodies: int

    @staticmethod
    def from_native(o) -> "AdaptiveCollisionStatistics":
        return AdaptiveCollisionStatistics(o.maximum_speed_pixels_per_second, o.maximum_penetration_pixels,
                                           o.substeps, o.maximum_contact_count, o.dense_contact_bodies)


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
        self.boundedSpaceSupported = GravityPipeline.supportsWalls(config.gravity)
        self._cfg = cfg

    # -- the trajectory preview (K/GravityPipeline.kt:514-602) ---------------
    @property
    def trajectoryVertexCount(self) -> int:
        return self.config.trajectory.steps * 6

    def updateTrajectory(self, activeBodies: int, originX: float, originY: float, velocityX: float,
                         velocityY: float) -> None:
        native.check(self._lib.gs_update_trajectory(self._h, int(activeBodies), float(originX),
                                                    float(originY), float(velocityX), float(velocityY)))

    def trajectory(self) -> np.ndarray:
        """The preview's vertices, (steps * 6, 4): x, y, dash coordinate px,
        path fraction (-1 for an unused vertex)."""
        n = self.trajectoryVertexCount
        out = np.empty((n, 4), np.float32)
        native.check(self._lib.gs_download_trajectory(self._h, 0, n, native.ptr(out)))
        return out

    def trajectoryView(self) -> tuple[int, int]:
        """bindTrajectoryForRender: the vertex buffer's device address and vertex count."""
        v = ctypes.c_void_p()
        n = ctypes.c_int()
        native.check(self._lib.gs_trajectory_view(self._h, ctypes.byref(v), ctypes.byref(n)))
        return v.value or 0, n.value

    def measureAdaptiveCollisions(self) -> AdaptiveCollisionStatistics:
        """K/GravityPipeline.kt:425-445: blocking end-of-run diagnostic of the
        adaptive collision state after the last step."""
        out = native.GsAdaptiveStatistics()
        native.check(self._lib.gs_measure_adaptive_collisions(self._h, 0, ctypes.byref(out)))
        return AdaptiveCollisionStatistics.from_native(out)

    # -- launch shapes (no reference counterpart) ----------------------------
    @property
    def tuning(self):
        """The launch shapes of the step (gs_tuning); no choice changes a result."""
        return tuning.get(self._lib, self._h)

    @tuning.setter
    def tuning(self, value) -> None:
        tuning.set(self._lib, self._h, value)

    def autotune(self, activeBodies: int, steps: int = 3, cache=None):
        """Chooses the launch shapes for `activeBodies` active bodies (see
        gravity_amd.tuning.autotune); the simulation's state is unchanged."""
        native.check(self._lib.gs_batch_set_active(self._h, 0, int(activeBodies)))
        return tuning.autotune(self._lib, self._h, tuning.workload_key([self._cfg], [activeBodies]),
                               steps, cache)

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
        """K/GravityPipeline.kt:147-151: adaptive collisions choose their own."""
        native.check(self._lib.gs_set_substeps(self._h, int(value)))
        k = self.config.collisions
        self._substeps = k.maximumSubsteps if k.adaptive else min(max(int(value), 1), k.maximumSubsteps)

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
        """K/GravityPipeline.kt:180"""
        count = ctypes.c_int(0)
        native.check(self._lib.gs_projectile_body_count(self._h, float(diameterPixels), ctypes.byref(count)))
        return count.value

    def spawnProjectile(self, firstBody: int, centerX: float, centerY: float, velocityX: float,
                        velocityY: float, diameterPixels: float) -> int:
        """K/GravityPipeline.kt:621-646: packed disc of loose unit-mass bodies;
        returns the bodies written (0 when they do not fit the capacity)."""
        spawned = ctypes.c_int(0)
        native.check(self._lib.gs_spawn_projectile(self._h, int(firstBody), float(centerX), float(centerY),
                                                   float(velocityX), float(velocityY), float(diameterPixels),
                                                   ctypes.byref(spawned)))
        return spawned.value

    # -- glued blocks (K/RigidBlocks.kt) ------------------------------------
    def glueBlock(self, firstBody: int, offsets: np.ndarray, centerX: float, centerY: float) -> int:
        """RigidBlocks.add (K/RigidBlocks.kt:280-301): bodies firstBody.. at
        centre + offsets (px, uploaded already, materials naming the block)
        glued as one block; returns its group index. The RIGID contact passes
        are not built: stepping a pipeline with a block raises."""
        xy = np.ascontiguousarray(offsets, dtype=np.float32).reshape(-1, 2)
        group = ctypes.c_int(-1)
        native.check(self._lib.gs_rigid_glue(self._h, int(firstBody), len(xy), native.ptr(xy), float(centerX),
                                             float(centerY), ctypes.byref(group)))
        return group.value

    def spawnBlock(self, firstBody: int, minX: float, minY: float, maxX: float, maxY: float,
                   density: float | None = None, velocityX: float = 0.0, velocityY: float = 0.0,
                   spin: float = 0.0) -> int:
        """spawnBlock without a mask (K/GravityPipeline.kt:647-717): the
        rectangle filled with a hexagonal packing glued into one block;
        returns the bodies written (0 when they do not fit)."""
        if density is None:
            density = self.config.projectiles.blockDensity
        spawned = ctypes.c_int(0)
        native.check(self._lib.gs_spawn_block(self._h, int(firstBody), float(minX), float(minY), float(maxX),
                                              float(maxY), float(density), float(velocityX), float(velocityY),
                                              float(spin), ctypes.byref(spawned)))
        return spawned.value

    def projectBlocks(self, contacts: bool) -> None:
        """RigidBlocks.project (K/RigidBlocks.kt:309-357) for blocks nothing breaks."""
        native.check(self._lib.gs_rigid_project(self._h, int(bool(contacts))))

    def blockSize(self) -> tuple[int, int, int]:
        """(group indices in use, tiles in use, member slots)."""
        out = (ctypes.c_int * 3)()
        native.check(self._lib.gs_rigid_size(self._h, out))
        return out[0], out[1], out[2]

    def blockRead(self, what: int, first: int, count: int, dtype=np.int32) -> np.ndarray:
        """32-bit words of a block buffer (native.GS_RIGID_*)."""
        out = np.zeros(count, dtype=dtype)
        assert out.itemsize == 4
        native.check(self._lib.gs_rigid_read(self._h, int(what), int(first), int(count), native.ptr(out)))
        return out

    def blockWrite(self, what: int, first: int, words: np.ndarray) -> None:
        """Writes 32-bit words into the groups or the contact memory."""
        w = np.ascontiguousarray(words)
        assert w.itemsize == 4
        native.check(self._lib.gs_rigid_write(self._h, int(what), int(first), w.size, native.ptr(w)))

    # -- render view (bindBodiesForRender, K/GravityPipeline.kt:509-512) -----
    def renderView(self) -> tuple[int, int]:
        """Device addresses of the current bodies (vec4) and pressures (float)
        buffers, valid until the next step, upload or sort."""
        b, p = ctypes.c_void_p(), ctypes.c_void_p()
        native.check(self._lib.gs_render_view(self._h, ctypes.byref(b), ctypes.byref(p)))
        return int(b.value or 0), int(p.value or 0)

    def renderCopy(self, count: int, bodies_device: int | None, pressures_device: int | None) -> None:
        """Queues device-to-device copies of `count` bodies and pressures into
        caller device memory (asynchronous on the pipeline's stream)."""
        native.check(self._lib.gs_render_copy(self._h, int(count), bodies_device, pressures_device))

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
        out = native.GsClusterStatistics()
        native.check(self._lib.gs_measure_cluster(self._h, int(activeBodies), ctypes.byref(out)))
        return ClusterStatistics(
            centerX=out.center_x, centerY=out.center_y, covarianceAreaPixels=out.covariance_area_pixels,
            equivalentDiameterPixels=out.equivalent_diameter_pixels, core90AreaPixels=out.core90_area_pixels,
            core99AreaPixels=out.core99_area_pixels, maxSpeedPixelsPerSecond=out.max_speed_pixels_per_second,
            averageSpeedPixelsPerSecond=out.average_speed_pixels_per_second, nonFiniteBodies=out.non_finite_bodies)

    @staticmethod
    def supportsWalls(gravity) -> bool:
        """K/GravityPipeline.kt:1628"""
        return bool(native.load().gs_supports_walls(int(gravity.gridWidth), int(gravity.gridHeight)))
