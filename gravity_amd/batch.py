"""GravityBatch: many independent simulations stepped as one (the batched
ensemble of the C ABI, gs_batch_*). The reference has no counterpart; each
member behaves exactly like a GravityPipeline of the same configuration and
state, which the tests check bit for bit."""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import native, tuning
from .config import GravityConfig


class GravityBatch:
    """Simulation i has configs[i], viewports[i] = (width, height) in px,
    walls if bounded[i], and starts from scenarioStates[i] (an (n, 4) array of
    normalised x, y, vx, vy, or None). All members share dt, the substep
    count, the sort interval and collisionsEnabled."""

    def __init__(self, configs: Sequence[GravityConfig], viewports: Sequence[tuple[int, int]],
                 scenarioStates: Sequence[np.ndarray | None] | None = None,
                 bounded: Sequence[bool] | None = None):
        k = len(configs)
        if k < 1 or len(viewports) != k:
            raise ValueError("one viewport per configuration")
        bounded = list(bounded) if bounded is not None else [c.start.walls for c in configs]
        states = list(scenarioStates) if scenarioStates is not None else [None] * k
        self._lib = native.load()
        cfgs = (native.GsConfig * k)()
        keep = []
        ptrs = (ctypes.c_void_p * k)()
        counts = (ctypes.c_int * k)()
        for i, (cfg, (w, h)) in enumerate(zip(configs, viewports)):
            cfgs[i] = cfg.native(w, h, bounded[i])
            if states[i] is not None:
                s = np.ascontiguousarray(states[i], dtype=np.float32).reshape(-1)
                keep.append(s)
                ptrs[i] = s.ctypes.data
                counts[i] = s.size // 4
        handle = ctypes.c_void_p()
        native.check(self._lib.gs_batch_create(cfgs, k, ptrs, counts, ctypes.byref(handle)))
        self._h = handle
        self.size = k
        self.viewports = [tuple(v) for v in viewports]
        self._cfgs = [cfgs[i] for i in range(k)]
        self._active = [int(counts[i]) for i in range(k)]

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

    def setActive(self, index: int, activeBodies: int) -> None:
        native.check(self._lib.gs_batch_set_active(self._h, int(index), int(activeBodies)))
        self._active[int(index)] = int(activeBodies)

    def measureAdaptiveCollisions(self, index: int):
        """The adaptive collision state of member `index` after the last step
        (GravityPipeline.measureAdaptiveCollisions)."""
        from .pipeline import AdaptiveCollisionStatistics
        out = native.GsAdaptiveStatistics()
        native.check(self._lib.gs_measure_adaptive_collisions(self._h, int(index), ctypes.byref(out)))
        return AdaptiveCollisionStatistics.from_native(out)

    @property
    def tuning(self) -> tuning.Tuning:
        """The launch shapes of the batch's step (gs_tuning)."""
        return tuning.get(self._lib, self._h)

    @tuning.setter
    def tuning(self, value: tuning.Tuning) -> None:
        tuning.set(self._lib, self._h, value)

    def autotune(self, steps: int = 3, cache=None):
        """Chooses the launch shapes for the batch's current workload (see
        gravity_amd.tuning.autotune); the members' states are unchanged."""
        return tuning.autotune(self._lib, self._h, tuning.workload_key(self._cfgs, self._active),
                               steps, cache)

    def upload(self, index: int, bodies: np.ndarray, materials: np.ndarray | None = None,
               first: int = 0) -> None:
        b = np.ascontiguousarray(bodies, dtype=np.float32).reshape(-1, 4)
        m = None if materials is None else np.ascontiguousarray(materials, dtype=np.float32).reshape(-1, 4)
        native.check(self._lib.gs_batch_upload(self._h, int(index), int(first), len(b),
                                               native.ptr(b), native.ptr(m)))

    def _download(self, index: int, what: int, count: int, per: int, dtype) -> np.ndarray:
        out = np.empty(count * per, dtype=dtype)
        native.check(self._lib.gs_batch_download(self._h, int(index), what, 0, int(count),
                                                 native.ptr(out)))
        return out

    def bodies(self, index: int, count: int) -> np.ndarray:
        return self._download(index, native.GS_BODIES_XYVV, count, 4, np.float32).reshape(count, 4)

    def acceleration(self, index: int, count: int) -> np.ndarray:
        return self._download(index, native.GS_ACCEL, count, 2, np.float32).reshape(count, 2)

    def pressures(self, index: int, count: int) -> np.ndarray:
        return self._download(index, native.GS_PRESSURES, count, 1, np.float32)

    def initializeAcceleration(self) -> None:
        native.check(self._lib.gs_batch_initialize_acceleration(self._h))

    def step(self) -> None:
        native.check(self._lib.gs_batch_step(self._h))

    def synchronize(self) -> None:
        native.check(self._lib.gs_synchronize(self._h))
