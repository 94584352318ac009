"""Launch-shape tuning of the step (gs_tuning, gs_autotune). The reference
has no counterpart: its workgroup sizes are fixed in its shaders. No choice
changes a result, only the time a step takes.

``autotune`` times the handle's own workload under each candidate on the
device (the state is copied first and put back, so the simulations go on as
if untuned) and can keep its choice in a JSON cache, keyed by the workload:
the configurations, the viewports, the active counts and the GPU
architecture the library was built for."""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
from dataclasses import asdict, dataclass
from pathlib import Path

from . import native


@dataclass(frozen=True)
class Tuning:
    bigSegmentLimit: int = 0
    forwardItemBodies: int = 4096
    depositRuns: int = 0
    scanDirectBlocks: int = 1024

    def native(self) -> native.GsTuning:
        return native.GsTuning(self.bigSegmentLimit, self.forwardItemBodies, self.depositRuns,
                               self.scanDirectBlocks)

    @staticmethod
    def from_native(t: native.GsTuning) -> "Tuning":
        return Tuning(t.big_segment_limit, t.forward_item_bodies, t.deposit_runs,
                      t.scan_direct_blocks)


def get(lib, handle) -> Tuning:
    out = native.GsTuning()
    native.check(lib.gs_get_tuning(handle, ctypes.byref(out)))
    return Tuning.from_native(out)


def set(lib, handle, tuning: Tuning) -> None:  # noqa: A001 (mirrors gs_set_tuning)
    t = tuning.native()
    native.check(lib.gs_set_tuning(handle, ctypes.byref(t)))


def workload_key(configs: list, actives: list[int]) -> str:
    """The cache key of a workload: a digest of the native configurations
    (which hold the viewports and walls), the active counts and the target
    the library is built for (gfx950)."""
    h = hashlib.sha256()
    h.update(b"gfx950")
    for cfg, n in zip(configs, actives):
        h.update(bytes(cfg))
        h.update(int(n).to_bytes(8, "little", signed=True))
    return h.hexdigest()


def autotune(lib, handle, key: str, steps: int = 3, cache: str | os.PathLike | None = None):
    """Tunes the handle: with a cache that holds the key, applies the stored
    choice without timing anything; otherwise times the candidates and, with
    a cache, stores the choice. Returns (tuning, ms) where ms is None for a
    cached choice, else {"chosen": ms per step, "start": ms per step}."""
    path = Path(cache) if cache is not None else None
    table = {}
    if path is not None and path.exists():
        table = json.loads(path.read_text())
        if key in table:
            t = Tuning(**table[key])
            set(lib, handle, t)
            return t, None
    chosen = native.GsTuning()
    ms = (ctypes.c_double * 2)()
    native.check(lib.gs_autotune(handle, int(steps), ctypes.byref(chosen), ms))
    t = Tuning.from_native(chosen)
    if path is not None:
        table[key] = asdict(t)
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(path.suffix + ".tmp")
        tmp.write_text(json.dumps(table, indent=1, sort_keys=True))
        os.replace(tmp, path)
    return t, {"chosen": ms[0], "start": ms[1]}
