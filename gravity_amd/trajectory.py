"""TrajectoryUpdateScheduler (K/TrajectoryUpdateScheduler.kt:1-28): while a
drag goes on, the preview (GravityPipeline.updateTrajectory) is recomputed at
most updatesPerSecond times a second; a new drag may update at once. Host
logic only."""
from __future__ import annotations


class TrajectoryUpdateScheduler:
    def __init__(self, updatesPerSecond: int):
        if updatesPerSecond <= 0:
            raise ValueError("Failed requirement.")
        self._interval = 1_000_000_000 // int(updatesPerSecond)
        self._last = 0
        self._updated = False

    def shouldUpdate(self, nowNanos: int) -> bool:
        if self._updated and nowNanos - self._last < self._interval:
            return False
        self._last = int(nowNanos)
        self._updated = True
        return True

    def reset(self) -> None:
        self._last = 0
        self._updated = False
