"""SubstepGovernor (K/SubstepGovernor.kt:1-48): keeps a frame within its GPU
budget. The collision substeps take most of a step, so while a frame takes
longer than the budget the next steps run proportionally fewer, down to half
of the maximum; below RECOVERY of the budget they come back one at a time.
The measured time is a few frames old, so after each change the governor
waits SETTLE_FRAMES frames for it to show. A budget of 0 keeps every substep.

Host logic only: its output feeds GravityPipeline.collisionSubsteps
(gs_set_substeps), which adaptive collisions ignore (they choose their own).
"""
from __future__ import annotations

import math


class SubstepGovernor:
    RECOVERY = 0.8
    SETTLE_FRAMES = 3

    def __init__(self, maximum: int, budgetMillis: float):
        self.maximum = int(maximum)
        self.budgetMillis = float(budgetMillis)
        self.minimum = (self.maximum + 1) // 2
        self._settling = 0
        self.substeps = self.maximum

    def update(self, gpuMillis: float) -> int:
        """The substeps of the next step, after a frame measured at gpuMillis."""
        if self.budgetMillis <= 0.0 or math.isnan(gpuMillis):
            return self.substeps
        if self._settling > 0:
            self._settling -= 1
            return self.substeps
        if gpuMillis > self.budgetMillis:
            nxt = min(self.substeps - 1, math.floor(self.substeps * self.budgetMillis / gpuMillis))
        elif gpuMillis < self.budgetMillis * self.RECOVERY:
            nxt = self.substeps + 1
        else:
            nxt = self.substeps
        nxt = max(self.minimum, min(self.maximum, nxt))
        if nxt != self.substeps:
            self.substeps = nxt
            self._settling = self.SETTLE_FRAMES
        return self.substeps

    def reset(self) -> None:
        self.substeps = self.maximum
        self._settling = 0
