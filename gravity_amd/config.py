"""GravityConfig mirror (K/GravityConfig.kt): options, defaults, validation.

Only the groups the particle step reads are modelled (particles, gravity,
collisions, liquid, start); ``parse`` accepts the reference's ``--key=value``
options with the same names, environment fallbacks and error messages, and
ignores (after validating the key) options of the interactive shell.
"""
from __future__ import annotations

import dataclasses
import math
import os
from dataclasses import dataclass, field

from . import native

REFERENCE_DIAMETER_PIXELS = 1.12838  # 2/sqrt(pi), K/GravityConfig.kt:735
MAXIMUM_DIAMETER_PIXELS = 16.0

# K/GravityConfig.kt:52-137
SUPPORTED_KEYS = frozenset("""
bodies capacity grid gravity softening near-radius-sigma near-radius-min near-radius-max
near-exact-bodies near-home-bodies near-min-distance-fraction cic-window-floor field-gravity dt
collisions diameter collision-neighbor-radius restitution friction coincident-distance-epsilon
adaptive-collisions collision-cache collision-skin sleep sleep-speed gpu-budget-ms sort-interval
adaptive-displacement-fraction adaptive-penetration-fraction adaptive-dense-contact-count
adaptive-dense-body-fraction adaptive-dense-min-bodies pressure-max target-gpu-ms time-scale-min
time-scale-scroll fps-update-ms console-report-ms gpu-smoothing point-size overlay-characters
overlay-margin-px overlay-line-height-px trajectory-steps trajectory-max-dt trajectory-segment-px
trajectory-path-diagonals trajectory-updates-per-second trajectory-min-dash-px projectile-diameter
projectile-drag-scale projectile-min-drag-px projectile-size-step projectile-max-diameter
projectile-plus-bodies block-density block-friction block-restitution block-strength scenario walls
mutual-gravity field liquid liquid-kernel liquid-iterations liquid-viscosity preview-period-px frames
square-test disc-test seed profile benchmark warmup-frames benchmark-frames benchmark-projectile-x
benchmark-projectile-y benchmark-projectile-diameter benchmark-projectile-vx benchmark-projectile-vy
fft-validate profiler-ring profiler-segments profile-history timer-ring
""".split())


def _require(ok: bool, message: str) -> None:
    if not ok:
        raise ValueError(message)


def _pow2(v: int) -> bool:
    return v > 0 and v & (v - 1) == 0


@dataclass(frozen=True)
class ParticleConfig:
    capacity: int = 400_000
    initialBodies: int = 50_000
    seed: int | None = None
    spatialSortInterval: int = 2

    def __post_init__(self):
        _require(self.capacity > 0, "capacity must be positive")
        _require(self.spatialSortInterval >= 0, "sort interval must be non-negative")
        _require(1 <= self.initialBodies <= self.capacity, "initial body count must be in 1..capacity")


@dataclass(frozen=True)
class GravityFieldConfig:
    gridWidth: int = 512
    gridHeight: int = 256
    gravitationalConstant: float = 0.035
    softeningCells: float = 0.85
    localSearchSigmaMultiplier: float = 3.0
    minimumNeighborRadiusCells: int = 1
    maximumNeighborRadiusCells: int = 6
    exactNeighborBodyLimit: int = 2
    maximumHomeBodyInteractions: int = 8
    minimumDistanceCellFraction: float = 0.05
    minimumAssignmentWindowPower: float = 0.08
    uniformFieldAcceleration: float = 300.0
    # Grids below the reference's 128 (K/GravityConfig.kt:657-662) serve the
    # batched ensembles; the kernels accept them.
    minimumGrid: int = dataclasses.field(default=128, compare=False, repr=False)

    def __post_init__(self):
        lo = self.minimumGrid
        _require(lo <= self.gridWidth <= 2048 and _pow2(self.gridWidth),
                 f"grid width must be a power of two in {lo}..2048")
        _require(lo <= self.gridHeight <= 2048 and _pow2(self.gridHeight),
                 f"grid height must be a power of two in {lo}..2048")
        _require(self.gravitationalConstant >= 0 and math.isfinite(self.gravitationalConstant), "gravity")
        _require(self.softeningCells > 0 and math.isfinite(self.softeningCells), "softening")
        _require(self.minimumNeighborRadiusCells >= 1, "near-radius-min")
        _require(self.maximumNeighborRadiusCells >= self.minimumNeighborRadiusCells, "near-radius-max")
        _require(self.exactNeighborBodyLimit >= 1, "near-exact-bodies")
        _require(self.maximumHomeBodyInteractions >= 1, "near-home-bodies")
        _require(0 < self.minimumAssignmentWindowPower <= 1, "cic-window-floor")


@dataclass(frozen=True)
class CollisionConfig:
    timeStep: float = 1.0 / 120.0
    maximumSubsteps: int = 16
    contactDiameterPixels: float = REFERENCE_DIAMETER_PIXELS
    neighborCellRadius: int = 1
    restitutionCoefficient: float = 0.0
    friction: float = 0.0
    coincidentDistanceSquaredEpsilon: float = 1e-12
    adaptive: bool = False
    neighborCache: bool = True
    neighborSkinPixels: float = 0.5
    sleep: bool = True
    sleepSpeedPixels: float = 6.0
    gpuBudgetMillis: float = 8.5
    allowedDisplacementFraction: float = 0.25
    allowedPenetrationFraction: float = 0.125
    denseContactThreshold: int = 4
    denseBodyFraction: float = 0.01
    minimumDenseBodies: int = 256

    def __post_init__(self):
        _require(self.timeStep > 0 and math.isfinite(self.timeStep), "dt")
        _require(1 <= self.maximumSubsteps <= 16, "collision substeps must be in 1..16")
        _require(0.5 <= self.contactDiameterPixels <= MAXIMUM_DIAMETER_PIXELS,
                 f"contact diameter must be in 0.5..{MAXIMUM_DIAMETER_PIXELS} pixels")
        _require(1 <= self.neighborCellRadius <= 4, "collision-neighbor-radius")
        _require(0 <= self.restitutionCoefficient <= 1, "restitution")
        _require(self.friction >= 0 and math.isfinite(self.friction), "friction")
        _require(self.coincidentDistanceSquaredEpsilon > 0, "coincident-distance-epsilon")
        _require(self.neighborSkinPixels > 0, "collision-skin")
        _require(self.sleepSpeedPixels >= 0, "sleep-speed")
        _require(self.gpuBudgetMillis >= 0 and math.isfinite(self.gpuBudgetMillis), "gpu-budget-ms")
        _require(0 <= self.allowedDisplacementFraction <= 1, "adaptive-displacement-fraction")
        _require(0 <= self.allowedPenetrationFraction <= 1, "adaptive-penetration-fraction")
        _require(self.denseContactThreshold >= 1, "adaptive-dense-contact-count")
        _require(0 <= self.denseBodyFraction <= 1, "adaptive-dense-body-fraction")
        _require(self.minimumDenseBodies >= 1, "adaptive-dense-min-bodies")


@dataclass(frozen=True)
class LiquidConfig:
    kernelRadiusDiameters: float = 2.0
    iterations: int = 12
    viscosity: float = 0.1

    def __post_init__(self):
        _require(1.5 <= self.kernelRadiusDiameters <= 4, "liquid kernel must be 1.5..4 diameters")
        _require(1 <= self.iterations <= 16, "liquid iterations must be in 1..16")
        _require(0 <= self.viscosity <= 1, "liquid viscosity must be in 0..1")


@dataclass(frozen=True)
class TrajectoryConfig:
    """TrajectoryConfig (K/GravityConfig.kt:739-758): the preview path."""
    steps: int = 1024
    maximumTimeStep: float = 1.0
    targetSegmentPixels: float = 6.0
    maximumPathViewportDiagonals: float = 1.0
    updatesPerSecond: int = 30
    minimumDashPixels: float = 3.0

    def __post_init__(self):
        _require(1 <= self.steps <= 65536, "trajectory-steps")
        _require(self.maximumTimeStep > 0 and math.isfinite(self.maximumTimeStep), "trajectory-max-dt")
        _require(self.targetSegmentPixels > 0 and math.isfinite(self.targetSegmentPixels),
                 "trajectory-segment-px")
        _require(self.maximumPathViewportDiagonals > 0 and math.isfinite(self.maximumPathViewportDiagonals),
                 "trajectory-path-diagonals")
        _require(1 <= self.updatesPerSecond <= 1000, "trajectory-updates-per-second")
        _require(self.minimumDashPixels > 0 and math.isfinite(self.minimumDashPixels), "trajectory-min-dash-px")


@dataclass(frozen=True)
class ProjectileConfig:
    """The block fields of ProjectileConfig (K/GravityConfig.kt:760-810)."""
    blockDensity: float = 0.5
    blockFriction: float = 0.5
    blockRestitution: float = 0.2
    blockStrength: float = 400.0

    def __post_init__(self):
        _require(0.05 <= self.blockDensity <= 10.0, "block density must be in 0.05..10.0")
        _require(self.blockFriction >= 0 and math.isfinite(self.blockFriction), "block friction must be non-negative")
        _require(0.0 <= self.blockRestitution <= 1.0, "block restitution must be in 0..1")
        _require(self.blockStrength > 0 and not math.isnan(self.blockStrength), "block strength must be positive")


@dataclass(frozen=True)
class StartConfig:
    """StartConfig.UNIFORM, the start benchmarks were measured with."""
    walls: bool = False
    mutualGravity: bool = True
    liquid: bool = False


@dataclass(frozen=True)
class GravityConfig:
    particles: ParticleConfig = field(default_factory=ParticleConfig)
    gravity: GravityFieldConfig = field(default_factory=GravityFieldConfig)
    collisions: CollisionConfig = field(default_factory=CollisionConfig)
    liquid: LiquidConfig = field(default_factory=LiquidConfig)
    start: StartConfig = field(default_factory=StartConfig)
    trajectory: TrajectoryConfig = field(default_factory=TrajectoryConfig)
    projectiles: ProjectileConfig = field(default_factory=ProjectileConfig)
    # Extension: False runs plain drifts instead of contact passes.
    collisionsEnabled: bool = True

    def __post_init__(self):
        _require(self.particles.initialBodies <= self.particles.capacity,
                 "initial body count must not exceed capacity")

    @staticmethod
    def parse(args: list[str], env: dict[str, str] | None = None) -> "GravityConfig":
        """GravityConfig.parse (K/GravityConfig.kt:141-581)."""
        env = dict(os.environ) if env is None else env
        values: dict[str, str] = {}
        for argument in args:
            _require(argument.startswith("--") and "=" in argument, f"Unknown argument: {argument}")
            key, value = argument[2:].split("=", 1)
            _require(key in SUPPORTED_KEYS, f"Unknown option: --{key}")
            values[key] = value

        def string(key, envkey=None):
            return values.get(key, env.get(envkey) if envkey else None)

        def integer(key, default, envkey=None):
            v = string(key, envkey)
            if v is None:
                return default
            try:
                return int(v)
            except ValueError:
                raise ValueError(f"Invalid integer for --{key}: {v}") from None

        def number(key, default, envkey=None):
            v = string(key, envkey)
            if v is None:
                return default
            try:
                return float(v)
            except ValueError:
                raise ValueError(f"Invalid float for --{key}: {v}") from None

        def boolean(key, default):
            v = values.get(key)
            if v is None:
                return default
            if v.lower() in ("true", "1", "yes", "on"):
                return True
            if v.lower() in ("false", "0", "no", "off"):
                return False
            raise ValueError(f"Invalid boolean for --{key}: {v}")

        d = GravityConfig()
        point = number("point-size", 1.0)
        _require(point > 0 and math.isfinite(point), f"Invalid --point-size: {point}")
        scale = point

        def scaled_count(count):
            return max(1, round(count / (scale * scale)))

        grid = string("grid", "GRAVITY_GRID")
        gw, gh = d.gravity.gridWidth, d.gravity.gridHeight
        if grid is not None:
            if "x" in grid.lower():
                a, b = grid.lower().split("x", 1)
                try:
                    gw = int(a)
                except ValueError:
                    raise ValueError(f"Invalid grid width: {a}") from None
                try:
                    gh = int(b)
                except ValueError:
                    raise ValueError(f"Invalid grid height: {b}") from None
            else:
                try:
                    gw = gh = int(grid)
                except ValueError:
                    raise ValueError(f"Invalid grid: {grid}") from None
        capacity = integer("capacity", d.particles.capacity, "GRAVITY_CAPACITY")
        seed = values.get("seed")
        particles = ParticleConfig(
            capacity=capacity,
            initialBodies=min(integer("bodies", scaled_count(d.particles.initialBodies), "GRAVITY_BODIES"), capacity),
            seed=int(seed) if seed is not None else None,
            spatialSortInterval=integer("sort-interval", d.particles.spatialSortInterval),
        )
        gravity = GravityFieldConfig(
            gridWidth=gw, gridHeight=gh,
            gravitationalConstant=number("gravity", d.gravity.gravitationalConstant, "GRAVITY_G"),
            softeningCells=number("softening", d.gravity.softeningCells, "GRAVITY_SOFTENING"),
            localSearchSigmaMultiplier=number("near-radius-sigma", d.gravity.localSearchSigmaMultiplier),
            minimumNeighborRadiusCells=integer("near-radius-min", d.gravity.minimumNeighborRadiusCells),
            maximumNeighborRadiusCells=integer("near-radius-max", d.gravity.maximumNeighborRadiusCells),
            exactNeighborBodyLimit=integer("near-exact-bodies", d.gravity.exactNeighborBodyLimit),
            maximumHomeBodyInteractions=integer("near-home-bodies", d.gravity.maximumHomeBodyInteractions),
            minimumDistanceCellFraction=number("near-min-distance-fraction", d.gravity.minimumDistanceCellFraction),
            minimumAssignmentWindowPower=number("cic-window-floor", d.gravity.minimumAssignmentWindowPower),
            uniformFieldAcceleration=number("field-gravity", d.gravity.uniformFieldAcceleration),
        )
        collisions = CollisionConfig(
            timeStep=number("dt", d.collisions.timeStep, "GRAVITY_DT"),
            maximumSubsteps=integer("collisions", d.collisions.maximumSubsteps, "GRAVITY_COLLISIONS"),
            contactDiameterPixels=number("diameter", d.collisions.contactDiameterPixels * scale, "GRAVITY_DIAMETER"),
            neighborCellRadius=integer("collision-neighbor-radius", d.collisions.neighborCellRadius),
            restitutionCoefficient=number("restitution", d.collisions.restitutionCoefficient),
            friction=number("friction", d.collisions.friction),
            coincidentDistanceSquaredEpsilon=number("coincident-distance-epsilon",
                                                    d.collisions.coincidentDistanceSquaredEpsilon),
            adaptive=boolean("adaptive-collisions", d.collisions.adaptive),
            neighborCache=boolean("collision-cache", d.collisions.neighborCache),
            neighborSkinPixels=number("collision-skin", d.collisions.neighborSkinPixels * scale),
            sleep=boolean("sleep", d.collisions.sleep),
            sleepSpeedPixels=number("sleep-speed", d.collisions.sleepSpeedPixels * scale),
            gpuBudgetMillis=number("gpu-budget-ms", d.collisions.gpuBudgetMillis),
            allowedDisplacementFraction=number("adaptive-displacement-fraction",
                                               d.collisions.allowedDisplacementFraction),
            allowedPenetrationFraction=number("adaptive-penetration-fraction",
                                              d.collisions.allowedPenetrationFraction),
            denseContactThreshold=integer("adaptive-dense-contact-count", d.collisions.denseContactThreshold),
            denseBodyFraction=number("adaptive-dense-body-fraction", d.collisions.denseBodyFraction),
            minimumDenseBodies=integer("adaptive-dense-min-bodies", d.collisions.minimumDenseBodies),
        )
        liquid = LiquidConfig(
            kernelRadiusDiameters=number("liquid-kernel", d.liquid.kernelRadiusDiameters),
            iterations=integer("liquid-iterations", d.liquid.iterations),
            viscosity=number("liquid-viscosity", d.liquid.viscosity),
        )
        start = StartConfig(
            walls=boolean("walls", False),
            mutualGravity=boolean("mutual-gravity", True),
            liquid=boolean("liquid", False),
        )
        trajectory = TrajectoryConfig(
            steps=integer("trajectory-steps", d.trajectory.steps),
            maximumTimeStep=number("trajectory-max-dt", d.trajectory.maximumTimeStep),
            targetSegmentPixels=number("trajectory-segment-px", d.trajectory.targetSegmentPixels),
            maximumPathViewportDiagonals=number("trajectory-path-diagonals",
                                                d.trajectory.maximumPathViewportDiagonals),
            updatesPerSecond=integer("trajectory-updates-per-second", d.trajectory.updatesPerSecond),
            minimumDashPixels=number("trajectory-min-dash-px", d.trajectory.minimumDashPixels),
        )
        projectiles = ProjectileConfig(
            blockDensity=number("block-density", d.projectiles.blockDensity),
            blockFriction=number("block-friction", d.projectiles.blockFriction),
            blockRestitution=number("block-restitution", d.projectiles.blockRestitution),
            blockStrength=number("block-strength", d.projectiles.blockStrength),
        )
        return GravityConfig(particles=particles, gravity=gravity, collisions=collisions, liquid=liquid,
                             start=start, trajectory=trajectory, projectiles=projectiles)

    def native(self, width: int, height: int, bounded: bool) -> native.GsConfig:
        """The flat C struct gs_create takes."""
        c = native.GsConfig()
        c.viewport_width, c.viewport_height = int(width), int(height)
        c.bounded = int(bounded)
        c.mutual_gravity = int(self.start.mutualGravity)
        c.liquid = int(self.start.liquid)
        c.capacity = self.particles.capacity
        c.spatial_sort_interval = self.particles.spatialSortInterval
        g = self.gravity
        c.grid_width, c.grid_height = g.gridWidth, g.gridHeight
        c.gravitational_constant = g.gravitationalConstant
        c.softening_cells = g.softeningCells
        c.local_search_sigma_multiplier = g.localSearchSigmaMultiplier
        c.minimum_neighbor_radius_cells = g.minimumNeighborRadiusCells
        c.maximum_neighbor_radius_cells = g.maximumNeighborRadiusCells
        c.exact_neighbor_body_limit = g.exactNeighborBodyLimit
        c.maximum_home_body_interactions = g.maximumHomeBodyInteractions
        c.minimum_distance_cell_fraction = g.minimumDistanceCellFraction
        c.minimum_assignment_window_power = g.minimumAssignmentWindowPower
        k = self.collisions
        c.time_step = k.timeStep
        c.maximum_substeps = k.maximumSubsteps
        c.contact_diameter_pixels = k.contactDiameterPixels
        c.neighbor_cell_radius = k.neighborCellRadius
        c.restitution_coefficient = k.restitutionCoefficient
        c.friction = k.friction
        c.coincident_distance_squared_epsilon = k.coincidentDistanceSquaredEpsilon
        c.neighbor_skin_pixels = k.neighborSkinPixels
        c.sleep = int(k.sleep)
        c.sleep_speed_pixels = k.sleepSpeedPixels
        lq = self.liquid
        c.kernel_radius_diameters = lq.kernelRadiusDiameters
        c.liquid_iterations = lq.iterations
        c.liquid_viscosity = lq.viscosity
        c.collisions_enabled = int(self.collisionsEnabled)
        c.adaptive_collisions = int(k.adaptive)
        c.allowed_displacement_fraction = k.allowedDisplacementFraction
        c.allowed_penetration_fraction = k.allowedPenetrationFraction
        c.dense_contact_threshold = k.denseContactThreshold
        c.dense_body_fraction = k.denseBodyFraction
        c.minimum_dense_bodies = k.minimumDenseBodies
        t = self.trajectory
        c.trajectory_steps = t.steps
        c.trajectory_maximum_time_step = t.maximumTimeStep
        c.trajectory_target_segment_pixels = t.targetSegmentPixels
        c.trajectory_maximum_path_diagonals = t.maximumPathViewportDiagonals
        c.trajectory_minimum_dash_pixels = t.minimumDashPixels
        c.block_friction = self.projectiles.blockFriction
        return c


def replace(obj, **kw):
    return dataclasses.replace(obj, **kw)
