"""GravityConfig mirror (T/GravityConfigTest.kt) and the C ABI's argument
checks, which run before any GPU call."""
import ctypes
import re
from pathlib import Path

import pytest

from gravity_amd import GravityConfig
from gravity_amd.config import CollisionConfig, GravityFieldConfig, ParticleConfig

ROOT = Path(__file__).resolve().parents[1]


def test_parser_uses_constructor_defaults():
    assert GravityConfig.parse([], {}) == GravityConfig()


def test_command_line_overrides_environment():
    cfg = GravityConfig.parse(
        ["--bodies=120000", "--capacity=200000", "--grid=1024x512", "--near-exact-bodies=3",
         "--collisions=6", "--diameter=1.2"],
        {"GRAVITY_BODIES": "1000", "GRAVITY_CAPACITY": "5", "GRAVITY_GRID": "128", "GRAVITY_COLLISIONS": "2"})
    assert cfg.particles.initialBodies == 120_000
    assert cfg.particles.capacity == 200_000
    assert (cfg.gravity.gridWidth, cfg.gravity.gridHeight) == (1024, 512)
    assert cfg.gravity.exactNeighborBodyLimit == 3
    assert cfg.collisions.maximumSubsteps == 6
    assert abs(cfg.collisions.contactDiameterPixels - 1.2) < 1e-9


def test_environment_fills_what_the_command_line_leaves():
    cfg = GravityConfig.parse([], {"GRAVITY_GRID": "256", "GRAVITY_DT": "0.01"})
    assert (cfg.gravity.gridWidth, cfg.gravity.gridHeight) == (256, 256)
    assert cfg.collisions.timeStep == 0.01


def test_point_size_scales_every_size_default():
    base = GravityConfig()
    cfg = GravityConfig.parse(["--point-size=3"], {})
    assert abs(cfg.collisions.contactDiameterPixels - base.collisions.contactDiameterPixels * 3) < 1e-5
    assert abs(cfg.collisions.neighborSkinPixels - base.collisions.neighborSkinPixels * 3) < 1e-5
    assert abs(base.particles.initialBodies / 9 - cfg.particles.initialBodies) <= 1
    explicit = GravityConfig.parse(["--point-size=3", "--diameter=2", "--bodies=1000"], {})
    assert explicit.collisions.contactDiameterPixels == 2
    assert explicit.particles.initialBodies == 1000


def test_body_count_is_clipped_to_capacity():
    assert GravityConfig.parse(["--bodies=1000", "--capacity=500"], {}).particles.initialBodies == 500


@pytest.mark.parametrize("args, message", [
    (["--grid=300"], "grid width must be a power of two"),
    (["--grid=512x100"], "grid height must be a power of two"),
    (["--collisions=17"], "collision substeps must be in 1..16"),
    (["--diameter=0.1"], "contact diameter must be in 0.5"),
    (["--liquid-iterations=0"], "liquid iterations must be in 1..16"),
    (["--typo=1"], "Unknown option: --typo"),
    (["bodies=5"], "Unknown argument: bodies=5"),
    (["--bodies=lots"], "Invalid integer for --bodies: lots"),
    (["--sleep=maybe"], "Invalid boolean for --sleep: maybe"),
])
def test_invalid_configuration_is_rejected(args, message):
    with pytest.raises(ValueError, match=re.escape(message)):
        GravityConfig.parse(args, {})


def test_nested_validation():
    with pytest.raises(ValueError):
        ParticleConfig(capacity=0)
    with pytest.raises(ValueError):
        GravityFieldConfig(minimumAssignmentWindowPower=0.0)
    with pytest.raises(ValueError):
        CollisionConfig(restitutionCoefficient=2.0)


def declared_symbols():
    text = "\n".join(p.read_text() for p in (ROOT / "include").glob("*.h"))
    return sorted(set(re.findall(r"\b(gs_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from gravity_amd import native
    lib = native.load()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(native.SIGNATURES), "the ctypes table covers the header"


def test_config_struct_round_trips_the_c_defaults():
    from gravity_amd import native
    c = native.defaults()
    assert (c.viewport_width, c.viewport_height, c.capacity) == (3440, 1440, 400000)
    assert (c.grid_width, c.grid_height, c.maximum_substeps) == (512, 256, 16)
    assert abs(c.contact_diameter_pixels - 1.12838) < 1e-6
    assert abs(c.gravitational_constant - 0.035) < 1e-9
    assert c.collisions_enabled == 1 and c.spatial_sort_interval == 2
    py = GravityConfig().native(3440, 1440, False)
    for name, _ in native.GsConfig._fields_:
        if name in ("mutual_gravity", "sleep", "liquid", "bounded", "field_acceleration_x", "field_acceleration_y"):
            continue
        assert abs(getattr(py, name) - getattr(c, name)) < 1e-6, name


def test_create_rejects_bad_arguments_before_touching_the_gpu():
    from gravity_amd import native
    lib = native.load()
    c = native.defaults()
    c.grid_width = 300
    h = ctypes.c_void_p()
    assert lib.gs_create(ctypes.byref(c), None, 0, ctypes.byref(h)) == native.GS_EINVAL
    assert b"grid width must be a power of two" in lib.gs_last_error()
    c = native.defaults()
    c.bounded = 1
    c.grid_width = 2048
    assert lib.gs_create(ctypes.byref(c), None, 0, ctypes.byref(h)) == native.GS_EINVAL
    assert b"Walls need an FFT grid" in lib.gs_last_error()
    c = native.defaults()
    assert lib.gs_create(ctypes.byref(c), None, -1, ctypes.byref(h)) == native.GS_EINVAL
    assert lib.gs_step(None, 1) == native.GS_EINVAL


def test_batch_rejects_mixed_time_steps_before_touching_the_gpu():
    from gravity_amd import CollisionConfig, GravityBatch, GravityConfig, ParticleConfig
    a = GravityConfig(particles=ParticleConfig(capacity=100, initialBodies=1))
    b = GravityConfig(particles=ParticleConfig(capacity=100, initialBodies=1),
                      collisions=CollisionConfig(timeStep=1.0 / 60.0))
    with pytest.raises(ValueError, match="shares dt"):
        GravityBatch([a, b], [(256, 256), (256, 256)])


def test_batch_rejects_mixed_collision_schedules_before_touching_the_gpu():
    # Every member steps through the same schedule: contact passes for all
    # or plain drifts for all.
    from gravity_amd import GravityBatch, GravityConfig, ParticleConfig
    a = GravityConfig(particles=ParticleConfig(capacity=100, initialBodies=1))
    b = GravityConfig(particles=ParticleConfig(capacity=100, initialBodies=1), collisionsEnabled=False)
    with pytest.raises(ValueError, match="shares collisionsEnabled"):
        GravityBatch([a, b], [(256, 256), (256, 256)])


def test_adaptive_collision_options():
    # K/GravityConfig.kt:72-83, 336-372, 686-705, 725-729
    cfg = GravityConfig.parse(["--adaptive-collisions=true", "--adaptive-displacement-fraction=0.5",
                               "--adaptive-penetration-fraction=0.25", "--adaptive-dense-contact-count=6",
                               "--adaptive-dense-body-fraction=0.02", "--adaptive-dense-min-bodies=64",
                               "--gpu-budget-ms=12"], {})
    k = cfg.collisions
    assert k.adaptive and k.allowedDisplacementFraction == 0.5 and k.allowedPenetrationFraction == 0.25
    assert (k.denseContactThreshold, k.denseBodyFraction, k.minimumDenseBodies) == (6, 0.02, 64)
    assert k.gpuBudgetMillis == 12.0
    c = cfg.native(640, 320, False)
    assert c.adaptive_collisions == 1 and c.dense_contact_threshold == 6 and c.minimum_dense_bodies == 64
    assert abs(c.allowed_displacement_fraction - 0.5) < 1e-7 and abs(c.dense_body_fraction - 0.02) < 1e-7
    d = CollisionConfig()
    assert not d.adaptive and (d.allowedDisplacementFraction, d.allowedPenetrationFraction) == (0.25, 0.125)
    assert (d.denseContactThreshold, d.denseBodyFraction, d.minimumDenseBodies) == (4, 0.01, 256)
    for bad in (["--adaptive-displacement-fraction=1.5"], ["--adaptive-penetration-fraction=-0.1"],
                ["--adaptive-dense-contact-count=0"], ["--adaptive-dense-body-fraction=2"],
                ["--adaptive-dense-min-bodies=0"], ["--gpu-budget-ms=-1"]):
        with pytest.raises(ValueError):
            GravityConfig.parse(bad, {})


def test_library_validates_the_adaptive_fields_before_touching_the_gpu():
    from gravity_amd import native
    lib = native.load()
    h = ctypes.c_void_p()
    for field, value, message in [("allowed_displacement_fraction", 1.5, b"adaptive-displacement-fraction"),
                                  ("allowed_penetration_fraction", -1.0, b"adaptive-penetration-fraction"),
                                  ("dense_contact_threshold", 0, b"adaptive-dense-contact-count"),
                                  ("dense_body_fraction", 2.0, b"adaptive-dense-body-fraction"),
                                  ("minimum_dense_bodies", 0, b"adaptive-dense-min-bodies")]:
        c = native.defaults()
        setattr(c, field, value)
        assert lib.gs_create(ctypes.byref(c), None, 0, ctypes.byref(h)) == native.GS_EINVAL, field
        assert message in lib.gs_last_error(), field


def test_batch_rejects_mixed_adaptive_collisions_before_touching_the_gpu():
    from gravity_amd import GravityBatch, GravityConfig, ParticleConfig
    a = GravityConfig(particles=ParticleConfig(capacity=100, initialBodies=1))
    b = GravityConfig(particles=ParticleConfig(capacity=100, initialBodies=1),
                      collisions=CollisionConfig(adaptive=True))
    with pytest.raises(ValueError, match="shares adaptive collisions"):
        GravityBatch([a, b], [(256, 256), (256, 256)])
