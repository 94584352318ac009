"""Glued blocks on the MI355X against the oracle (tests/test_rigid.py pins
the oracle): the block buffers RigidBlocks.add uploads (K/RigidBlocks.kt:
779-857), the rigid projection (S/rigid_reduce.comp, S/rigid_apply.comp)
bit for bit over several projections on the torus and in walls, with guard,
support and contact memory posed as the RIGID contact passes leave them,
the members' remap after a spatial sort (S/rigid_remap.comp), and the
refusal to step a pipeline whose RIGID contact passes are not built."""
import numpy as np
import pytest

from blocks import GROUP_WORDS, KEYS, Layout, hex_lattice, place
from gravity_amd import CollisionConfig, GravityConfig, GravityFieldConfig, GravityPipeline, ParticleConfig
from gravity_amd import native, scenarios

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
DIAMETER = 1.12838
SUB = np.float32(np.float32(1.0 / 120.0) / np.float32(16))


def _scene(bounded, sizes=(100, 700, 5000), loose=1000, seed=3):
    """Loose bodies, then blocks of hex lattices at separate centres, each
    turned, moving and spinning; returns the pipeline, the layout and the
    bodies."""
    rng = np.random.default_rng(seed)
    cap = loose + sum(sizes) + 500
    bodies, mats = scenarios.uniform(loose, W, H, seed=seed)
    cfg = GravityConfig(particles=ParticleConfig(capacity=cap, initialBodies=loose),
                        gravity=GravityFieldConfig(gridWidth=256, gridHeight=128),
                        collisions=CollisionConfig(contactDiameterPixels=DIAMETER))
    pipe = GravityPipeline(cfg, W, H, scenarioState=bodies, bounded=bounded)
    lay = Layout(cap)
    all_bodies = [bodies]
    all_mats = [mats]
    first = loose
    centers = [(300.0, 300.0), (1000.0, 540.0), (1600.0 if bounded else 1915.0, 700.0)]
    for k, n in enumerate(sizes):
        offsets = hex_lattice(n)
        cx, cy = centers[k]
        block = place(offsets, (cx, cy), 0.0, (rng.uniform(-40, 40), rng.uniform(-40, 40)), 0.0, W, H, bounded)
        m = np.zeros((n, 4), np.float32)
        m[:, 0], m[:, 2], m[:, 3] = 0.5, 1.0, float(k)
        all_bodies.append(block)
        all_mats.append(m)
        pipe.upload(block, m, first=first)
        g = pipe.glueBlock(first, offsets, cx, cy)
        assert g == lay.glue(first, offsets, cx, cy) == k
        first += n
    state = np.concatenate(all_bodies)
    # the members jostled as a contact pass leaves them: turned, pushed, sped up
    moved = state.copy()
    moved[loose:, :2] += rng.normal(0, 0.05, (len(state) - loose, 2)).astype(np.float32) / np.float32([W, H])
    moved[loose:, 2:] += rng.normal(0, 3.0, (len(state) - loose, 2)).astype(np.float32) / np.float32([W, H])
    if not bounded:
        moved[:, :2] -= np.floor(moved[:, :2])
    pipe.upload(moved, np.concatenate(all_mats))
    return pipe, lay, moved.copy()


def _read_all(pipe, lay):
    groups = pipe.blockRead(native.GS_RIGID_GROUPS, 0, lay.groups.size, np.uint32)
    memory = pipe.blockRead(native.GS_RIGID_CONTACT_MEMORY, 0, 5 * KEYS, np.int32)
    return groups, memory


@pytest.mark.parametrize("bounded", [False, True])
def test_glued_blocks_hold_the_uploaded_layout(bounded):
    pipe, lay, _ = _scene(bounded)
    with pipe:
        groups, tiles, slots = pipe.blockSize()
        assert (groups, tiles, slots) == (lay.group_count, lay.tile_count, lay.member_cap)
        got_groups, memory = _read_all(pipe, lay)
        members = pipe.blockRead(native.GS_RIGID_MEMBERS, 0, lay.member_cap, np.int32)
        rest = pipe.blockRead(native.GS_RIGID_REST, 0, 2 * lay.member_cap, np.float32)
        tile_groups = pipe.blockRead(native.GS_RIGID_TILE_GROUPS, 0, lay.tile_count, np.int32)
        shapes = pipe.blockRead(native.GS_RIGID_SHAPES, 0, lay.value_top, np.float32)
    assert np.array_equal(got_groups, lay.groups)
    assert np.array_equal(members, lay.members)
    assert np.array_equal(rest.view(np.uint32), lay.rest.view(np.uint32))
    assert np.array_equal(tile_groups, lay.tile_groups[:lay.tile_count])
    assert np.array_equal(shapes.view(np.uint32), lay.shapes[:lay.value_top].view(np.uint32))
    assert not memory.any()


def _pose_contact_state(pipe, lay, rng):
    """Guard and support requests and contact memory as the RIGID passes
    write them, set alike on the device and in the oracle's layout."""
    for g in range(lay.group_count):
        words = lay.ints(g)
        words[28:36] = rng.integers(0, 3 * 65536, 8)
        pipe.blockWrite(native.GS_RIGID_GROUPS, GROUP_WORDS * g + 28, words[28:36].copy())
    mem = lay.memory
    keys = rng.choice(KEYS, 4000, replace=False)
    mem[keys] = rng.integers(0, 1 << 26, len(keys))                     # impacts
    mem[KEYS + keys] = rng.integers(0, 1 << 20, len(keys))              # next impacts
    mem[2 * KEYS + keys] = rng.integers(0, 3, len(keys))                # states
    mem[3 * KEYS + keys] = rng.integers(0, 4, len(keys))                # idle passes
    mem[4 * KEYS + keys] = rng.integers(0, 9, len(keys))                # pair contacts
    pipe.blockWrite(native.GS_RIGID_CONTACT_MEMORY, 0, mem)


@pytest.mark.parametrize("bounded", [False, True])
def test_projection_matches_the_oracle_bit_for_bit(bounded):
    rng = np.random.default_rng(7 + int(bounded))
    pipe, lay, state = _scene(bounded)
    n = len(state)
    radius = np.float32(0.5) * np.float32(DIAMETER)
    with pipe:
        for round_, contacts in enumerate((True, False, True, True)):
            print(f"projection {round_} contacts={contacts} bounded={bounded}", flush=True)
            if round_ == 2:
                _pose_contact_state(pipe, lay, rng)
            pipe.projectBlocks(contacts)
            lay.project(state, W, H, bounded, contacts, SUB, radius, block_friction=0.5)
            got = pipe.bodies(n)
            groups, memory = _read_all(pipe, lay)
            assert np.array_equal(got.view(np.uint32), state.view(np.uint32)), round_
            assert np.array_equal(groups, lay.groups), round_
            assert np.array_equal(memory, lay.memory), round_
            # the members move again before the next projection
            state[1000:, :2] += rng.normal(0, 0.02, (n - 1000, 2)).astype(np.float32) / np.float32([W, H])
            if not bounded:
                state[:, :2] -= np.floor(state[:, :2])
            pipe.upload(state, None)
    # the loose bodies were never touched
    assert np.array_equal(got[:1000].view(np.uint32), state[:1000].view(np.uint32))


def test_members_follow_their_bodies_through_a_sort():
    pipe, lay, state = _scene(False)
    n = len(state)
    with pipe:
        before = pipe.blockRead(native.GS_RIGID_MEMBERS, 0, lay.member_cap, np.int32)
        pipe.sortBodies(n)
        dest = pipe.sortDestinations(n)
        after = pipe.blockRead(native.GS_RIGID_MEMBERS, 0, lay.member_cap, np.int32)
        sorted_bodies = pipe.bodies(n)
        pipe.projectBlocks(False)
        projected = pipe.bodies(n)
    listed = before >= 0
    assert np.array_equal(after[listed], dest[before[listed]].astype(np.int32))
    assert np.array_equal(after[~listed], before[~listed])
    assert np.array_equal(sorted_bodies[after[listed]].view(np.uint32), state[before[listed]].view(np.uint32))
    # the projection after the sort moves the same bodies as before it
    lay.project(state, W, H, False, False, SUB, np.float32(0.5) * np.float32(DIAMETER))
    assert np.array_equal(projected[after[listed]].view(np.uint32), state[before[listed]].view(np.uint32))


def test_members_beyond_the_sorted_bodies_keep_their_indices():
    pipe, lay, state = _scene(False)
    with pipe:
        before = pipe.blockRead(native.GS_RIGID_MEMBERS, 0, lay.member_cap, np.int32)
        pipe.sortBodies(1000)          # the loose bodies only; every member lies beyond them
        after = pipe.blockRead(native.GS_RIGID_MEMBERS, 0, lay.member_cap, np.int32)
        bodies = pipe.bodies(len(state))
    assert np.array_equal(after, before)
    assert np.array_equal(bodies[1000:].view(np.uint32), state[1000:].view(np.uint32))


def test_a_pipeline_with_blocks_refuses_to_step():
    pipe, lay, state = _scene(True, sizes=(300,), loose=500)
    with pipe:
        with pytest.raises(native.GravityError, match="RIGID"):
            pipe.step(len(state))
        with pytest.raises(ValueError):
            pipe.glueBlock(len(state) + 10_000, hex_lattice(10), 100.0, 100.0)
        with pytest.raises(ValueError):
            pipe.blockWrite(native.GS_RIGID_MEMBERS, 0, np.zeros(4, np.int32))
        with pytest.raises(ValueError):
            pipe.blockWrite(native.GS_RIGID_GROUPS, 10, np.zeros(4, np.int32))


@pytest.mark.parametrize("bounded", [False, True])
def test_spawned_block_is_the_restated_upload_and_glue(bounded):
    from gravity_amd.pipeline import block_offsets
    f32 = np.float32
    loose = 500
    bodies, mats = scenarios.uniform(loose, W, H, seed=9)
    cfg = GravityConfig(particles=ParticleConfig(capacity=6000, initialBodies=loose),
                        gravity=GravityFieldConfig(gridWidth=256, gridHeight=128),
                        collisions=CollisionConfig(contactDiameterPixels=DIAMETER))
    minX, minY, maxX, maxY = 1880.25, 100.5, 1950.0, 160.0   # across the seam on the torus, clamped in walls
    density, vx, vy, spin = 0.75, 30.0, -12.0, 0.4
    with GravityPipeline(cfg, W, H, scenarioState=bodies, bounded=bounded) as pipe:
        pipe.upload(bodies, mats)
        k = pipe.spawnBlock(loose, minX, minY, maxX, maxY, density, vx, vy, spin)
        got = pipe.bodies(loose + k)[loose:]
        got_mats = pipe.materials(loose + k)[loose:]
        groups = pipe.blockRead(native.GS_RIGID_GROUPS, 0, 40, np.uint32)
        assert pipe.spawnBlock(6000 - 10, 0.0, 0.0, 50.0, 50.0) == 0      # does not fit
        assert pipe.spawnBlock(loose + k, 10.0, 10.0, 10.4, 10.4) == 0    # narrower than a body
    lattice = block_offsets(DIAMETER, f32(maxX) - f32(minX), f32(maxY) - f32(minY))
    assert k == len(lattice) > 0
    mean = (lattice.astype(np.float64).sum(axis=0).astype(f32) / f32(k)).astype(f32)
    offsets = (lattice - mean).astype(f32)
    cx = f32(f32(0.5) * f32(f32(minX) + f32(maxX))) + mean[0]
    cy = f32(f32(0.5) * f32(f32(minY) + f32(maxY))) + mean[1]
    px = (cx + offsets[:, 0]).astype(f32)
    py = (cy + offsets[:, 1]).astype(f32)
    if bounded:
        px = np.minimum(np.maximum(px, f32(0)), f32(W) * f32(0.99999994))
        py = np.minimum(np.maximum(py, f32(0)), f32(H) * f32(0.99999994))
    else:
        px = (px - np.floor(px / f32(W)) * f32(W)).astype(f32)
        py = (py - np.floor(py / f32(H)) * f32(H)).astype(f32)
    want = np.stack([px / f32(W), py / f32(H), (f32(vx) - f32(spin) * offsets[:, 1]) / f32(W),
                     (f32(vy) + f32(spin) * offsets[:, 0]) / f32(H)], 1).astype(f32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.all(got_mats == np.array([density, 0.5, 1.0, 0.0], f32))
    lay = Layout(6000)
    lay.glue(loose, offsets, cx, cy)
    assert np.array_equal(groups, lay.groups[:40])
