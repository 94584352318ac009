"""Glued blocks for the rigid projection tests: lattices of member offsets,
the block buffers laid out as RigidBlocks.add uploads them
(K/RigidBlocks.kt:280-301, 779-857; the shape region holding the hull only),
and the rigid-motion helpers the pins use."""
from __future__ import annotations

import numpy as np

from oracle import oracle

TILE = oracle.RIGID_TILE
KEYS = oracle.RIGID_CONTACT_KEYS
PARTNERS = oracle.RIGID_CONTACT_PARTNERS
GROUP_WORDS = 40


def hex_lattice(n: int, spacing: float = 1.12838) -> np.ndarray:
    """n offsets (px) of a hexagonal packing, rows of one height, about (0, 0)."""
    cols = max(1, int(np.ceil(np.sqrt(n * 1.2))))
    pts = []
    row = 0
    while len(pts) < n:
        y = row * spacing * np.sqrt(3.0) / 2.0
        shift = 0.5 * spacing if row % 2 else 0.0
        for c in range(cols):
            if len(pts) == n:
                break
            pts.append((c * spacing + shift, y))
        row += 1
    xy = np.asarray(pts, np.float64)
    xy -= xy.mean(axis=0)
    return xy.astype(np.float32)


class Layout:
    """The block buffers of simulation 0, as the library lays them out."""

    def __init__(self, capacity: int):
        self.member_cap = (capacity + oracle.RIGID_MAXIMUM_GROUPS * (TILE - 1) + TILE - 1) // TILE * TILE
        self.tile_cap = self.member_cap // TILE
        self.groups = np.zeros(oracle.RIGID_MAXIMUM_GROUPS * GROUP_WORDS, np.uint32)
        self.memory = np.zeros(5 * KEYS, np.int32)
        self.shapes = np.zeros(max(1 << 18, 2 * capacity), np.float32)
        self.members = np.full(self.member_cap, -1, np.int32)
        self.rest = np.zeros(2 * self.member_cap, np.float32)
        self.tile_groups = np.zeros(self.tile_cap, np.int32)
        self.group_count = 0
        self.tile_count = 0
        self.value_top = 0

    def floats(self, group: int) -> np.ndarray:
        return self.groups[GROUP_WORDS * group: GROUP_WORDS * (group + 1)].view(np.float32)

    def ints(self, group: int) -> np.ndarray:
        return self.groups[GROUP_WORDS * group: GROUP_WORDS * (group + 1)].view(np.int32)

    def glue(self, first_body: int, offsets: np.ndarray, cx: float, cy: float) -> int:
        offsets = np.asarray(offsets, np.float32).reshape(-1, 2)
        n = len(offsets)
        g = self.group_count
        tiles = (n + TILE - 1) // TILE
        rest, centroid, inertia, bounds = oracle.rigid_prepare(offsets)
        corners = oracle.rigid_hull(rest[:, 0], rest[:, 1])
        first = self.tile_count * TILE
        self.members[first:first + n] = first_body + np.arange(n, dtype=np.int32)
        self.rest[2 * first:2 * (first + n)] = rest.reshape(-1)
        self.tile_groups[self.tile_count:self.tile_count + tiles] = g
        hull = rest[corners].reshape(-1)
        self.shapes[self.value_top:self.value_top + hull.size] = hull
        f, i = self.floats(g), self.ints(g)
        f[0:4] = (cx, cy, cx, cy)
        f[4:8] = (0.0, 0.0, 1.0, 0.0)
        f[8:12] = (0.0, np.float32(inertia), float(n), 0.0)
        i[12:16] = (self.tile_count, tiles, self.value_top, len(corners))
        f[16:20] = 0.0
        i[20:24] = (self.value_top, 0, 0, 0)
        f[24:28] = bounds
        i[28:36] = 0
        f[36:40] = (cx, cy, 1.0, 0.0)
        self.group_count = g + 1
        self.tile_count += tiles
        self.value_top += hull.size
        return g

    def project(self, bodies: np.ndarray, width, height, bounded, contacts, substep_dt, body_radius,
                block_friction=0.5) -> None:
        oracle.rigid_project(width, height, bounded, contacts, self.group_count, self.tile_count, self.groups,
                             self.memory, self.shapes, self.members, self.rest, self.tile_groups, bodies,
                             substep_dt=substep_dt, body_radius=body_radius, block_friction=block_friction)


def rotate(theta: float, xy: np.ndarray) -> np.ndarray:
    c, s = np.cos(theta), np.sin(theta)
    x, y = xy[:, 0].astype(np.float64), xy[:, 1].astype(np.float64)
    return np.stack([c * x - s * y, s * x + c * y], axis=1)


def place(offsets_rest, center, theta, velocity, spin, width, height, bounded):
    """Bodies (normalised vec4) of members at centre + R(theta) r0 moving with
    velocity + spin x (R(theta) r0), in FP64 then rounded."""
    arm = rotate(theta, offsets_rest)
    pos = np.asarray(center, np.float64) + arm
    vel = np.asarray(velocity, np.float64) + spin * np.stack([-arm[:, 1], arm[:, 0]], axis=1)
    if not bounded:
        pos = np.mod(pos, [width, height])
    out = np.zeros((len(arm), 4), np.float32)
    out[:, 0] = pos[:, 0] / width
    out[:, 1] = pos[:, 1] / height
    out[:, 2] = vel[:, 0] / width
    out[:, 3] = vel[:, 1] / height
    return out
