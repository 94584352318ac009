"""Scenes of the reference's GPU tests, restated as data
(T/CollisionShaderIntegrationTest.kt:21-35, T/SpatialSortIntegrationTest.kt:23-33).
Random positions come from numpy's seeded generator instead of
kotlin.random (whose stream nothing in the reference pins); the same arrays go
to the HIP path and to the oracle."""
from __future__ import annotations

import numpy as np

WIDTH, HEIGHT = 3440, 1440
DIAMETER = np.float32(1.12838)


def scene(rows, width=WIDTH, height=HEIGHT) -> np.ndarray:
    """Pixel rows (x, y, vx, vy) -> normalised float32 bodies, as scene() does:
    value / (width or height) in float."""
    a = np.asarray(rows, dtype=np.float32).reshape(-1, 4)
    out = np.empty_like(a)
    out[:, 0] = a[:, 0] / np.float32(width)
    out[:, 1] = a[:, 1] / np.float32(height)
    out[:, 2] = a[:, 2] / np.float32(width)
    out[:, 3] = a[:, 3] / np.float32(height)
    return out


def collision_scenes(seed=7235) -> dict[str, np.ndarray]:
    rng = np.random.default_rng(seed)
    d = float(DIAMETER)
    scenes = {}
    iso = np.zeros((128, 4), dtype=np.float32)
    iso[:, 0] = rng.random(128, dtype=np.float32)
    iso[:, 1] = rng.random(128, dtype=np.float32)
    scenes["isolated"] = iso
    scenes["pair"] = scene([(100 + i * 0.8, 100, 10 if i == 0 else -10, 0) for i in range(2)])
    scenes["contact-cache overflow"] = scene(
        [(500 + (i % 5) * 0.12, 500 + (i // 5) * 0.12, i, -i) for i in range(20)])
    scenes["neighbor-cache overflow"] = scene(
        [(700 + (i % 13) * 0.025, 600 + (i // 13) * 0.025, i, -i) for i in range(130)])
    scenes["coincident"] = scene([(700, 600, 0, 0)] * 130)
    scenes["seams"] = scene([(0.1 if i % 2 == 0 else WIDTH - 0.1, 0.1 if i % 4 < 2 else HEIGHT - 0.1,
                              i * 11, -i * 7) for i in range(16)])
    scenes["square"] = scene([(1000 + i % 16, 700 + i // 16, 0, 0) for i in range(256)])
    gaps = [np.nextafter(DIAMETER, np.float32(0)), DIAMETER, np.nextafter(DIAMETER, np.float32(10))]
    scenes["contact threshold"] = scene(
        [(np.float32(10 + (i // 2) * 4) + np.float32((i % 2)) * gaps[(i // 2) % 3], 10, 0, 0) for i in range(128)])
    return scenes


def sort_scene(count=6000, seed=4242):
    """T/SpatialSortIntegrationTest.kt:23-33: a dense block, coincident
    bodies, both seams, the far corner, then uniform bodies; shuffled."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(count):
        if i < 2500:
            rows.append((1200 + (i % 50) * 0.9, 700 + (i // 50) * 0.9, i, -i))
        elif i < 2600:
            rows.append((300, 300, 1, 2))
        elif i < 2700:
            rows.append((0 if i % 2 == 0 else WIDTH - 0.01, (i % 97) * 14, -3, 3))
        elif i < 2710:
            rows.append((WIDTH * 0.99999994, HEIGHT * 0.99999994, 0, 0))
        else:
            rows.append((rng.random() * WIDTH, rng.random() * HEIGHT, rng.random(), 0))
    b = scene(rows)
    return b[rng.permutation(count)]


def brute_force_contacts(bodies: np.ndarray, width=WIDTH, height=HEIGHT, diameter=DIAMETER, bounded=False):
    """O(N^2) contact sets with the float predicate of the reference's tests
    (T/CollisionShaderIntegrationTest.kt:422-435), lower index first."""
    f = np.float32
    n = len(bodies)
    W, H = f(width), f(height)
    x = bodies[:, 0].astype(np.float32)
    y = bodies[:, 1].astype(np.float32)
    i, j = np.triu_indices(n, 1)
    dx = x[i] * W - x[j] * W
    dy = y[i] * H - y[j] * H
    if not bounded:
        dx = dx - np.floor(dx / W + f(0.5)) * W
        dy = dy - np.floor(dy / H + f(0.5)) * H
    d2 = dx * dx + dy * dy
    hit = d2 < f(diameter) * f(diameter)
    out = [[] for _ in range(n)]
    for a, b in zip(i[hit], j[hit]):
        out[a].append(int(b))
        out[b].append(int(a))
    return [sorted(c) for c in out]
