"""Seeded starting states of the benchmark workloads (the reference's
Scenario / UniformParticleInitializer role, K/Scenario.kt,
K/UniformParticleInitializer.kt). The reference draws from
kotlin.random.Random; these use numpy's PCG64 with the stated seeds, and the
same arrays are uploaded to every code under test.

Every generator returns (bodies, materials): float32 arrays of shape (n, 4)
in the reference's units, vec4(x/W, y/H, vx/W, vy/H) and
vec4(mass, friction, solid, group).
"""
from __future__ import annotations

import math

import numpy as np

MAXIMUM_UNIT = np.float32(0.99999994)


def _materials(n: int, friction: float = 0.0) -> np.ndarray:
    m = np.zeros((n, 4), dtype=np.float32)
    m[:, 0] = 1.0
    m[:, 1] = friction
    return m


def _pack(x_px, y_px, vx_px, vy_px, width, height, bounded) -> np.ndarray:
    f = np.float32
    W, H = f(width), f(height)
    b = np.empty((len(x_px), 4), dtype=np.float32)
    x = np.asarray(x_px, dtype=np.float32) / W
    y = np.asarray(y_px, dtype=np.float32) / H
    if bounded:
        x = np.clip(x, f(0), MAXIMUM_UNIT)
        y = np.clip(y, f(0), MAXIMUM_UNIT)
    else:
        x = x - np.floor(x)
        y = y - np.floor(y)
        x[x >= 1] = 0
        y[y >= 1] = 0
    b[:, 0], b[:, 1] = x, y
    b[:, 2] = np.asarray(vx_px, dtype=np.float32) / W
    b[:, 3] = np.asarray(vy_px, dtype=np.float32) / H
    return b


def uniform(n: int, width: int, height: int, sigma_px: float = 5.0, seed: int = 1):
    """Uniform positions on the torus, zero-mean Gaussian velocities of
    sigma px/s per axis (the gravity-only P3M workload)."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, width, n)
    y = rng.uniform(0, height, n)
    v = rng.normal(0.0, sigma_px, (n, 2))
    v -= v.mean(axis=0)
    return _pack(x, y, v[:, 0], v[:, 1], width, height, False), _materials(n)


def exponential_disc(n: int, width: int, height: int, scale_px: float = 200.0, dispersion: float = 0.02,
                     gravity: float = 0.035, friction: float = 0.2, seed: int = 7, bounded: bool = True):
    """Exponential disc (surface density ~ exp(-r / scale)) on circular orbits
    from the enclosed mass (K/Scenario.kt:49-54, 207-247): v^2 = G f H^2 in
    the reference's 2D normalisation, with `dispersion` Gaussian scatter in
    the tangential and radial speeds. Truncated to stay 20 px inside the
    shorter half extent."""
    rng = np.random.default_rng(seed)
    rmax = 0.5 * min(width, height) - 20.0
    r = np.empty(0)
    while r.size < n:
        draw = rng.gamma(2.0, scale_px, 2 * n)
        r = np.concatenate([r, draw[draw <= rmax]])
    r = r[:n]
    phi = rng.uniform(0.0, 2 * math.pi, n)

    def enclosed(x):
        return 1.0 - (1.0 + x) * np.exp(-x)

    frac = enclosed(r / scale_px) / enclosed(rmax / scale_px)
    effective = frac
    if not bounded:
        effective = np.maximum(frac - math.pi * r * r / (width * height), 0.0)
    speed = np.sqrt(gravity * effective) * height
    tangential = speed * (1.0 + dispersion * rng.standard_normal(n))
    radial = dispersion * speed * rng.standard_normal(n)
    cx, cy = 0.5 * width, 0.5 * height
    x = cx + r * np.cos(phi)
    y = cy + r * np.sin(phi)
    vx = np.cos(phi) * radial - np.sin(phi) * tangential
    vy = np.sin(phi) * radial + np.cos(phi) * tangential
    vx -= vx.mean()
    vy -= vy.mean()
    return _pack(x, y, vx, vy, width, height, bounded), _materials(n, friction)


def plummer(n: int, width: int, height: int, a_px: float = 50.0, gravity: float = 0.035, seed: int = 100):
    """Plummer sphere projected on the plane (surface density ~ (1 + r^2/a^2)^-2)
    with isotropic Gaussian velocities at the 2D virial speed
    <v^2> = G H^2 / 2 of the reference's normalisation, centred on the torus."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, 0.999, n)
    r = a_px * np.sqrt(u / (1.0 - u))
    r = np.minimum(r, 0.45 * min(width, height))
    phi = rng.uniform(0.0, 2 * math.pi, n)
    sigma = math.sqrt(gravity / 2.0) * height / math.sqrt(2.0)
    v = rng.normal(0.0, sigma, (n, 2))
    v -= v.mean(axis=0)
    x = 0.5 * width + r * np.cos(phi)
    y = 0.5 * height + r * np.sin(phi)
    return _pack(x, y, v[:, 0], v[:, 1], width, height, False), _materials(n)


def liquid_and_sand(n: int, width: int, height: int, diameter: float = 1.0746, seed: int = 3):
    """60% fluid bodies in a uniform random layer on the floor at one body per
    px^2, 40% hard grains (solid, no block) in a triangular heap above it."""
    rng = np.random.default_rng(seed)
    nl = int(round(0.6 * n))
    ns = n - nl
    depth = nl / width
    lx = rng.uniform(0.5 * diameter, width - 0.5 * diameter, nl)
    ly = rng.uniform(0.5 * diameter, depth, nl)
    # Heap: isosceles triangle of area ns px^2 with a 2:1 base, above the layer.
    hh = math.sqrt(ns / 2.0)
    base = 4.0 * hh
    sx, sy = [], []
    while len(sx) < ns:
        px = rng.uniform(-0.5 * base, 0.5 * base, 2 * ns)
        py = rng.uniform(0.0, hh, 2 * ns)
        keep = np.abs(px) <= 0.5 * base * (1.0 - py / hh)
        sx.extend(px[keep]); sy.extend(py[keep])
    sx = np.array(sx[:ns]) + 0.5 * width
    sy = np.array(sy[:ns]) + depth + 2.0 * diameter
    x = np.concatenate([lx, sx])
    y = np.concatenate([ly, sy])
    bodies = _pack(x, y, np.zeros(n), np.zeros(n), width, height, True)
    mats = _materials(n)
    mats[nl:, 2] = 1.0   # solid
    mats[nl:, 3] = -1.0  # a loose grain, no block
    return bodies, mats


def targets(n: int, seed: int) -> np.ndarray:
    """Seeded uniform target positions (normalised) of the adjoint loss."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, (n, 2)).astype(np.float32)


def ensemble(count: int = 64, first_seed: int = 100, width: int = 1024, height: int = 1024):
    """The batched-ensemble workload: `count` independent torus simulations,
    seeds first_seed.., each with N = floor(10^U(3, 5)) bodies in a Plummer
    sphere (a = 50 px, virial velocities) and a square P3M grid of
    2^U{6..9} cells a side. Returns (N, grid, bodies, materials) per member."""
    out = []
    for k in range(count):
        seed = first_seed + k
        rng = np.random.default_rng(seed)
        n = int(math.floor(10.0 ** rng.uniform(3.0, 5.0)))
        grid = int(2 ** rng.integers(6, 10))
        bodies, mats = plummer(n, width, height, a_px=50.0, seed=seed)
        out.append((n, grid, bodies, mats))
    return out
