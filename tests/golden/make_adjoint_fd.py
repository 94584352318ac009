"""Writes tests/golden/adjoint_fd_100k.npz: FP64 central finite differences
of the adjoint workload's loss with respect to 64 initial-velocity
components, from the oracle's FP64 gravity-only leapfrog (orc_grav64_loss).

Workload: 100,000 bodies uniform on a 1440 x 720 torus (scenarios.uniform,
seed 1), P3M grid 512 x 256, G = 0.035, dt = 1/120 s, 64 steps, no contact
passes; loss = mean |minimum image(x_T - target)|^2 in px^2 with seeded
targets (scenarios.targets(n, 12)). Each difference perturbs one component
by +-h, h = 1e-3 px/s, and is formed from the two runs' final positions body
by body, (d+^2 - d-^2) = (d+ - d-)(d+ + d-), so the loss's own summation
rounding does not swamp the 1e-11 relative change of the loss.

Run from the repository root: python tests/golden/make_adjoint_fd.py
(about six minutes on eight cores).
"""
from __future__ import annotations

import multiprocessing as mp
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

N, W, H, STEPS, H_PX = 100_000, 1440, 720, 64, 1e-3
OUT = os.path.join(ROOT, "tests", "golden", "adjoint_fd_100k.npz")


def workload():
    from gravity_amd import scenarios
    bodies, mats = scenarios.uniform(N, W, H, seed=1)
    return bodies, mats, scenarios.targets(N, 12)


def params():
    from oracle import oracle as orc
    return orc.params(W, H, collide=False, substeps=1, grid=(512, 256), sort_interval=0)


def final_positions(dv):
    from oracle import oracle as orc
    bodies, mats, targets = workload()
    _, final = orc.grav64_loss(params(), bodies, mats, STEPS, targets, dv, final=True)
    return final


def one(job):
    j, c = job
    S = (W, H)[c]
    dv = np.zeros((N, 2))
    dv[j, c] = H_PX / S
    plus = final_positions(dv)
    minus = final_positions(-dv)
    _, _, targets = workload()
    view = np.array([W, H], np.float64)

    def offsets(x):
        d = (x - targets) * view
        return d - np.floor(d / view + 0.5) * view

    dp, dm = offsets(plus), offsets(minus)
    diff = ((dp - dm) * (dp + dm)).sum()
    return diff / N / (2.0 * H_PX / S)


def main():
    from oracle import oracle as orc
    orc.build()
    bodies, mats, targets = workload()
    loss, final = orc.grav64_loss(params(), bodies, mats, STEPS, targets, final=True)
    rng = np.random.default_rng(0)
    sample = rng.choice(N, 32, replace=False)
    jobs = [(int(j), c) for j in sample for c in (0, 1)]
    with mp.Pool(min(8, os.cpu_count() or 1)) as pool:
        fd = np.array(pool.map(one, jobs))
    view = np.array([W, H], np.float64)
    d = (final - targets) * view
    d -= np.floor(d / view + 0.5) * view
    ballistic = np.array([2.0 * d[j, c] * view[c] / N * STEPS / 120.0 for j, c in jobs])
    np.savez(OUT, bodies=np.array([j for j, _ in jobs], np.int64), axes=np.array([c for _, c in jobs], np.int64),
             fd=fd, ballistic=ballistic, loss=np.float64(loss),
             final=final[[j for j, _ in jobs]], h_px=np.float64(H_PX), steps=np.int64(STEPS))
    print(f"wrote {OUT}: loss {loss:.6f}, |fd| {np.linalg.norm(fd):.4e}, "
          f"gravity share {np.linalg.norm(fd - ballistic) / np.linalg.norm(fd):.3f}")


if __name__ == "__main__":
    main()
