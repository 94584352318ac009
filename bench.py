"""Benchmark of the MI355X particle step: body-steps/s at 400k bodies
(P3M + collisions), one GPU.

Workload: 400,000 FP32 bodies in walls on a 2880 x 1440 viewport, an
exponential-disc galaxy (scale radius 200 px, circular velocities from the
enclosed mass, 2% Gaussian dispersion), contact diameter 1 px, P3M grid
1024 x 512 (2048 x 1024 mesh under walls), 16 collision substeps, friction
0.2, seed 7. Inputs are resident in HBM before the timed region.

Prints one JSON line with the throughput, the roofline of the dominant
kernel (per-kernel HIP events on the library's stream) and a CPU baseline
(the oracle, timed on a bounded sample on the host).
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WIDTH, HEIGHT = 2880, 1440
N_BODIES = 400_000
GRID = (1024, 512)
SUBSTEPS = 16
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per launch of each kernel, per unit it processes: the
# logical FP32 SoA arrays a launch must read and write once (neighbour
# gathers and list traffic excluded).
#   solve : one contact pass of the byte model, read x, y, vx, vy, mass
#           (20 B), write x, y, vx, vy (16 B)                       36 B/body
#   count : read x, y (8 B), write the contact count (4 B)            12 B/body
#   kick  : read x, y, vx, vy (16 B), write vx, vy, ax, ay (16 B)   32 B/body
#   fft_columns : read + write one complex mesh value (16 B) and
#                 read the spectral kernel (8 B)                    24 B/node
#   fft_rows_forward : write one complex mesh value (8 B)            8 B/node
#   fft_rows_inverse : read + write one complex mesh value (16 B)   16 B/node
BYTES_PER_UNIT = {"count": 12, "solve": 36, "kick": 32, "drift_insert": 32, "fft_columns": 24,
                  "fft_rows_forward": 8, "fft_rows_inverse": 16}


def build_pipeline(n=N_BODIES):
    from gravity_amd import CollisionConfig, GravityConfig, GravityFieldConfig, GravityPipeline, ParticleConfig
    from gravity_amd import scenarios
    bodies, mats = scenarios.exponential_disc(n, WIDTH, HEIGHT, scale_px=200.0, dispersion=0.02,
                                              friction=0.2, seed=7)
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n, seed=7),
                        gravity=GravityFieldConfig(gridWidth=GRID[0], gridHeight=GRID[1]),
                        collisions=CollisionConfig(maximumSubsteps=SUBSTEPS, contactDiameterPixels=1.0,
                                                   friction=0.2, sleep=False))
    pipe = GravityPipeline(cfg, WIDTH, HEIGHT, scenarioState=bodies, bounded=True)
    pipe.upload(bodies, mats)
    pipe.initializeAcceleration(n)
    return pipe, bodies, mats


def units_per_launch(name, pipe, n):
    q = pipe.query()
    if name.startswith("fft"):
        mw, mh = q["meshWidth"], q["meshHeight"]
        if name == "fft_columns":
            return mw * mh
        return mw * (q["gridHeight"] + 2)  # rows the walls leave live
    return n


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "traffic_galaxy400k.md")


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE, scaled by a plain copy of known size), or None.
    A templated kernel's row may carry its template argument (k_kick<3>)."""
    if not os.path.exists(TRAFFIC_SUMMARY):
        return None
    symbol = "k_" + kernel
    for line in open(TRAFFIC_SUMMARY):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0].split("<")[0] == symbol:
            return float(cells[4])
    return None


ENSEMBLE_VIEWPORT = (1024, 1024)


def ensemble_configs():
    from gravity_amd import CollisionConfig, GravityConfig, GravityFieldConfig, ParticleConfig
    from gravity_amd import scenarios
    members = scenarios.ensemble(64, 100, *ENSEMBLE_VIEWPORT)
    cfgs = [GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n, seed=100 + k),
                          gravity=GravityFieldConfig(gridWidth=g, gridHeight=g, minimumGrid=16),
                          collisions=CollisionConfig(maximumSubsteps=SUBSTEPS, sleep=False))
            for k, (n, g, _, _) in enumerate(members)]
    return members, cfgs


ENSEMBLE_RELAX_STEPS = 40


def batched_leg(steps, warmup, plain=True, tune=True):
    """The ensemble of 64 torus simulations stepped as one batch (one graph
    launch per step) and, for comparison (unless plain is False), as 64 plain
    pipelines stepped one after another. Both rates are whole-ensemble
    body-steps/s.

    Both forms are autotuned first (the batch on its workload, each plain
    pipeline on its own; the tuners put the states back). Two windows are
    timed: `steps` steps after `warmup` from the initial spheres, whose
    overlapped cores keep the contact passes up to three times busier than
    later, and `steps` steps from step ENSEMBLE_RELAX_STEPS on, once the cores
    have spread out."""
    from gravity_amd import GravityBatch, GravityPipeline
    members, cfgs = ensemble_configs()
    total = sum(m[0] for m in members)
    relax_gap = max(ENSEMBLE_RELAX_STEPS - warmup - steps, 0)

    def windows(advance, sync):
        for _ in range(warmup):
            advance()
        sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            advance()
        sync()
        early = time.perf_counter() - t0
        for _ in range(relax_gap):
            advance()
        sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            advance()
        sync()
        return early, time.perf_counter() - t0

    with GravityBatch(cfgs, [ENSEMBLE_VIEWPORT] * len(cfgs), [m[2] for m in members],
                      bounded=[False] * len(cfgs)) as batch:
        for k, (n, _, b, mat) in enumerate(members):
            batch.upload(k, b, mat)
            batch.setActive(k, n)
        batch.initializeAcceleration()
        chosen, tuned_ms = batch.autotune(steps=2) if tune else (batch.tuning, None)
        t_batch, t_batch_relaxed = windows(batch.step, batch.synchronize)
    rate = total * steps / t_batch
    relaxed = {"from_step": warmup + steps + relax_gap, "value": total * steps / t_batch_relaxed,
               "ms_per_step": t_batch_relaxed * 1e3 / steps}
    record = {"workload": "ensemble-64: 64 torus simulations on 1024x1024, N = floor(10^U(3,5)), "
                          "grids 2^U{6..9} squared, Plummer spheres a = 50 px, seeds 100..163",
              "sims": len(members), "bodies": total, "steps": steps, "warmup": warmup,
              "value": rate, "unit": "body-steps/s", "ms_per_step": t_batch * 1e3 / steps,
              "tuning": dataclasses.asdict(chosen),
              "autotune_ms_per_step": tuned_ms, "relaxed": relaxed}
    if not plain:
        return record
    pipes = []
    try:
        for cfg, (n, _, b, mat) in zip(cfgs, members):
            p = GravityPipeline(cfg, *ENSEMBLE_VIEWPORT, scenarioState=b, bounded=False)
            p.upload(b, mat)
            p.initializeAcceleration(n)
            if tune:
                p.autotune(n, steps=2)
            pipes.append(p)

        def step_all():
            for p, m in zip(pipes, members):
                p.step(m[0])

        def sync_all():
            for p in pipes:
                p.synchronize()

        t_plain, t_plain_relaxed = windows(step_all, sync_all)
    finally:
        for p in pipes:
            p.close()
    loop = total * steps / t_plain
    record.update({"plain_loop": loop, "plain_ms_per_step": t_plain * 1e3 / steps,
                   "speedup_vs_plain": rate / loop})
    relaxed.update({"plain_loop": total * steps / t_plain_relaxed,
                    "plain_ms_per_step": t_plain_relaxed * 1e3 / steps,
                    "speedup_vs_plain": t_plain_relaxed / t_batch_relaxed})
    return record


def p3m_only_leg(steps=100, warmup=10):
    """The P3M-only workload: 50,000 bodies uniform on a 1440x720 torus
    (velocities 5 px/s Gaussian, seed 1), grid 512x256, mutual gravity,
    no contact passes, dt 1/120 s; `steps` graph-replayed steps."""
    from gravity_amd import scenarios
    from gravity_amd import CollisionConfig, GravityConfig, GravityFieldConfig, GravityPipeline, ParticleConfig
    n, W, H = 50_000, 1440, 720
    bodies, mats = scenarios.uniform(n, W, H, seed=1)
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n, seed=1),
                        gravity=GravityFieldConfig(gridWidth=512, gridHeight=256),
                        collisions=CollisionConfig(maximumSubsteps=1), collisionsEnabled=False)
    with GravityPipeline(cfg, W, H, scenarioState=bodies) as pipe:
        pipe.upload(bodies, mats)
        pipe.initializeAcceleration(n)
        for _ in range(warmup):
            pipe.step(n)
        pipe.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            pipe.step(n)
        pipe.synchronize()
        t = (time.perf_counter() - t0) / steps
        finite = bool(np.isfinite(pipe.bodies(n)).all())
    return {"workload": "p3m-50k: 50,000 bodies uniform on a 1440x720 torus, grid 512x256, "
                        "gravity only, dt 1/120, seed 1",
            "bodies": n, "steps": steps, "ms_per_step": t * 1e3, "value": n / t,
            "unit": "body-steps/s", "finite": finite}


def rigid_leg(projections=200, warmup=10):
    """Glued blocks: 400,000 bodies in 64 blocks of 6,250 members (hexagonal
    packings at the contact diameter) on a 2880x1440 walled viewport, each
    moving and spinning; `projections` rigid projections after a contact pass
    (tile sums, fit, snap), timed back to back. Algorithmic bytes: 28 B per
    member slot read by the tile sums (body 16, member 4, rest offset 8) and
    28 by the snap (member and rest offset read, body written)."""
    from gravity_amd import CollisionConfig, GravityConfig, GravityFieldConfig, GravityPipeline, ParticleConfig
    blocks, per, W, H, d = 64, 6250, 2880, 1440, 1.12838
    n = blocks * per
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n, seed=1),
                        gravity=GravityFieldConfig(gridWidth=1024, gridHeight=512),
                        collisions=CollisionConfig(contactDiameterPixels=d))
    rng = np.random.default_rng(1)
    cols = int(np.ceil(np.sqrt(per * 1.2)))
    k = np.arange(per)
    row, col = k // cols, k % cols
    offsets = np.stack([(col + 0.5 * (row % 2)) * d, row * d * np.sqrt(3.0) / 2], 1)
    offsets = (offsets - offsets.mean(axis=0)).astype(np.float32)
    state = np.zeros((n, 4), np.float32)
    mats = np.zeros((n, 4), np.float32)
    centres = [((i % 8 + 0.5) * W / 8, (i // 8 + 0.5) * H / 8) for i in range(blocks)]
    for i, (cx, cy) in enumerate(centres):
        v = rng.uniform(-40, 40, 2)
        spin = rng.uniform(-0.5, 0.5)
        pos = np.array([cx, cy]) + offsets
        vel = v + spin * np.stack([-offsets[:, 1], offsets[:, 0]], 1)
        state[i * per:(i + 1) * per] = np.concatenate([pos / [W, H], vel / [W, H]], 1)
        mats[i * per:(i + 1) * per] = (0.5, 0.5, 1.0, float(i))
    with GravityPipeline(cfg, W, H, scenarioState=state, bounded=True) as pipe:
        pipe.upload(state, mats)
        for i, (cx, cy) in enumerate(centres):
            pipe.glueBlock(i * per, offsets, cx, cy)
        for _ in range(warmup):
            pipe.projectBlocks(True)
        pipe.synchronize()
        t0 = time.perf_counter()
        for _ in range(projections):
            pipe.projectBlocks(True)
        pipe.synchronize()
        t = (time.perf_counter() - t0) / projections
        finite = bool(np.isfinite(pipe.bodies(n)).all())
        _, tiles, _ = pipe.blockSize()
    alg = 56 * tiles * 256
    return {"workload": "rigid-400k: 64 glued blocks of 6,250 members, walls, 2880x1440",
            "bodies": n, "projections": projections, "us_per_projection": t * 1e6,
            "value": n / t, "unit": "member-projections/s", "algorithmic_bytes": alg,
            "achieved_GBs": alg / t / 1e9, "launches_per_projection": 3, "finite": finite}


def force_error_leg(sample=1000):
    """FP64 direct-summation force error of the HIP path and of the oracle on
    the galaxy workload's initial state (walls: the isolated 2D sum
    a_i = sum_j G m_j d / (M max(r^2, r_min^2)), x scaled by the aspect),
    over `sample` seeded bodies: RMS of |a - a_ref| over RMS |a_ref|."""
    from oracle import oracle as orc
    pipe, bodies, mats = build_pipeline()
    n = len(bodies)
    gpu = pipe.acceleration(n).astype(np.float64)
    pipe.close()
    sim = orc.Sim(orc.params(WIDTH, HEIGHT, bounded=True, grid=GRID, diameter=1.0, substeps=SUBSTEPS),
                  bodies, mats)
    sim.initialize_acceleration()
    cpu = sim.state()[1].astype(np.float64)
    sim.close()
    aspect = WIDTH / HEIGHT
    cw, ch = aspect / GRID[0], 1.0 / GRID[1]
    min_d2 = (0.05 * min(cw, ch)) ** 2
    pts = bodies[:, :2].astype(np.float64) * np.array([aspect, 1.0])
    m = mats[:, 0].astype(np.float64)
    pick = np.random.default_rng(11).choice(n, sample, replace=False)
    ref = np.zeros((sample, 2))
    for a in range(0, sample, 50):
        rows = pick[a:a + 50]
        d = pts[None, :, :] - pts[rows, None, :]
        r2 = np.maximum((d * d).sum(-1), min_d2)
        r2[np.arange(len(rows)), rows] = np.inf
        ref[a:a + 50] = 0.035 / n * ((m / r2)[..., None] * d).sum(1)
    ref /= np.array([aspect, 1.0])

    def rms(a):
        e = (a[pick] - ref) * np.array([aspect, 1.0])
        r = ref * np.array([aspect, 1.0])
        return float(np.sqrt((e ** 2).sum(1).mean() / (r ** 2).sum(1).mean()))

    diff = (gpu - cpu) * np.array([aspect, 1.0])
    return {"workload": "galaxy-400k initial state, walls, grid 1024x512",
            "reference": "FP64 direct sum (isolated, minimum-distance floor)", "sample": sample,
            "hip_rms_relative": rms(gpu), "oracle_rms_relative": rms(cpu),
            "hip_vs_oracle_max_relative": float(np.abs(diff).max() /
                                                np.abs(cpu * np.array([aspect, 1.0])).max())}


def adjoint_leg(steps=64, every=8, warmup=5):
    """The adjoint workload: 100,000 bodies uniform on a 1440x720 torus, P3M
    grid 512x256, gravity only; 64 leapfrog steps recorded with a checkpoint
    every 8, then the VJP of the mean squared distance of the final positions
    to seeded targets with respect to the initial velocities (recomputing
    each checkpoint interval). The cost ratio is the VJP's time per step over
    a plain forward step's."""
    from gravity_amd import CollisionConfig, GravityConfig, GravityFieldConfig, GravityPipeline
    from gravity_amd import ParticleConfig, scenarios
    n, W, H = 100_000, 1440, 720
    bodies, mats = scenarios.uniform(n, W, H, seed=1)
    targets = scenarios.targets(n, 12)
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n, seed=1),
                        gravity=GravityFieldConfig(gridWidth=512, gridHeight=256),
                        collisions=CollisionConfig(maximumSubsteps=1), collisionsEnabled=False)
    with GravityPipeline(cfg, W, H, scenarioState=bodies) as pipe:
        pipe.upload(bodies, mats)
        pipe.initializeAcceleration(n)
        for _ in range(warmup):
            pipe.step(n)
        pipe.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            pipe.step(n)
        pipe.synchronize()
        t_fwd = (time.perf_counter() - t0) / steps
        # The first recording of this shape allocates the checkpoints; the
        # timed one reuses them.
        pipe.upload(bodies, mats)
        pipe.initializeAcceleration(n)
        pipe.synchronize()
        t0 = time.perf_counter()
        pipe.adjointRecord(n, steps, every)
        pipe.synchronize()
        t_setup = time.perf_counter() - t0
        pipe.upload(bodies, mats)
        pipe.initializeAcceleration(n)
        pipe.synchronize()
        t0 = time.perf_counter()
        pipe.adjointRecord(n, steps, every)
        pipe.synchronize()
        t_rec = time.perf_counter() - t0
        final = pipe.bodies(n)
        where = pipe.adjointOrder(n)
        view = np.array([W, H], np.float64)
        d = (final[where, :2].astype(np.float64) - targets) * view
        d -= np.floor(d / view + 0.5) * view
        dldx = np.zeros((n, 2), np.float32)
        dldx[where] = (2.0 * d * view / n).astype(np.float32)
        t0 = time.perf_counter()
        _, dv0 = pipe.adjointVjp(dldx, n)
        t_vjp = time.perf_counter() - t0
    return {"workload": "adjoint-100k: 100,000 bodies uniform on a 1440x720 torus, grid 512x256, "
                        "gravity only, 64 steps, checkpoint every 8, loss mean |x_T - target|^2",
            "bodies": n, "steps": steps, "checkpoint_every": every,
            "forward_ms_per_step": t_fwd * 1e3, "record_ms_per_step": t_rec * 1e3 / steps,
            "first_record_ms": t_setup * 1e3,
            "vjp_ms_per_step": t_vjp * 1e3 / steps,
            "cost_vs_forward": (t_vjp / steps) / t_fwd,
            "value": n * steps / (t_rec + t_vjp), "unit": "body-steps/s (record + VJP)",
            "gradient_finite": bool(np.isfinite(dv0).all())}


def liquid_leg(steps=40, warmup=5):
    """The liquid-and-sand workload: 400,000 bodies in walls on 2880x1440, a
    uniform field of -500 px/s^2 in y, no mutual gravity; 60% fluid bodies
    in a layer at one body per px^2 and 40% hard grains in a heap
    (scenarios.liquid_and_sand, seed 3), contact diameter 1.0746 px, 16
    substeps, 12 density iterations and XSPH viscosity per step. Reports the
    graph-replayed step and the liquid stage's share of it."""
    from gravity_amd import (CollisionConfig, GravityConfig, GravityFieldConfig, GravityPipeline,
                             LiquidConfig, ParticleConfig, StartConfig, scenarios)
    n, W, H, d = 400_000, 2880, 1440, 1.0746
    bodies, mats = scenarios.liquid_and_sand(n, W, H, diameter=d, seed=3)
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n, seed=3),
                        gravity=GravityFieldConfig(gridWidth=1024, gridHeight=512),
                        collisions=CollisionConfig(contactDiameterPixels=d, maximumSubsteps=SUBSTEPS),
                        liquid=LiquidConfig(iterations=12, viscosity=0.1),
                        start=StartConfig(walls=True, mutualGravity=False, liquid=True))
    with GravityPipeline(cfg, W, H, scenarioState=bodies, bounded=True) as pipe:
        pipe.upload(bodies, mats)
        pipe.fieldAccelerationY = -500.0
        pipe.initializeAcceleration(n)
        for _ in range(warmup):
            pipe.step(n)
        pipe.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            pipe.step(n)
        pipe.synchronize()
        t = (time.perf_counter() - t0) / steps
        stats = pipe.profile(2)
        finite = bool(np.isfinite(pipe.bodies(n)).all())
    total = sum(ms for _, _, ms in stats)
    liquid = sum(ms for name, _, ms in stats if name.startswith("liquid"))
    return {"workload": "liquid-sand-400k: walls, 2880x1440, field -500 px/s^2, 60% PBF fluid "
                        "layer at 1 body/px^2, 40% grains heap, diameter 1.0746 px, 16 substeps, "
                        "12 density iterations, seed 3",
            "bodies": n, "steps": steps, "ms_per_step": t * 1e3, "value": n / t,
            "unit": "body-steps/s", "liquid_stage_fraction": liquid / total if total else None,
            "finite": finite}


def cpu_baseline(bodies, mats, steps=2):
    """The oracle (single-threaded C restatement) on `steps` steps of the same workload."""
    from oracle import oracle as orc
    p = orc.params(WIDTH, HEIGHT, bounded=True, grid=GRID, diameter=1.0, substeps=SUBSTEPS)
    sim = orc.Sim(p, bodies, mats)
    sim.initialize_acceleration()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    dt = time.perf_counter() - t0
    sim.close()
    return len(bodies) * steps / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=8)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=20,
                    help="steps of the batched-ensemble leg (0 skips it)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-liquid", action="store_true", help="skip the liquid-and-sand leg")
    ap.add_argument("--no-p3m", action="store_true", help="skip the P3M-only leg")
    ap.add_argument("--no-force-error", action="store_true", help="skip the FP64 force-error leg")
    ap.add_argument("--no-rigid", action="store_true", help="skip the rigid-projection leg")
    ap.add_argument("--eager", action="store_true",
                    help="launch every kernel eagerly instead of through captured graphs")
    ap.add_argument("--leg", choices=("batched", "batch", "adjoint", "liquid", "p3m", "force", "rigid"),
                    help="run only this leg and print its record (batch: the batched ensemble "
                         "without the plain loop), for kernel profiles of one workload")
    ap.add_argument("--leg-steps", type=int, default=0,
                    help="timed steps of the liquid or P3M-only leg run with --leg (0: its default)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 5 times (the known-size kernel "
                         "that scales the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        src = torch.ones(args.calibrate_copy << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": src.numel() * 4}))
        return

    if args.leg:
        bw = max(2, args.warmup // 4)
        k = args.leg_steps
        legs = {"batched": lambda: batched_leg(args.batched_steps, bw),
                "batch": lambda: batched_leg(args.batched_steps, bw, plain=False),
                "adjoint": adjoint_leg,
                "liquid": (lambda: liquid_leg(k, 2)) if k else liquid_leg,
                "p3m": (lambda: p3m_only_leg(k, 2)) if k else p3m_only_leg,
                "force": force_error_leg,
                "rigid": (lambda: rigid_leg(k, 2)) if k else rigid_leg}
        print(json.dumps({args.leg: legs[args.leg]()}))
        return

    pipe, bodies, mats = build_pipeline()
    if args.eager:
        pipe.setGraphs(False)
    n = N_BODIES
    print(f"[bench] warmup {args.warmup} steps", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        pipe.step(n)
    pipe.synchronize()
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pipe.step(n)
    pipe.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = n * args.steps / elapsed

    print(f"[bench] profiling {args.profile_steps} steps with per-kernel events", file=sys.stderr, flush=True)
    stats = pipe.profile(args.profile_steps)
    per_kernel = {name: (launches, ms) for name, launches, ms in stats}
    dominant = max(per_kernel, key=lambda k: per_kernel[k][1])
    launches, total_ms = per_kernel[dominant]
    eager_s = total_ms / launches / 1e3
    # The timed steps are graph replays, whose kernels run back to back: the
    # step time beyond the kernels' own (dispatch-stamped) time is the
    # launches' turnaround, shared evenly over the step's launches. A kernel's
    # graph-replayed time is its own plus that share, which is what a kernel
    # trace of the replays reports (each launch starts as the one before ends).
    launches_per_step = sum(c for c, _ in per_kernel.values()) / args.profile_steps
    kernel_ms_per_step = sum(ms for _, ms in per_kernel.values()) / args.profile_steps
    turnaround_s = max(ms_per_step - kernel_ms_per_step, 0.0) / 1e3 / launches_per_step
    avg_s = eager_s + (0.0 if args.eager else turnaround_s)
    units = units_per_launch(dominant, pipe, n)
    alg_bytes = BYTES_PER_UNIT.get(dominant, 0) * units
    achieved = alg_bytes / avg_s / 1e9
    b = pipe.bodies(n)
    finite = bool(np.isfinite(b).all())
    pipe.close()

    batched = None
    if args.batched_steps > 0:
        print(f"[bench] batched ensemble: {args.batched_steps} steps, batch and plain loop",
              file=sys.stderr, flush=True)
        batched = batched_leg(args.batched_steps, max(2, args.warmup // 4))

    p3m = None
    if not args.no_p3m:
        print("[bench] P3M only: 50k bodies, 100 steps", file=sys.stderr, flush=True)
        p3m = p3m_only_leg()

    force = None
    if not args.no_force_error:
        print("[bench] FP64 force error of the galaxy's initial state", file=sys.stderr, flush=True)
        force = force_error_leg()

    adjoint = None
    if not args.no_adjoint:
        print("[bench] adjoint: 64 steps recorded, VJP", file=sys.stderr, flush=True)
        adjoint = adjoint_leg()

    liquid = None
    if not args.no_liquid:
        print("[bench] liquid and sand: 40 steps", file=sys.stderr, flush=True)
        liquid = liquid_leg()

    rigid = None
    if not args.no_rigid:
        print("[bench] rigid projection: 64 blocks, 400k members", file=sys.stderr, flush=True)
        rigid = rigid_leg()

    cpu = None
    if not args.no_cpu:
        print(f"[bench] CPU baseline: oracle, {args.cpu_steps} steps", file=sys.stderr, flush=True)
        rate, secs = cpu_baseline(bodies, mats, args.cpu_steps)
        cpu = {"value": rate, "unit": "body-steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the same 400k workload ({secs:.1f} s, single thread)"}

    line = {
        "metric": "body-steps/s at 400k bodies (P3M+collisions); batched & adjoint vs plain, 1 GPU",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "galaxy-400k: exponential disc, walls, 2880x1440, grid 1024x512, "
                               "16 substeps, friction 0.2, diameter 1 px, seed 7",
                   "bodies": n, "grid": list(GRID), "substeps": SUBSTEPS, "bounded": True,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dominant),
                     "avg_launch_us": avg_s * 1e6, "bytes_per_launch": alg_bytes,
                     "timing": "eager" if args.eager else "graph-replayed",
                     "dispatch_stamped_us": eager_s * 1e6, "turnaround_us": turnaround_s * 1e6},
        "kernels_ms_per_step": {k: v[1] / args.profile_steps for k, v in per_kernel.items()},
        "finite": finite,
        "batched": batched,
        "adjoint": adjoint,
        "liquid": liquid,
        "p3m_only": p3m,
        "force_error": force,
        "rigid": rigid,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
