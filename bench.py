#!/usr/bin/env python3
"""Headline benchmark: explicit APIC MPM step on one MI355X.

Workload: 3D FP32, 256^3 cells on [0,1]^3, twelve non-overlapping cubes with
sides uniform in [0.08, 0.2] placed in [0.1, 0.9]^3 (box seed 46767, chosen
so the draw holds ~8.0M particles), 8 jittered particles per cell (seed 1234),
fixed-corotated with per-box E log-uniform in [1e3, 1e5] and nu uniform in
[0.2, 0.4], density 1000, gravity -9.8 y, sticky floor below y = 0.05,
dt = 2e-5.  A step is one MPM_DRIVER::Advance_One_Time_Step over all particles
with inputs resident in HBM.

Prints one JSON line (see --help); per-phase GPU times come from HIP events
recorded on the stream the kernels run on, inside the timed region.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from physbam_amd import mpmx  # noqa: E402

METRIC = "particle-steps/s, 8M-particle 3D APIC MPM; fwd+adjoint / fwd time ratio"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E spec peak
BOX_SEED = 46767
SEED = 1234
CELLS = 256
DT = 2e-5

# Algorithmic HBM bytes (FP32, 3D).  Per particle: P2G reads X, V, F, B, m,
# V0, mu, lambda (112 B); G2P reads X, F (48 B) and writes X, V, F, B (96 B).
# Per active grid node: P2G result write 16 B; grid update read 16 + write
# 12 B; G2P read 12 B.
# g2p2g (the fused G2P of one step and P2G of the next): read X, F, m, V0, mu,
# lambda (64 B), write X, V, F, B (96 B); per node read v (12 B), write m, p (16 B)
BYTES_PER_PARTICLE = {"p2g": 112.0, "g2p": 144.0, "g2p2g": 160.0, "grid": 0.0, "step": 256.0}
BYTES_PER_NODE = {"p2g": 16.0, "g2p": 12.0, "g2p2g": 28.0, "grid": 28.0, "step": 56.0}


def headline_boxes():
    lo, hi, E, nu = mpmx.random_boxes(BOX_SEED, nboxes=12, side_min=0.08, side_max=0.2, region_lo=0.1,
                                      region_hi=0.9, E_min=1e3, E_max=1e5, nu_min=0.2, nu_max=0.4, cubes=True)
    return [dict(type=mpmx.SHAPE_BOX, a=lo[i], b=hi[i], density=1000.0, E=E[i], nu=nu[i], v=[0.0, 0.0, 0.0])
            for i in range(len(lo))]


def floor_wall():
    # STANDARD_TESTS_BASE::Add_Walls geometry for the -y side: the domain
    # thickened by 7 cells, cut at y = 0.05.
    h = 1.0 / CELLS
    lo = [-7 * h, -7 * h, -7 * h]
    hi = [1 + 7 * h, 0.05, 1 + 7 * h]
    return lo, hi


def make_context(P):
    ctx = mpmx.Context(3, CELLS, [0.0] * 3, [1.0] * 3, precision=mpmx.FP32)
    ctx.set_constitutive(mpmx.FIXED_COROTATED)
    ctx.set_gravity([0.0, -9.8, 0.0])
    lo, hi = floor_wall()
    ctx.add_collision_box(mpmx.STICK, 0.0, lo, hi)
    ctx.set_particles(P["X"], P["V"], P["F"], P["B"], P["mass"], P["volume"], P["mu"], P["lam"], P["valid"])
    return ctx


def cpu_baseline(shapes, log):
    """Reference CPU step (the PhysBAM MPM_DRIVER built under oracle/_ref) on a
    bounded sample of the same workload: the first box of the scene alone at
    the same resolution, particles per cell, materials, floor and dt."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
    P = mpmx.seed_particles(3, CELLS, [0.0] * 3, [1.0] * 3, shapes[:1], ppc_axis=2, seed=SEED,
                            precision=mpmx.FP32)
    n = int(P["X"].shape[0])
    lo, hi = floor_wall()
    sc = dict(dim=3, prec=4, model=0, N=[CELLS] * 3, lo=[0.0] * 3, hi=[1.0] * 3, gravity=[0.0, -9.8, 0.0], dt=DT,
              neo_threshold=0.25, walls=[dict(type=mpmx.STICK, friction=0.0, lo=lo, hi=hi)], X=P["X"], V=P["V"],
              F=P["F"], B=P["B"], mass=P["mass"], volume=P["volume"], mu=P["mu"], valid=P["valid"],
              **{"lambda": P["lam"]})
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    warm, steps = 1, 3
    if os.access(harness, os.X_OK):
        sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
        import scene_io

        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "sample.bin")
            scene_io.write_scene(path, sc, 0, 0, threads)
            log(f"cpu baseline: reference MPM_DRIVER, {n} particles, {threads} threads")
            r = subprocess.run([harness, "time", path, str(threads), str(warm), str(steps)], capture_output=True,
                               text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"reference harness failed: {r.stderr[-400:]}")
        res = json.loads(r.stdout.strip().splitlines()[-1])
        return {"value": res["particle_steps_per_s"], "unit": "particle-steps/s", "cores": threads,
                "kind": "reference",
                "sample": f"one cube of the 12-box scene ({n} particles, 256^3 grid, fixed-corotated, sticky floor, "
                          f"dt 2e-5), {warm} warm-up + {steps} timed MPM_DRIVER::Advance_One_Time_Step calls, "
                          f"FP32, OpenMP threads={threads}"}
    # Fallback: the oracle restatement, serial.
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import mpmtest

    log(f"cpu baseline: oracle restatement, {n} particles, 1 thread")
    mpmtest.oracle_run(sc, 1, prec=4, want_grid=False)
    t0 = time.perf_counter()
    mpmtest.oracle_run(sc, 1, prec=4, want_grid=False)
    sec = time.perf_counter() - t0
    return {"value": n / sec, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"one cube of the 12-box scene ({n} particles), 1 step of the serial oracle, FP32"}


def batched(log, nscenes=256, steps=100, seed=BOX_SEED):
    """Batched configuration: independent 3D scenes in one context, grid N
    uniform in {32, 48, 64, 96}, one sphere of radius U[0.1, 0.3] at the centre,
    8 ppc, E log-U[1e3, 1e5], nu U[0.2, 0.4], fixed corotated, gravity, FP32,
    dt 1e-4; every step advances all scenes with the same launches."""
    rng = np.random.default_rng(seed)
    grids, parts = [], []
    t0 = time.time()
    for s in range(nscenes):
        N = int(rng.choice([32, 48, 64, 96]))
        r = float(rng.uniform(0.1, 0.3))
        E = float(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
        nu = float(rng.uniform(0.2, 0.4))
        sph = [dict(type=mpmx.SHAPE_SPHERE, a=[0.5, 0.5, 0.5], b=[r, 0.0, 0.0], density=1000.0, E=E, nu=nu,
                    v=[0.0] * 3)]
        grids.append((N, [0.0] * 3, [1.0] * 3))
        parts.append(mpmx.seed_particles(3, N, [0.0] * 3, [1.0] * 3, sph, ppc_axis=2, seed=SEED + s,
                                         precision=mpmx.FP32))
    counts = [int(P["X"].shape[0]) for P in parts]
    n = sum(counts)
    cat = lambda k: np.concatenate([P[k] for P in parts])
    log(f"batched: {nscenes} scenes, {n} particles ({min(counts)}..{max(counts)}), seeded in {time.time() - t0:.1f}s")
    ctx = mpmx.Context.batch(3, grids, precision=mpmx.FP32)
    ctx.set_constitutive(mpmx.FIXED_COROTATED)
    ctx.set_gravity([0.0, -9.8, 0.0])
    ctx.set_batch_particles(counts, cat("X"), cat("V"), cat("F"), cat("B"), cat("mass"), cat("volume"), cat("mu"),
                            cat("lam"), cat("valid"))
    del parts
    dt = 1e-4
    ctx.step(dt, 5)
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.step(dt, steps)
    sec = time.perf_counter() - t1
    ctx.close()
    return {"value": n * steps / sec, "unit": "particle-steps/s", "ms_per_step": sec * 1e3 / steps,
            "scenes": nscenes, "particles": n, "particles_min": min(counts), "particles_max": max(counts),
            "steps": steps,
            "workload": f"{nscenes} independent 3D scenes, N in {{32,48,64,96}}^3, one sphere r U[0.1,0.3], 8 ppc, "
                        "fixed-corotated, E log-U[1e3,1e5], nu U[0.2,0.4], gravity, dt 1e-4, FP32, one context"}


def adjoint_ratio(log, steps=200, checkpoint=20):
    """Differentiable configuration: 64^3 grid, one 0.3^3 box of 8 ppc
    (~56.6k particles), fixed corotated E 1e4, nu 0.3, gravity, dt 1e-4,
    200 steps, checkpoint every 20, FP32; loss |mean x(T) - (0.7, 0.3, 0.5)|^2.
    Returns the time of a taped forward + reverse pass over the time of a
    plain forward pass of the same steps, both after an untimed warm-up of
    each (as in a training loop past its first iteration), each the median of
    three runs."""
    cells, dt = 64, 1e-4
    box = [dict(type=mpmx.SHAPE_BOX, a=[0.35] * 3, b=[0.65] * 3, density=1000.0, E=1e4, nu=0.3, v=[0.0] * 3)]
    P = mpmx.seed_particles(3, cells, [0.0] * 3, [1.0] * 3, box, ppc_axis=2, seed=SEED, precision=mpmx.FP32)
    n = int(P["X"].shape[0])
    ctx = mpmx.Context(3, cells, [0.0] * 3, [1.0] * 3, precision=mpmx.FP32)
    ctx.set_constitutive(mpmx.FIXED_COROTATED)
    ctx.set_gravity([0.0, -9.8, 0.0])
    target = np.array([0.7, 0.3, 0.5])

    def load():
        ctx.set_particles(P["X"], P["V"], P["F"], P["B"], P["mass"], P["volume"], P["mu"], P["lam"], P["valid"])

    load()
    ctx.step(dt, 20)  # warm-up
    # one untimed taped run as a training loop's first iteration: it sizes the
    # tape storage (checkpoints, segment states and grids), kept for later tapes
    load()
    ctx.tape_begin(checkpoint)
    ctx.step(dt, steps)
    ctx.backward(dX=np.zeros((n, 3)))
    def forward():
        load()
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.step(dt, steps)
        ctx.synchronize()
        return time.perf_counter() - t0

    def forward_adjoint():
        load()
        ctx.step(dt, 1)
        load()
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.tape_begin(checkpoint)
        ctx.step(dt, steps)
        X = ctx.get_particles()["X"].astype(np.float64)
        dX = np.broadcast_to(2 * (X.mean(0) - target) / n, X.shape)
        g = ctx.backward(dX=dX)
        return time.perf_counter() - t0, g

    # the median of three runs of each (the runs last tens of milliseconds)
    t_fwd = float(np.median([forward() for _ in range(3)]))
    runs = [forward_adjoint() for _ in range(3)]
    t_fa = float(np.median([t for t, _ in runs]))
    g = runs[-1][1]
    ctx.close()
    assert np.isfinite(g["V"]).all() and np.isfinite(g["mu"]).all()
    log(f"adjoint: {n} particles, forward {t_fwd * 1e3:.1f} ms, forward+adjoint {t_fa * 1e3:.1f} ms")
    return {"ratio": t_fa / t_fwd, "forward_ms_per_step": t_fwd * 1e3 / steps,
            "forward_adjoint_ms_per_step": t_fa * 1e3 / steps, "particles": n, "steps": steps,
            "checkpoint_every": checkpoint,
            "workload": "64^3 grid, one 0.3^3 box at 8 ppc, fixed-corotated, gravity, dt 1e-4, FP32, "
                        "loss |mean x(T) - (0.7, 0.3, 0.5)|^2, gradients w.r.t. v(0) and Lame parameters"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the reference CPU timing")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward+adjoint / forward ratio")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-scenes throughput")
    ap.add_argument("--reorder-interval", type=int, default=16)
    ap.add_argument("--rebin-fraction", type=float, default=1.0 / 128,
                    help="rebuild the bins when more than this fraction of particles left their binned tile")
    ap.add_argument("--rebin-interval", type=int, default=0, help="also rebuild the bins every K steps (0: never)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only run the forward+adjoint / forward measurement and print its JSON")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (calibrates HBM byte counters)")
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)

    if args.copy_probe:
        import torch

        n = args.copy_probe * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log(f"copy probe: 3 copies of {args.copy_probe} MiB")
        return

    if args.adjoint_only:
        print(json.dumps(adjoint_ratio(log)), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)

    t0 = time.time()
    shapes = headline_boxes()
    P = mpmx.seed_particles(3, CELLS, [0.0] * 3, [1.0] * 3, shapes, ppc_axis=2, seed=SEED, precision=mpmx.FP32)
    n = int(P["X"].shape[0])
    log(f"scene: {len(shapes)} boxes, {n} particles, seeded in {time.time() - t0:.1f}s")

    ctx = make_context(P)
    del P
    ctx.set_reorder_interval(args.reorder_interval)
    ctx.set_rebin_policy(args.rebin_fraction, args.rebin_interval)
    ctx.step_async(DT, args.warmup)
    ctx.synchronize()
    log(f"warm-up: {args.warmup} steps done")

    ctx.enable_timing(True)
    ctx.reset_timing()
    rebins0, _ = ctx.bin_stats()
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.step_async(DT, args.steps)
    ctx.synchronize()
    t2 = time.perf_counter()
    sec = t2 - t1
    log(f"timed: {args.steps} steps in {sec * 1e3:.2f} ms")

    rebins1, movers_last = ctx.bin_stats()
    phases = {}
    for ph in ("rebin", "bin", "p2g", "grid", "g2p", "g2p2g", "g2p2g_kernel", "reorder", "step"):
        tot, cnt = ctx.phase_time(ph)
        phases[ph] = (tot, cnt)
    ctx.enable_timing(False)
    mass, _ = ctx.get_grid()
    nodes = int(np.count_nonzero(mass))
    _, _, nvalid = ctx.particle_stats()
    del mass
    ctx.close()

    ms_per_step = sec * 1e3 / args.steps
    value = n * args.steps / sec
    # dominant kernel by accumulated time
    kern = max(("p2g", "grid", "g2p", "g2p2g"), key=lambda k: phases[k][0])
    # the g2p2g phase includes the mover side paths: time its kernel alone
    ktot, kcnt = phases["g2p2g_kernel" if kern == "g2p2g" else kern]
    avg_s = ktot / max(kcnt, 1) * 1e-3
    kbytes = BYTES_PER_PARTICLE[kern] * n + BYTES_PER_NODE[kern] * nodes
    achieved = kbytes / avg_s / 1e9
    step_bytes = BYTES_PER_PARTICLE["step"] * n + BYTES_PER_NODE["step"] * nodes
    step_gbs = step_bytes / (sec / args.steps) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.isfile(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        if kern in tr:
            traffic = tr[kern]["bytes_per_launch"]

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "12-box 3D APIC MPM, 256^3 grid, 8 ppc, fixed-corotated, sticky floor, dt 2e-5",
                   "particles": n, "valid_particles_after": nvalid, "active_nodes": nodes, "grid": [CELLS] * 3,
                   "box_seed": BOX_SEED, "seed": SEED, "parallelism": "single GPU",
                   "reorder_interval": args.reorder_interval, "rebin_fraction": args.rebin_fraction,
                   "rebin_interval": args.rebin_interval, "rebins_timed": rebins1 - rebins0,
                   "movers_last_step": movers_last},
        "roofline": {"bound": "hbm", "kernel": kern, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": kbytes, "avg_launch_ms": avg_s * 1e3},
        "step_roofline": {"achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                          "bytes_per_step": step_bytes},
        "phase_ms_per_step": {k: v[0] / args.steps for k, v in phases.items()},
        "phase_launches": {k: v[1] for k, v in phases.items()},
        "adjoint_ratio": None,
        "cpu_baseline": None,
    }
    if not args.no_batched:
        try:
            out["batched"] = batched(log)
        except Exception as e:  # report, never hide the GPU number
            log(f"batched run failed: {e}")
    if not args.no_adjoint:
        try:
            out["adjoint_ratio"] = adjoint_ratio(log)
        except Exception as e:  # report, never hide the GPU number
            log(f"adjoint ratio failed: {e}")
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(shapes, log)
        except Exception as e:  # report, never hide the GPU number
            log(f"cpu baseline failed: {e}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
