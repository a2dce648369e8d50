"""A batch of independent scenes with different grids in one context
(mpmx_create_batch): every scene must evolve exactly as it does alone, which
the oracle checks scene by scene (FP64 to 1e-8 after ten steps; FP32 against
the FP64 oracle to 1e-4)."""
import numpy as np
import pytest

import mpmtest as mt
from physbam_amd import mpmx

pytestmark = pytest.mark.gpu


def scenes(dim, prec, walls):
    out = []
    Ns = [12, 20, 16, 9] if dim == 3 else [24, 40, 32, 17]
    for s, N in enumerate(Ns):
        c = np.array([0.5, 0.45 + 0.05 * s, 0.5])[:dim]
        r = 0.18 + 0.04 * s
        shape = lambda X, c=c, r=r: ((X - c) ** 2).sum(1) < r ** 2
        w = (mt.wall_boxes(dim, [N] * dim, [0] * dim, [1] * dim, flags=(1 << (2 * dim)) - 1, inset=0.2,
                           type_=mt.SLIP, friction=0.0) if walls else [])
        out.append(mt.make_scene(dim, N, prec=prec, model=mt.COROTATED, shape=shape,
                                 v0=[0.5 - 0.3 * s, -2.0, 0.2 * s][:dim], vrand=1.0, brand=1e-3, frand=0.1,
                                 walls=w, dt=1e-3, per_particle_E=True, seed=100 + s))
    return out


@pytest.mark.parametrize("dim,prec,walls", [(3, 8, True), (2, 8, False), (3, 4, True)])
def test_batch_matches_each_scene_alone(dim, prec, walls):
    scs = scenes(dim, prec, walls)
    steps = 10
    counts = [sc["X"].shape[0] for sc in scs]
    print("batch", dim, prec, "particles", counts, flush=True)
    cat = lambda k: np.concatenate([sc[k] for sc in scs])
    ctx = mpmx.Context.batch(dim, [(sc["N"], sc["lo"], sc["hi"]) for sc in scs], precision=prec)
    try:
        assert mpmx.lib().mpmx_scene_count(ctx._h) == len(scs)
        ctx.set_constitutive(mt.COROTATED)
        ctx.set_gravity(scs[0]["gravity"])
        for w in scs[0]["walls"]:
            ctx.add_collision_box(w["type"], w["friction"], w["lo"][:dim], w["hi"][:dim])
        ctx.set_reorder_interval(3)
        ctx.set_rebin_policy(1.0 / 128, 4)
        ctx.set_batch_particles(counts, cat("X"), cat("V"), cat("F"), cat("B"), cat("mass"), cat("volume"),
                                cat("mu"), cat("lambda"), cat("valid"))
        ctx.step(scs[0]["dt"], steps)
        gpu = ctx.get_particles()
        with pytest.raises(mpmx.MpmxError):
            ctx.get_grid()
    finally:
        ctx.close()
    off = 0
    for s, sc in enumerate(scs):
        n = counts[s]
        orc = mt.oracle_run(sc, steps, prec=8, want_grid=False)
        part = {k: gpu[k][off:off + n] for k in ("X", "V", "F", "B", "valid")}
        assert np.array_equal(part["valid"], orc["valid"]), s
        for k in ("X", "V", "F", "B"):
            err = np.abs(part[k] - orc[k]).max() / max(np.abs(orc[k]).max(), 1e-30)
            print("scene", s, k, f"{err:.2e}", flush=True)
            assert err <= (1e-8 if prec == 8 else 1e-4), (s, k, err)
        off += n


@pytest.mark.parametrize("graph", [False, True])
def test_batch_with_empty_scenes(graph):
    """Scenes without particles inside a batch (first, middle and last) leave
    the others as they are alone (FP64, ten steps, to 1e-8 against the
    oracle), on the eager and the graph path."""
    scs = scenes(3, 8, True)[:2]
    empty = dict(scs[0], N=10)
    grids = [(10, empty["lo"], empty["hi"]), (scs[0]["N"], scs[0]["lo"], scs[0]["hi"]), (10, empty["lo"], empty["hi"]),
             (scs[1]["N"], scs[1]["lo"], scs[1]["hi"]), (10, empty["lo"], empty["hi"])]
    n0, n1 = scs[0]["X"].shape[0], scs[1]["X"].shape[0]
    counts = [0, n0, 0, n1, 0]
    cat = lambda k: np.concatenate([sc[k] for sc in scs])
    ctx = mpmx.Context.batch(3, grids, precision=8)
    try:
        ctx.set_constitutive(mt.COROTATED)
        ctx.set_gravity(scs[0]["gravity"])
        for w in scs[0]["walls"]:
            ctx.add_collision_box(w["type"], w["friction"], w["lo"], w["hi"])
        ctx.set_launch_config(graph=graph)
        ctx.set_batch_particles(counts, cat("X"), cat("V"), cat("F"), cat("B"), cat("mass"), cat("volume"),
                                cat("mu"), cat("lambda"), cat("valid"))
        ctx.step(scs[0]["dt"], 10)
        gpu = ctx.get_particles()
    finally:
        ctx.close()
    off = 0
    for s, sc in enumerate(scs):
        n = sc["X"].shape[0]
        orc = mt.oracle_run(sc, 10, prec=8, want_grid=False)
        assert np.array_equal(gpu["valid"][off:off + n], orc["valid"]), s
        for k in ("X", "V", "F", "B"):
            err = np.abs(gpu[k][off:off + n] - orc[k]).max() / max(np.abs(orc[k]).max(), 1e-30)
            print("empty-scene batch", graph, s, k, f"{err:.2e}", flush=True)
            assert err <= 1e-8, (s, k, err)
        off += n


@pytest.fixture(scope="module")
def batched_workload():
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench

    print("seeding the 256 batched scenes", flush=True)
    sizes, parts = bench.batched_scenes()
    return sizes, parts


@pytest.mark.timeout(900)
@pytest.mark.parametrize("graph", [False, True])
def test_batched_workload_at_full_shape(batched_workload, graph):
    """The batched benchmark workload itself: 256 scenes of 32^3..96^3 cells,
    one sphere each (the generator the benchmark times), fixed corotated,
    gravity, FP32, dt 1e-4, in one context.  After 3 steps (12 with graph:
    four one by one, then two hipGraph replays of four fused steps) the
    smallest and the largest scene of each grid size match their own FP64
    oracle runs to 1e-4 (X, V, F; B by the velocity its error carries into the
    next P2G; valid flags equal); after 10 steps (16) no particle is lost and
    each scene's momentum is M g t (no walls: the internal forces and the
    APIC transfers add no momentum), to 1e-5 of M |g| t."""
    sizes, parts = batched_workload
    counts = [int(P["X"].shape[0]) for P in parts]
    n = sum(counts)
    print(f"{len(parts)} scenes, {n} particles ({min(counts)}..{max(counts)})", flush=True)
    cat = lambda k: np.concatenate([P[k] for P in parts])
    dt, g = 1e-4, np.array([0.0, -9.8, 0.0])
    ctx = mpmx.Context.batch(3, [(N, [0.0] * 3, [1.0] * 3) for N in sizes], precision=mpmx.FP32)
    with ctx:
        ctx.set_constitutive(mpmx.FIXED_COROTATED)
        ctx.set_gravity(list(g))
        ctx.set_batch_particles(counts, cat("X"), cat("V"), cat("F"), cat("B"), cat("mass"), cat("volume"),
                                cat("mu"), cat("lam"), cat("valid"))
        ctx.set_launch_config(graph=graph)
        if graph:  # four steps one by one, then two graph replays of four
            ctx.step(dt, 4)
            ctx.step(dt, 8)
            ncmp = 12
            assert ctx.phase_time("graph")[1] == 2
        else:
            ctx.step(dt, 3)
            ncmp = 3
        at_cmp = ctx.get_particles()
        nend = ncmp + (4 if graph else 7)
        ctx.step(dt, nend - ncmp)
        at_end = ctx.get_particles()
    offs = np.concatenate([[0], np.cumsum(counts)])
    picks = set()
    for N in (32, 48, 64, 96):
        ids = [s for s in range(len(sizes)) if sizes[s] == N]
        assert ids, N
        picks.add(min(ids, key=lambda s: counts[s]))
        picks.add(max(ids, key=lambda s: counts[s]))
    assert len(picks) == 8
    for s in sorted(picks):
        P = parts[s]
        sc = dict(dim=3, prec=4, model=mpmx.FIXED_COROTATED, N=[sizes[s]] * 3, lo=[0.0] * 3, hi=[1.0] * 3,
                  gravity=list(g), dt=dt, neo_threshold=0.25, walls=[], X=P["X"], V=P["V"], F=P["F"], B=P["B"],
                  mass=P["mass"], volume=P["volume"], mu=P["mu"], valid=P["valid"], **{"lambda": P["lam"]})
        orc = mt.oracle_run(sc, ncmp, prec=8, want_grid=False)
        part = {k: at_cmp[k][offs[s]:offs[s + 1]] for k in ("X", "V", "F", "B", "valid")}
        assert np.array_equal(part["valid"], orc["valid"]), s
        errs = {k: float(np.abs(part[k] - orc[k]).max() / max(np.abs(orc[k]).max(), 1e-30))
                for k in ("X", "V", "F")}
        # the spheres start at rest and fall freely, so their exact B is ~0:
        # B is measured by the velocity its error carries into the next P2G
        errs["B"] = mt.b_velocity_error(part["B"], orc["B"], orc["V"], 1.0 / sizes[s])
        print(f"scene {s} (N {sizes[s]}, {counts[s]} particles):", {k: f"{v:.1e}" for k, v in errs.items()},
              flush=True)
        for k, v in errs.items():
            assert v <= 1e-4, (s, k, v)
    assert np.all(at_end["valid"] == 1)
    t = nend * dt
    for s in range(len(parts)):
        m = parts[s]["mass"].astype(np.float64)
        V = at_end["V"][offs[s]:offs[s + 1]].astype(np.float64)
        M = m.sum()
        err = np.abs((m[:, None] * V).sum(0) - M * g * t).max() / (M * 9.8 * t)
        assert err <= 1e-5, (s, err)
