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
