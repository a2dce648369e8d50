"""torch.autograd.Function over the GPU step and its adjoint (test harness).

MPMRollout.apply(V0, mu, lam, sim) runs `sim.steps` steps of dt from the
scene's initial state with initial velocity V0 and Lame parameters mu, lam
(per particle), recording a checkpoint every `sim.checkpoint` steps, and
returns the final positions; its backward calls mpmx_backward with dL/dX and
returns dL/dV0, dL/dmu, dL/dlam.
"""
import numpy as np
import torch

from physbam_amd import mpmx


class Rollout:
    def __init__(self, sc, steps, checkpoint):
        self.sc, self.steps, self.checkpoint = sc, steps, checkpoint
        d = sc["dim"]
        self.ctx = mpmx.Context(d, sc["N"], sc["lo"], sc["hi"], precision=sc["prec"])
        self.ctx.set_constitutive(sc["model"], sc.get("neo_threshold", 0.25))
        self.ctx.set_gravity(sc["gravity"])
        for w in sc["walls"]:
            self.ctx.add_collision_box(w["type"], w["friction"], w["lo"][:d], w["hi"][:d])

    def close(self):
        self.ctx.close()


class MPMRollout(torch.autograd.Function):
    @staticmethod
    def forward(fctx, V0, mu, lam, sim):
        sc = sim.sc
        sim.ctx.set_particles(sc["X"], V0.detach().cpu().numpy(), sc["F"], sc["B"], sc["mass"], sc["volume"],
                              mu.detach().cpu().numpy(), lam.detach().cpu().numpy(), sc["valid"])
        # record a tape only when a gradient is wanted: a forward under
        # torch.no_grad() (a finite-difference probe) leaves no open tape behind
        if any(fctx.needs_input_grad):
            sim.ctx.tape_begin(sim.checkpoint)
        sim.ctx.step(sc["dt"], sim.steps)
        fctx.sim = sim
        X = sim.ctx.get_particles()["X"]
        return torch.tensor(np.asarray(X, np.float64))

    @staticmethod
    def backward(fctx, gX):
        g = fctx.sim.ctx.backward(dX=gX.detach().cpu().numpy())
        t = lambda a: torch.tensor(np.asarray(a, np.float64))
        return t(g["V"]), t(g["mu"]), t(g["lambda"]), None
