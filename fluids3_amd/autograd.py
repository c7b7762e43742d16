"""torch.autograd.Function over the C ABI's reverse mode (for checks).

`rollout_loss(v0, visc, k, system, scene, nsteps, every, target)` restarts
`system` on `scene` with viscosity `visc` and internal stiffness `k`, sets the
initial velocity and evaluation velocity of every particle to v0 (n x 3, by
particle id), runs nsteps steps with checkpoints every `every` steps and
returns L = (1/n) sum |x_T - x*|^2 as a float64 scalar tensor.  Its backward
calls fx_adjoint_backward: dL/dv0 = dL/dvel0 + dL/dveval0, dL/dvisc, dL/dk.
Torch is only the carrier of the values here; all work runs in libfluidsx.
"""
import copy

import numpy as np
import torch

from . import abi
from .rollout import target_positions  # noqa: F401  (re-exported for the checks)


class RolloutLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, visc, k, system, scene, nsteps, every, target):
        sc = copy.copy(scene)
        sc.visc = float(visc)
        sc.int_stiff = float(k)
        system.Restart(sc)
        v = v0.detach().to("cpu", torch.float32).numpy()
        system.Commit(abi.FVEL, v)
        system.Commit(abi.FVEVAL, v)
        system.AdjointConfigure(int(nsteps), int(every))
        system.AdjointSetTarget(np.asarray(target, np.float32))
        loss, _ = system.AdjointForward()
        ctx.system = system
        ctx.v0_device = v0.device
        return torch.tensor(float(loss[0]), dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        g = ctx.system.AdjointBackward()
        s = float(gout)
        gv0 = torch.from_numpy((g["gvel"].astype(np.float64) + g["gveval"]) * s).to(torch.float32)
        return (gv0.to(ctx.v0_device), torch.tensor(g["gvisc"][0] * s, dtype=torch.float64),
                torch.tensor(g["gk"][0] * s, dtype=torch.float64), None, None, None, None, None)


def rollout_loss(v0, visc, k, system, scene, nsteps, every, target):
    return RolloutLoss.apply(v0, visc, k, system, scene, nsteps, every, target)

