"""Helpers for differentiable rollouts over the C ABI (no torch needed)."""
import copy

import numpy as np

from . import abi


def target_positions(system, scene, nsteps, visc=0.30, k=1.2):
    """Final positions (n x 3 float32, by particle id) of an nsteps rollout of
    `scene` at viscosity visc and internal stiffness k: the fitting target of
    the differentiable dam break (a rollout at other parameters)."""
    sc = copy.copy(scene)
    sc.visc = float(visc)
    sc.int_stiff = float(k)
    system.Restart(sc)
    system.Run(int(nsteps))
    pos = system.Retrieve(abi.FPOS)
    ids = system.Retrieve(abi.FID).astype(np.int64)
    out = np.empty_like(pos)
    out[ids] = pos
    return out
