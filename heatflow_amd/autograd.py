"""torch.autograd.Function over the hand-written adjoint of the HeatFlow time step.

    T_final = heatflow_run(kappa, power, sim, nsteps=200, ckpt_every=20)
    loss = ((T_final[probes] - targets) ** 2).sum()
    loss.backward()          # kappa.grad, power.grad from the reverse-time adjoint

kappa: per-cell conductivity (W/(m K)), power: per-cell heater power P (W) entering as a power
density P / volume (the superset per-cell input of the C-ABI), both float64 tensors of length N
(on the GPU or the host).  The forward runs `nsteps` time steps from T^0 = T^{-1} = t_init on
the library's device context, in the context's precision mode (FP64, or FP32-inner / FP64-refined)
and tolerance, and keeps (T^n, T^{n-1}) checkpoints every `ckpt_every` steps; backward re-runs
each segment from its checkpoint and sweeps the adjoint backwards (hf_forward_checkpointed /
hf_adjoint in include/heatflow_hip.h).  ckpt_every = 1 keeps every state instead (nothing is
re-run); 0 picks 1 when the states fit in free device memory, else ceil(sqrt(nsteps)).

The checkpoints live in the context, so one `Simulation` holds one pending graph: backward of a
graph whose forward was followed by another forward on the same context raises RuntimeError
(and the library itself refuses after the fields, the operator or the solver changed).
"""
from __future__ import annotations

import numpy as np
import torch

from .evolution import Simulation


class HeatFlowRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kappa: torch.Tensor, power: torch.Tensor, sim: Simulation, itime0: int,
                nsteps: int, ckpt_every: int, t_init: float):
        vol = torch.as_tensor(sim.system.cells.volume, dtype=torch.float64, device=power.device)
        qd = power.detach().double() / vol
        sim.set_fields(kappa.detach(), qd)
        sim.sparse_Hmatrix()
        sim.initial_evolve(t_init)
        sim.forward_checkpointed(itime0, nsteps, ckpt_every)
        ctx.generation = sim.forward_generation
        out = torch.empty(sim.n, dtype=torch.float64, device=kappa.device)
        if out.is_cuda:
            torch.cuda.synchronize(out.device)   # the library writes on its own stream
        sim.state_into(out)
        ctx.sim = sim
        ctx.dev = kappa.device
        return out

    @staticmethod
    def backward(ctx, g_t: torch.Tensor):
        sim = ctx.sim
        if sim.forward_generation != ctx.generation:
            raise RuntimeError("heatflow_run: another forward ran on this Simulation since this "
                               "graph's forward; its checkpoints are gone (one pending graph "
                               "per Simulation)")
        gk = torch.empty(sim.n, dtype=torch.float64, device=ctx.dev)
        gp = torch.empty(sim.n, dtype=torch.float64, device=ctx.dev)
        if gk.is_cuda:
            torch.cuda.synchronize(gk.device)
        sim.adjoint(g_t.detach().double().contiguous(), gk, gp)
        return gk, gp, None, None, None, None, None


def heatflow_run(kappa: torch.Tensor, power: torch.Tensor, sim: Simulation, nsteps: int,
                 ckpt_every: int = 20, itime0: int = 1, t_init: float | None = None):
    """Final temperature field after `nsteps` steps, differentiable in kappa and power."""
    if t_init is None:
        t_init = sim.system.params.t_system
    return HeatFlowRun.apply(kappa, power, sim, itime0, nsteps, ckpt_every, t_init)
