"""torch.autograd.Function over the discrete-adjoint entry point.

`PulseObjective.apply(env, spec)` returns J[p] = sum_k W_k |c_{k,p}(T)|^2
for a batch of piecewise-linear pulse envelopes env (P x K samples, any
device); backward returns dJ/d env from the same run (the forward call
performs the adjoint sweep, since it reuses the forward's checkpoints).
The HIP engine does all the work; torch only carries the tensors.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .engine import Propagator, Pulse


@dataclass
class ObjectiveSpec:
    prop: Propagator          # propagator with P pulses on the system
    x0: np.ndarray            # initial state (2N doubles), same for all
    w: float                  # carrier frequency (a.u.)
    T: float                  # pulse length = PWL support (a.u.)
    W: np.ndarray             # objective weights (N)
    gauge: str = "v"
    h0: float = 0.05
    tol: float = 1e-10
    ckpt_every: int = 32
    max_steps: int = 100000


class PulseObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, env: torch.Tensor, spec: ObjectiveSpec):
        e = env.detach().to("cpu", torch.float64).numpy()
        P = spec.prop.P
        if e.shape[0] != P:
            raise ValueError("env must hold one envelope per pulse")
        for p in range(P):
            spec.prop.set_state(spec.x0, p)
            spec.prop.set_pulse(Pulse.pwl(e[p], spec.w, spec.T, spec.gauge), p)
        J, grad, _, _, st = spec.prop.adjoint(
            0.0, spec.T, spec.h0, spec.W, atol=spec.tol, rtol=spec.tol,
            max_steps=spec.max_steps, ckpt_every=spec.ckpt_every)
        ctx.save_for_backward(torch.from_numpy(grad[:, :e.shape[1]]).to(env))
        ctx.stats = st
        return torch.from_numpy(J).to(env)

    @staticmethod
    def backward(ctx, gJ: torch.Tensor):
        (grad,) = ctx.saved_tensors
        return gJ[:, None] * grad, None
