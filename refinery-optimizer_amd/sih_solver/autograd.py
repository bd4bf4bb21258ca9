"""Differentiable batched LP layer: torch.autograd.Function over the GPU engine.

Forward: the enhanced PDHG (FP32 iterates, FP64 checks and refinement) solves
every LP of the batch; the output is the concatenated primal solutions.
Backward: the adjoint kernels (mpdhg_adjoint_batch) return dL/dc and dL/db,
where b shifts each row's bounds (rl + b, ru + b; equality rows rl = ru = b
move as one).  Tensors may live on any device: they cross to host memory at
the C boundary, which takes plain host pointers.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np
import torch

from . import GpuPdhgConfig, Problem, SolutionStatus, adjoint_batch, solve_batch


def _shifted(p: Problem, c: np.ndarray, b: np.ndarray) -> Problem:
    q = Problem(p.name)
    q._A = p.A
    q.c = c
    q.col_lower, q.col_upper = p.col_lower, p.col_upper
    q.row_lower = p.row_lower + b
    q.row_upper = p.row_upper + b
    return q


class LpBatchSolve(torch.autograd.Function):
    """x* = argmin_x c^T x s.t. rl + b <= A x <= ru + b, l <= x <= u, per LP."""

    @staticmethod
    def forward(ctx, c_all: torch.Tensor, b_all: torch.Tensor, problems: Sequence[Problem],
                config: GpuPdhgConfig, active_tol: float = 1e-7):
        ns = [p.num_cols() for p in problems]
        ms = [p.num_rows() for p in problems]
        c_np = c_all.detach().to("cpu", torch.float64).numpy()
        b_np = b_all.detach().to("cpu", torch.float64).numpy()
        co = np.concatenate([[0], np.cumsum(ns)])
        ro = np.concatenate([[0], np.cumsum(ms)])
        shifted = [_shifted(p, c_np[co[k]:co[k + 1]], b_np[ro[k]:ro[k + 1]])
                   for k, p in enumerate(problems)]
        sols = solve_batch(shifted, config)
        bad = [k for k, s in enumerate(sols) if s.status != SolutionStatus.Optimal]
        if bad:
            raise RuntimeError(f"LP layer: {len(bad)} LPs did not reach the tolerance (first: {bad[0]})")
        ctx.problems = shifted
        ctx.sols = sols
        ctx.offsets = (co, ro)
        ctx.active_tol = active_tol
        ctx.out_dev = (c_all.device, c_all.dtype)
        x = np.concatenate([s.x for s in sols])
        return torch.from_numpy(x).to(device=c_all.device, dtype=c_all.dtype)

    @staticmethod
    def backward(ctx, grad_x: torch.Tensor):
        co, ro = ctx.offsets
        w = grad_x.detach().to("cpu", torch.float64).numpy()
        res = adjoint_batch(ctx.problems, [s.x for s in ctx.sols], [s.row_duals for s in ctx.sols],
                            [w[co[k]:co[k + 1]] for k in range(len(ctx.problems))],
                            active_tol=ctx.active_tol)
        dev, dt = ctx.out_dev
        gc = torch.from_numpy(np.concatenate([r.dl_dc for r in res])).to(device=dev, dtype=dt)
        gb = torch.from_numpy(np.concatenate([r.dl_db for r in res])).to(device=dev, dtype=dt)
        ctx.last_adjoint = res
        return gc, gb, None, None, None


def lp_layer(problems: Sequence[Problem], c_all: torch.Tensor, b_all: torch.Tensor,
             config: GpuPdhgConfig = None, active_tol: float = 1e-7) -> torch.Tensor:
    """Differentiable solve of a batch of LPs (see LpBatchSolve)."""
    if config is None:
        config = GpuPdhgConfig.enhanced(tol=1e-8)
    return LpBatchSolve.apply(c_all, b_all, list(problems), config, active_tol)
