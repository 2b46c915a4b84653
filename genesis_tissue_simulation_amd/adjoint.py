"""Per-tet compliance gradients of an XPBD rollout: the hand-written adjoint kernels behind torch.autograd.

The rollout is `nsub` substeps of the undamped substep (`params.iterations` Gauss-Seidel iterations) with coupled
Neo-Hookean tets (deviatoric compliance c per tet, volume compliance c/10) and lattice edges, driven by a per-node
force field, with no contact and no cutting: gradients are not taken across cut events. Many blocks run together,
one workgroup per block. The forward kernel keeps (x, v) every `ckpt` substeps; the backward re-runs each segment
from its checkpoint with a tape of every projection's inputs and reverses it substep by substep, iteration by
iteration, colour by colour. The re-run of one segment runs on a side stream while the next-later segment is
reversed, so two sets of workgroups are busy at once.

There is no CPU path: the library must be built and a GPU present.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np
import torch

from ._lib import GtsAdjointView, check, load
from .batched import Batch
from .params import TET_NEOHOOKEAN
from .tissue import current_stream, step_params


class ComplianceFit:
    """Blocks given as `AdjointInstance`s (grid, params, dt, initial positions, force, mask); all share dt, substeps
    and the rollout length. `rollout(c)` maps per-tet compliance (generation order, blocks back to back) to the node
    positions after `nsub` substeps, differentiably."""

    def __init__(self, instances, nsub=50, ckpt=10, device="cuda:0", block_size=1024, workgroups_per_cu=1,
                 lds_nodes=None):
        if nsub <= 0 or ckpt <= 0:
            raise ValueError("nsub and ckpt must be positive")
        p0 = instances[0].params
        for inst in instances:
            p = inst.params
            if (p.tet_mode != TET_NEOHOOKEAN or p.iterations < 1 or p.iterations != p0.iterations or
                    p.substeps != p0.substeps or inst.dt != instances[0].dt or p.strain_damping or p.bulk_damping or
                    p.beta_s or p.damping or p.rayleigh or p.omega != 1.0 or not p.edge_ks > 0):
                raise ValueError("the adjoint covers the undamped Neo-Hookean substep with edges; all blocks must "
                                 "share dt, substeps and iterations")
        self.lib = load()
        self.device = torch.device(device)
        self.nsub, self.ckpt = int(nsub), int(ckpt)
        self.n_ckpt = (self.nsub + self.ckpt - 1) // self.ckpt
        self.params = p0
        self.dt = float(instances[0].dt)
        self.batch = Batch([(i.grid, i.params, None, None, i.initial_positions) for i in instances], p0,
                           gravity=(0.0, 0.0, 0.0), device=device, block_size=block_size, lds_nodes=lds_nodes,
                           workgroups_per_cu=workgroups_per_cu, use_graph=False)
        b = self.batch
        dev = self.device
        sN, sE, sT = int(b.node_off[-1]), int(b.edge_off[-1]), int(b.tet_off[-1])
        self.sN, self.sE, self.sT = sN, sE, sT
        self.pos0 = b.pos.clone()
        force = np.zeros((sN, 4), dtype=np.float32)
        mask = np.zeros(sN, dtype=np.float32)
        perm = np.zeros(sT, dtype=np.int64)
        for k, inst in enumerate(instances):
            n0, n1 = int(b.node_off[k]), int(b.node_off[k + 1])
            force[n0:n1, :3] = inst.force
            mask[n0:n1] = inst.mask
            t0 = int(b.tet_off[k])
            gen = b.tissues[k].topology["tet_gen"]
            perm[t0:t0 + len(gen)] = t0 + gen.astype(np.int64)
        self.force = torch.as_tensor(force).to(dev)
        self.mask = torch.as_tensor(mask).to(dev)
        self.perm = torch.as_tensor(perm).to(dev)      # colour-major slot -> generation index
        f4 = lambda n: torch.zeros((max(n, 1), 4), dtype=torch.float32, device=dev)
        f1 = lambda n: torch.zeros(max(n, 1), dtype=torch.float32, device=dev)
        slots = 2 * self.ckpt * p0.iterations  # two tape halves: one re-taped while the other is reversed
        self.compliance = f1(sT)
        self.alpha = f1(2 * sT)
        self.checkpoints = f4(2 * self.n_ckpt * sN)
        self.tet_tape = f4(slots * 4 * sT)
        self.edge_tape = f4(slots * 2 * sE)
        self.lambda_tape = f1(slots * (2 * sT + sE) if p0.iterations > 1 else 0)
        self.xbar, self.vbar, self.prevbar = f4(sN), f4(sN), f4(sN)
        self.abar = f1(2 * sT)
        self.lambar = f1(2 * sT + sE)
        self.grad = f1(sT)
        self.version = 0
        self._view()

    def _view(self):
        self.view = GtsAdjointView(
            self.batch.view, self.sN, self.sT, self.sE, self.force.data_ptr(), self.compliance.data_ptr(),
            self.alpha.data_ptr(), self.checkpoints.data_ptr(), self.tet_tape.data_ptr(), self.edge_tape.data_ptr(),
            self.lambda_tape.data_ptr(), self.xbar.data_ptr(), self.vbar.data_ptr(), self.prevbar.data_ptr(),
            self.abar.data_ptr(), self.lambar.data_ptr(), self.grad.data_ptr())

    def configure(self, block_size=1024, workgroups_per_cu=1, lds_nodes=None):
        """Launch shape: threads per workgroup (the backward uses at most 512), workgroups per CU and the LDS node
        capacity: the forward stages a block's positions in LDS when it has at most lds_nodes nodes, the backward its
        positions and position adjoints when it has at most lds_nodes / 2."""
        self.batch.configure(block_size, lds_nodes, workgroups_per_cu)
        self._view()

    @property
    def n_tets(self):
        return self.sT

    def _step_params(self):
        return step_params(self.params, self.dt)

    def forward_raw(self, compliance):
        """Rollout from the initial state with per-tet compliance (generation order); returns end positions
        [sum N, 3] (a view of the pooled state)."""
        c = torch.as_tensor(compliance, dtype=torch.float32, device=self.device).reshape(-1)
        if c.numel() != self.sT:
            raise ValueError(f"compliance has {c.numel()} entries, the blocks have {self.sT} tets")
        if not bool((c > 0).all()):
            raise ValueError("compliance must be positive")
        self.compliance[:self.sT].copy_(c[self.perm])
        self.batch.pos.copy_(self.pos0)
        self.batch.vel.zero_()
        p = self._step_params()
        check("gts_adjoint_forward", self.lib.gts_adjoint_forward(ct.byref(self.view), ct.byref(p), self.nsub,
                                                                  self.ckpt, current_stream()))
        self.version += 1
        return self.batch.pos[:, :3]

    def backward_raw(self, grad_positions, grad_velocities=None):
        """Gradient of a scalar with respect to the compliance of the last forward_raw, given its gradients with
        respect to the end positions (and velocities) [sum N, 3]; returned in generation order. The backward reuses
        the forward's checkpoints and per-tet alphas, so it must follow the forward it differentiates."""
        self.xbar.zero_()
        self.vbar.zero_()
        self.xbar[:self.sN, :3].copy_(grad_positions)
        if grad_velocities is not None:
            self.vbar[:self.sN, :3].copy_(grad_velocities)
        p = self._step_params()
        check("gts_adjoint_backward", self.lib.gts_adjoint_backward(ct.byref(self.view), ct.byref(p), self.nsub,
                                                                    self.ckpt, current_stream()))
        self.version += 1
        out = torch.empty(self.sT, dtype=torch.float32, device=self.device)
        out[self.perm] = self.grad[:self.sT]
        return out

    def rollout(self, compliance):
        """Differentiable: end positions [sum N, 3] after nsub substeps as a function of per-tet compliance."""
        return _Rollout.apply(compliance, self)

    def loss(self, positions, target):
        """L2 error on the masked (surface) nodes."""
        d = positions - target
        return (self.mask[:, None] * d * d).sum()


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, compliance, fit):
        x = fit.forward_raw(compliance.detach()).clone()
        ctx.fit = fit
        ctx.version = fit.version
        return x

    @staticmethod
    def backward(ctx, grad_x):
        fit = ctx.fit
        if fit.version != ctx.version:
            raise RuntimeError("the rollout's checkpoints were overwritten by a later forward: backward each rollout "
                               "before running the next")
        return fit.backward_raw(grad_x.contiguous()), None
