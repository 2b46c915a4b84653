"""Autotuners. The batched persistent launch: threads per workgroup, LDS budget per workgroup (which sets how many
workgroups share a CU and how much of each instance is staged in LDS), workgroup count and substeps (colours) per
launch, measured on the actual batch. The single-instance step: threads per workgroup of the colour kernels,
measured on a scratch copy of the tissue.

Each candidate re-plans the LPT packing for its workgroup count, runs a few untimed steps (which also capture its
step graph), then times steps between device synchronisations. The state the tuning steps advance is restored
afterwards, so tuning does not change the simulation.
"""
from __future__ import annotations

import time

import torch


def candidates(max_lds_bytes, substeps=10):
    """Threads per workgroup x workgroups per CU (which splits the LDS) with the whole step in one launch, and the
    shapes with one workgroup per CU at 512 and 1024 threads with 1, 2 and half the substeps per launch (the colours
    per launch are that many substeps' colours)."""
    out = []
    for wpc in (1, 2, 4):
        lds_nodes = max_lds_bytes // 16 // wpc
        for block in (256, 512, 1024):
            if block * wpc > 2048:
                continue
            out.append(dict(block_size=block, workgroups_per_cu=wpc, lds_nodes=lds_nodes, substeps_per_launch=0))
    per_launch = sorted({k for k in (1, 2, max(1, substeps // 2)) if k < substeps})
    for block in (512, 1024):
        for k in per_launch:
            out.append(dict(block_size=block, workgroups_per_cu=1, lds_nodes=max_lds_bytes // 16,
                            substeps_per_launch=k))
    return out


def tune(batch, dt, steps=5, warmup=1, space=None):
    """Time every candidate launch shape on `batch` and leave the fastest configured. Returns (best, table)."""
    # everything a step advances, the particle frames included (each step's frame update blends the previous ones)
    saved = {name: getattr(batch, name).clone()
             for name in ("pos", "prev", "vel", "lambda_edge", "lambda_tet", "rot", "rot_prev")}
    space = space or candidates(int(batch.lib.gts_batched_lds_bytes()), int(batch.params.substeps))
    table = []
    for cand in space:
        batch.configure(cand["block_size"], cand["lds_nodes"], cand["workgroups_per_cu"],
                        substeps_per_launch=cand["substeps_per_launch"])
        for _ in range(warmup):
            batch.step(dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.step(dt)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        table.append(dict(cand, n_workgroups=batch.n_workgroups, ms_per_step=ms))
        for name, t in saved.items():
            getattr(batch, name).copy_(t)
    best = min(table, key=lambda r: r["ms_per_step"])
    batch.configure(best["block_size"], best["lds_nodes"], best["workgroups_per_cu"],
                    substeps_per_launch=best["substeps_per_launch"])
    torch.cuda.synchronize()
    return best, table


def tune_plain(make_tissue, step, steps=3, warmup=2, blocks=(32, 64, 128, 256)):
    """Colour-kernel workgroup size of the single-instance step: each candidate runs `warmup` untimed and `steps`
    timed steps of `step(tissue, k)` on a fresh tissue from `make_tissue(colour_block)`. Returns (best, table)."""
    table = []
    for cb in blocks:
        t = make_tissue(cb)
        for k in range(warmup):
            step(t, k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(warmup, warmup + steps):
            step(t, k)
        torch.cuda.synchronize()
        table.append(dict(colour_block=cb, ms_per_step=1e3 * (time.perf_counter() - t0) / steps))
        del t
    return min(table, key=lambda r: r["ms_per_step"]), table
