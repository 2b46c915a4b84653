"""Autotuner of the batched persistent launch: threads per workgroup, LDS budget per workgroup (which sets how many
workgroups share a CU and which instances are staged in LDS) and workgroup count, measured on the actual batch.

Each candidate re-plans the LPT packing for its workgroup count, runs a few untimed steps (which also capture its
step graph), then times steps between device synchronisations. The state the tuning steps advance is restored
afterwards, so tuning does not change the simulation.
"""
from __future__ import annotations

import time

import torch


def candidates(max_lds_bytes):
    out = []
    for wpc in (1, 2, 4):
        lds_nodes = max_lds_bytes // 16 // wpc
        for block in (256, 512, 1024):
            if block * wpc > 2048:
                continue
            out.append(dict(block_size=block, workgroups_per_cu=wpc, lds_nodes=lds_nodes))
    return out


def tune(batch, dt, steps=3, warmup=1, space=None):
    """Time every candidate launch shape on `batch` and leave the fastest configured. Returns (best, table)."""
    saved = {name: getattr(batch, name).clone() for name in ("pos", "prev", "vel", "lambda_edge", "lambda_tet")}
    space = space or candidates(int(batch.lib.gts_batched_lds_bytes()))
    table = []
    for cand in space:
        batch.configure(cand["block_size"], cand["lds_nodes"], cand["workgroups_per_cu"])
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
    batch.configure(best["block_size"], best["lds_nodes"], best["workgroups_per_cu"])
    torch.cuda.synchronize()
    return best, table
