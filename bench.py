"""Throughput of the XPBD tissue substep on one MI355X: tet-substeps/s.

`python bench.py [--steps K] [--warmup W]` runs the kidney-scale single instance (an ellipsoid of ~250k Kuhn tets,
20 substeps x 2 iterations per step, grasper contact, a Bezier blade cutting once per step) for W untimed and K
timed steps on cuda:0, and prints one JSON line. `value` is live tets x substeps / second over the timed steps, with
all state resident on the GPU. The dominant kernel (the tet colour pass) is timed per launch with HIP events on its
stream over the same steps; `roofline.achieved` is its algorithmic bytes per tet update (the per-tet-substep figure
(52 N + 8 T + 4 E)/T) over its measured duration. `cpu_baseline` is the CPU oracle of the same step, timed on this
host's cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_tet_substep(active_nodes, tets, edges):
    """Compulsory HBM traffic of one substep per tet: node position + velocity read and write in FP32 and the inverse
    mass (52 B/node), per-tet compliance and liveness (8 B/tet), edge liveness (4 B/edge)."""
    return (52.0 * active_nodes + 8.0 * tets + 4.0 * edges) / tets


def load_traffic(name):
    path = ROOT / "profiles" / name
    if path.exists():
        try:
            return json.loads(path.read_text())
        except ValueError:
            return None
    return None


def run_gpu(w, steps, warmup):
    import ctypes as ct

    import torch

    from genesis_tissue_simulation_amd._lib import check
    from genesis_tissue_simulation_amd.tissue import Tissue, swept_cut_tool

    kd, kh = w.stiffness()
    t = Tissue(w.grid, w.params, kd, kh, w.initial_positions)
    t.set_gravity(w.gravity)

    def one(step):
        tool = w.tool(step) if w.tool else None
        t.step(w.dt, tool=tool, tool_armed=tool is not None)
        if w.blade:
            t.detect_cut(swept_cut_tool(w.blade(step), w.blade(step + 1), w.blade_thickness))

    for s in range(warmup):
        one(s)
    torch.cuda.synchronize()
    live0 = int(t.liveness()[0].sum())
    t0 = time.perf_counter()
    for s in range(warmup, warmup + steps):
        one(s)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    live1 = int(t.liveness()[0].sum())
    # Per-launch timing of the tet colour kernel with HIP events on its stream. Event records cannot sit inside the
    # captured step graph, so these steps issue the same kernels directly; the kernels and their arguments are the
    # same as in the graph.
    lib = t.lib
    check("gts_profile_begin", lib.gts_profile_begin())
    for s in range(warmup + steps, warmup + steps + max(1, min(steps, 3))):
        one(s)
    total_ms, launches, units = ct.c_double(), ct.c_longlong(), ct.c_longlong()
    check("gts_profile_end", lib.gts_profile_end(ct.byref(total_ms), ct.byref(launches), ct.byref(units)))
    torch.cuda.synchronize()
    pos = t.positions()
    if not np.isfinite(pos).all():
        raise RuntimeError("non-finite positions: simulation stopped")
    return dict(elapsed=elapsed, live=(live0 + live1) / 2, tet_ms=total_ms.value, launches=launches.value,
                units=units.value, E=t.E, T=t.T, active=int((w.grid.active != 0).sum()))


def run_batched(steps, warmup, count=1024):
    """The batched workload: 1024 blocks of 2k-100k tets, each cut once, stepped in one persistent launch per step
    with the launch shape picked by the autotuner; the step kernel is timed with events on its stream."""
    import torch

    from genesis_tissue_simulation_amd import autotune
    from genesis_tissue_simulation_amd.batched import Batch
    from genesis_tissue_simulation_amd.workloads import batched_blocks

    t0 = time.perf_counter()
    w = batched_blocks(count)
    b = Batch(w.instances, w.params, gravity=w.gravity)
    for k, cut in enumerate(w.cuts):
        b.cut(k, cut)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    best, table = autotune.tune(b, w.dt)
    for _ in range(warmup):
        b.step(w.dt)
    torch.cuda.synchronize()
    live = b.live_tets()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    for _ in range(steps):
        b.step(w.dt)
    stop.record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kernel_ms = start.elapsed_time(stop)
    pos = b.pos[:, :3]
    if not bool(torch.isfinite(pos).all()):
        raise RuntimeError("non-finite batched positions")
    active = sum(int((inst[0].active != 0).sum()) for inst in w.instances)
    return dict(elapsed=elapsed, kernel_ms=kernel_ms, live=live, T=b.total_tets, E=int(b.edge_off[-1]),
                active=active, substeps=w.params.substeps, best=best, table=table, setup_s=setup, count=count,
                workload=w)


def algorithmic_bytes_per_adjoint_tet_substep(nodes, tets, edges):
    """Compulsory HBM traffic of one substep of the rollout and its reverse per tet: the forward's 52 B/node,
    8 B/tet, 4 B/edge, plus the reverse's position and velocity adjoints read and written (48 B/node), per-tet
    compliance and the two alpha adjoints read and written (20 B/tet) and edge liveness (4 B/edge)."""
    return (100.0 * nodes + 28.0 * tets + 8.0 * edges) / tets


def run_adjoint(iters, warmup, count=64, tets=20_000, nsub=50, ckpt=10):
    """The compliance fit: `count` blocks of about `tets` tets, a rollout of `nsub` substeps under the seeded point
    loads, gradient of the surface-node L2 loss by the adjoint kernels with checkpoints every `ckpt` substeps.
    One iteration is a forward rollout plus its backward through torch.autograd; the adjoint kernels are timed with
    events on their stream."""
    import torch

    from genesis_tissue_simulation_amd.adjoint import ComplianceFit
    from genesis_tissue_simulation_amd.workloads import adjoint_instance

    t0 = time.perf_counter()
    insts = [adjoint_instance(tets=tets, seed=k) for k in range(count)]
    fit = ComplianceFit(insts, nsub=nsub, ckpt=ckpt)
    hidden = torch.as_tensor(np.concatenate([i.compliance_hidden for i in insts]).astype(np.float32))
    c0 = torch.as_tensor(np.concatenate([i.compliance_init for i in insts]).astype(np.float32))
    target = fit.forward_raw(hidden).clone()
    c = c0.clone().cuda().requires_grad_(True)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0

    def one():
        c.grad = None
        loss = fit.loss(fit.rollout(c), target)
        loss.backward()
        return loss

    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    for _ in range(iters):
        one()
    stop.record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # block 0 at the stated compliance and at a x1e5 softer one, for the comparison with the FP64 oracle
    n0, t1 = len(insts[0].grid.positions), int(fit.batch.tet_off[1])
    checks = {}
    for scale in (1.0, 1e5):
        tgt = fit.forward_raw(hidden * scale).clone()
        x_hidden = tgt[:n0].cpu().numpy()
        x = fit.forward_raw(c0 * scale).clone()
        g = fit.backward_raw((x - tgt) * (2 * fit.mask[:, None]))
        checks[scale] = dict(x_hidden=x_hidden, target=tgt[:n0].cpu().numpy(), grad=g[:t1].cpu().numpy())
    return dict(elapsed=elapsed, kernel_ms=start.elapsed_time(stop), T=fit.sT, N=fit.sN, E=fit.sE, count=count,
                nsub=nsub, ckpt=ckpt, setup_s=setup, inst0=insts[0], hidden0=hidden[:t1].numpy(),
                init0=c0[:t1].numpy(), checks=checks,
                launch=dict(block_size=fit.batch.block_size, n_workgroups=fit.batch.n_workgroups))


def adjoint_fp64_checks(a):
    """Block 0 against the FP64 oracle: relative end-position error of the FP32 rollout at the hidden compliance,
    and the FP32 adjoint gradient against the FP64 adjoint gradient (itself pinned by finite differences), at the
    stated compliance and at a x1e5 softer one."""
    import ctypes as ct

    from oracle import oracle as orc

    inst = a["inst0"]
    L = orc.lib()
    f = L.oracle_adjoint_f64
    f.argtypes = [ct.c_void_p, ct.c_float, ct.c_int] + [ct.c_void_p] * 5
    f.restype = ct.c_double
    ptr = lambda arr: np.ascontiguousarray(arr, dtype=np.float32).ctypes.data_as(ct.c_void_p)
    N = len(inst.grid.positions)
    out = {}
    for scale, chk in a["checks"].items():
        o = orc.Oracle(inst.grid, inst.params, inst.initial_positions, double=True)
        f(o.h, np.float32(inst.dt), a["nsub"], ptr(inst.force), ptr(a["hidden0"] * scale), ptr(inst.mask),
          ptr(np.zeros((N, 3))), None)
        x64 = o.state()[0]
        err = float(np.abs(chk["x_hidden"] - x64).max() / np.abs(x64).max())
        o = orc.Oracle(inst.grid, inst.params, inst.initial_positions, double=True)
        g64 = np.zeros(len(a["init0"]))
        f(o.h, np.float32(inst.dt), a["nsub"], ptr(inst.force), ptr(a["init0"] * scale), ptr(inst.mask),
          ptr(chk["target"]), g64.ctypes.data_as(ct.c_void_p))
        g32 = chk["grad"].astype(np.float64)
        fin = np.isfinite(g32)
        entry = {"rel_err_positions": err, "grad_finite_fraction": float(fin.mean())}
        if fin.all():
            entry["grad_cosine"] = float(g32 @ g64 / (np.linalg.norm(g32) * np.linalg.norm(g64)))
            entry["grad_rel_l2"] = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
        out["stated" if scale == 1.0 else f"soft_x{scale:g}"] = entry
    return out


def run_cpu_batched(w, threads, sample=8):
    from oracle import oracle as orc

    used = orc.set_threads(threads)
    sims = []
    for grid, p, kd, kh, init in w.instances[:sample]:
        o = orc.Oracle(grid, p, init)
        o.set_tet_stiffness(kd, kh)
        o.set_gravity(w.gravity)
        sims.append(o)
    t0 = time.perf_counter()
    for o in sims:
        o.step(w.dt)
    elapsed = time.perf_counter() - t0
    tets = sum(o.T for o in sims)
    return dict(value=tets * w.params.substeps / elapsed, threads=used, tets=tets)


def run_cpu(w, steps, threads):
    from genesis_tissue_simulation_amd.tissue import swept_cut_tool
    from oracle import oracle as orc

    orc.build()
    used = orc.set_threads(threads)
    kd, kh = w.stiffness()
    o = orc.Oracle(w.grid, w.params, w.initial_positions)
    o.set_tet_stiffness(kd, kh)
    o.set_gravity(w.gravity)
    t0 = time.perf_counter()
    for s in range(steps):
        if w.tool:
            o.set_tool(w.tool(s), tool_armed=True, reset=(s == 0))
        o.step(w.dt)
        if w.blade:
            o.detect_cut(swept_cut_tool(w.blade(s), w.blade(s + 1), w.blade_thickness))
    elapsed = time.perf_counter() - t0
    live = int(o.liveness()[0].sum())
    return dict(elapsed=elapsed, live=live, threads=used)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--batched-count", type=int, default=1024)
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--adjoint-iters", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the tissue path runs on a single MI355X and has no multi-GPU mode "
              "(instances shard by LPT packing across GPUs by design only)", file=sys.stderr)
        sys.exit(2)

    from genesis_tissue_simulation_amd.workloads import kidney_250k

    w = kidney_250k()
    g = run_gpu(w, args.steps, args.warmup)
    substeps = w.params.substeps
    value = g["live"] * substeps * args.steps / g["elapsed"]
    b_alg = algorithmic_bytes_per_tet_substep(g["active"], g["T"], g["E"])
    avg_launch_ms = g["tet_ms"] / max(g["launches"], 1)
    achieved = b_alg * g["units"] / (g["tet_ms"] * 1e-3) / 1e9 if g["tet_ms"] > 0 else 0.0
    traffic = load_traffic("tet_colour_traffic_kidney.json")
    line = {
        "metric": "tet-substeps/s plain vs batched vs adjoint @ <=1e-4 rel err vs FP64",
        "value": value,
        "unit": "tet-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * g["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "kidney-250k plain single instance", "tets": g["T"], "live_tets": g["live"],
                   "active_nodes": g["active"], "substeps": substeps, "iterations": w.params.iterations,
                   "dt": w.dt, "contact": "capsule grasper", "cut": "swept-quad blade, once per step"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic.get("bytes_per_launch") if traffic else None,
                     "kernel": "k_tets (one colour of the tet pass)", "avg_launch_us": 1e3 * avg_launch_ms,
                     "launches": g["launches"], "bytes_per_tet_update": b_alg},
    }
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    if not args.skip_batched:
        bt = run_batched(args.batched_steps, 2, args.batched_count)
        bv = bt["live"] * bt["substeps"] * args.batched_steps / bt["elapsed"]
        bb = algorithmic_bytes_per_tet_substep(bt["active"], bt["T"], bt["E"])
        bach = bb * bt["live"] * bt["substeps"] * args.batched_steps / (bt["kernel_ms"] * 1e-3) / 1e9
        line["batched"] = {
            "value": bv, "unit": "tet-substeps/s", "vs_plain": bv / value,
            "ms_per_step": 1e3 * bt["elapsed"] / args.batched_steps,
            "config": {"workload": f"{bt['count']} blocks, tets log-uniform 2k-100k, one straight cut each",
                       "tets": bt["T"], "live_tets": bt["live"], "substeps": bt["substeps"], "iterations": 1,
                       "launch": {k: bt["best"][k] for k in ("block_size", "workgroups_per_cu", "lds_nodes",
                                                               "n_workgroups")}},
            "roofline": {"bound": "hbm", "achieved": bach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": bach / HBM_PEAK_GBS, "kernel": "k_batched_step (one launch per step)",
                         "avg_launch_us": 1e3 * bt["kernel_ms"] / args.batched_steps, "bytes_per_tet_substep": bb,
                         "traffic": None},
            "autotune": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()} for r in bt["table"]],
        }
        if not args.no_cpu:
            cb = run_cpu_batched(bt["workload"], threads)
            line["batched"]["cpu_baseline"] = {"value": cb["value"], "unit": "tet-substeps/s", "cores": cb["threads"],
                                               "kind": "port", "sample": f"1 step of the first 8 instances "
                                                                         f"({cb['tets']} tets)"}
    if not args.skip_adjoint:
        a = run_adjoint(args.adjoint_iters, 1)
        av = a["T"] * a["nsub"] * args.adjoint_iters / a["elapsed"]
        ab = algorithmic_bytes_per_adjoint_tet_substep(a["N"], a["T"], a["E"])
        aach = ab * a["T"] * a["nsub"] * args.adjoint_iters / (a["kernel_ms"] * 1e-3) / 1e9
        line["adjoint"] = {
            "value": av, "unit": "tet-substeps/s (rollout + gradient)", "vs_plain": av / value,
            "ms_per_iteration": 1e3 * a["elapsed"] / args.adjoint_iters,
            "config": {"workload": f"{a['count']} blocks of ~20k tets, compliance fit, {a['nsub']} substeps, "
                                   f"checkpoint every {a['ckpt']}", "tets": a["T"], "substeps": a["nsub"],
                       "checkpoint": a["ckpt"], "launch": a["launch"]},
            "rel_err_vs_fp64": None,
            "roofline": {"bound": "hbm", "achieved": aach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": aach / HBM_PEAK_GBS,
                         "kernel": "k_adjoint_forward + k_adjoint_backward (one launch each per iteration)",
                         "avg_iteration_us": 1e3 * a["kernel_ms"] / args.adjoint_iters,
                         "bytes_per_tet_substep": ab, "traffic": None},
        }
        checks = adjoint_fp64_checks(a)
        line["adjoint"]["rel_err_vs_fp64"] = checks["stated"]["rel_err_positions"]
        line["adjoint"]["fp64_checks"] = checks
    if not args.no_cpu:
        c = run_cpu(w, args.cpu_steps, threads)
        line["cpu_baseline"] = {
            "value": c["live"] * substeps * args.cpu_steps / c["elapsed"], "unit": "tet-substeps/s",
            "cores": c["threads"], "kind": "port",
            "sample": f"{args.cpu_steps} steps of the same kidney-250k workload (CPU oracle, FP32, OpenMP over the "
                      f"members of a colour)"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
