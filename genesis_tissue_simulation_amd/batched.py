"""Many independent tissue instances of uneven size, stepped together in one persistent launch per step.

Pooled device buffers hold every instance back to back; `GtsBatchInstance` records index them. Each instance also
has an ordinary single-instance view (`Tissue`) over its slice of the pooled buffers, used for what happens between
steps: cutting, liveness refresh, compaction and readback.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np
import torch

from . import _lib
from ._lib import GtsBatchInstance, GtsBatchView, check, load
from .params import SolverParams
from .tissue import Tissue, build_topology, current_stream, step_params


def device_cus():
    return int(torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count)


class Batch:
    """Instances given as (grid, params, k_dev, k_hyd, initial_positions); all share `params.substeps`,
    `iterations`, `tet_mode` and damping; each carries its own edge stiffness (`params.edge_ks` > 0)."""

    def __init__(self, instances, params: SolverParams, gravity=(0.0, -981.0, 0.0), device="cuda:0",
                 block_size=512, lds_nodes=None, workgroups_per_cu=1, use_graph=True):
        self.lib = load()
        self.device = torch.device(device)
        self.params = params
        self.gravity = tuple(float(g) for g in gravity)
        dev = self.device
        topos = [build_topology(inst[0], bends=False) for inst in instances]  # the batched step has no bending
        Ns = [len(inst[0].positions) for inst in instances]
        Es = [len(t["edge_desc"]) for t in topos]
        Ts = [len(t["tet_desc"]) for t in topos]
        self.node_off = np.concatenate(([0], np.cumsum(Ns)))
        self.edge_off = np.concatenate(([0], np.cumsum(Es)))
        self.tet_off = np.concatenate(([0], np.cumsum(Ts)))
        sN, sE, sT = int(self.node_off[-1]), int(self.edge_off[-1]), int(self.tet_off[-1])
        self.pos = torch.zeros((sN, 4), dtype=torch.float32, device=dev)
        self.prev = torch.zeros((sN, 4), dtype=torch.float32, device=dev)
        self.vel = torch.zeros((sN, 4), dtype=torch.float32, device=dev)
        self.mass = torch.zeros(sN, dtype=torch.float32, device=dev)
        self.edge_desc = torch.zeros(max(sE, 1), dtype=torch.int32, device=dev)
        self.tet_desc = torch.zeros(max(sT, 1), dtype=torch.int32, device=dev)
        self.k_dev = torch.zeros(max(sT, 1), dtype=torch.float32, device=dev)
        self.k_hyd = torch.zeros(max(sT, 1), dtype=torch.float32, device=dev)
        self.lambda_edge = torch.zeros(max(sE, 1), dtype=torch.float32, device=dev)
        self.lambda_tet = torch.zeros(max(2 * sT, 1), dtype=torch.float32, device=dev)
        self.alpha = torch.zeros(max(2 * sT, 1), dtype=torch.float32, device=dev)
        # links, flags and particle frames of every instance, for the frame update that ends each step
        self.nbr = torch.zeros(6 * sN, dtype=torch.int32, device=dev)
        self.active = torch.zeros(sN, dtype=torch.int32, device=dev)
        self.pinned = torch.zeros(sN, dtype=torch.int32, device=dev)
        self.rot = torch.zeros(9 * sN, dtype=torch.float32, device=dev)
        self.rot_prev = torch.zeros(9 * sN, dtype=torch.float32, device=dev)
        self.node_base = torch.as_tensor(np.repeat(self.node_off[:-1], Ns).astype(np.int32)).to(dev)
        self.tissues = []
        records = (GtsBatchInstance * len(instances))()
        colour_rows = []
        coff = 0
        for k, ((grid, p_k, kd, kh, init), topo) in enumerate(zip(instances, topos)):
            n0, n1 = int(self.node_off[k]), int(self.node_off[k + 1])
            e0, e1 = int(self.edge_off[k]), int(self.edge_off[k + 1])
            t0, t1 = int(self.tet_off[k]), int(self.tet_off[k + 1])
            storage = dict(pos=self.pos[n0:n1], prev=self.prev[n0:n1], vel=self.vel[n0:n1], mass=self.mass[n0:n1],
                           edge_desc=self.edge_desc[e0:e1], edge_lambda=self.lambda_edge[e0:e1],
                           tet_desc=self.tet_desc[t0:t1], tet_k_dev=self.k_dev[t0:t1],
                           tet_k_hyd=self.k_hyd[t0:t1], tet_lambda=self.lambda_tet[2 * t0:2 * t1],
                           nbr=self.nbr[6 * n0:6 * n1], active=self.active[n0:n1], pinned=self.pinned[n0:n1],
                           rot=self.rot[9 * n0:9 * n1], rot_prev=self.rot_prev[9 * n0:9 * n1])
            self.tissues.append(Tissue(grid, p_k, kd, kh, init, device=device, topology=topo, storage=storage,
                                       use_graph=False))
            r = records[k]
            r.node_off, r.n_nodes, r.edge_off, r.n_edges, r.tet_off, r.n_tets = n0, n1 - n0, e0, e1 - e0, t0, t1 - t0
            nx, ny = grid.dims[0] + 1, grid.dims[1] + 1
            r.nx, r.nxy = nx, nx * ny
            r.n_edge_colours = len(topo["edge_offsets"]) - 1
            r.n_tet_colours = len(topo["tet_offsets"]) - 1
            r.colour_off = coff
            r.edge_ks = float(p_k.edge_ks)
            r.rest_len = (ct.c_float * 3)(*[float(x) for x in topo["axis_rest_length"]])
            r.inv_rest = (ct.c_float * 54)(*topo["type_inv_rest"].astype(np.float32).reshape(-1))
            row = np.concatenate((topo["edge_offsets"], topo["tet_offsets"])).astype(np.int32)
            colour_rows.append(row)
            coff += len(row)
        self.records = records
        raw = np.frombuffer(bytes(records), dtype=np.uint8)
        self.instances = torch.as_tensor(raw.copy()).to(dev)
        self.colour_table = torch.as_tensor(np.concatenate(colour_rows)).to(dev)
        self.costs = np.array([T + 0.25 * E + 0.5 * N for N, E, T in zip(Ns, Es, Ts)], dtype=np.float64)
        self.Ns, self.Es, self.Ts = Ns, Es, Ts
        self.use_graph = use_graph
        self.configure(block_size, lds_nodes, workgroups_per_cu)

    def configure(self, block_size=512, lds_nodes=None, workgroups_per_cu=1, n_workgroups=None,
                  substeps_per_launch=0):
        """Launch shape: threads per workgroup, LDS node capacity (a larger instance stages that many of its nodes and
        keeps the rest in global memory), substeps per persistent launch (0: the whole step) and
        workgroup count, with the instances LPT-packed onto the workgroups."""
        max_lds = int(self.lib.gts_batched_lds_bytes())
        if lds_nodes is None:
            lds_nodes = max_lds // 16 // max(1, int(workgroups_per_cu))
        if lds_nodes * 16 > max_lds:
            raise ValueError(f"lds_nodes {lds_nodes} exceeds the {max_lds} B of LDS a workgroup may use")
        if n_workgroups is None:
            n_workgroups = min(len(self.tissues), device_cus() * int(workgroups_per_cu))
        n_workgroups = max(1, min(int(n_workgroups), len(self.tissues)))
        plan = np.zeros(n_workgroups + 1 + len(self.tissues), dtype=np.int32)
        check("gts_batched_plan", self.lib.gts_batched_plan(self.costs.ctypes.data_as(ct.c_void_p),
                                                            len(self.tissues), n_workgroups,
                                                            plan.ctypes.data_as(ct.c_void_p)))
        self.work_offsets = torch.as_tensor(plan[:n_workgroups + 1].copy()).to(self.device)
        self.work_items = torch.as_tensor(plan[n_workgroups + 1:].copy()).to(self.device)
        self.block_size, self.lds_nodes, self.n_workgroups = int(block_size), int(lds_nodes), n_workgroups
        self.substeps_per_launch = int(substeps_per_launch)
        self.view = GtsBatchView(
            len(self.tissues), self.instances.data_ptr(), self.colour_table.data_ptr(), self.work_offsets.data_ptr(),
            self.work_items.data_ptr(), n_workgroups, self.block_size, self.lds_nodes, int(bool(self.use_graph)),
            self.substeps_per_launch, (ct.c_float * 3)(*self.gravity), self.pos.data_ptr(), self.prev.data_ptr(), self.vel.data_ptr(),
            self.mass.data_ptr(), self.edge_desc.data_ptr(), self.tet_desc.data_ptr(), self.k_dev.data_ptr(),
            self.k_hyd.data_ptr(), self.lambda_edge.data_ptr(), self.lambda_tet.data_ptr(), self.alpha.data_ptr(),
            int(self.node_off[-1]), self.nbr.data_ptr(), self.active.data_ptr(), self.pinned.data_ptr(),
            self.node_base.data_ptr(), self.rot.data_ptr(), self.rot_prev.data_ptr())
        return self

    def step(self, dt, params=None):
        p = step_params(params or self.params, dt)
        check("gts_batched_step", self.lib.gts_batched_step(ct.byref(self.view), ct.byref(p), current_stream()))

    def cut(self, k, cut_tool):
        self.tissues[k].detect_cut(cut_tool)

    def positions(self, k):
        return self.tissues[k].positions()

    def velocities(self, k):
        return self.tissues[k].velocities()

    def live_tets(self):
        return int((self.tet_desc[:int(self.tet_off[-1])] & 1).sum().item())

    @property
    def total_tets(self):
        return int(self.tet_off[-1])
