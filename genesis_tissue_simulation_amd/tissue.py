"""One tissue instance on the GPU: device buffers as PyTorch-ROCm tensors, computation in libgts_xpbd.so.

`Tissue` is the host-side owner of everything the reference keeps in its physics / xpbd / cut modules for one
session (`physics.cu:214-315`, `xpbd/xpbd_solver.cu:1412-1848`, `cut.cu:1105-1206`), but without process-global
state: several instances can live side by side.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np
import torch

from . import _lib
from ._lib import (GtsBatchInstance, GtsBatchView, GtsCutTool, GtsLattice, GtsStepParams, GtsTissueView, GtsToolState,
                   GtsTopology, check, load, v3)
from .grid import CUT_POINT_DTYPE
from .params import SolverParams

TOOL_FIELDS3 = ("shaftA", "shaftB", "jawHinge", "jawAxis", "jawMid", "jawNormal", "bladeA", "bladeB")
TOOL_FIELDS1 = ("shaftR", "jawHalfWidth", "jawLen", "jawGap", "grasp01", "bladeR")


def tool_state(tool):
    """GtsToolState from a dict of XpbdToolState fields (missing ones are zero)."""
    t = GtsToolState()
    tool = tool or {}
    for name in TOOL_FIELDS3:
        setattr(t, name, v3(tool.get(name, (0, 0, 0))))
    for name in TOOL_FIELDS1:
        setattr(t, name, float(tool.get(name, 0.0)))
    return t


def swept_cut_tool(prev_segment, segment, thickness):
    """Swept plane of a straight blade moving from `prev_segment` (Sprev, Eprev) to `segment` (S, E), as the cut
    descriptor CutToolDesc: T1 = (Sprev, Eprev, E), T2 = (Sprev, E, S), unit normal and AABB (`cut.h:25-38`)."""
    sp, ep = (np.asarray(p, dtype=np.float32) for p in prev_segment)
    s, e = (np.asarray(p, dtype=np.float32) for p in segment)
    n = np.cross(ep - sp, s - sp).astype(np.float32)
    length = float(np.linalg.norm(n))
    valid = length > 1e-8
    n = (n / length).astype(np.float32) if valid else np.zeros(3, dtype=np.float32)
    pts = np.stack((sp, ep, e, s))
    return dict(t1v0=sp, t1v1=ep, t1v2=e, t2v0=sp, t2v1=e, t2v2=s, ncut=n, d=float(thickness), valid=int(valid),
                aabbMin=pts.min(axis=0), aabbMax=pts.max(axis=0))


def cut_tool_struct(tool):
    t = GtsCutTool()
    for name in ("t1v0", "t1v1", "t1v2", "t2v0", "t2v1", "t2v2", "ncut", "aabbMin", "aabbMax"):
        setattr(t, name, v3(tool[name]))
    t.d = float(tool["d"])
    t.valid = int(tool["valid"])
    return t


def step_params(params: SolverParams, dt):
    p = params
    return GtsStepParams(float(dt), int(p.substeps), int(p.iterations), int(p.tet_mode), float(p.omega),
                         float(p.edge_ks), float(p.young), float(p.beta_s), float(p.strain_damping),
                         float(p.bulk_damping), float(p.rayleigh), float(p.damping), int(bool(p.contact_enabled)),
                         float(p.mu_s), float(p.mu_d), float(p.jaw_thickness), float(p.contact_radius),
                         float(p.jaw_gap_closed), float(p.contact_compliance), int(bool(p.precise_strain)),
                         float(p.guard_min_volume), int(p.guard_passes), float(p.tensile_strength),
                         float(p.bend_kb), float(p.beta_b))


def build_topology(grid, bends=True):
    """Host topology of a lattice via gts_build_topology; arrays in device order plus per-type rest data. With
    `bends`, also the bend elements and their colouring (the bending constraints need them)."""
    lib = load()
    nx, ny, nz = (d + 1 for d in grid.dims)
    N = nx * ny * nz
    V = int(np.prod(grid.dims))
    rest = np.ascontiguousarray(grid.positions, dtype=np.float32)
    active = np.ascontiguousarray(grid.active, dtype=np.int32)
    nbr = np.ascontiguousarray(grid.neighbors, dtype=np.int32)
    rest_nbr = np.ascontiguousarray(grid.rest_neighbors, dtype=np.float32)
    lat = GtsLattice(*grid.dims, rest.ctypes.data, active.ctypes.data, nbr.ctypes.data, rest_nbr.ctypes.data)
    edge_desc = np.zeros(3 * N, dtype=np.uint32)
    edge_gen = np.zeros(3 * N, dtype=np.int32)
    eoff = np.zeros(_lib.GTS_MAX_COLOURS + 1, dtype=np.int32)
    tet_desc = np.zeros(6 * V, dtype=np.uint32)
    tet_gen = np.zeros(6 * V, dtype=np.int32)
    toff = np.zeros(_lib.GTS_MAX_COLOURS + 1, dtype=np.int32)
    topo = GtsTopology(edge_desc.ctypes.data, edge_gen.ctypes.data, eoff.ctypes.data, tet_desc.ctypes.data,
                       tet_gen.ctypes.data, toff.ctypes.data)
    if bends:
        bend_desc = np.zeros(15 * N, dtype=np.uint32)
        bend_gen = np.zeros(15 * N, dtype=np.int32)
        boff = np.zeros(_lib.GTS_MAX_COLOURS + 1, dtype=np.int32)
        topo.bend_desc, topo.bend_gen_index = bend_desc.ctypes.data, bend_gen.ctypes.data
        topo.bend_colour_offsets = boff.ctypes.data
    check("gts_build_topology", lib.gts_build_topology(ct.byref(lat), ct.byref(topo)))
    E, T, B = topo.n_edges, topo.n_tets, topo.n_bends
    extra = {}
    if bends:
        extra = dict(bend_desc=bend_desc[:B].copy(), bend_gen=bend_gen[:B].copy(),
                     bend_offsets=boff[:topo.n_bend_colours + 1].copy())
    return dict(**extra, edge_desc=edge_desc[:E].copy(), edge_gen=edge_gen[:E].copy(),
                edge_offsets=eoff[:topo.n_edge_colours + 1].copy(), tet_desc=tet_desc[:T].copy(),
                tet_gen=tet_gen[:T].copy(), tet_offsets=toff[:topo.n_tet_colours + 1].copy(),
                type_inv_rest=np.array(topo.type_inv_rest, dtype=np.float64).reshape(6, 3, 3),
                type_volume=np.array(topo.type_volume, dtype=np.float64),
                axis_rest_length=np.array(topo.axis_rest_length, dtype=np.float32))


def _ptr(t):
    return ct.c_void_p(t.data_ptr())


def current_stream():
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


class Tissue:
    """Device state of one tissue and the HIP entry points that advance and cut it."""

    def __init__(self, grid, params: SolverParams, k_dev=None, k_hyd=None, initial_positions=None,
                 device="cuda:0", topology=None, use_graph=True, max_substeps=64, storage=None, colour_block=0,
                 xcd_order=True, persistent=False, tiles=None):
        """`storage` optionally supplies pre-allocated device tensors (pos, prev, vel, mass, active, pinned, nbr,
        edge_desc, edge_lambda, tet_desc, tet_k_dev, tet_k_hyd, tet_lambda, rot, rot_prev) -- views into the pooled
        buffers of a batch. `colour_block` is
        the workgroup size of the edge / tet colour kernels (0: the library default); `xcd_order` gives each XCD one
        contiguous range of every colour (same results). `persistent`: a small tissue (every node in one workgroup's
        LDS, at most PERSISTENT_COLOUR_TETS tets per colour) steps without tool contact, guard, bending, edge damping or
        cut tool as one persistent workgroup running the whole step (the batched kernel over this tissue alone):
        colours are separated by workgroup barriers instead of launches, with the same results. `tiles`: (region,
        kmax[, edges]) runs each tet sweep (and with edges the edge sweep) as tiled launches of up to kmax colours
        over regions of region^3 cells (see set_tiles); None keeps one launch per colour."""
        self.lib = load()
        if not torch.cuda.is_available():
            raise _lib.NativeError("no GPU visible: the HIP tissue path has no CPU fallback")
        self.grid, self.params, self.device = grid, params, torch.device(device)
        topo = topology or build_topology(grid)
        self.topology = topo
        N = len(grid.positions)
        V = int(np.prod(grid.dims))
        E, T = len(topo["edge_desc"]), len(topo["tet_desc"])
        Eg = len(grid.edges)
        self.N, self.E, self.T, self.Eg, self.V = N, E, T, Eg, V
        # current tet slot -> generation index and current edge slot -> original colour-order slot (compaction
        # removes dead entries; the readbacks report over the original lists)
        self.T0, self.E0 = T, E
        dev = self.device
        self.tet_slot_gen = torch.as_tensor(np.asarray(topo["tet_gen"], dtype=np.int64)).to(dev)
        self.edge_slot_orig = torch.arange(E, dtype=torch.int64, device=dev)
        self.pooled = storage is not None

        def f32(a, shape=None):
            x = torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32))
            return (x.reshape(shape) if shape else x).to(dev)

        def i32(a):
            return torch.as_tensor(np.ascontiguousarray(np.asarray(a).astype(np.int64).astype(np.int32)
                                                        if np.asarray(a).dtype != np.uint32 else
                                                        np.asarray(a).view(np.int32))).to(dev)

        init = grid.positions if initial_positions is None else initial_positions
        pos = np.zeros((N, 4), dtype=np.float32)
        pos[:, :3] = init
        st = storage or {}
        gen = topo["tet_gen"]
        k_dev = np.ones(T, dtype=np.float32) if k_dev is None else np.asarray(k_dev, dtype=np.float32)
        k_hyd = np.ones(T, dtype=np.float32) if k_hyd is None else np.asarray(k_hyd, dtype=np.float32)

        def take(name, make):
            t = st.get(name)
            if t is None:
                return make()
            t.copy_(make())
            return t

        self.pos = take("pos", lambda: f32(pos))
        self.prev = take("prev", lambda: f32(pos))
        self.vel = take("vel", lambda: torch.zeros((N, 4), dtype=torch.float32, device=dev))
        self.force = torch.zeros((N, 4), dtype=torch.float32, device=dev)
        self.mass = take("mass", lambda: f32(grid.mass))
        self.active = take("active", lambda: i32(grid.active))
        self.pinned = take("pinned", lambda: i32(grid.pinned))
        self.inside = i32(grid.inside)
        self.nbr = take("nbr", lambda: i32(grid.neighbors.reshape(-1)))
        self.bend_alive = torch.zeros(N, dtype=torch.int32, device=dev)
        self.edge_desc = take("edge_desc", lambda: i32(topo["edge_desc"]))
        self.edge_lambda = take("edge_lambda", lambda: torch.zeros(max(E, 1), dtype=torch.float32, device=dev))
        self.slot_support = torch.zeros(6 * N, dtype=torch.int32, device=dev)
        self.tet_desc = take("tet_desc", lambda: i32(topo["tet_desc"]))
        self.tet_k_dev = take("tet_k_dev", lambda: f32(k_dev[gen]))
        self.tet_k_hyd = take("tet_k_hyd", lambda: f32(k_hyd[gen]))
        self.tet_alpha = torch.zeros(2 * max(T, 1), dtype=torch.float32, device=dev)
        self.tet_lambda = take("tet_lambda", lambda: torch.zeros(2 * max(T, 1), dtype=torch.float32, device=dev))
        self.grid_edges = i32(grid.edges.reshape(-1) if Eg else np.zeros(4, dtype=np.int32))
        self.occupied = i32(grid.occupied)
        self.cut_mask = torch.zeros(V, dtype=torch.int32, device=dev)
        self.edge_cut_t = torch.full((6 * N,), -1.0, dtype=torch.float32, device=dev)
        self.edge_skin_t = f32(grid.edge_skin_fraction())
        self.rest = f32(np.asarray(grid.positions, dtype=np.float32).reshape(-1))
        self.rot = take("rot", lambda: torch.zeros(9 * N, dtype=torch.float32, device=dev))
        self.rot_prev = take("rot_prev", lambda: torch.zeros(9 * N, dtype=torch.float32, device=dev))
        self.prev_corner = torch.zeros(3 * N, dtype=torch.float32, device=dev)
        self.cut_point_capacity = 4 * max(Eg, 1)
        self.cut_points = torch.zeros(8 * self.cut_point_capacity, dtype=torch.float32, device=dev)
        self.counters = torch.zeros(8, dtype=torch.int32, device=dev)
        scratch = int(self.lib.gts_scratch_floats(N, E, T, Eg))
        self.scratch = torch.zeros(scratch, dtype=torch.float32, device=dev)
        self.reaction = torch.zeros(4, dtype=torch.float32, device=dev)
        self.max_substeps = int(max_substeps)
        self.step_consts = torch.zeros(self.max_substeps * _lib.GTS_STEP_CONST_FLOATS, dtype=torch.float32,
                                       device=dev)
        self.edge_offsets = np.ascontiguousarray(topo["edge_offsets"], dtype=np.int32)
        self.tet_offsets = np.ascontiguousarray(topo["tet_offsets"], dtype=np.int32)
        inv_rest = topo["type_inv_rest"].astype(np.float32).reshape(-1)
        self.view = GtsTissueView(
            N, E, T, Eg, *grid.dims, len(self.edge_offsets) - 1, len(self.tet_offsets) - 1,
            self.edge_offsets.ctypes.data, self.tet_offsets.ctypes.data, (ct.c_float * 54)(*inv_rest),
            v3(topo["axis_rest_length"]), float(grid.cell_size[0]),
            *[t.data_ptr() for t in (self.pos, self.prev, self.vel, self.force, self.mass, self.active, self.pinned,
                                     self.inside, self.nbr, self.bend_alive, self.edge_desc, self.edge_lambda,
                                     self.slot_support, self.tet_desc, self.tet_k_dev, self.tet_k_hyd,
                                     self.tet_alpha, self.tet_lambda, self.grid_edges, self.occupied, self.cut_mask,
                                     self.edge_cut_t, self.edge_skin_t, self.rest, self.rot, self.rot_prev,
                                     self.prev_corner, self.cut_points)],
            self.cut_point_capacity, self.counters.data_ptr(), self.scratch.data_ptr(), self.reaction.data_ptr(),
            self.step_consts.data_ptr(), self.max_substeps, int(bool(use_graph)), int(colour_block))
        # the tear law's grasp ring [2N] (flags, dilation scratch), per cut edge exemption flags and decay ticks, and
        # the neighbour table as built (cut.cu:68-83, :1143-1147, :1190)
        self.grasp_ring = torch.zeros(2 * N, dtype=torch.int32, device=dev)
        self.edge_exempt = torch.zeros(2 * max(Eg, 1), dtype=torch.int32, device=dev)
        self.nbr_init = self.nbr.clone()
        v = self.view
        v.grasp_ring, v.edge_exempt = self.grasp_ring.data_ptr(), self.edge_exempt.data_ptr()
        v.nbr_init = self.nbr_init.data_ptr()
        v.xcd_order = int(bool(xcd_order))
        # bend elements (solved while params.bend_kb > 0)
        self.bend_offsets = np.ascontiguousarray(topo.get("bend_offsets", np.zeros(1)), dtype=np.int32)
        self.B = len(topo.get("bend_desc", ()))
        self.bend_desc = i32(topo["bend_desc"]) if self.B else torch.zeros(1, dtype=torch.int32, device=dev)
        self.bend_lambda = torch.zeros(max(self.B, 1), dtype=torch.float32, device=dev)
        v.n_bends, v.n_bend_colours = self.B, len(self.bend_offsets) - 1
        v.bend_colour_offsets = self.bend_offsets.ctypes.data
        v.bend_desc, v.bend_lambda = self.bend_desc.data_ptr(), self.bend_lambda.data_ptr()
        self._tool_from = None
        self.persistent = bool(persistent)
        self._solo = None             # the one-instance batched view of the persistent step, and what it was built for
        self._liveness_fresh = False  # liveness derived since initialisation (the persistent step does not derive it)
        self.gravity = (0.0, 0.0, 0.0)
        self._tiles = None            # the tiled sweep's plan and its device arrays
        self._staging = None          # pinned host buffers of observation()
        self.cut_tool = None          # LCS_SetTool's tool, arming the CCD ribbon tick of every substep
        self.emit_ncut = np.array([0.0, 1.0, 0.0], dtype=np.float32)
        self.tear_tau = 0.0
        self.contact_model = 1
        self.grasp_event = 0
        check("gts_init_state", self.lib.gts_init_state(ct.byref(self.view), current_stream()))
        if tiles is not None:
            self.set_tiles(*tiles)

    def set_tiles(self, region=8, kmax=3, edges=False):
        """Run each tet sweep of an undamped, unguarded step as tiled launches: up to `kmax` consecutive colours per
        launch, one workgroup per region of region^3 lattice cells (tiling.build_tile_plan), and with `edges` each
        undamped edge sweep without the fallback compliance as launches of up to min(3, kmax - 1) edge colours over the
        same regions; the results are the per-colour sweeps'. region=None turns it off. The plan holds for the current tet and edge
        lists: after a compaction the step takes one launch per colour again until set_tiles is called anew."""
        from . import tiling
        v = self.view
        if region is None:
            self._tiles = None
            v.tile_groups = 0
            return None
        plan = tiling.build_tile_plan(self.grid.dims, self.tet_desc[:self.T].cpu().numpy().view(np.uint32),
                                      self.tet_offsets, region=region, kmax=kmax,
                                      edge_desc=self.edge_desc[:self.E].cpu().numpy().view(np.uint32) if edges else None,
                                      edge_offsets=self.edge_offsets if edges else None)
        if plan is None:
            self._tiles = None
            v.tile_groups = 0
            return None
        dev = self.device
        n_entries = len(plan["entries"])
        keep = dict(plan=plan, groups=np.ascontiguousarray(plan["groups"].reshape(-1)),
                    entries=torch.as_tensor(plan["entries"].reshape(-1)).to(dev),
                    writes=torch.as_tensor(plan["writes"].reshape(-1)).to(dev),
                    pos_alt=self.pos.clone(),
                    inputs=torch.zeros(4 * n_entries, dtype=torch.float32, device=dev),
                    lam=torch.zeros(2 * n_entries, dtype=torch.float32, device=dev))
        self._tiles = keep
        v.tile_groups = len(plan["groups"])
        v.tile_kmax, v.tile_box, v.tile_tets, v.tile_edges = plan["kmax"], plan["box"], plan["tets"], plan["edges"]
        v.tile_group_info = keep["groups"].ctypes.data
        v.tile_entries = keep["entries"].data_ptr()
        v.tile_writes = keep["writes"].data_ptr()
        v.pos_alt = keep["pos_alt"].data_ptr()
        v.tile_inputs, v.tile_lambda = keep["inputs"].data_ptr(), keep["lam"].data_ptr()
        v.tile_region, v.tile_rx, v.tile_ry = plan["region"], plan["rx"], plan["ry"]
        keep["dense"] = np.ascontiguousarray(plan["dense_ids"], dtype=np.int32)
        v.tile_dense_ids = keep["dense"].ctypes.data
        return plan

    def set_contact_model(self, model, grasp=None):
        """xpbd_set_contact_model: 0 = legacy grasp (capture, slab ring, core pins, break check) with the CCD gate,
        1 = friction contact. `grasp` is a params.GraspParams; the grasp state starts cleared."""
        from .params import GraspParams
        if model not in (0, 1):
            raise ValueError("contact model must be 0 (legacy grasp) or 1 (friction contact)")
        self.contact_model = int(model)
        gp = grasp or GraspParams()
        self.grasp_params = _lib.GtsGraspParams(*[float(getattr(gp, n)) for n in _lib.GRASP_FIELDS],
                                                int(gp.exempt_rings))
        self.grasp_ring.zero_()  # a new contact model starts with no captured corners, so no ring
        if model == 0:
            dev, N = self.device, self.N
            z = lambda n, dt=torch.float32: torch.zeros(n, dtype=dt, device=dev)
            self.grasp_buffers = dict(grasped=z(N, torch.int32), core=z(N, torch.int32), core_offset=z(3 * N),
                                      anchor=z(2 * N), slab_lambda=z(2 * N), core_lambda=z(4 * N),
                                      core_force_ema=z(N), gate_lambda=z(2 * N), capture_scratch=z(4 * N))
            b = self.grasp_buffers
            self.grasp = _lib.GtsGraspState(*[b[k].data_ptr() for k in (
                "grasped", "core", "core_offset", "anchor", "slab_lambda", "core_lambda", "core_force_ema",
                "gate_lambda", "capture_scratch")], 0)

    def grasp_state(self):
        """The legacy grasp state: the latch, grasped and core flags, core force EMA, ring anchors (u, w) and the
        previous substep's slab multiplier."""
        b = self.grasp_buffers
        return dict(grasping=bool(self.grasp.grasping), grasped=b["grasped"].cpu().numpy(),
                    core=b["core"].cpu().numpy(), force_ema=b["core_force_ema"].cpu().numpy(),
                    anchor=b["anchor"].cpu().numpy().reshape(-1, 2),
                    slab_lambda_prev=b["slab_lambda"][self.N:].cpu().numpy())

    def set_colour_block(self, threads):
        """Workgroup size of the edge / tet colour kernels (a multiple of 64 up to 1024; 0: the default)."""
        self.view.colour_block = int(threads)

    # ---------------------------------------------------------------- stepping
    def set_gravity(self, g):
        self.gravity = tuple(float(x) for x in g)
        check("gts_set_gravity", self.lib.gts_set_gravity(ct.byref(self.view), *(float(x) for x in g),
                                                          current_stream()))

    # A colour of at most this many tets takes at most two waves per SIMD in one 1024-thread workgroup, about 2.6 us of
    # projection issue, below the 3.6-3.9 us a colour launch costs on MI355X; larger colours launch.
    PERSISTENT_COLOUR_TETS = 512

    def _persistent_view(self, p, tool_armed, blade_armed):
        """The one-instance batched view over this tissue's own buffers when the step qualifies, else None."""
        if not self.persistent or self.pooled or self.contact_model == 0 or self.cut_tool is not None:
            return None
        if p.contact_enabled and (tool_armed or blade_armed):
            return None
        if p.guard_passes or p.bend_kb > 0 or p.beta_s != 0 or not p.edge_ks > 0 or self.T <= 0:
            return None
        lds_nodes = int(self.lib.gts_batched_lds_bytes()) // 16
        if self.N > lds_nodes or int(np.diff(self.tet_offsets).max(initial=0)) > self.PERSISTENT_COLOUR_TETS:
            return None
        key = (self.T, self.E, self.tet_offsets.tobytes(), self.edge_offsets.tobytes(), float(p.edge_ks))
        if self._solo is None or self._solo[0] != key:
            dev, topo, grid = self.device, self.topology, self.grid
            nx, ny = grid.dims[0] + 1, grid.dims[1] + 1
            rec = GtsBatchInstance(0, self.N, 0, self.E, 0, self.T, nx, nx * ny, len(self.edge_offsets) - 1,
                                   len(self.tet_offsets) - 1, 0, 0, float(p.edge_ks),
                                   (ct.c_float * 3)(*[float(x) for x in topo["axis_rest_length"]]),
                                   (ct.c_float * 54)(*topo["type_inv_rest"].astype(np.float32).reshape(-1)))
            raw = np.frombuffer(bytes(rec), dtype=np.uint8).copy()
            keep = dict(instances=torch.as_tensor(raw).to(dev),
                        colours=torch.as_tensor(np.concatenate((self.edge_offsets, self.tet_offsets))
                                                .astype(np.int32)).to(dev),
                        work=torch.tensor([0, 1, 0], dtype=torch.int32, device=dev),
                        base=torch.zeros(self.N, dtype=torch.int32, device=dev))
            # one kernel and the frame update per step: launched directly (a graph would be keyed on the gravity,
            # which a caller may ramp every step)
            view = GtsBatchView(
                1, keep["instances"].data_ptr(), keep["colours"].data_ptr(), keep["work"].data_ptr(),
                keep["work"].data_ptr() + 8, 1, 1024, self.N, 0, 0, (ct.c_float * 3)(0.0, 0.0, 0.0),
                self.pos.data_ptr(), self.prev.data_ptr(), self.vel.data_ptr(), self.mass.data_ptr(),
                self.edge_desc.data_ptr(), self.tet_desc.data_ptr(), self.tet_k_dev.data_ptr(),
                self.tet_k_hyd.data_ptr(), self.edge_lambda.data_ptr(), self.tet_lambda.data_ptr(),
                self.tet_alpha.data_ptr(), self.N, self.nbr.data_ptr(), self.active.data_ptr(),
                self.pinned.data_ptr(), keep["base"].data_ptr(), self.rot.data_ptr(), self.rot_prev.data_ptr())
            self._solo = (key, view, keep)
        view = self._solo[1]
        for a in range(3):  # the batched kernel applies gravity * mass, as k_set_gravity's force does
            view.gravity[a] = self.gravity[a]
        if not self._liveness_fresh:
            check("gts_refresh_liveness", self.lib.gts_refresh_liveness(ct.byref(self.view), current_stream()))
            self._liveness_fresh = True
        return self._solo[1]

    def _guard_buffers(self):
        """The tissue guard's device buffers (xpbd_set_tissue_guard's, xpbd_solver.cu:1900-1912), allocated on the
        first guarded step: fibre multipliers, candidate targets, the node flags and each node's incident tet slots."""
        if getattr(self, "fiber_lambda", None) is not None:
            return
        dev, N, T = self.device, self.N, self.T
        self.fiber_lambda = torch.zeros(6 * max(T, 1), dtype=torch.float32, device=dev)
        self.feasible_target = torch.zeros((N, 4), dtype=torch.float32, device=dev)
        self.feasible_needed = torch.zeros(N + 1, dtype=torch.int32, device=dev)
        self._incidence()
        v = self.view
        v.fiber_lambda, v.feasible_target = self.fiber_lambda.data_ptr(), self.feasible_target.data_ptr()
        v.feasible_needed = self.feasible_needed.data_ptr()

    def _incidence(self):
        """CSR lists of the tet slots around each node, from the current tet descriptors (corner-0 node << 4 | Kuhn
        type << 1)."""
        from .grid import KUHN_TETS
        dev, N, T = self.device, self.N, self.T
        nx, ny = self.grid.dims[0] + 1, self.grid.dims[1] + 1
        bits = torch.as_tensor(KUHN_TETS, dtype=torch.int64, device=dev)
        corner = (bits & 1) + ((bits >> 1) & 1) * nx + ((bits >> 2) & 1) * nx * ny   # [6, 4] node offsets
        desc = self.tet_desc[:T].to(torch.int64) & 0xFFFFFFFF
        ids = ((desc >> 4)[:, None] + corner[(desc >> 1) & 7]).reshape(-1)
        order = torch.sort(ids, stable=True).indices
        self.incident_tets = (order // 4).to(torch.int32) if T else torch.zeros(1, dtype=torch.int32, device=dev)
        counts = torch.bincount(ids, minlength=N)
        self.incident_offsets = torch.cat((torch.zeros(1, dtype=torch.int64, device=dev),
                                           torch.cumsum(counts, 0))).to(torch.int32)
        self.view.incident_offsets = self.incident_offsets.data_ptr()
        self.view.incident_tets = self.incident_tets.data_ptr()

    def step(self, dt, tool=None, tool_from=None, tool_armed=False, blade_armed=False, params=None):
        """One outer step; the tool pose moves from `tool_from` (default: the previous step's `tool`) to `tool`."""
        p = step_params(params or self.params, dt)
        if p.guard_passes > 0:
            self._guard_buffers()
        to = tool_state(tool)
        frm = tool_state(tool_from) if tool_from is not None else (self._tool_from or to)
        if self.contact_model == 0:
            # xpbd_set_tool's legacy grasp latch for the new pose, then the legacy step
            self.grasp_event = 0
            if tool is not None:
                event = ct.c_int(0)
                check("gts_grasp_set_tool", self.lib.gts_grasp_set_tool(
                    ct.byref(self.view), ct.byref(self.grasp), ct.byref(self.grasp_params), ct.byref(to),
                    ct.byref(event), current_stream()))
                self.grasp_event = event.value
                if event.value == 1:
                    frm = to  # the capture offsets belong to this pose
            if self.cut_tool is not None:  # physics_step's ribbon tick runs whatever the contact model
                check("gts_step_legacy_cutting", self.lib.gts_step_legacy_cutting(
                    ct.byref(self.view), ct.byref(p), ct.byref(self.grasp_params), ct.byref(self.grasp),
                    ct.byref(frm), ct.byref(to), int(tool_armed), ct.byref(self.cut_tool),
                    self.emit_ncut.ctypes.data_as(ct.c_void_p), float(self.tear_tau), current_stream()))
            else:
                check("gts_step_legacy", self.lib.gts_step_legacy(
                    ct.byref(self.view), ct.byref(p), ct.byref(self.grasp_params), ct.byref(self.grasp),
                    ct.byref(frm), ct.byref(to), int(tool_armed), current_stream()))
        elif self.cut_tool is not None:
            check("gts_step_cutting", self.lib.gts_step_cutting(
                ct.byref(self.view), ct.byref(p), ct.byref(frm), ct.byref(to), int(tool_armed), int(blade_armed),
                ct.byref(self.cut_tool), self.emit_ncut.ctypes.data_as(ct.c_void_p), float(self.tear_tau),
                current_stream()))
        else:
            solo = self._persistent_view(p, tool_armed, blade_armed)
            if solo is not None:
                check("gts_batched_step", self.lib.gts_batched_step(ct.byref(solo), ct.byref(p), current_stream()))
            else:
                check("gts_step", self.lib.gts_step(ct.byref(self.view), ct.byref(p), ct.byref(frm), ct.byref(to),
                                                    int(tool_armed), int(blade_armed), current_stream()))
                self._liveness_fresh = True
        self._tool_from = to

    def set_cut_tool(self, cut_tool, tear_tau=0.0):
        """LCS_SetTool (cut_set_tool, cut.cu:1208-1226): from the next step on, every substep ends with the CCD ribbon
        tick against this tool's current rod (and the tear law when tear_tau >= 1); the corner snapshot is taken when a
        tool is first set. None disarms it."""
        if cut_tool is None:
            self.cut_tool = None
            return
        t = cut_tool_struct(cut_tool)
        if t.valid:
            self.emit_ncut = np.ascontiguousarray(np.asarray(cut_tool["ncut"], dtype=np.float32))
        if self.cut_tool is None:
            check("gts_snapshot_corners", self.lib.gts_snapshot_corners(ct.byref(self.view), current_stream()))
        self.cut_tool = t
        self.tear_tau = float(tear_tau)

    def detect_cut(self, cut_tool):
        t = cut_tool_struct(cut_tool)
        check("gts_detect_cut", self.lib.gts_detect_cut(ct.byref(self.view), ct.byref(t), current_stream()))

    def detect_cut_ccd(self, S, E, emit_ncut, tool_d, prev_positions):
        pp = torch.as_tensor(np.ascontiguousarray(prev_positions, dtype=np.float32)).reshape(-1).to(self.device)
        vs = [np.ascontiguousarray(v, dtype=np.float32) for v in (S, E, emit_ncut)]
        check("gts_detect_cut_ccd", self.lib.gts_detect_cut_ccd(
            ct.byref(self.view), *[v.ctypes.data_as(ct.c_void_p) for v in vs], float(tool_d), _ptr(pp),
            current_stream()))
        torch.cuda.current_stream().synchronize()

    def finalize(self):
        """LCS_Finalize's frame update after a cut: the particle frames from the current positions and links
        (physics_compute_rotations). step() already ends with one, as LCS_Step does."""
        check("gts_compute_rotations", self.lib.gts_compute_rotations(ct.byref(self.view), current_stream()))

    def rotations(self):
        """Particle frames, [N, 3, 3] rows of R."""
        return self.rot.cpu().numpy().reshape(-1, 3, 3)

    def tear_exempt(self):
        """The tear law's grasp ring [N] and each cut edge's exemption flag and release-decay ticks [Eg]."""
        Eg = self.Eg
        ex = self.edge_exempt.cpu().numpy()
        return self.grasp_ring[:self.N].cpu().numpy(), ex[:Eg], ex[Eg:2 * Eg]

    def tear(self, tau, tool_d):
        """Tear law: every intact link stretched past tau times its rest length is cut at its midpoint, then the
        links are severed (k_TearOverstretch + k_SeverLinks). tau < 1 turns it off, as in the reference."""
        check("gts_tear_overstretch", self.lib.gts_tear_overstretch(ct.byref(self.view), float(tau), float(tool_d),
                                                                    current_stream()))

    # ---------------------------------------------------------------- surface
    def init_surface(self, tri_capacity=None, qef_iters=8):
        """recon_init (dc_recon.cu:555-658): upload the grid's intersections, surface edges and voxel corners, build
        the surface once on the rest lattice with no cut state, and keep its feature points as the rest snapshot
        that later rebuilds advect. `tri_capacity` defaults to NativeTissue's (native_backend.py:291)."""
        g, dev, N, V = self.grid, self.device, self.N, self.V
        Se, I = len(g.surface_edges), len(g.intersections)
        if V and int(np.max(g.counts)) > 12:
            raise ValueError("a voxel lists at most one surface crossing per edge (12)")
        cap = int(tri_capacity or 6 * Se + 1024 + 2 * self.Eg)
        fc = 1
        while fc < 2 * cap:
            fc *= 2
        nc, items = 2 * Se + 12 * N, max(Se + 6 * N, cap)
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32)).to(dev)
        zf = lambda n: torch.zeros(max(int(n), 1), dtype=torch.float32, device=dev)
        zi = lambda n: torch.zeros(max(int(n), 1), dtype=torch.int32, device=dev)
        edge_of = np.full(3 * N, -1, dtype=np.int32)
        if self.Eg:
            e = g.edges.astype(np.int64)
            nx, ny = g.dims[0] + 1, g.dims[1] + 1
            edge_of[3 * (e[:, 1] + nx * (e[:, 2] + ny * e[:, 3])) + e[:, 0]] = np.arange(self.Eg, dtype=np.int32)
        isect = np.ascontiguousarray(g.intersections).view(np.float32).reshape(-1) if I else np.zeros(10, np.float32)
        self.surf = dict(
            voxel_corners=i32(g.corners.reshape(-1)), isect_offset=i32(g.offsets), isect_count=i32(g.counts),
            isect=torch.as_tensor(isect.copy()).to(dev), surface_edges=i32(g.surface_edges.reshape(-1) if Se else [0]),
            conn=i32(g.connectivity.reshape(-1)), edge_of=i32(edge_of), rest_fp=zf(4 * V),
            isect_world=zf(3 * I), fp=zf(4 * V), cut_fp=zf(32 * V), edge_points=zi(2 * self.Eg).fill_(-1),
            candidates=zi(3 * nc), item_mask=zi(items), tri=zi(3 * cap), tri_work=zi(3 * cap), face_table=zi(4 * fc),
            normal_fixed=zi(27 * V), scan=zi(-(-items // 256) + 1), counts=zi(4), out_pos=zf(9 * cap),
            out_nrm=zf(9 * cap), out_aux=zf(12 * cap))
        names = [f[0] for f in _lib.GtsSurfaceView._fields_[6:]]
        self.surface_view = _lib.GtsSurfaceView(V, I, Se, cap, int(qef_iters), fc,
                                                *[self.surf[n].data_ptr() for n in names])
        # the rest build reads the rest lattice, whatever the tissue's initial positions
        rest4 = torch.zeros((N, 4), dtype=torch.float32, device=dev)
        rest4[:, :3] = self.rest.reshape(N, 3)
        at_rest = GtsTissueView.from_buffer_copy(self.view)
        at_rest.pos = rest4.data_ptr()
        check("gts_surface_build", self.lib.gts_surface_build(ct.byref(at_rest), ct.byref(self.surface_view), 0, 0,
                                                              current_stream()))
        self.surf["rest_fp"].copy_(self.surf["fp"])
        torch.cuda.current_stream().synchronize()

    def rebuild_surface(self):
        """The surface rebuild of LCS_Finalize (recon_rebuild, dc_recon.cu:736-741) from the current positions,
        frames, links and cut state."""
        if getattr(self, "surface_view", None) is None:
            self.init_surface()
        check("gts_surface_build", self.lib.gts_surface_build(ct.byref(self.view), ct.byref(self.surface_view), 1, 1,
                                                              current_stream()))

    def material_quality(self, out=None):
        """LCS_GetMaterialQuality (xpbd_get_material_quality, xpbd_solver.cu:2133-2138) as a device tensor [2]: the
        smallest det F over the live tets (seeded at 1) and how many have det F <= 0 or a non-finite det F."""
        out = torch.empty(2, dtype=torch.float32, device=self.device) if out is None else out
        check("gts_material_quality", self.lib.gts_material_quality(ct.byref(self.view), out.data_ptr(),
                                                                    current_stream()))
        return out

    def surface_depth(self, tool, kinds=0b0111, out=None):
        """LCS_SurfaceDepth (xpbd_surface_depth, xpbd_solver.cu:2046-2059): the deepest penetration of the tool's
        shapes into the last rebuilt surface, as a 0-d device tensor (0 when none). `kinds` selects them by bit: 1 the
        shaft capsule, 2 and 4 the jaw boxes (with the step's jaw_thickness and jaw_gap_closed), 8 the blade capsule."""
        if getattr(self, "surface_view", None) is None:
            raise RuntimeError("the surface has not been built")
        out = torch.empty((), dtype=torch.float32, device=self.device) if out is None else out
        check("gts_surface_depth", self.lib.gts_surface_depth(
            ct.byref(self.surface_view), ct.byref(tool_state(tool)), int(kinds), float(self.params.jaw_thickness),
            float(self.params.jaw_gap_closed), out.data_ptr(), current_stream()))
        return out

    def surface_overflowed(self):
        """Whether the last rebuild's triangles exceeded the capacity (cleanInternalFaces then fails,
        surface_cleanup.cuh:38)."""
        return int(self.surf["counts"][2].item()) > 3 * self.surface_view.tri_capacity

    def surface(self):
        """LCS_GetSurfaceCounts / LCS_GetSurface / LCS_GetSurfaceAux (recon_get_counts / recon_get_surface /
        recon_get_surface_aux, dc_recon.cu:660-712) of the last rebuild: the triangle soup's positions [K, 3],
        normals [K, 3], indices [K / 3, 3] (0..K-1) and aux [K, 4] (rest anchor, wall flag), plus the vertex tags,
        the feature points, the index count before the cleanup and the index count the triangles asked for."""
        s = self.surf
        k = int(s["counts"][0].item())
        return dict(pos=s["out_pos"][:3 * k].reshape(-1, 3).cpu().numpy(),
                    nrm=s["out_nrm"][:3 * k].reshape(-1, 3).cpu().numpy(),
                    indices=np.arange(k, dtype=np.int32).reshape(-1, 3),
                    aux=s["out_aux"][:4 * k].reshape(-1, 4).cpu().numpy(),
                    tags=s["tri"][:k].cpu().numpy(), fp=s["fp"].reshape(-1, 4).cpu().numpy(),
                    cut_fp=s["cut_fp"].reshape(-1, 4).cpu().numpy(), listed=int(s["counts"][1].item()),
                    asked=int(s["counts"][2].item()))

    # ---------------------------------------------------------------- readback
    def observation(self, known_revision=None, with_core=False):
        """The readbacks of one Observation in one round trip (LCS_GetCornerPos, LCS_GetTopologyRevision,
        LCS_GetContactReaction and, with `with_core`, LCS_GetGraspCoreFlags; plugin_api.cpp:170-176, 211-235): copies
        on the current stream into pinned host buffers allocated once, one wait. The neighbour table (LCS_GetNbrIdx)
        is read in a second round trip only when the revision differs from `known_revision`, as the reference's
        observe does (native_backend.py:396-403); otherwise `neighbors` is None. The arrays returned are fresh."""
        st = self._staging
        if st is None:
            def pin(n, dtype):
                return torch.empty(n, dtype=dtype, pin_memory=True)
            st = self._staging = dict(pos=pin((self.N, 3), torch.float32), counters=pin(8, torch.int32),
                                      reaction=pin(4, torch.float32), nbr=pin(6 * self.N, torch.int32),
                                      core=pin(self.N, torch.int32),
                                      xyz=torch.empty((self.N, 3), dtype=torch.float32, device=self.device))
        stream = torch.cuda.current_stream()
        # positions packed to xyz rows on the device first: one contiguous copy each way, no strided host copy
        st["xyz"].copy_(self.pos[:, :3])
        st["pos"].copy_(st["xyz"], non_blocking=True)
        st["counters"].copy_(self.counters, non_blocking=True)
        st["reaction"].copy_(self.reaction, non_blocking=True)
        if with_core:
            st["core"].copy_(self.grasp_buffers["core"], non_blocking=True)
        stream.synchronize()
        revision = int(st["counters"].numpy().view(np.uint32)[0])
        r = st["reaction"].numpy()
        out = dict(positions=st["pos"].numpy().copy(), revision=revision,
                   reaction=(r[:3].astype(np.float64), int(round(float(r[3])))),
                   core=st["core"].numpy().copy() if with_core else None, neighbors=None)
        if revision != known_revision:
            st["nbr"].copy_(self.nbr, non_blocking=True)
            stream.synchronize()
            out["neighbors"] = st["nbr"].numpy().reshape(-1, 6).copy()
        return out

    def positions(self):
        return self.pos[:, :3].cpu().numpy()

    def velocities(self):
        return self.vel[:, :3].cpu().numpy()

    def inverse_mass(self):
        return self.pos[:, 3].cpu().numpy()

    def set_state(self, pos, vel):
        self.pos[:, :3] = torch.as_tensor(np.asarray(pos, dtype=np.float32), device=self.device)
        self.vel[:, :3] = torch.as_tensor(np.asarray(vel, dtype=np.float32), device=self.device)

    def neighbors(self):
        return self.nbr.cpu().numpy().reshape(-1, 6)

    def revision(self):
        return int(self.counters[0].item()) & 0xffffffff

    def reaction_force(self):
        r = self.reaction.cpu().numpy()
        return r[:3].astype(np.float64), int(round(float(r[3])))

    def cut_state(self):
        counters = self.counters.cpu().numpy().view(np.uint32)
        n = int(min(counters[1], self.cut_point_capacity))
        raw = self.cut_points[:8 * n].cpu().numpy() if n else np.zeros(0, dtype=np.float32)
        points = raw.view(CUT_POINT_DTYPE).copy()     # the device records are CutPointGpu (cut.cuh:37-46)
        bend = self.bend_alive.cpu().numpy().view(np.uint32)
        bend_alive = ((bend[:, None] >> np.arange(15, dtype=np.uint32)) & 1).astype(np.int32)
        return dict(neighbors=self.neighbors(), cut_mask=self.cut_mask.cpu().numpy().view(np.uint32),
                    edge_cut_t=self.edge_cut_t.cpu().numpy().reshape(-1, 6), bend_alive=bend_alive,
                    revision=int(counters[0]), cut_point_counter=int(counters[1]), tear_count=int(counters[4]),
                    cut_points=points)

    def liveness(self):
        """Tet liveness in generation order and edge liveness in colour order over the original (uncompacted) lists,
        as the CPU checker reports them; removed entries read 0."""
        tet = torch.zeros(max(self.T0, 1), dtype=torch.int32, device=self.device)
        tet[self.tet_slot_gen] = self.tet_desc[:self.T] & 1
        edge = torch.zeros(max(self.E0, 1), dtype=torch.int32, device=self.device)
        edge[self.edge_slot_orig] = self.edge_desc[:self.E] & 1
        return tet[:self.T0].cpu().numpy(), edge[:self.E0].cpu().numpy()

    def multipliers(self):
        """(edge multipliers, tet deviatoric, tet hydrostatic) after the last substep, in the orders liveness()
        uses; removed entries read 0."""
        edge = torch.zeros(max(self.E0, 1), dtype=torch.float32, device=self.device)
        edge[self.edge_slot_orig] = self.edge_lambda[:self.E]
        dev = torch.zeros(max(self.T0, 1), dtype=torch.float32, device=self.device)
        hyd = torch.zeros_like(dev)
        dev[self.tet_slot_gen] = self.tet_lambda[:self.T]
        hyd[self.tet_slot_gen] = self.tet_lambda[self.T:2 * self.T]
        return edge[:self.E0].cpu().numpy(), dev[:self.T0].cpu().numpy(), hyd[:self.T0].cpu().numpy()

    def compact(self, apply=False, readback=True):
        """Stable remap of the live topology: (tet_remap in colour order, edge_remap, node_remap, counts), relative
        to the current lists. With `apply`, the dead tets and edges are removed: the live ones move to their remapped
        slots, the colour ranges shrink to their live counts and every later step and cut runs over the live lists
        only (node storage keeps its lattice addressing). With readback=False only the counts come back to the host
        (the remaps stay on the device, as the first three entries of the returned tuple)."""
        dev = self.device
        tr = torch.empty(max(self.T, 1), dtype=torch.int32, device=dev)
        er = torch.empty(max(self.E, 1), dtype=torch.int32, device=dev)
        nr = torch.empty(self.N, dtype=torch.int32, device=dev)
        counts = torch.zeros(3, dtype=torch.int32, device=dev)
        used = torch.empty(self.N, dtype=torch.int32, device=dev)
        check("gts_compact_topology", self.lib.gts_compact_topology(
            ct.byref(self.view), _ptr(tr), _ptr(er), _ptr(nr), _ptr(counts), _ptr(used), current_stream()))
        T, E = self.T, self.E
        counts_h = counts.cpu().numpy()
        out = ((tr[:T].cpu().numpy(), er[:E].cpu().numpy(), nr.cpu().numpy(), counts_h) if readback else
               (tr[:T], er[:E], nr, counts_h))
        if apply:
            self._apply_compaction(tr, er, int(counts_h[0]), int(counts_h[1]))
        return out

    def _apply_compaction(self, tr, er, new_t, new_e):
        if self.pooled:
            raise ValueError("a tissue in a batch's pooled buffers cannot be compacted in place")
        dev = self.device
        T, E = self.T, self.E
        tet_desc = torch.zeros(max(new_t, 1), dtype=torch.int32, device=dev)
        k_dev = torch.zeros(max(new_t, 1), dtype=torch.float32, device=dev)
        k_hyd = torch.zeros(max(new_t, 1), dtype=torch.float32, device=dev)
        edge_desc = torch.zeros(max(new_e, 1), dtype=torch.int32, device=dev)
        check("gts_apply_compaction", self.lib.gts_apply_compaction(
            ct.byref(self.view), _ptr(tr), _ptr(er), _ptr(tet_desc), _ptr(k_dev), _ptr(k_hyd), _ptr(edge_desc),
            current_stream()))
        # live count before each colour boundary, and the readback order of the survivors, on the device
        live_t, live_e = tr[:T] >= 0, er[:E] >= 0
        kept_t = torch.cat((torch.zeros(1, dtype=torch.int64, device=dev), torch.cumsum(live_t, 0)))
        kept_e = torch.cat((torch.zeros(1, dtype=torch.int64, device=dev), torch.cumsum(live_e, 0)))
        bounds = torch.cat((kept_t[torch.as_tensor(self.tet_offsets, dtype=torch.int64, device=dev)],
                            kept_e[torch.as_tensor(self.edge_offsets, dtype=torch.int64, device=dev)])).cpu().numpy()
        nt = len(self.tet_offsets)
        self.tet_offsets = np.ascontiguousarray(bounds[:nt], dtype=np.int32)
        self.edge_offsets = np.ascontiguousarray(bounds[nt:], dtype=np.int32)
        self.tet_slot_gen = self.tet_slot_gen[live_t]
        self.edge_slot_orig = self.edge_slot_orig[live_e]
        self.tet_desc, self.tet_k_dev, self.tet_k_hyd, self.edge_desc = tet_desc, k_dev, k_hyd, edge_desc
        self.T, self.E = new_t, new_e
        v = self.view
        v.n_tets, v.n_edges = new_t, new_e
        v.tet_colour_offsets, v.edge_colour_offsets = self.tet_offsets.ctypes.data, self.edge_offsets.ctypes.data
        v.tet_desc, v.tet_k_dev, v.tet_k_hyd = tet_desc.data_ptr(), k_dev.data_ptr(), k_hyd.data_ptr()
        v.edge_desc = edge_desc.data_ptr()
        if getattr(self, "fiber_lambda", None) is not None:  # the guard's lists follow the new tet slots
            self._incidence()
