"""ctypes binding of libgts_xpbd.so (declarations in include/gts_xpbd.h).

Loading fails loudly when the library has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as ct
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libgts_xpbd.so"

F = ct.c_float
I = ct.c_int
P = ct.c_void_p
V3 = F * 3

GTS_ABI_VERSION = 2
GTS_MAX_COLOURS = 64
GTS_STEP_CONST_FLOATS = 184
STATUS = {0: "ok", -1000: "not ready", -1002: "bad dt", -1003: "colour overflow", -1004: "bad argument",
          -1005: "irregular rest lattice", -3004: "capacity exceeded"}


class GtsLattice(ct.Structure):
    _fields_ = [("vx", I), ("vy", I), ("vz", I), ("rest", P), ("active", P), ("nbr", P), ("rest_nbr", P)]


class GtsTopology(ct.Structure):
    _fields_ = [("edge_desc", P), ("edge_gen_index", P), ("edge_colour_offsets", P), ("tet_desc", P),
                ("tet_gen_index", P), ("tet_colour_offsets", P), ("type_inv_rest", ct.c_double * 54),
                ("type_volume", ct.c_double * 6), ("axis_rest_length", F * 3), ("n_edges", I), ("n_tets", I),
                ("n_edge_colours", I), ("n_tet_colours", I), ("bend_desc", P), ("bend_gen_index", P),
                ("bend_colour_offsets", P), ("n_bends", I), ("n_bend_colours", I)]


class GtsTissueView(ct.Structure):
    _fields_ = [("n_nodes", I), ("n_edges", I), ("n_tets", I), ("n_grid_edges", I), ("vx", I), ("vy", I),
                ("vz", I), ("n_edge_colours", I), ("n_tet_colours", I), ("edge_colour_offsets", P),
                ("tet_colour_offsets", P), ("inv_rest", F * 54), ("axis_rest_length", F * 3),
                ("voxel_length", F), ("pos", P), ("prev", P), ("vel", P), ("force", P), ("mass", P),
                ("active", P), ("pinned", P), ("inside", P), ("nbr", P), ("bend_alive", P), ("edge_desc", P),
                ("edge_lambda", P), ("slot_support", P), ("tet_desc", P), ("tet_k_dev", P), ("tet_k_hyd", P),
                ("tet_alpha", P), ("tet_lambda", P), ("grid_edges", P), ("occupied", P), ("cut_mask", P),
                ("edge_cut_t", P), ("edge_skin_t", P), ("rest", P), ("rot", P), ("rot_prev", P), ("prev_corner", P), ("cut_points", P), ("cut_point_capacity", I),
                ("counters", P), ("scratch", P), ("reaction", P), ("step_consts", P), ("step_const_capacity", I),
                ("use_graph", I), ("colour_block", I), ("fiber_lambda", P), ("feasible_target", P),
                ("feasible_needed", P), ("incident_offsets", P), ("incident_tets", P), ("grasp_ring", P),
                ("edge_exempt", P), ("nbr_init", P), ("xcd_order", I), ("n_bends", I), ("n_bend_colours", I),
                ("bend_colour_offsets", P), ("bend_desc", P), ("bend_lambda", P), ("tile_groups", I),
                ("tile_kmax", I), ("tile_box", I), ("tile_tets", I), ("tile_group_info", P),
                ("tile_entries", P), ("tile_writes", P), ("pos_alt", P), ("tile_inputs", P),
                ("tile_lambda", P), ("tile_region", I), ("tile_rx", I), ("tile_ry", I), ("tile_dense_ids", P),
                ("tile_edges", I)]


class GtsStepParams(ct.Structure):
    _fields_ = [("dt", F), ("substeps", I), ("iterations", I), ("tet_mode", I), ("omega", F), ("edge_ks", F),
                ("young", F), ("beta_s", F), ("strain_damping", F), ("bulk_damping", F), ("rayleigh", F),
                ("damping", F), ("contact_enabled", I), ("mu_s", F), ("mu_d", F), ("jaw_thickness", F),
                ("contact_radius", F), ("jaw_gap_closed", F), ("contact_compliance", F), ("precise_strain", I),
                ("guard_min_volume", F), ("guard_passes", I), ("tensile_strength", F), ("bend_kb", F),
                ("beta_b", F)]


class GtsToolState(ct.Structure):
    _fields_ = [("shaftA", V3), ("shaftB", V3), ("shaftR", F), ("jawHinge", V3), ("jawAxis", V3), ("jawMid", V3),
                ("jawNormal", V3), ("jawHalfWidth", F), ("jawLen", F), ("jawGap", F), ("grasp01", F),
                ("bladeA", V3), ("bladeB", V3), ("bladeR", F)]


class GtsCutTool(ct.Structure):
    _fields_ = [("t1v0", V3), ("t1v1", V3), ("t1v2", V3), ("t2v0", V3), ("t2v1", V3), ("t2v2", V3), ("ncut", V3),
                ("d", F), ("valid", I), ("aabbMin", V3), ("aabbMax", V3)]


class GtsBatchInstance(ct.Structure):
    _fields_ = [("node_off", I), ("n_nodes", I), ("edge_off", I), ("n_edges", I), ("tet_off", I), ("n_tets", I),
                ("nx", I), ("nxy", I), ("n_edge_colours", I), ("n_tet_colours", I), ("colour_off", I),
                ("reserved", I), ("edge_ks", F), ("rest_len", F * 3), ("inv_rest", F * 54)]


class GtsBatchView(ct.Structure):
    _fields_ = [("n_instances", I), ("instances", P), ("colour_table", P), ("work_offsets", P), ("work_items", P),
                ("n_workgroups", I), ("block_size", I), ("lds_nodes", I), ("use_graph", I),
                ("substeps_per_launch", I), ("gravity", F * 3),
                ("pos", P), ("prev", P), ("vel", P), ("mass", P), ("edge_desc", P), ("tet_desc", P), ("k_dev", P),
                ("k_hyd", P), ("lambda_edge", P), ("lambda_tet", P), ("alpha", P), ("total_nodes", ct.c_longlong),
                ("nbr", P), ("active", P), ("pinned", P), ("node_base", P), ("rot", P), ("rot_prev", P)]


class GtsAdjointView(ct.Structure):
    _fields_ = [("batch", GtsBatchView), ("total_nodes", ct.c_longlong), ("total_tets", ct.c_longlong),
                ("total_edges", ct.c_longlong), ("force", P), ("compliance", P), ("alpha", P), ("checkpoints", P),
                ("tet_tape", P), ("edge_tape", P), ("lambda_tape", P), ("xbar", P), ("vbar", P), ("prevbar", P),
                ("abar", P), ("lambar", P), ("grad", P)]


GRASP_FIELDS = ("capture", "release", "closed_gap", "core_fraction", "friction", "contact_compliance", "skin",
                "ccd_velocity", "ccd_displacement", "ccd_omega", "grasp_omega", "core_compliance", "break_force",
                "grasp_displacement")


class GtsGraspParams(ct.Structure):
    _fields_ = [(name, F) for name in GRASP_FIELDS] + [("exempt_rings", I)]


class GtsGraspState(ct.Structure):
    _fields_ = [("grasped", P), ("core", P), ("core_offset", P), ("anchor", P), ("slab_lambda", P),
                ("core_lambda", P), ("core_force_ema", P), ("gate_lambda", P), ("capture_scratch", P),
                ("grasping", I)]


class GtsSurfaceView(ct.Structure):
    _fields_ = [("n_voxels", I), ("n_isect", I), ("n_surface_edges", I), ("tri_capacity", I), ("qef_iters", I),
                ("face_capacity", I)] + [(name, P) for name in (
                    "voxel_corners", "isect_offset", "isect_count", "isect", "surface_edges", "conn", "edge_of",
                    "rest_fp", "isect_world", "fp", "cut_fp", "edge_points", "candidates", "item_mask", "tri",
                    "tri_work", "face_table", "normal_fixed", "scan", "counts", "out_pos", "out_nrm", "out_aux")]


STRUCTS = (GtsLattice, GtsTopology, GtsTissueView, GtsStepParams, GtsToolState, GtsCutTool, GtsBatchInstance,
           GtsBatchView, GtsAdjointView, GtsGraspParams, GtsGraspState, GtsSurfaceView)

# name -> (argtypes, restype); the exported C ABI
SIGNATURES = {
    "gts_abi_version": ((), I),
    "gts_struct_sizes": ((P,), I),
    "gts_build_topology": ((P, P), I),
    "gts_iso_pow": ((F,), F),
    "gts_log": ((F,), F),
    "gts_acos": ((F,), F),
    "gts_scratch_floats": ((I, I, I, I), ct.c_longlong),
    "gts_init_state": ((P, P), I),
    "gts_set_gravity": ((P, F, F, F, P), I),
    "gts_refresh_liveness": ((P, P), I),
    "gts_step": ((P, P, P, P, I, I, P), I),
    "gts_detect_cut": ((P, P, P), I),
    "gts_detect_cut_static": ((P, P, P), I),
    "gts_detect_cut_static_floats": ((I,), ct.c_longlong),
    "gts_detect_cut_ccd": ((P, P, P, P, F, P, P), I),
    "gts_tear_overstretch": ((P, F, F, P), I),
    "gts_compute_rotations": ((P, P), I),
    "gts_step_cutting": ((P, P, P, P, I, I, P, P, F, P), I),
    "gts_snapshot_corners": ((P, P), I),
    "gts_grasp_set_tool": ((P, P, P, P, P, P), I),
    "gts_step_legacy": ((P, P, P, P, P, P, I, P), I),
    "gts_step_legacy_cutting": ((P, P, P, P, P, P, I, P, P, F, P), I),
    "gts_compact_topology": ((P, P, P, P, P, P, P), I),
    "gts_apply_compaction": ((P, P, P, P, P, P, P, P), I),
    "gts_profile_begin": ((), I),
    "gts_profile_end": ((P, P, P), I),
    "gts_batched_plan": ((P, I, I, P), I),
    "gts_batched_step": ((P, P, P), I),
    "gts_batched_lds_bytes": ((), I),
    "gts_stream_copy": ((P, P, ct.c_longlong, P), I),
    "gts_adjoint_forward": ((P, P, I, I, P), I),
    "gts_adjoint_backward": ((P, P, I, I, P), I),
    "gts_surface_build": ((P, P, I, I, P), I),
    "gts_surface_depth": ((P, P, I, F, F, P, P), I),
    "gts_material_quality": ((P, P, P), I),
}

_lib = None


class NativeError(RuntimeError):
    pass


def load():
    """The loaded library; raises if it is missing or was built for another ABI."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise NativeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() or "
                          "python -m genesis_tissue_simulation_amd._build")
    lib = ct.CDLL(str(LIB_PATH))
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes, fn.restype = list(args), res
    if lib.gts_abi_version() != GTS_ABI_VERSION:
        raise NativeError("libgts_xpbd.so ABI version mismatch")
    sizes = np.zeros(len(STRUCTS), dtype=np.int32)
    lib.gts_struct_sizes(sizes.ctypes.data_as(P))
    for size, st in zip(sizes, STRUCTS):
        if int(size) != ct.sizeof(st):
            raise NativeError(f"{st.__name__}: library size {int(size)}, Python size {ct.sizeof(st)}")
    _lib = lib
    return lib


def check(name, status):
    if status != 0:
        raise NativeError(f"{name} failed with native status {status} ({STATUS.get(status, 'HIP error')})")


def v3(values):
    return V3(*[float(v) for v in values])
