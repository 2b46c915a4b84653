"""Drop-in counterpart of the reference's Python-facing tissue API (`endoscopic_tissue/native_backend.py`).

`NativeTissue(grid, settings, dll_path, initial_command)` with `.step(Command) -> Observation`, `.observe()`,
`.cut_edge_fractions()`, `.material_quality()`, `.contact_depth_m()`, `.knife_body_depth_m()`, `.close()` and the
context manager keep the reference's names, argument meaning and error behaviour (`native_backend.py:84-196,
:212-483`). Underneath, the legacy XPBD substep and the lattice cut run as HIP kernels on cuda:0
(libgts_xpbd.so), with device buffers held as PyTorch-ROCm tensors.

What the step does, per reference call:
  * `LCS_SetGraspTool`: the grasper (shaft capsule + two jaw plates) from the command (`native_backend.py:341-360`);
  * `LCS_SetGravity`: gravity ramped over 0.5 s (`:379-380`);
  * `LCS_Step`: `Settings.substeps` substeps of the legacy branch (`xpbd/xpbd_solver.cu:3460-3721`) with the
    material parameters `initialize` sends (`:289-325`): fallback edges only (ks = 0), guarded Neo-Hookean tets,
    strain damping, bulk damping 0.01 s. The tool acts through the contact model `Settings.contact_model` selects
    (`LCS_SetContactModel`, `:318`): 0 is the legacy grasp (capture / release latch, slab ring with Coulomb anchors,
    core pins, break check) with the CCD gate, using the `GraspParams` `initialize` sends (`:314-317`); 1 and 2 are
    the friction contact of the lattice nodes, Coulomb 0.6 / 0.4 (model 2's persistent anchors and embedded grip
    are not part of this path);
  * the knife: a kinematic capsule whose swept surface between consecutive commands cuts the lattice
    (`LCS_SetTool` + `LCS_DetectCut`, `plugin_api.cpp:145-155`) instead of the force-driven pressure blade;
  * `LCS_Finalize`: the particle frames, then the rebuild of the dual-contouring surface with its cut walls
    (`plugin_api.cpp:161-169`), which `surface()` reads (`LCS_GetSurface*`, `native_backend.py:434-451`).
Out of this path: scissors and the unified VBD/PCG branch.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .grid import METERS_PER_UNIT, TissueGrid
from .params import TET_GUARDED, SolverParams, legacy_grasp_params, neohookean_stiffness

# Analytic instrument dimensions (cm); the reference derives them from the instrument meshes
# (`instrument_geometry.py`), which this path does not load.
GRASPER_SHAFT_LENGTH = 3.0
GRASPER_SHAFT_RADIUS = 0.4


@dataclass(frozen=True)
class Settings:
    """Field-for-field the reference settings (`native_backend.py:84-116`)."""
    dt: float = 1 / 30
    substeps: int = 3
    young_pa: float = 6000
    poisson: float = 0.25
    strain_damping_s: float = 0.04
    density_kg_m3: float = 1050
    jaw_length: float = 1.99
    jaw_width: float = 0.52
    jaw_open_gap: float = 1.0
    jaw_axis: tuple = (0.0, -0.8, 0.6)
    jaw_normal: tuple = (1.0, 0.0, 0.0)
    blade_length: float = 1.6
    blade_radius: float = 0.04
    blade_axis: tuple = (2 ** -.5, 0.0, 2 ** -.5)
    start_force_n: float = 0.2
    cut_force_n: float = 0.12
    blade_drive_stiffness: float = 1200
    blade_mass_kg: float = 0.02
    blade_friction: float = 0.15
    blade_preload_n: float = 1.0
    contact_refine_passes: int = 8
    has_knife_body: bool = True
    instrument_model: str = "davinci"
    cutting_tool: str = "knife"
    scissors_jaw_stiffness: float = 150
    scissors_squeeze_cm: float = .2
    contact_model: int = 2
    # constraint iterations per substep (the reference sends 1, physics.cu:281-295)
    iterations: int = 1


@dataclass(frozen=True)
class Command:
    """Field-for-field the reference command (`native_backend.py:119-130`)."""
    hinge: tuple = (-1.4, 3.384, -2.288)
    closure: float = 0.0
    blade_center: tuple = (0.2, 2.2, -0.6)
    blade_enabled: bool = True
    scissors_closure: float = 0.0
    scissors_cutting: bool = False
    scissors_pitch_deg: float = 19.9
    scissors_roll_deg: float = 0.0
    scissors_yaw_deg: float = 0.0


def scissors_frame(pitch_deg, roll_deg=0.0, yaw_deg=0.0):
    """The scissors' axis and opening normal in native coordinates (`native_backend.py:140-152`): at zero angles the
    axis points along -x and the first blade opens along +y; the roll turns the normal about the axis towards +z, the
    pitch then tilts the axis downwards (about +z), and the yaw turns the result about +y towards +z."""
    r, p, y = np.radians((roll_deg, pitch_deg, yaw_deg))

    def about_x(v, a):
        return np.array((v[0], np.cos(a) * v[1] - np.sin(a) * v[2], np.sin(a) * v[1] + np.cos(a) * v[2]))

    def about_z(v, a):
        return np.array((np.cos(a) * v[0] - np.sin(a) * v[1], np.sin(a) * v[0] + np.cos(a) * v[1], v[2]))

    def about_y(v, a):
        return np.array((np.cos(a) * v[0] + np.sin(a) * v[2], v[1], -np.sin(a) * v[0] + np.cos(a) * v[2]))

    def turn(v):
        return about_y(about_z(about_x(v, r), p), y)

    return turn(np.array((-1.0, 0.0, 0.0))), turn(np.array((0.0, 1.0, 0.0)))


@dataclass(frozen=True)
class ScissorsLoads:
    """The reference's scissors blade loads (`native_backend.py:155-166`), in newtons. The pressure-fracture scissors
    are outside the legacy path, so observations here never carry one; the type is kept for callers that name it."""
    left_n: float
    right_n: float
    fracture_left_n: float
    fracture_right_n: float
    fracture_total_n: float
    fractures: int
    is_squeezing: bool


@dataclass(frozen=True)
class ScissorsPose:
    """The reference's solved scissors pose (`native_backend.py:169-180`); `Observation.scissors` is always None on the
    legacy path."""
    hinge: np.ndarray
    axis: np.ndarray
    normal: np.ndarray
    gap: float
    radius: float
    edges: np.ndarray
    pressure_n: float
    phase: int
    contacts: int
    loads: ScissorsLoads


@dataclass
class Observation:
    """Field-for-field the reference observation (`native_backend.py:182-196`)."""
    nodes: np.ndarray
    neighbors: np.ndarray
    grip_flags: np.ndarray
    hinge: np.ndarray
    jaw_gap: float
    grip_points: int
    grip_error_m: float
    blade: np.ndarray
    reaction_n: np.ndarray
    scissors: ScissorsPose | None = None
    jaw_load_n: tuple = (0.0, 0.0)


def solver_params(settings: Settings) -> SolverParams:
    """The solver parameters `NativeTissue.initialize` sends (`native_backend.py:289-325`), with the tissue guard of
    LCS_SetTissueGuard(0.15, 1, ...) and the material of LCS_SetTissueMaterial(80, 0.01) (`native_backend.py:308-309`)."""
    return SolverParams(substeps=int(settings.substeps), iterations=int(settings.iterations), tet_mode=TET_GUARDED,
                        omega=1.0, edge_ks=0.0, young=float(settings.young_pa * METERS_PER_UNIT), beta_s=0.0,
                        strain_damping=float(settings.strain_damping_s), bulk_damping=0.01, rayleigh=0.0,
                        damping=0.0, contact_enabled=True, mu_s=0.6, mu_d=0.4, jaw_thickness=0.1,
                        contact_radius=0.01, jaw_gap_closed=0.0, contact_compliance=1e-8, guard_min_volume=0.15,
                        guard_passes=1, tensile_strength=80.0)


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


@dataclass
class Surface:
    """Field-for-field the reference's reconstructed surface (`native_backend.py:199-204`): a triangle soup of
    vertices and normals f32[K, 3], triangles i32[K / 3, 3] over them and anchors f32[K, 4] (rest anchor, wall flag)."""
    vertices: np.ndarray
    normals: np.ndarray
    triangles: np.ndarray
    anchors: np.ndarray


class NativeTissue:
    """Owns one tissue on the GPU; like the reference, one session per process (close before constructing another)."""
    active_session = False

    def __init__(self, grid: TissueGrid, settings=Settings(), dll_path=None, initial_command=Command()):
        if NativeTissue.active_session:
            raise RuntimeError("the tissue solver is single-session; use separate processes for separate environments")
        self.grid, self.settings = grid, settings
        self.initial_command = initial_command
        if settings.cutting_tool not in ("knife", "scissors"):
            raise ValueError(f"Unknown cutting tool: {settings.cutting_tool}")
        if settings.cutting_tool == "scissors":
            raise ValueError("scissors are outside this path; use the knife")
        from . import _lib
        if dll_path is not None and Path(dll_path).resolve() != _lib.LIB_PATH.resolve():
            raise ValueError(f"{dll_path}: this backend loads {_lib.LIB_PATH}")
        self.dll_path = _lib.LIB_PATH
        self.blade_axis = _unit(settings.blade_axis)
        self.closed = True
        NativeTissue.active_session = True
        self.closed = False
        try:
            self.initialize()
        except (RuntimeError, ValueError, OSError):
            self.close()
            raise

    def initialize(self):
        from .tissue import Tissue
        g, s = self.grid, self.settings
        self.params = solver_params(s)
        from .tissue import build_topology
        topo = build_topology(g)
        tet_type = ((topo["tet_desc"] >> 1) & 7)[np.argsort(topo["tet_gen"])]
        k_dev, k_hyd = neohookean_stiffness(s.young_pa * METERS_PER_UNIT, s.poisson, topo["type_volume"], tet_type)
        self.tissue = Tissue(g, self.params, k_dev, k_hyd, g.positions, topology=topo)
        # LCS_Init's surface capacity and rest build (`native_backend.py:291-296`, dc_recon.cu:555-658)
        self.tri_capacity = 6 * len(g.surface_edges) + 1024 + 2 * len(g.edges)
        self.tissue.init_surface(tri_capacity=self.tri_capacity, qef_iters=8)
        self.legacy_grasp = int(s.contact_model) == 0
        if self.legacy_grasp:
            self.tissue.set_contact_model(0, legacy_grasp_params(g.spacing))
        self._tool = None
        self._blade = None
        self.command(self.initial_command)
        self.tissue.set_gravity((0.0, 0.0, 0.0))
        self._advance()
        self.neighbors, self.topology_revision = None, None
        self.tick = 0

    def command(self, command: Command):
        s = self.settings
        if not np.isfinite((*command.hinge, *command.blade_center, command.closure)).all():
            raise ValueError("Tool command contains non-finite values")
        if not 0 <= command.closure <= 1 or not 0 <= command.scissors_closure <= 1:
            raise ValueError("Jaw closure must be in [0,1]")
        hinge = np.array(command.hinge, dtype=np.float64)
        axis = np.array(s.jaw_axis, dtype=np.float64)
        self._tool = dict(shaftA=hinge - GRASPER_SHAFT_LENGTH * axis, shaftB=hinge, shaftR=GRASPER_SHAFT_RADIUS,
                          jawHinge=hinge, jawAxis=axis, jawMid=hinge + s.jaw_length / 2 * axis,
                          jawNormal=np.array(s.jaw_normal, dtype=np.float64), jawHalfWidth=s.jaw_width / 2,
                          jawLen=s.jaw_length, jawGap=s.jaw_open_gap * (1 - command.closure),
                          grasp01=command.closure)
        center = np.array(command.blade_center, dtype=np.float64)
        offset = self.blade_axis * (s.blade_length / 2)
        self._blade_next = (center - offset, center + offset) if command.blade_enabled else None
        self._command = command

    def _advance(self):
        """LCS_Step (the step and its frame update), the knife's LCS_DetectCut, then LCS_Finalize: the frame update
        and the surface rebuild, which fails like the reference's when its triangles overflow the capacity."""
        self.tissue.step(self.settings.dt, tool=self._tool, tool_armed=True)
        nxt = self._blade_next
        if nxt is not None and self._blade is not None:
            from .tissue import swept_cut_tool
            tool = swept_cut_tool(self._blade, nxt, 2 * self.settings.blade_radius)
            if tool["valid"]:
                self.tissue.detect_cut(tool)
        self._blade = nxt
        self.tissue.finalize()
        self.tissue.rebuild_surface()
        if self.tissue.surface_overflowed():
            raise RuntimeError("LCS_Finalize failed with native status -1")

    def step(self, command: Command):
        if self.closed:
            raise RuntimeError("Native session is closed")
        self.command(command)
        ramp = min(1.0, (self.tick + 1) * self.settings.dt / 0.5)
        self.tissue.set_gravity((0.0, -9.81 / METERS_PER_UNIT * ramp, 0.0))
        self._advance()
        self.tick += 1
        return self.observe()

    def detect_cut(self, cut_tool):
        """LCS_SetTool + LCS_DetectCut: cut the deformed lattice with a swept quad (see tissue.swept_cut_tool)."""
        if self.closed:
            raise RuntimeError("Native session is closed")
        self.tissue.detect_cut(cut_tool)

    def observe(self):
        t = self.tissue
        ob = t.observation(self.topology_revision, with_core=self.legacy_grasp)
        nodes, revision = ob["positions"], ob["revision"]
        if ob["neighbors"] is not None:
            neighbors = ob["neighbors"]
            neighbors.flags.writeable = False
            self.neighbors, self.topology_revision = neighbors, revision
        n = len(nodes)
        blade = np.zeros(10, dtype=np.float32)
        if self._blade is not None:
            blade[0:3], blade[3:6] = self._blade[0], self._blade[1]
            blade[6] = self.settings.blade_radius
        force, contacts = ob["reaction"]
        reaction = (force * METERS_PER_UNIT).astype(np.float32)
        hinge = np.array(self._command.hinge, dtype=np.float32)
        gap = float(self.settings.jaw_open_gap * (1 - self._command.closure))
        if not np.isfinite(nodes).all() or not np.isfinite(reaction).all():
            raise RuntimeError("Non-finite native state; simulation stopped")
        # LCS_GetGraspCoreFlags: the core corners of the legacy grasp; the friction model captures none
        grip = ob["core"] if self.legacy_grasp else np.zeros(n, dtype=np.int32)
        return Observation(nodes, self.neighbors, grip, hinge, gap, 0, 0.0, blade, reaction)

    def cut_edge_fractions(self):
        """First blade fraction of each directed lattice edge, f32[N, 6], -1 where uncut (`cut.cu:159-162`)."""
        return self.tissue.edge_cut_t.cpu().numpy().reshape(-1, 6)

    def compact_topology(self, apply=True):
        """Stable remap of the live tets (colour order), edges and nodes after cuts, plus their counts; with `apply`
        (the default) the dead tets and edges are removed from the device lists."""
        return self.tissue.compact(apply=apply)

    def material_quality(self):
        """LCS_GetMaterialQuality (`native_backend.py:463-466`): f32[2] = [the smallest det F over the live tets, at
        most 1; the number of live tets that are inverted (det F <= 0) or whose det F is not finite]."""
        if self.closed:
            raise RuntimeError("Native session is closed")
        return self.tissue.material_quality().cpu().numpy()

    def contact_depth_m(self):
        """LCS_SurfaceDepth (`native_backend.py:453-456` -> xpbd_surface_depth, xpbd_solver.cu:2046-2059): the deepest
        penetration of the grasper's shaft capsule and jaw boxes and, while the knife is enabled, of its edge capsule
        (queryBladeDepth) into the reconstructed surface, in metres; 0 before the first tool command (the tool is
        unarmed). The knife capsule sits at the commanded pose: the reference's pressure-blade body follows the command
        through its drive dynamics, which this path does not simulate, so the two agree once the body has caught up
        with the command."""
        if self.closed:
            raise RuntimeError("Native session is closed")
        if self._tool is None:
            return 0.0
        tool, kinds = dict(self._tool), 0b0111
        if self._blade is not None:
            tool.update(bladeA=self._blade[0], bladeB=self._blade[1], bladeR=self.settings.blade_radius)
            kinds |= 0b1000
        return float(self.tissue.surface_depth(tool, kinds=kinds).item()) * METERS_PER_UNIT

    def knife_body_depth_m(self):
        """LCS_KnifeBodyDepth (`native_backend.py:468-471` -> xpbd_knife_body_depth, pressure_blade.cuh:50-57) measures
        the knife's two back boxes on the pressure-blade body, a body this path does not simulate (its dynamics and
        shapes belong to the pressure blade). Rather than report another quantity under the reference's name, this
        raises."""
        if self.closed:
            raise RuntimeError("Native session is closed")
        raise NotImplementedError("knife_body_depth_m needs the pressure-blade body (knife back boxes), which this "
                                  "backend does not simulate; contact_depth_m includes the knife's edge capsule")

    def surface(self):
        """The surface LCS_Finalize last rebuilt (`native_backend.py:434-451`), with the reference's checks: an empty
        or ragged soup, the cut point capacity (4 per cut edge) and non-finite values raise RuntimeError."""
        if self.closed:
            raise RuntimeError("Native session is closed")
        s = self.tissue.surface()
        n = len(s["pos"])
        if n <= 0 or n % 3:
            raise RuntimeError("Native surface is empty or has an invalid triangle count")
        cut_points = int(self.tissue.counters[1].item()) & 0xffffffff
        if cut_points > 4 * len(self.grid.edges) or s["asked"] > 3 * self.tri_capacity:
            raise RuntimeError("Native cut/surface capacity exceeded")
        if not np.isfinite(s["pos"]).all() or not np.isfinite(s["nrm"]).all():
            raise RuntimeError("Non-finite reconstructed surface")
        return Surface(s["pos"], s["nrm"], s["indices"], s["aux"])

    def close(self):
        if not self.closed:
            self.tissue = None
            self.closed = True
            NativeTissue.active_session = False

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc, traceback):
        self.close()
