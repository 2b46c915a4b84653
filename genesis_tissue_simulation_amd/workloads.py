"""The benchmark workloads: lattices, materials, tools and schedules, all synthetic and seeded.

Units are the reference's native ones (cm, kg, s; gravity -9.81/0.01 cm/s^2, `native_backend.py:380`).

Material mapping. The reference's tet constraint is the coupled Neo-Hookean 2x2 block of `k_solve_tet`
(`xpbd/xpbd_solver.cu:749-829`) with deviatoric compliance 1/(mu V) and hydrostatic compliance 1/(kappa V). A
workload's "tet compliance" c is the deviatoric compliance of each tet, and its volume compliance is c/10 (the
kappa/mu ratio of Poisson's ratio 0.45, the reference's design value, `xpbd/xpbd_solver.h:62`). The 6144-tet block
states edge compliance 1e-7 and volume compliance 1e-8, i.e. c = 1e-7. Without the deviatoric row an axis-edge
lattice has no first-order shear stiffness and the block does not come to rest.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .grid import Ellipsoid, block_grid, block_nodes_for_tets, build_grid
from .params import TET_NEOHOOKEAN, SolverParams, volume_stiffness

GRAVITY = (0.0, -9.81 / 0.01, 0.0)
VOLUME_RATIO = 10.0


@dataclass
class Workload:
    name: str
    grid: object
    params: SolverParams
    dt: float
    steps: int
    compliance: np.ndarray          # per-tet deviatoric compliance, generation order
    initial_positions: np.ndarray
    gravity: tuple = GRAVITY
    blade: object = None            # callable step -> (S, E) segment, or None
    blade_thickness: float = 0.0
    tool: object = None             # callable step -> tool dict, or None
    gravity_ramp_s: float = 0.0     # gravity ramped linearly over this time, as NativeTissue.step does
    blades: list = field(default_factory=list)  # further blades (callable step -> (S, E)), same thickness
    info: dict = field(default_factory=dict)

    def gravity_at(self, step):
        """Gravity of step `step` (0-based): the reference's ramp min(1, (step + 1) dt / 0.5 s)
        (`native_backend.py:379-380`) when gravity_ramp_s is set, else constant."""
        if self.gravity_ramp_s <= 0:
            return self.gravity
        ramp = min(1.0, (step + 1) * self.dt / self.gravity_ramp_s)
        return tuple(g * ramp for g in self.gravity)

    def stiffness(self):
        """(k_dev, k_hyd) per tet in generation order: 1/c and 1/(c/10)."""
        c = np.asarray(self.compliance, dtype=np.float64)
        return volume_stiffness(c), volume_stiffness(c / VOLUME_RATIO)


def tet_count(grid):
    """Kuhn tets of the all-active cells of a lattice (six per cell)."""
    nx, ny, nz = (d + 1 for d in grid.dims)
    a = grid.active.reshape(nz, ny, nx) != 0
    cells = (a[:-1, :-1, :-1] & a[:-1, :-1, 1:] & a[:-1, 1:, :-1] & a[:-1, 1:, 1:] &
             a[1:, :-1, :-1] & a[1:, :-1, 1:] & a[1:, 1:, :-1] & a[1:, 1:, 1:])
    return 6 * int(cells.sum())


def jittered(grid, rng, sigma):
    """Lattice positions plus N(0, sigma) per coordinate on every free node; pinned nodes stay exact."""
    pos = grid.positions.astype(np.float64) + rng.normal(0.0, sigma, grid.positions.shape)
    pos = pos.astype(np.float32)
    pos[grid.pinned != 0] = grid.positions[grid.pinned != 0]
    return pos


def block_6144(seed=0, spacing=0.25, steps=200, substeps=10, iterations=1, precise=False):
    """The CPU-runnable block: 17x9x9 nodes, 6144 tets, edge compliance 1e-7, volume compliance 1e-8, bottom face
    pinned, dt = 1/60 with 10 substeps of 1 iteration, 200 steps; a planar blade of half-width 2 cells at mid-height
    sweeps +x at half a cell per step. Initial positions carry seeded N(0, 1e-4 cell) jitter. `precise`: the tet
    invariants formed from F - I instead of the reference's direct forms."""
    rng = np.random.default_rng(seed)
    grid = block_grid((17, 9, 9), spacing)
    L = spacing
    params = SolverParams(substeps=substeps, iterations=iterations, tet_mode=TET_NEOHOOKEAN, edge_ks=1e7,
                          precise_strain=precise)
    T = tet_count(grid)
    zc = 4 * L

    def blade(step):
        x = -L + 0.5 * L * step
        return (np.array((x, 4 * L, zc - 2 * L), np.float32), np.array((x, 4 * L, zc + 2 * L), np.float32))

    return Workload("block-6144", grid, params, 1 / 60, steps, np.full(T, 1e-7), jittered(grid, rng, 1e-4 * L),
                    blade=blade, blade_thickness=0.1 * L, info=dict(nodes=(17, 9, 9), spacing=L))


def ellipsoid_spacing_for(tets, half_extents=(6.0, 3.0, 2.5)):
    """Lattice spacing that puts about `tets` Kuhn tets in the ellipsoid's active cells."""
    lo, hi = 0.1, 0.7
    for _ in range(30):
        mid = 0.5 * (lo + hi)
        n = tet_count(build_grid(mid, shape=Ellipsoid(half_extents)))
        if n > tets:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def kidney_250k(seed=0, spacing=None, steps=300, substeps=20, iterations=2, tets=250_000, precise=False):
    """Kidney-scale single instance: an ellipsoid of semi-axes 6 x 3 x 2.5 cm rasterised at the spacing that gives
    about 250k tets (1.72 mm; a 0.5 mm lattice would give 9M); per-tet compliance log-normal (median 1e-7,
    sigma 0.5); a capsule grasper of radius 2 mm pressing 3 mm into the top; a blade on a seeded cubic Bezier over
    the top surface cutting for 300 steps; 20 substeps of 2 iterations (`precise`: the tet invariants formed from
    F - I instead of the reference's direct forms). Gravity ramps in over 0.5 s as in
    NativeTissue.step: applied at once, the 30-cell-tall lattice falls onto its pinned bed faster than two
    Gauss-Seidel sweeps per substep can carry the support up, and the bottom tets invert (in FP64 as in FP32)."""
    rng = np.random.default_rng(seed)
    shape = Ellipsoid((6.0, 3.0, 2.5))
    h = spacing if spacing is not None else 0.1716
    grid = build_grid(h, shape=shape)
    T = tet_count(grid)
    compliance = np.exp(rng.normal(np.log(1e-7), 0.5, T))
    params = SolverParams(substeps=substeps, iterations=iterations, tet_mode=TET_NEOHOOKEAN, edge_ks=1e7,
                          contact_enabled=True, contact_radius=0.0, jaw_thickness=0.1, precise_strain=precise)
    top = 3.0
    # grasper capsule along x above the top, lowered 3 mm into the surface over the first 60 steps
    r = 0.2

    def tool(step):
        depth = 0.3 * min(1.0, step / 60.0)
        y = top + r - depth
        return dict(shaftA=(-1.5, y, 0.0), shaftB=(1.5, y, 0.0), shaftR=r, jawAxis=(0, 0, 1), jawNormal=(1, 0, 0))

    # blade: cubic Bezier with 4 seeded control points on the top surface; a vertical segment 0.6 cm tall
    u = rng.uniform(-0.7, 0.7, size=(4, 2))
    ctrl = np.stack((u[:, 0] * 6.0, np.zeros(4), u[:, 1] * 2.5), axis=1)
    ctrl[:, 0] = np.sort(ctrl[:, 0])
    for c in ctrl:
        c[1] = 3.0 * np.sqrt(max(0.0, 1 - (c[0] / 6.0) ** 2 - (c[2] / 2.5) ** 2))

    def blade(step):
        s = min(1.0, step / steps)
        b = ((1 - s) ** 3) * ctrl[0] + 3 * ((1 - s) ** 2) * s * ctrl[1] + 3 * (1 - s) * s * s * ctrl[2] + s ** 3 * ctrl[3]
        return (np.array(b + (0, 0.3, 0), np.float32), np.array(b - (0, 0.6, 0), np.float32))

    return Workload("kidney-250k", grid, params, 1 / 30, steps, compliance, jittered(grid, rng, 0.0),
                    blade=blade, blade_thickness=0.02, tool=tool, gravity_ramp_s=0.5,
                    info=dict(spacing=h, half_extents=(6.0, 3.0, 2.5)))


def cutting_block(nodes=(112, 56, 56), spacing=0.25, seed=0, steps=500, substeps=20, iterations=2, n_blades=3,
                  depth_fraction=0.35):
    """The large continuous-cutting block: 112 x 56 x 56 nodes (111 x 55 x 55 cells, 2.01M Kuhn tets; the 161 x 81 x 81
    lattice of the same description would hold 6.14M), tet compliance 1e-7 (volume 1e-8) and edge compliance 1e-7,
    bottom face pinned, gravity ramped over 0.5 s, 20 substeps of 2 iterations at dt = 1/60. Three seeded vertical
    blades of depth 0.35 x the height cross the top face along seeded straight paths over `steps` steps, cutting
    every step; the blocks are compacted every step by the caller."""
    rng = np.random.default_rng(seed)
    grid = block_grid(nodes, spacing)
    T = tet_count(grid)
    params = SolverParams(substeps=substeps, iterations=iterations, tet_mode=TET_NEOHOOKEAN, edge_ks=1e7)
    lo = grid.positions.min(axis=0).astype(np.float64)
    hi = grid.positions.max(axis=0).astype(np.float64)
    top, depth = hi[1] + 0.5 * spacing, depth_fraction * (hi[1] - lo[1])
    paths = []
    for _ in range(n_blades):
        a = lo[[0, 2]] + rng.uniform(0.05, 0.95, 2) * (hi - lo)[[0, 2]]
        b = lo[[0, 2]] + rng.uniform(0.05, 0.95, 2) * (hi - lo)[[0, 2]]
        paths.append((a, b))

    def blade_of(path):
        a, b = path

        def blade(step):
            p = a + (b - a) * min(1.0, step / steps)
            return (np.array((p[0], top, p[1]), np.float32), np.array((p[0], hi[1] - depth, p[1]), np.float32))
        return blade

    blades = [blade_of(p) for p in paths]
    return Workload("cutting-block-2m", grid, params, 1 / 60, steps, np.full(T, 1e-7),
                    jittered(grid, rng, 1e-4 * spacing), blade=blades[0], blade_thickness=0.02,
                    gravity_ramp_s=0.5, blades=blades[1:], info=dict(nodes=nodes, spacing=spacing, paths=paths))


def all_blades(w):
    """Every blade of a workload: the first, then the further ones."""
    return ([w.blade] if w.blade else []) + list(w.blades)


@dataclass
class AdjointInstance:
    grid: object
    params: SolverParams
    dt: float
    initial_positions: np.ndarray
    force: np.ndarray            # [N, 3] gravity * mass plus the point loads (native force units)
    mask: np.ndarray             # [N] 1 on the surface nodes the loss sees
    compliance_hidden: np.ndarray
    compliance_init: np.ndarray


def adjoint_instance(nodes=None, tets=20_000, seed=0, spacing=1.0, substeps=50, edge_compliance=1e-7, iterations=1):
    """One block of the compliance fit: hidden per-tet compliance log-normal (median 1e-7, sigma 0.7), point loads
    N(0, 0.5) N per component on 5% of the top nodes, loss on the free surface nodes, learnable compliance initialised
    at the median. The rollout's 50 substeps are one 1/60 s frame (substep 1/3000 s); at ten substeps per frame the
    one-iteration solver cannot hold a 0.5 N load on a 1 g node and the rollout is chaotic (FP32 and FP64 end
    positions differ by 8%)."""
    rng = np.random.default_rng(seed)
    grid = block_grid(nodes or block_nodes_for_tets(tets), spacing)
    T = tet_count(grid)
    hidden = np.exp(rng.normal(np.log(1e-7), 0.7, T))
    params = SolverParams(substeps=substeps, iterations=iterations, tet_mode=TET_NEOHOOKEAN,
                          edge_ks=1.0 / edge_compliance)
    ijk = grid.positions / grid.cell_size
    top = np.flatnonzero(np.isclose(ijk[:, 1], ijk[:, 1].max()))
    loaded = rng.choice(top, max(1, int(round(0.05 * len(top)))), replace=False)
    force = (np.asarray(GRAVITY, dtype=np.float64)[None, :] * grid.mass[:, None].astype(np.float64))
    force[loaded] += rng.normal(0.0, 0.5, (len(loaded), 3)) / 0.01
    lo, hi = ijk.min(axis=0), ijk.max(axis=0)
    surface = (np.isclose(ijk, lo) | np.isclose(ijk, hi)).any(axis=1) & (grid.pinned == 0)
    return AdjointInstance(grid, params, 1 / 60, grid.positions.copy(), force.astype(np.float32),
                           surface.astype(np.float32), hidden, np.full(T, 1e-7))


@dataclass
class BatchWorkload:
    name: str
    instances: list        # (grid, params, k_dev, k_hyd, initial_positions) per instance
    cuts: list             # one swept cut per instance (cut descriptor dict)
    params: SolverParams
    dt: float
    gravity: tuple = GRAVITY
    info: dict = field(default_factory=dict)


def straight_cut(grid, rng, depth_fraction=0.4, thickness=0.02):
    """A seeded straight cut over the top face of a block: start point and direction uniform over the face, a
    vertical blade reaching `depth_fraction` of the block height, swept across the face in one stroke."""
    from .tissue import swept_cut_tool
    lo = grid.positions.min(axis=0).astype(np.float64)
    hi = grid.positions.max(axis=0).astype(np.float64)
    start = lo[[0, 2]] + rng.uniform(0, 1, 2) * (hi - lo)[[0, 2]]
    angle = rng.uniform(0, 2 * np.pi)
    length = float(np.linalg.norm((hi - lo)[[0, 2]]))
    end = start + length * np.array((np.cos(angle), np.sin(angle)))
    top, depth = hi[1] + 0.5 * grid.cell_size[1], (hi[1] - lo[1]) * depth_fraction
    seg = lambda p: ((p[0], top, p[1]), (p[0], hi[1] - depth, p[1]))
    return swept_cut_tool(seg(start), seg(end), thickness)


def batched_blocks(count=1024, seed=0, tmin=2000, tmax=100_000, spacing=0.25, substeps=10, iterations=1,
                   precise=False):
    """Many independent blocks: tet counts log-uniform in [tmin, tmax], 2:1:1 cells, per-instance compliance
    log-uniform in [1e-8, 1e-6] (edges and tets), one seeded straight cut per instance across the top face; a
    10-substep step of one iteration at dt = 1/60 (`precise`: the tet invariants formed from F - I)."""
    rng = np.random.default_rng(seed)
    base = SolverParams(substeps=substeps, iterations=iterations, tet_mode=TET_NEOHOOKEAN, edge_ks=1e7,
                        precise_strain=precise)
    instances, cuts = [], []
    for _ in range(count):
        t = float(np.exp(rng.uniform(np.log(tmin), np.log(tmax))))
        grid = block_grid(block_nodes_for_tets(t), spacing)
        c = float(np.exp(rng.uniform(np.log(1e-8), np.log(1e-6))))
        T = tet_count(grid)
        p = base.with_(edge_ks=float(np.float32(1.0 / c)))
        instances.append((grid, p, volume_stiffness(np.full(T, c)), volume_stiffness(np.full(T, c / VOLUME_RATIO)),
                          jittered(grid, rng, 1e-4 * spacing)))
        cuts.append(straight_cut(grid, rng))
    return BatchWorkload("batched-1024", instances, cuts, base, 1 / 60,
                         info=dict(count=count, tmin=tmin, tmax=tmax, spacing=spacing))
