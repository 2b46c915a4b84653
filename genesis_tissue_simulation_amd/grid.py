"""Voxel lattice of a tissue: node positions, flags, lumped mass, six-neighbour links and cut edges.

This is the host-side input generator of the XPBD substep. It restates the behaviour of the reference lattice
builder (`endoscopic_tissue/grid.py:176-281`, `TissueGrid` at `:41-61`) so that a grid built here can be handed to
`NativeTissue` exactly as the reference's is, and adds the regular no-shell blocks and the ellipsoid the benchmark
workloads are defined on.

Conventions (shared with the reference):
  * native units are cm, y is up; node id = x + nx * (y + ny * z) (`grid.py:64-66`);
  * neighbour slots 0..5 are -x, +x, -y, +y, -z, +z (`xpbd_solver.h:132-133`);
  * the eight voxel corners follow the figure order of `CORNER_OFFSETS` below; the Kuhn tetrahedra use the bit order
    `i = dx | dy << 1 | dz << 2` (`xpbd_solver.cu:1624`).
"""
from __future__ import annotations

from dataclasses import dataclass
from itertools import combinations

import numpy as np

# Voxel corner order of the dual-contouring tables (`grid.py:11-13`).
CORNER_OFFSETS = np.array(((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                           (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)), dtype=np.int32)
# The twelve voxel edges as corner pairs in that order (`grid.py:14-16`).
EDGE_CORNERS = np.array(((0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6),
                         (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)), dtype=np.int32)
# Kuhn split of a cell around its main diagonal 0-7, corners in bit order (`xpbd_solver.cu:1624`).
KUHN_TETS = np.array(((0, 1, 3, 7), (0, 3, 2, 7), (0, 2, 6, 7),
                      (0, 6, 4, 7), (0, 4, 5, 7), (0, 5, 1, 7)), dtype=np.int32)
# Figure-order corner for each bit-order corner index.
BIT_TO_FIGURE = np.array((0, 1, 3, 2, 4, 5, 7, 6), dtype=np.int32)
ISECT_DTYPE = np.dtype([("position", "<f4", 3), ("normal", "<f4", 3),
                        ("a", "<i4"), ("b", "<i4"), ("t", "<f4"), ("inside_a", "<i4")])
BEND_DTYPE = np.dtype([("i", "<i4"), ("j", "<i4"), ("k", "<i4"), ("theta", "<f4"),
                       ("alive", "<i4"), ("padding", "<i4", 3)])
# CutPointGpu (`cut.cuh:37-46`), 32 bytes: owner-frame offset on the rest edge, owner node, cut edge, owner-frame normal.
CUT_POINT_DTYPE = np.dtype([("offset", "<f4", 3), ("owner", "<i4"), ("edge", "<i4"), ("normal", "<f4", 3)])
METERS_PER_UNIT = 0.01
BLOCK_SIZE = (6.0, 1.4, 4.0)
# The 15 unordered slot pairs of a node's bend elements, in lexicographic order (`grid.py:257-260`).
BEND_SLOT_PAIRS = np.array(list(combinations(range(6), 2)), dtype=np.int32)


@dataclass
class TissueGrid:
    """Field-for-field the reference lattice record (`grid.py:41-61`)."""
    spacing: float
    cell_size: np.ndarray
    dims: tuple
    origin: np.ndarray
    positions: np.ndarray
    inside: np.ndarray
    active: np.ndarray
    pinned: np.ndarray
    mass: np.ndarray
    neighbors: np.ndarray
    rest_neighbors: np.ndarray
    bends: np.ndarray
    corners: np.ndarray
    occupied: np.ndarray
    intersections: np.ndarray
    offsets: np.ndarray
    counts: np.ndarray
    surface_edges: np.ndarray
    edges: np.ndarray
    connectivity: np.ndarray

    @property
    def node_dims(self):
        return tuple(int(d) + 1 for d in self.dims)

    def edge_skin_fraction(self):
        """Directed-edge surface crossing fraction, f32[6N], -1 where a lattice edge does not cross the skin.

        The reference uploads the same table in `dc_recon.cu:593-599`: slot(a->b) holds t and slot(b->a) 1-t.
        """
        skin = np.full(6 * len(self.positions), -1.0, dtype=np.float32)
        if len(self.intersections):
            a = self.intersections["a"].astype(np.int64)
            b = self.intersections["b"].astype(np.int64)
            t = self.intersections["t"].astype(np.float32)
            skin[directed_slot(a, b, self.dims)] = t
            skin[directed_slot(b, a, self.dims)] = np.float32(1) - t
        return skin


def directed_slot(a, b, dims):
    """Slot of the directed lattice edge a->b in a [6N] table (`cut.cuh:25-28`)."""
    d = np.asarray(b) - np.asarray(a)
    ad = np.abs(d)
    axis = np.where(ad == 1, 0, np.where(ad == dims[0] + 1, 1, 2))
    return 6 * np.asarray(a) + 2 * axis + (d > 0)


def lattice_coordinates(node_dims):
    """Integer (x, y, z) of every lattice point, x fastest (`grid.py:64-66`)."""
    nx, ny, nz = node_dims
    z, y, x = np.meshgrid(np.arange(nz, dtype=np.int32), np.arange(ny, dtype=np.int32),
                          np.arange(nx, dtype=np.int32), indexing="ij")
    return np.stack((x.ravel(), y.ravel(), z.ravel()), axis=1)


def connectivity_table():
    """For each of the 4096 cut masks of a voxel's 12 edges: component count, then each corner's component id.

    Components are numbered in the order of their lowest corner (`grid.py:69-82`). Computed here by union-find over
    the uncut edges of every mask.
    """
    table = np.zeros((4096, 9), dtype=np.int32)
    for mask in range(4096):
        parent = list(range(8))

        def root(v):
            while parent[v] != v:
                v = parent[v]
            return v
        for bit, (a, b) in enumerate(EDGE_CORNERS):
            if not mask >> bit & 1:
                ra, rb = root(int(a)), root(int(b))
                if ra != rb:
                    parent[max(ra, rb)] = min(ra, rb)
        roots = [root(v) for v in range(8)]
        ids = {}
        for v in range(8):
            ids.setdefault(roots[v], len(ids))
        table[mask, 0] = len(ids)
        table[mask, 1:] = [ids[r] for r in roots]
    return table


@dataclass(frozen=True)
class Tab:
    """Box grown from a positive face of the block (`grid.py:27-38`)."""
    low: tuple
    high: tuple
    attached: int


def _exit_along_axis(points, axis, direction, lows, highs):
    """Where rays from interior points along +-axis leave a union of open boxes (`grid.py:85-106`).

    Boxes whose open cross-section contains the ray and that overlap along it chain into one interval, so the exit is
    the far end of the chain that contains the start.
    """
    n = len(points)
    exits = np.empty(n, dtype=np.float64)
    for r in range(n):
        p = points[r]
        s = float(direction[r]) * float(p[axis])
        spans = []
        for lo, hi in zip(lows, highs):
            others = [k for k in range(3) if k != axis]
            if all(lo[k] < p[k] < hi[k] for k in others):
                if direction[r] > 0:
                    spans.append((float(lo[axis]), float(hi[axis])))
                else:
                    spans.append((-float(hi[axis]), -float(lo[axis])))
        reach = -np.inf
        for near, far in spans:
            if near < s < far:
                reach = max(reach, far)
        for _ in range(len(spans)):
            for near, far in spans:
                if near < reach < far:
                    reach = max(reach, far)
        if not np.isfinite(reach):
            raise AssertionError("an interior point lies outside every box")
        exits[r] = direction[r] * reach
    return exits


@dataclass(frozen=True)
class BoxUnion:
    """Axis-aligned boxes in cm: the block centred on the origin plus an optional tab (`grid.py:112-160`)."""
    lows: np.ndarray
    highs: np.ndarray
    half: np.ndarray

    @property
    def low(self):
        return self.lows.min(axis=0)

    @property
    def high(self):
        return self.highs.max(axis=0)

    def inside(self, positions):
        p = positions[:, None, :]
        return ((self.lows < p) & (p < self.highs)).all(axis=2).any(axis=1)

    def boundary(self, positions, a, b, inside):
        axis = np.abs(positions[b] - positions[a]).argmax(axis=1)
        interior = np.where(inside[a] != 0, a, b)
        exterior = np.where(inside[a] != 0, b, a)
        rows = np.arange(len(a))
        direction = np.sign(positions[exterior, axis] - positions[interior, axis])
        starts = positions[interior]
        exit_coord = np.array([_exit_along_axis(starts[r:r + 1], int(axis[r]), direction[r:r + 1],
                                                self.lows, self.highs)[0] for r in rows])
        t = (exit_coord - positions[a, axis]) / (positions[b, axis] - positions[a, axis])
        normals = np.zeros((len(a), 3), dtype=np.float32)
        normals[rows, axis] = direction
        return t, normals

    def bed(self, positions):
        return positions[:, 1] <= -self.half[1]


def implicit_boundary(distance, positions, a, b):
    """Zero crossing of a signed distance on each lattice edge a->b by 40 bisections, plus the outward normal from a
    central-difference gradient (`grid.py:163-173`)."""
    pa, pb = positions[a].astype(np.float64), positions[b].astype(np.float64)
    lo, hi = np.zeros(len(a)), np.ones(len(a))
    sign_a = np.sign(distance(pa))
    for _ in range(40):
        mid = (lo + hi) / 2
        same = np.sign(distance(pa + (pb - pa) * mid[:, None])) == sign_a
        lo, hi = np.where(same, mid, lo), np.where(same, hi, mid)
    t = (lo + hi) / 2
    hit = pa + (pb - pa) * t[:, None]
    step = 1e-4
    gradient = np.stack([distance(hit + step * e) - distance(hit - step * e) for e in np.eye(3)], axis=1)
    return t, (gradient / np.linalg.norm(gradient, axis=1, keepdims=True)).astype(np.float32)


def ellipsoid_distance(points, half_extents):
    """Signed distance bound to an origin-centred ellipsoid, exact on its surface (`kidney.py:15-20`)."""
    k0 = np.linalg.norm(points / half_extents, axis=1)
    k1 = np.maximum(np.linalg.norm(points / np.square(half_extents), axis=1), 1e-12)
    return k0 * (k0 - 1) / k1


@dataclass(frozen=True)
class Ellipsoid:
    """Ellipsoidal organ of `half_extents` cm centred on the origin, resting on a bed: nodes below `bed_height` are
    pinned. This is the kidney-scale benchmark body; the reference's kidney (`kidney.py:23-81`) is an ellipsoid plus
    a tumour lobe over the same implicit-shape protocol."""
    half_extents: tuple = (6.0, 3.0, 2.5)
    bed_fraction: float = 0.15

    @property
    def low(self):
        return -np.array(self.half_extents, dtype=np.float32)

    @property
    def high(self):
        return np.array(self.half_extents, dtype=np.float32)

    @property
    def bed_height(self):
        return -self.half_extents[1] * (1 - 2 * self.bed_fraction)

    def distance(self, points):
        return ellipsoid_distance(points, np.array(self.half_extents, dtype=np.float64))

    def inside(self, positions):
        return self.distance(positions.astype(np.float64)) < 0

    def boundary(self, positions, a, b, inside):
        return implicit_boundary(self.distance, positions, a, b)

    def bed(self, positions):
        return positions[:, 1] < self.bed_height


def _assemble(spacing, cell_size, dims, origin, positions, inside, bed_mask, density, shape=None,
              keep_all_active=False):
    """Everything of a TissueGrid that follows from the lattice geometry and the inside flags."""
    nc = tuple(int(d) + 1 for d in dims)
    ijk = lattice_coordinates(nc)
    strides = np.array((1, nc[0], nc[0] * nc[1]), dtype=np.int64)
    voxel_ijk = lattice_coordinates(tuple(int(d) for d in dims))
    corners = ((voxel_ijk[:, None, :].astype(np.int64) + CORNER_OFFSETS) @ strides).astype(np.int32)
    occupied = inside[corners].any(axis=1).astype(np.int32)
    active = np.zeros(len(positions), dtype=np.int32)
    active[corners[occupied != 0]] = 1
    if keep_all_active:
        active[:] = 1
    pinned = (bed_mask & (active != 0)).astype(np.int32)

    # Voxel edges whose ends disagree on inside carry a surface crossing (`grid.py:224-236`).
    ca = corners[:, EDGE_CORNERS[:, 0]]
    cb = corners[:, EDGE_CORNERS[:, 1]]
    crossing = inside[ca] != inside[cb]
    counts = crossing.sum(axis=1, dtype=np.int32)
    offsets = (np.cumsum(counts, dtype=np.int32) - counts).astype(np.int32)
    a, b = ca[crossing], cb[crossing]
    intersections = np.empty(len(a), dtype=ISECT_DTYPE)
    if len(a):
        t, normals = shape.boundary(positions, a, b, inside)
        hit = positions[a] + (positions[b] - positions[a]) * t[:, None]
        intersections["position"], intersections["normal"] = hit, normals
        intersections["a"], intersections["b"] = a, b
        intersections["t"], intersections["inside_a"] = t, inside[a]

    neighbors = np.full((len(positions), 6), -1, dtype=np.int32)
    rest = np.zeros((len(positions), 6, 3), dtype=np.float32)
    edge_parts, surface_parts = [], []
    for axis in range(3):
        below = ijk[:, axis] < dims[axis]
        lower = np.flatnonzero(below)
        upper = lower + strides[axis]
        linked = (active[lower] != 0) & (active[upper] != 0)
        neighbors[lower[linked], 2 * axis + 1] = upper[linked]
        neighbors[upper[linked], 2 * axis] = lower[linked]
        rest[lower[linked], 2 * axis + 1, axis] = cell_size[axis]
        rest[upper[linked], 2 * axis, axis] = -cell_size[axis]
        # Cut edges need all four incident voxels in bounds (`grid.py:249-255`, `cut.cuh:140-144`).
        interior = below.copy()
        for other in range(3):
            if other != axis:
                interior &= (ijk[:, other] > 0) & (ijk[:, other] < dims[other])
        base = np.flatnonzero(interior)
        edge = np.column_stack((np.full(len(base), axis), ijk[base])).astype(np.int32)
        edge_parts.append(edge)
        cross = inside[base] != inside[base + strides[axis]]
        surface = np.zeros((int(np.count_nonzero(cross)), 8), dtype=np.int32)
        surface[:, :4] = edge[cross]
        surface[:, 4] = inside[base[cross]]
        surface_parts.append(surface)

    bends = np.zeros((len(positions), 15), dtype=BEND_DTYPE)
    bends["i"] = np.arange(len(positions), dtype=np.int32)[:, None]
    bends["j"] = neighbors[:, BEND_SLOT_PAIRS[:, 0]]
    bends["k"] = neighbors[:, BEND_SLOT_PAIRS[:, 1]]
    bends["alive"] = (bends["j"] >= 0) & (bends["k"] >= 0)
    straight = BEND_SLOT_PAIRS[:, 0] // 2 == BEND_SLOT_PAIRS[:, 1] // 2
    bends["theta"] = np.where(straight, np.pi, np.pi / 2)

    # Each Kuhn tet of a supported cell lumps a quarter of its mass (cell/6) on each corner: rho*h^3/24 per
    # tet corner (`grid.py:269-277`).
    bit_corners = corners[:, BIT_TO_FIGURE]
    supported = np.all(active[bit_corners] != 0, axis=1)
    tet_nodes = bit_corners[supported][:, KUHN_TETS].reshape(-1)
    mass = (np.bincount(tet_nodes, minlength=len(positions)) * density *
            np.prod(cell_size * METERS_PER_UNIT) / 24).astype(np.float32)
    if not np.all(mass[active != 0] > 0):
        raise ValueError("an active node belongs to no supported cell")
    return TissueGrid(spacing, cell_size, tuple(int(d) for d in dims), origin, positions, inside, active, pinned, mass,
                      neighbors, rest, bends, corners, occupied, intersections, offsets, counts,
                      np.concatenate(surface_parts), np.concatenate(edge_parts), connectivity_table_cached())


_CONNECTIVITY = None


def connectivity_table_cached():
    global _CONNECTIVITY
    if _CONNECTIVITY is None:
        _CONNECTIVITY = connectivity_table()
    return _CONNECTIVITY.copy()


def build_grid(spacing=0.6, size=BLOCK_SIZE, density=1000.0, vertical_spacing=None, tab=None, shape=None):
    """Lattice of the block of `size` (with an optional `tab`) or of an implicit `shape`, in a two-cell air shell.

    Same contract as the reference builder (`grid.py:176-281`): a shape gives `low`, `high`, `inside(positions)`,
    `boundary(positions, a, b, inside)` and `bed(positions)`.
    """
    if not np.isfinite(spacing) or not 0.1 <= spacing <= 0.7:
        raise ValueError("spacing must be 0.1..0.7 native units (1..7 mm)")
    cell_size = np.full(3, spacing, dtype=np.float32)
    if vertical_spacing is not None:
        if not np.isfinite(vertical_spacing) or not .01 <= vertical_spacing <= spacing:
            raise ValueError("vertical spacing must be 0.01..spacing native units")
        cell_size[1] = vertical_spacing
    if shape is None:
        half = np.array(size, dtype=np.float32) / 2
        lows, highs = [-half], [half]
        if tab is not None:
            low, high = np.array(tab.low, dtype=np.float32), np.array(tab.high, dtype=np.float32)
            if not (np.isfinite((low, high)).all() and tab.attached in (0, 1, 2)
                    and np.all(high - low >= 2 * cell_size - 1e-6) and abs(low[tab.attached] - half[tab.attached]) < 1e-6):
                raise ValueError("tab must span two cells per axis and grow from a positive face of the block")
            for bound, face in ((low, -half), (low, half), (high, -half), (high, half)):
                flush = np.abs(bound - face) < 1e-6
                bound[flush] = face[flush]
            low[tab.attached] = half[tab.attached] - cell_size[tab.attached]
            lows.append(low)
            highs.append(high)
        shape = BoxUnion(np.array(lows), np.array(highs), half)
    low, high = shape.low, shape.high
    dims = tuple(int(v) for v in np.ceil((high - low) / cell_size).astype(int) + 4)
    origin = (low + high) / 2 - np.array(dims, dtype=np.float32) * cell_size / 2
    nc = tuple(d + 1 for d in dims)
    ijk = lattice_coordinates(nc)
    positions = np.ascontiguousarray(origin + ijk * cell_size, dtype=np.float32)
    inside = shape.inside(positions).astype(np.int32)
    return _assemble(spacing, cell_size, dims, origin, positions, inside, shape.bed(positions), density, shape)


def block_grid(nodes=(17, 9, 9), spacing=0.25, density=1000.0, origin=(0.0, 0.0, 0.0)):
    """A solid block of `nodes` lattice points with no air shell: every node active and inside, the y = 0 layer
    pinned. Spacing is not limited to the reference builder's 1..7 mm range."""
    nodes = tuple(int(n) for n in nodes)
    if min(nodes) < 2:
        raise ValueError("a block needs at least two nodes per axis")
    cell_size = np.full(3, spacing, dtype=np.float32)
    dims = tuple(n - 1 for n in nodes)
    origin = np.array(origin, dtype=np.float32)
    ijk = lattice_coordinates(nodes)
    positions = np.ascontiguousarray(origin + ijk * cell_size, dtype=np.float32)
    inside = np.ones(len(positions), dtype=np.int32)
    return _assemble(float(spacing), cell_size, dims, origin, positions, inside, ijk[:, 1] == 0, density)


def block_nodes_for_tets(tets, aspect=(2, 1, 1)):
    """Node dims of a block of about `tets` Kuhn tets (6 per cell) with the given cell aspect ratio."""
    cells = tets / 6.0
    unit = (cells / float(np.prod(aspect))) ** (1 / 3)
    dims = [max(1, int(round(a * unit))) for a in aspect]
    return tuple(d + 1 for d in dims)


def to_genesis(points):
    """Native (x, y, z) cm to Genesis (x, -z, y) m (`grid.py:284-285`)."""
    rotation = np.array(((1, 0, 0), (0, 0, -1), (0, 1, 0)), dtype=np.float32)
    return points @ rotation.T * METERS_PER_UNIT
