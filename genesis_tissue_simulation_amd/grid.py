"""Voxel lattice of a tissue: node positions, flags, lumped mass, six-neighbour links and cut edges.

The host-side input of the XPBD substep. `build_grid` takes the place of the reference lattice builder
(`endoscopic_tissue/grid.py:176-281`; the record `TissueGrid` at `:41-61`): a grid built here can be handed to
`NativeTissue` as the reference's is. `block_grid` and `Ellipsoid` add the shell-free blocks and the ellipsoidal organ
the benchmark workloads are defined on.

The lattice is handled as a 3-D array of nodes (z, y, x; x fastest in the flat numbering) and a 3-D array of voxels;
occupancy, activity, links, cut edges and the lumped mass are shifted-slice reductions over those arrays.

Formats kept from the reference, because other code reads them (paths relative to endoscopic_tissue/):
  * native units are cm, y is up; node id = x + nx * (y + ny * z) (`grid.py:64-66`);
  * neighbour slots 0..5 are -x, +x, -y, +y, -z, +z (`native/vendor/cuda/xpbd/xpbd_solver.h:132-133`);
  * voxel corners in the figure order of CORNER_OFFSETS and voxel edges in EDGE_CORNERS order (`grid.py:11-16`), the
    order of the dual-contouring tables; the Kuhn tets use the bit order i = dx | dy << 1 | dz << 2
    (`xpbd_solver.cu:1624`);
  * the 40-byte intersection, 32-byte bend and 32-byte cut-point records; cut edges listed axis by axis in node order,
    intersections voxel by voxel in edge order;
  * the lattice placement: the tissue's bounds rounded up to whole cells, two air cells added on every side, centred.
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
# Air cells kept around the tissue's bounding box on every side.
SHELL_CELLS = 2
# How many of a cell's six Kuhn tets hold each bit-order corner: the diagonal ends 0 and 7 are in all six, every other
# corner in two. Each tet lumps a quarter of its mass (the cell's sixth) on each of its corners.
KUHN_MULTIPLICITY = np.bincount(KUHN_TETS.ravel(), minlength=8).astype(np.int64)


@dataclass
class TissueGrid:
    """The lattice record, field for field the reference's (`grid.py:41-61`)."""
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


_CONNECTIVITY = None


def connectivity_table_cached():
    global _CONNECTIVITY
    if _CONNECTIVITY is None:
        _CONNECTIVITY = connectivity_table()
    return _CONNECTIVITY.copy()


# ------------------------------------------------------------------------------------------------------ tissue shapes
# A shape gives its bounds `low` and `high`, `inside(positions)`, `boundary(positions, a, b, inside)` (for lattice
# edges a -> b with one end inside: the fraction along a -> b where the edge leaves the tissue, and the outward unit
# normal there) and `bed(positions)`, the nodes pinned where they are active -- the protocol of `grid.py:176-186`.

@dataclass(frozen=True)
class Tab:
    """Box of tissue grown from a positive face of the block, in cm (`grid.py:27-38`): corners `low` and `high`,
    its `low` face along axis `attached` on the block's positive face of that axis."""
    low: tuple
    high: tuple
    attached: int


@dataclass(frozen=True)
class BoxUnion:
    """Tissue made of axis-aligned open boxes in cm: the block centred on the origin, then an optional tab; the bed is
    the block's bottom face (`half` is the block's half extent)."""
    lows: np.ndarray
    highs: np.ndarray
    half: np.ndarray

    @property
    def low(self):
        return self.lows.min(axis=0)

    @property
    def high(self):
        return self.highs.max(axis=0)

    def _contains(self, points, box):
        return np.all((self.lows[box] < points) & (points < self.highs[box]), axis=-1)

    def inside(self, positions):
        result = np.zeros(len(positions), dtype=bool)
        for box in range(len(self.lows)):
            result |= self._contains(positions, box)
        return result

    def boundary(self, positions, a, b, inside):
        """Each edge runs along one lattice axis from its inside end. Along that line the union is open, so the first
        point where it leaves the tissue is a box bound: the nearest bound ahead of the inside end at which the line is
        in no box's interior. The normal is the edge axis, pointing out."""
        a, b = np.asarray(a), np.asarray(b)
        axis = np.abs(positions[b] - positions[a]).argmax(axis=1)
        inner = np.where(inside[a] != 0, a, b)
        outer = np.where(inside[a] != 0, b, a)
        sign = np.sign(positions[outer, axis] - positions[inner, axis])
        exit = np.empty(len(a), dtype=np.float32)
        for ax in range(3):
            rows = np.flatnonzero(axis == ax)
            if len(rows) == 0:
                continue
            start = positions[inner[rows]]
            bounds = np.unique(np.concatenate((self.lows[:, ax], self.highs[:, ax]))).astype(np.float32)
            others = [k for k in range(3) if k != ax]
            # the line's cross-section lies in box j's open face; bound c lies in box j's open span along the axis
            section = np.stack([np.all((self.lows[j, others] < start[:, others]) &
                                       (start[:, others] < self.highs[j, others]), axis=1)
                                for j in range(len(self.lows))], axis=1)
            span = (self.lows[:, ax, None] < bounds[None, :]) & (bounds[None, :] < self.highs[:, ax, None])
            covered = (section[:, :, None] & span[None, :, :]).any(axis=1)
            s = sign[rows, None]
            ahead = s * bounds[None, :] > s * start[:, ax, None]
            key = np.where(ahead & ~covered, s * bounds[None, :], np.inf)
            if not np.isfinite(key.min(axis=1)).all():
                raise AssertionError("an inside node lies in no box")
            exit[rows] = bounds[key.argmin(axis=1)]
        pa, pb = positions[a, axis], positions[b, axis]
        t = (exit - pa) / (pb - pa)
        normals = np.zeros((len(a), 3), dtype=np.float32)
        normals[np.arange(len(a)), axis] = sign
        return t, normals

    def bed(self, positions):
        return positions[:, 1] <= -self.half[1]


def implicit_boundary(distance, positions, a, b):
    """Crossing of a signed distance (negative inside) on each lattice edge a -> b, for shapes given only by their
    distance: a bracketed Illinois (modified regula falsi) iteration to the float64 resolution of the fraction, then
    the outward unit normal from central differences of the distance over a thousandth of the edge."""
    pa, pb = positions[a].astype(np.float64), positions[b].astype(np.float64)
    d = pb - pa
    f = lambda t: distance(pa + d * t[:, None])
    lo, hi = np.zeros(len(a)), np.ones(len(a))
    flo, fhi = f(lo), f(hi)
    for _ in range(200):
        t = np.where(fhi != flo, (lo * fhi - hi * flo) / (fhi - flo), 0.5 * (lo + hi))
        t = np.where((t > lo) & (t < hi), t, 0.5 * (lo + hi))
        ft = f(t)
        left = np.sign(ft) == np.sign(flo)
        # Illinois: halve the value at the end that stays, so the bracket shrinks from both sides
        fhi = np.where(left, fhi * 0.5, ft)
        flo = np.where(left, ft, flo * 0.5)
        lo, hi = np.where(left, t, lo), np.where(left, hi, t)
        done = (hi - lo <= 4 * np.finfo(np.float64).eps) | (ft == 0)
        if done.all():
            break
    t = 0.5 * (lo + hi)
    hit = pa + d * t[:, None]
    step = 1e-3 * np.linalg.norm(d, axis=1, keepdims=True)
    gradient = np.stack([distance(hit + step * e) - distance(hit - step * e) for e in np.eye(3)], axis=1)
    return t, (gradient / np.linalg.norm(gradient, axis=1, keepdims=True)).astype(np.float32)


def ellipsoid_distance(points, half_extents):
    """Signed distance estimate to an origin-centred ellipsoid, exact on its surface: k0 (k0 - 1) / k1 with
    k0 = |p / r| and k1 = |p / r^2|."""
    k0 = np.linalg.norm(points / half_extents, axis=1)
    k1 = np.maximum(np.linalg.norm(points / np.square(half_extents), axis=1), 1e-12)
    return k0 * (k0 - 1) / k1


@dataclass(frozen=True)
class Ellipsoid:
    """Ellipsoidal organ of `half_extents` cm centred on the origin, resting on a bed: nodes below `bed_height` are
    pinned. This is the kidney-scale benchmark body (the reference's kidney, `kidney.py:23-81`, is an ellipsoid plus a
    tumour lobe over the same shape protocol). Inside is |p / r| < 1; an edge's crossing solves the quadratic
    |(pa + t d) / r|^2 = 1 and its normal is the ellipsoid's, p / r^2 normalised."""
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

    def _level(self, points):
        r = np.array(self.half_extents, dtype=np.float64)
        return np.square(np.asarray(points, dtype=np.float64) / r).sum(axis=1)

    def distance(self, points):
        return ellipsoid_distance(points, np.array(self.half_extents, dtype=np.float64))

    def inside(self, positions):
        return self._level(positions) < 1.0

    def boundary(self, positions, a, b, inside):
        r = np.array(self.half_extents, dtype=np.float64)
        pa = positions[a].astype(np.float64) / r
        pd = (positions[b].astype(np.float64) - positions[a].astype(np.float64)) / r
        qa, qb, qc = (pd * pd).sum(axis=1), 2 * (pa * pd).sum(axis=1), (pa * pa).sum(axis=1) - 1
        root = np.sqrt(np.maximum(qb * qb - 4 * qa * qc, 0.0))
        w = -0.5 * (qb + np.copysign(root, qb))  # the two roots are w / qa and qc / w, without cancellation
        t1 = w / qa
        t2 = np.divide(qc, w, out=np.full_like(w, np.inf), where=w != 0)
        t = np.where((t1 >= 0) & (t1 <= 1), t1, t2)
        hit = positions[a].astype(np.float64) + (positions[b].astype(np.float64) - positions[a]) * t[:, None]
        n = hit / np.square(r)
        return t, (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)

    def bed(self, positions):
        return positions[:, 1] < self.bed_height


# ------------------------------------------------------------------------------------------------------ lattice build

class _Blocks:
    """3-D views of a lattice of `dims` voxels: node and voxel arrays in (z, y, x) order, x fastest when flat."""

    def __init__(self, dims):
        self.dims = tuple(int(d) for d in dims)
        vx, vy, vz = self.dims
        self.nodes3 = (vz + 1, vy + 1, vx + 1)
        self.voxels3 = (vz, vy, vx)
        self.strides = np.array((1, vx + 1, (vx + 1) * (vy + 1)), dtype=np.int64)
        self.ids = np.arange(int(np.prod(self.nodes3)), dtype=np.int64).reshape(self.nodes3)

    def corner_view(self, nodes3, bit):
        """The nodes at bit-order corner `bit` of every voxel, as a voxel-shaped view of a node array."""
        vz, vy, vx = self.voxels3
        dx, dy, dz = bit & 1, (bit >> 1) & 1, (bit >> 2) & 1
        return nodes3[dz:dz + vz, dy:dy + vy, dx:dx + vx]

    def pairs(self, axis):
        """Slices of the lower and upper node of every lattice edge along `axis` (axis 0 = x = last array axis)."""
        k = 2 - axis
        lower = [slice(None)] * 3
        upper = [slice(None)] * 3
        lower[k], upper[k] = slice(0, -1), slice(1, None)
        return tuple(lower), tuple(upper)


def _occupancy(blocks, inside3):
    """Occupied voxels (a corner inside), active nodes (a corner of an occupied voxel) and supported voxels (all eight
    corners active)."""
    occupied = np.zeros(blocks.voxels3, dtype=bool)
    for bit in range(8):
        occupied |= blocks.corner_view(inside3, bit)
    active = np.zeros(blocks.nodes3, dtype=bool)
    for bit in range(8):
        blocks.corner_view(active, bit)[...] |= occupied
    supported = np.ones(blocks.voxels3, dtype=bool)
    for bit in range(8):
        supported &= blocks.corner_view(active, bit)
    return occupied, active, supported


def _links(blocks, active3, cell_size):
    """Six-neighbour table and rest vectors of the links between active nodes."""
    n = blocks.ids.size
    neighbors = np.full((n, 6), -1, dtype=np.int32)
    rest = np.zeros((n, 6, 3), dtype=np.float32)
    for axis in range(3):
        lo, hi = blocks.pairs(axis)
        linked = active3[lo] & active3[hi]
        lower, upper = blocks.ids[lo][linked], blocks.ids[hi][linked]
        neighbors[lower, 2 * axis + 1] = upper
        neighbors[upper, 2 * axis] = lower
        rest[lower, 2 * axis + 1, axis] = cell_size[axis]
        rest[upper, 2 * axis, axis] = -cell_size[axis]
    return neighbors, rest


def _cut_edges(blocks, inside):
    """Lattice edges that may be cut -- those with all four incident voxels in the lattice (`cut.cuh:140-144`) -- as
    (axis, x, y, z) records axis by axis in node order, and the ones that cross the skin as 32-byte records with the
    inside flag of their lower end."""
    edge_parts, surface_parts = [], []
    vx, vy, vz = blocks.dims
    nx, ny = vx + 1, vy + 1
    for axis in range(3):
        keep = np.ones(blocks.nodes3, dtype=bool)
        for other in range(3):
            k = 2 - other
            sl = [slice(None)] * 3
            if other == axis:
                sl[k] = slice(blocks.dims[other], None)  # no upper end
                keep[tuple(sl)] = False
            else:
                sl[k] = 0
                keep[tuple(sl)] = False
                sl[k] = blocks.dims[other]
                keep[tuple(sl)] = False
        base = np.flatnonzero(keep.ravel())
        x, y, z = base % nx, (base // nx) % ny, base // (nx * ny)
        edge = np.column_stack((np.full(len(base), axis), x, y, z)).astype(np.int32)
        edge_parts.append(edge)
        cross = inside[base] != inside[base + blocks.strides[axis]]
        surface = np.zeros((int(cross.sum()), 8), dtype=np.int32)
        surface[:, :4] = edge[cross]
        surface[:, 4] = inside[base[cross]]
        surface_parts.append(surface)
    return np.concatenate(edge_parts), np.concatenate(surface_parts)


def _bends(neighbors):
    """Fifteen bend records per node over its slot pairs: j, k the neighbours, alive where both exist, rest angle pi
    for the two slots of one axis and pi/2 otherwise (`grid.py:257-268`)."""
    n = len(neighbors)
    bends = np.zeros((n, 15), dtype=BEND_DTYPE)
    bends["i"] = np.arange(n, dtype=np.int32)[:, None]
    bends["j"] = neighbors[:, BEND_SLOT_PAIRS[:, 0]]
    bends["k"] = neighbors[:, BEND_SLOT_PAIRS[:, 1]]
    bends["alive"] = (bends["j"] >= 0) & (bends["k"] >= 0)
    same_axis = (BEND_SLOT_PAIRS[:, 0] >> 1) == (BEND_SLOT_PAIRS[:, 1] >> 1)
    bends["theta"] = np.where(same_axis, np.pi, np.pi / 2)
    return bends


def _lumped_mass(blocks, supported, cell_size, density):
    """Mass of each node: every supported cell's six Kuhn tets lump rho V_cell / 24 on each of their corners
    (`grid.py:269-277`), so a cell gives its diagonal ends 6/24 and its other corners 2/24 of its mass."""
    count = np.zeros(blocks.nodes3, dtype=np.int64)
    for bit in range(8):
        blocks.corner_view(count, bit)[...] += KUHN_MULTIPLICITY[bit] * supported
    cell_mass = density * float(np.prod(cell_size.astype(np.float64) * METERS_PER_UNIT))
    return (count.ravel() * (cell_mass / 24.0)).astype(np.float32)


def _lattice(spacing, cell_size, dims, origin, positions, inside, bed, density, shape=None):
    """The TissueGrid of a placed lattice from its inside flags."""
    blocks = _Blocks(dims)
    inside = np.asarray(inside, dtype=np.int32)
    occupied3, active3, supported = _occupancy(blocks, inside.reshape(blocks.nodes3) != 0)
    active = active3.ravel().astype(np.int32)
    pinned = (np.asarray(bed, dtype=bool) & (active != 0)).astype(np.int32)
    vz, vy, vx = blocks.voxels3
    z, y, x = np.meshgrid(np.arange(vz), np.arange(vy), np.arange(vx), indexing="ij")
    base = (x + (vx + 1) * (y + (vy + 1) * z)).ravel()
    corners = (base[:, None] + CORNER_OFFSETS.astype(np.int64) @ blocks.strides).astype(np.int32)

    # the skin crossings of every voxel's edges, voxel by voxel in EDGE_CORNERS order
    end_a, end_b = corners[:, EDGE_CORNERS[:, 0]], corners[:, EDGE_CORNERS[:, 1]]
    crossing = inside[end_a] != inside[end_b]
    counts = crossing.sum(axis=1).astype(np.int32)
    offsets = np.concatenate(([0], np.cumsum(counts)[:-1])).astype(np.int32)
    a, b = end_a[crossing], end_b[crossing]
    intersections = np.zeros(len(a), dtype=ISECT_DTYPE)
    if len(a):
        t, normals = shape.boundary(positions, a, b, inside)
        intersections["position"] = positions[a] + (positions[b] - positions[a]) * t[:, None]
        intersections["normal"] = normals
        intersections["a"], intersections["b"], intersections["t"] = a, b, t
        intersections["inside_a"] = inside[a]

    neighbors, rest = _links(blocks, active3, cell_size)
    edges, surface_edges = _cut_edges(blocks, inside)
    mass = _lumped_mass(blocks, supported, cell_size, density)
    if not np.all(mass[active != 0] > 0):
        raise ValueError("an active node belongs to no supported cell")
    return TissueGrid(spacing, cell_size, tuple(int(d) for d in dims), origin, positions, inside, active, pinned, mass,
                      neighbors, rest, _bends(neighbors), corners, occupied3.ravel().astype(np.int32), intersections,
                      offsets, counts, surface_edges, edges, connectivity_table_cached())


def _placement(low, high, cell_size):
    """Voxel dims and origin of the lattice around bounds low..high: the extent in whole cells plus SHELL_CELLS of air
    on each side, centred on the bounds' centre."""
    dims = tuple(int(v) for v in np.ceil((high - low) / cell_size).astype(int) + 2 * SHELL_CELLS)
    origin = (low + high) / 2 - np.array(dims, dtype=np.float32) * cell_size / 2
    return dims, origin


def _node_positions(origin, dims, cell_size):
    index = lattice_coordinates(tuple(d + 1 for d in dims))
    return np.ascontiguousarray(origin + index * cell_size, dtype=np.float32)


def _tab_box(tab, half, cell_size):
    """The tab's box: its faces within 1e-6 cm of a block face moved onto that face (one node layer for both), and its
    attached face pushed one cell into the block, so that the nodes on the shared face lie inside the union."""
    low = np.array(tab.low, dtype=np.float32)
    high = np.array(tab.high, dtype=np.float32)
    ok = (np.isfinite(low).all() and np.isfinite(high).all() and tab.attached in (0, 1, 2)
          and np.all(high - low >= 2 * cell_size - 1e-6) and abs(low[tab.attached] - half[tab.attached]) < 1e-6)
    if not ok:
        raise ValueError(f"invalid tab {tab}: it must be at least two cells thick along every axis and start on a "
                         "positive face of the block")
    faces = np.stack((-half, half))
    for bound in (low, high):
        near = np.abs(bound[None, :] - faces) < 1e-6
        for side in range(2):
            bound[near[side]] = faces[side][near[side]]
    low[tab.attached] = half[tab.attached] - cell_size[tab.attached]
    return low, high


def build_grid(spacing=0.6, size=BLOCK_SIZE, density=1000.0, vertical_spacing=None, tab=None, shape=None):
    """Lattice of the block of `size` (with an optional `tab`) or of an implicit `shape`, in a two-cell air shell
    (the contract of `grid.py:176-186`: lattice spacing 1..7 mm, a vertical spacing between 0.1 mm and the spacing)."""
    if not np.isfinite(spacing) or not 0.1 <= spacing <= 0.7:
        raise ValueError(f"lattice spacing {spacing} cm is outside 0.1..0.7 cm (1..7 mm)")
    cell_size = np.full(3, spacing, dtype=np.float32)
    if vertical_spacing is not None:
        if not np.isfinite(vertical_spacing) or not .01 <= vertical_spacing <= spacing:
            raise ValueError(f"vertical spacing {vertical_spacing} cm is outside 0.01 cm..the lattice spacing")
        cell_size[1] = vertical_spacing
    if shape is None:
        half = np.array(size, dtype=np.float32) / 2
        boxes = [(-half, half)]
        if tab is not None:
            boxes.append(_tab_box(tab, half, cell_size))
        shape = BoxUnion(np.array([lo for lo, _ in boxes]), np.array([hi for _, hi in boxes]), half)
    dims, origin = _placement(shape.low, shape.high, cell_size)
    positions = _node_positions(origin, dims, cell_size)
    inside = shape.inside(positions).astype(np.int32)
    return _lattice(spacing, cell_size, dims, origin, positions, inside, shape.bed(positions), density, shape)


def block_grid(nodes=(17, 9, 9), spacing=0.25, density=1000.0, origin=(0.0, 0.0, 0.0)):
    """A solid block of `nodes` lattice points with no air shell: every node active and inside, the y = 0 layer
    pinned. Spacing is not limited to the reference builder's 1..7 mm range."""
    nodes = tuple(int(n) for n in nodes)
    if min(nodes) < 2:
        raise ValueError("a block needs at least two nodes per axis")
    cell_size = np.full(3, spacing, dtype=np.float32)
    dims = tuple(n - 1 for n in nodes)
    origin = np.array(origin, dtype=np.float32)
    positions = _node_positions(origin, dims, cell_size)
    inside = np.ones(len(positions), dtype=np.int32)
    return _lattice(float(spacing), cell_size, dims, origin, positions, inside, lattice_coordinates(nodes)[:, 1] == 0,
                    density)


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
