"""Tile plans for the temporally blocked tet sweep: several consecutive tet colours in one launch.

The Gauss-Seidel sweep runs its tet colours one after another; a launch per colour costs about 4 us on MI355X whatever
its size. A tiled launch takes K consecutive colours c0 .. c0+K-1 over cubic regions of S x S x S lattice cells, one
workgroup per region. The workgroup loads the nodes of its region and a halo into LDS, then projects, pass by pass,
  pass p:  the tets of colour c0+p whose cell lies within K-1-p cells of the region,
so that every tet a later pass reads through its nodes has already been projected in this workgroup's LDS: a tet of
colour c0+p touches nodes whose colour-(c0+p-1) tets lie within one more cell. Halo tets are projected by every region
that needs them, from the same inputs with the same operations, so each copy equals the owner's. The owner -- the
region holding the tet's cell -- writes the tet's multipliers; each node is written by the region holding the last
tet of the launch's colours that touches it (whose LDS copy is final), or, untouched, by the region holding the node.

Launches read positions from one buffer and write them to the other (a region's halo overlaps its neighbours' writes),
so a sweep takes an even number of launches. Every input a region needs sits at an address it can compute from its
index, so that all its loads go out in one round: its entries, write list and header at fixed strides, its box origin
from its dense id (passed with the launch), and per entry a copy of the tet's inputs (gathered once a step) and of its
multipliers (each region keeps the copies of the tets it projects, halo ones included, which stay equal to the owner's).
The per-node sequence of updates is the per-colour sweep's, so the results are bit for bit those of one launch per
colour.
"""
from __future__ import annotations

import numpy as np

# bit-order Kuhn corners (i = dx | dy << 1 | dz << 2) of the six types, as the kernels index them
KUHN_BITS = ((0, 1, 3, 7), (0, 3, 2, 7), (0, 2, 6, 7), (0, 6, 4, 7), (0, 4, 5, 7), (0, 5, 1, 7))
HDR = 8  # region header ints: box origin x, y, z, entry offset, write offset, write count, dense region id, 0; pass
# counts follow
MAX_REGIONS = 640  # regions a launch may take: their dense ids travel as kernel arguments (4 KiB in all)


def group_sizes(colours, kmax):
    """Consecutive colour groups of at most `kmax`, as few as possible but an even number of them (each launch swaps
    the position buffers, so a sweep must end where it began)."""
    if colours <= 0:
        return []
    n = -(-colours // kmax)
    if n % 2:
        n += 1
    if n > colours:  # one colour: a sweep of one launch cannot end where it began
        return None
    base, extra = divmod(colours, n)
    return [base + 1] * extra + [base] * (n - extra)


def build_tile_plan(dims, tet_desc, tet_offsets, region=8, kmax=3, edge_desc=None, edge_offsets=None, edge_kmax=3):
    """Plan of the tiled sweep for a lattice of `dims` cells and its colour-major tet descriptors; with the edge
    descriptors and colour offsets, also of the edge sweep (groups of at most min(edge_kmax, kmax - 1) edge colours,
    after the tet groups, marked kind 1).

    Returns a dict of numpy arrays:
      groups  int32 [G, 16]: first colour, colours, first region header, regions, box width, halo of the box, entry
              rows a region (the sum of the pass capacities), first entry, write capacity (the most nodes any region
              writes), first write, then the capacity of each pass p < 4 (the most tets any region's pass p holds;
              0 past the group's colours), kind (0 tets, 1 edges), 0
      headers int32 [R, HDR + kmax]: box origin (x, y, z), entry offset, write offset, write count, dense region id, 0,
              then the entry count of each pass (a description of the regions for the host; the kernel derives all
              of it from the region's index, its dense id and the rows' padding marks)
      entries int32 [M, 2]: tet slot | 1 << 30 | owned << 31, the tet's corner-0 index in the region's LDS box (padding
              rows are zero: bit 30 clear). The regions of a group take (entry rows) rows each, in header order from
              the group's first entry, pass p from row cap_0 + .. + cap_(p-1) of its block, count rows of it used
      writes  int32 [W, 2]: LDS box index, node id; each region of a group takes write-capacity rows in header order
              from the group's first write (padding rows: node id -1)
      dense_ids int32 [R per launch]: the dense id r = rx + RX (ry + RY rz) of each region, in header order (the
              same in every launch)
    with region, kmax, box, halo, RX, RY and the tet and edge counts, or None when the sweep cannot be tiled (fewer
    than two colours, or more than MAX_REGIONS regions).

    An edge (node i, axis a, node j = i + e_a) belongs to the cell at node i, clamped into the lattice's cells, so its
    nodes are corners of that cell. Its halo is node-based: a region's pass p must hold every edge touching the nodes
    N_p the later passes read, where N_last is the region's node range [qS, qS + S] on each axis and each pass of an
    axis-a colour widens it by one node along a (an edge reaches one node further along its own axis only)."""
    vx, vy, vz = (int(d) for d in dims)
    nx, ny = vx + 1, vy + 1
    nxy = nx * ny
    desc = np.asarray(tet_desc, dtype=np.uint32)
    offs = np.asarray(tet_offsets, dtype=np.int64)
    sizes = group_sizes(len(offs) - 1, kmax)
    if not sizes:
        return None
    S, H = int(region), int(kmax) - 1
    B = S + 2 * H + 1
    RX, RY, RZ = -(-vx // S), -(-vy // S), -(-vz // S)
    base = (desc >> 4).astype(np.int64)
    ttype = ((desc >> 1) & 7).astype(np.int64)
    cx, cy, cz = base % nx, (base // nx) % ny, base // nxy
    corner = np.array([[(c & 1) + ((c >> 1) & 1) * nx + ((c >> 2) & 1) * nxy for c in k] for k in KUHN_BITS])
    ids = base[:, None] + corner[ttype]
    all_nodes = np.unique(ids)
    if edge_desc is not None:
        ed = np.asarray(edge_desc, dtype=np.uint32)
        eoffs = np.asarray(edge_offsets, dtype=np.int64)
        ei = (ed >> 4).astype(np.int64)
        eaxis = ((ed >> 1) & 3).astype(np.int64)
        ej = ei + np.array([1, nx, nxy], dtype=np.int64)[np.minimum(eaxis, 2)]
        all_nodes = np.union1d(all_nodes, np.concatenate((ei, ej)))
    ncx, ncy, ncz = all_nodes % nx, (all_nodes // nx) % ny, all_nodes // nxy
    node_region = (np.minimum(ncx // S, RX - 1) + RX * (np.minimum(ncy // S, RY - 1) + RY * np.minimum(ncz // S, RZ - 1)))
    offsets27 = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    per_group = []  # (first colour, colours, rows, node writers)
    c0 = 0
    for k in sizes:
        rows = []  # (region, pass, slot | owned, local base)
        writer = np.full(nx * ny * (vz + 1), -1, dtype=np.int64)
        for p in range(k):
            c = c0 + p
            lo, hi = int(offs[c]), int(offs[c + 1])
            if hi <= lo:
                continue
            h = k - 1 - p
            slot = np.arange(lo, hi, dtype=np.int64)
            x, y, z = cx[lo:hi], cy[lo:hi], cz[lo:hi]
            rx, ry, rz = x // S, y // S, z // S
            for dx, dy, dz in offsets27:
                qx, qy, qz = rx + dx, ry + dy, rz + dz
                ok = (qx >= 0) & (qx < RX) & (qy >= 0) & (qy < RY) & (qz >= 0) & (qz < RZ)
                ok &= (x >= qx * S - h) & (x < (qx + 1) * S + h)
                ok &= (y >= qy * S - h) & (y < (qy + 1) * S + h)
                ok &= (z >= qz * S - h) & (z < (qz + 1) * S + h)
                if not ok.any():
                    continue
                r = qx[ok] + RX * (qy[ok] + RY * qz[ok])
                lb = (x[ok] - (qx[ok] * S - H)) + B * ((y[ok] - (qy[ok] * S - H)) + B * (z[ok] - (qz[ok] * S - H)))
                owned = 1 if (dx, dy, dz) == (0, 0, 0) else 0
                rows.append(np.stack((r, np.full(len(r), p), slot[ok] | (owned << 31) | (1 << 30), lb), axis=1))
            # the last colour of the launch touching a node decides its writer: the region of that tet's cell
            own = rx + RX * (ry + RY * rz)
            for q in range(4):
                writer[ids[lo:hi, q]] = own
        rows = np.concatenate(rows) if rows else np.zeros((0, 4), dtype=np.int64)
        rows = rows[np.lexsort((rows[:, 1], rows[:, 0]))]
        # nodes of any tet that no tet of these colours touches are carried by the region holding them
        wr = writer[all_nodes]
        wr = np.where(wr < 0, node_region, wr)
        per_group.append((c0, k, rows, wr, 0))
        c0 += k
    ke = min(int(edge_kmax), H)
    if edge_desc is not None and ke >= 1 and len(eoffs) > 1:
        ecolours = len(eoffs) - 1
        esizes = group_sizes(ecolours, ke)
        if not esizes:
            return None
        ex, ey, ez = ei % nx, (ei // nx) % ny, ei // nxy
        erx, ery, erz = np.minimum(ex, vx - 1) // S, np.minimum(ey, vy - 1) // S, np.minimum(ez, vz - 1) // S
        axes_of = [set(np.unique(eaxis[eoffs[c]:eoffs[c + 1]]).tolist()) for c in range(ecolours)]
        c0 = 0
        for k in esizes:
            rows = []
            writer = np.full(nx * ny * (vz + 1), -1, dtype=np.int64)
            for p in range(k):
                c = c0 + p
                lo, hi = int(eoffs[c]), int(eoffs[c + 1])
                if hi <= lo:
                    continue
                grow = [sum(b in axes_of[c0 + q] for q in range(p + 1, k)) for b in range(3)]
                slot = np.arange(lo, hi, dtype=np.int64)
                u = (ex[lo:hi], ey[lo:hi], ez[lo:hi])
                a = eaxis[lo:hi]
                rx, ry, rz = erx[lo:hi], ery[lo:hi], erz[lo:hi]
                for dx, dy, dz in offsets27:
                    q = (rx + dx, ry + dy, rz + dz)
                    ok = (q[0] >= 0) & (q[0] < RX) & (q[1] >= 0) & (q[1] < RY) & (q[2] >= 0) & (q[2] < RZ)
                    for b in range(3):  # touches N_p: node i in it, or node j along the edge's own axis
                        ok &= (u[b] >= q[b] * S - grow[b] - (a == b)) & (u[b] <= q[b] * S + S + grow[b])
                    if not ok.any():
                        continue
                    loc = [u[b][ok] - (q[b][ok] * S - H) for b in range(3)]
                    for b in range(3):  # both nodes inside the box
                        if np.any(loc[b] < 0) or np.any(loc[b] + (a[ok] == b) >= B):
                            raise ValueError("edge halo exceeds the box")
                    r = q[0][ok] + RX * (q[1][ok] + RY * q[2][ok])
                    lb = loc[0] + B * (loc[1] + B * loc[2])
                    owned = 1 if (dx, dy, dz) == (0, 0, 0) else 0
                    rows.append(np.stack((r, np.full(len(r), p), slot[ok] | (owned << 31) | (1 << 30), lb), axis=1))
                own = rx + RX * (ry + RY * rz)
                writer[ei[lo:hi]] = own
                writer[ej[lo:hi]] = own
            rows = np.concatenate(rows) if rows else np.zeros((0, 4), dtype=np.int64)
            rows = rows[np.lexsort((rows[:, 1], rows[:, 0]))]
            wr = writer[all_nodes]
            wr = np.where(wr < 0, node_region, wr)
            per_group.append((c0, k, rows, wr, 1))
            c0 += k
    # Every launch takes the same regions in the same order -- those owning a tet or writing a node in any launch (a
    # region owning nothing and writing nothing computes nothing that is kept) -- so that a region lands on the same
    # XCD in every launch and finds what it wrote last launch in that XCD's L2. A region idle in one launch has no
    # entries and no writes there and returns at once.
    regions = np.unique(np.concatenate([np.concatenate((rows[rows[:, 2] >= (1 << 31), 0], wr))
                                        for _, _, rows, wr, _ in per_group]))
    if len(regions) > MAX_REGIONS:  # a launch's table of dense ids
        return None
    groups, headers, entries, writes = [], [], [], []
    n_entries = n_writes = 0
    for c0, k, rows, wr, kind in per_group:
        active = np.union1d(rows[rows[:, 2] >= (1 << 31), 0], wr)
        rows = rows[np.isin(rows[:, 0], active)]
        g_first, g_entry, g_write = len(headers), n_entries, n_writes
        per_region = []
        for r in regions:
            sel = rows[rows[:, 0] == r]
            per_region.append((sel, np.bincount(sel[:, 1].astype(np.int64), minlength=k), all_nodes[wr == r]))
        caps = [max(1, max(int(c[p]) for _, c, _ in per_region)) for p in range(k)]
        starts = np.concatenate(([0], np.cumsum(caps)))
        stride = int(starts[-1])
        wcap = max(1, max(len(nd) for _, _, nd in per_region))
        for r, (sel, counts, nodes) in zip(regions, per_region):
            rx, ry, rz = r % RX, (r // RX) % RY, r // (RX * RY)
            ox, oy, oz = rx * S - H, ry * S - H, rz * S - H
            lx, ly, lz = nodes % nx - ox, (nodes // nx) % ny - oy, nodes // nxy - oz
            local = lx + B * (ly + B * lz)
            hdr = np.zeros(HDR + kmax, dtype=np.int64)
            hdr[:7] = (ox, oy, oz, n_entries, n_writes, len(nodes), r)
            hdr[HDR:HDR + k] = counts
            headers.append(hdr)
            padded = np.zeros((stride, 2), dtype=np.int64)
            pos = 0
            for p in range(k):  # sel is sorted by pass
                padded[starts[p]:starts[p] + counts[p]] = sel[pos:pos + counts[p], 2:4]
                pos += counts[p]
            entries.append(padded)
            wpad = np.zeros((wcap, 2), dtype=np.int64)
            wpad[:, 1] = -1  # padding: no node
            wpad[:len(nodes)] = np.stack((local, nodes), axis=1)
            writes.append(wpad)
            n_entries += stride
            n_writes += wcap
        groups.append((c0, k, g_first, len(regions), B, H, stride, g_entry, wcap, g_write,
                       *(caps + [0] * (4 - k)), kind, 0))
    as32 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int64).astype(np.int32))
    ent = np.concatenate(entries) if entries else np.zeros((0, 2), dtype=np.int64)
    # slot | 1 << 30 | owned << 31 wraps to a negative int32, which the kernel reads back as the same 32 bits
    ent32 = np.ascontiguousarray(ent.astype(np.uint64).astype(np.uint32).view(np.int32).reshape(-1, 2))
    return dict(groups=as32(groups), headers=as32(np.stack(headers)), entries=ent32,
                writes=as32(np.concatenate(writes)), dense_ids=as32(regions), region=S, kmax=int(kmax), box=B,
                halo=H, rx=RX, ry=RY, tets=int(len(desc)), edges=int(len(ed)) if edge_desc is not None else 0)
