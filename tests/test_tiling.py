"""The tiled tet sweep's plan (genesis_tissue_simulation_amd/tiling.py), checked by brute force on small lattices (no
GPU): the properties that make a launch of several colours compute what one launch per colour computes.

  * The groups cover the colours in order, at most kmax each, in an even number of launches.
  * Every launch takes the same regions in the same order (the plan's dense ids, from which the box origins follow),
    each region's entries and writes at fixed strides. Every tet of a group's colours is owned by exactly one
    region, the one holding its cell, and listed in a region's pass p exactly when its cell lies within k-1-p cells of
    the region, unless the region owns no tet and writes no node in that launch (then it lists nothing).
  * Halo completeness: for every region and pass p > 0, every tet of colour c0+p-1 that shares a node with a tet of
    the region's pass p is in the region's pass p-1 (so every node a pass reads is already final in LDS).
  * Every node of every tet is written by exactly one region per launch: the region of the last tet of the launch's
    colours touching it, else the region holding the node; the LDS box indices of entries and writes are in range and
    name the right nodes.
"""
import numpy as np
import pytest

from genesis_tissue_simulation_amd import tiling
from genesis_tissue_simulation_amd.grid import block_grid


def topo_of(g):
    from genesis_tissue_simulation_amd.tissue import build_topology
    return build_topology(g, bends=False)


@pytest.mark.parametrize("nodes,region,kmax", [((10, 7, 6), 3, 2), ((10, 7, 6), 3, 3), ((13, 9, 8), 4, 4),
                                               ((9, 9, 9), 2, 3)])
def test_plan_properties(nodes, region, kmax):
    g = block_grid(nodes, 0.25)
    g.active[np.arange(len(g.active)) % 11 == 3] = 0  # holes: cells without tets
    topo = topo_of(g)
    desc, offs = topo["tet_desc"], topo["tet_offsets"]
    plan = tiling.build_tile_plan(g.dims, desc, offs, region=region, kmax=kmax)
    vx, vy, vz = g.dims
    nx, ny = vx + 1, vy + 1
    nxy = nx * ny
    S, B, H = plan["region"], plan["box"], plan["halo"]
    RX, RY = -(-vx // S), -(-vy // S)
    base = (desc >> 4).astype(np.int64)
    typ = ((desc >> 1) & 7).astype(np.int64)
    cell = np.stack((base % nx, (base // nx) % ny, base // nxy), axis=1)
    corner = np.array([[(c & 1) + ((c >> 1) & 1) * nx + ((c >> 2) & 1) * nxy for c in k] for k in tiling.KUHN_BITS])
    ids = base[:, None] + corner[typ]
    colour = np.repeat(np.arange(len(offs) - 1), np.diff(offs))
    tet_nodes = np.unique(ids)
    groups = plan["groups"]
    assert len(groups) % 2 == 0 and groups[0, 0] == 0
    assert np.all(groups[1:, 0] == groups[:-1, 0] + groups[:-1, 1]) and groups[-1, 0] + groups[-1, 1] == len(offs) - 1
    assert np.all((groups[:, 1] >= 1) & (groups[:, 1] <= kmax))
    hdrs, ents, wrs = plan["headers"], plan["entries"].view(np.uint32), plan["writes"]
    # every launch takes the same regions in the same order: the plan's dense ids
    origins = [hdrs[f:f + n, :3] for f, n in zip(groups[:, 2], groups[:, 3])]
    assert all(np.array_equal(o, origins[0]) for o in origins)
    dense = plan["dense_ids"]
    assert len(dense) <= tiling.MAX_REGIONS and np.all(np.diff(dense) > 0)
    RXp, RYp = plan["rx"], plan["ry"]
    for f, n in zip(groups[:, 2], groups[:, 3]):
        np.testing.assert_array_equal(hdrs[f:f + n, 6], dense)
        # the box origin follows from the dense id alone
        d = dense.astype(np.int64)
        o = np.stack((d % RXp * S - H, d // RXp % RYp * S - H, d // (RXp * RYp) * S - H), axis=1)
        np.testing.assert_array_equal(hdrs[f:f + n, :3], o)
    assert groups.shape[1] == 16
    for c0, k, first, nreg, box, halo, stride, entry0, wcap, write0, *rest in groups:
        caps = rest[:k]
        starts = np.concatenate(([0], np.cumsum(caps)))
        assert stride == starts[-1] and all(c >= 1 for c in caps) and all(c == 0 for c in rest[k:])
        assert box == B and halo == H
        assert hdrs[first, 3] == entry0
        owned_count = np.zeros(len(desc), dtype=np.int64)
        written = np.zeros(nx * ny * (vz + 1), dtype=np.int64)
        lists = {}
        regions_idle = 0
        for r in range(first, first + nreg):
            h = hdrs[r]
            o = h[:3].astype(np.int64)
            region_xyz = (o + H) // S
            rid = region_xyz[0] + RX * (region_xyz[1] + RY * region_xyz[2])
            e = int(h[3])
            assert e == entry0 + (r - first) * stride  # found without the header
            assert np.all(h[tiling.HDR:tiling.HDR + k] <= caps)
            per_pass = []
            # a region that owns no tet of these colours and writes no node is idle in this launch: no entries
            mine_here = np.all(cell // S == region_xyz, axis=1) & (colour >= c0) & (colour < c0 + k)
            idle = not mine_here.any() and int(h[5]) == 0
            if idle:
                assert np.all(h[tiling.HDR:tiling.HDR + k] == 0)
                regions_idle += 1
            for p in range(k):
                cnt = int(h[tiling.HDR + p])
                rows = ents[e + starts[p]:e + starts[p] + cnt]
                assert np.all(ents[e + starts[p] + cnt:e + starts[p + 1]] == 0)
                assert np.all((rows[:, 0] >> 30) & 1)  # real rows carry bit 30
                slots = (rows[:, 0] & 0x3fffffff).astype(np.int64)
                owned = (rows[:, 0] >> 31).astype(bool)
                assert np.all(colour[slots] == c0 + p)
                lo, hi = region_xyz * S - (k - 1 - p), (region_xyz + 1) * S + (k - 1 - p)
                inside = np.all((cell >= lo) & (cell < hi), axis=1) & (colour == c0 + p)
                np.testing.assert_array_equal(np.sort(slots), [] if idle else np.flatnonzero(inside))
                np.testing.assert_array_equal(owned, np.all(cell[slots] // S == region_xyz, axis=1))
                owned_count[slots[owned]] += 1
                # the LDS box index of corner 0
                lb = rows[:, 1].astype(np.int64)
                rel = cell[slots] - o
                np.testing.assert_array_equal(lb, rel[:, 0] + B * (rel[:, 1] + B * rel[:, 2]))
                assert np.all((rel >= 0) & (rel + 1 < B))
                per_pass.append(set(slots.tolist()))
            lists[rid] = per_pass
            # halo completeness
            for p in range(1, k):
                need = set()
                nodes_p = set(ids[list(per_pass[p])].reshape(-1).tolist()) if per_pass[p] else set()
                prev = np.flatnonzero(colour == c0 + p - 1)
                for t in prev:
                    if nodes_p.intersection(ids[t].tolist()):
                        need.add(int(t))
                assert need <= per_pass[p - 1], (rid, p)
            w0, wn = int(h[4]), int(h[5])
            assert w0 == write0 + (r - first) * wcap and wn <= wcap  # found without the header
            assert np.all(wrs[w0 + wn:w0 + wcap, 1] == -1)
            w = wrs[w0:w0 + wn].astype(np.int64)
            written[w[:, 1]] += 1
            rel = np.stack((w[:, 1] % nx, (w[:, 1] // nx) % ny, w[:, 1] // nxy), axis=1) - o
            assert np.all((rel >= 0) & (rel < B))
            np.testing.assert_array_equal(w[:, 0], rel[:, 0] + B * (rel[:, 1] + B * rel[:, 2]))
            # a written node touched by the launch's colours: this region projects its last toucher
            for node in w[:, 1]:
                touch = np.flatnonzero(np.any(ids == node, axis=1) & (colour >= c0) & (colour < c0 + k))
                if len(touch):
                    last = touch[np.argmax(colour[touch])]
                    assert last in per_pass[colour[last] - c0]
                    assert np.all(cell[last] // S == region_xyz)
        mine = (colour >= c0) & (colour < c0 + k)
        assert np.all(owned_count[mine] == 1) and np.all(owned_count[~mine] == 0)
        assert np.all(written[tet_nodes] == 1)
        others = np.ones(len(written), bool)
        others[tet_nodes] = False
        assert np.all(written[others] == 0)


def test_group_sizes_are_even_and_bounded():
    for colours in range(2, 70):
        for kmax in (1, 2, 3, 4, 5):
            sizes = tiling.group_sizes(colours, kmax)
            if kmax == 1 and colours % 2:  # one colour a launch, an odd number of them: no even split
                assert sizes is None
                continue
            assert sum(sizes) == colours and len(sizes) % 2 == 0 and max(sizes) <= kmax
            assert len(sizes) <= -(-colours // kmax) + 1
    assert tiling.group_sizes(1, 3) is None


def _mix(a, b, s, side):
    """A non-commutative integer update of an edge's node value from both its nodes and its slot (wrapping uint64)."""
    with np.errstate(over="ignore"):
        return (a * np.uint64(0x9E3779B97F4A7C15) + b * np.uint64(0xC2B2AE3D27D4EB4F) + s * np.uint64(2 * side + 3)
                + np.uint64(side + 1)) ^ (a >> np.uint64(29))


@pytest.mark.parametrize("nodes,region,kmax", [((10, 7, 6), 3, 4), ((13, 9, 8), 4, 4), ((11, 8, 9), 3, 3),
                                               ((10, 8, 9), 2, 2), ((12, 10, 9), 5, 4)])
def test_edge_groups_replay_the_colour_sweep(nodes, region, kmax):
    """The plan's edge launches, replayed on the CPU with each region's box, passes and write list, leave every node
    where the edge colours applied one after another leave it; every edge is owned once, by its cell's region."""
    g = block_grid(nodes, 0.25)
    g.active[np.arange(len(g.active)) % 7 == 2] = 0
    topo = topo_of(g)
    ed, eoffs = topo["edge_desc"].view(np.uint32), topo["edge_offsets"]
    plan = tiling.build_tile_plan(g.dims, topo["tet_desc"], topo["tet_offsets"], region=region, kmax=kmax,
                                  edge_desc=ed, edge_offsets=eoffs)
    vx, vy, vz = g.dims
    nx, ny = vx + 1, vy + 1
    nxy = nx * ny
    S, B, H, RX, RY = plan["region"], plan["box"], plan["halo"], plan["rx"], plan["ry"]
    groups = plan["groups"]
    eg = groups[groups[:, 14] == 1]
    assert plan["edges"] == len(ed)
    if H < 1:
        assert len(eg) == 0
        return
    assert len(eg) % 2 == 0 and eg[0, 0] == 0 and eg[-1, 0] + eg[-1, 1] == len(eoffs) - 1
    assert np.all(eg[1:, 0] == eg[:-1, 0] + eg[:-1, 1]) and np.all(eg[:, 1] <= min(3, H))
    ei = (ed >> 4).astype(np.int64)
    ej = ei + np.array([1, nx, nxy])[(ed >> 1) & 3]
    rng = np.random.default_rng(5)
    vals = rng.integers(0, 2**63, size=nx * ny * (vz + 1), dtype=np.uint64)
    ents = plan["entries"].view(np.uint32).reshape(-1, 2)
    wrs = plan["writes"]
    dense = plan["dense_ids"]
    for gi in eg:
        c0, k, regions, stride, e0, wcap, w0 = gi[0], gi[1], gi[3], gi[6], gi[7], gi[8], gi[9]
        caps = gi[10:10 + k]
        starts = np.concatenate(([0], np.cumsum(caps)))
        want = vals.copy()  # the colour sweep, one colour at a time (a colour's edges share no node)
        for c in range(c0, c0 + k):
            sl = np.arange(eoffs[c], eoffs[c + 1])
            a, b = want[ei[sl]], want[ej[sl]]
            want[ei[sl]], want[ej[sl]] = _mix(a, b, sl.astype(np.uint64), 0), _mix(b, a, sl.astype(np.uint64), 1)
        got = np.full_like(vals, 0)
        written = np.zeros(len(vals), dtype=np.int64)
        owners = np.zeros(len(ed), dtype=np.int64)
        for rl in range(regions):
            d = int(dense[rl])
            ox, oy, oz = (d % RX) * S - H, ((d // RX) % RY) * S - H, (d // (RX * RY)) * S - H
            lx, ly, lz = np.meshgrid(np.arange(B), np.arange(B), np.arange(B), indexing="ij")
            x, y, z = (ox + lx).T.ravel(), (oy + ly).T.ravel(), (oz + lz).T.ravel()  # box index lx + B (ly + B lz)
            inb = (x >= 0) & (x < nx) & (y >= 0) & (y < ny) & (z >= 0) & (z <= vz)
            box = np.zeros(B ** 3, dtype=np.uint64)
            box[inb] = vals[(x + nx * y + nxy * z)[inb]]
            for p in range(k):
                rows = ents[e0 + rl * stride + starts[p]:e0 + rl * stride + starts[p] + caps[p]]
                rows = rows[(rows[:, 0] >> 30) & 1 == 1]
                sl = (rows[:, 0] & 0x3fffffff).astype(np.int64)
                owners[sl[(rows[:, 0] >> 31) == 1]] += 1
                assert np.all(eoffs[c0 + p] <= sl) and np.all(sl < eoffs[c0 + p + 1])
                li = rows[:, 1].astype(np.int64)
                lj = li + np.array([1, B, B * B])[(ed[sl] >> 1) & 3]
                assert np.all((x[li] + nx * y[li] + nxy * z[li]) == ei[sl]) and np.all(inb[li])
                assert np.all((x[lj] + nx * y[lj] + nxy * z[lj]) == ej[sl]) and np.all(inb[lj])
                a, b = box[li], box[lj]
                box[li], box[lj] = _mix(a, b, sl.astype(np.uint64), 0), _mix(b, a, sl.astype(np.uint64), 1)
            w = wrs[w0 + rl * wcap:w0 + (rl + 1) * wcap]
            w = w[w[:, 1] >= 0]
            got[w[:, 1]] = box[w[:, 0]]
            written[w[:, 1]] += 1
        touched = np.unique(np.concatenate((ei, ej)))
        assert np.all(written[touched] == 1)
        np.testing.assert_array_equal(got[touched], want[touched])
        sel = np.arange(eoffs[c0], eoffs[c0 + k])
        assert np.all(owners[sel] == 1)
        cx = np.minimum(ei[sel] % nx, vx - 1) // S
        assert cx.min() >= 0
        vals = want
