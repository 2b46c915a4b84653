"""The records this path exchanges in the reference's layouts, against the reference's definitions, no GPU:
IsectGpu 40 B (`common.cuh:23-33`), GridEdgeGpu 32 B for the surface edges (`common.cuh:36-42`), GridEdgeGpu2 16 B for
the cut edges (`cut.cuh:58-63`), Conn4096Gpu 36 B (`cut.cuh:51-56`), CutPointGpu 32 B (`cut.cuh:37-46`) and
BendPairGpuC 32 B (`cut.cu:18`): record sizes, field offsets and the contents the grid builder puts there."""
import numpy as np
import pytest

from genesis_tissue_simulation_amd.grid import (BEND_DTYPE, CUT_POINT_DTYPE, EDGE_CORNERS, ISECT_DTYPE, build_grid,
                                                directed_slot)


@pytest.fixture(scope="module")
def grid():
    return build_grid(spacing=0.6)


def offsets(dtype):
    return {name: dtype.fields[name][1] for name in dtype.names}


def test_isect_records_are_isectgpu(grid):
    assert ISECT_DTYPE.itemsize == 40
    assert offsets(ISECT_DTYPE) == dict(position=0, normal=12, a=24, b=28, t=32, inside_a=36)
    it = grid.intersections
    assert len(it) and it.dtype == ISECT_DTYPE
    # globalRest = lerp(a, b, t) on the rest lattice, the normal a unit vector, insideA the flag of corner a
    x = grid.positions.astype(np.float64)
    hit = x[it["a"]] + (x[it["b"]] - x[it["a"]]) * it["t"][:, None].astype(np.float64)
    np.testing.assert_allclose(it["position"], hit, atol=1e-5)
    np.testing.assert_allclose(np.linalg.norm(it["normal"], axis=1), 1.0, atol=1e-5)
    np.testing.assert_array_equal(it["inside_a"], grid.inside[it["a"]])
    assert (grid.inside[it["a"]] != grid.inside[it["b"]]).all()
    # per-voxel lists: offsets are the running sum of counts and each voxel lists crossings of its own edges
    np.testing.assert_array_equal(grid.offsets, np.concatenate(([0], np.cumsum(grid.counts)[:-1])))
    assert grid.counts.max() <= 12
    v = int(np.flatnonzero(grid.counts)[0])
    own = {tuple(sorted(e)) for e in grid.corners[v][EDGE_CORNERS]}
    for r in it[grid.offsets[v]:grid.offsets[v] + grid.counts[v]]:
        assert tuple(sorted((int(r["a"]), int(r["b"])))) in own


def test_edge_records_are_gridedgegpu_and_gridedgegpu2(grid):
    se, ge = grid.surface_edges, grid.edges
    assert se.dtype == np.int32 and se.shape[1] * 4 == 32      # axis, base xyz, insideA, 3 pad
    assert ge.dtype == np.int32 and ge.shape[1] * 4 == 16      # axis, base xyz
    assert (se[:, 5:] == 0).all() and set(np.unique(se[:, 0])) <= {0, 1, 2}
    nx, ny = grid.dims[0] + 1, grid.dims[1] + 1
    base = se[:, 1] + nx * (se[:, 2] + ny * se[:, 3])
    other = base + np.array([1, nx, nx * ny])[se[:, 0]]
    np.testing.assert_array_equal(se[:, 4], grid.inside[base])
    assert (grid.inside[base] != grid.inside[other]).all()
    # every surface edge is a cut edge (its four voxels in bounds), and the skin table holds its crossing
    cut = {tuple(r) for r in ge.tolist()}
    assert all(tuple(r) in cut for r in se[:, :4].tolist())
    skin = grid.edge_skin_fraction()
    assert (skin[directed_slot(base, other, grid.dims)] >= 0).all()


def test_connectivity_rows_are_conn4096gpu(grid):
    c = grid.connectivity
    assert c.dtype == np.int32 and c.shape == (4096, 9)       # compCount, then the component of corners 0..7
    np.testing.assert_array_equal(c[0], [1, 0, 0, 0, 0, 0, 0, 0, 0])
    np.testing.assert_array_equal(c[4095], [8, 0, 1, 2, 3, 4, 5, 6, 7])
    assert ((c[:, 1:] >= 0) & (c[:, 1:] < c[:, :1])).all()


def test_cut_point_and_bend_records():
    assert CUT_POINT_DTYPE.itemsize == 32
    assert offsets(CUT_POINT_DTYPE) == dict(offset=0, owner=12, edge=16, normal=20)
    raw = np.arange(16, dtype=np.float32)
    raw[3], raw[4] = np.array([7, 9], dtype=np.int32).view(np.float32)
    p = raw.view(CUT_POINT_DTYPE)
    assert len(p) == 2 and p["owner"][0] == 7 and p["edge"][0] == 9
    np.testing.assert_array_equal(p["normal"][1], [13, 14, 15])
    assert BEND_DTYPE.itemsize == 32
    assert offsets(BEND_DTYPE)["theta"] == 12 and offsets(BEND_DTYPE)["alive"] == 16
