"""Parity of the HIP tissue guard (tensile reinforcement after every tet sweep, minimum-volume safeguard at the end of
every substep) with the FP32 oracle (GPU).

Bar: bit-exact positions and velocities. The thrown block inverts tets in its candidate positions, so the repair
and the feasible filter both run (the oracle reports the filtered nodes), over one and three substeps (different
parity-class phases) and one and two tensile passes; every live tet ends at J >= the minimum. The friction contact
path and a compacted tissue (incidence lists rebuilt over the new tet slots) run guarded as well.
"""
import numpy as np
import pytest

from genesis_tissue_simulation_amd.grid import block_grid
from genesis_tissue_simulation_amd.params import SolverParams, TET_GUARDED, neohookean_stiffness

pytestmark = pytest.mark.gpu


def thrown(substeps, passes, speed=100.0, seed=5, pinned=False):
    g = block_grid((9, 6, 6), 0.25)
    g.pinned[:] = 0
    if pinned:
        g.pinned[g.positions[:, 1] < 0.1] = 1
    p = SolverParams(substeps=substeps, iterations=2, tet_mode=TET_GUARDED, edge_ks=0.0, young=500.0,
                     strain_damping=0.002, bulk_damping=0.01, guard_min_volume=0.15, guard_passes=passes,
                     tensile_strength=80.0)
    rng = np.random.default_rng(seed)
    vel = rng.normal(0, speed, (len(g.positions), 3)).astype(np.float32)
    return g, p, vel


def pair(g, p, vel, use_graph=True):
    from genesis_tissue_simulation_amd.tissue import Tissue
    from oracle.oracle import Oracle
    from genesis_tissue_simulation_amd.tissue import build_topology
    topo = build_topology(g)
    tet_type = ((topo["tet_desc"] >> 1) & 7)[np.argsort(topo["tet_gen"])]
    kd, kh = neohookean_stiffness(500.0, 0.3, topo["type_volume"], tet_type)
    o = Oracle(g, p, g.positions)
    o.set_tet_stiffness(kd, kh)
    o.set_gravity((0, -981.0, 0))
    o.set_state(g.positions, vel)
    t = Tissue(g, p, kd, kh, g.positions, topology=topo, use_graph=use_graph)
    t.set_gravity((0, -981.0, 0))
    t.set_state(g.positions, vel)
    return o, t


def min_live_j(o, pos):
    topo = o.topology()
    B, _ = o.type_data()
    tet, _ = o.liveness()
    ids = topo["tet_ids"]
    x = np.asarray(pos, np.float64)
    Ds = np.stack([x[ids[:, k]] - x[ids[:, 0]] for k in (1, 2, 3)], axis=2)
    J = np.linalg.det(Ds @ B[topo["tet_type"]])
    return J[tet.astype(bool)].min()


@pytest.mark.parametrize("substeps,passes,pinned", [(1, 1, False), (3, 2, False), (1, 1, True), (3, 1, True)])
def test_thrown_block_guard_bit_exact(substeps, passes, pinned):
    g, p, vel = thrown(substeps, passes, pinned=pinned)
    o, t = pair(g, p, vel)
    filtered = 0
    for s in range(4):
        print(f"guard parity: substeps {substeps} passes {passes} pinned {pinned}, step {s}", flush=True)
        o.step(1 / 60)
        t.step(1 / 60)
        filtered += int(o.feasible().sum())
        po, vo = o.state()
        np.testing.assert_array_equal(t.positions(), po)
        np.testing.assert_array_equal(t.velocities(), vo)
        assert min_live_j(o, po) >= 0.15 - 1e-4
    print(f"  nodes through the feasible filter (last substeps): {filtered}")
    if substeps == 1:
        assert filtered > 0


def test_guard_with_friction_contact_and_no_graph():
    g, p, vel = thrown(2, 1, speed=20.0, seed=9, pinned=True)
    p = p.with_(contact_enabled=True, contact_radius=0.01)
    o, t = pair(g, p, vel, use_graph=False)
    L = 0.25
    contacts = 0
    for s in range(5):
        print(f"guard + contact parity, step {s}", flush=True)
        depth = 0.3 * L * (s + 1)
        tool = dict(shaftA=(-L, 5 * L - depth, 2.5 * L), shaftB=(9 * L, 5 * L - depth, 2.5 * L), shaftR=0.4 * L,
                    jawHinge=(4 * L, 5 * L - depth, 0.5 * L), jawAxis=(0, 0, 1), jawNormal=(1, 0, 0),
                    jawHalfWidth=0.5 * L, jawLen=2 * L, jawGap=0.2 * L)
        o.set_tool(tool, tool_armed=True, reset=(s == 0))
        o.step(1 / 60)
        t.step(1 / 60, tool=tool, tool_armed=True)
        np.testing.assert_array_equal(t.positions(), o.state()[0])
        np.testing.assert_array_equal(t.velocities(), o.state()[1])
        contacts += o.reaction()[1]
    assert contacts > 0


def test_guard_after_compaction():
    """Cut the block in two, compact the GPU tissue (its incidence lists follow the new tet slots) and keep stepping
    guarded: the oracle steps its uncompacted lists, whose dead tets it skips."""
    from genesis_tissue_simulation_amd.tissue import swept_cut_tool
    g, p, vel = thrown(2, 1, speed=30.0, seed=13)
    o, t = pair(g, p, vel)
    y0, y1 = -1.0, 3.0
    tool = swept_cut_tool(((1.05, y0, -1.0), (1.05, y0, 3.0)), ((1.05, y1, -1.0), (1.05, y1, 3.0)), 0.05)
    o.step(1 / 60)
    t.step(1 / 60)
    o.detect_cut(tool)
    t.detect_cut(tool)
    tr, er, nr, counts = t.compact(apply=True)
    assert counts[0] < o.T
    for s in range(3):
        print(f"guard after compaction, step {s}", flush=True)
        o.step(1 / 60)
        t.step(1 / 60)
        np.testing.assert_array_equal(t.positions(), o.state()[0])
        np.testing.assert_array_equal(t.velocities(), o.state()[1])
