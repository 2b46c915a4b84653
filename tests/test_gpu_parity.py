"""Parity of the HIP substep and cut with the CPU oracle on the same seeded inputs (GPU).

Bars: cut topology (severed links, cut masks, directed cut fractions, bend pairs, revision, cut points) bit-exact;
positions within 1e-5 relative (max |dx| / max |x|) after 100 substeps. The kernels and the FP32 oracle evaluate the
same IEEE operation sequence without contraction, so positions are also checked for exact equality; the FP64 oracle
error is reported.
"""
import numpy as np
import pytest

from genesis_tissue_simulation_amd.params import TET_GUARDED, TET_NEOHOOKEAN, TET_VOLUME
from genesis_tissue_simulation_amd.workloads import block_6144, tet_count
from genesis_tissue_simulation_amd.grid import block_grid

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def make_pair(w, params=None, double=False, positions=None):
    from genesis_tissue_simulation_amd.tissue import Tissue
    from oracle.oracle import Oracle
    p = params or w.params
    init = w.initial_positions if positions is None else positions
    kd, kh = w.stiffness()
    o = Oracle(w.grid, p, init, double=double)
    o.set_tet_stiffness(kd, kh)
    o.set_gravity(w.gravity)
    t = Tissue(w.grid, p, kd, kh, init)
    t.set_gravity(w.gravity)
    return o, t


def advance(pairs, w, steps, cut=True, start=0):
    from genesis_tissue_simulation_amd.tissue import swept_cut_tool
    for s in range(start, start + steps):
        tool = swept_cut_tool(w.blade(s), w.blade(s + 1), w.blade_thickness) if cut and w.blade else None
        for sim in pairs:
            sim.step(w.dt)
            if tool is not None:
                sim.detect_cut(tool)


def rel_err(a, b):
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).max() /
                 np.abs(np.asarray(b, np.float64)).max())


def assert_cut_equal(ct, co):
    np.testing.assert_array_equal(ct["neighbors"], co["neighbors"])
    np.testing.assert_array_equal(ct["cut_mask"], co["cut_mask"])
    np.testing.assert_array_equal(ct["edge_cut_t"], co["edge_cut_t"])
    np.testing.assert_array_equal(ct["bend_alive"], co["bend_alive"])
    assert ct["revision"] == co["revision"]
    assert ct["cut_point_counter"] == co["cut_point_counter"]
    np.testing.assert_array_equal(ct["cut_points"], co["cut_points"])


def test_block_100_substeps_matches_oracle():
    w = block_6144()
    o, t = make_pair(w)
    o64, _ = make_pair(w, double=True)
    advance([o, t, o64], w, 10, cut=False)
    po, vo = o.state()
    pt, vt = t.positions(), t.velocities()
    err = rel_err(pt, po)
    err64 = rel_err(pt, o64.state()[0])
    print(f"block 100 substeps: rel err vs FP32 oracle {err:.3e}, vs FP64 oracle {err64:.3e}")
    assert err <= REL_TOL
    assert err64 <= 1e-4
    np.testing.assert_array_equal(pt, po)
    np.testing.assert_array_equal(vt, vo)


def test_block_cut_topology_bit_exact():
    w = block_6144()
    o, t = make_pair(w)
    advance([o, t], w, 40, cut=True)
    co, ct = o.cut_state(), t.cut_state()
    assert co["revision"] > 0, "the blade must sever links"
    assert_cut_equal(ct, co)
    lo, lt = o.liveness(), t.liveness()
    np.testing.assert_array_equal(lt[0], lo[0])
    np.testing.assert_array_equal(lt[1], lo[1])
    assert rel_err(t.positions(), o.state()[0]) <= REL_TOL
    print(f"cut: revision {co['revision']}, cut points {co['cut_point_counter']}, live tets {lo[0].sum()}")


def test_cut_unit_parity_on_identical_positions():
    """The cut kernels on exactly the same deformed positions as the oracle (no stepping in between)."""
    w = block_6144()
    rng = np.random.default_rng(3)
    pos = w.initial_positions + rng.normal(0, 0.02 * 0.25, w.initial_positions.shape).astype(np.float32)
    o, t = make_pair(w, positions=pos)
    from genesis_tissue_simulation_amd.tissue import swept_cut_tool
    for s in range(0, 36, 3):
        tool = swept_cut_tool(w.blade(s), w.blade(s + 3), w.blade_thickness)
        o.detect_cut(tool)
        t.detect_cut(tool)
    assert_cut_equal(t.cut_state(), o.cut_state())


@pytest.mark.parametrize("precise", [False, True])
@pytest.mark.parametrize("mode", [TET_VOLUME, TET_NEOHOOKEAN, TET_GUARDED])
@pytest.mark.parametrize("iterations", [1, 2])
def test_tet_modes_and_iterations(mode, iterations, precise):
    """Every tet form, undamped and damped (bulk damping with the guarded form), in the reference's direct forms and
    in the precise ones: bit-exact with the oracle (the guarded log is the same own sequence on both sides)."""
    w = block_6144()
    damped = iterations == 2
    p = w.params.with_(tet_mode=mode, iterations=iterations, strain_damping=0.01 if damped else 0.0,
                       bulk_damping=0.005 if damped and mode == TET_GUARDED else 0.0, precise_strain=precise)
    o, t = make_pair(w, params=p)
    o64, _ = make_pair(w, params=p, double=True)
    advance([o, t, o64], w, 3, cut=False)
    err = rel_err(t.positions(), o.state()[0])
    print(f"mode {mode} iterations {iterations} precise {precise}: rel err {err:.3e}, "
          f"vs FP64 {rel_err(t.positions(), o64.state()[0]):.3e}")
    np.testing.assert_array_equal(t.positions(), o.state()[0])
    np.testing.assert_array_equal(t.velocities(), o.state()[1])
    if iterations > 1:  # with one iteration nothing reads the multipliers and the kernels store none
        for a, b in zip(t.multipliers(), o.multipliers()):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("iterations", [2, 3])
def test_multipliers_of_rows_that_cannot_move(iterations):
    """Rows whose every node is pinned project nothing; the substep's reset leaves their multipliers at 0 whatever
    the buffers held before (xpbd_solver.cu:3534-3539), and every other multiplier matches the oracle bit for bit."""
    import torch
    w = block_6144()
    grid = w.grid
    pinned = grid.pinned.astype(bool)
    o, t = make_pair(w, params=w.params.with_(iterations=iterations))
    with torch.no_grad():
        t.edge_lambda.fill_(7.0)
        t.tet_lambda.fill_(-3.0)
    advance([o, t], w, 1, cut=False)
    ge, gd, gh = t.multipliers()
    oe, od, oh = o.multipliers()
    np.testing.assert_array_equal(ge, oe)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(gh, oh)
    e_top = o.topology()
    both_pinned = pinned[e_top["edge_i"]] & pinned[e_top["edge_j"]]
    assert both_pinned.any(), "the block has no edge between two pinned nodes"
    assert (ge[both_pinned] == 0).all()
    np.testing.assert_array_equal(t.positions(), o.state()[0])


def test_fallback_edges_and_damping():
    w = block_6144()
    p = w.params.with_(edge_ks=0.0, young=60.0, beta_s=0.5, rayleigh=2.0, damping=0.01)
    o, t = make_pair(w, params=p)
    advance([o, t], w, 8, cut=True)
    np.testing.assert_array_equal(t.positions(), o.state()[0])
    np.testing.assert_array_equal(t.velocities(), o.state()[1])
    assert_cut_equal(t.cut_state(), o.cut_state())


def test_tool_contact_parity():
    w = block_6144()
    p = w.params.with_(contact_enabled=True, contact_radius=0.01, jaw_gap_closed=0.02)
    o, t = make_pair(w, params=p)
    L = 0.25
    for s in range(6):
        depth = 0.4 * L * (s + 1) / 6
        tool = dict(shaftA=(-L, 8 * L + 0.3 * L - depth, 4 * L), shaftB=(17 * L, 8 * L + 0.3 * L - depth, 4 * L),
                    shaftR=0.3 * L, jawHinge=(8 * L, 8 * L + 0.2 * L - depth, 2 * L), jawAxis=(0, 0, 1),
                    jawNormal=(1, 0, 0), jawHalfWidth=0.5 * L, jawLen=2 * L, jawGap=0.2 * L,
                    bladeA=(2 * L, 8 * L + 0.1 * L - depth, 6 * L), bladeB=(6 * L, 8 * L + 0.1 * L - depth, 6 * L),
                    bladeR=0.1 * L)
        o.set_tool(tool, tool_armed=True, blade_armed=True, reset=(s == 0))
        o.step(w.dt)
        t.step(w.dt, tool=tool, tool_armed=True, blade_armed=True)
    err = rel_err(t.positions(), o.state()[0])
    ro, no = o.reaction()
    rt, nt = t.reaction_force()
    print(f"contact: rel err {err:.3e}, reaction oracle {ro} gpu {rt}, contacts {no} {nt}")
    assert no > 0
    np.testing.assert_array_equal(t.positions(), o.state()[0])
    assert nt == no
    np.testing.assert_allclose(rt, ro, rtol=1e-4, atol=1e-6 * max(1.0, np.abs(ro).max()))


def test_compaction_matches_oracle():
    w = block_6144()
    o, t = make_pair(w)
    advance([o, t], w, 30, cut=True)
    tro, ero, nro, co = o.compact()
    trt, ert, nrt, ct = t.compact()
    np.testing.assert_array_equal(trt, tro)
    np.testing.assert_array_equal(ert, ero)
    np.testing.assert_array_equal(nrt, nro)
    np.testing.assert_array_equal(ct, co)
    live = trt >= 0
    assert np.array_equal(np.sort(trt[live]), np.arange(live.sum()))


def test_ccd_cut_parity():
    """CCD of edges that moved one cell along +x during the tick, against a static blade along z."""
    w = block_6144()
    o, t = make_pair(w)
    L = 0.25
    advance([o, t], w, 2, cut=False)
    cur = t.positions()
    prev = (cur - np.array((1.0 * L, 0.0, 0.0), np.float32)).astype(np.float32)
    S, E = np.array((4.1 * L, 7.5 * L, 1 * L), np.float32), np.array((4.1 * L, 7.5 * L, 7 * L), np.float32)
    o.detect_cut_ccd(S, E, (1, 0, 0), 0.02, prev)
    t.detect_cut_ccd(S, E, (1, 0, 0), 0.02, prev)
    co, ct = o.cut_state(), t.cut_state()
    assert co["revision"] > 0
    assert_cut_equal(ct, co)


def test_cutting_everything_and_compacting_to_nothing():
    """Horizontal blades through every half-cell layer of the default lattice (the reference's box-and-tab scene,
    with its air shell, so every tissue link has a cut record) sever every vertical link: every tet dies, the
    compaction leaves an empty tet list (and the horizontal edges), and stepping the remains (free nodes hanging on
    edges, pinned ones still) goes on matching the oracle, which never compacts. A blade that misses the tissue
    changes nothing."""
    from genesis_tissue_simulation_amd.grid import build_grid
    from genesis_tissue_simulation_amd.tissue import Tissue, swept_cut_tool
    from oracle.oracle import Oracle

    grid = build_grid(0.6)
    T = tet_count(grid)
    w = block_6144()
    params = w.params
    kd = np.full(T, 1e7, dtype=np.float32)
    o = Oracle(grid, params, grid.positions)
    o.set_tet_stiffness(kd, kd * 10)
    o.set_gravity(w.gravity)
    t = Tissue(grid, params, kd, kd * 10, grid.positions)
    t.set_gravity(w.gravity)
    act = grid.active != 0
    lo, hi = grid.positions[act].min(axis=0), grid.positions[act].max(axis=0)
    L = float(grid.cell_size[1])
    miss = swept_cut_tool(((lo[0] - 5, hi[1] + 5, lo[2]), (lo[0] - 5, hi[1] + 5, hi[2])),
                          ((hi[0] + 5, hi[1] + 5, lo[2]), (hi[0] + 5, hi[1] + 5, hi[2])), 0.02)
    for sim in (o, t):
        sim.detect_cut(miss)
    assert t.revision() == 0 and t.compact(apply=True)[3][0] == T
    for y in np.arange(lo[1] + 0.5 * L, hi[1], L):
        cut = swept_cut_tool(((lo[0] - L, y, lo[2] - L), (lo[0] - L, y, hi[2] + L)),
                             ((hi[0] + L, y, lo[2] - L), (hi[0] + L, y, hi[2] + L)), 0.02)
        for sim in (o, t):
            sim.detect_cut(cut)
    assert_cut_equal(t.cut_state(), o.cut_state())
    tr, er, nr, counts = t.compact(apply=True)
    _, _, _, ocounts = o.compact()
    np.testing.assert_array_equal(counts, ocounts)
    assert counts[0] == 0 and t.T == 0 and counts[1] > 0
    for _ in range(3):
        for sim in (o, t):
            sim.step(w.dt)
    np.testing.assert_array_equal(t.positions(), o.state()[0])
    for a, b in zip(t.liveness(), o.liveness()):
        np.testing.assert_array_equal(a, b)
    assert t.compact(apply=True)[3][0] == 0


def test_particle_frames_and_cut_points_in_the_owner_frame():
    """A sheared and rotated block: the frames after steps and after extra updates (LCS_Finalize), and the cut points
    recorded in them, equal the oracle's bit for bit."""
    from genesis_tissue_simulation_amd.tissue import Tissue, swept_cut_tool
    from oracle.oracle import Oracle

    w = block_6144()
    c, s_ = np.cos(0.6), np.sin(0.6)
    A = np.array([[c, -s_, 0.1], [s_, c, 0.0], [0.05, 0.0, 1.0]])
    x0 = w.grid.positions.astype(np.float64)
    centre = x0.mean(axis=0)
    x = ((x0 - centre) @ A.T + centre).astype(np.float32)
    kd, kh = w.stiffness()
    o = Oracle(w.grid, w.params, x)
    o.set_tet_stiffness(kd, kh)
    t = Tissue(w.grid, w.params, kd, kh, x)
    for sim in (o, t):
        sim.set_gravity(w.gravity)
    L = 0.25
    world = lambda p: (np.asarray(p, np.float64) - centre) @ A.T + centre
    for s in range(4):
        print(f"frames parity: step {s}", flush=True)
        o.step(w.dt)
        t.step(w.dt)
        np.testing.assert_array_equal(t.rotations(), o.rotations())
        o.compute_rotations()
        t.finalize()
        np.testing.assert_array_equal(t.rotations(), o.rotations())
    tool = swept_cut_tool((world((-L, 4 * L, -L)), world((-L, 4 * L, 9 * L))),
                          (world((17 * L, 4 * L, -L)), world((17 * L, 4 * L, 9 * L))), 0.02)
    o.detect_cut(tool)
    t.detect_cut(tool)
    rot = o.rotations()
    assert np.abs(rot - np.eye(3)).max() > 0.3
    assert_cut_equal(t.cut_state(), o.cut_state())
    assert o.cut_state()["cut_point_counter"] > 0


def test_xcd_order_changes_nothing():
    """The XCD-aware workgroup order of the colour and node kernels is a bijection of the items: with it on and off
    the step (tool contact, damping, cuts) gives the same bits."""
    from genesis_tissue_simulation_amd.tissue import Tissue
    w = block_6144()
    p = w.params.with_(contact_enabled=True, strain_damping=0.01, beta_s=0.5, iterations=2)
    kd, kh = w.stiffness()
    sims = []
    for order in (False, True):
        t = Tissue(w.grid, p, kd, kh, w.initial_positions, xcd_order=order)
        t.set_gravity(w.gravity)
        sims.append(t)
    L = 0.25
    for s in range(6):
        tool = dict(shaftA=(-L, 8 * L - 0.1 * s * L, 4 * L), shaftB=(17 * L, 8 * L - 0.1 * s * L, 4 * L),
                    shaftR=0.3 * L)
        for t in sims:
            t.step(w.dt, tool=tool, tool_armed=True)
        advance(sims, w, 1, cut=True, start=s)
    a, b = sims
    np.testing.assert_array_equal(a.positions(), b.positions())
    np.testing.assert_array_equal(a.velocities(), b.velocities())
    assert_cut_equal(a.cut_state(), b.cut_state())
