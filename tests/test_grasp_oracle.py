"""Pins of the oracle's legacy grasp (contact model 0) against what the mechanics fixes (no GPU).

The reference ships no vectors for the grasp kernels (`k_MarkGraspCandidates`, `k_SolveSlabCoulomb`, `k_SolveCorePin`,
`k_CoreBreakCheck`, `k_SolveCCDContact`, xpbd_solver.cu:930-1199) and its only native self-test covers the tear law.
The pins below use lattices of free particles: the tets and edges get compliances near 1e35, so the material moves
nothing and each corner answers to the grasp alone, where XPBD has closed forms.
"""
import numpy as np
import pytest

from genesis_tissue_simulation_amd.grid import block_grid
from genesis_tissue_simulation_amd.params import TET_VOLUME, GraspParams, SolverParams, legacy_grasp_params
from genesis_tissue_simulation_amd.workloads import block_6144

DT = float(np.float32(1 / 60))  # the step length as the oracle receives it (float)


def free_particles(oracle_lib, gp, gravity=(0.0, 0.0, 0.0), double=True, substeps=1, damping=0.0, positions=None):
    g = block_grid(nodes=(4, 4, 4), spacing=0.25)
    g.pinned[:] = 0
    p = SolverParams(substeps=substeps, iterations=1, tet_mode=TET_VOLUME, edge_ks=1e-30, damping=damping)
    o = oracle_lib.Oracle(g, p, positions, double=double)
    o.set_tet_stiffness(np.full(o.T, 1e-30, np.float32), np.full(o.T, 1e-30, np.float32))
    o.set_gravity(gravity)
    o.set_contact_model(0, gp)
    return g, o


def jaw_tool(hinge, grasp01=1.0, gap=0.0, shaft=((5.0, 5.0, 5.0), (5.0, 5.0, 6.0)), shaft_r=0.01, half_width=0.3,
             length=0.8):
    return dict(shaftA=shaft[0], shaftB=shaft[1], shaftR=shaft_r, jawHinge=hinge, jawAxis=(0, 0, 1),
                jawNormal=(0, 1, 0), jawHalfWidth=half_width, jawLen=length, jawGap=gap, grasp01=grasp01)


def jaw_coords(x, hinge):
    """(u, v, n) in the frame axis = z, normal = y, w = axis x normal = -x."""
    rel = np.asarray(x, np.float64) - np.asarray(hinge, np.float64)
    return rel[:, 2], -rel[:, 0], rel[:, 1]


STIFF = GraspParams(core_compliance=0.0, grasp_omega=1.0, grasp_displacement=1e9, skin=0.1)
HINGE = np.array([0.25, 0.5, 0.0])


def test_capture_takes_the_jaw_box_and_the_ellipse_core(oracle_lib):
    """The captured set is every live unpinned corner inside the jaw box widened by the skin; the core is the
    corners inside the normalised ellipse of the set's own u / v extents."""
    w = block_6144()
    o = oracle_lib.Oracle(w.grid, w.params, w.initial_positions)
    o.set_tet_stiffness(*w.stiffness())
    o.set_gravity(w.gravity)
    gp = legacy_grasp_params(0.25)
    o.set_contact_model(0, gp)
    for _ in range(3):
        o.step(w.dt)
    x = o.state()[0].astype(np.float64)
    hinge = (2.0, 1.95, 0.4)
    assert o.set_tool(jaw_tool(hinge, half_width=0.55, length=1.1), reset=True) == 1
    gs = o.grasp_state()
    u, v, n = jaw_coords(x, hinge)
    band = 0.5 * max(0.0, gp.closed_gap) + gp.skin
    free = (w.grid.active != 0) & (w.grid.pinned == 0)
    inside = free & (u >= 0) & (u <= 1.1) & (np.abs(v) <= 0.55) & (np.abs(n) <= band)
    margin = np.minimum.reduce([np.abs(u), np.abs(u - 1.1), np.abs(np.abs(v) - 0.55), np.abs(np.abs(n) - band)])
    sure = inside & (margin > 1e-5)
    assert gs["grasping"] and sure.sum() >= 8
    assert (gs["grasped"][sure] == 1).all() and (gs["grasped"][~inside & (margin > 1e-5)] == 0).all()
    ids = np.nonzero(gs["grasped"])[0]
    uc, vc = u[ids], v[ids]
    nu = (uc - 0.5 * (uc.min() + uc.max())) / max(0.5 * (uc.max() - uc.min()), 1e-6)
    nv = (vc - 0.5 * (vc.min() + vc.max())) / max(0.5 * (vc.max() - vc.min()), 1e-6)
    d2 = nu * nu + nv * nv
    core = d2 < gp.core_fraction ** 2
    if not core.any():
        core[np.argmin(d2)] = True
    clear = np.abs(d2 - gp.core_fraction ** 2) > 1e-5
    np.testing.assert_array_equal(gs["core"][ids][clear], core[clear].astype(np.int32))
    assert gs["core"].sum() >= 1 and not (gs["core"] & (1 - gs["grasped"])).any()
    np.testing.assert_allclose(gs["anchor"][ids], np.stack((u[ids], v[ids]), axis=1), atol=1e-6)


def test_release_damps_the_grasped_velocity_and_clears_the_sets(oracle_lib):
    g, o = free_particles(oracle_lib, STIFF, gravity=(0.0, -981.0, 0.0))
    assert o.set_tool(jaw_tool(HINGE), reset=True) == 1
    o.step(DT)
    gs = o.grasp_state()
    grasped = gs["grasped"].astype(bool)
    vel_before = o.state()[1]
    assert o.set_tool(jaw_tool(HINGE, grasp01=0.4)) == 0       # between the thresholds: nothing
    assert o.set_tool(jaw_tool(HINGE, grasp01=0.2)) == 2
    vel_after = o.state()[1]
    np.testing.assert_allclose(vel_after[grasped], float(np.float32(0.1)) * vel_before[grasped], rtol=1e-12)
    np.testing.assert_array_equal(vel_after[~grasped], vel_before[~grasped])
    gs = o.grasp_state()
    assert not gs["grasping"] and gs["grasped"].sum() == 0 and gs["core"].sum() == 0


def test_stiff_core_pin_holds_the_weight_and_the_break_check_measures_it(oracle_lib):
    """A stiff pin cancels gravity's h^2 g each substep, so |lambda| / h^2 is the corner's weight m g, and the
    break-check EMA with its one-second time constant reads m g (1 - e^-t)."""
    g, o = free_particles(oracle_lib, STIFF, gravity=(0.0, -981.0, 0.0), substeps=4)
    x0 = o.state()[0].copy()
    o.set_tool(jaw_tool(HINGE), reset=True)
    core = np.nonzero(o.grasp_state()["core"])[0]
    assert len(core) >= 1
    for s in range(1, 91):
        o.step(DT)
        if s in (1, 30, 90):
            ema = o.grasp_state()["force_ema"][core]
            np.testing.assert_allclose(ema, g.mass[core].astype(np.float64) * 981.0 * (1 - np.exp(-s * DT)), rtol=1e-9)
    np.testing.assert_allclose(o.state()[0][core], x0[core], atol=1e-12)


def test_core_breaks_when_the_ema_passes_the_break_force(oracle_lib):
    weight = None
    for s_break in (10, 25):
        g, o = free_particles(oracle_lib, STIFF, gravity=(0.0, -981.0, 0.0))
        o.set_tool(jaw_tool(HINGE), reset=True)
        core = np.nonzero(o.grasp_state()["core"])[0]
        weight = float(g.mass[core].min()) * 981.0
        # between the EMA of step s_break - 1 and of step s_break
        f = weight * 0.5 * ((1 - np.exp(-(s_break - 1) * DT)) + (1 - np.exp(-s_break * DT)))
        o.set_contact_model(0, STIFF.with_(break_force=float(f)))
        o.set_tool(jaw_tool(HINGE), reset=True)
        for s in range(1, s_break + 2):
            o.step(DT)
            gs = o.grasp_state()
            held = gs["core"][core].astype(bool)
            assert held.all() == (s < s_break), s
        assert not gs["grasped"][core].any() and gs["grasping"]
        assert (o.state()[1][core, 1] < 0).all()                   # the broken corners fall


def test_compliant_core_pin_converges_geometrically(oracle_lib):
    """With omega = 1 a compliant pin lands each substep on the XPBD fixed point x0 + w (t - x0) / (w + alpha): the
    error to the target shrinks by alpha / (w + alpha) per substep (velocity killed, no gravity)."""
    g = block_grid(nodes=(4, 4, 4), spacing=0.25)
    c = float(np.float32(0.5 * DT ** 2 / g.mass.max()))    # alpha ~ w / 2: a visible contraction per substep
    gp = STIFF.with_(core_compliance=c)
    g, o = free_particles(oracle_lib, gp, damping=1.0)
    o.set_tool(jaw_tool(HINGE), reset=True)
    core = np.nonzero(o.grasp_state()["core"])[0]
    x0 = o.state()[0][core].copy()
    shift = np.array([0.01, 0.02, -0.005])
    moved = jaw_tool(HINGE + shift)
    o.set_tool(moved)                      # the step moves the tool to the new pose
    o.step(DT)
    # the tool fields are float: the target moves by the difference of the rounded hinges
    target = x0 + (np.float32(HINGE + shift).astype(np.float64) - np.float32(HINGE).astype(np.float64))
    alpha = c / (DT * DT)
    wm = 1.0 / g.mass[core].astype(np.float64)
    ratio = alpha / (wm + alpha)
    assert (ratio > 0.05).all()
    err = o.state()[0][core] - target
    for _ in range(3):
        o.set_tool(moved)
        o.step(DT)
        err_next = o.state()[0][core] - target
        np.testing.assert_allclose(err_next, err * ratio[:, None], rtol=1e-9, atol=1e-15)
        err = err_next
    assert np.abs(err).max() > 0


@pytest.mark.parametrize("mu,sticks", [(3.0, True), (0.5, False)])
def test_ring_corners_follow_coulomb(oracle_lib, mu, sticks):
    """Gravity pressing the ring into the slab (y) and along it (-x) equally: the slab's normal force m g_n makes
    the friction cone mu g_n h^2 against the tangential pull g_t h^2, so corners with mu = 3 stick at their anchors
    and corners with mu = 0.5 slide, the anchor dragged along at their in-plane position."""
    gp = STIFF.with_(friction=mu, contact_compliance=0.0)
    g, o = free_particles(oracle_lib, gp, gravity=(-10.0, -10.0, 0.0))
    hinge = np.array([0.375, 0.5, -0.1])    # every ring corner well inside the box along the jaw axis
    o.set_tool(jaw_tool(hinge, half_width=1.0, length=1.0), reset=True)
    gs = o.grasp_state()
    ring = np.nonzero(gs["grasped"] & (1 - gs["core"]))[0]
    assert len(ring) >= 8
    for _ in range(3):
        o.step(DT)
    x3 = o.state()[0][ring].copy()
    for _ in range(20):
        o.step(DT)
    x = o.state()[0][ring]
    u, v, n = jaw_coords(x, hinge)
    assert np.abs(n).max() < 1e-9                                   # the slab holds them in the band
    if sticks:
        np.testing.assert_allclose(x, x3, atol=1e-12)
    else:
        assert (x[:, 0] < x3[:, 0] - 1e-3).all()                    # sliding along -x
        np.testing.assert_allclose(o.grasp_state()["anchor"][ring], np.stack((u, v), axis=1), atol=1e-12)


def _ccd_setup(oracle_lib, gp, speed):
    """Free particles moving with `speed` toward a shaft capsule along z through the lattice centre line
    x = 0.4, y = 0.4 of radius 0.2; jaws far away and open."""
    g = block_grid(nodes=(4, 4, 4), spacing=0.25)
    x = g.positions.astype(np.float64)
    axis_pt = np.stack((np.full(len(x), 0.4), np.full(len(x), 0.4), x[:, 2]), axis=1)
    d = x - axis_pt
    r = np.linalg.norm(d, axis=1)
    vel = (-speed * d / r[:, None]).astype(np.float32).astype(np.float64)   # the state is handed over in float
    gg, o = free_particles(oracle_lib, gp)
    o.set_state(x, vel)
    tool = jaw_tool((9.0, 9.0, 9.0), grasp01=0.0, gap=0.1, shaft=((0.4, 0.4, -1.0), (0.4, 0.4, 2.0)), shaft_r=0.2)
    o.set_tool(tool, reset=True)
    o.step(DT)
    predicted = x + DT * vel
    return o, predicted


def test_ccd_gate_pushes_fast_corners_to_the_capsule_surface(oracle_lib):
    gp = STIFF.with_(ccd_omega=1.0, contact_compliance=0.0, ccd_displacement=1e9, ccd_velocity=0.5)
    o, pred = _ccd_setup(oracle_lib, gp, speed=2.0)
    x = o.state()[0]
    rp = np.hypot(pred[:, 0] - 0.4, pred[:, 1] - 0.4)
    r = np.hypot(x[:, 0] - 0.4, x[:, 1] - 0.4)
    inside = rp < 0.2 - 1e-9
    assert inside.sum() >= 4 and (~inside).sum() >= 4
    np.testing.assert_allclose(r[inside], 0.2, atol=1e-12)
    np.testing.assert_allclose(x[~inside], pred[~inside], atol=1e-12)


def test_ccd_gate_ignores_slow_corners_and_caps_the_push(oracle_lib):
    gp = STIFF.with_(ccd_omega=1.0, contact_compliance=0.0, ccd_displacement=1e9, ccd_velocity=0.5)
    o, pred = _ccd_setup(oracle_lib, gp, speed=0.3)
    np.testing.assert_allclose(o.state()[0], pred, atol=1e-12)
    cap = float(np.float32(0.01))            # the parameters are float
    o, pred = _ccd_setup(oracle_lib, gp.with_(ccd_displacement=cap), speed=2.0)
    x = o.state()[0]
    moved = np.linalg.norm(x - pred, axis=1)
    rp = np.hypot(pred[:, 0] - 0.4, pred[:, 1] - 0.4)
    deep = rp < 0.2 - cap - 1e-6
    assert deep.sum() >= 2
    np.testing.assert_allclose(moved[deep], cap, rtol=1e-9)
    assert (moved <= cap * (1 + 1e-9)).all()


def test_ccd_jaw_slab_pushes_free_corners_back_to_the_gap(oracle_lib):
    """Corners inside the jaw box beyond half the gap, moving toward the mid-plane faster than the threshold, end on
    the gap's edge; captured corners are left to the grasp."""
    gp = STIFF.with_(ccd_omega=1.0, contact_compliance=0.0, ccd_displacement=1e9, ccd_velocity=0.5)
    g = block_grid(nodes=(4, 4, 4), spacing=0.25)
    x = g.positions.astype(np.float64)
    hinge = np.array([0.4, 0.4, -0.1])
    vel = np.zeros_like(x)
    vel[:, 1] = -2.0 * np.sign(x[:, 1] - hinge[1])                  # toward the mid-plane y = 0.4
    _, o = free_particles(oracle_lib, gp)
    o.set_state(x, vel)
    o.set_tool(jaw_tool(hinge, grasp01=0.0, gap=0.2, half_width=0.5, length=1.0), reset=True)
    o.step(DT)
    pred = x + DT * vel
    u, v, n = jaw_coords(pred, hinge)
    box = (u >= 0) & (u <= 1.0) & (np.abs(v) <= 0.5)
    hit = box & (np.abs(n) > 0.1 + 1e-9)
    assert hit.sum() >= 4
    xn = o.state()[0]
    _, _, n_after = jaw_coords(xn, hinge)
    np.testing.assert_allclose(np.abs(n_after[hit]), 0.1, atol=1e-12)
    np.testing.assert_allclose(xn[~hit], pred[~hit], atol=1e-12)


def test_fp32_grasp_tracks_fp64(oracle_lib):
    """Capture, hold and lift on the C1 block: FP32 stays within 1e-4 of FP64. The CCD gate is switched off here
    (threshold 1e30): its velocity threshold is a jump in the inputs, so a corner approaching at nearly the threshold
    speed is corrected in one precision and not in the other, and no FP32 bound holds across it."""
    w = block_6144()
    sims = []
    for double in (False, True):
        o = oracle_lib.Oracle(w.grid, w.params, w.initial_positions, double=double)
        o.set_tet_stiffness(*w.stiffness())
        o.set_gravity(w.gravity)
        o.set_contact_model(0, legacy_grasp_params(0.25).with_(ccd_velocity=1e30))
        sims.append(o)
    for s in range(12):
        hinge = (2.0, 1.95 + 0.02 * max(0, s - 3), 0.4)
        for o in sims:
            o.set_tool(jaw_tool(hinge, grasp01=1.0 if s >= 2 else 0.0, half_width=0.55, length=1.1), reset=(s == 0))
            o.step(w.dt)
    g32, g64 = sims[0].grasp_state(), sims[1].grasp_state()
    np.testing.assert_array_equal(g32["grasped"], g64["grasped"])
    assert g32["grasping"] and g32["grasped"].sum() >= 8
    p32, p64 = sims[0].state()[0], sims[1].state()[0]
    assert np.abs(p32 - p64).max() / np.abs(p64).max() < 1e-4
