"""Pins of the oracle's surface depth (oracle_surface_depth: the grasper part of LCS_SurfaceDepth, k_SurfaceDepth with
triangleWitness, surface_contact.cuh:169-198, 662-674), no GPU: against a flat sheet of triangles the depth of a
capsule or a jaw box pressed into it is the press depth; a shape clear of the sheet, or a soup outside its bounds,
gives zero; the deepest of several shapes counts."""
import numpy as np

from oracle import oracle as orc


def sheet(y=0.0, half=3.0, n=6):
    """A flat square sheet of 2 n^2 triangles in the plane y, as a soup [K, 3]."""
    xs = np.linspace(-half, half, n + 1)
    tris = []
    for i in range(n):
        for j in range(n):
            a, b = (xs[i], y, xs[j]), (xs[i + 1], y, xs[j])
            c, d = (xs[i + 1], y, xs[j + 1]), (xs[i], y, xs[j + 1])
            tris += [a, b, c, a, c, d]
    return np.array(tris, dtype=np.float32)


def tool(shaft_y=5.0, shaft_r=0.4, hinge_y=5.0, gap=0.5, closure=0.0):
    """A horizontal shaft along x at height shaft_y, and jaws hinged at (1, hinge_y, 1) pointing along +z, opening in
    y (normal along y), far from each other unless placed."""
    return dict(shaftA=(-1.0, shaft_y, 0.2), shaftB=(1.0, shaft_y, 0.2), shaftR=shaft_r,
                jawHinge=(1.0, hinge_y, 1.0), jawAxis=(0.0, 0.0, 1.0), jawMid=(1.0, hinge_y, 1.5),
                jawNormal=(0.0, 1.0, 0.0), jawHalfWidth=0.25, jawLen=1.0, jawGap=gap, grasp01=closure)


def test_capsule_pressed_into_a_sheet_gives_its_press_depth():
    s = sheet()
    for d in (0.05, 0.1, 0.25):
        depth = orc.surface_depth(tool(shaft_y=0.4 - d), 0.1, 0.0, s)
        assert abs(depth - d) < 2e-6, (d, depth)


def test_clear_shapes_and_distant_soups_give_zero():
    s = sheet()
    assert orc.surface_depth(tool(shaft_y=0.41), 0.1, 0.0, s) == 0.0
    assert orc.surface_depth(tool(shaft_y=0.2), 0.1, 0.0, s + np.float32(50.0)) == 0.0
    assert orc.surface_depth(tool(shaft_y=0.2), 0.1, 0.0, np.zeros((0, 3), np.float32)) == 0.0


def test_a_jaw_box_pressed_into_a_sheet_gives_its_press_depth():
    """Closed jaws (gap 0) hinged above the sheet: the lower jaw box (thickness 0.1, below the hinge by its half
    thickness plus nothing) reaches hinge - 0.1; lowering the hinge by d below 0.1 presses it d into the sheet."""
    s = sheet()
    for d in (0.02, 0.05):
        depth = orc.surface_depth(tool(hinge_y=0.1 - d, gap=0.0, closure=1.0), 0.1, 0.0, s)
        assert abs(depth - d) < 2e-6, (d, depth)


def test_the_deepest_shape_counts():
    s = sheet()
    both = orc.surface_depth(tool(shaft_y=0.4 - 0.2, hinge_y=0.1 - 0.05, gap=0.0, closure=1.0), 0.1, 0.0, s)
    assert abs(both - 0.2) < 2e-6


def test_only_the_kinds_asked_for_count():
    s = sheet()
    t = tool(shaft_y=0.4 - 0.2, hinge_y=0.1 - 0.05, gap=0.0, closure=1.0)
    t.update(bladeA=(-2.0, 0.1, -1.0), bladeB=(-2.0, 0.1, 1.0), bladeR=0.4)
    assert abs(orc.surface_depth(t, 0.1, 0.0, s, kinds=0b0001) - 0.2) < 2e-6
    assert abs(orc.surface_depth(t, 0.1, 0.0, s, kinds=0b0110) - 0.05) < 2e-6
    assert abs(orc.surface_depth(t, 0.1, 0.0, s, kinds=orc.BLADE_KIND) - 0.3) < 2e-6
    assert orc.surface_depth(t, 0.1, 0.0, s, kinds=0) == 0.0


def closest_point_distance(p, a, b, c):
    """Distance from p to triangle abc by brute force over a fine barycentric grid refined around its minimum."""
    best = np.inf
    u = np.linspace(0, 1, 801)
    U, W = np.meshgrid(u, u)
    keep = U + W <= 1
    U, W = U[keep], W[keep]
    q = a + np.outer(U, b - a) + np.outer(W, c - a)
    best = np.linalg.norm(q - p, axis=1).min()
    return best


def test_a_sphere_against_a_lone_triangle_gives_radius_minus_its_distance():
    """A capsule of zero length is a sphere: its depth into one triangle is r minus the distance from its centre to
    the triangle. The witness is a point of the triangle, so the depth never exceeds that; for triangles of the size
    of a lattice cell (here under 0.3 across, the sphere 0.1-0.5 off the plane) the Frank-Wolfe steps with their
    golden-section searches land within a few thousandths of it, wherever the closest point falls (interior, edge,
    corner). On large triangles twelve steps leave the interior optimum short, as the reference's do."""
    rng = np.random.default_rng(7)
    for case in range(24):
        a, b, c = rng.normal(size=(3, 3)) * (0.1 if case < 16 else 1.5)
        w = rng.dirichlet((1, 1, 1)) * (1.6 if case % 2 else 1.0) - (0.2 if case % 2 else 0.0)
        nrm = np.cross(b - a, c - a)
        nrm /= np.linalg.norm(nrm)
        centre = a * w[0] + b * w[1] + c * w[2] + nrm * rng.uniform(0.1, 0.5)
        exact = closest_point_distance(centre, a, b, c)
        r = float(exact + 0.3)
        t = tool(shaft_y=50.0)
        t.update(shaftA=tuple(centre), shaftB=tuple(centre), shaftR=r)
        soup = np.array([a, b, c], dtype=np.float32)
        depth = orc.surface_depth(t, 0.1, 0.0, soup, kinds=0b0001)
        assert depth <= 0.3 + 1e-5, (case, depth)
        if case < 16:
            assert abs(depth - 0.3) < 3e-3, (case, depth)
