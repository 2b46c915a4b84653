"""The oracle's swept tool contact and friction solve, pinned by brute force and closed forms (no GPU).

  * the first-touch time of a point moving past a translating box against a fine march with bisection;
  * conservative advancement against a moving capsule and a turning box: a contact exactly when the march finds a
    touch, its target on the inflated surface of the end pose, its normal outward there;
  * one friction pass: the normal solve puts a node on the target plane with lambda_N >= 0 and never pulls; the
    tangential solve cancels the slip below mu_s lambda_N (stick) and moves the node back by exactly mu_d lambda_N w
    above it (slide).
"""
import ctypes as ct

import numpy as np
import pytest


def _bind(L):
    D = ct.c_double
    L.oracle_box_touch_time_f64.argtypes = [ct.c_void_p, D, ct.c_void_p, ct.c_void_p, ct.c_void_p]
    L.oracle_trace_contact_f64.argtypes = [ct.c_void_p, ct.c_void_p, D, ct.c_void_p, ct.c_void_p, ct.c_void_p]
    L.oracle_friction_update_f64.argtypes = [ct.c_void_p, D, D, D, D, ct.c_void_p, ct.c_void_p, ct.c_void_p]
    L.oracle_tool_gap_f64.argtypes = [ct.c_void_p, ct.c_void_p]
    L.oracle_tool_gap_f64.restype = D
    return L


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(ct.c_void_p)


def rotation(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def box(center, R, half):
    v = np.zeros(23)
    v[0:3], v[3:6], v[6:9], v[9:12], v[12:15] = center, R[:, 0], R[:, 1], R[:, 2], half
    return v


def capsule(a, b, r):
    v = np.zeros(23)
    v[15:18], v[18:21], v[21], v[22] = a, b, r, 1
    return v


def box_gap(center, R, half, x):
    """Distance of x to the box (negative inside), in closed form."""
    loc = (np.asarray(x) - center) @ R
    excess = np.abs(loc) - half
    out = np.linalg.norm(np.maximum(excess, 0))
    return out if (excess > 0).any() else excess.max()


def first_touch(f, n=20001):
    """First t in [0, 1] with f(t) <= 0 by a march and 60 bisections, or None."""
    ts = np.linspace(0, 1, n)
    vals = np.array([f(t) for t in ts])
    hit = np.flatnonzero(vals <= 0)
    if len(hit) == 0:
        return None
    k = hit[0]
    if k == 0:
        return 0.0
    lo, hi = ts[k - 1], ts[k]
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        lo, hi = (lo, mid) if f(mid) <= 0 else (mid, hi)
    return hi


def test_box_touch_time_against_a_march(oracle_lib):
    """Random oriented boxes, radii and straight paths (some missing, some starting inside the inflated box)."""
    L = _bind(oracle_lib.lib())
    rng = np.random.default_rng(4)
    hits = misses = 0
    for trial in range(60):
        R = rotation(rng)
        half = rng.uniform(0.1, 1.0, 3)
        r = rng.uniform(0.0, 0.2)
        a = rng.normal(0, 1.5, 3)
        b = rng.normal(0, 1.5, 3) if trial % 5 else a + rng.normal(0, 0.05, 3)
        t = np.zeros(1)
        got = L.oracle_box_touch_time_f64(P(box(np.zeros(3), R, half)), r, P(a), P(b), t.ctypes.data_as(ct.c_void_p))
        expect = first_touch(lambda s: box_gap(np.zeros(3), R, half, a + (b - a) * s) - r)
        assert bool(got) == (expect is not None), (trial, got, expect)
        if got:
            hits += 1
            assert abs(t[0] - expect) <= 2e-9, (trial, t[0], expect)
        else:
            misses += 1
    assert hits > 15 and misses > 5


@pytest.mark.parametrize("kind", ["capsule", "turning_box"])
def test_conservative_advancement_against_a_march(oracle_lib, kind):
    """A moving capsule or a turning box sweeping past nodes on straight paths: a contact exactly when the march over
    the interpolated poses finds the node within the radius (up to the advancement's tolerance), and then a target on
    the end pose's inflated surface with the outward normal there."""
    L = _bind(oracle_lib.lib())
    rng = np.random.default_rng(7 if kind == "capsule" else 8)
    radius = 0.05
    found = 0
    for trial in range(24):
        if kind == "capsule":
            a0 = rng.normal(0, 0.5, 3)
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            shift = rng.normal(0, 0.4, 3)
            frm, to = capsule(a0, a0 + d, 0.2), capsule(a0 + shift, a0 + d + shift, 0.2)

            def gap(s):
                fa = a0 + shift * s
                seg = d
                x = x0 + (x1 - x0) * s
                u = np.clip(np.dot(x - fa, seg) / np.dot(seg, seg), 0, 1)
                return np.linalg.norm(x - (fa + seg * u)) - 0.2 - radius
            reach = 0.7
            centre = a0 + 0.5 * d + 0.5 * shift
        else:
            half = np.array((0.5, 0.1, 0.3))
            R0 = rotation(rng)
            ang = rng.uniform(0.05, 0.3)
            ax = rng.normal(size=3)
            ax /= np.linalg.norm(ax)
            K = np.array(((0, -ax[2], ax[1]), (ax[2], 0, -ax[0]), (-ax[1], ax[0], 0)))
            R1 = (np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K) @ R0
            frm, to = box(np.zeros(3), R0, half), box(np.zeros(3), R1, half)

            def pose(s):  # the oracle's interpolation: lerp the frame, re-orthonormalise u and n, w = u x n
                u = (1 - s) * R0[:, 0] + s * R1[:, 0]
                u /= np.linalg.norm(u)
                n = (1 - s) * R0[:, 1] + s * R1[:, 1]
                n = n - u * np.dot(u, n)
                n /= np.linalg.norm(n)
                return np.stack((u, n, np.cross(u, n)), axis=1)

            def gap(s):
                return box_gap(np.zeros(3), pose(s), half, x0 + (x1 - x0) * s) - radius
            reach = 0.8
            centre = np.zeros(3)
        x0 = centre + rng.normal(0, 0.45, 3)
        # half the paths aim through the tool's middle, the others wander
        x1 = x0 + (centre - x0) * 1.3 + rng.normal(0, 0.05, 3) if trial % 2 else x0 + rng.normal(0, 0.3, 3)
        out = np.zeros(9)
        valid = L.oracle_trace_contact_f64(P(frm), P(to), radius, P(x0), P(x1), out.ctypes.data_as(ct.c_void_p))
        touch = first_touch(gap)
        tol = max(1e-6, 1e-4 * (reach + radius))
        if valid:
            found += 1
            # the march agrees that the node comes within the radius plus the tolerance
            assert first_touch(lambda s: gap(s) - 1.01 * tol) is not None, trial
            normal, target = out[0:3], out[3:6]
            assert abs(np.linalg.norm(normal) - 1) < 1e-9
            assert abs(L.oracle_tool_gap_f64(P(to), P(target)) - radius) < 1e-7, trial
            probe = target + 1e-4 * normal
            assert L.oracle_tool_gap_f64(P(to), P(probe)) > radius, trial  # outward
        else:
            assert touch is None, (trial, touch)
    assert found >= 5


def test_normal_solve_lands_on_the_target_plane_and_never_pulls(oracle_lib):
    L = _bind(oracle_lib.lib())
    normal = np.array((0.0, 1.0, 0.0))
    rec = np.concatenate((normal, np.zeros(3), np.zeros(3)))
    for depth, w in ((0.02, 1000.0), (0.5, 2.0), (-0.03, 50.0)):  # below the plane, then above it
        x = np.array((0.1, -depth, 0.2))
        prev = x.copy()
        lam = np.zeros(4)
        L.oracle_friction_update_f64(P(rec), w, 0.0, 0.6, 0.4, P(prev), x.ctypes.data_as(ct.c_void_p),
                                     lam.ctypes.data_as(ct.c_void_p))
        assert lam[0] >= 0
        if depth > 0:
            assert abs(x[1]) < 1e-12 and abs(lam[0] - depth / w) < 1e-15
        else:
            assert lam[0] == 0 and x[1] == pytest.approx(-depth)  # separated: untouched
    # with compliance alpha the multiplier is depth / (w + alpha), the node stops short of the plane
    x, lam = np.array((0.0, -0.02, 0.0)), np.zeros(4)
    L.oracle_friction_update_f64(P(rec), 100.0, 300.0, 0.6, 0.4, P(x.copy()), x.ctypes.data_as(ct.c_void_p),
                                 lam.ctypes.data_as(ct.c_void_p))
    assert abs(lam[0] - 0.02 / 400.0) < 1e-15 and abs(x[1] - (-0.02 + 100.0 * 0.02 / 400.0)) < 1e-15


@pytest.mark.parametrize("slip", [0.001, 0.05])
def test_tangential_solve_sticks_below_mu_s_and_slides_at_mu_d(oracle_lib, slip):
    """A node pressed 0.01 into a tool that moves 0.02 along x while the node drifts `slip` further along x: below
    mu_s lambda_N the pass cancels the relative slip (stick); above it the node moves back by mu_d lambda_N w (slide)."""
    L = _bind(oracle_lib.lib())
    w, mus, mud, depth = 10.0, 0.6, 0.4, 0.01
    motion = np.array((0.02, 0.0, 0.0))
    rec = np.concatenate(((0.0, 1.0, 0.0), np.zeros(3), motion))
    prev = np.array((0.0, -depth, 0.0))
    x = prev + motion + np.array((slip, 0.0, 0.0))
    lam = np.zeros(4)
    L.oracle_friction_update_f64(P(rec), w, 0.0, mus, mud, P(prev), x.ctypes.data_as(ct.c_void_p),
                                 lam.ctypes.data_as(ct.c_void_p))
    lam_n = depth / w
    assert abs(lam[0] - lam_n) < 1e-15 and abs(x[1]) < 1e-15
    if slip / w <= mus * lam_n:  # stick: the node follows the tool exactly
        assert abs(x[0] - motion[0]) < 1e-15
        np.testing.assert_allclose(lam[1:], (-slip / w, 0, 0), atol=1e-18)
    else:  # slide: kinetic friction of mu_d lambda_N, against the slip
        assert abs(x[0] - (motion[0] + slip - mud * lam_n * w)) < 1e-15
        np.testing.assert_allclose(lam[1:], (-mud * lam_n, 0, 0), atol=1e-18)
