// Where a lattice node, moving from x0 to x1 during a substep, first touches a tool shape that moves from pose `from`
// to pose `to`, and the contact record the friction solve uses: the outward normal, the target point on the inflated
// surface and the tool's own motion at the touch point.
//
// Reference: contact/friction_contact.cuh:40-142 (shapeDistance, shapeLerp, rotateShaftVector, witnessAt, pointBoxTime,
// traceToolContact). The method is the reference's -- closest-point distances to a capsule or an oriented box, a box
// that only translates swept exactly (its squared distance to a moving point is piecewise quadratic in time), anything
// else by conservative advancement -- formulated here around a surface query record, the box's local frame and a
// piece-local time variable. The CPU checker restates this formulation step for step, so both round alike.
#pragma once
#include "../../include/gts_xpbd.h"
#include "gts_common.h"

// The tool pose of a C-ABI tool state (XpbdToolState, xpbd_solver.h:223-229, plus the blade capsule).
inline ToolPose pose_of(const GtsToolState* t) {
    ToolPose p;
    p.shaftA = ld3(t->shaftA); p.shaftB = ld3(t->shaftB); p.shaftR = t->shaftR;
    p.jawHinge = ld3(t->jawHinge); p.jawAxis = ld3(t->jawAxis); p.jawMid = ld3(t->jawMid);
    p.jawNormal = ld3(t->jawNormal);
    p.jawHalfWidth = t->jawHalfWidth; p.jawLen = t->jawLen; p.jawGap = t->jawGap; p.grasp01 = t->grasp01;
    p.bladeA = ld3(t->bladeA); p.bladeB = ld3(t->bladeB); p.bladeR = t->bladeR;
    return p;
}

struct ContactRecord {
    f3 normal, target, motion;
    int valid;
};

// A point against a shape: signed gap (negative inside), outward unit normal, nearest point of the shape's surface.
struct ToolHit {
    float gap;
    f3 normal, foot;
};

// A unit vector across the spine direction, built on the coordinate axis least aligned with it.
__device__ __forceinline__ f3 across_spine(f3 spine) {
    f3 ax = fabsf(spine.x) <= fabsf(spine.y) && fabsf(spine.x) <= fabsf(spine.z)
                ? mk3(1, 0, 0)
                : (fabsf(spine.y) <= fabsf(spine.z) ? mk3(0, 1, 0) : mk3(0, 0, 1));
    return unit_or(cross3(spine, ax), mk3(0, 0, 1));
}

// Parameter of the spine point nearest x, clamped to the segment.
__device__ __forceinline__ float spine_param(const Shape& s, f3 x) {
    f3 spine = s.b - s.a;
    return fminf(1.f, fmaxf(0.f, dot3(x - s.a, spine) / fmaxf(dot3(spine, spine), 1e-20f)));
}

__device__ __forceinline__ ToolHit capsule_hit(const Shape& s, f3 x) {
    f3 spine = s.b - s.a;
    f3 base = s.a + spine * spine_param(s, x);
    f3 off = x - base;
    float dist = len3(off);
    ToolHit h;
    // A node on the spine: the offset is projection roundoff along the spine, not a direction out of the capsule.
    h.normal = dist > 1e-6f * fmaxf(1.f, len3(spine)) ? off * (1.f / dist) : across_spine(spine);
    h.foot = base + h.normal * s.radius;
    h.gap = dist - s.radius;
    return h;
}

__device__ __forceinline__ ToolHit box_hit(const Shape& s, f3 x) {
    const f3 r = x - s.center;
    const float l0 = dot3(r, s.u), l1 = dot3(r, s.n), l2 = dot3(r, s.w);  // local coordinates
    const float x0 = fabsf(l0) - s.half.x, x1 = fabsf(l1) - s.half.y, x2 = fabsf(l2) - s.half.z;  // beyond each face
    ToolHit h;
    if (x0 > 0.f || x1 > 0.f || x2 > 0.f) {
        // outside: the offset from the box, assembled in the box frame from the local excesses
        f3 off = s.u * copysignf(fmaxf(x0, 0.f), l0) + s.n * copysignf(fmaxf(x1, 0.f), l1) +
                 s.w * copysignf(fmaxf(x2, 0.f), l2);
        float dist = len3(off);
        h.normal = dist > 0.f ? off * (1.f / dist) : s.n;
        h.foot = s.center + s.u * fminf(s.half.x, fmaxf(-s.half.x, l0)) + s.n * fminf(s.half.y, fmaxf(-s.half.y, l1)) +
                 s.w * fminf(s.half.z, fmaxf(-s.half.z, l2));
        h.gap = dist;
        return h;
    }
    // inside: out through the face of least penetration (the first such axis on ties)
    float depth = x0, side = l0;
    f3 ax = s.u;
    if (x1 > depth) { depth = x1; side = l1; ax = s.n; }
    if (x2 > depth) { depth = x2; side = l2; ax = s.w; }
    h.normal = ax * (side >= 0.f ? 1.f : -1.f);
    h.foot = x - h.normal * depth;
    h.gap = depth;
    return h;
}

__device__ __forceinline__ ToolHit tool_hit(const Shape& s, f3 x) { return s.capsule ? capsule_hit(s, x) : box_hit(s, x); }

// The shape part way between two poses: points interpolated, the frame interpolated and made orthonormal again.
__device__ __forceinline__ Shape pose_between(const Shape& from, const Shape& to, float t) {
    Shape s = to;
    const float k = 1.f - t;
    s.center = from.center * k + to.center * t;
    s.a = from.a * k + to.a * t;
    s.b = from.b * k + to.b * t;
    s.u = unit_or(from.u * k + to.u * t, to.u);
    f3 n = from.n * k + to.n * t;
    s.n = unit_or(n - s.u * dot3(s.u, n), to.n);
    s.w = cross3(s.u, s.n);
    return s;
}

// The rotation taking the spine direction of `from` onto that of `to`, applied to r (Rodrigues' formula with the
// sine folded into the cross product; a half turn about a perpendicular when the spines are opposite).
__device__ __forceinline__ f3 swing(const Shape& from, const Shape& to, f3 r) {
    f3 p = from.b - from.a, q = to.b - to.a;
    if (len3(p) < 1e-8f || len3(q) < 1e-8f) return r;
    p = unit_or(p, mk3(0, 0, 1));
    q = unit_or(q, p);
    const float c = fminf(1.f, fmaxf(-1.f, dot3(p, q)));
    if (c < -0.9999f) {
        f3 half_turn = unit_or(cross3(p, fabsf(p.x) < 0.9f ? mk3(1, 0, 0) : mk3(0, 1, 0)), mk3(0, 0, 1));
        return half_turn * (2.f * dot3(half_turn, r)) - r;
    }
    const f3 v = cross3(p, q);
    const f3 vr = cross3(v, r);
    return r + vr + cross3(v, vr) * (1.f / (1.f + c));
}

// Where the material point of the tool at p (on pose `from`) sits on pose `to`: along the spine by the same
// parameter, swung with it; in a box by the same local coordinates.
__device__ __forceinline__ f3 carry(const Shape& from, const Shape& to, f3 p) {
    if (from.capsule) {
        const float t = spine_param(from, p);
        const f3 spine = from.b - from.a;
        return to.a + (to.b - to.a) * t + swing(from, to, p - from.a - spine * t);
    }
    const f3 r = p - from.center;
    return to.center + to.u * dot3(r, from.u) + to.n * dot3(r, from.n) + to.w * dot3(r, from.w);
}

// First time in [0, 1] at which a point moving from a to b (both relative to the box centre) comes within `radius`
// of a box that does not rotate. Along the path the squared distance to the box is a quadratic in time between the
// times at which a local coordinate crosses a face plane; each piece is solved in its own time variable s = t - lo:
// D(lo + s) - radius^2 = A s^2 + B s + C with C the value at lo. The square roots run in double as the reference's
// solve does.
__device__ __forceinline__ void order2(float& a, float& b) {
    const float lo = fminf(a, b), hi = fmaxf(a, b);
    a = lo;
    b = hi;
}

__device__ __forceinline__ bool box_touch_time(const Shape& box, float radius, f3 a, f3 b, float& time) {
    const f3 d = b - a;
    const float p[3] = {dot3(a, box.u), dot3(a, box.n), dot3(a, box.w)};
    const float v[3] = {dot3(d, box.u), dot3(d, box.n), dot3(d, box.w)};
    const float h[3] = {box.half.x, box.half.y, box.half.z};
    // piece boundaries: 0, the face-plane crossings inside (0, 1), 1; an absent crossing is 2, past the end
    float cut[8];
    cut[0] = 0.f;
    cut[7] = 1.f;
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const float t = ((side ? h[k] : -h[k]) - p[k]) / v[k];
            cut[1 + 2 * k + side] = fabsf(v[k]) > 1e-12f && t > 0.f && t < 1.f ? t : 2.f;
        }
    // Batcher's odd-even merge network for eight values, in registers
    order2(cut[0], cut[1]); order2(cut[2], cut[3]); order2(cut[4], cut[5]); order2(cut[6], cut[7]);
    order2(cut[0], cut[2]); order2(cut[1], cut[3]); order2(cut[4], cut[6]); order2(cut[5], cut[7]);
    order2(cut[1], cut[2]); order2(cut[5], cut[6]);
    order2(cut[0], cut[4]); order2(cut[1], cut[5]); order2(cut[2], cut[6]); order2(cut[3], cut[7]);
    order2(cut[2], cut[4]); order2(cut[3], cut[5]);
    order2(cut[1], cut[2]); order2(cut[3], cut[4]); order2(cut[5], cut[6]);
#pragma unroll
    for (int i = 0; i < 7; i++) {
        const float lo = cut[i], hi = cut[i + 1];
        if (hi > 1.f) break;
        const float mid = 0.5f * (lo + hi);
        double A = 0.0, B = 0.0, C = -(double)radius * radius;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float at_mid = p[k] + v[k] * mid;
            if (fabsf(at_mid) <= h[k]) continue;  // inside this slab over the whole piece
            const double e = (double)(p[k] - copysignf(h[k], at_mid)) + (double)v[k] * lo;  // excess at lo
            A += (double)v[k] * v[k];
            B += 2.0 * e * v[k];
            C += e * e;
        }
        if (C <= 1e-14) { time = lo; return true; }
        if (A <= 1e-24) continue;
        const double disc = B * B - 4.0 * A * C;
        if (disc < 0.0) continue;
        const double s = (-B - sqrt(disc)) / (2.0 * A);
        if (s >= -1e-8 && s <= (double)(hi - lo) + 1e-8) {
            time = fminf(hi, fmaxf(lo, (float)((double)lo + s)));
            return true;
        }
    }
    return false;
}

// A direction at pose `at` as the tool carries it to pose `to`: swung with a capsule's spine, re-expressed in a box's
// frame.
__device__ __forceinline__ f3 carry_direction(const Shape& at, const Shape& to, f3 n) {
    if (to.capsule) return swing(at, to, n);
    return unit_or(to.u * dot3(n, at.u) + to.n * dot3(n, at.n) + to.w * dot3(n, at.w), n);
}

// The contact record of a touch at pose `hit` with surface point `foot`: the tool moves that point from its place on
// `from` to its place on `to` over the substep.
__device__ __forceinline__ ContactRecord record_at(const Shape& from, const Shape& hit, const Shape& to, f3 normal,
                                                   f3 foot, float radius) {
    ContactRecord c;
    const f3 end = carry(hit, to, foot);
    c.normal = normal;
    c.target = end + normal * radius;
    c.motion = end - carry(hit, from, foot);
    c.valid = 1;
    return c;
}

__device__ __forceinline__ ContactRecord trace_tool_contact(const Shape& from, const Shape& to, float radius, f3 x0,
                                                            f3 x1) {
    ContactRecord none;
    none.normal = none.target = none.motion = mk3(0, 0, 0);
    none.valid = 0;
    // a box whose frame does not turn over the substep: the exact first-touch time
    if (!to.capsule && len3(to.u - from.u) + len3(to.n - from.n) < 1e-7f) {
        float time = 0.f;
        if (!box_touch_time(to, radius, x0 - from.center, x1 - to.center, time)) return none;
        const Shape hit = pose_between(from, to, time);
        const ToolHit q = tool_hit(hit, x0 * (1.f - time) + x1 * time);
        return record_at(from, hit, to, q.normal, q.foot, radius);
    }
    // otherwise conservative advancement: the gap closes at most at `speed` per unit time, so stepping by a fraction
    // of gap / speed cannot step over a touch
    float speed = len3(x1 - x0);
    if (to.capsule) speed += fmaxf(len3(to.a - from.a), len3(to.b - from.b));
    else speed += len3(to.center - from.center) + 4.f * len3(to.half) * (len3(to.u - from.u) + len3(to.n - from.n));
    // a swept bounding sphere of the tool rejects far nodes at once
    const f3 mid0 = to.capsule ? (from.a + from.b) * 0.5f : from.center;
    const f3 mid1 = to.capsule ? (to.a + to.b) * 0.5f : to.center;
    const float reach = to.capsule ? fmaxf(len3(from.b - from.a), len3(to.b - to.a)) * 0.5f + to.radius : len3(to.half);
    if (len3(x0 - mid0) > reach + radius + len3(x1 - x0) + len3(mid1 - mid0) + 1e-5f) return none;
    const float tol = fmaxf(1e-6f, 1e-4f * (reach + radius));
    double t = 0.0;  // time accumulates in double: grazing paths take many small steps
    while (t < 1.0) {
        const Shape at = pose_between(from, to, (float)t);
        const ToolHit q = tool_hit(at, x0 * (1.f - (float)t) + x1 * (float)t);
        const float gap = q.gap - radius;
        if (gap <= tol) return record_at(from, at, to, carry_direction(at, to, q.normal), q.foot, radius);
        if (speed <= 1e-12f) return none;
        const float step = 0.8f * gap / speed;
        if ((float)t + step >= 1.f) {
            const ToolHit e = tool_hit(to, x1);
            if (e.gap - radius > tol) return none;
            return record_at(from, to, to, carry_direction(to, to, e.normal), e.foot, radius);
        }
        t += step;
    }
    return none;
}
