// Swept tool contact of lattice nodes (reference: contact/friction_contact.cuh:40-169), device side.
// Same operation sequence as the CPU checker's restatement, in single precision except the piecewise-quadratic box
// time, which the reference evaluates in double.
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

__device__ __forceinline__ float shape_distance(const Shape& s, f3 x, f3& normal, f3& surface) {
    if (s.capsule) {
        f3 ab = s.b - s.a;
        float tt = fminf(1.f, fmaxf(0.f, dot3(x - s.a, ab) / fmaxf(dot3(ab, ab), 1e-20f)));
        f3 d = x - (s.a + ab * tt);
        float l = len3(d);
        f3 fallback = unit_or(cross3(ab, fabsf(ab.x) < fabsf(ab.y) ? mk3(1, 0, 0) : mk3(0, 1, 0)), mk3(0, 0, 1));
        normal = l > 1e-6f * fmaxf(1.f, len3(ab)) ? d * (1.f / l) : fallback;
        surface = s.a + ab * tt + normal * s.radius;
        return l - s.radius;
    }
    f3 r = x - s.center;
    float v[3] = {dot3(r, s.u), dot3(r, s.n), dot3(r, s.w)};
    float h[3] = {s.half.x, s.half.y, s.half.z};
    float q[3], cl[3];
    for (int k = 0; k < 3; k++) {
        q[k] = fabsf(v[k]) - h[k];
        cl[k] = fminf(h[k], fmaxf(-h[k], v[k]));
    }
    f3 closest = s.center + s.u * cl[0] + s.n * cl[1] + s.w * cl[2];
    f3 d = s.u * copysignf(fmaxf(q[0], 0.f), v[0]) + s.n * copysignf(fmaxf(q[1], 0.f), v[1]) +
           s.w * copysignf(fmaxf(q[2], 0.f), v[2]);
    float len = len3(d);
    if (q[0] > 0 || q[1] > 0 || q[2] > 0) {
        normal = len > 0 ? d * (1.f / len) : s.n;
        surface = closest;
        return len;
    }
    int k = q[1] > q[0] ? 1 : 0;
    if (q[2] > q[k]) k = 2;
    f3 ax = k == 0 ? s.u : (k == 1 ? s.n : s.w);
    normal = ax * (v[k] >= 0 ? 1.f : -1.f);
    surface = x - normal * q[k];
    return q[k];
}

__device__ __forceinline__ Shape shape_lerp(const Shape& a, const Shape& b, float t) {
    Shape s = b;
    s.center = a.center * (1.f - t) + b.center * t;
    s.a = a.a * (1.f - t) + b.a * t;
    s.b = a.b * (1.f - t) + b.b * t;
    s.u = unit_or(a.u * (1.f - t) + b.u * t, b.u);
    s.n = a.n * (1.f - t) + b.n * t;
    s.n = unit_or(s.n - s.u * dot3(s.u, s.n), b.n);
    s.w = cross3(s.u, s.n);
    return s;
}

__device__ __forceinline__ f3 rotate_shaft_vector(const Shape& from, const Shape& to, f3 r) {
    f3 a = from.b - from.a, b = to.b - to.a;
    if (len3(a) < 1e-8f || len3(b) < 1e-8f) return r;
    a = unit_or(a, mk3(0, 0, 1));
    b = unit_or(b, a);
    float c = fminf(1.f, fmaxf(-1.f, dot3(a, b)));
    f3 v = cross3(a, b);
    if (c < -0.9999f) {
        f3 u = unit_or(cross3(a, fabsf(a.x) < 0.9f ? mk3(1, 0, 0) : mk3(0, 1, 0)), mk3(0, 0, 1));
        return u * (2.f * dot3(u, r)) - r;
    }
    return r + cross3(v, r) + cross3(v, cross3(v, r)) * (1.f / (1.f + c));
}

__device__ __forceinline__ f3 witness_at(const Shape& from, const Shape& to, f3 p) {
    if (from.capsule) {
        f3 ab = from.b - from.a;
        float t = fminf(1.f, fmaxf(0.f, dot3(p - from.a, ab) / fmaxf(dot3(ab, ab), 1e-20f)));
        return to.a + (to.b - to.a) * t + rotate_shaft_vector(from, to, p - from.a - ab * t);
    }
    f3 r = p - from.center;
    return to.center + to.u * dot3(r, from.u) + to.n * dot3(r, from.n) + to.w * dot3(r, from.w);
}

__device__ __forceinline__ bool point_box_time(const Shape& box, float radius, f3 a, f3 b, float& time) {
    float p[3], v[3], half[3], cuts[8];
    cuts[0] = 0.f;
    cuts[1] = 1.f;
    int count = 2;
    for (int k = 0; k < 3; k++) {
        f3 ax = k == 0 ? box.u : (k == 1 ? box.n : box.w);
        p[k] = dot3(a, ax);
        v[k] = dot3(b - a, ax);
        half[k] = comp3(box.half, k);
        if (fabsf(v[k]) > 1e-12f)
            for (int side = -1; side <= 1; side += 2) {
                float t = ((float)side * half[k] - p[k]) / v[k];
                if (t > 0 && t < 1) cuts[count++] = t;
            }
    }
    for (int i = 1; i < count; i++) {
        float x = cuts[i];
        int j = i;
        while (j > 0 && cuts[j - 1] > x) { cuts[j] = cuts[j - 1]; --j; }
        cuts[j] = x;
    }
    for (int i = 0; i < count - 1; i++) {
        double A = 0, B = 0, C = -(double)radius * radius;
        float lo = cuts[i], hi = cuts[i + 1], mid = 0.5f * (lo + hi);
        for (int k = 0; k < 3; k++) {
            float q = p[k] + v[k] * mid;
            if (fabsf(q) <= half[k]) continue;
            double d = p[k] - copysignf(half[k], q);
            A += (double)v[k] * v[k];
            B += 2 * d * v[k];
            C += d * d;
        }
        double atLo = (A * lo + B) * lo + C;
        if (atLo <= 1e-14) { time = lo; return true; }
        if (A <= 1e-24) continue;
        double disc = B * B - 4 * A * C;
        if (disc < 0) continue;
        double root = (-B - sqrt(disc)) / (2 * A);
        if (root >= lo - 1e-8 && root <= hi + 1e-8) { time = fminf(hi, fmaxf(lo, (float)root)); return true; }
    }
    return false;
}

__device__ __forceinline__ ContactRecord trace_tool_contact(const Shape& from, const Shape& to, float radius, f3 x0,
                                                            f3 x1) {
    ContactRecord out;
    out.normal = out.target = out.motion = mk3(0, 0, 0);
    out.valid = 0;
    f3 normal = mk3(0, 0, 0), surface = mk3(0, 0, 0);
    if (!to.capsule && len3(to.u - from.u) + len3(to.n - from.n) < 1e-7f) {
        float time = 0;
        if (!point_box_time(to, radius, x0 - from.center, x1 - to.center, time)) return out;
        Shape hit = shape_lerp(from, to, time);
        shape_distance(hit, x0 * (1.f - time) + x1 * time, normal, surface);
        f3 end = witness_at(hit, to, surface), start = witness_at(hit, from, surface);
        out.normal = normal; out.target = end + normal * radius; out.motion = end - start; out.valid = 1;
        return out;
    }
    float speed = len3(x1 - x0);
    if (to.capsule) speed += fmaxf(len3(to.a - from.a), len3(to.b - from.b));
    else speed += len3(to.center - from.center) + 4.f * len3(to.half) * (len3(to.u - from.u) + len3(to.n - from.n));
    f3 center0 = to.capsule ? (from.a + from.b) * 0.5f : from.center;
    f3 center1 = to.capsule ? (to.a + to.b) * 0.5f : to.center;
    float bound = to.capsule ? fmaxf(len3(from.b - from.a), len3(to.b - to.a)) * 0.5f + to.radius : len3(to.half);
    if (len3(x0 - center0) > bound + radius + len3(x1 - x0) + len3(center1 - center0) + 1e-5f) return out;
    float tol = fmaxf(1e-6f, 1e-4f * (bound + radius));
    double t = 0;
    Shape hit = from;
    bool found = false;
    while (t < 1.) {
        hit = shape_lerp(from, to, (float)t);
        float d = shape_distance(hit, x0 * (1.f - (float)t) + x1 * (float)t, normal, surface) - radius;
        if (d <= tol) { found = true; break; }
        if (speed <= 1e-12f) break;
        float step = 0.8f * d / speed;
        if ((float)t + step >= 1.f) {
            t = 1;
            hit = to;
            float de = shape_distance(to, x1, normal, surface) - radius;
            found = de <= tol;
            break;
        }
        t += step;
    }
    if (!found) return out;
    f3 end = witness_at(hit, to, surface), start = witness_at(hit, from, surface);
    if (to.capsule) normal = rotate_shaft_vector(hit, to, normal);
    else normal = unit_or(to.u * dot3(normal, hit.u) + to.n * dot3(normal, hit.n) + to.w * dot3(normal, hit.w), normal);
    out.normal = normal; out.target = end + normal * radius; out.motion = end - start; out.valid = 1;
    return out;
}
