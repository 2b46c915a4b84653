// Shared arithmetic of the HIP tissue kernels and their host-side set-up code.
//
// Every function here is written so that host and device evaluate the same sequence of IEEE single-precision
// operations: the library is compiled with -ffp-contract=off (no multiply-add fusion) and correctly rounded
// division and square root, so a kernel and a host call of the same function round identically.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GTS_HD __host__ __device__ __forceinline__

struct f3 {
    float x, y, z;
};
GTS_HD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
GTS_HD f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
GTS_HD f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
GTS_HD f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
GTS_HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
GTS_HD float len3(f3 a) { return sqrtf(dot3(a, a)); }
GTS_HD f3 cross3(f3 a, f3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
GTS_HD float comp3(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
GTS_HD f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }

// Kuhn split of a cell around its main diagonal 0-7, corners in bit order i = dx | dy<<1 | dz<<2
// (reference: xpbd/xpbd_solver.cu:1624). Packed 4 x 3 bits per type.
GTS_HD int kuhn_corner(int type, int k) {
    // {0,1,3,7},{0,3,2,7},{0,2,6,7},{0,6,4,7},{0,4,5,7},{0,5,1,7}
    const uint32_t packed[6] = {0u | 1u << 3 | 3u << 6 | 7u << 9, 0u | 3u << 3 | 2u << 6 | 7u << 9,
                                0u | 2u << 3 | 6u << 6 | 7u << 9, 0u | 6u << 3 | 4u << 6 | 7u << 9,
                                0u | 4u << 3 | 5u << 6 | 7u << 9, 0u | 5u << 3 | 1u << 6 | 7u << 9};
    return (int)((packed[type] >> (3 * k)) & 7u);
}

// Node offset of bit-order corner c in a lattice with nx, nxy points per row / layer.
GTS_HD int corner_offset(int c, int nx, int nxy) { return (c & 1) + ((c >> 1) & 1) * nx + ((c >> 2) & 1) * nxy; }

// J^(-2/3) for J >= 0.05: a bit-level cube-root seed, three Newton steps c -= (c^3 - J)/(3c^2), then 1/c^2. The
// reference evaluates powf(J, -2/3) (xpbd/xpbd_solver.cu:773); this fixed IEEE sequence stays within a few ulp of it
// and is reproduced exactly by the CPU checker.
GTS_HD float iso_pow(float J) {
    uint32_t i = __builtin_bit_cast(uint32_t, J);
    i = i / 3u + 0x2a514067u;
    float c = __builtin_bit_cast(float, i);
    for (int k = 0; k < 3; k++) {
        float c2 = c * c;
        c = c - (c2 * c - J) / (3.f * c2);
    }
    return 1.f / (c * c);
}

// Natural log for the guarded bulk strain log J (the reference calls logf, xpbd/xpbd_solver.cu:780): split x = m 2^e
// with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with s = (m-1)/(m+1) summed to s^11, e ln2 in two parts. A fixed
// IEEE sequence, so the CPU checker reproduces it exactly; within about 1 ulp of the exact log for normal x > 0.
GTS_HD float own_log(float x) {
    uint32_t i = __builtin_bit_cast(uint32_t, x);
    int e = (int)((i >> 23) & 0xffu) - 127;
    float m = __builtin_bit_cast(float, (i & 0x7fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    float s = (m - 1.f) / (m + 1.f);
    float s2 = s * s;
    float p = s2 * (0.333333343f + s2 * (0.2f + s2 * (0.142857149f + s2 * (0.111111112f + s2 * 0.0909090936f))));
    float lm = 2.f * s + 2.f * s * p;
    float fe = (float)e;
    return fe * 0.693145752f + (fe * 1.42860677e-06f + lm);
}

// ---------------------------------------------------------------------------------------------------------------
// Tool shapes and the swept node contact (reference: contact/friction_contact.cuh:17-169).
// ---------------------------------------------------------------------------------------------------------------
struct Shape {
    f3 center, u, n, w, half, a, b;
    float radius;
    int capsule;
};
struct ToolPose {
    f3 shaftA, shaftB;
    float shaftR;
    f3 jawHinge, jawAxis, jawMid, jawNormal;
    float jawHalfWidth, jawLen, jawGap, grasp01;
    f3 bladeA, bladeB;
    float bladeR;
};

GTS_HD f3 unit_or(f3 x, f3 fallback) {
    float l = len3(x);
    return l > 1e-8f ? x * (1.f / l) : fallback;
}

GTS_HD ToolPose lerp_tool(const ToolPose& from, const ToolPose& to, float t) {
    ToolPose r = to;
    r.jawHinge = from.jawHinge * (1.f - t) + to.jawHinge * t;
    r.jawAxis = from.jawAxis * (1.f - t) + to.jawAxis * t;
    r.jawNormal = from.jawNormal * (1.f - t) + to.jawNormal * t;
    r.shaftA = from.shaftA * (1.f - t) + to.shaftA * t;
    r.shaftB = from.shaftB * (1.f - t) + to.shaftB * t;
    r.jawGap = from.jawGap * (1.f - t) + to.jawGap * t;
    r.bladeA = from.bladeA * (1.f - t) + to.bladeA * t;
    r.bladeB = from.bladeB * (1.f - t) + to.bladeB * t;
    return r;
}

// kind 0: shaft capsule; 1, 2: the two jaw plates; 3: the blade capsule.
GTS_HD Shape tool_shape(const ToolPose& t, int kind, float thickness, float minGap) {
    Shape s = {};
    s.capsule = (kind == 0 || kind == 3);
    if (kind == 0) { s.a = t.shaftA; s.b = t.shaftB; s.radius = t.shaftR; return s; }
    if (kind == 3) { s.a = t.bladeA; s.b = t.bladeB; s.radius = t.bladeR; return s; }
    f3 u = unit_or(t.jawAxis, mk3(0, 0, 1));
    f3 n = t.jawNormal - u * dot3(u, t.jawNormal);
    n = unit_or(n, fabsf(u.y) < 0.9f ? unit_or(cross3(u, mk3(0, 1, 0)), mk3(1, 0, 0)) : mk3(1, 0, 0));
    float side = kind == 1 ? 1.f : -1.f;
    float sn = fminf(0.95f, fmaxf(0.f, t.jawGap) / (2.f * fmaxf(t.jawLen, 1e-6f)));
    float cs = sqrtf(1.f - sn * sn);
    s.u = u * cs + n * (side * sn);
    s.n = n * cs - u * (side * sn);
    s.w = cross3(s.u, s.n);
    s.half = mk3(t.jawLen * 0.5f, thickness * 0.5f, t.jawHalfWidth);
    s.center = t.jawHinge + s.u * s.half.x + s.n * (side * (s.half.y + 0.5f * minGap));
    return s;
}

GTS_HD bool shape_present(const Shape& s) { return s.capsule ? s.radius > 0.f : (s.half.x > 0.f && s.half.z > 0.f); }

struct StepShapes {
    Shape from[4], to[4];
};
