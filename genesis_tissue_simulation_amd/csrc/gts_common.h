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
// (reference: xpbd/xpbd_solver.cu:1624):
//   {0,1,3,7},{0,3,2,7},{0,2,6,7},{0,6,4,7},{0,4,5,7},{0,5,1,7}
// Every type starts at 0 and ends at 7; its two middle corners are packed 6 bits a type into one immediate, so that
// a kernel finds a tet's nodes without a table in memory (a load that would wait on the tet's descriptor).
GTS_HD int kuhn_corner(int type, int k) {
    if (k == 0) return 0;
    if (k == 3) return 7;
    return (int)((0x36c9b24d9ull >> (6 * type + 3 * (k - 1))) & 7u);
}

// Node offset of bit-order corner c in a lattice with nx, nxy points per row / layer.
GTS_HD int corner_offset(int c, int nx, int nxy) { return (c & 1) + ((c >> 1) & 1) * nx + ((c >> 2) & 1) * nxy; }

// J^(-2/3) for J >= 0.05: y = J^(-1/3) from a bit-level seed (K - bits(J) / 3, within 3.5%) and three Newton steps
// y += y (1 - J y^3) / 3 on y^-3 = J, which need no division, then y^2. The reference evaluates powf(J, -2/3)
// (xpbd/xpbd_solver.cu:773); this fixed IEEE sequence stays within 3 ulp of the exact power and is reproduced exactly
// by the CPU checker.
GTS_HD float iso_pow(float J) {
    uint32_t i = __builtin_bit_cast(uint32_t, J);
    float y = __builtin_bit_cast(float, 0x54a23280u - i / 3u);
    for (int k = 0; k < 3; k++) {
        const float r = 1.f - J * (y * y * y);
        y = y + y * r * 0.333333343f;
    }
    return y * y;
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

// arccos for the bend angle (the reference calls acosf, xpbd/xpbd_solver.cu:553), from the Maclaurin series of
// arcsin, asin x = x + sum_n (2n)! / (4^n (n!)^2 (2n + 1)) x^(2n+1), summed to n = 11 for |x| <= 1/2:
//   |c| <= 1/2:  acos c = pi/2 - asin c;
//   c > 1/2:     acos c = 2 asin sqrt((1 - c)/2);   c < -1/2: acos c = pi - 2 asin sqrt((1 + c)/2)
// (1 -+ c and the halving are exact there). pi and pi/2 enter as a float head plus the float of their residual. A
// fixed IEEE sequence the CPU checker repeats; within 1.1 ulp of the exact arccos on [-1, 1].
GTS_HD float own_acos(float c) {
    const float a = fabsf(c);
    const bool mid = a <= 0.5f;
    const float x = mid ? c : sqrtf((1.f - a) * 0.5f);
    const float x2 = mid ? c * c : (1.f - a) * 0.5f;
    float p = 0.007312526f;           // n = 11: 88179/12058624
    p = 0.008390335f + x2 * p;        // 46189/5505024
    p = 0.009761609f + x2 * p;        // 12155/1245184
    p = 0.011551801f + x2 * p;        // 6435/557056
    p = 0.013964844f + x2 * p;        // 143/10240
    p = 0.017352764f + x2 * p;        // 231/13312
    p = 0.022372158f + x2 * p;        // 63/2816
    p = 0.030381944f + x2 * p;        // 35/1152
    p = 0.044642858f + x2 * p;        // 5/112
    p = 0.075f + x2 * p;              // 3/40
    p = 0.16666667f + x2 * p;         // n = 1: 1/6
    const float tail = x * x2 * p;    // asin x - x
    if (mid) return 1.57079637f - (x + (tail - -4.37113883e-08f));  // pi/2 = 1.57079637f + -4.37113883e-08f
    const float twice = 2.f * (x + tail);
    return c > 0.f ? twice : 3.14159274f - (twice - -8.74227766e-08f);  // pi = 3.14159274f + -8.74227766e-08f
}

// Bend slot pair p = 0..14 of a node: slots (a, b), a < b, in lexicographic order (grid.py:257-260).
GTS_HD int bend_slot_a(int p) { return (int)((0x11B491248000ull >> (3 * p)) & 7ull); }
GTS_HD int bend_slot_b(int p) { return (int)((0x16CB1D8D58D1ull >> (3 * p)) & 7ull); }

// ---------------------------------------------------------------------------------------------------------------
// Precise strain invariants (no reference counterpart; selected by GtsStepParams.precise_strain).
//
// The reference forms J^(-2/3) I1 - 3 and J - 1 from F itself. In FP32 both then carry a rounding noise of about one
// ulp of 3 (2.4e-7) whatever the strain, which is larger than the deviatoric q = O(strain^2) below strains of ~5e-4:
// C_D = sqrt(q) and its gradient direction become noise. The same quantities are formed here from the displacement
// gradient H = F - I = (e_k - r_k) Dm^-1, r_k the rest edges of the Kuhn type (e_k - r_k is exact where an edge is
// near its rest vector), with every cancellation of lower-order terms carried out symbolically:
//   a = J - 1 = tr H + e,  e = tr cof H + det H          b = I1 - 3 = 2 tr H + |H|^2
//   u = J^(-2/3) - 1 = -2a/3 + u2,  u2 = 5a^2/9 + u3     (u3: the binomial series from a^3 on)
//   q = 3u + b + ub = 2 |dev sym H|^2 + 2 tr H e + 5e^2/3 - 2 det H + 3 u3 - 2a|H|^2/3 + u2 b
// (the O(H) terms of 3u + b and the O(H^2) terms besides 2 |dev sym H|^2 cancel identically), and
//   Ghat = J^(-2/3) F - (J^(-2/3) I1 / 3J) cof F = J^(-2/3) [H + H^T - tr H I - cof H - m cof F],  m = (b - 3a) / 3J
// using cof(I + H) = (1 + tr H) I - H^T + cof H. Ghat is half the gradient of q in F; the deviatoric constraint
// gradient of corner k is Ghat Dm^-T row k over C_D. The residual Dm Dm^-1 - I of the float Dm^-1 (~1e-8) is left
// out: it is below the resolution of F in FP32.
// ---------------------------------------------------------------------------------------------------------------

// u = (1 + a)^(-2/3) - 1 with its tails u2 = u + 2a/3 and u3 = u2 - 5a^2/9: the binomial series for |a| <= 1/16 (to
// a^8; the first omitted term is below 2e-10 of u), else -a r^3 (1 + r) / (1 + r + r^2) with r = (1 + a)^(-1/3) =
// sqrt(iso), from r^2 - 1 = (r - 1)(r + 1) and 1 - r^3 = a r^3. `iso` is J^(-2/3) of the same J = 1 + a.
GTS_HD float iso_minus_one(float a, float iso, float& u2, float& u3) {
    if (fabsf(a) <= 0.0625f) {
        float s = 0.36418906f;           // 21505/59049: binomial(-2/3, 8), then down to binomial(-2/3, 3)
        s = -0.38002336f + a * s;        // -7480/19683
        s = 0.39902455f + a * s;         // 2618/6561
        s = -0.42249656f + a * s;        // -308/729
        s = 0.4526749f + a * s;          // 110/243
        s = -0.49382716f + a * s;        // -40/81
        u3 = a * a * a * s;
        u2 = 0.5555556f * a * a + u3;    // 5/9
        return -0.6666667f * a + u2;     // -2/3
    }
    float r = sqrtf(iso);
    float u = -a * iso * r * (1.f + r) / (1.f + r + iso);
    u2 = u + 0.6666667f * a;
    u3 = u2 - 0.5555556f * a * a;
    return u;
}

// log(1 + a) without the rounding of 1 + a: 2 atanh(s), s = a / (2 + a), for |a| < 1/4 (|s| < 1/7, series to s^11),
// else own_log(1 + a).
GTS_HD float own_log1p(float a) {
    if (fabsf(a) < 0.25f) {
        float s = a / (2.f + a);
        float s2 = s * s;
        float p = s2 * (0.333333343f + s2 * (0.2f + s2 * (0.142857149f + s2 * (0.111111112f + s2 * 0.0909090936f))));
        return 2.f * s + 2.f * s * p;
    }
    return own_log(1.f + a);
}

struct PreciseStrain {
    f3 H[3], cH[3];                 // columns of H and of cof H
    float trH, e, det, nH2, a, b;   // tr H, tr cof H + det H, det H, |H|^2, J - 1, I1 - 3
};

// H = d Dm^-1 from the edge displacements d_k = e_k - r_k (columns built as F is: H[c] = sum_k d_k (Dm^-1)_kc).
GTS_HD PreciseStrain precise_strain(const f3 d[3], const f3 br[3]) {
    PreciseStrain s;
    for (int c = 0; c < 3; c++) s.H[c] = d[0] * comp3(br[0], c) + d[1] * comp3(br[1], c) + d[2] * comp3(br[2], c);
    s.trH = s.H[0].x + s.H[1].y + s.H[2].z;
    s.cH[0] = cross3(s.H[1], s.H[2]);
    s.cH[1] = cross3(s.H[2], s.H[0]);
    s.cH[2] = cross3(s.H[0], s.H[1]);
    s.det = dot3(s.H[0], s.cH[0]);
    s.e = (s.cH[0].x + s.cH[1].y + s.cH[2].z) + s.det;
    s.a = s.trH + s.e;
    s.nH2 = dot3(s.H[0], s.H[0]) + dot3(s.H[1], s.H[1]) + dot3(s.H[2], s.H[2]);
    s.b = 2.f * s.trH + s.nH2;
    return s;
}

// F = I + H, cof F = (1 + tr H) I - H^T + cof H and J = 1 + (J - 1) from the precise strain: the same functions of the
// corners as the direct forms, without forming F, its cofactors and determinant a second time.
GTS_HD void precise_frame(const PreciseStrain& s, f3 F[3], f3 cof[3], float& J) {
    const f3* H = s.H;
    const float c1 = 1.f + s.trH;
    F[0] = mk3(H[0].x + 1.f, H[0].y, H[0].z);
    F[1] = mk3(H[1].x, H[1].y + 1.f, H[1].z);
    F[2] = mk3(H[2].x, H[2].y, H[2].z + 1.f);
    cof[0] = mk3(s.cH[0].x + c1 - H[0].x, s.cH[0].y - H[1].x, s.cH[0].z - H[2].x);
    cof[1] = mk3(s.cH[1].x - H[0].y, s.cH[1].y + c1 - H[1].y, s.cH[1].z - H[2].y);
    cof[2] = mk3(s.cH[2].x - H[0].z, s.cH[2].y - H[1].z, s.cH[2].z + c1 - H[2].z);
    J = 1.f + s.a;
}

// q = J^(-2/3) I1 - 3 (unclamped) and the columns of Ghat at pJ = max(J, .05), iso = pJ^(-2/3). Below the J clamp
// (jlive false) pJ is a constant and iso - 1 is far from 0: q = 3u + b + ub with u = iso - 1, Ghat = iso F as the
// reference's gradient there.
GTS_HD float precise_dev(const PreciseStrain& s, const f3 F[3], const f3 cof[3], bool jlive, float pJ, float iso,
                         f3 Gh[3]) {
    if (!jlive) {
        float u = iso - 1.f;
        for (int c = 0; c < 3; c++) Gh[c] = F[c] * iso;
        return 3.f * u + s.b + u * s.b;
    }
    float u2, u3;
    iso_minus_one(s.a, iso, u2, u3);
    const float t3 = s.trH * 0.333333343f;  // FP32 1/3
    float d0 = s.H[0].x - t3, d1 = s.H[1].y - t3, d2 = s.H[2].z - t3;
    float s01 = s.H[1].x + s.H[0].y, s02 = s.H[2].x + s.H[0].z, s12 = s.H[2].y + s.H[1].z;
    float q2 = 2.f * (d0 * d0 + d1 * d1 + d2 * d2) + (s01 * s01 + s02 * s02 + s12 * s12);
    float rest = 2.f * s.trH * s.e + 1.6666666f * s.e * s.e - 2.f * s.det + 3.f * u3 - 0.6666667f * s.a * s.nH2 +
                 u2 * s.b;
    float m = (s.b - 3.f * s.a) / (3.f * pJ);
    // column c of H + H^T - tr H I - cof H - m cof F, tr H taken off the diagonal entry only
    const f3* H = s.H;
    f3 S0 = mk3(H[0].x + H[0].x - s.trH, H[0].y + H[1].x, H[0].z + H[2].x);
    f3 S1 = mk3(H[1].x + H[0].y, H[1].y + H[1].y - s.trH, H[1].z + H[2].y);
    f3 S2 = mk3(H[2].x + H[0].z, H[2].y + H[1].z, H[2].z + H[2].z - s.trH);
    Gh[0] = (S0 - s.cH[0] - cof[0] * m) * iso;
    Gh[1] = (S1 - s.cH[1] - cof[1] * m) * iso;
    Gh[2] = (S2 - s.cH[2] - cof[2] * m) * iso;
    return q2 + rest;
}

// Rest edge k (1..3) of a Kuhn type: the bit offsets of its corner times the axis rest lengths.
GTS_HD f3 kuhn_rest_edge(int type, int k, const float* rest_len) {
    int c = kuhn_corner(type, k);
    return mk3((c & 1) ? rest_len[0] : 0.f, ((c >> 1) & 1) ? rest_len[1] : 0.f, ((c >> 2) & 1) ? rest_len[2] : 0.f);
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
