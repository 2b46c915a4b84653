// CPU oracle of the endoscopic_tissue legacy XPBD substep and swept-quad cut.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker the HIP path is compared with; nothing in the product links,
// loads or calls it. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
//
// It is a serial restatement of the reference algorithm, templated on the real type (float = parity oracle and the
// timed CPU baseline, double = accuracy oracle). Every function cites the reference file:line it follows; paths are
// relative to the reference checkout (endoscopic_tissue/native/vendor/cuda unless noted). Compile with
// -ffp-contract=off so that every product and sum is rounded on its own, as written.
//
// Reference structure restated here:
//   topology     xpbd_bind edge dedup (xpbd/xpbd_solver.cu:1445-1463) + greedy edge colouring (:1213-1249);
//                xpbd_build_tets Kuhn split, Dm^-1 per type, greedy tet colouring (:1600-1848)
//   frame        xpbd_begin_frame (:2481-2514), k_ComputeTetAlpha (:578-587)
//   substep      xpbd_substep legacy branch order (:3460-3721): tool lerp, k_RefreshLiveness (:429-466),
//                fallback edge support (:469-479), k_predict (:393-412), lambda reset, k_solve_edges (:484-524),
//                k_solve_tet (:749-829), friction tool contact (contact/friction_contact.cuh:17-169),
//                k_update_velocity / k_perp_damp_edges / k_apply_velscale (:871-919); k_solve_bending (:534-573)
//                after the edges while kb > 0, over the live bend pairs coloured by xpbd_set_bending (:1344-1356)
//   cut          k_DetectCutQuad (cut.cu:421-474), MollerTrumbore (cut.cuh:162-185), MarkAndEmit/EmitCutPoint
//                (cut.cu:186-254), k_DetectCutCCD (cut.cu:486-592), SeverLink/k_SeverLinks (cut.cu:719-746)
//   compaction   build-side stable remap of live tets / edges / nodes (no reference counterpart)
//   adjoint      reverse of the volume-only substep (no reference counterpart; pinned by finite differences)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <class R> struct V3 { R x, y, z; };
template <class R> inline V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return mk(a.x * s, a.y * s, a.z * s); }
template <class R> inline R dot3(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R len3(V3<R> a) { return std::sqrt(dot3(a, a)); }
template <class R> inline V3<R> cross3(V3<R> a, V3<R> b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> inline R comp(V3<R> v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// J^(-2/3) of k_solve_tet (the reference calls powf, xpbd_solver.cu:773). The float oracle and the HIP kernel both
// evaluate it with this fixed sequence of IEEE operations -- a bit-level seed of J^(-1/3), three division-free Newton
// steps y += y (1 - J y^3) / 3, then y^2 -- so the two agree bit for bit; it stays within 3 ulp of the exact power for
// J >= 0.05. The double oracle uses the library power.
inline float isoPow(float J) {
    uint32_t i;
    std::memcpy(&i, &J, 4);
    i = 0x54a23280u - i / 3u;
    float y;
    std::memcpy(&y, &i, 4);
    for (int k = 0; k < 3; k++) {
        float y3 = y * y * y;
        float r = 1.f - J * y3;
        y = y + y * r * 0.333333343f;
    }
    return y * y;
}
inline double isoPow(double J) { return std::pow(J, -2.0 / 3.0); }

// log of the guarded bulk strain (the reference calls logf, xpbd_solver.cu:780). The float oracle and the kernel share
// this fixed sequence: x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)) summed to s^11, e ln2 in
// two parts. The double oracle uses the library log.
inline float guardLog(float x) {
    uint32_t i;
    std::memcpy(&i, &x, 4);
    int e = (int)((i >> 23) & 0xffu) - 127;
    uint32_t mb = (i & 0x7fffffu) | 0x3f800000u;
    float m;
    std::memcpy(&m, &mb, 4);
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
inline double guardLog(double x) { return std::log(x); }
template <class R> inline R fmin_(R a, R b) { return std::fmin(a, b); }
template <class R> inline R fmax_(R a, R b) { return std::fmax(a, b); }

// Kuhn split of a cell around the diagonal 0-7, corners in bit order i = dx | dy<<1 | dz<<2 (xpbd_solver.cu:1624).
const int KUHN[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7}, {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};
// Bend slot pairs in lexicographic order (grid.py:257-260).
const int BEND_PAIRS[15][2] = {{0, 1}, {0, 2}, {0, 3}, {0, 4}, {0, 5}, {1, 2}, {1, 3}, {1, 4},
                               {1, 5}, {2, 3}, {2, 4}, {2, 5}, {3, 4}, {3, 5}, {4, 5}};

}  // namespace

extern "C" {

// Solver description; every field is the value the reference derives from PhysicsDesc / XpbdParams /
// XpbdGraspParams / the contact model (xpbd_solver.h:38-75, physics.cu:281-295, xpbd_solver.cu:2200-2210).
struct OracleDesc {
    int vx, vy, vz;            // voxel dims; nodes are (vx+1)(vy+1)(vz+1)
    int grid_edge_count;       // lattice edges that may be cut (four incident voxels in bounds)
    int substeps, iters;       // N small steps per frame, constraint iterations per substep
    int tet_mode;              // 0: volume row only (C = J-1); 1: coupled Neo-Hookean, C_H = J-1; 2: guarded log J
    float omega;               // SOR factor applied to lambda and positions alike
    float edge_ks;             // > 0: every live edge with compliance 1/ks; <= 0: fallback edges only
    float young;               // Young's modulus of the fallback edge law alpha~ = 1/(E h^2 L)
    float beta_s;              // edge Kelvin-Voigt damping stiffness
    float strain_damping;      // tet Kelvin damping retardation time, s
    float bulk_damping;        // guarded-mode bulk damping time, s
    float rayleigh;            // global velocity damping, 1/s
    float damping;             // legacy per-substep velocity kill
    int contact_enabled;       // friction tool contact of lattice nodes (friction_contact.cuh:143-169)
    float mu_s, mu_d;          // Coulomb coefficients
    float jaw_thickness;       // jaw plate thickness
    float contact_radius;      // node radius of the swept tests
    float jaw_gap_closed;      // closed-gap floor of the jaw boxes
    float contact_compliance;  // contact compliance, scaled by 1/h^2
    float tool_voxel;          // cut: lattice spacing used as AABB pad and rest edge length (voxelL)
    int cut_point_capacity;
    int precise_strain;        // FP32: tet invariants from F - I (the kernels' precise_strain); 0: the direct forms
    // tissue guard and material (xpbd_set_tissue_guard / xpbd_set_tissue_material, xpbd_solver.cu:1895-1915,
    // :2165-2171): with guard_passes > 0, tensile reinforcement after every tet sweep and the minimum-volume
    // safeguard at the end of every substep
    float guard_min_volume;    // minimum J of a live tet, 0 .. 0.95
    int guard_passes;          // tensile reinforcement passes per tet sweep, 0 .. 4 (0: no guard)
    float tensile_strength;    // fibre energy strength (the reference's default 80)
    // bending (PhysicsDesc kb / betaB; xpbd_begin_frame, xpbd_solver.cu:2496-2499): kb <= 0 turns it off
    float bend_kb;
    float beta_b;
};

// Tool pose (XpbdToolState, xpbd_solver.h:223-229) plus the optional blade capsule; 27 floats.
struct OracleTool {
    float shaftA[3], shaftB[3], shaftR;
    float jawHinge[3], jawAxis[3], jawMid[3], jawNormal[3];
    float jawHalfWidth, jawLen, jawGap, grasp01;
    float bladeA[3], bladeB[3], bladeR;
};

// XpbdGraspParams (xpbd_solver.h:250-289), in the field order of native_backend.py's GraspParams: the legacy contact
// model's capture/release thresholds, ring friction, core pins, break force and CCD gate.
struct OracleGrasp {
    float capture, release;       // grasp01 hysteresis
    float closed_gap;             // floor of the jaw gap
    float core_fraction;          // normalised-ellipse core fraction
    float friction;               // Coulomb mu of the ring
    float contact_compliance;     // slab and CCD gate compliance, scaled by 1/h^2
    float skin;                   // capture band widening
    float ccd_velocity, ccd_displacement, ccd_omega;
    float grasp_omega;            // under-relaxation of the slab ring and the core pins
    float core_compliance;        // core pin compliance, scaled by 1/h^2
    float break_force;            // threshold of the core force EMA
    float grasp_displacement;     // per-substep cap of the grasp corrections
    int exempt_rings;             // tear-law ring dilation (not used on this path)
};

// Swept cut plane (CutToolDesc, cut.h:30-38).
struct OracleCutTool {
    float t1v0[3], t1v1[3], t1v2[3];
    float t2v0[3], t2v1[3], t2v2[3];
    float ncut[3];
    float d;
    int valid;
    float aabbMin[3], aabbMax[3];
};

}  // extern "C"

namespace {

template <class R> struct Shape {  // ContactShape (friction_contact.cuh:3-7)
    V3<R> center, u, n, w, half, a, b;
    R radius;
    int capsule;
};
template <class R> struct Record { V3<R> normal, target, motion; int valid; };

template <class R> struct ToolR {
    V3<R> shaftA, shaftB; R shaftR;
    V3<R> jawHinge, jawAxis, jawMid, jawNormal;
    R jawHalfWidth, jawLen, jawGap, grasp01;
    V3<R> bladeA, bladeB; R bladeR;
};

template <class R> V3<R> v3(const float* p) { return mk<R>((R)p[0], (R)p[1], (R)p[2]); }

template <class R> ToolR<R> toolFrom(const OracleTool& t) {
    ToolR<R> r;
    r.shaftA = v3<R>(t.shaftA); r.shaftB = v3<R>(t.shaftB); r.shaftR = (R)t.shaftR;
    r.jawHinge = v3<R>(t.jawHinge); r.jawAxis = v3<R>(t.jawAxis); r.jawMid = v3<R>(t.jawMid);
    r.jawNormal = v3<R>(t.jawNormal);
    r.jawHalfWidth = (R)t.jawHalfWidth; r.jawLen = (R)t.jawLen; r.jawGap = (R)t.jawGap; r.grasp01 = (R)t.grasp01;
    r.bladeA = v3<R>(t.bladeA); r.bladeB = v3<R>(t.bladeB); r.bladeR = (R)t.bladeR;
    return r;
}

template <class R> inline V3<R> unitOr(V3<R> x, V3<R> fallback) {  // contactUnit (friction_contact.cuh:14-16)
    R l = len3(x);
    return l > (R)1e-8f ? x * ((R)1 / l) : fallback;
}

// contactShape (friction_contact.cuh:17-39); kind 0 shaft capsule, 1/2 jaw boxes, 3 the blade capsule.
template <class R> Shape<R> contactShape(const ToolR<R>& t, int kind, R thickness, R minGap) {
    Shape<R> s{};
    s.capsule = (kind == 0 || kind == 3);
    if (kind == 0) { s.a = t.shaftA; s.b = t.shaftB; s.radius = t.shaftR; return s; }
    if (kind == 3) { s.a = t.bladeA; s.b = t.bladeB; s.radius = t.bladeR; return s; }
    V3<R> u = unitOr(t.jawAxis, mk<R>(0, 0, 1));
    V3<R> n = t.jawNormal - u * dot3(u, t.jawNormal);
    n = unitOr(n, std::fabs(u.y) < (R)0.9f ? unitOr(cross3(u, mk<R>(0, 1, 0)), mk<R>(1, 0, 0)) : mk<R>(1, 0, 0));
    R side = kind == 1 ? (R)1 : (R)-1;
    R sn = fmin_<R>((R)0.95f, fmax_<R>((R)0, t.jawGap) / ((R)2 * fmax_<R>(t.jawLen, (R)1e-6f)));
    R cs = std::sqrt((R)1 - sn * sn);
    s.u = u * cs + n * (side * sn);
    s.n = n * cs - u * (side * sn);
    s.w = cross3(s.u, s.n);
    s.half = mk<R>(t.jawLen * (R)0.5f, thickness * (R)0.5f, t.jawHalfWidth);
    s.center = t.jawHinge + s.u * s.half.x + s.n * (side * (s.half.y + (R)0.5f * minGap));
    return s;
}

// The swept tool contact (contact/friction_contact.cuh:40-142) restated in the kernels' formulation: a surface query
// record (gap, outward normal, nearest surface point), the box's local frame, pose interpolation, the carry of tool
// material points between poses, the exact first-touch time of a translating box solved piece by piece in a
// piece-local time, and conservative advancement otherwise.
template <class R> struct Hit { R gap; V3<R> normal, foot; };

template <class R> V3<R> acrossSpine(V3<R> sp) {  // a unit vector across the spine, on its least aligned axis
    V3<R> ax = std::fabs(sp.x) <= std::fabs(sp.y) && std::fabs(sp.x) <= std::fabs(sp.z)
                   ? mk<R>(1, 0, 0)
                   : (std::fabs(sp.y) <= std::fabs(sp.z) ? mk<R>(0, 1, 0) : mk<R>(0, 0, 1));
    return unitOr(cross3(sp, ax), mk<R>(0, 0, 1));
}
template <class R> R spineParam(const Shape<R>& s, V3<R> x) {
    V3<R> sp = s.b - s.a;
    return fmin_<R>((R)1, fmax_<R>((R)0, dot3(x - s.a, sp) / fmax_<R>(dot3(sp, sp), (R)1e-20f)));
}
template <class R> Hit<R> capsuleHit(const Shape<R>& s, V3<R> x) {
    V3<R> sp = s.b - s.a;
    V3<R> base = s.a + sp * spineParam(s, x);
    V3<R> off = x - base;
    R dist = len3(off);
    Hit<R> h;
    h.normal = dist > (R)1e-6f * fmax_<R>((R)1, len3(sp)) ? off * ((R)1 / dist) : acrossSpine(sp);
    h.foot = base + h.normal * s.radius;
    h.gap = dist - s.radius;
    return h;
}
template <class R> Hit<R> boxHit(const Shape<R>& s, V3<R> x) {
    V3<R> r = x - s.center;
    R l0 = dot3(r, s.u), l1 = dot3(r, s.n), l2 = dot3(r, s.w);
    R x0 = std::fabs(l0) - s.half.x, x1 = std::fabs(l1) - s.half.y, x2 = std::fabs(l2) - s.half.z;
    Hit<R> h;
    if (x0 > 0 || x1 > 0 || x2 > 0) {
        V3<R> off = s.u * std::copysign(fmax_<R>(x0, (R)0), l0) + s.n * std::copysign(fmax_<R>(x1, (R)0), l1) +
                    s.w * std::copysign(fmax_<R>(x2, (R)0), l2);
        R dist = len3(off);
        h.normal = dist > 0 ? off * ((R)1 / dist) : s.n;
        h.foot = s.center + s.u * fmin_<R>(s.half.x, fmax_<R>(-s.half.x, l0)) +
                 s.n * fmin_<R>(s.half.y, fmax_<R>(-s.half.y, l1)) + s.w * fmin_<R>(s.half.z, fmax_<R>(-s.half.z, l2));
        h.gap = dist;
        return h;
    }
    R depth = x0, side = l0;
    V3<R> ax = s.u;
    if (x1 > depth) { depth = x1; side = l1; ax = s.n; }
    if (x2 > depth) { depth = x2; side = l2; ax = s.w; }
    h.normal = ax * (side >= 0 ? (R)1 : (R)-1);
    h.foot = x - h.normal * depth;
    h.gap = depth;
    return h;
}
template <class R> Hit<R> toolHit(const Shape<R>& s, V3<R> x) { return s.capsule ? capsuleHit(s, x) : boxHit(s, x); }

// The signed gap of x to a shape, its outward normal and nearest surface point (the surface-depth witness's query).
template <class R> R shapeDistance(const Shape<R>& s, V3<R> x, V3<R>& normal, V3<R>& surface) {
    Hit<R> h = toolHit(s, x);
    normal = h.normal;
    surface = h.foot;
    return h.gap;
}

// shapeBounds (surface_contact.cuh:229-235): the axis-aligned box of a shape, widened by radius.
inline void shapeBoundsF(const Shape<float>& s, float radius, V3<float>& lo, V3<float>& hi) {
    V3<float> c = s.capsule ? (s.a + s.b) * .5f : s.center;
    V3<float> e = s.capsule ? mk(std::fabs(s.b.x - s.a.x) * .5f + s.radius, std::fabs(s.b.y - s.a.y) * .5f + s.radius,
                                 std::fabs(s.b.z - s.a.z) * .5f + s.radius)
                            : mk(std::fabs(s.u.x) * s.half.x + std::fabs(s.n.x) * s.half.y + std::fabs(s.w.x) * s.half.z,
                                 std::fabs(s.u.y) * s.half.x + std::fabs(s.n.y) * s.half.y + std::fabs(s.w.y) * s.half.z,
                                 std::fabs(s.u.z) * s.half.x + std::fabs(s.n.z) * s.half.y + std::fabs(s.w.z) * s.half.z);
    e = e + mk(radius, radius, radius);
    lo = c - e;
    hi = c + e;
}

// triangleWitness (surface_contact.cuh:169-198): the barycentric point of a triangle deepest in a shape, by
// Frank-Wolfe steps on the simplex with an eight-reduction golden-section line search, keeping the best of the
// centroid, the corners and the iterates.
inline V3<float> triangleWitnessF(const Shape<float>& sh, V3<float> a, V3<float> b, V3<float> c) {
    V3<float> q[3] = {a, b, c}, bar = mk(1.f / 3, 1.f / 3, 1.f / 3), best = bar, n, s;
    float bestD = shapeDistance(sh, (a + b + c) * (1.f / 3), n, s);
    for (int k = 0; k < 3; k++) {
        float d = shapeDistance(sh, q[k], n, s);
        if (d < bestD) { bestD = d; best = mk((float)(k == 0), (float)(k == 1), (float)(k == 2)); }
    }
    for (int it = 0; it < 12; it++) {
        V3<float> p = a * bar.x + b * bar.y + c * bar.z;
        shapeDistance(sh, p, n, s);
        int k = dot3(n, b) < dot3(n, a) ? 1 : 0;
        if (dot3(n, c) < dot3(n, q[k])) k = 2;
        V3<float> target = mk((float)(k == 0), (float)(k == 1), (float)(k == 2)), dir = q[k] - p;
        float lo = 0, hi = 1, l = .381966f, r = .618034f;
        float dl = shapeDistance(sh, p + dir * l, n, s), dr = shapeDistance(sh, p + dir * r, n, s);
        for (int j = 0; j < 8; j++) {
            if (dl < dr) {
                hi = r; r = l; dr = dl; l = lo + (hi - lo) * .381966f;
                if (j < 7) dl = shapeDistance(sh, p + dir * l, n, s);
            } else {
                lo = l; l = r; dl = dr; r = lo + (hi - lo) * .618034f;
                if (j < 7) dr = shapeDistance(sh, p + dir * r, n, s);
            }
        }
        float t = .5f * (lo + hi);
        bar = bar * (1 - t) + target * t;
        float d = shapeDistance(sh, a * bar.x + b * bar.y + c * bar.z, n, s);
        if (d < bestD) { bestD = d; best = bar; }
    }
    return best;
}

template <class R> Shape<R> poseBetween(const Shape<R>& from, const Shape<R>& to, R t) {
    Shape<R> s = to;
    const R k = (R)1 - t;
    s.center = from.center * k + to.center * t;
    s.a = from.a * k + to.a * t;
    s.b = from.b * k + to.b * t;
    s.u = unitOr(from.u * k + to.u * t, to.u);
    V3<R> n = from.n * k + to.n * t;
    s.n = unitOr(n - s.u * dot3(s.u, n), to.n);
    s.w = cross3(s.u, s.n);
    return s;
}
// The rotation taking from's spine direction onto to's, applied to r (a half turn when they are opposite).
template <class R> V3<R> swing(const Shape<R>& from, const Shape<R>& to, V3<R> r) {
    V3<R> p = from.b - from.a, q = to.b - to.a;
    if (len3(p) < (R)1e-8f || len3(q) < (R)1e-8f) return r;
    p = unitOr(p, mk<R>(0, 0, 1));
    q = unitOr(q, p);
    R c = fmin_<R>((R)1, fmax_<R>((R)-1, dot3(p, q)));
    if (c < (R)-0.9999f) {
        V3<R> ht = unitOr(cross3(p, std::fabs(p.x) < (R)0.9f ? mk<R>(1, 0, 0) : mk<R>(0, 1, 0)), mk<R>(0, 0, 1));
        return ht * ((R)2 * dot3(ht, r)) - r;
    }
    V3<R> v = cross3(p, q);
    V3<R> vr = cross3(v, r);
    return r + vr + cross3(v, vr) * ((R)1 / ((R)1 + c));
}
// Where the tool's material point at p on pose `from` sits on pose `to`.
template <class R> V3<R> carry(const Shape<R>& from, const Shape<R>& to, V3<R> p) {
    if (from.capsule) {
        R t = spineParam(from, p);
        V3<R> sp = from.b - from.a;
        return to.a + (to.b - to.a) * t + swing(from, to, p - from.a - sp * t);
    }
    V3<R> r = p - from.center;
    return to.center + to.u * dot3(r, from.u) + to.n * dot3(r, from.n) + to.w * dot3(r, from.w);
}
template <class R> V3<R> carryDirection(const Shape<R>& at, const Shape<R>& to, V3<R> n) {
    if (to.capsule) return swing(at, to, n);
    return unitOr(to.u * dot3(n, at.u) + to.n * dot3(n, at.n) + to.w * dot3(n, at.w), n);
}

// First time in [0, 1] at which a point moving a -> b (relative to the box centre) comes within `radius` of a box
// that does not rotate: pieces between the times a local coordinate crosses a face plane, each a quadratic in the
// piece-local time s = t - lo with its value at lo as constant term, solved in double.
template <class R> bool boxTouchTime(const Shape<R>& box, R radius, V3<R> a, V3<R> b, R& time) {
    V3<R> d = b - a;
    R p[3] = {dot3(a, box.u), dot3(a, box.n), dot3(a, box.w)};
    R v[3] = {dot3(d, box.u), dot3(d, box.n), dot3(d, box.w)};
    R h[3] = {box.half.x, box.half.y, box.half.z};
    std::vector<R> cut = {(R)0, (R)1};  // 0, the crossings inside (0, 1), 1
    for (int k = 0; k < 3; k++)
        for (int side = 0; side < 2; side++) {
            R t = ((side ? h[k] : -h[k]) - p[k]) / v[k];
            if (std::fabs(v[k]) > (R)1e-12f && t > 0 && t < 1) cut.push_back(t);
        }
    std::sort(cut.begin(), cut.end());
    for (size_t i = 0; i + 1 < cut.size(); i++) {
        R lo = cut[i], hi = cut[i + 1], mid = (R)0.5f * (lo + hi);
        double A = 0, B = 0, C = -(double)radius * radius;
        for (int k = 0; k < 3; k++) {
            R atMid = p[k] + v[k] * mid;
            if (std::fabs(atMid) <= h[k]) continue;
            double e = (double)(p[k] - std::copysign(h[k], atMid)) + (double)v[k] * lo;
            A += (double)v[k] * v[k];
            B += 2.0 * e * v[k];
            C += e * e;
        }
        if (C <= 1e-14) { time = lo; return true; }
        if (A <= 1e-24) continue;
        double disc = B * B - 4.0 * A * C;
        if (disc < 0) continue;
        double sroot = (-B - std::sqrt(disc)) / (2.0 * A);
        if (sroot >= -1e-8 && sroot <= (double)(hi - lo) + 1e-8) {
            time = fmin_<R>(hi, fmax_<R>(lo, (R)((double)lo + sroot)));
            return true;
        }
    }
    return false;
}

template <class R> Record<R> recordAt(const Shape<R>& from, const Shape<R>& hit, const Shape<R>& to, V3<R> normal,
                                      V3<R> foot, R radius) {
    Record<R> c;
    V3<R> end = carry(hit, to, foot);
    c.normal = normal;
    c.target = end + normal * radius;
    c.motion = end - carry(hit, from, foot);
    c.valid = 1;
    return c;
}

// The node's swept contact with a moving shape: the exact first touch for a translating box, conservative
// advancement (steps of 0.8 gap / speed, time in double) otherwise.
template <class R> Record<R> traceToolContact(const Shape<R>& from, const Shape<R>& to, R radius, V3<R> x0, V3<R> x1) {
    Record<R> none{};
    if (!to.capsule && len3(to.u - from.u) + len3(to.n - from.n) < (R)1e-7f) {
        R time = 0;
        if (!boxTouchTime(to, radius, x0 - from.center, x1 - to.center, time)) return none;
        Shape<R> hit = poseBetween(from, to, time);
        Hit<R> q = toolHit(hit, x0 * ((R)1 - time) + x1 * time);
        return recordAt(from, hit, to, q.normal, q.foot, radius);
    }
    R speed = len3(x1 - x0);
    if (to.capsule) speed += fmax_<R>(len3(to.a - from.a), len3(to.b - from.b));
    else speed += len3(to.center - from.center) + (R)4.f * len3(to.half) * (len3(to.u - from.u) + len3(to.n - from.n));
    V3<R> mid0 = to.capsule ? (from.a + from.b) * (R)0.5f : from.center;
    V3<R> mid1 = to.capsule ? (to.a + to.b) * (R)0.5f : to.center;
    R reach = to.capsule ? fmax_<R>(len3(from.b - from.a), len3(to.b - to.a)) * (R)0.5f + to.radius : len3(to.half);
    if (len3(x0 - mid0) > reach + radius + len3(x1 - x0) + len3(mid1 - mid0) + (R)1e-5f) return none;
    R tol = fmax_<R>((R)1e-6f, (R)1e-4f * (reach + radius));
    double t = 0;
    while (t < 1.0) {
        Shape<R> at = poseBetween(from, to, (R)t);
        Hit<R> q = toolHit(at, x0 * ((R)1 - (R)t) + x1 * (R)t);
        R gap = q.gap - radius;
        if (gap <= tol) return recordAt(from, at, to, carryDirection(at, to, q.normal), q.foot, radius);
        if (speed <= (R)1e-12f) return none;
        R step = (R)0.8f * gap / speed;
        if ((R)t + step >= (R)1) {
            Hit<R> e = toolHit(to, x1);
            if (e.gap - radius > tol) return none;
            return recordAt(from, to, to, carryDirection(to, to, e.normal), e.foot, radius);
        }
        t += step;
    }
    return none;
}

// One pass of k_SolveFrictionContacts (friction_contact.cuh:150-168) for one contact of a node: the unilateral normal
// multiplier (compliance alpha, clamped at 0), then the tangential multiplier against the slip relative to the tool's
// motion, kept whole below mu_s lambda_N (stick) and scaled to mu_d lambda_N above it (slide).
template <class R>
void frictionUpdate(const Record<R>& c, R w, R alpha, R muS, R muD, V3<R> prev, V3<R>& x, R& lamN, V3<R>& lamT) {
    R C = dot3(x - c.target, c.normal), old = lamN;
    R next = fmax_<R>((R)0, old + (-C - alpha * old) / (w + alpha));
    x = x + c.normal * (w * (next - old));
    lamN = next;
    V3<R> slip = x - prev - c.motion;
    slip = slip - c.normal * dot3(slip, c.normal);
    V3<R> oldT = lamT;
    V3<R> trial = oldT - slip * ((R)1 / w);
    R len = len3(trial);
    V3<R> nextT = trial;
    if (len > muS * next) nextT = len > (R)1e-15f ? trial * (muD * next / len) : mk<R>(0, 0, 0);
    x = x + (nextT - oldT) * w;
    lamT = nextT;
}

// Möller–Trumbore of the deformed edge against one swept triangle (cut.cuh:162-185).
template <class R> bool mollerTrumbore(V3<R> O, V3<R> D, V3<R> V0, V3<R> V1, V3<R> V2, R& t, R& u, R& v) {
    V3<R> L1 = V1 - V0, L2 = V2 - V0, T = V0 - O;
    R det = dot3(cross3(L2, D), L1);
    t = 0; u = 0; v = 0;
    if (std::fabs(det) < (R)1e-12f) return false;
    R inv = (R)1 / det;
    t = dot3(cross3(L2, T), L1) * inv;
    u = -dot3(cross3(L2, D), T) * inv;
    v = dot3(cross3(D, T), L1) * inv;
    const R EPS = (R)1e-5f;
    return u >= -EPS && v >= -EPS && u + v <= (R)1 + EPS && t > 0 && t <= (R)1;
}

struct CutPoint {  // CutPointGpu (cut.cuh:37-46): 32 bytes, rotation taken as identity
    float off[3];
    int owner;
    int edge;
    float nrm[3];
};

template <class R> struct Tissue {
    OracleDesc d{};
    int N = 0, V = 0, E = 0, T = 0, nx = 0, ny = 0, nz = 0;
    // node state (AoS, as the reference's float3 buffers)
    std::vector<V3<R>> pos, prev, vel, ext, rest;
    // particle frames (RotGpu rows r0, r1, r2 per node) and their previous values (physics.cu:76-77)
    std::vector<V3<R>> rot, rotPrev;
    std::vector<R> mass, invMass;
    std::vector<int> active, pinned, inside, nbr, nbr0;
    // bend pairs (grid.py BEND_DTYPE / BendPairGpuC): only (j, k, alive) are read on this path
    std::vector<int> bendJ, bendK, bendAlive;
    // bend elements: the pairs alive at build, colour-ordered (xpbd_bind :1466-1478, xpbd_set_bending :1344-1378);
    // bsrc is the pair's index 15 i + p, whose alive flag gates the element (bendActive, :446)
    std::vector<int> bi, bj, bk, bsrc, bcolorOff, bcolorCnt;
    std::vector<R> btheta, blambda;
    R alphaB = 0, betaGammaHB = 0;
    bool doBend = false;
    // edges, colour-ordered (xpbd_solver.cu:1230-1249)
    std::vector<int> ei, ej, eowner, eactive, ecolorOff, ecolorCnt, egenIndex;
    std::vector<R> erest, elambda;
    std::vector<int> support;  // [6N] tet support of each owner slot (xpbd_solver.cu:469-479)
    // tets, generation order (cell z,y,x major, Kuhn type minor) plus the colour-major list (:1740-1768)
    std::vector<int> tids, ttype, towner, tactive, tcolor, tcolorOff, tcolorCnt, tcolorOrder;
    std::vector<R> kDev, kHyd, alphaDev, alphaHyd, lamDev, lamHyd;
    R typeB[6][3][3];
    V3<R> typeRk[6][3];
    double typeV[6];
    // frame constants (xpbd_begin_frame)
    R h = 0, alphaTilde = 0, velScale = 1, betaGammaH = 0;
    bool frameReady = false;
    int substepIndex = 0;
    // tool
    ToolR<R> tool{}, toolFrom{};
    bool toolArmed = false, bladeArmed = false;
    R reaction[3] = {0, 0, 0};
    int reactionContacts = 0;
    // cut (cut.cu:1105-1206)
    std::vector<int> gridEdges;           // [4*Eg] axis, x, y, z
    std::vector<uint32_t> occupied, cutMask;
    std::vector<float> edgeCutT, edgeSkinT;
    std::vector<CutPoint> cutPoints;
    uint32_t cutPointCounter = 0, revision = 0, refreshedRevision = 0xffffffffu;
    uint32_t tearCount = 0;  // edges the tear law marked (the reference's dbg[15])
    // the cut tool of LCS_SetTool and the CCD ribbon tick it arms (cut.cu:1208-1226, :1303-1431)
    bool cutToolSet = false, havePrevCorner = false;
    OracleCutTool cutTool{};
    V3<R> emitNcut = V3<R>{0, 1, 0};
    float tearTau = 0.f;
    std::vector<V3<R>> prevCorner;
    // legacy grasp (contact model 0, xpbd_solver.cu:250-275): the host latch and the per-node state
    int contactModel = 1;         // 0: legacy grasp and CCD gate; 1: friction contact
    OracleGrasp gp{};
    bool grasping = false;
    std::vector<int> grasped, core;
    std::vector<R> offU, offN, offW, anchU, anchW, slab, slabPrev, ema, jawLam, shaftLam;
    std::vector<V3<R>> coreLam;
    // tissue guard (xpbd_solver.cu:588-741): fibre multipliers, the substep's start and candidate positions, the
    // nodes that need the feasible filter and each node's incident tets
    std::vector<R> fiberLambda;
    std::vector<V3<R>> feasStart, feasTarget;
    std::vector<int> needed, incOff, incTets;
    // the tear law's grasp ring and release decay (cut.cu:68-83): per node, per grid edge
    std::vector<int> graspRing, edgeWasExempt, edgeDecayTicks;
    int status = 0;

    int cid(int x, int y, int z) const { return x + nx * (y + ny * z); }
};

// Greedy first-free colouring on a 64-bit per-vertex mask (xpbd_solver.cu:1213-1230, :1740-1751).
template <int K>
int greedyColour(int count, int vertices, const std::vector<int>& ids, std::vector<int>& colour, int& numColours) {
    std::vector<unsigned long long> vmask(vertices, 0ull);
    colour.assign(count, 0);
    numColours = 0;
    for (int e = 0; e < count; e++) {
        unsigned long long used = 0;
        for (int k = 0; k < K; k++) used |= vmask[ids[K * e + k]];
        int c = 0;
        while (c < 64 && (used & (1ull << c))) c++;
        if (c >= 64) return -1003;  // XPBD_ERR_COLOR_OVERFLOW
        colour[e] = c;
        for (int k = 0; k < K; k++) vmask[ids[K * e + k]] |= (1ull << c);
        if (c + 1 > numColours) numColours = c + 1;
    }
    return 0;
}

template <class R> bool inv3(const double m[3][3], double out[3][3]) {  // host double inverse (xpbd_solver.cu:1580-1598)
    double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    if (std::fabs(det) < 1e-30) return false;
    double id = 1.0 / det;
    out[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
    out[0][1] = -(m[0][1] * m[2][2] - m[0][2] * m[2][1]) * id;
    out[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    out[1][0] = -(m[1][0] * m[2][2] - m[1][2] * m[2][0]) * id;
    out[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    out[1][2] = -(m[0][0] * m[1][2] - m[0][2] * m[1][0]) * id;
    out[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
    out[2][1] = -(m[0][0] * m[2][1] - m[0][1] * m[2][0]) * id;
    out[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
    return true;
}

template <class R> int buildTopology(Tissue<R>& s) {
    // Edges: one per positive slot, p ascending, axis 0..2; L0 = |restNbr| (xpbd_solver.cu:1445-1463).
    std::vector<int> ids, owner;
    std::vector<R> rl;
    for (int p = 0; p < s.N; p++)
        for (int axis = 0; axis < 3; axis++) {
            int slot = 2 * axis + 1, n = s.nbr[6 * p + slot];
            if (n < 0) continue;
            ids.push_back(p);
            ids.push_back(n);
            owner.push_back(6 * p + slot);
        }
    s.E = (int)owner.size();
    std::vector<int> colour;
    int nc = 0;
    if (greedyColour<2>(s.E, s.N, ids, colour, nc)) return -1003;
    s.ecolorCnt.assign(nc, 0);
    for (int e = 0; e < s.E; e++) s.ecolorCnt[colour[e]]++;
    s.ecolorOff.assign(nc, 0);
    for (int c = 1; c < nc; c++) s.ecolorOff[c] = s.ecolorOff[c - 1] + s.ecolorCnt[c - 1];
    s.ei.resize(s.E); s.ej.resize(s.E); s.eowner.resize(s.E); s.erest.resize(s.E); s.egenIndex.resize(s.E);
    std::vector<int> cursor = s.ecolorOff;
    for (int e = 0; e < s.E; e++) {
        int dst = cursor[colour[e]]++;
        s.ei[dst] = ids[2 * e];
        s.ej[dst] = ids[2 * e + 1];
        s.eowner[dst] = owner[e];
        s.egenIndex[dst] = e;
    }
    s.eactive.assign(s.E, 1);
    s.elambda.assign(s.E, 0);
    s.support.assign(6 * s.N, 0);

    // Kuhn tets of all-active cells; Dm^-1 and V per type from the rest lattice spacing in double
    // (xpbd_solver.cu:1646-1660), each rest tet checked against its type (:1701, :1785-1792).
    double spacing[3] = {((double)s.rest[s.cid(s.d.vx, 0, 0)].x - (double)s.rest[0].x) / s.d.vx,
                         ((double)s.rest[s.cid(0, s.d.vy, 0)].y - (double)s.rest[0].y) / s.d.vy,
                         ((double)s.rest[s.cid(0, 0, s.d.vz)].z - (double)s.rest[0].z) / s.d.vz};
    double tB[6][3][3];
    for (int ti = 0; ti < 6; ti++) {
        double Dm[3][3];
        for (int k = 1; k < 4; k++)
            for (int a = 0; a < 3; a++) Dm[a][k - 1] = ((KUHN[ti][k] >> a) & 1) * spacing[a];
        if (!inv3<R>(Dm, tB[ti])) return -1005;
        double det = Dm[0][0] * (Dm[1][1] * Dm[2][2] - Dm[1][2] * Dm[2][1]) -
                     Dm[0][1] * (Dm[1][0] * Dm[2][2] - Dm[1][2] * Dm[2][0]) +
                     Dm[0][2] * (Dm[1][0] * Dm[2][1] - Dm[1][1] * Dm[2][0]);
        s.typeV[ti] = std::fabs(det) / 6.0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) s.typeB[ti][a][b] = (R)(float)tB[ti][a][b];
    }
    s.tids.clear(); s.ttype.clear(); s.towner.clear();
    for (int z = 0; z < s.d.vz; z++)
        for (int y = 0; y < s.d.vy; y++)
            for (int x = 0; x < s.d.vx; x++) {
                bool all = true;
                for (int li = 0; li < 8 && all; li++)
                    if (!s.active[s.cid(x + (li & 1), y + ((li >> 1) & 1), z + ((li >> 2) & 1))]) all = false;
                if (!all) continue;
                for (int ti = 0; ti < 6; ti++) {
                    int g[4];
                    for (int k = 0; k < 4; k++) {
                        int li = KUHN[ti][k];
                        g[k] = s.cid(x + (li & 1), y + ((li >> 1) & 1), z + ((li >> 2) & 1));
                    }
                    double e[3][3];
                    for (int k = 0; k < 3; k++) {
                        e[k][0] = (double)s.rest[g[k + 1]].x - (double)s.rest[g[0]].x;
                        e[k][1] = (double)s.rest[g[k + 1]].y - (double)s.rest[g[0]].y;
                        e[k][2] = (double)s.rest[g[k + 1]].z - (double)s.rest[g[0]].z;
                    }
                    double Dm[3][3] = {{e[0][0], e[1][0], e[2][0]}, {e[0][1], e[1][1], e[2][1]}, {e[0][2], e[1][2], e[2][2]}};
                    double B[3][3];
                    if (!inv3<R>(Dm, B)) continue;
                    double scale = 0, dev = 0;
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) {
                            scale = std::fmax(scale, std::fabs(tB[ti][a][b]));
                            dev = std::fmax(dev, std::fabs(B[a][b] - tB[ti][a][b]));
                        }
                    if (dev > 1e-4 * scale) return -1005;
                    int axes = 0;
                    int own[3];
                    for (int a = 0; a < 4 && axes < 3; a++)
                        for (int b = a + 1; b < 4 && axes < 3; b++) {
                            int lu = KUHN[ti][a], lv = KUHN[ti][b], diff = lu ^ lv;
                            if (diff != 1 && diff != 2 && diff != 4) continue;
                            int axis = diff == 1 ? 0 : (diff == 2 ? 1 : 2);
                            int lower = ((lu >> axis) & 1) == 0 ? lu : lv;
                            int lg = s.cid(x + (lower & 1), y + ((lower >> 1) & 1), z + ((lower >> 2) & 1));
                            own[axes++] = 6 * lg + 2 * axis + 1;
                        }
                    for (int k = 0; k < 4; k++) s.tids.push_back(g[k]);
                    s.ttype.push_back(ti);
                    for (int k = 0; k < 3; k++) s.towner.push_back(own[k]);
                }
            }
    s.T = (int)s.ttype.size();
    int tnc = 0;
    if (greedyColour<4>(s.T, s.N, s.tids, s.tcolor, tnc)) return -1003;
    s.tcolorCnt.assign(tnc, 0);
    for (int e = 0; e < s.T; e++) s.tcolorCnt[s.tcolor[e]]++;
    s.tcolorOff.assign(tnc, 0);
    for (int c = 1; c < tnc; c++) s.tcolorOff[c] = s.tcolorOff[c - 1] + s.tcolorCnt[c - 1];
    s.tcolorOrder.assign(s.T, 0);
    std::vector<int> tcur = s.tcolorOff;
    for (int e = 0; e < s.T; e++) s.tcolorOrder[tcur[s.tcolor[e]]++] = e;
    s.tactive.assign(s.T, 1);
    s.kDev.assign(s.T, (R)1); s.kHyd.assign(s.T, (R)1);
    s.alphaDev.assign(s.T, 0); s.alphaHyd.assign(s.T, 0);
    s.lamDev.assign(s.T, 0); s.lamHyd.assign(s.T, 0);
    // the guard's buffers and incidence lists (xpbd_set_tissue_guard, xpbd_solver.cu:1900-1912): tets in index order
    s.fiberLambda.assign(6 * (size_t)s.T, 0);
    s.feasStart.assign(s.N, V3<R>{0, 0, 0});
    s.feasTarget.assign(s.N, V3<R>{0, 0, 0});
    s.needed.assign(s.N, 0);
    s.incOff.assign(s.N + 1, 0);
    for (int i : s.tids) s.incOff[i + 1]++;
    for (int i = 1; i <= s.N; i++) s.incOff[i] += s.incOff[i - 1];
    s.incTets.assign(4 * (size_t)s.T, 0);
    std::vector<int> slot(s.incOff.begin(), s.incOff.end() - 1);
    for (int t = 0; t < s.T; t++)
        for (int k = 0; k < 4; k++) s.incTets[slot[s.tids[4 * t + k]]++] = t;
    // Bend elements: every pair alive at bind time, apex-major, with rest angle pi for opposite slots and pi/2
    // otherwise (grid.py:262-268), coloured over their three nodes (xpbd_set_bending, :1344-1378).
    std::vector<int> bids, bsrc;
    std::vector<R> bth;
    for (int i = 0; i < s.N; i++)
        for (int p = 0; p < 15; p++) {
            if (!s.bendAlive[15 * i + p]) continue;
            bids.push_back(i);
            bids.push_back(s.bendJ[15 * i + p]);
            bids.push_back(s.bendK[15 * i + p]);
            bsrc.push_back(15 * i + p);
            bool opposite = BEND_PAIRS[p][0] / 2 == BEND_PAIRS[p][1] / 2;
            bth.push_back((R)(opposite ? (float)3.14159265358979323846 : (float)(3.14159265358979323846 / 2)));
        }
    const int B = (int)bsrc.size();
    std::vector<int> bcol;
    int bnc = 0;
    if (greedyColour<3>(B, s.N, bids, bcol, bnc)) return -1003;
    s.bcolorCnt.assign(bnc, 0);
    for (int e = 0; e < B; e++) s.bcolorCnt[bcol[e]]++;
    s.bcolorOff.assign(bnc, 0);
    for (int c = 1; c < bnc; c++) s.bcolorOff[c] = s.bcolorOff[c - 1] + s.bcolorCnt[c - 1];
    s.bi.resize(B); s.bj.resize(B); s.bk.resize(B); s.bsrc.resize(B); s.btheta.resize(B);
    std::vector<int> bcur = s.bcolorOff;
    for (int e = 0; e < B; e++) {
        int dst = bcur[bcol[e]]++;
        s.bi[dst] = bids[3 * e]; s.bj[dst] = bids[3 * e + 1]; s.bk[dst] = bids[3 * e + 2];
        s.bsrc[dst] = bsrc[e];
        s.btheta[dst] = bth[e];
    }
    s.blambda.assign(B, 0);
    return 0;
}

// k_RefreshLiveness (xpbd_solver.cu:429-466) + fallback support (:469-479); run when the topology revision moved.
template <class R> void refreshLiveness(Tissue<R>& s) {
    if (s.refreshedRevision == s.revision) return;
    s.refreshedRevision = s.revision;
    for (int e = 0; e < s.E; e++)
        s.eactive[e] = (s.nbr[s.eowner[e]] >= 0 && s.active[s.ei[e]] != 0 && s.active[s.ej[e]] != 0) ? 1 : 0;
    for (int t = 0; t < s.T; t++) {
        bool live = s.nbr[s.towner[3 * t]] >= 0 && s.nbr[s.towner[3 * t + 1]] >= 0 && s.nbr[s.towner[3 * t + 2]] >= 0;
        for (int k = 0; k < 4; k++) live = live && s.active[s.tids[4 * t + k]] != 0;
        s.tactive[t] = live ? 1 : 0;
    }
    std::fill(s.support.begin(), s.support.end(), 0);
    for (int t = 0; t < s.T; t++)
        if (s.tactive[t])
            for (int k = 0; k < 3; k++) s.support[s.towner[3 * t + k]] = 1;
}

// xpbd_begin_frame (xpbd_solver.cu:2481-2514) + k_ComputeTetAlpha (:578-587).
template <class R> int beginFrame(Tissue<R>& s, R dt) {
    int N = s.d.substeps > 0 ? s.d.substeps : 1;
    R h = dt / (R)N;
    if (!(h > 0) || !std::isfinite(h)) return -1002;
    R ks = (R)s.d.edge_ks;
    R at = ks > 0 ? (R)1 / (ks * h * h) : (R)0;
    s.velScale = ((R)1 - (R)s.d.damping) / ((R)1 + (R)s.d.rayleigh * h);
    s.betaGammaH = at * (R)s.d.beta_s * h;
    s.alphaTilde = at;
    s.h = h;
    s.doBend = !s.bi.empty() && s.d.bend_kb > 0;  // g_doBendF (:2496-2499)
    s.alphaB = s.doBend ? (R)1 / ((R)s.d.bend_kb * h * h) : (R)0;
    s.betaGammaHB = s.alphaB * (R)s.d.beta_b * h;
    R invH2 = (R)1 / (h * h);
    for (int t = 0; t < s.T; t++) {
        s.alphaDev[t] = invH2 / s.kDev[t];
        s.alphaHyd[t] = invH2 / s.kHyd[t];
    }
    s.substepIndex = 0;
    s.frameReady = true;
    return 0;
}

// k_solve_edges (xpbd_solver.cu:484-524) as a pure function of one edge: positions xi, xj (updated), their substep
// starts pi, pj, inverse masses, rest length, compliance alphaTilde, Kelvin-Voigt factor gamma, relaxation omega and
// the multiplier lam (updated). Returns false where the reference returns without writing.
template <class R>
bool edgeProject(V3<R>& xi, V3<R>& xj, const V3<R>& pi, const V3<R>& pj, R wi, R wj, R rest, R alphaTilde, R gamma,
                 R omega, R& lam) {
    R wsum = wi + wj;
    if (wsum <= 0) return false;
    V3<R> dd = xi - xj;
    R L = len3(dd);
    if (L < (R)1e-9f) return false;
    V3<R> n = dd * ((R)1 / L);
    R C = L - rest;
    R gradDx = dot3(n, (xi - pi) - (xj - pj));
    R dl = (-C - alphaTilde * lam - gamma * gradDx) / (((R)1 + gamma) * wsum + alphaTilde);
    R applied = omega * dl;
    lam = lam + applied;
    xi = xi + n * (wi * applied);
    xj = xj - n * (wj * applied);
    return true;
}

// k_solve_edges on edge e of the tissue; under the fallback (no edge stiffness, :469-479) an edge whose owner cell
// keeps a live tet is skipped and the others take alphaTilde = 1 / (E L h^2).
template <class R> void solveEdge(Tissue<R>& s, int e, R alphaTilde, bool fallback, R fallbackInvEh2, bool useLambda) {
    // the substep's reset (xpbd_solver.cu:3534) leaves every row that does not project at 0
    if (s.eactive[e] == 0 || (fallback && s.support[s.eowner[e]])) {
        if (!useLambda) s.elambda[e] = 0;
        return;
    }
    if (fallback) alphaTilde = fallbackInvEh2 / s.erest[e];
    int i = s.ei[e], j = s.ej[e];
    R lam = useLambda ? s.elambda[e] : (R)0;
    if (edgeProject(s.pos[i], s.pos[j], s.prev[i], s.prev[j], s.invMass[i], s.invMass[j], s.erest[e], alphaTilde,
                    s.betaGammaH, (R)s.d.omega, lam))
        s.elambda[e] = lam;
    else if (!useLambda)
        s.elambda[e] = 0;  // the substep's reset (xpbd_solver.cu:3534): a row that writes nothing keeps lambda = 0
}

// arccos of the bend angle. The reference calls acosf (xpbd_solver.cu:553); the float oracle and the kernel share one
// fixed sequence: the Maclaurin series of arcsin to x^23 on |x| <= 1/2, with acos c = pi/2 - asin c there and
// acos c = 2 asin sqrt((1 - c)/2), pi - 2 asin sqrt((1 + c)/2) beyond, pi and pi/2 as a float head plus the float of
// the residual. The double oracle uses the library arccos.
inline float bendAcos(float c) {
    const float a = std::fabs(c);
    const bool mid = a <= 0.5f;
    const float x = mid ? c : std::sqrt((1.f - a) * 0.5f);
    const float x2 = mid ? c * c : (1.f - a) * 0.5f;
    // (2n)! / (4^n (n!)^2 (2n + 1)), n = 11 down to 1
    const float coef[11] = {0.007312526f, 0.008390335f, 0.009761609f, 0.011551801f, 0.013964844f, 0.017352764f,
                            0.022372158f, 0.030381944f, 0.044642858f, 0.075f, 0.16666667f};
    float p = coef[0];
    for (int n = 1; n < 11; n++) p = coef[n] + x2 * p;
    const float tail = x * x2 * p;
    if (mid) return 1.57079637f - (x + (tail - -4.37113883e-08f));
    const float twice = 2.f * (x + tail);
    return c > 0.f ? twice : 3.14159274f - (twice - -8.74227766e-08f);
}
inline double bendAcos(double c) { return std::acos(c); }

// k_solve_bending (xpbd_solver.cu:534-573) as a pure function of one element: apex xi, arms xj, xk (updated), their
// substep starts, inverse masses, rest angle theta0, compliance alphaB, Kelvin-Voigt factor gamma, relaxation omega
// and the multiplier lam (updated). C = theta - theta0; the corner directions are the reference's g_j = n_v - c n_u,
// g_k = n_u - c n_v, g_i = -(g_j + g_k). Returns false where the reference returns without writing.
template <class R>
bool bendProject(V3<R>& xi, V3<R>& xj, V3<R>& xk, const V3<R>& pi, const V3<R>& pj, const V3<R>& pk, R wi, R wj, R wk,
                 R theta0, R alphaB, R gamma, R omega, R& lam) {
    if (wi + wj + wk <= 0) return false;
    V3<R> u = xi - xj;
    R Lu = len3(u);
    V3<R> v = xi - xk;
    R Lv = len3(v);
    if (Lu < (R)1e-9f || Lv < (R)1e-9f) return false;
    V3<R> nu = u * ((R)1 / Lu), nv = v * ((R)1 / Lv);
    R c = fmax_<R>((R)-1, fmin_<R>((R)1, dot3(nu, nv)));
    R sn = std::sqrt(fmax_<R>((R)1 - c * c, (R)0));
    if (sn < (R)1e-4f) return false;
    R C = bendAcos(c) - theta0;
    V3<R> gj = nv - nu * c;
    V3<R> gk = nu - nv * c;
    V3<R> gi = (gj + gk) * (R)-1;
    R wg = wi * dot3(gi, gi) + wj * dot3(gj, gj) + wk * dot3(gk, gk);
    R gradDx = dot3(gi, xi - pi) + dot3(gj, xj - pj) + dot3(gk, xk - pk);
    R denom = ((R)1 + gamma) * wg + alphaB;
    if (denom < (R)1e-12f) return false;
    R dl = (-C - alphaB * lam - gamma * gradDx) / denom;
    R applied = omega * dl;
    lam = lam + applied;
    xi = xi + gi * (wi * applied);
    xj = xj + gj * (wj * applied);
    xk = xk + gk * (wk * applied);
    return true;
}

// k_solve_bending on element e of the tissue, gated by its pair's alive flag.
template <class R> void solveBend(Tissue<R>& s, int e, bool useLambda) {
    if (!s.bendAlive[s.bsrc[e]]) {
        if (!useLambda) s.blambda[e] = 0;  // the substep's reset (:3535)
        return;
    }
    int i = s.bi[e], j = s.bj[e], k = s.bk[e];
    R lam = useLambda ? s.blambda[e] : (R)0;
    if (bendProject(s.pos[i], s.pos[j], s.pos[k], s.prev[i], s.prev[j], s.prev[k], s.invMass[i], s.invMass[j],
                    s.invMass[k], s.btheta[e], s.alphaB, s.betaGammaHB, (R)s.d.omega, lam))
        s.blambda[e] = lam;
    else if (!useLambda)
        s.blambda[e] = 0;  // the substep's reset (xpbd_solver.cu:3535)
}

// k_perp_damp_edges (xpbd_solver.cu:890-911) on one edge: the part of the endpoints' relative velocity across the
// edge, times hBeta = h beta_s, is taken off i and given to j in proportion to their inverse masses.
template <class R>
void perpDamp(V3<R>& vi, V3<R>& vj, const V3<R>& xi, const V3<R>& xj, R wi, R wj, R hBeta) {
    if (wi + wj <= 0) return;
    V3<R> dd = xi - xj;
    R L = len3(dd);
    if (L < (R)1e-9f) return;
    V3<R> n = dd * ((R)1 / L);
    V3<R> vrel = vi - vj;
    V3<R> vp = vrel - n * dot3(vrel, n);
    V3<R> imp = vp * hBeta;
    if (wi > 0) vi = vi - imp * wi;
    if (wj > 0) vj = vj + imp * wj;
}

// ---- precise tet invariants (no reference counterpart; OracleDesc.precise_strain). The FP32 path restates the
// kernels' sequence (gts_common.h: iso_minus_one, own_log1p, precise_strain, precise_dev) op for op; the FP64 oracle
// keeps the reference's direct forms, which FP64 resolves, and is the accuracy reference for both.
// J - 1, I1 - 3, J^(-2/3) I1 - 3 and the deviatoric gradient Ghat = J^(-2/3) F - (J^(-2/3) I1 / 3J) cof F are formed
// from H = F - I = (e_k - r_k) Dm^-1 so that their O(1) parts never cancel (see the kernels' header for the identities).
inline float isoMinusOne(float a, float iso, float& u2, float& u3) {  // u = (1 + a)^(-2/3) - 1, u2, u3 its tails
    if (std::fabs(a) <= 0.0625f) {  // binomial series to a^8
        float s = 0.36418906f;
        s = -0.38002336f + a * s;
        s = 0.39902455f + a * s;
        s = -0.42249656f + a * s;
        s = 0.4526749f + a * s;
        s = -0.49382716f + a * s;
        u3 = a * a * a * s;
        u2 = 0.5555556f * a * a + u3;
        return -0.6666667f * a + u2;
    }
    float r = std::sqrt(iso);  // (1 + a)^(-1/3); r^2 - 1 = -a r^3 (1 + r) / (1 + r + r^2)
    float u = -a * iso * r * (1.f + r) / (1.f + r + iso);
    u2 = u + 0.6666667f * a;
    u3 = u2 - 0.5555556f * a * a;
    return u;
}
inline float log1pOwn(float a) {  // 2 atanh(a / (2 + a)) near 0, else the guarded log of 1 + a
    if (std::fabs(a) < 0.25f) {
        float s = a / (2.f + a);
        float s2 = s * s;
        float p = s2 * (0.333333343f + s2 * (0.2f + s2 * (0.142857149f + s2 * (0.111111112f + s2 * 0.0909090936f))));
        return 2.f * s + 2.f * s * p;
    }
    return guardLog(1.f + a);
}
struct Strain32 {
    V3<float> H[3], cH[3];
    float trH, e, det, nH2, a, b;  // tr H, tr cof H + det H, det H, |H|^2, J - 1, I1 - 3
};
inline Strain32 preciseStrain(const V3<float> d[3], const V3<float> br[3]) {
    Strain32 s;
    for (int c = 0; c < 3; c++) s.H[c] = d[0] * comp(br[0], c) + d[1] * comp(br[1], c) + d[2] * comp(br[2], c);
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
// F = I + H, cof F = (1 + tr H) I - H^T + cof H, J = 1 + (J - 1): the precise path's F, its cofactors and determinant.
inline void preciseFrame(const Strain32& s, V3<float> F[3], V3<float> cof[3], float& J) {
    const V3<float>* H = s.H;
    const float c1 = 1.f + s.trH;
    F[0] = mk(H[0].x + 1.f, H[0].y, H[0].z);
    F[1] = mk(H[1].x, H[1].y + 1.f, H[1].z);
    F[2] = mk(H[2].x, H[2].y, H[2].z + 1.f);
    cof[0] = mk(s.cH[0].x + c1 - H[0].x, s.cH[0].y - H[1].x, s.cH[0].z - H[2].x);
    cof[1] = mk(s.cH[1].x - H[0].y, s.cH[1].y + c1 - H[1].y, s.cH[1].z - H[2].y);
    cof[2] = mk(s.cH[2].x - H[0].z, s.cH[2].y - H[1].z, s.cH[2].z + c1 - H[2].z);
    J = 1.f + s.a;
}
// q = J^(-2/3) I1 - 3 (unclamped) and the columns of Ghat, at pJ = max(J, .05) and iso = pJ^(-2/3):
// q = 2 |dev sym H|^2 + 2 tr H e + 5e^2/3 - 2 det H + 3 u3 - 2a|H|^2/3 + u2 b above the J clamp.
inline float preciseDev(const Strain32& s, const V3<float> F[3], const V3<float> cof[3], bool jlive, float pJ,
                        float iso, V3<float> Gh[3]) {
    if (!jlive) {
        float u = iso - 1.f;
        for (int c = 0; c < 3; c++) Gh[c] = F[c] * iso;
        return 3.f * u + s.b + u * s.b;
    }
    float u2, u3;
    isoMinusOne(s.a, iso, u2, u3);
    const float t3 = s.trH * 0.333333343f;  // FP32 1/3
    float d0 = s.H[0].x - t3, d1 = s.H[1].y - t3, d2 = s.H[2].z - t3;
    float s01 = s.H[1].x + s.H[0].y, s02 = s.H[2].x + s.H[0].z, s12 = s.H[2].y + s.H[1].z;
    float q2 = 2.f * (d0 * d0 + d1 * d1 + d2 * d2) + (s01 * s01 + s02 * s02 + s12 * s12);
    float rest = 2.f * s.trH * s.e + 1.6666666f * s.e * s.e - 2.f * s.det + 3.f * u3 - 0.6666667f * s.a * s.nH2 +
                 u2 * s.b;
    float m = (s.b - 3.f * s.a) / (3.f * pJ);
    // column c of H + H^T - tr H I - cof H - m cof F, tr H taken off the diagonal entry only
    const V3<float>* H = s.H;
    V3<float> S0 = mk(H[0].x + H[0].x - s.trH, H[0].y + H[1].x, H[0].z + H[2].x);
    V3<float> S1 = mk(H[1].x + H[0].y, H[1].y + H[1].y - s.trH, H[1].z + H[2].y);
    V3<float> S2 = mk(H[2].x + H[0].z, H[2].y + H[1].z, H[2].z + H[2].z - s.trH);
    Gh[0] = (S0 - s.cH[0] - cof[0] * m) * iso;
    Gh[1] = (S1 - s.cH[1] - cof[1] * m) * iso;
    Gh[2] = (S2 - s.cH[2] - cof[2] * m) * iso;
    return q2 + rest;
}

// k_solve_tet (xpbd_solver.cu:749-829) on one tet. Mode 0 keeps the hydrostatic row alone (the deviatoric row and
// its multiplier are absent, which is the 2x2 block with gd = 0); modes 1/2 are the coupled block, legacy / guarded.
// k_solve_tet (xpbd_solver.cu:749-829) on one tet as a pure function of its inputs. Mode 0 keeps the hydrostatic row
// alone (the deviatoric row and its multiplier are absent: the 2x2 block with gd = 0); modes 1/2 are the coupled block,
// legacy / guarded. Returns false where the reference returns without writing. With the rest edges rk of the tet's
// type (FP32 only), the invariants take the precise forms above; without, the reference's direct forms.
template <class R>
bool tetProject(int mode, const V3<R> p[4], const V3<R> pv[4], const R w[4], const V3<R> br[3], R ad, R ah,
                R& lamD, R& lamH, R dampingOverH, R bulkDampingOverH, R omega, V3<R> out[4], const V3<R>* rk = nullptr) {
    V3<R> e1 = p[1] - p[0], e2 = p[2] - p[0], e3 = p[3] - p[0];
    V3<R> F[3], cof[3];
    R J;
    // precise invariants: FP32 with the type's rest edges rk (the FP64 oracle passes none); F, cof F and J follow
    // from H there
    bool precise = false;
    Strain32 st{};
    if constexpr (std::is_same<R, float>::value) {
        if (rk) {
            precise = true;
            V3<float> d[3] = {e1 - rk[0], e2 - rk[1], e3 - rk[2]};
            st = preciseStrain(d, br);
            preciseFrame(st, F, cof, J);
        }
    }
    if (!precise) {
        for (int c = 0; c < 3; c++) F[c] = e1 * comp(br[0], c) + e2 * comp(br[1], c) + e3 * comp(br[2], c);
        cof[0] = cross3(F[1], F[2]);
        cof[1] = cross3(F[2], F[0]);
        cof[2] = cross3(F[0], F[1]);
        J = dot3(F[0], cof[0]);
    }
    const bool safeguarded = mode == 2;
    V3<R> gh[4], gd[4];
    for (int k = 1; k < 4; k++) {
        V3<R> b = br[k - 1];
        gh[k] = cof[0] * b.x + cof[1] * b.y + cof[2] * b.z;
    }
    R scale = (R)1 + dampingOverH;
    if (mode == 0) {
        R CH = J - (R)1;
        if constexpr (std::is_same<R, float>::value)
            if (precise) CH = st.a;
        gh[0] = (gh[1] + gh[2] + gh[3]) * (R)-1;
        R bulkScale = scale;
        R effectiveAh = ah / bulkScale;
        R vh = 0;
        R schur = effectiveAh;
        for (int k = 0; k < 4; k++) vh += dot3(gh[k], p[k] - pv[k]);
        for (int k = 0; k < 4; k++) schur += w[k] * dot3(gh[k], gh[k]);
        R rh = (-CH - ah * lamH - (bulkScale - (R)1) * vh) / bulkScale;
        if (!(schur > 0)) return false;
        R dlH = omega * rh / schur;
        if (!std::isfinite(dlH)) return false;
        lamH = lamH + dlH;
        for (int k = 0; k < 4; k++) out[k] = p[k] + gh[k] * dlH * w[k];
        return true;
    }
    R I1 = precise ? (R)0 : dot3(F[0], F[0]) + dot3(F[1], F[1]) + dot3(F[2], F[2]);  // the direct forms' only
    R positiveJ = fmax_<R>(J, (R)0.05f);
    R iso = isoPow(positiveJ);
    const bool jlive = J > (R)0.05f;
    R q = fmax_<R>(iso * I1 - (R)3, (R)0);
    V3<R> Gh[3];
    if constexpr (std::is_same<R, float>::value)
        if (precise) q = fmax_<R>(preciseDev(st, F, cof, jlive, positiveJ, iso, Gh), (R)0);
    R CD = std::sqrt(fmax_<R>(q, (R)1e-12f));
    R devScale = (R)1 / CD;
    R volumeGradient = jlive ? iso * I1 / ((R)3 * positiveJ) : (R)0;
    R CH = safeguarded ? guardLog(positiveJ) + fmin_<R>((R)0, (J - (R)0.05f) / (R)0.05f) : J - (R)1;
    if constexpr (std::is_same<R, float>::value)
        if (precise)
            CH = safeguarded ? (jlive ? log1pOwn(st.a) : guardLog(positiveJ)) + fmin_<R>(0.f, (J - 0.05f) / 0.05f)
                             : st.a;
    for (int k = 1; k < 4; k++) {
        V3<R> b = br[k - 1];
        if (precise) gd[k] = (Gh[0] * b.x + Gh[1] * b.y + Gh[2] * b.z) * devScale;
        else gd[k] = ((F[0] * b.x + F[1] * b.y + F[2] * b.z) * iso - gh[k] * volumeGradient) * devScale;
        if (safeguarded) gh[k] = gh[k] * ((R)1 / positiveJ);
    }
    gd[0] = (gd[1] + gd[2] + gd[3]) * (R)-1;
    gh[0] = (gh[1] + gh[2] + gh[3]) * (R)-1;
    R dd = 0, dh = 0, vd = 0, vh = 0;
    for (int k = 0; k < 4; k++) {
        dd += w[k] * dot3(gd[k], gd[k]);
        dh += w[k] * dot3(gd[k], gh[k]);
        V3<R> dx = p[k] - pv[k];
        vd += dot3(gd[k], dx);
        vh += dot3(gh[k], dx);
    }
    R bulkScale = safeguarded ? (R)1 + bulkDampingOverH : scale;
    R effectiveAd = ad / scale, effectiveAh = ah / bulkScale;
    R A = dd + effectiveAd, Bc = dh, gamma = Bc / A;
    R schur = effectiveAh + effectiveAd * gamma * gamma;
    for (int k = 0; k < 4; k++) {
        V3<R> residual = gh[k] - gd[k] * gamma;
        schur += w[k] * dot3(residual, residual);
    }
    R rd = (-CD - ad * lamD - dampingOverH * vd) / scale;
    R rh = (-CH - ah * lamH - (bulkScale - (R)1) * vh) / bulkScale;
    if (!(A > 0) || !(schur > 0)) return false;
    R dlH = omega * (rh - gamma * rd) / schur, dlD = (omega * rd - Bc * dlH) / A;
    if (!std::isfinite(dlD) || !std::isfinite(dlH)) return false;
    if (safeguarded) {
        R largest = 0;
        for (int k = 0; k < 4; k++) largest = fmax_<R>(largest, len3((gd[k] * dlD + gh[k] * dlH) * w[k]));
        R limit = (R)0.15f / fmax_<R>(len3(br[0]), fmax_<R>(len3(br[1]), len3(br[2])));
        R accepted = largest > limit ? limit / largest : (R)1;
        dlD *= accepted;
        dlH *= accepted;
    }
    lamD = lamD + dlD;
    lamH = lamH + dlH;
    for (int k = 0; k < 4; k++) out[k] = p[k] + (gd[k] * dlD + gh[k] * dlH) * w[k];
    return true;
}

template <class R> void solveTet(Tissue<R>& s, int e, R dampingOverH, R bulkDampingOverH, bool useLambda) {
    if (s.tactive[e] == 0) {  // the substep's reset (xpbd_solver.cu:3536-3539)
        if (!useLambda) {
            s.lamDev[e] = 0;
            s.lamHyd[e] = 0;
        }
        return;
    }
    const int* id = &s.tids[4 * e];
    V3<R> p[4], pv[4], out[4];
    R w[4];
    for (int k = 0; k < 4; k++) { p[k] = s.pos[id[k]]; pv[k] = s.prev[id[k]]; w[k] = s.invMass[id[k]]; }
    const R(*B)[3] = s.typeB[s.ttype[e]];
    V3<R> br[3] = {mk<R>(B[0][0], B[0][1], B[0][2]), mk<R>(B[1][0], B[1][1], B[1][2]), mk<R>(B[2][0], B[2][1], B[2][2])};
    R lamD = useLambda ? s.lamDev[e] : (R)0, lamH = useLambda ? s.lamHyd[e] : (R)0;
    if (!tetProject<R>(s.d.tet_mode, p, pv, w, br, s.alphaDev[e], s.alphaHyd[e], lamD, lamH, dampingOverH,
                       bulkDampingOverH, (R)s.d.omega, out, s.d.precise_strain ? s.typeRk[s.ttype[e]] : nullptr)) {
        if (!useLambda) {  // the substep's reset (xpbd_solver.cu:3536-3539): a row that writes nothing keeps 0
            s.lamDev[e] = 0;
            s.lamHyd[e] = 0;
        }
        return;
    }
    s.lamDev[e] = lamD;
    s.lamHyd[e] = lamH;
    for (int k = 0; k < 4; k++) s.pos[id[k]] = out[k];
}

// ---- tissue guard passes (xpbd_solver.cu:588-741) ------------------------------------------------------------------
template <class R> void typeRows(const Tissue<R>& s, int e, V3<R> b[3]) {
    const R(*B)[3] = s.typeB[s.ttype[e]];
    for (int r = 0; r < 3; r++) b[r] = mk<R>(B[r][0], B[r][1], B[r][2]);
}

// projectTensile (xpbd_solver.cu:591-620) as a pure function of one tet: on its six edges (i < k, in order) the
// finite tensile energy of strength S, C = s^3 with s = max(L / L0 - 1, 0) and compliance alphaDev / S; the rest
// edges are the columns of Dm = B^-1, formed from the Dm^-1 rows b. lam: the six fibre multipliers, updated.
template <class R> void tensileProject(V3<R> p[4], const R w[4], const V3<R> b[3], R alphaDev, R strength, R lam[6]) {
    if (!(strength > 0)) return;
    R detB = dot3(b[0], cross3(b[1], b[2]));
    V3<R> rest[4] = {mk<R>(0, 0, 0), cross3(b[1], b[2]) * ((R)1 / detB), cross3(b[2], b[0]) * ((R)1 / detB),
                     cross3(b[0], b[1]) * ((R)1 / detB)};
    int fiber = 0;
    for (int i = 0; i < 4; i++)
        for (int k = i + 1; k < 4; k++, fiber++) {
            V3<R> edge = p[k] - p[i];
            R length = len3(edge), restLength = len3(rest[k] - rest[i]);
            R strain = fmax_<R>(length / restLength - (R)1, (R)0), gradient = (R)3 * strain * strain / restLength;
            R alpha = alphaDev / fmax_<R>(strength, (R)1e-10f);
            if (strain > 0 && length > 0) {
                R dl = (-strain * strain * strain - alpha * lam[fiber]) / ((w[i] + w[k]) * gradient * gradient + alpha);
                V3<R> g = edge * (gradient / length);
                lam[fiber] += dl;
                p[i] = p[i] - g * (w[i] * dl);
                p[k] = p[k] + g * (w[k] * dl);
            }
        }
}

template <class R> void tensilePass(Tissue<R>& s, int e, R strength) {  // k_TensileReinforcement on tet e
    if (!s.tactive[e]) return;
    const int* id = &s.tids[4 * e];
    V3<R> p[4], b[3];
    R w[4];
    for (int k = 0; k < 4; k++) { p[k] = s.pos[id[k]]; w[k] = s.invMass[id[k]]; }
    typeRows(s, e, b);
    tensileProject(p, w, b, s.alphaDev[e], strength, &s.fiberLambda[6 * (size_t)e]);
    for (int k = 0; k < 4; k++) s.pos[id[k]] = p[k];
}

template <class R> R tetJ(const Tissue<R>& s, int e, const std::vector<V3<R>>& x) {  // k_CheckCandidateVolume's J
    const int* id = &s.tids[4 * e];
    V3<R> b[3];
    typeRows(s, e, b);
    V3<R> p = x[id[0]], u = x[id[1]] - p, v = x[id[2]] - p, w = x[id[3]] - p;
    return dot3(u, cross3(v, w)) * dot3(b[0], cross3(b[1], b[2]));
}

// k_RepairMinimumVolume (:694-709) as a pure function of one tet: below minimum + 1e-4 it is pushed to minimum +
// 2e-4 by one rigid Newton step on J along w-weighted gradients. Returns false where it does not write.
template <class R> bool repairProject(V3<R> p[4], const R w[4], const V3<R> b[3], R minimum) {
    V3<R> u = p[1] - p[0], v = p[2] - p[0], x = p[3] - p[0];
    R detB = dot3(b[0], cross3(b[1], b[2]));
    R J = dot3(u, cross3(v, x)) * detB;
    if (!std::isfinite(J) || J >= minimum + (R)1e-4f) return false;
    V3<R> g[4];
    g[1] = cross3(v, x) * detB;
    g[2] = cross3(x, u) * detB;
    g[3] = cross3(u, v) * detB;
    g[0] = (g[1] + g[2] + g[3]) * (R)-1;
    R den = 0;
    for (int k = 0; k < 4; k++) den += w[k] * dot3(g[k], g[k]);
    if (den < (R)1e-20f) return false;
    R dl = (minimum + (R)2e-4f - J) / den;
    for (int k = 0; k < 4; k++) p[k] = p[k] + g[k] * (w[k] * dl);
    return true;
}

template <class R> void repairVolume(Tissue<R>& s, int e, R minimum) {
    if (!s.tactive[e]) return;
    const int* id = &s.tids[4 * e];
    V3<R> p[4], b[3];
    R w[4];
    for (int k = 0; k < 4; k++) { p[k] = s.pos[id[k]]; w[k] = s.invMass[id[k]]; }
    typeRows(s, e, b);
    if (repairProject(p, w, b, minimum))
        for (int k = 0; k < 4; k++) s.pos[id[k]] = p[k];
}

// k_NeedFeasible (:663-680): J is multi-affine in the four corners' interpolation parameters, so the 16 corner
// configurations between the substep start and the candidate bound every partial move; a cheap Lipschitz bound
// certifies most tets first. Uncertified tets mark their corners.
template <class R> void needFeasible(Tissue<R>& s, int e, R minimum) {
    if (!s.tactive[e]) return;
    const int* id = &s.tids[4 * e];
    V3<R> x[4], v[4];
    for (int k = 0; k < 4; k++) { x[k] = s.feasStart[id[k]]; v[k] = s.pos[id[k]] - x[k]; }
    V3<R> b[3];
    typeRows(s, e, b);
    R detB = dot3(b[0], cross3(b[1], b[2]));
    V3<R> u = x[1] - x[0], vv = x[2] - x[0], w = x[3] - x[0];
    R r0 = len3(v[0]), r1 = r0 + len3(v[1]), r2 = r0 + len3(v[2]), r3 = r0 + len3(v[3]);
    R bound = r1 * len3(cross3(vv, w)) + r2 * len3(cross3(w, u)) + r3 * len3(cross3(u, vv)) + r1 * r2 * len3(w) +
              r1 * r3 * len3(vv) + r2 * r3 * len3(u) + r1 * r2 * r3;
    if (dot3(u, cross3(vv, w)) * detB - bound * std::fabs(detB) >= minimum + (R)1e-5f) return;
    bool unsafe = false;
    for (int mask = 0; mask < 16; mask++) {
        V3<R> p[4];
        for (int k = 0; k < 4; k++) p[k] = x[k] + v[k] * (R)((mask >> k) & 1);
        R j = dot3(p[1] - p[0], cross3(p[2] - p[0], p[3] - p[0])) * detB;
        if (!std::isfinite(j) || j < minimum + (R)1e-5f) { unsafe = true; break; }
    }
    if (unsafe)
        for (int k = 0; k < 4; k++) s.needed[id[k]] = 1;
}

// k_FeasibleVertex (:634-661) on node i: its move from the substep start to the candidate, cut back so that no
// live incident tet drops below min(minimum, max(0, J)) -- with the other corners fixed, J is linear in this
// node's displacement -- with a 0.95 margin.
template <class R> void feasibleVertex(Tissue<R>& s, int i, R minimum) {
    if (!s.needed[i]) return;
    V3<R> delta = s.feasTarget[i] - s.pos[i];
    R accepted = 1;
    if (!std::isfinite(delta.x + delta.y + delta.z)) return;
    for (int k = s.incOff[i]; k < s.incOff[i + 1]; k++) {
        int t = s.incTets[k];
        if (!s.tactive[t]) continue;
        const int* id = &s.tids[4 * t];
        V3<R> a = s.pos[id[0]], u = s.pos[id[1]] - a, v = s.pos[id[2]] - a, w = s.pos[id[3]] - a;
        V3<R> b[3];
        typeRows(s, t, b);
        R detB = dot3(b[0], cross3(b[1], b[2]));
        R j = dot3(u, cross3(v, w)) * detB;
        V3<R> grad[3] = {cross3(v, w), cross3(w, u), cross3(u, v)};
        V3<R> g = i == id[0] ? (grad[0] + grad[1] + grad[2]) * (R)-1
                             : (i == id[1] ? grad[0] : (i == id[2] ? grad[1] : grad[2]));
        R dj = dot3(g, delta) * detB;
        if (dj < 0) {
            R floor = fmin_<R>(minimum, fmax_<R>((R)0, j));
            if (j + dj < floor) accepted = fmin_<R>(accepted, fmax_<R>((R)0, (R)0.95f * (j - floor) / (-dj)));
        }
    }
    s.pos[i] = s.pos[i] + delta * accepted;
}

// finishFeasible (:718-741) at the end of a substep: the candidate's live tets are checked against the minimum J;
// if any fails, two colour sweeps of the volume repair, a second check, and if still unsafe the corners of the
// uncertified tets go back to the substep start and move toward their candidate one parity colour at a time (eight
// lattice parity classes, so no two nodes of a class share a tet), starting from the substep's phase.
template <class R> void finishFeasible(Tissue<R>& s) {
    if (s.d.guard_passes <= 0 || s.T <= 0) return;
    const R minimum = (R)s.d.guard_min_volume;
    const int phase = s.substepIndex & 7;
    std::fill(s.needed.begin(), s.needed.end(), 0);
    auto check = [&]() {
        for (int t = 0; t < s.T; t++) {
            if (!s.tactive[t]) continue;
            R J = tetJ(s, t, s.pos);
            if (!std::isfinite(J) || J < minimum) return true;
        }
        return false;
    };
    if (check())
        for (int pass = 0; pass < 2; pass++)
            for (size_t c = 0; c < s.tcolorCnt.size(); c++)
                for (int q = 0; q < s.tcolorCnt[c]; q++) repairVolume(s, s.tcolorOrder[s.tcolorOff[c] + q], minimum);
    if (!check()) return;
    for (int t = 0; t < s.T; t++) needFeasible(s, t, minimum);
    for (int i = 0; i < s.N; i++)
        if (s.needed[i]) { s.feasTarget[i] = s.pos[i]; s.pos[i] = s.feasStart[i]; }
    for (int k = 0; k < 8; k++) {
        int colour = (k + phase) & 7;
        for (int z = (colour >> 2) & 1; z < s.nz; z += 2)
            for (int y = (colour >> 1) & 1; y < s.ny; y += 2)
                for (int x = colour & 1; x < s.nx; x += 2) feasibleVertex(s, s.cid(x, y, z), minimum);
    }
}

template <class R> ToolR<R> lerpTool(const ToolR<R>& from, const ToolR<R>& to, R t) {  // xpbd_solver.cu:3467-3483
    ToolR<R> r = to;
    r.jawHinge = from.jawHinge * ((R)1 - t) + to.jawHinge * t;
    r.jawAxis = from.jawAxis * ((R)1 - t) + to.jawAxis * t;
    r.jawNormal = from.jawNormal * ((R)1 - t) + to.jawNormal * t;
    r.shaftA = from.shaftA * ((R)1 - t) + to.shaftA * t;
    r.shaftB = from.shaftB * ((R)1 - t) + to.shaftB * t;
    r.jawGap = from.jawGap * ((R)1 - t) + to.jawGap * t;
    r.bladeA = from.bladeA * ((R)1 - t) + to.bladeA * t;
    r.bladeB = from.bladeB * ((R)1 - t) + to.bladeB * t;
    return r;
}

// ---- legacy grasp (contact model 0): capture/release, slab ring with Coulomb anchors, core pins, break check, CCD gate

// The jaw frame of a substep's interpolated pose as xpbd_substep derives it for the grasp kernels
// (xpbd_solver.cu:3572-3580): unit axis, normal made orthogonal to it, w = axis x normal, half the floored gap.
template <class R> struct GraspFrame {
    V3<R> hinge, axis, normal, w, shaftA, shaftB;
    R shaftR, jawLen, jawHalfWidth, gapHalf;
};

template <class R> GraspFrame<R> graspFrame(const ToolR<R>& t, R closedGap) {
    GraspFrame<R> f;
    f.axis = len3(t.jawAxis) > (R)1e-9f ? t.jawAxis * ((R)1 / len3(t.jawAxis)) : mk<R>(0, 0, 1);
    V3<R> n = len3(t.jawNormal) > (R)1e-9f ? t.jawNormal * ((R)1 / len3(t.jawNormal)) : mk<R>(0, 1, 0);
    n = n - f.axis * dot3(n, f.axis);
    if (len3(n) < (R)1e-7f) n = cross3(f.axis, std::fabs(f.axis.x) < (R)0.9f ? mk<R>(1, 0, 0) : mk<R>(0, 1, 0));
    f.normal = n * ((R)1 / len3(n));
    f.w = cross3(f.axis, f.normal);
    f.gapHalf = (R)0.5f * fmax_<R>(t.jawGap, closedGap);
    f.hinge = t.jawHinge;
    f.shaftA = t.shaftA;
    f.shaftB = t.shaftB;
    f.shaftR = t.shaftR;
    f.jawLen = t.jawLen;
    f.jawHalfWidth = t.jawHalfWidth;
    return f;
}

// xpbd_set_contact_model's reset of the grasp state (xpbd_solver.cu:2200-2207).
template <class R> void setContactModel(Tissue<R>& s, int model, const OracleGrasp& gp) {
    s.contactModel = model;
    s.gp = gp;
    s.grasping = false;
    s.grasped.assign(s.N, 0);
    s.core.assign(s.N, 0);
    for (auto* v : {&s.offU, &s.offN, &s.offW, &s.anchU, &s.anchW, &s.slab, &s.slabPrev, &s.ema, &s.jawLam,
                    &s.shaftLam})
        v->assign(s.N, (R)0);
    s.coreLam.assign(s.N, mk<R>(0, 0, 0));
}

// xpbd_set_tool's legacy branch (xpbd_solver.cu:2318-2341, :2372-2470), run after the new pose is stored: capture on
// grasp01 rising past the capture threshold, release on it falling below the release threshold. Returns 1 on capture,
// 2 on release, 0 otherwise. A capture makes the next step start at the new pose.
template <class R> int graspSetTool(Tissue<R>& s) {
    const ToolR<R>& t = s.tool;
    R alen = len3(t.jawAxis);
    V3<R> axis = alen > (R)1e-9f ? t.jawAxis * ((R)1 / alen) : mk<R>(0, 0, 1);
    R nlen = len3(t.jawNormal);
    V3<R> normal = nlen > (R)1e-9f ? t.jawNormal * ((R)1 / nlen) : mk<R>(0, 1, 0);
    V3<R> w = cross3(axis, normal);
    R wlen = len3(w);
    w = wlen > (R)1e-9f ? w * ((R)1 / wlen) : mk<R>(1, 0, 0);
    R gapHalf = (R)0.5f * fmax_<R>(t.jawGap, (R)s.gp.closed_gap);
    bool nowAbove = t.grasp01 > (R)s.gp.capture, nowBelow = t.grasp01 < (R)s.gp.release;
    if (!s.grasping && nowAbove) {
        // k_MarkGraspCandidates (:930-951): live, unpinned corners inside the jaw box widened by the capture skin
        std::vector<int> idxs;
        std::vector<R> U(s.N, (R)0), Vv(s.N, (R)0), Nn(s.N, (R)0);
        for (int c = 0; c < s.N; c++) {
            if (s.active[c] == 0 || s.pinned[c] != 0) continue;
            V3<R> rel = s.pos[c] - t.jawHinge;
            R u = dot3(rel, axis), v = dot3(rel, w), n = dot3(rel, normal);
            if (u < 0 || u > t.jawLen) continue;
            if (std::fabs(v) > t.jawHalfWidth) continue;
            if (std::fabs(n) > gapHalf + (R)s.gp.skin) continue;
            idxs.push_back(c);
            U[c] = u; Vv[c] = v; Nn[c] = n;
        }
        std::fill(s.grasped.begin(), s.grasped.end(), 0);
        std::fill(s.core.begin(), s.core.end(), 0);
        for (auto* v : {&s.offU, &s.offN, &s.offW, &s.anchU, &s.anchW}) std::fill(v->begin(), v->end(), (R)0);
        if (!idxs.empty()) {
            // normalised-ellipse core over the captured set's own u / v extents (:2405-2431)
            R uMin = (R)1e30f, uMax = (R)-1e30f, vMin = (R)1e30f, vMax = (R)-1e30f;
            for (int c : idxs) {
                uMin = fmin_<R>(uMin, U[c]); uMax = fmax_<R>(uMax, U[c]);
                vMin = fmin_<R>(vMin, Vv[c]); vMax = fmax_<R>(vMax, Vv[c]);
            }
            R uMid = (R)0.5f * (uMin + uMax), uHalf = fmax_<R>((R)0.5f * (uMax - uMin), (R)1e-6f);
            R vMid = (R)0.5f * (vMin + vMax), vHalf = fmax_<R>((R)0.5f * (vMax - vMin), (R)1e-6f);
            R frac = (R)s.gp.core_fraction;
            int bestIdx = -1;
            R bestD2 = (R)1e30f;
            for (int c : idxs) {
                s.grasped[c] = 1;
                R nu = (U[c] - uMid) / uHalf, nv = (Vv[c] - vMid) / vHalf;
                R d2 = nu * nu + nv * nv;
                if (d2 < frac * frac) s.core[c] = 1;
                if (d2 < bestD2) { bestD2 = d2; bestIdx = c; }
            }
            bool anyCore = false;
            for (int c : idxs) anyCore = anyCore || s.core[c] != 0;
            if (!anyCore && bestIdx >= 0) s.core[bestIdx] = 1;  // a grasp is never core-less
            for (int c : idxs) {
                s.offU[c] = U[c]; s.offN[c] = Nn[c]; s.offW[c] = Vv[c];
                s.anchU[c] = U[c]; s.anchW[c] = Vv[c];
            }
        }
        std::fill(s.slab.begin(), s.slab.end(), (R)0);
        std::fill(s.slabPrev.begin(), s.slabPrev.end(), (R)0);
        std::fill(s.ema.begin(), s.ema.end(), (R)0);
        s.grasping = true;
        s.toolFrom = s.tool;
        return 1;
    }
    if (s.grasping && nowBelow) {
        for (int c = 0; c < s.N; c++)  // k_ReleaseDampVelocity (:956-961), then the sets are cleared
            if (s.grasped[c] != 0) s.vel[c] = s.vel[c] * (R)0.1f;
        std::fill(s.grasped.begin(), s.grasped.end(), 0);
        std::fill(s.core.begin(), s.core.end(), 0);
        s.grasping = false;
        return 2;
    }
    return 0;
}

// k_SolveSlabCoulomb (xpbd_solver.cu:1131-1199) for a ring corner, k_SolveCorePin (:960-977) for a core corner.
template <class R> void graspPass(Tissue<R>& s, int c, const GraspFrame<R>& f, R h, R contactAlpha, R coreAlpha) {
    R wc = s.invMass[c];
    if (!(wc > 0)) return;
    R omega = (R)s.gp.grasp_omega, maxDisp = (R)s.gp.grasp_displacement;
    V3<R> x = s.pos[c];
    if (s.grasped[c] != 0 && s.core[c] == 0) {
        V3<R> rel = x - f.hinge;
        R u = dot3(rel, f.axis), v = dot3(rel, f.w), n = dot3(rel, f.normal);
        if (u >= 0 && u <= f.jawLen && std::fabs(v) <= f.jawHalfWidth) {
            R absN = std::fabs(n);
            if (absN > f.gapHalf) {
                R Ctrue = f.gapHalf - absN;
                R sgn = n >= 0 ? (R)1 : (R)-1;
                R denom = wc + contactAlpha;
                R dLambda = (-Ctrue - contactAlpha * s.slab[c]) / denom;
                R applied = omega * dLambda;
                R corrMag = wc * applied;
                if (std::fabs(corrMag) > maxDisp) {
                    R capScale = maxDisp / std::fabs(corrMag);
                    applied *= capScale;
                    corrMag *= capScale;
                }
                s.slab[c] += applied;
                x = x - f.normal * (sgn * corrMag);
                rel = x - f.hinge;
                n = dot3(rel, f.normal);
            }
        }
        R Fn = std::fabs(s.slabPrev[c]) / (h * h);
        V3<R> anchorWorld = f.hinge + f.axis * s.anchU[c] + f.w * s.anchW[c] + f.normal * n;
        V3<R> tangential = x - anchorWorld;
        R dist = len3(tangential);
        R coneRadius = (R)s.gp.friction * Fn * (h * h * wc);
        if (dist <= coneRadius + (R)1e-9f) {  // stick: compliant, capped pull to the anchor
            V3<R> corr = tangential * omega;
            R corrLen = len3(corr);
            if (corrLen > maxDisp) corr = corr * (maxDisp / corrLen);
            x = x - corr;
        } else {  // slip: move by the cone radius and drag the anchor along
            V3<R> dir = tangential * ((R)1 / dist);
            x = x - dir * coneRadius;
            V3<R> newAnchor = anchorWorld + dir * (dist - coneRadius);
            V3<R> newRel = newAnchor - f.hinge;
            s.anchU[c] = dot3(newRel, f.axis);
            s.anchW[c] = dot3(newRel, f.w);
        }
        s.pos[c] = x;
    } else if (s.core[c] != 0) {
        V3<R> target = f.hinge + f.axis * s.offU[c] + f.normal * s.offN[c] + f.w * s.offW[c];
        V3<R> applied = ((x - target) + s.coreLam[c] * coreAlpha) * (-omega / (wc + coreAlpha));
        R corrLen = len3(applied) * wc;
        if (corrLen > maxDisp) applied = applied * (maxDisp / corrLen);
        s.coreLam[c] = s.coreLam[c] + applied;
        s.pos[c] = x + applied * wc;
    }
}

// k_CoreBreakCheck (xpbd_solver.cu:982-1000): a core corner whose one-second EMA of |lambda| / h^2 passes the break
// force leaves the grasp for good.
template <class R> void coreBreak(Tissue<R>& s, int c, R h, R kEMA) {
    if (s.core[c] == 0) return;
    R raw = len3(s.coreLam[c]) / (h * h);
    s.ema[c] += (raw - s.ema[c]) * kEMA;
    if (s.ema[c] > (R)s.gp.break_force) {
        s.core[c] = 0;
        s.grasped[c] = 0;
    }
}

// k_SolveCCDContact (xpbd_solver.cu:1049-1123): the anti-tunnelling gate against the shaft capsule (every corner) and
// the jaw slab (corners not grasped), firing only for a corner approaching faster than the threshold.
template <class R> void ccdGate(Tissue<R>& s, int c, const GraspFrame<R>& f, R contactAlpha) {
    R wc = s.invMass[c];
    if (!(wc > 0)) return;
    R omega = (R)s.gp.ccd_omega, thresh = (R)s.gp.ccd_velocity, maxDisp = (R)s.gp.ccd_displacement;
    V3<R> x = s.pos[c], vel = s.vel[c];
    {
        V3<R> ab = f.shaftB - f.shaftA;
        R ab2 = dot3(ab, ab);
        R tt = ab2 > (R)1e-12f ? dot3(x - f.shaftA, ab) / ab2 : (R)0;
        tt = fmax_<R>((R)0, fmin_<R>((R)1, tt));
        V3<R> closest = f.shaftA + ab * tt;
        V3<R> dvec = x - closest;
        R dist = len3(dvec);
        R C = dist - f.shaftR;
        if (C < 0) {
            V3<R> n = dist > (R)1e-9f ? dvec * ((R)1 / dist) : mk<R>(0, 1, 0);
            R vApproach = -dot3(vel, n);
            if (vApproach > thresh) {
                R denom = wc + contactAlpha;
                R dLambda = (-C - contactAlpha * s.shaftLam[c]) / denom;
                R applied = omega * dLambda;
                V3<R> corr = n * (wc * applied);
                R corrLen = len3(corr);
                if (corrLen > maxDisp) {
                    R capScale = maxDisp / corrLen;
                    applied *= capScale;
                    corr = corr * capScale;
                }
                s.shaftLam[c] += applied;
                x = x + corr;
            }
        }
    }
    if (s.grasped[c] == 0) {
        V3<R> rel = x - f.hinge;
        R u = dot3(rel, f.axis), v = dot3(rel, f.w), n = dot3(rel, f.normal);
        if (u >= 0 && u <= f.jawLen && std::fabs(v) <= f.jawHalfWidth) {
            R absN = std::fabs(n);
            if (absN > f.gapHalf) {
                R Ctrue = f.gapHalf - absN;
                R sgn = n >= 0 ? (R)1 : (R)-1;
                V3<R> nOut = f.normal * sgn;
                R vApproach = -dot3(vel, nOut);
                if (vApproach > thresh) {
                    R denom = wc + contactAlpha;
                    R dLambda = (-Ctrue - contactAlpha * s.jawLam[c]) / denom;
                    R applied = omega * dLambda;
                    R corrMag = wc * applied;
                    if (std::fabs(corrMag) > maxDisp) {
                        R capScale = maxDisp / std::fabs(corrMag);
                        applied *= capScale;
                        corrMag *= capScale;
                    }
                    s.jawLam[c] += applied;
                    x = x - f.normal * (sgn * corrMag);
                }
            }
        }
    }
    s.pos[c] = x;
}

// One substep of the legacy branch (xpbd_solver.cu:3460-3721): with the legacy contact model, K = 2
// material passes while grasping, each followed by the grasp ring and core pins, then the break check and the CCD
// gate; with the friction model, friction tool contact of lattice nodes (:3659-3665, friction_contact.cuh:143-169);
// with the tissue guard, tensile reinforcement after every tet sweep and the minimum-volume safeguard before the
// velocity update.
template <class R> int substep(Tissue<R>& s) {
    if (!s.frameReady) return -1001;
    R h = s.h;
    int Nsub = s.d.substeps > 0 ? s.d.substeps : 1;
    R t = fmin_<R>((R)1, (R)(++s.substepIndex) / (R)Nsub);
    R previousT = fmax_<R>((R)0, t - (R)1 / (R)Nsub);
    ToolR<R> previousTool = lerpTool(s.toolFrom, s.tool, previousT);
    ToolR<R> stepTool = lerpTool(s.toolFrom, s.tool, t);
    refreshLiveness(s);
    const int guardPasses = s.d.guard_passes > 0 ? s.d.guard_passes : 0;
    if (guardPasses) s.feasStart = s.pos;  // beginFeasible (:684-686)
    // k_predict (:393-412)
    for (int p = 0; p < s.N; p++) {
        s.prev[p] = s.pos[p];
        if (s.invMass[p] > 0) {
            V3<R> a = s.ext[p] * s.invMass[p];
            V3<R> v = s.vel[p] + a * h;
            s.vel[p] = v;
            s.pos[p] = s.pos[p] + v * h;
        }
    }
    const bool strain = s.T > 0;
    const bool fallback = strain && s.d.edge_ks <= 0;
    const bool edges = s.d.edge_ks > 0 || fallback;
    R fallbackInvEh2 = fallback ? (R)1 / ((R)s.d.young * h * h) : (R)0;
    R dampingOverH = fmax_<R>((R)0, (R)s.d.strain_damping) / h;
    R bulkOverH = h > 0 ? (R)s.d.bulk_damping / h : (R)0;
    int iters = s.d.iters > 0 ? s.d.iters : 1;
    const bool legacy = s.contactModel == 0;
    const bool grasping = legacy && s.grasping;
    // the legacy multipliers (:3538-3545): the CCD gate's while the tool is armed, the grasp's while grasping
    if (legacy && s.toolArmed) {
        std::fill(s.jawLam.begin(), s.jawLam.end(), (R)0);
        std::fill(s.shaftLam.begin(), s.shaftLam.end(), (R)0);
    }
    if (grasping) {
        std::fill(s.slab.begin(), s.slab.end(), (R)0);
        std::fill(s.coreLam.begin(), s.coreLam.end(), mk<R>(0, 0, 0));
    }
    if (guardPasses) std::fill(s.fiberLambda.begin(), s.fiberLambda.end(), (R)0);  // (:3557)
    GraspFrame<R> frame{};
    R contactAlpha = 0, coreAlpha = 0;
    if (legacy) {
        frame = graspFrame(stepTool, (R)s.gp.closed_gap);
        contactAlpha = (R)s.gp.contact_compliance / (h * h);
        coreAlpha = (R)s.gp.core_compliance / (h * h);
    }
    const int K = grasping ? 2 : 1;  // a second material pass only while genuinely grasping (:3571)
    for (int k = 0; k < K; k++) {
        for (int it = 0; it < iters; it++) {
            bool useLambda = k > 0 || it > 0;  // the multipliers were reset at the start of the substep (:3534-3563)
            if (edges)
                for (size_t c = 0; c < s.ecolorCnt.size(); c++) {
                    int off = s.ecolorOff[c], cnt = s.ecolorCnt[c];
#pragma omp parallel for schedule(static) if (cnt > 4096)
                    for (int q = 0; q < cnt; q++)
                        solveEdge(s, off + q, s.alphaTilde, fallback, fallbackInvEh2, useLambda);
                }
            if (s.doBend)  // then the bending angle constraints (:3590-3596)
                for (size_t c = 0; c < s.bcolorCnt.size(); c++) {
                    int off = s.bcolorOff[c], cnt = s.bcolorCnt[c];
#pragma omp parallel for schedule(static) if (cnt > 4096)
                    for (int q = 0; q < cnt; q++) solveBend(s, off + q, useLambda);
                }
            if (strain)
                for (size_t c = 0; c < s.tcolorCnt.size(); c++) {
                    int off = s.tcolorOff[c], cnt = s.tcolorCnt[c];
#pragma omp parallel for schedule(static) if (cnt > 2048)
                    for (int q = 0; q < cnt; q++)
                        solveTet(s, s.tcolorOrder[off + q], dampingOverH, bulkOverH, useLambda);
                }
            // k_TensileReinforcement after the tet sweep, inside solveTetGraph (:2531-2533)
            if (strain)
                for (int pass = 0; pass < guardPasses; pass++)
                    for (size_t c = 0; c < s.tcolorCnt.size(); c++) {
                        int off = s.tcolorOff[c], cnt = s.tcolorCnt[c];
#pragma omp parallel for schedule(static) if (cnt > 2048)
                        for (int q = 0; q < cnt; q++)
                            tensilePass(s, s.tcolorOrder[off + q], (R)s.d.tensile_strength);
                    }
        }
        if (grasping)  // k_SolveSlabCoulomb, then k_SolveCorePin (:3603-3610); each touches only its own corner
            for (int c = 0; c < s.N; c++) graspPass(s, c, frame, h, contactAlpha, coreAlpha);
    }
    if (grasping) {  // k_CoreBreakCheck once after the coupling passes (:3612-3617)
        R kEMA = -std::expm1(-h);
        for (int c = 0; c < s.N; c++) coreBreak(s, c, h, kEMA);
    }
    if (legacy && s.toolArmed)  // k_SolveCCDContact, once, last (:3618-3624)
        for (int c = 0; c < s.N; c++) ccdGate(s, c, frame, contactAlpha);
    // friction contact: k_FindToolContacts x kinds, k_SolveFrictionContacts x2, k_NodeReaction (:3659-3665)
    if (!legacy && s.d.contact_enabled && (s.toolArmed || s.bladeArmed)) {
        const int kinds = 4;
        Shape<R> from[kinds], to[kinds];
        for (int kind = 0; kind < kinds; kind++) {
            from[kind] = contactShape(previousTool, kind, (R)s.d.jaw_thickness, (R)s.d.jaw_gap_closed);
            to[kind] = contactShape(stepTool, kind, (R)s.d.jaw_thickness, (R)s.d.jaw_gap_closed);
        }
        R alpha = (R)s.d.contact_compliance / (h * h);
        R muS = (R)s.d.mu_s, muD = (R)s.d.mu_d, radius = (R)s.d.contact_radius;
        R scaleR = (R)1 / (h * h * fmax_<R>((R)1, (R)Nsub));
        for (int i = 0; i < s.N; i++) {
            if (!(s.invMass[i] > 0)) continue;
            Record<R> rec[kinds];
            R lamN[kinds];
            V3<R> lamT[kinds];
            for (int kind = 0; kind < kinds; kind++) {
                rec[kind] = Record<R>{};
                lamN[kind] = 0;
                lamT[kind] = mk<R>(0, 0, 0);
                bool armed = kind == 3 ? s.bladeArmed : s.toolArmed;
                bool present = to[kind].capsule ? to[kind].radius > 0 : (to[kind].half.x > 0 && to[kind].half.z > 0);
                if (armed && present) rec[kind] = traceToolContact(from[kind], to[kind], radius, s.prev[i], s.pos[i]);
            }
            R w = s.invMass[i];
            for (int pass = 0; pass < 2; pass++)
                for (int kind = 0; kind < kinds; kind++)
                    if (rec[kind].valid)
                        frictionUpdate(rec[kind], w, alpha, muS, muD, s.prev[i], s.pos[i], lamN[kind], lamT[kind]);
            for (int kind = 0; kind < kinds; kind++)
                if (rec[kind].valid) {  // k_NodeReaction: force of the tissue on the tool
                    V3<R> f = (rec[kind].normal * lamN[kind] + lamT[kind]) * (-scaleR);
                    s.reaction[0] += f.x; s.reaction[1] += f.y; s.reaction[2] += f.z;
                    if (lamN[kind] > 0) s.reactionContacts++;
                }
        }
    }
    finishFeasible(s);  // (:3678)
    // k_update_velocity, [k_perp_damp_edges], k_apply_velscale (:871-919)
    R invH = (R)1 / h;
    for (int p = 0; p < s.N; p++) s.vel[p] = s.invMass[p] > 0 ? (s.pos[p] - s.prev[p]) * invH : mk<R>(0, 0, 0);
    if (grasping) s.slabPrev = s.slab;  // the next substep's normal force (:3710-3711)
    if (s.d.beta_s > 0)
        for (size_t c = 0; c < s.ecolorCnt.size(); c++)
            for (int k = 0; k < s.ecolorCnt[c]; k++) {
                int e = s.ecolorOff[c] + k;
                if (s.eactive[e] == 0) continue;
                int i = s.ei[e], j = s.ej[e];
                perpDamp(s.vel[i], s.vel[j], s.pos[i], s.pos[j], s.invMass[i], s.invMass[j], h * (R)s.d.beta_s);
            }
    for (int p = 0; p < s.N; p++) s.vel[p] = s.invMass[p] > 0 ? s.vel[p] * s.velScale : mk<R>(0, 0, 0);
    if (s.substepIndex >= Nsub) s.toolFrom = s.tool;
    return 0;
}

// 3x3 matrices as three rows (physics.cu:27-64).
template <class R> struct M3 { V3<R> r0, r1, r2; };
template <class R> R det3(const M3<R>& m) {
    return m.r0.x * (m.r1.y * m.r2.z - m.r1.z * m.r2.y) - m.r0.y * (m.r1.x * m.r2.z - m.r1.z * m.r2.x) +
           m.r0.z * (m.r1.x * m.r2.y - m.r1.y * m.r2.x);
}
template <class R> M3<R> inverse3(const M3<R>& m) {  // adjugate times 1/det
    R invDet = (R)1 / det3(m);
    M3<R> a;
    a.r0 = mk<R>((m.r1.y * m.r2.z - m.r1.z * m.r2.y), -(m.r0.y * m.r2.z - m.r0.z * m.r2.y),
                 (m.r0.y * m.r1.z - m.r0.z * m.r1.y));
    a.r1 = mk<R>(-(m.r1.x * m.r2.z - m.r1.z * m.r2.x), (m.r0.x * m.r2.z - m.r0.z * m.r2.x),
                 -(m.r0.x * m.r1.z - m.r0.z * m.r1.x));
    a.r2 = mk<R>((m.r1.x * m.r2.y - m.r1.y * m.r2.x), -(m.r0.x * m.r2.y - m.r0.y * m.r2.x),
                 (m.r0.x * m.r1.y - m.r0.y * m.r1.x));
    return M3<R>{a.r0 * invDet, a.r1 * invDet, a.r2 * invDet};
}
template <class R> M3<R> transpose3(const M3<R>& m) {
    return M3<R>{mk<R>(m.r0.x, m.r1.x, m.r2.x), mk<R>(m.r0.y, m.r1.y, m.r2.y), mk<R>(m.r0.z, m.r1.z, m.r2.z)};
}
template <class R> V3<R> divs(V3<R> a, R d) { return mk<R>(a.x / d, a.y / d, a.z / d); }

// k_ComputeParticleRot (physics.cu:109-205) for every node, after the current frames became the previous ones
// (physics_compute_rotations, physics.cu:399-411): per axis the through-particle edge direction (one-sided where a
// neighbour is gone) blended with the mean previous axis of the live neighbours, a single lost axis recovered as the
// cross product of the other two, R = polar(M) by 16 Higham steps; identity for pinned / inactive nodes and for
// degenerate, reflected or unconverged frames.
template <class R> void computeRotations(Tissue<R>& s) {
    s.rotPrev = s.rot;
    const V3<R> I0 = mk<R>(1, 0, 0), I1 = mk<R>(0, 1, 0), I2 = mk<R>(0, 0, 1);
    for (int c = 0; c < s.N; c++) {
        V3<R>* out = &s.rot[3 * c];
        out[0] = I0; out[1] = I1; out[2] = I2;
        if (s.pinned[c] != 0 || s.active[c] == 0) continue;
        V3<R> axis[3];
        bool have[3] = {false, false, false};
        for (int k = 0; k < 3; k++) {
            int np = s.nbr[6 * c + 2 * k + 1], nm = s.nbr[6 * c + 2 * k];
            bool okp = np >= 0 && s.active[np] != 0, okm = nm >= 0 && s.active[nm] != 0;
            V3<R> e;
            if (okp && okm) e = s.pos[np] - s.pos[nm];
            else if (okp) e = s.pos[np] - s.pos[c];
            else if (okm) e = s.pos[c] - s.pos[nm];
            else continue;
            R len = len3(e);
            if (len < (R)1e-12f) continue;
            V3<R> dir = divs(e, len);
            V3<R> psum = mk<R>(0, 0, 0);
            int pc = 0;
            for (int q = 0; q < 6; q++) {
                int n = s.nbr[6 * c + q];
                if (n >= 0 && s.active[n] != 0) {
                    const V3<R>* rp = &s.rotPrev[3 * n];
                    psum = psum + mk<R>(comp(rp[0], k), comp(rp[1], k), comp(rp[2], k));
                    pc++;
                }
            }
            if (pc > 0) {
                R pl = len3(psum);
                if (pl > (R)1e-6f) {
                    V3<R> blended = dir + divs(psum, pl);
                    R dl = len3(blended);
                    if (dl > (R)0.25f) dir = divs(blended, dl);
                }
            }
            axis[k] = dir;
            have[k] = true;
        }
        int miss = -1, mc = 0;
        for (int k = 0; k < 3; k++)
            if (!have[k]) { miss = k; mc++; }
        if (mc == 1) {
            V3<R> cx = cross3(axis[(miss + 1) % 3], axis[(miss + 2) % 3]);
            R l = len3(cx);
            if (l > (R)1e-9f) { axis[miss] = divs(cx, l); have[miss] = true; }
        }
        if (!(have[0] && have[1] && have[2])) continue;
        M3<R> M{mk<R>(axis[0].x, axis[1].x, axis[2].x), mk<R>(axis[0].y, axis[1].y, axis[2].y),
                mk<R>(axis[0].z, axis[1].z, axis[2].z)};
        if (det3(M) <= (R)1e-4f) continue;
        M3<R> Rm = M;
        for (int it = 0; it < 16; it++) {
            M3<R> t = transpose3(inverse3(Rm));
            Rm.r0 = (Rm.r0 + t.r0) * (R)0.5f;
            Rm.r1 = (Rm.r1 + t.r1) * (R)0.5f;
            Rm.r2 = (Rm.r2 + t.r2) * (R)0.5f;
        }
        V3<R> c0 = mk<R>(Rm.r0.x, Rm.r1.x, Rm.r2.x), c1 = mk<R>(Rm.r0.y, Rm.r1.y, Rm.r2.y),
              c2 = mk<R>(Rm.r0.z, Rm.r1.z, Rm.r2.z);
        R ortho = fmax_<R>(fmax_<R>(std::fabs(dot3(c0, c0) - (R)1), std::fabs(dot3(c1, c1) - (R)1)),
                           std::fabs(dot3(c2, c2) - (R)1));
        ortho = fmax_<R>(ortho, fmax_<R>(std::fabs(dot3(c0, c1)), fmax_<R>(std::fabs(dot3(c0, c2)),
                                                                            std::fabs(dot3(c1, c2)))));
        if (ortho > (R)1e-3f) continue;
        out[0] = Rm.r0; out[1] = Rm.r1; out[2] = Rm.r2;
    }
}

// mulRtv (cut.cuh:75-78): R^T v with R as rows.
template <class R> V3<R> mulRtv(const V3<R>* rows, V3<R> v) {
    return rows[0] * v.x + rows[1] * v.y + rows[2] * v.z;
}

template <class R> void ribbonTick(Tissue<R>& s);

// LCS_Step (plugin_api.cpp:125-134): physics_step (physics.cu:331-394) with the CCD ribbon tick after every substep
// while a cut tool is set, the frames recomputed after a tick that emitted cut points (but the last), then the
// particle frames.
template <class R> int step(Tissue<R>& s, R dt) {
    int rc = beginFrame(s, dt);
    if (rc) return rc;
    int N = s.d.substeps > 0 ? s.d.substeps : 1;
    s.reaction[0] = s.reaction[1] = s.reaction[2] = 0;
    s.reactionContacts = 0;
    for (int n = 0; n < N; n++) {
        rc = substep(s);
        if (rc) return rc;
        if (s.cutToolSet) {
            uint32_t before = s.cutPointCounter;
            ribbonTick(s);
            if (s.cutPointCounter != before && n + 1 < N) computeRotations(s);
        }
    }
    computeRotations(s);
    return 0;
}

// ES_VoxelOffset / ES_LocalEdge (cut.cuh:126-137).
const int ES_OFF[3][4][3] = {{{0, 0, 0}, {0, -1, 0}, {0, -1, -1}, {0, 0, -1}},
                             {{0, 0, 0}, {0, 0, -1}, {-1, 0, -1}, {-1, 0, 0}},
                             {{0, 0, 0}, {-1, 0, 0}, {-1, -1, 0}, {0, -1, 0}}};
const int ES_EDGE[3][4] = {{0, 2, 6, 4}, {3, 7, 5, 1}, {8, 9, 10, 11}};

template <class R> bool voxIn(const Tissue<R>& s, int x, int y, int z) {
    return x >= 0 && y >= 0 && z >= 0 && x < s.d.vx && y < s.d.vy && z < s.d.vz;
}
template <class R> int voxId(const Tissue<R>& s, int x, int y, int z) { return x + s.d.vx * (y + s.d.vy * z); }

inline int directedSlot(int a, int b, int vx) {  // DirectedEdgeSlot (cut.cuh:25-28)
    int dlt = b - a, ad = dlt < 0 ? -dlt : dlt;
    int axis = ad == 1 ? 0 : (ad == vx + 1 ? 1 : 2);
    return 6 * a + 2 * axis + (dlt > 0 ? 1 : 0);
}

inline bool cutHitsMaterial(float t, float skinT, int insideA, int insideB) {  // CutHitsMaterial (cut.cuh:17-22)
    if (insideA && insideB) return true;
    if (!insideA && !insideB) return false;
    if (!(skinT >= 0.f && skinT <= 1.f)) return false;
    return insideA ? t <= skinT : t >= skinT;
}

// MarkAndEmit + EmitCutPoint (cut.cu:186-254), with the owners' particle frames of the last computeRotations.
template <class R> void markAndEmit(Tissue<R>& s, int e, int axis, const int bc[3], int idA, int idB, R t, V3<R> hit,
                                    V3<R> ncut, R toolD) {
    bool firstCut = false, sawAny = false, occupiedEdge = false;
    for (int k = 0; k < 4; k++) {
        int x = bc[0] + ES_OFF[axis][k][0], y = bc[1] + ES_OFF[axis][k][1], z = bc[2] + ES_OFF[axis][k][2];
        if (!voxIn(s, x, y, z)) continue;
        int vid = voxId(s, x, y, z);
        if (s.occupied[vid] == 0u) continue;
        occupiedEdge = true;
        uint32_t bit = 1u << ES_EDGE[axis][k];
        uint32_t old = s.cutMask[vid];
        s.cutMask[vid] = old | bit;
        if (!sawAny) { sawAny = true; firstCut = (old & bit) == 0u; }
    }
    if (!(firstCut && occupiedEdge)) return;
    s.edgeCutT[directedSlot(idA, idB, s.d.vx)] = (float)t;
    s.edgeCutT[directedSlot(idB, idA, s.d.vx)] = (float)((R)1 - t);
    if (!cutHitsMaterial((float)t, s.edgeSkinT[directedSlot(idA, idB, s.d.vx)], s.inside[idA], s.inside[idB])) return;
    // rest edge from the corner rest positions (materialRestDevice = recon_corner_rest, cut.cu:248, :1181-1183),
    // evaluated in float as the reference's float3 buffer holds them
    V3<float> ra = mk<float>((float)s.rest[idA].x, (float)s.rest[idA].y, (float)s.rest[idA].z);
    V3<float> rb = mk<float>((float)s.rest[idB].x, (float)s.rest[idB].y, (float)s.rest[idB].z);
    V3<float> rf = rb - ra;
    V3<R> restEdge = mk<R>((R)rf.x, (R)rf.y, (R)rf.z);
    V3<R> along[2] = {restEdge * t, restEdge * (t - (R)1)};
    int owners[2] = {idA, idB};
    for (int k = 0; k < 2; k++) {
        int p = owners[k];
        R side = dot3(s.pos[p] - hit, ncut);
        side = side > 0 ? (R)1 : (side < 0 ? (R)-1 : (R)0);
        if (side == 0) side = 1;
        uint32_t slot = s.cutPointCounter++;
        if (slot >= (uint32_t)s.d.cut_point_capacity) continue;
        // the gap and the normal in the owner's material frame, R^T applied (EmitCutPoint, cut.cu:201-208)
        V3<R> gap = mulRtv(&s.rot[3 * p], ncut * (side * (toolD * (R)0.5f)));
        V3<R> off = along[k] + gap;
        V3<R> nrm = mulRtv(&s.rot[3 * p], ncut * side);
        CutPoint cp;
        cp.off[0] = (float)off.x; cp.off[1] = (float)off.y; cp.off[2] = (float)off.z;
        cp.owner = p;
        cp.edge = e;
        cp.nrm[0] = (float)nrm.x; cp.nrm[1] = (float)nrm.y; cp.nrm[2] = (float)nrm.z;
        s.cutPoints.push_back(cp);
    }
}

// SeverLink (cut.cu:719-731).
template <class R> void severLink(Tissue<R>& s, int a, int b, int axis) {
    if (s.nbr[6 * a + 2 * axis + 1] >= 0) s.revision++;
    s.nbr[6 * a + 2 * axis + 1] = -1;
    s.nbr[6 * b + 2 * axis] = -1;
    for (int p = 0; p < 15; p++) {
        int ia = 15 * a + p;
        if (s.bendAlive[ia] != 0 && (s.bendJ[ia] == b || s.bendK[ia] == b)) s.bendAlive[ia] = 0;
        int ib = 15 * b + p;
        if (s.bendAlive[ib] != 0 && (s.bendJ[ib] == a || s.bendK[ib] == a)) s.bendAlive[ib] = 0;
    }
}

template <class R> void severLinks(Tissue<R>& s) {  // k_SeverLinks (cut.cu:732-746), edges in list order
    int Eg = s.d.grid_edge_count;
    for (int e = 0; e < Eg; e++) {
        int axis = s.gridEdges[4 * e], bc[3] = {s.gridEdges[4 * e + 1], s.gridEdges[4 * e + 2], s.gridEdges[4 * e + 3]};
        bool valid = true, cut = false;
        for (int k = 0; k < 4; k++) {
            int x = bc[0] + ES_OFF[axis][k][0], y = bc[1] + ES_OFF[axis][k][1], z = bc[2] + ES_OFF[axis][k][2];
            if (!voxIn(s, x, y, z)) { valid = false; continue; }
            if ((s.cutMask[voxId(s, x, y, z)] >> ES_EDGE[axis][k]) & 1u) cut = true;
        }
        if (!valid || !cut) continue;
        int a = s.cid(bc[0], bc[1], bc[2]);
        int b = s.cid(bc[0] + (axis == 0), bc[1] + (axis == 1), bc[2] + (axis == 2));
        severLink(s, a, b, axis);
    }
}

// k_TearOverstretch (cut.cu:607-681), then k_SeverLinks, as cut_ribbon_tick runs them (cut.cu:1397-1405): an intact
// link between active corners, not yet marked and with an occupied stencil voxel, that is longer than tau times its
// rest length is marked at its midpoint with the plane normal to it. tau < 1 or NaN disables the law (cut.cu:1199-1202).
// A link with a grasp-ring corner is exempt, and a released one stays exempt while it decays (above). Each edge reads
// and sets only its own mask bits and exemption state, so the serial edge order decides nothing but the cut point
// slots.
// cut_recompute_grasp_ring (cut.cu:1513-1528), after every tool update (physics_set_tool, physics.cu:436-447): the
// captured corners, dilated by exempt_rings hops over live links (k_DilateRingHop), minus every corner that has lost a
// link since the lattice was built (k_CarveRimBridge): a cut-rim corner keeps the tear law.
template <class R> void recomputeGraspRing(Tissue<R>& s) {
    const int hops = s.gp.exempt_rings > 0 ? s.gp.exempt_rings : 0;
    std::vector<int> ring(s.N, 0), next(s.N, 0);
    if ((int)s.grasped.size() == s.N) ring = s.grasped;
    for (int hop = 0; hop < hops; hop++) {
        for (int c = 0; c < s.N; c++) {
            int v = ring[c] != 0 ? 1 : 0;
            for (int k = 0; k < 6 && !v; k++) {
                int n = s.nbr[6 * c + k];
                if (n >= 0 && ring[n] != 0) v = 1;
            }
            next[c] = v;
        }
        ring.swap(next);
    }
    for (int c = 0; c < s.N; c++) {
        if (ring[c] == 0) continue;
        for (int k = 0; k < 6; k++)
            if (s.nbr0[6 * c + k] >= 0 && s.nbr[6 * c + k] < 0) { ring[c] = 0; break; }
    }
    s.graspRing = ring;
}

// The tear law's exemption (cut.cu:647-673): an edge with a ring corner is exempt and (re)starts its release
// decay; once out of the ring it stays exempt while longer than 0.8 tau L, for up to 2000 ticks.
constexpr int TEAR_DECAY_TICKS = 2000;
template <class R> void tearOverstretch(Tissue<R>& s, float tau, R toolD, bool sever = true) {
    if (!(tau >= 1.0f)) return;
    float voxelL = s.d.tool_voxel;
    float tearLen2 = (tau * voxelL) * (tau * voxelL);
    int Eg = s.d.grid_edge_count;
    for (int e = 0; e < Eg; e++) {
        int axis = s.gridEdges[4 * e], bc[3] = {s.gridEdges[4 * e + 1], s.gridEdges[4 * e + 2], s.gridEdges[4 * e + 3]};
        int idA = s.cid(bc[0], bc[1], bc[2]);
        int idB = s.cid(bc[0] + (axis == 0), bc[1] + (axis == 1), bc[2] + (axis == 2));
        if (s.nbr[6 * idA + 2 * axis + 1] < 0) continue;
        if (s.active[idA] == 0 || s.active[idB] == 0) continue;
        bool cut = false, anyOcc = false;
        for (int k = 0; k < 4; k++) {
            int x = bc[0] + ES_OFF[axis][k][0], y = bc[1] + ES_OFF[axis][k][1], z = bc[2] + ES_OFF[axis][k][2];
            if (!voxIn(s, x, y, z)) continue;
            int vid = voxId(s, x, y, z);
            if ((s.cutMask[vid] >> ES_EDGE[axis][k]) & 1u) cut = true;
            if (s.occupied[vid] != 0u) anyOcc = true;
        }
        if (cut || !anyOcc) continue;
        V3<R> A = s.pos[idA], d = s.pos[idB] - A;
        R len2 = dot3(d, d);
        V3<float> ra = mk<float>((float)s.rest[idA].x, (float)s.rest[idA].y, (float)s.rest[idA].z);
        V3<float> rb = mk<float>((float)s.rest[idB].x, (float)s.rest[idB].y, (float)s.rest[idB].z);
        V3<float> r = rb - ra;
        float limit = tearLen2 * (dot3(r, r) / (voxelL * voxelL));
        if (s.graspRing[idA] != 0 || s.graspRing[idB] != 0) {
            s.edgeWasExempt[e] = 1;
            s.edgeDecayTicks[e] = 0;
            continue;
        }
        if (s.edgeWasExempt[e] != 0) {
            float decayLen2 = 0.64f * limit;
            if (len2 > (R)decayLen2) {
                int ticks = s.edgeDecayTicks[e] + 1;
                s.edgeDecayTicks[e] = ticks;
                if (ticks < TEAR_DECAY_TICKS) continue;
            }
            s.edgeWasExempt[e] = 0;
            s.edgeDecayTicks[e] = 0;
        }
        if (len2 <= (R)limit) continue;
        R len = std::sqrt(len2);
        V3<R> ncut = mk<R>(d.x / len, d.y / len, d.z / len);
        V3<R> hit = A + d * (R)0.5f;
        s.tearCount++;
        markAndEmit(s, e, axis, bc, idA, idB, (R)0.5f, hit, ncut, toolD);
    }
    if (sever) severLinks(s);
}

// k_DetectCutQuad (cut.cu:421-474) + k_SeverLinks: cut_detect (cut.cu:1235-1254).
template <class R> void detectCutQuad(Tissue<R>& s, const OracleCutTool& tl) {
    if (!tl.valid) return;
    V3<R> t1v0 = v3<R>(tl.t1v0), t1v1 = v3<R>(tl.t1v1), t1v2 = v3<R>(tl.t1v2);
    V3<R> t2v0 = v3<R>(tl.t2v0), t2v1 = v3<R>(tl.t2v1), t2v2 = v3<R>(tl.t2v2);
    V3<R> ncut = v3<R>(tl.ncut), lo = v3<R>(tl.aabbMin), hi = v3<R>(tl.aabbMax);
    R pad = (R)s.d.tool_voxel;
    int Eg = s.d.grid_edge_count;
    for (int e = 0; e < Eg; e++) {
        int axis = s.gridEdges[4 * e], bc[3] = {s.gridEdges[4 * e + 1], s.gridEdges[4 * e + 2], s.gridEdges[4 * e + 3]};
        int idA = s.cid(bc[0], bc[1], bc[2]);
        int idB = s.cid(bc[0] + (axis == 0), bc[1] + (axis == 1), bc[2] + (axis == 2));
        V3<R> O = s.pos[idA], D = s.pos[idB] - O;
        V3<R> eMin = mk<R>(fmin_<R>(O.x, O.x + D.x), fmin_<R>(O.y, O.y + D.y), fmin_<R>(O.z, O.z + D.z));
        V3<R> eMax = mk<R>(fmax_<R>(O.x, O.x + D.x), fmax_<R>(O.y, O.y + D.y), fmax_<R>(O.z, O.z + D.z));
        if (eMax.x < lo.x - pad || eMin.x > hi.x + pad || eMax.y < lo.y - pad || eMin.y > hi.y + pad ||
            eMax.z < lo.z - pad || eMin.z > hi.z + pad)
            continue;
        R ta, ua, va, tb, ub, vb;
        bool h1 = mollerTrumbore(O, D, t1v0, t1v1, t1v2, ta, ua, va);
        bool h2 = mollerTrumbore(O, D, t2v0, t2v1, t2v2, tb, ub, vb);
        if (!h1 && !h2) continue;
        R t = (h1 && h2) ? fmin_<R>(ta, tb) : (h1 ? ta : tb);
        t = fmin_<R>(fmax_<R>(t, (R)0), (R)1);
        V3<R> hit = O + D * t;
        markAndEmit(s, e, axis, bc, idA, idB, t, hit, ncut, (R)tl.d);
    }
    severLinks(s);
}

// k_DetectCutCCD (cut.cu:486-592) of the moving edges against a static blade segment S->E, then k_SeverLinks.
template <class R> void detectCutCCD(Tissue<R>& s, V3<R> S, V3<R> E, V3<R> emitNcut, R toolD,
                                     const std::vector<V3<R>>& prevPos, bool sever = true) {
    R pad = (R)s.d.tool_voxel;
    int Eg = s.d.grid_edge_count;
    for (int e = 0; e < Eg; e++) {
        int axis = s.gridEdges[4 * e], bc[3] = {s.gridEdges[4 * e + 1], s.gridEdges[4 * e + 2], s.gridEdges[4 * e + 3]};
        int idA = s.cid(bc[0], bc[1], bc[2]);
        int idB = s.cid(bc[0] + (axis == 0), bc[1] + (axis == 1), bc[2] + (axis == 2));
        if (s.nbr[6 * idA + 2 * axis + 1] < 0) continue;
        V3<R> a0 = prevPos[idA], a1 = s.pos[idA], b0 = prevPos[idB], b1 = s.pos[idB];
        V3<R> sMin = mk<R>(fmin_<R>(S.x, E.x) - pad, fmin_<R>(S.y, E.y) - pad, fmin_<R>(S.z, E.z) - pad);
        V3<R> sMax = mk<R>(fmax_<R>(S.x, E.x) + pad, fmax_<R>(S.y, E.y) + pad, fmax_<R>(S.z, E.z) + pad);
        V3<R> eMin = mk<R>(fmin_<R>(fmin_<R>(a0.x, a1.x), fmin_<R>(b0.x, b1.x)),
                           fmin_<R>(fmin_<R>(a0.y, a1.y), fmin_<R>(b0.y, b1.y)),
                           fmin_<R>(fmin_<R>(a0.z, a1.z), fmin_<R>(b0.z, b1.z)));
        V3<R> eMax = mk<R>(fmax_<R>(fmax_<R>(a0.x, a1.x), fmax_<R>(b0.x, b1.x)),
                           fmax_<R>(fmax_<R>(a0.y, a1.y), fmax_<R>(b0.y, b1.y)),
                           fmax_<R>(fmax_<R>(a0.z, a1.z), fmax_<R>(b0.z, b1.z)));
        if (eMax.x < sMin.x || eMin.x > sMax.x || eMax.y < sMin.y || eMin.y > sMax.y || eMax.z < sMin.z ||
            eMin.z > sMax.z)
            continue;
        V3<R> D = E - S, P0 = a0 - S, Pv = a1 - a0, Q0 = b0 - S, Qv = b1 - b0;
        R c0 = dot3(cross3(P0, Q0), D);
        R c1 = dot3(cross3(P0, Qv) + cross3(Pv, Q0), D);
        R c2 = dot3(cross3(Pv, Qv), D);
        R roots[2];
        int nr = 0;
        if (std::fabs(c2) < (R)1e-14f) {
            if (std::fabs(c1) > (R)1e-14f) {
                R r = -c0 / c1;
                if (r >= 0 && r <= 1) roots[nr++] = r;
            }
        } else {
            R disc = c1 * c1 - (R)4 * c2 * c0;
            if (disc >= 0) {
                R sq = std::sqrt(disc);
                R q = (R)-0.5f * (c1 + (c1 >= 0 ? sq : -sq));
                R r0 = q / c2, r1 = std::fabs(q) > (R)1e-20f ? c0 / q : r0;
                if (r0 > r1) std::swap(r0, r1);
                if (r0 >= 0 && r0 <= 1) roots[nr++] = r0;
                if (r1 >= 0 && r1 <= 1 && r1 != r0) roots[nr++] = r1;
            }
        }
        for (int ri = 0; ri < nr; ri++) {
            R tc = roots[ri];
            V3<R> a = a0 + Pv * tc, b = b0 + Qv * tc, e1 = b - a, n2 = cross3(e1, D);
            R n2l = dot3(n2, n2);
            if (n2l < (R)1e-16f) continue;
            V3<R> w = S - a;
            R u = dot3(cross3(w, D), n2) / n2l;
            R v = dot3(cross3(w, e1), n2) / n2l;
            const R EPS = (R)2e-2f;
            if (u <= 0 || u > 1 || v < -EPS || v > (R)1 + EPS) continue;
            V3<R> hit = a1 + (b1 - a1) * u;
            V3<R> relV = (Pv + Qv) * (R)0.5f;
            V3<R> nHit = cross3(D, relV);
            R nLen = len3(nHit);
            V3<R> nEmit = nLen > (R)1e-8f ? nHit * ((R)1 / nLen) : emitNcut;
            markAndEmit(s, e, axis, bc, idA, idB, u, hit, nEmit, toolD);
            break;
        }
    }
    if (sever) severLinks(s);
}

// cut_ribbon_tick without the pressure blade (cut.cu:1303-1431): the first tick after the tool was set only takes the
// corner snapshot; every later one runs the CCD of the edges moving from the snapshot to the current positions against
// the tool's current rod S = t2v2, E = t1v2, then the tear law, one sever, and rolls the snapshot.
template <class R> void ribbonTick(Tissue<R>& s) {
    if (!s.havePrevCorner) {
        s.prevCorner = s.pos;
        s.havePrevCorner = true;
        return;
    }
    V3<R> S = v3<R>(s.cutTool.t2v2), E = v3<R>(s.cutTool.t1v2);
    detectCutCCD(s, S, E, s.emitNcut, (R)s.cutTool.d, s.prevCorner, false);
    tearOverstretch(s, s.tearTau, (R)s.cutTool.d, false);
    severLinks(s);
    s.prevCorner = s.pos;
}

// cut_set_tool (cut.cu:1208-1226): the tool, the last valid cut normal, and the corner snapshot on activation.
template <class R> void setCutTool(Tissue<R>& s, const OracleCutTool& t) {
    s.cutTool = t;
    if (t.valid != 0) s.emitNcut = v3<R>(t.ncut);
    if (!s.cutToolSet) {
        s.prevCorner = s.pos;
        s.havePrevCorner = true;
    }
    s.cutToolSet = true;
}

// Stable compaction of the live topology: old -> new index, -1 for removed (build side, no reference counterpart).
// Tets are numbered in colour-major order, edges in colour order, nodes in id order; a node survives when it is
// active and some live tet or live edge references it.
template <class R> void compact(Tissue<R>& s, int* tetRemap, int* edgeRemap, int* nodeRemap, int* counts) {
    refreshLiveness(s);
    std::vector<int> used(s.N, 0);
    int nt = 0;
    for (int k = 0; k < s.T; k++) {
        int t = s.tcolorOrder[k];
        if (s.tactive[t]) {
            tetRemap[k] = nt++;
            for (int j = 0; j < 4; j++) used[s.tids[4 * t + j]] = 1;
        } else tetRemap[k] = -1;
    }
    int ne = 0;
    for (int e = 0; e < s.E; e++) {
        if (s.eactive[e]) {
            edgeRemap[e] = ne++;
            used[s.ei[e]] = used[s.ej[e]] = 1;
        } else edgeRemap[e] = -1;
    }
    int nn = 0;
    for (int p = 0; p < s.N; p++) nodeRemap[p] = (used[p] && s.active[p]) ? nn++ : -1;
    counts[0] = nt; counts[1] = ne; counts[2] = nn;
}

template <class R> Tissue<R>* create(const OracleDesc* d, const float* rest, const float* pos, const float* mass,
                                     const int* pinned, const int* active, const int* inside, const int* nbr,
                                     const int* occupied, const int* gridEdges, const float* edgeSkin,
                                     const float* restNbr) {
    auto* s = new Tissue<R>();
    s->d = *d;
    s->nx = d->vx + 1; s->ny = d->vy + 1; s->nz = d->vz + 1;
    s->N = s->nx * s->ny * s->nz;
    s->V = d->vx * d->vy * d->vz;
    int N = s->N;
    s->pos.resize(N); s->prev.resize(N); s->vel.resize(N); s->ext.resize(N); s->rest.resize(N);
    s->rot.resize(3 * (size_t)N); s->rotPrev.resize(3 * (size_t)N);
    for (int i = 0; i < N; i++) {  // identity frames (physics.cu:260-268)
        s->rot[3 * i] = mk<R>(1, 0, 0); s->rot[3 * i + 1] = mk<R>(0, 1, 0); s->rot[3 * i + 2] = mk<R>(0, 0, 1);
    }
    s->rotPrev = s->rot;
    s->mass.resize(N); s->invMass.resize(N);
    for (int i = 0; i < N; i++) {
        s->rest[i] = v3<R>(rest + 3 * i);
        s->pos[i] = v3<R>(pos + 3 * i);
        s->prev[i] = s->pos[i];
        s->vel[i] = mk<R>(0, 0, 0);
        s->ext[i] = mk<R>(0, 0, 0);
        s->mass[i] = (R)mass[i];
    }
    s->active.assign(active, active + N);
    s->pinned.assign(pinned, pinned + N);
    s->inside.assign(inside, inside + N);
    s->nbr.assign(nbr, nbr + 6 * N);
    s->nbr0 = s->nbr;
    for (int i = 0; i < N; i++)  // invMass rule of k_RefreshLiveness; active/pinned never change on this path
        s->invMass[i] = (s->pinned[i] != 0 || s->active[i] == 0) ? (R)0 : (R)1 / s->mass[i];
    s->bendJ.resize(15 * N); s->bendK.resize(15 * N); s->bendAlive.resize(15 * N);
    for (int i = 0; i < N; i++)
        for (int p = 0; p < 15; p++) {
            int j = s->nbr[6 * i + BEND_PAIRS[p][0]], k = s->nbr[6 * i + BEND_PAIRS[p][1]];
            s->bendJ[15 * i + p] = j;
            s->bendK[15 * i + p] = k;
            s->bendAlive[15 * i + p] = (j >= 0 && k >= 0) ? 1 : 0;
        }
    s->gridEdges.assign(gridEdges, gridEdges + 4 * d->grid_edge_count);
    s->occupied.assign(occupied, occupied + s->V);
    s->cutMask.assign(s->V, 0u);
    s->edgeCutT.assign(6 * N, -1.f);
    s->edgeSkinT.assign(edgeSkin, edgeSkin + 6 * N);
    s->graspRing.assign(N, 0);
    s->edgeWasExempt.assign(d->grid_edge_count, 0);
    s->edgeDecayTicks.assign(d->grid_edge_count, 0);
    s->status = buildTopology(*s);
    for (int e = 0; e < s->E; e++) {  // L0 = sqrtf(|restNbr[6p+s]|^2) in float (xpbd_solver.cu:1458)
        const float* r = restNbr + 3 * s->eowner[e];
        s->erest[e] = (R)std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    }
    // rest edges of the Kuhn types for the precise invariants: bit offsets times the rest length of each axis (the
    // first edge's L0 along it, as the library's topology reports axis_rest_length; the lattice spacing if none)
    float axisLen[3] = {(float)(((double)s->rest[s->cid(s->d.vx, 0, 0)].x - (double)s->rest[0].x) / s->d.vx),
                        (float)(((double)s->rest[s->cid(0, s->d.vy, 0)].y - (double)s->rest[0].y) / s->d.vy),
                        (float)(((double)s->rest[s->cid(0, 0, s->d.vz)].z - (double)s->rest[0].z) / s->d.vz)};
    bool seen[3] = {false, false, false};
    for (int e = 0; e < s->E; e++) {
        int axis = (s->eowner[e] % 6) / 2;
        if (!seen[axis]) { axisLen[axis] = (float)s->erest[e]; seen[axis] = true; }
    }
    for (int ti = 0; ti < 6; ti++)
        for (int k = 0; k < 3; k++) {
            int c = KUHN[ti][k + 1];
            s->typeRk[ti][k] = mk<R>((c & 1) ? (R)axisLen[0] : (R)0, ((c >> 1) & 1) ? (R)axisLen[1] : (R)0,
                                     ((c >> 2) & 1) ? (R)axisLen[2] : (R)0);
        }
    return s;
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// C entry points (ctypes). The float handle is the parity oracle and the timed CPU baseline, the double handle the
// accuracy oracle. Inputs are float arrays as the product receives them.
// ---------------------------------------------------------------------------------------------------------------
#define ORACLE_API(R, SUF)                                                                                           \
    extern "C" void* oracle_create_##SUF(const OracleDesc* d, const float* rest, const float* pos, const float* mass, \
                                         const int* pinned, const int* active, const int* inside, const int* nbr,     \
                                         const int* occupied, const int* gridEdges, const float* edgeSkin,            \
                                         const float* restNbr) {                                                      \
        return create<R>(d, rest, pos, mass, pinned, active, inside, nbr, occupied, gridEdges, edgeSkin, restNbr);    \
    }                                                                                                                 \
    extern "C" int oracle_status_##SUF(void* h) { return ((Tissue<R>*)h)->status; }                                   \
    extern "C" void oracle_destroy_##SUF(void* h) { delete (Tissue<R>*)h; }                                           \
    extern "C" void oracle_counts_##SUF(void* h, int* out) {                                                          \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        out[0] = s->N; out[1] = s->E; out[2] = s->T;                                                                  \
        out[3] = (int)s->ecolorCnt.size(); out[4] = (int)s->tcolorCnt.size();                                         \
    }                                                                                                                 \
    /* bend elements: count, colours, then per colour slot the pair index 15 i + p, and the colour counts */          \
    extern "C" void oracle_bends_##SUF(void* h, int* counts, int* src, int* colorCnt) {                               \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        counts[0] = (int)s->bsrc.size(); counts[1] = (int)s->bcolorCnt.size();                                        \
        if (src) std::memcpy(src, s->bsrc.data(), sizeof(int) * s->bsrc.size());                                      \
        if (colorCnt)                                                                                                 \
            for (size_t c = 0; c < s->bcolorCnt.size(); c++) colorCnt[c] = s->bcolorCnt[c];                           \
    }                                                                                                                 \
    extern "C" void oracle_topology_##SUF(void* h, int* ei, int* ej, int* eowner, int* ecolorCnt, int* tids,          \
                                          int* ttype, int* towner, int* tcolor, int* tcolorCnt) {                     \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        for (int e = 0; e < s->E; e++) { ei[e] = s->ei[e]; ej[e] = s->ej[e]; eowner[e] = s->eowner[e]; }              \
        for (size_t c = 0; c < s->ecolorCnt.size(); c++) ecolorCnt[c] = s->ecolorCnt[c];                              \
        std::memcpy(tids, s->tids.data(), sizeof(int) * 4 * s->T);                                                    \
        std::memcpy(ttype, s->ttype.data(), sizeof(int) * s->T);                                                      \
        std::memcpy(towner, s->towner.data(), sizeof(int) * 3 * s->T);                                                \
        std::memcpy(tcolor, s->tcolor.data(), sizeof(int) * s->T);                                                    \
        for (size_t c = 0; c < s->tcolorCnt.size(); c++) tcolorCnt[c] = s->tcolorCnt[c];                             \
    }                                                                                                                 \
    extern "C" void oracle_type_data_##SUF(void* h, double* typeB, double* typeV) {                                   \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        for (int t = 0; t < 6; t++) {                                                                                 \
            typeV[t] = s->typeV[t];                                                                                   \
            for (int a = 0; a < 9; a++) typeB[9 * t + a] = (double)s->typeB[t][a / 3][a % 3];                         \
        }                                                                                                             \
    }                                                                                                                 \
    extern "C" void oracle_set_tet_stiffness_##SUF(void* h, const float* kDev, const float* kHyd) {                   \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        for (int t = 0; t < s->T; t++) { s->kDev[t] = (R)kDev[t]; s->kHyd[t] = (R)kHyd[t]; }                          \
    }                                                                                                                 \
    extern "C" void oracle_set_gravity_##SUF(void* h, float gx, float gy, float gz) {                                 \
        Tissue<R>* s = (Tissue<R>*)h; /* k_SetGravity: force = acceleration * mass (physics.cu:317-329) */            \
        for (int i = 0; i < s->N; i++) s->ext[i] = mk<R>((R)gx, (R)gy, (R)gz) * s->mass[i];                            \
    }                                                                                                                 \
    extern "C" int oracle_set_tool_##SUF(void* h, const OracleTool* t, int toolArmed, int bladeArmed, int reset) {    \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        s->tool = toolFrom<R>(*t);                                                                                    \
        if (reset) s->toolFrom = s->tool;                                                                             \
        s->toolArmed = toolArmed != 0; s->bladeArmed = bladeArmed != 0;                                               \
        int event = s->contactModel == 0 ? graspSetTool(*s) : 0;                                                      \
        recomputeGraspRing(*s);                                                                                       \
        return event;                                                                                                 \
    }                                                                                                                 \
    extern "C" void oracle_set_contact_model_##SUF(void* h, int model, const OracleGrasp* gp) {                       \
        setContactModel(*(Tissue<R>*)h, model, *gp);                                                                  \
    }                                                                                                                 \
    extern "C" int oracle_get_grasp_##SUF(void* h, int* grasped, int* core, R* ema, R* anchor, R* slabPrev) {         \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        if ((int)s->grasped.size() != s->N) return -1;                                                                \
        for (int i = 0; i < s->N; i++) {                                                                              \
            grasped[i] = s->grasped[i]; core[i] = s->core[i]; ema[i] = s->ema[i]; slabPrev[i] = s->slabPrev[i];       \
            anchor[2 * i] = s->anchU[i]; anchor[2 * i + 1] = s->anchW[i];                                             \
        }                                                                                                             \
        return s->grasping ? 1 : 0;                                                                                   \
    }                                                                                                                 \
    extern "C" int oracle_step_##SUF(void* h, float dt) { return step(*(Tissue<R>*)h, (R)dt); }                       \
    extern "C" void oracle_detect_cut_##SUF(void* h, const OracleCutTool* t) { detectCutQuad(*(Tissue<R>*)h, *t); }   \
    extern "C" void oracle_detect_cut_ccd_##SUF(void* h, const float* S, const float* E, const float* ncut,           \
                                                float toolD, const float* prevPos) {                                  \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        std::vector<V3<R>> pp(s->N);                                                                                  \
        for (int i = 0; i < s->N; i++) pp[i] = v3<R>(prevPos + 3 * i);                                                \
        detectCutCCD(*s, v3<R>(S), v3<R>(E), v3<R>(ncut), (R)toolD, pp);                                              \
    }                                                                                                                 \
    extern "C" void oracle_tear_##SUF(void* h, float tau, float toolD) { tearOverstretch(*(Tissue<R>*)h, tau, (R)toolD); } \
    extern "C" void oracle_set_cut_tool_##SUF(void* h, const OracleCutTool* t, float tearTau) {                      \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        if (!t) { s->cutToolSet = false; s->havePrevCorner = false; return; }                                          \
        setCutTool(*s, *t);                                                                                           \
        s->tearTau = tearTau;                                                                                         \
    }                                                                                                                 \
    extern "C" void oracle_compute_rotations_##SUF(void* h) { computeRotations(*(Tissue<R>*)h); }                    \
    extern "C" void oracle_get_rotations_##SUF(void* h, R* out) {                                                     \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        for (size_t i = 0; i < s->rot.size(); i++) {                                                                  \
            out[3 * i] = s->rot[i].x; out[3 * i + 1] = s->rot[i].y; out[3 * i + 2] = s->rot[i].z;                     \
        }                                                                                                             \
    }                                                                                                                 \
    extern "C" void oracle_refresh_##SUF(void* h) { refreshLiveness(*(Tissue<R>*)h); }                                \
    extern "C" void oracle_get_state_##SUF(void* h, R* pos, R* vel) {                                                 \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        for (int i = 0; i < s->N; i++) {                                                                              \
            pos[3 * i] = s->pos[i].x; pos[3 * i + 1] = s->pos[i].y; pos[3 * i + 2] = s->pos[i].z;                     \
            vel[3 * i] = s->vel[i].x; vel[3 * i + 1] = s->vel[i].y; vel[3 * i + 2] = s->vel[i].z;                     \
        }                                                                                                             \
    }                                                                                                                 \
    extern "C" void oracle_set_state_##SUF(void* h, const float* pos, const float* vel) {                             \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        for (int i = 0; i < s->N; i++) { s->pos[i] = v3<R>(pos + 3 * i); s->vel[i] = v3<R>(vel + 3 * i); }             \
    }                                                                                                                 \
    extern "C" void oracle_get_reaction_##SUF(void* h, double* out) {                                                 \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        out[0] = s->reaction[0]; out[1] = s->reaction[1]; out[2] = s->reaction[2]; out[3] = s->reactionContacts;      \
    }                                                                                                                 \
    extern "C" unsigned oracle_get_cut_##SUF(void* h, int* nbr, unsigned* cutMask, float* edgeCutT,                   \
                                             int* bendAlive, unsigned* counters) {                                    \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        std::memcpy(nbr, s->nbr.data(), sizeof(int) * 6 * s->N);                                                      \
        std::memcpy(cutMask, s->cutMask.data(), sizeof(unsigned) * s->V);                                             \
        std::memcpy(edgeCutT, s->edgeCutT.data(), sizeof(float) * 6 * s->N);                                          \
        std::memcpy(bendAlive, s->bendAlive.data(), sizeof(int) * 15 * s->N);                                         \
        counters[0] = s->revision; counters[1] = s->cutPointCounter; counters[2] = s->tearCount;                      \
        return (unsigned)s->cutPoints.size();                                                                         \
    }                                                                                                                 \
    extern "C" void oracle_get_cut_points_##SUF(void* h, void* out) {                                                 \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        std::memcpy(out, s->cutPoints.data(), sizeof(CutPoint) * s->cutPoints.size());                                \
    }                                                                                                                 \
    extern "C" void oracle_get_liveness_##SUF(void* h, int* tetActiveGen, int* edgeActive) {                          \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        refreshLiveness(*s);                                                                                          \
        std::memcpy(tetActiveGen, s->tactive.data(), sizeof(int) * s->T);                                             \
        std::memcpy(edgeActive, s->eactive.data(), sizeof(int) * s->E);                                               \
    }                                                                                                                 \
    /* the tear law's grasp ring [N] and per grid edge exemption flag and decay ticks [Eg]; null arrays are skipped */ \
    extern "C" void oracle_set_tear_exempt_##SUF(void* h, const int* ring, const int* was, const int* ticks) {       \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        if (ring) s->graspRing.assign(ring, ring + s->N);                                                             \
        if (was) s->edgeWasExempt.assign(was, was + s->d.grid_edge_count);                                            \
        if (ticks) s->edgeDecayTicks.assign(ticks, ticks + s->d.grid_edge_count);                                     \
    }                                                                                                                 \
    extern "C" void oracle_get_tear_exempt_##SUF(void* h, int* ring, int* was, int* ticks) {                         \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        if (ring) std::copy(s->graspRing.begin(), s->graspRing.end(), ring);                                          \
        if (was) std::copy(s->edgeWasExempt.begin(), s->edgeWasExempt.end(), was);                                    \
        if (ticks) std::copy(s->edgeDecayTicks.begin(), s->edgeDecayTicks.end(), ticks);                              \
    }                                                                                                                 \
    /* the nodes the last minimum-volume safeguard sent through the feasible filter; returns their count */         \
    extern "C" int oracle_get_feasible_##SUF(void* h, int* needed) {                                                  \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        int n = 0;                                                                                                    \
        for (int i = 0; i < s->N; i++) { if (needed) needed[i] = s->needed[i]; n += s->needed[i] ? 1 : 0; }           \
        return n;                                                                                                     \
    }                                                                                                                 \
    /* the edge multipliers [E] and the tets' deviatoric and hydrostatic multipliers [T] after the last substep */     \
    extern "C" void oracle_get_multipliers_##SUF(void* h, R* edge, R* dev, R* hyd) {                                  \
        Tissue<R>* s = (Tissue<R>*)h;                                                                                 \
        std::memcpy(edge, s->elambda.data(), sizeof(R) * s->E);                                                       \
        std::memcpy(dev, s->lamDev.data(), sizeof(R) * s->T);                                                         \
        std::memcpy(hyd, s->lamHyd.data(), sizeof(R) * s->T);                                                         \
    }                                                                                                                 \
    extern "C" void oracle_compact_##SUF(void* h, int* tetRemap, int* edgeRemap, int* nodeRemap, int* counts) {       \
        compact(*(Tissue<R>*)h, tetRemap, edgeRemap, nodeRemap, counts);                                              \
    }

ORACLE_API(float, f32)
ORACLE_API(double, f64)

// One outer step of each of n independent instances, the instances spread over the threads (the colour loops inside
// each step then run on one thread): the CPU counterpart of the batched launch.
#define ORACLE_MANY(R, SUF)                                                                                           \
    extern "C" int oracle_step_many_##SUF(void** h, int n, float dt) {                                               \
        int bad = 0;                                                                                                   \
        _Pragma("omp parallel for schedule(dynamic, 1) reduction(|| : bad)") for (int k = 0; k < n; k++) bad =         \
            bad || step(*(Tissue<R>*)h[k], (R)dt) != 0;                                                                \
        return bad ? -1 : 0;                                                                                           \
    }
ORACLE_MANY(float, f32)
ORACLE_MANY(double, f64)

extern "C" int oracle_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

// ---------------------------------------------------------------------------------------------------------------
// Unit hooks for the checker's own tests: single projections and primitives on raw double inputs.
// ---------------------------------------------------------------------------------------------------------------
extern "C" int oracle_tet_project_f64(int mode, const double* p12, const double* prev12, const double* w4,
                                      const double* B9, double ad, double ah, double* lam2, double dampingOverH,
                                      double bulkOverH, double omega, double* out12) {
    V3<double> p[4], pv[4], out[4], br[3];
    double w[4];
    for (int k = 0; k < 4; k++) {
        p[k] = mk(p12[3 * k], p12[3 * k + 1], p12[3 * k + 2]);
        pv[k] = mk(prev12[3 * k], prev12[3 * k + 1], prev12[3 * k + 2]);
        w[k] = w4[k];
    }
    for (int r = 0; r < 3; r++) br[r] = mk(B9[3 * r], B9[3 * r + 1], B9[3 * r + 2]);
    bool ok = tetProject<double>(mode, p, pv, w, br, ad, ah, lam2[0], lam2[1], dampingOverH, bulkOverH, omega, out);
    for (int k = 0; k < 4; k++) { out12[3 * k] = out[k].x; out12[3 * k + 1] = out[k].y; out12[3 * k + 2] = out[k].z; }
    return ok ? 1 : 0;
}
// One edge projection: x6 = (xi, xj) updated, p6 their substep starts, lam updated; returns 0 where it does not write.
extern "C" int oracle_edge_project_f64(double* x6, const double* p6, double wi, double wj, double rest,
                                       double alphaTilde, double gamma, double omega, double* lam) {
    V3<double> xi = mk(x6[0], x6[1], x6[2]), xj = mk(x6[3], x6[4], x6[5]);
    V3<double> pi = mk(p6[0], p6[1], p6[2]), pj = mk(p6[3], p6[4], p6[5]);
    bool ok = edgeProject(xi, xj, pi, pj, wi, wj, rest, alphaTilde, gamma, omega, *lam);
    x6[0] = xi.x; x6[1] = xi.y; x6[2] = xi.z; x6[3] = xj.x; x6[4] = xj.y; x6[5] = xj.z;
    return ok ? 1 : 0;
}
// One bend projection: x9 = (xi, xj, xk) updated, p9 their substep starts, w3 inverse masses, lam updated; returns 0
// where it does not write.
extern "C" int oracle_bend_project_f64(double* x9, const double* p9, const double* w3, double theta0, double alphaB,
                                       double gamma, double omega, double* lam) {
    V3<double> x[3], p[3];
    for (int k = 0; k < 3; k++) {
        x[k] = mk(x9[3 * k], x9[3 * k + 1], x9[3 * k + 2]);
        p[k] = mk(p9[3 * k], p9[3 * k + 1], p9[3 * k + 2]);
    }
    bool ok = bendProject(x[0], x[1], x[2], p[0], p[1], p[2], w3[0], w3[1], w3[2], theta0, alphaB, gamma, omega, *lam);
    for (int k = 0; k < 3; k++) { x9[3 * k] = x[k].x; x9[3 * k + 1] = x[k].y; x9[3 * k + 2] = x[k].z; }
    return ok ? 1 : 0;
}
// The float oracle's arccos of the bend angle.
extern "C" float oracle_acos_f32(float c) { return bendAcos(c); }
// One edge's perpendicular velocity damping: v6 = (vi, vj) updated at positions x6.
extern "C" void oracle_perp_damp_f64(double* v6, const double* x6, double wi, double wj, double hBeta) {
    V3<double> vi = mk(v6[0], v6[1], v6[2]), vj = mk(v6[3], v6[4], v6[5]);
    perpDamp(vi, vj, mk(x6[0], x6[1], x6[2]), mk(x6[3], x6[4], x6[5]), wi, wj, hBeta);
    v6[0] = vi.x; v6[1] = vi.y; v6[2] = vi.z; v6[3] = vj.x; v6[4] = vj.y; v6[5] = vj.z;
}
// One tet's tensile reinforcement: p12 corners (updated), w4 inverse masses, B9 Dm^-1 rows, lam6 fibre multipliers
// (updated).
extern "C" void oracle_tensile_project_f64(double* p12, const double* w4, const double* B9, double alphaDev,
                                           double strength, double* lam6) {
    V3<double> p[4], b[3];
    for (int k = 0; k < 4; k++) p[k] = mk(p12[3 * k], p12[3 * k + 1], p12[3 * k + 2]);
    for (int r = 0; r < 3; r++) b[r] = mk(B9[3 * r], B9[3 * r + 1], B9[3 * r + 2]);
    tensileProject(p, w4, b, alphaDev, strength, lam6);
    for (int k = 0; k < 4; k++) { p12[3 * k] = p[k].x; p12[3 * k + 1] = p[k].y; p12[3 * k + 2] = p[k].z; }
}
// One tet's minimum-volume repair; returns 0 where it does not write.
extern "C" int oracle_repair_project_f64(double* p12, const double* w4, const double* B9, double minimum) {
    V3<double> p[4], b[3];
    for (int k = 0; k < 4; k++) p[k] = mk(p12[3 * k], p12[3 * k + 1], p12[3 * k + 2]);
    for (int r = 0; r < 3; r++) b[r] = mk(B9[3 * r], B9[3 * r + 1], B9[3 * r + 2]);
    bool ok = repairProject(p, w4, b, minimum);
    for (int k = 0; k < 4; k++) { p12[3 * k] = p[k].x; p12[3 * k + 1] = p[k].y; p12[3 * k + 2] = p[k].z; }
    return ok ? 1 : 0;
}
// One tet projection in FP32 (rk9: the type's rest edges for the precise invariants, or null for the direct forms).
extern "C" int oracle_tet_project_f32(int mode, const float* p12, const float* prev12, const float* w4,
                                      const float* B9, const float* rk9, float ad, float ah, float* lam2,
                                      float dampingOverH, float bulkOverH, float omega, float* out12) {
    V3<float> p[4], pv[4], out[4], br[3], rk[3];
    float w[4];
    for (int k = 0; k < 4; k++) {
        p[k] = mk(p12[3 * k], p12[3 * k + 1], p12[3 * k + 2]);
        pv[k] = mk(prev12[3 * k], prev12[3 * k + 1], prev12[3 * k + 2]);
        w[k] = w4[k];
    }
    for (int r = 0; r < 3; r++) br[r] = mk(B9[3 * r], B9[3 * r + 1], B9[3 * r + 2]);
    if (rk9)
        for (int r = 0; r < 3; r++) rk[r] = mk(rk9[3 * r], rk9[3 * r + 1], rk9[3 * r + 2]);
    bool ok = tetProject<float>(mode, p, pv, w, br, ad, ah, lam2[0], lam2[1], dampingOverH, bulkOverH, omega, out,
                                rk9 ? rk : nullptr);
    for (int k = 0; k < 4; k++) { out12[3 * k] = out[k].x; out12[3 * k + 1] = out[k].y; out12[3 * k + 2] = out[k].z; }
    return ok ? 1 : 0;
}
// The precise invariants of one tet in FP32: out = [J - 1, I1 - 3, q, Ghat columns (9)], the tet's corners p12, its
// Dm^-1 rows B9 and rest edges rk9.
extern "C" void oracle_precise_invariants_f32(const float* p12, const float* B9, const float* rk9, float* out) {
    V3<float> p[4], br[3], d[3], F[3];
    for (int k = 0; k < 4; k++) p[k] = mk(p12[3 * k], p12[3 * k + 1], p12[3 * k + 2]);
    for (int r = 0; r < 3; r++) br[r] = mk(B9[3 * r], B9[3 * r + 1], B9[3 * r + 2]);
    for (int k = 0; k < 3; k++) d[k] = (p[k + 1] - p[0]) - mk(rk9[3 * k], rk9[3 * k + 1], rk9[3 * k + 2]);
    Strain32 st = preciseStrain(d, br);
    V3<float> cof[3];
    float J;
    preciseFrame(st, F, cof, J);
    float pJ = std::fmax(J, 0.05f);
    V3<float> Gh[3];
    float q = preciseDev(st, F, cof, J > 0.05f, pJ, isoPow(pJ), Gh);
    out[0] = st.a;
    out[1] = st.b;
    out[2] = q;
    for (int c = 0; c < 3; c++) { out[3 + 3 * c] = Gh[c].x; out[4 + 3 * c] = Gh[c].y; out[5 + 3 * c] = Gh[c].z; }
}
extern "C" float oracle_iso_minus_one_f32(float a, float iso, float* tails) {
    return isoMinusOne(a, iso, tails[0], tails[1]);
}
extern "C" float oracle_log1p_f32(float a) { return log1pOwn(a); }
extern "C" int oracle_moller_trumbore_f64(const double* O, const double* D, const double* V, double* tuv) {
    double t, u, v;
    bool hit = mollerTrumbore<double>(mk(O[0], O[1], O[2]), mk(D[0], D[1], D[2]), mk(V[0], V[1], V[2]),
                                      mk(V[3], V[4], V[5]), mk(V[6], V[7], V[8]), t, u, v);
    tuv[0] = t; tuv[1] = u; tuv[2] = v;
    return hit ? 1 : 0;
}
extern "C" float oracle_iso_pow_f32(float J) { return isoPow(J); }
extern "C" float oracle_log_f32(float x) { return guardLog(x); }
extern "C" double oracle_shape_distance_f64(const double* center, const double* u, const double* n, const double* half,
                                           const double* a, const double* b, double radius, int capsule,
                                           const double* x, double* normal) {
    Shape<double> s{};
    s.center = mk(center[0], center[1], center[2]);
    s.u = mk(u[0], u[1], u[2]);
    s.n = mk(n[0], n[1], n[2]);
    s.w = cross3(s.u, s.n);
    s.half = mk(half[0], half[1], half[2]);
    s.a = mk(a[0], a[1], a[2]);
    s.b = mk(b[0], b[1], b[2]);
    s.radius = radius;
    s.capsule = capsule;
    V3<double> nrm, surf;
    double d = shapeDistance(s, mk(x[0], x[1], x[2]), nrm, surf);
    normal[0] = nrm.x; normal[1] = nrm.y; normal[2] = nrm.z;
    return d;
}

// ---------------------------------------------------------------------------------------------------------------
// Adjoint of the substep for per-tet compliance gradients (no reference counterpart; pinned by finite differences).
//
// Forward: `nsub` substeps of the legacy branch with `iterations` Gauss-Seidel iterations (multipliers reset every
// substep, read from the second iteration on, as solveEdge / solveTet do), no damping, no contact, omega = 1,
// coupled Neo-Hookean tets (mode 1), node forces `force` (gravity * mass plus point loads) in place of the gravity
// field. Per-tet compliance c gives alpha~_D = (1/h^2)/(1/c) and alpha~_H = (1/h^2)/(10/c).
// Loss: sum over masked nodes of |x - target|^2 after the last substep. Backward: every substep in reverse, each
// re-run forward from its stored start state with a tape of the inputs (positions and multipliers) of every
// projection, then reversed iteration by iteration and colour by colour; the multipliers carry adjoints between the
// iterations of a substep. Gradients are not taken through the max(J, .05) and max(q, 1e-12) clamps (zero
// subgradient).
// ---------------------------------------------------------------------------------------------------------------
namespace {

template <class R> struct TetTape { V3<R> p[4]; R lamD, lamH; };
// The subgradient rule at the floor of C_D (the kernels always apply it). Off: the exact derivative through the
// clamp (the direction G / 1e-6 differentiated as it stands), which central differences of the FP64 forward see.
bool g_clampRule = true;
template <class R> struct EdgeTape { V3<R> xi, xj; R lam; };

// Reverse of tetProject (mode 1, no damping, omega = 1) at input positions p and multipliers lamD / lamH: given the
// adjoints of the outputs (ob for the positions, lDb / lHb for the multipliers, replaced by the adjoints of the input
// multipliers), accumulates the adjoint of the input positions (pb) and of alpha~_D / alpha~_H.
template <class R>
void tetProjectReverse(const V3<R> p[4], const R w[4], const V3<R> br[3], R ad, R ah, R lamD, R lamH,
                       const V3<R> ob[4], V3<R> pb[4], R& adb, R& ahb, R& lDb, R& lHb, const V3<R>* rk = nullptr) {
    V3<R> e1 = p[1] - p[0], e2 = p[2] - p[0], e3 = p[3] - p[0];
    V3<R> F[3], cof[3];
    R J, I1;
    // FP32: the precise invariants of the adjoint's forward, F, cof F and J from H (the FP64 oracle keeps the direct
    // forms)
    Strain32 st{};
    if constexpr (std::is_same<R, float>::value) {
        V3<float> d[3] = {e1 - rk[0], e2 - rk[1], e3 - rk[2]};
        st = preciseStrain(d, br);
        preciseFrame(st, F, cof, J);
        I1 = 3.f + st.b;
    } else {
        for (int c = 0; c < 3; c++) F[c] = e1 * comp(br[0], c) + e2 * comp(br[1], c) + e3 * comp(br[2], c);
        cof[0] = cross3(F[1], F[2]);
        cof[1] = cross3(F[2], F[0]);
        cof[2] = cross3(F[0], F[1]);
        J = dot3(F[0], cof[0]);
        I1 = dot3(F[0], F[0]) + dot3(F[1], F[1]) + dot3(F[2], F[2]);
    }
    V3<R> gh[4], gd[4], Fb[4], G[4];
    for (int k = 1; k < 4; k++) {
        V3<R> b = br[k - 1];
        gh[k] = cof[0] * b.x + cof[1] * b.y + cof[2] * b.z;
    }
    R pJ = fmax_<R>(J, (R)0.05f);
    R iso = isoPow(pJ);
    R qraw = iso * I1 - (R)3;
    bool jlive = J > (R)0.05f;
    R CH = J - (R)1;
    V3<R> Gh[3];
    if constexpr (std::is_same<R, float>::value) {
        qraw = preciseDev(st, F, cof, jlive, pJ, iso, Gh);
        CH = st.a;
    } else {
        R vgd = jlive ? iso * I1 / ((R)3 * pJ) : (R)0;
        for (int c = 0; c < 3; c++) Gh[c] = F[c] * iso - cof[c] * vgd;
    }
    R q = fmax_<R>(qraw, (R)0);
    R CD = std::sqrt(fmax_<R>(q, (R)1e-12f));
    R ds = (R)1 / CD;
    // at or below the floor of C_D the row's value and direction are held constant (no adjoint into F through them)
    const bool atFloor = !(q > (R)1e-12f), frozen = atFloor && g_clampRule;
    R vg = jlive ? iso * I1 / ((R)3 * pJ) : (R)0;
    for (int k = 1; k < 4; k++) {
        V3<R> b = br[k - 1];
        Fb[k] = F[0] * b.x + F[1] * b.y + F[2] * b.z;
        if constexpr (std::is_same<R, float>::value) G[k] = Gh[0] * b.x + Gh[1] * b.y + Gh[2] * b.z;
        else G[k] = Fb[k] * iso - gh[k] * vg;
        gd[k] = G[k] * ds;
    }
    gd[0] = (gd[1] + gd[2] + gd[3]) * (R)-1;
    gh[0] = (gh[1] + gh[2] + gh[3]) * (R)-1;
    R dd = 0, dh = 0;
    for (int k = 0; k < 4; k++) {
        dd += w[k] * dot3(gd[k], gd[k]);
        dh += w[k] * dot3(gd[k], gh[k]);
    }
    R A = dd + ad, gamma = dh / A;
    V3<R> Rk[4];
    R schur = ah + ad * gamma * gamma;
    for (int k = 0; k < 4; k++) {
        Rk[k] = gh[k] - gd[k] * gamma;
        schur += w[k] * dot3(Rk[k], Rk[k]);
    }
    R rd = -CD - ad * lamD, rh = -CH - ah * lamH;
    for (int k = 0; k < 4; k++) pb[k] = ob[k];
    if (!(A > 0) || !(schur > 0)) return;
    R dlH = (rh - gamma * rd) / schur, dlD = (rd - dh * dlH) / A;
    if (!std::isfinite(dlD) || !std::isfinite(dlH)) return;
    // out_k = p_k + (gd_k dlD + gh_k dlH) w_k; lamD' = lamD + dlD; lamH' = lamH + dlH
    V3<R> gdb[4], ghb[4];
    R dlDb = lDb, dlHb = lHb;
    for (int k = 0; k < 4; k++) {
        V3<R> u = ob[k] * w[k];
        gdb[k] = u * dlD;
        ghb[k] = u * dlH;
        dlDb += dot3(u, gd[k]);
        dlHb += dot3(u, gh[k]);
    }
    // dlD = (rd - dh dlH) / A
    R rdb = dlDb / A, dhb = -dlDb * dlH / A, Ab = -dlDb * dlD / A;
    dlHb += -dlDb * dh / A;
    // dlH = (rh - gamma rd) / schur
    R rhb = dlHb / schur, gammab = -dlHb * rd / schur, schurb = -dlHb * dlH / schur;
    rdb += -dlHb * gamma / schur;
    // rd = -CD - ad lamD; rh = -CH - ah lamH
    R CDb = -rdb, CHb = -rhb;
    adb += -lamD * rdb;
    ahb += -lamH * rhb;
    lDb += -ad * rdb;
    lHb += -ah * rhb;
    // schur = ah + ad gamma^2 + sum w |gh - gamma gd|^2
    ahb += schurb;
    adb += schurb * gamma * gamma;
    gammab += schurb * (R)2 * ad * gamma;
    for (int k = 0; k < 4; k++) {
        V3<R> Rb = Rk[k] * (schurb * (R)2 * w[k]);
        ghb[k] = ghb[k] + Rb;
        gdb[k] = gdb[k] - Rb * gamma;
        gammab += -dot3(Rb, gd[k]);
    }
    // gamma = dh / A
    dhb += gammab / A;
    Ab += -gammab * gamma / A;
    // A = dd + ad
    R ddb = Ab;
    adb += Ab;
    for (int k = 0; k < 4; k++) {
        gdb[k] = gdb[k] + gd[k] * ((R)2 * w[k] * ddb) + gh[k] * (w[k] * dhb);
        ghb[k] = ghb[k] + gd[k] * (w[k] * dhb);
    }
    for (int k = 1; k < 4; k++) {
        gdb[k] = gdb[k] - gdb[0];
        ghb[k] = ghb[k] - ghb[0];
    }
    // gd_k = G_k ds, G_k = Fb_k iso - gh_k vg
    R dsb = 0, isob = 0, vgb = 0;
    V3<R> Fbb[4] = {mk<R>(0, 0, 0), mk<R>(0, 0, 0), mk<R>(0, 0, 0), mk<R>(0, 0, 0)};
    if (!frozen)
        for (int k = 1; k < 4; k++) {
            V3<R> Gb = gdb[k] * ds;
            dsb += dot3(gdb[k], G[k]);
            Fbb[k] = Gb * iso;
            isob += dot3(Gb, Fb[k]);
            ghb[k] = ghb[k] - Gb * vg;
            vgb += -dot3(Gb, gh[k]);
        }
    V3<R> Fbar[3] = {mk<R>(0, 0, 0), mk<R>(0, 0, 0), mk<R>(0, 0, 0)};
    for (int k = 1; k < 4; k++) {
        V3<R> b = br[k - 1];
        Fbar[0] = Fbar[0] + Fbb[k] * b.x;
        Fbar[1] = Fbar[1] + Fbb[k] * b.y;
        Fbar[2] = Fbar[2] + Fbb[k] * b.z;
    }
    // ds = 1 / CD; CD = sqrt(q) above the floor; dq/dF = 2 Ghat
    CDb += -dsb / (CD * CD);
    R qb = atFloor ? (R)0 : CDb / ((R)2 * CD);
    if (!atFloor)
        for (int c = 0; c < 3; c++) Fbar[c] = Fbar[c] + Gh[c] * ((R)2 * qb);
    R I1b = 0;
    R pJb = 0;
    if (jlive) {
        isob += vgb * I1 / ((R)3 * pJ);
        I1b += vgb * iso / ((R)3 * pJ);
        pJb += -vgb * iso * I1 / ((R)3 * pJ * pJ);
    }
    pJb += isob * ((R)(-2.0 / 3.0)) * iso / pJ;
    R Jb = CHb + (jlive ? pJb : (R)0);
    for (int c = 0; c < 3; c++) Fbar[c] = Fbar[c] + F[c] * ((R)2 * I1b);
    V3<R> cofb[3] = {mk<R>(0, 0, 0), mk<R>(0, 0, 0), mk<R>(0, 0, 0)};
    for (int k = 1; k < 4; k++) {
        V3<R> b = br[k - 1];
        cofb[0] = cofb[0] + ghb[k] * b.x;
        cofb[1] = cofb[1] + ghb[k] * b.y;
        cofb[2] = cofb[2] + ghb[k] * b.z;
    }
    Fbar[0] = Fbar[0] + cof[0] * Jb;
    cofb[0] = cofb[0] + F[0] * Jb;
    Fbar[1] = Fbar[1] + cross3(F[2], cofb[0]);
    Fbar[2] = Fbar[2] + cross3(cofb[0], F[1]);
    Fbar[2] = Fbar[2] + cross3(F[0], cofb[1]);
    Fbar[0] = Fbar[0] + cross3(cofb[1], F[2]);
    Fbar[0] = Fbar[0] + cross3(F[1], cofb[2]);
    Fbar[1] = Fbar[1] + cross3(cofb[2], F[0]);
    for (int k = 0; k < 3; k++) {
        V3<R> eb = Fbar[0] * comp(br[k], 0) + Fbar[1] * comp(br[k], 1) + Fbar[2] * comp(br[k], 2);
        pb[k + 1] = pb[k + 1] + eb;
        pb[0] = pb[0] - eb;
    }
}

// Reverse of solveEdge (no damping, omega = 1) at inputs xi, xj and multiplier lam; lb holds the adjoint of the
// output multiplier on entry and of the input multiplier on return.
template <class R>
void edgeProjectReverse(V3<R> xi, V3<R> xj, R wi, R wj, R L0, R at, R lam, V3<R> oib, V3<R> ojb, V3<R>& xib,
                        V3<R>& xjb, R& lb) {
    xib = oib;
    xjb = ojb;
    R wsum = wi + wj;
    if (wsum <= 0) return;
    V3<R> d = xi - xj;
    R L = len3(d);
    if (L < (R)1e-9f) return;
    V3<R> n = d * ((R)1 / L);
    R C = L - L0;
    R kap = (R)1 / (wsum + at);
    R dl = (-C - at * lam) * kap;
    V3<R> m = oib * wi - ojb * wj;
    V3<R> nb = m * dl;
    R dlb = dot3(m, n) + lb;
    R Lb = -dlb * kap;
    lb += -at * kap * dlb;
    V3<R> db = (nb - n * dot3(nb, n)) * ((R)1 / L) + n * Lb;
    xib = xib + db;
    xjb = xjb - db;
}

// One forward substep (mode 1, `iterations` iterations) with an external force field; optionally records the tape
// of iteration `it` at tt[it * T + t] / et[it * E + e].
template <class R>
void adjointSubstep(Tissue<R>& s, const std::vector<V3<R>>& force, R h, int iterations, std::vector<TetTape<R>>* tt,
                    std::vector<EdgeTape<R>>* et) {
    for (int p = 0; p < s.N; p++) {
        s.prev[p] = s.pos[p];
        if (s.invMass[p] > 0) {
            V3<R> a = force[p] * s.invMass[p];
            V3<R> v = s.vel[p] + a * h;
            s.vel[p] = v;
            s.pos[p] = s.pos[p] + v * h;
        }
    }
    std::fill(s.elambda.begin(), s.elambda.end(), (R)0);
    std::fill(s.lamDev.begin(), s.lamDev.end(), (R)0);
    std::fill(s.lamHyd.begin(), s.lamHyd.end(), (R)0);
    for (int it = 0; it < iterations; it++) {
        const bool useLambda = it > 0;
        for (size_t c = 0; c < s.ecolorCnt.size(); c++)
            for (int k = 0; k < s.ecolorCnt[c]; k++) {
                int e = s.ecolorOff[c] + k;
                if (et) {
                    EdgeTape<R>& r = (*et)[(size_t)it * s.E + e];
                    r.xi = s.pos[s.ei[e]];
                    r.xj = s.pos[s.ej[e]];
                    r.lam = useLambda ? s.elambda[e] : (R)0;
                }
                solveEdge(s, e, s.alphaTilde, false, (R)0, useLambda);
            }
        for (size_t c = 0; c < s.tcolorCnt.size(); c++)
            for (int k = 0; k < s.tcolorCnt[c]; k++) {
                int t = s.tcolorOrder[s.tcolorOff[c] + k];
                if (tt) {
                    TetTape<R>& r = (*tt)[(size_t)it * s.T + t];
                    for (int j = 0; j < 4; j++) r.p[j] = s.pos[s.tids[4 * t + j]];
                    r.lamD = useLambda ? s.lamDev[t] : (R)0;
                    r.lamH = useLambda ? s.lamHyd[t] : (R)0;
                }
                solveTet(s, t, (R)0, (R)0, useLambda);
            }
    }
    R invH = (R)1 / h;
    for (int p = 0; p < s.N; p++) s.vel[p] = s.invMass[p] > 0 ? (s.pos[p] - s.prev[p]) * invH : mk<R>(0, 0, 0);
    for (int p = 0; p < s.N; p++) s.vel[p] = s.invMass[p] > 0 ? s.vel[p] * s.velScale : mk<R>(0, 0, 0);
}

template <class R>
double adjointRun(Tissue<R>& s, R dt, int nsub, int iterations, const float* force, const float* compliance,
                  const float* mask, const float* target, R* grad) {
    // frame constants as xpbd_begin_frame with substeps = d.substeps (h = dt / substeps); the FP32 forward takes the
    // precise invariants, as the reverse needs them
    s.d.precise_strain = 1;
    int Nsub = s.d.substeps > 0 ? s.d.substeps : 1;
    R h = dt / (R)Nsub;
    R ks = (R)s.d.edge_ks;
    s.alphaTilde = ks > 0 ? (R)1 / (ks * h * h) : (R)0;
    s.velScale = (R)1;
    s.betaGammaH = (R)0;
    s.h = h;
    R invH2 = (R)1 / (h * h);
    for (int t = 0; t < s.T; t++) {
        R kd = (R)1 / (R)compliance[t], kh = (R)1 / ((R)compliance[t] / (R)10);
        if (sizeof(R) == 4) { kd = (R)(float)(1.0 / (double)compliance[t]); kh = (R)(float)(1.0 / ((double)compliance[t] / 10.0)); }
        s.alphaDev[t] = invH2 / kd;
        s.alphaHyd[t] = invH2 / kh;
    }
    std::vector<V3<R>> F(s.N);
    for (int p = 0; p < s.N; p++) F[p] = v3<R>(force + 3 * p);
    refreshLiveness(s);
    std::vector<V3<R>> xs((size_t)(nsub + 1) * s.N), vs((size_t)(nsub + 1) * s.N);
    for (int n = 0; n < nsub; n++) {
        std::copy(s.pos.begin(), s.pos.end(), xs.begin() + (size_t)n * s.N);
        std::copy(s.vel.begin(), s.vel.end(), vs.begin() + (size_t)n * s.N);
        adjointSubstep<R>(s, F, h, iterations, nullptr, nullptr);
    }
    std::copy(s.pos.begin(), s.pos.end(), xs.begin() + (size_t)nsub * s.N);
    std::copy(s.vel.begin(), s.vel.end(), vs.begin() + (size_t)nsub * s.N);
    double loss = 0;
    std::vector<V3<R>> xb(s.N), vb(s.N);
    for (int p = 0; p < s.N; p++) {
        V3<R> d = s.pos[p] - v3<R>(target + 3 * p);
        loss += (double)mask[p] * (double)dot3(d, d);
        xb[p] = d * ((R)2 * (R)mask[p]);
        vb[p] = mk<R>(0, 0, 0);
    }
    if (grad) {
        std::vector<R> adb(s.T, 0), ahb(s.T, 0), lDb(s.T), lHb(s.T), leb(s.E);
        std::vector<TetTape<R>> tt((size_t)iterations * s.T);
        std::vector<EdgeTape<R>> et((size_t)iterations * s.E);
        std::vector<V3<R>> prevb(s.N);
        for (int n = nsub - 1; n >= 0; n--) {
            std::copy(xs.begin() + (size_t)n * s.N, xs.begin() + (size_t)(n + 1) * s.N, s.pos.begin());
            std::copy(vs.begin() + (size_t)n * s.N, vs.begin() + (size_t)(n + 1) * s.N, s.vel.begin());
            adjointSubstep<R>(s, F, h, iterations, &tt, &et);
            // velocity: v' = (x_end - prev) / h * velScale for w > 0
            R invH = (R)1 / h;
            for (int p = 0; p < s.N; p++) {
                if (s.invMass[p] > 0) {
                    V3<R> g = vb[p] * (invH * s.velScale);
                    xb[p] = xb[p] + g;
                    prevb[p] = g * (R)-1;
                } else prevb[p] = mk<R>(0, 0, 0);
            }
            // the multipliers of the last iteration reach nothing after the substep
            std::fill(lDb.begin(), lDb.end(), (R)0);
            std::fill(lHb.begin(), lHb.end(), (R)0);
            std::fill(leb.begin(), leb.end(), (R)0);
            for (int it = iterations - 1; it >= 0; it--) {
                // tets, colours in reverse
                for (int c = (int)s.tcolorCnt.size() - 1; c >= 0; c--)
                    for (int k = 0; k < s.tcolorCnt[c]; k++) {
                        int t = s.tcolorOrder[s.tcolorOff[c] + k];
                        if (!s.tactive[t]) continue;
                        const int* id = &s.tids[4 * t];
                        R w[4];
                        V3<R> ob[4], pb[4];
                        for (int j = 0; j < 4; j++) { w[j] = s.invMass[id[j]]; ob[j] = xb[id[j]]; }
                        const R(*B)[3] = s.typeB[s.ttype[t]];
                        V3<R> br[3] = {mk<R>(B[0][0], B[0][1], B[0][2]), mk<R>(B[1][0], B[1][1], B[1][2]),
                                       mk<R>(B[2][0], B[2][1], B[2][2])};
                        const TetTape<R>& r = tt[(size_t)it * s.T + t];
                        tetProjectReverse<R>(r.p, w, br, s.alphaDev[t], s.alphaHyd[t], r.lamD, r.lamH, ob, pb, adb[t],
                                             ahb[t], lDb[t], lHb[t], s.typeRk[s.ttype[t]]);
                        for (int j = 0; j < 4; j++) xb[id[j]] = pb[j];
                    }
                for (int c = (int)s.ecolorCnt.size() - 1; c >= 0; c--)
                    for (int k = 0; k < s.ecolorCnt[c]; k++) {
                        int e = s.ecolorOff[c] + k;
                        if (!s.eactive[e]) continue;
                        int i = s.ei[e], j = s.ej[e];
                        const EdgeTape<R>& r = et[(size_t)it * s.E + e];
                        V3<R> xib, xjb;
                        edgeProjectReverse<R>(r.xi, r.xj, s.invMass[i], s.invMass[j], s.erest[e], s.alphaTilde, r.lam,
                                              xb[i], xb[j], xib, xjb, leb[e]);
                        xb[i] = xib;
                        xb[j] = xjb;
                    }
                // the input multipliers of iteration 0 are the substep's reset zeros
            }
            // predict: xp = x + h (v + h a) for w > 0, prev = x
            for (int p = 0; p < s.N; p++) {
                V3<R> g = xb[p];
                xb[p] = g + prevb[p];
                vb[p] = s.invMass[p] > 0 ? g * h : mk<R>(0, 0, 0);
            }
        }
        for (int t = 0; t < s.T; t++) grad[t] = (adb[t] + ahb[t] / (R)10) * invH2;
    }
    // leave the state at the end of the forward run
    std::copy(xs.begin() + (size_t)nsub * s.N, xs.end(), s.pos.begin());
    std::copy(vs.begin() + (size_t)nsub * s.N, vs.end(), s.vel.begin());
    return loss;
}

}  // namespace

#define ORACLE_ADJ(R, SUF)                                                                                            \
    extern "C" double oracle_adjoint_##SUF(void* h, float dt, int nsub, int iterations, const float* force,          \
                                           const float* compliance, const float* mask, const float* target, R* grad) { \
        return adjointRun(*(Tissue<R>*)h, (R)dt, nsub, iterations, force, compliance, mask, target, grad);            \
    }                                                                                                                  \
    extern "C" void oracle_tet_reverse_##SUF(const R* p12, const R* w4, const R* B9, const R* rk9, R ad, R ah,      \
                                             const R* lam2, const R* ob12, R* lb2, R* pb12, R* abar2) {                \
        V3<R> p[4], ob[4], pb[4], br[3], rk[3];                                                                        \
        R w[4];                                                                                                        \
        for (int k = 0; k < 4; k++) {                                                                                  \
            p[k] = mk(p12[3 * k], p12[3 * k + 1], p12[3 * k + 2]);                                                     \
            ob[k] = mk(ob12[3 * k], ob12[3 * k + 1], ob12[3 * k + 2]);                                                 \
            w[k] = w4[k];                                                                                              \
        }                                                                                                              \
        for (int r = 0; r < 3; r++) br[r] = mk(B9[3 * r], B9[3 * r + 1], B9[3 * r + 2]);                              \
        for (int r = 0; r < 3; r++) rk[r] = rk9 ? mk(rk9[3 * r], rk9[3 * r + 1], rk9[3 * r + 2]) : mk<R>(0, 0, 0);     \
        R adb = 0, ahb = 0;                                                                                            \
        tetProjectReverse<R>(p, w, br, ad, ah, lam2[0], lam2[1], ob, pb, adb, ahb, lb2[0], lb2[1], rk);                \
        for (int k = 0; k < 4; k++) { pb12[3 * k] = pb[k].x; pb12[3 * k + 1] = pb[k].y; pb12[3 * k + 2] = pb[k].z; }   \
        abar2[0] = adb;                                                                                                \
        abar2[1] = ahb;                                                                                                \
    }

extern "C" void oracle_adjoint_clamp_rule(int on) { g_clampRule = on != 0; }
ORACLE_ADJ(float, f32)
ORACLE_ADJ(double, f64)

// LCS_GetMaterialQuality (xpbd_get_material_quality with k_MaterialQuality, xpbd_solver.cu:2122-2138): over the live
// tets (ids [4T] in the tet's corner order, Dm^-1 rows [9T], active [T]), out[0] = min det F seeded at 1, out[1] = the
// number with det F <= 0 or a non-finite det F, where F = [x1 - x0, x2 - x0, x3 - x0] Dm^-1 and det F = f0 . (f1 x f2).
extern "C" void oracle_material_quality(const float* pos, const int* ids, const float* rows, const int* active, int T,
                                        float* out) {
    out[0] = 1.f;
    out[1] = 0.f;
    for (int i = 0; i < T; i++) {
        if (!active[i]) continue;
        const int* id = ids + 4 * (size_t)i;
        const float* b = rows + 9 * (size_t)i;
        V3<float> a = v3<float>(pos + 3 * (size_t)id[0]);
        V3<float> e1 = v3<float>(pos + 3 * (size_t)id[1]) - a, e2 = v3<float>(pos + 3 * (size_t)id[2]) - a,
                  e3 = v3<float>(pos + 3 * (size_t)id[3]) - a;
        V3<float> f[3];
        for (int k = 0; k < 3; k++) f[k] = e1 * b[k] + e2 * b[3 + k] + e3 * b[6 + k];
        float j = dot3(f[0], cross3(f[1], f[2]));
        if (!std::isfinite(j)) { out[1] += 1.f; continue; }
        if (j < out[0]) out[0] = j;
        if (j <= 0.f) out[1] += 1.f;
    }
}

// LCS_SurfaceDepth for the grasper (xpbd_surface_depth, xpbd_solver.cu:2046-2059, with k_SurfaceDepth,
// surface_contact.cuh:662-674): the deepest penetration of the tool shapes whose bits are set in `kinds` (0 shaft
// capsule, 1 and 2 the jaw boxes, 3 the blade capsule) into the surface soup `pos` [3K] (K indices, three per
// triangle), 0 when none.
extern "C" float oracle_surface_depth(const OracleTool* tool, int kinds, float thickness, float minGap, const float* pos,
                                      int K) {
    ToolR<float> t = toolFrom<float>(*tool);
    float result = 0.f;
    for (int kind = 0; kind < 4; kind++) {
        if (!((kinds >> kind) & 1)) continue;
        Shape<float> sh = contactShape<float>(t, kind, thickness, minGap);
        if (sh.capsule ? sh.radius <= 0 : (sh.half.x <= 0 || sh.half.z <= 0)) continue;
        V3<float> lo, hi;
        shapeBoundsF(sh, 0.f, lo, hi);
        for (int i = 0; i < K / 3; i++) {
            V3<float> a = v3<float>(pos + 9 * (size_t)i), b = v3<float>(pos + 9 * (size_t)i + 3),
                      c = v3<float>(pos + 9 * (size_t)i + 6);
            bool out = false;
            for (int ax = 0; ax < 3; ax++) {
                float mx = std::fmax(comp(a, ax), std::fmax(comp(b, ax), comp(c, ax)));
                float mn = std::fmin(comp(a, ax), std::fmin(comp(b, ax), comp(c, ax)));
                if (mx < comp(lo, ax) || mn > comp(hi, ax)) out = true;
            }
            if (out) continue;
            V3<float> bar = triangleWitnessF(sh, a, b, c), n, s;
            float depth = std::fmax(0.f, -shapeDistance(sh, a * bar.x + b * bar.y + c * bar.z, n, s));
            result = std::fmax(result, depth);
        }
    }
    return result;
}

// ---- contact pieces in FP64 for the CPU pins: a shape is 23 doubles (center, u, n, w, half, a, b, radius, capsule)
namespace {
Shape<double> shapeOf(const double* v) {
    Shape<double> s{};
    s.center = mk(v[0], v[1], v[2]); s.u = mk(v[3], v[4], v[5]); s.n = mk(v[6], v[7], v[8]); s.w = mk(v[9], v[10], v[11]);
    s.half = mk(v[12], v[13], v[14]); s.a = mk(v[15], v[16], v[17]); s.b = mk(v[18], v[19], v[20]);
    s.radius = v[21]; s.capsule = v[22] != 0;
    return s;
}
}  // namespace
extern "C" int oracle_box_touch_time_f64(const double* box23, double radius, const double* a3, const double* b3,
                                         double* time) {
    return boxTouchTime<double>(shapeOf(box23), radius, mk(a3[0], a3[1], a3[2]), mk(b3[0], b3[1], b3[2]), *time) ? 1
                                                                                                                    : 0;
}
// out: normal (3), target (3), motion (3); returns valid
extern "C" int oracle_trace_contact_f64(const double* from23, const double* to23, double radius, const double* x0,
                                        const double* x1, double* out9) {
    Record<double> r = traceToolContact<double>(shapeOf(from23), shapeOf(to23), radius, mk(x0[0], x0[1], x0[2]),
                                                mk(x1[0], x1[1], x1[2]));
    V3<double> o[3] = {r.normal, r.target, r.motion};
    for (int k = 0; k < 3; k++) { out9[3 * k] = o[k].x; out9[3 * k + 1] = o[k].y; out9[3 * k + 2] = o[k].z; }
    return r.valid;
}
// one friction pass on x (3) with the contact record rec9 (normal, target, motion); lam4 = lambda_N then lambda_T (3)
extern "C" void oracle_friction_update_f64(const double* rec9, double w, double alpha, double muS, double muD,
                                           const double* prev3, double* x3, double* lam4) {
    Record<double> c;
    c.normal = mk(rec9[0], rec9[1], rec9[2]); c.target = mk(rec9[3], rec9[4], rec9[5]);
    c.motion = mk(rec9[6], rec9[7], rec9[8]); c.valid = 1;
    V3<double> x = mk(x3[0], x3[1], x3[2]), lt = mk(lam4[1], lam4[2], lam4[3]);
    frictionUpdate<double>(c, w, alpha, muS, muD, mk(prev3[0], prev3[1], prev3[2]), x, lam4[0], lt);
    x3[0] = x.x; x3[1] = x.y; x3[2] = x.z;
    lam4[1] = lt.x; lam4[2] = lt.y; lam4[3] = lt.z;
}
extern "C" double oracle_tool_gap_f64(const double* shape23, const double* x3) {
    return toolHit<double>(shapeOf(shape23), mk(x3[0], x3[1], x3[2])).gap;
}
