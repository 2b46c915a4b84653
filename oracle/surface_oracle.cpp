// CPU oracle of the surface reconstruction of the deformed, cut lattice (test infrastructure only: the tests, the
// smoke check and the CPU baseline call it as the checker; the product never does).
//
// Restates, single-threaded and in FP32 with the reference's operation order:
//   - the dual-contouring rebuild `recon_build_static` (dc_recon.cu:498-551): k_RebuildIsectWorld (:97-142),
//     k_FeaturePoints (:149-207), k_AdvectIntactSurface (:477-493), k_Stitch (:239-287), k_ClearNormals /
//     k_NormalsScatter / k_NormalsNormalize (:292-345, :441-451), k_ExpandVertices (:358-386);
//   - the cut feature-point chain `cut_fp_chain` (cut.cu:1434-1461): k_ClearCutFP, k_LookupConnectivity,
//     k_AccumulateCutFP (:761-836), k_ComputeComponentFP (:858-985); k_BuildCutTriangles with EmitCutQuad
//     (:988-1090); k_NormalizeCutFP (:1093-1100);
//   - the internal-face cleanup `cleanInternalFaces` (surface_cleanup.cuh:1-53);
//   - the helpers SkinOwnerA / CutHitsMaterial / DirectedEdgeSlot / SurfaceOwner (cut.cuh:16-31), the edge stencil
//     (cut.cuh:120-155) and the corner order (cut.cuh:90-109).
//
// Where the reference orders work with atomics this oracle fixes the sequential order, which the HIP kernels follow:
//   - cut points add into their (voxel, component) sums in cut point index order (the reference adds floats with
//     atomicAdd in arrival order);
//   - triangles are listed as the stitch emits them in surface edge order (two per edge), then the cut walls in node
//     order (six edges, up to two triangles each); the first tri_capacity are kept (the reference takes slots with
//     atomicAdd and drops what overflows);
//   - the cleanup keeps, for each face key with a nonzero net orientation count, the lowest index of the majority
//     orientation, in index order (the reference compacts in hash-table order).
// Normal sums are fixed-point integers in the reference, so their order does not matter.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

namespace {

struct F3 {
    float x, y, z;
};
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline F3 operator/(F3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float len(F3 a) { return std::sqrt(dot(a, a)); }
inline F3 lerp(F3 a, F3 b, float t) { return f3(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t); }
inline F3 ld(const float* p) { return f3(p[0], p[1], p[2]); }

const float NORMAL_SCALE = 1048576.0f;  // common.cuh:45
const float POS_SCALE = 16384.0f;       // cut.cuh:12

// Voxel corner offsets in figure order (cut.cuh:90-99).
const int CORNER[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
// The four voxels around a grid edge and the edge's local id in each (cut.cuh:120-132).
const int STENCIL[3][4][3] = {{{0, 0, 0}, {0, -1, 0}, {0, -1, -1}, {0, 0, -1}},
                              {{0, 0, 0}, {0, 0, -1}, {-1, 0, -1}, {-1, 0, 0}},
                              {{0, 0, 0}, {-1, 0, 0}, {-1, -1, 0}, {0, -1, 0}}};
const int LOCAL_EDGE[3][4] = {{0, 2, 6, 4}, {3, 7, 5, 1}, {8, 9, 10, 11}};

struct Isect {  // IsectGpu (common.cuh:23-33), 40 bytes
    float rest[3], normal[3];
    int a, b;
    float t;
    int insideA;
};
static_assert(sizeof(Isect) == 40, "IsectGpu is 40 bytes");

struct Desc {
    int n_nodes, n_voxels, n_isect, n_surface_edges, n_grid_edges, tri_capacity;
    int vx, vy, vz;
    int qef_iters;
    int cut_aware;  // 1: LCS_Finalize's rebuild with the cut buffers (recon_set_cut_buffers); 0: recon_init's build
    int have_rest;  // the rest surface snapshot exists (k_AdvectIntactSurface runs)
    int n_cut_points;
};

struct Lat {
    int vx, vy, vz, nx, ny;
    int corner(int x, int y, int z) const { return x + nx * (y + ny * z); }
    int voxel(int x, int y, int z) const { return x + vx * (y + vy * z); }
    void coord(int id, int& x, int& y, int& z) const {
        x = id % nx;
        y = (id / nx) % ny;
        z = id / (nx * ny);
    }
    bool voxIn(int x, int y, int z) const { return x >= 0 && y >= 0 && z >= 0 && x < vx && y < vy && z < vz; }
    int slot(int a, int b) const {  // DirectedEdgeSlot (cut.cuh:23-26)
        int d = b - a, ad = d < 0 ? -d : d;
        int axis = ad == 1 ? 0 : (ad == nx ? 1 : 2);
        return 6 * a + 2 * axis + (d > 0 ? 1 : 0);
    }
};

inline int skinOwnerA(float skinT, float cutT, int insideA) {  // cut.cuh:16-18
    return cutT >= 0.f && cutT <= 1.f && std::fabs(skinT - cutT) > 1e-6f ? (skinT < cutT ? 1 : 0) : insideA;
}
inline bool cutHitsMaterial(float t, float skinT, int insideA, int insideB) {  // cut.cuh:19-22
    if (insideA && insideB) return true;
    if (!insideA && !insideB) return false;
    if (!(skinT >= 0.f && skinT <= 1.f)) return false;
    return insideA ? t <= skinT : t >= skinT;
}
inline int localCornerOf(int cx, int cy, int cz, int vx, int vy, int vz) {  // cut.cuh:101-109
    int dx = cx - vx, dy = cy - vy, dz = cz - vz;
    for (int lc = 0; lc < 8; lc++)
        if (CORNER[lc][0] == dx && CORNER[lc][1] == dy && CORNER[lc][2] == dz) return lc;
    return -1;
}
inline F3 mulRv(const float* R, F3 v) { return f3(dot(ld(R), v), dot(ld(R + 3), v), dot(ld(R + 6), v)); }

struct Surface {
    const Desc& d;
    Lat L;
    const float *pos, *rest, *rot;
    const int *nbr, *active, *inside, *corners, *ioff, *icnt, *sedges, *conn, *gedges;
    const Isect* isect;
    const uint32_t *occupied, *mask;
    const float *cutT, *skinT, *points;
    const float* restFP;
    std::vector<F3> iw;
    std::vector<float> fp, cutFP;

    Surface(const Desc& d_) : d(d_) {}

    F3 P(int c) const { return ld(pos + 3 * c); }
    F3 Rst(int c) const { return ld(rest + 3 * c); }
    int comp(uint32_t m, int lc) const { return conn[9 * (int)m + 1 + lc]; }
    F3 fetch(int raw) const {  // FetchPos (dc_recon.cu:228-234)
        uint32_t u = (uint32_t)raw;
        int idx = (int)(u & 0x3FFFFFFFu);
        if (u & 0x80000000u) return ld(&cutFP[4 * (size_t)idx]);
        return ld(&fp[4 * (size_t)idx]);
    }
    int surfaceOwner(const Isect& it) const {  // SurfaceOwner (cut.cuh:29-31)
        float ct = d.cut_aware ? cutT[L.slot(it.a, it.b)] : -1.f;
        return skinOwnerA(it.t, ct, it.insideA) ? it.a : it.b;
    }

    // k_RebuildIsectWorld (dc_recon.cu:97-142)
    void isectWorld() {
        iw.assign(d.n_isect, f3(0, 0, 0));
        for (int e = 0; e < d.n_isect; e++) {
            const Isect& it = isect[e];
            int A = it.a, B = it.b;
            bool severed = false;
            if (d.cut_aware) {
                int ax, ay, az, bx, by, bz;
                L.coord(A, ax, ay, az);
                L.coord(B, bx, by, bz);
                int dxi = bx - ax, dyi = by - ay, dzi = bz - az;
                int axis = dxi != 0 ? 0 : (dyi != 0 ? 1 : 2);
                int sgn = (dxi + dyi + dzi) > 0 ? 1 : -1;
                severed = nbr[6 * A + 2 * axis + (sgn > 0 ? 1 : 0)] < 0;
            }
            if (!severed) {
                iw[e] = lerp(P(A), P(B), it.t);
                continue;
            }
            F3 dir = Rst(B) - Rst(A);
            if (surfaceOwner(it) == A) {
                F3 off = mulRv(rot + 9 * (size_t)A, dir);
                iw[e] = lerp(P(A), P(A) + off, it.t);
            } else {
                F3 off = mulRv(rot + 9 * (size_t)B, dir);
                iw[e] = lerp(P(B) - off, P(B), it.t);
            }
        }
    }

    // k_FeaturePoints (dc_recon.cu:149-207), then k_AdvectIntactSurface (dc_recon.cu:477-493) over it
    void featurePoints() {
        fp.assign(4 * (size_t)d.n_voxels, 0.f);
        for (int v = 0; v < d.n_voxels; v++) {
            int off = ioff[v], cnt = icnt[v];
            float* out = &fp[4 * (size_t)v];
            if (cnt > 0) {
                F3 x = f3(0, 0, 0);
                for (int s = 0; s < cnt; s++) x = x + iw[off + s];
                x = x / (float)cnt;
                F3 lo = P(corners[8 * v]), hi = lo;
                for (int c = 1; c < 8; c++) {
                    F3 p = P(corners[8 * v + c]);
                    lo = f3(std::fmin(lo.x, p.x), std::fmin(lo.y, p.y), std::fmin(lo.z, p.z));
                    hi = f3(std::fmax(hi.x, p.x), std::fmax(hi.y, p.y), std::fmax(hi.z, p.z));
                }
                for (int it = 0; it < d.qef_iters; it++) {
                    F3 F = f3(0, 0, 0);
                    for (int s = 0; s < cnt; s++) {
                        F3 n = ld(isect[off + s].normal);
                        F = F + n * dot(n, iw[off + s] - x);
                    }
                    float a = 0.1f * (1.0f - (float)it / (float)d.qef_iters);
                    F3 xn = x + F * a;
                    if (xn.x <= lo.x || xn.y <= lo.y || xn.z <= lo.z || xn.x >= hi.x || xn.y >= hi.y || xn.z >= hi.z)
                        break;
                    x = xn;
                }
                out[0] = x.x; out[1] = x.y; out[2] = x.z; out[3] = 1.f;
            }
            if (!d.have_rest) continue;
            const float* r4 = restFP + 4 * (size_t)v;
            if (r4[3] < .5f || (d.cut_aware && mask[v] != 0u)) continue;
            F3 a = Rst(corners[8 * v]), b = Rst(corners[8 * v + 6]), r = ld(r4);
            F3 uv = f3(std::fmin(1.f, std::fmax(0.f, (r.x - a.x) / (b.x - a.x))),
                       std::fmin(1.f, std::fmax(0.f, (r.y - a.y) / (b.y - a.y))),
                       std::fmin(1.f, std::fmax(0.f, (r.z - a.z) / (b.z - a.z))));
            F3 p = f3(0, 0, 0);
            for (int k = 0; k < 8; k++) {
                int x = k == 1 || k == 2 || k == 5 || k == 6, y = k == 2 || k == 3 || k == 6 || k == 7, z = k >= 4;
                float w = (x ? uv.x : 1 - uv.x) * (y ? uv.y : 1 - uv.y) * (z ? uv.z : 1 - uv.z);
                p = p + P(corners[8 * v + k]) * w;
            }
            out[0] = p.x; out[1] = p.y; out[2] = p.z; out[3] = 1.f;
        }
    }

    // k_ClearCutFP + k_AccumulateCutFP + k_ComputeComponentFP (cut.cu:761-985)
    void componentPoints() {
        const size_t S = 8 * (size_t)d.n_voxels;
        cutFP.assign(4 * S, 0.f);
        if (!d.cut_aware) return;
        std::vector<F3> accPos(S, f3(0, 0, 0));
        std::vector<int> accCnt(S, 0);
        std::vector<uint32_t> accNrm(3 * S, 0u);  // int sums, wrapping like atomicAdd
        for (int i = 0; i < d.n_cut_points; i++) {
            const float* cp = points + 8 * (size_t)i;
            int owner, edge;
            std::memcpy(&owner, cp + 3, 4);
            std::memcpy(&edge, cp + 4, 4);
            if (active[owner] == 0) continue;
            const float* R = rot + 9 * (size_t)owner;
            F3 nrmW = mulRv(R, ld(cp + 5));
            int ox, oy, oz;
            L.coord(owner, ox, oy, oz);
            const int* ge = gedges + 4 * (size_t)edge;
            for (int s = 0; s < 4; s++) {
                int x = ge[1] + STENCIL[ge[0]][s][0], y = ge[2] + STENCIL[ge[0]][s][1], z = ge[3] + STENCIL[ge[0]][s][2];
                if (!L.voxIn(x, y, z)) continue;
                int v = L.voxel(x, y, z);
                if (occupied[v] == 0u) continue;
                uint32_t m = mask[v] & 0xFFFu;
                if (m == 0u) continue;
                int lc = localCornerOf(ox, oy, oz, x, y, z);
                if (lc < 0) continue;
                int c = comp(m, lc);
                size_t slot = 8 * (size_t)v + c;
                int root = 0;
                while (root < 8 && comp(m, root) != c) ++root;
                F3 origin = P(L.corner(x + CORNER[root][0], y + CORNER[root][1], z + CORNER[root][2]));
                F3 local = (P(owner) - origin) + mulRv(R, ld(cp));
                accPos[slot] = accPos[slot] + local;
                accCnt[slot] += 1;
                accNrm[3 * slot + 0] += (uint32_t)(int)std::rint(nrmW.x * POS_SCALE);
                accNrm[3 * slot + 1] += (uint32_t)(int)std::rint(nrmW.y * POS_SCALE);
                accNrm[3 * slot + 2] += (uint32_t)(int)std::rint(nrmW.z * POS_SCALE);
            }
        }
        for (int v = 0; v < d.n_voxels; v++) {
            uint32_t m = mask[v] & 0xFFFu;
            if (m == 0u || occupied[v] == 0u) continue;
            int vxc = v % d.vx, vyc = (v / d.vx) % d.vy, vzc = v / (d.vx * d.vy);
            int compCount = conn[9 * (int)m];
            int off = ioff[v], cnt = icnt[v];
            F3 ext = Rst(corners[8 * v + 6]) - Rst(corners[8 * v]);
            F3 half = f3(std::fabs(ext.x) * .5f, std::fabs(ext.y) * .5f, std::fabs(ext.z) * .5f);
            for (int c = 0; c < compCount; c++) {
                size_t slot = 8 * (size_t)v + c;
                F3 bc = f3(0, 0, 0), origin = f3(0, 0, 0);
                int nCorn = 0;
                for (int lc = 0; lc < 8; lc++) {
                    if (comp(m, lc) != c) continue;
                    F3 p = P(corners[8 * v + lc]);
                    if (nCorn == 0) origin = p;
                    bc = bc + p;
                    nCorn++;
                }
                if (nCorn > 0) bc = bc / (float)nCorn;
                auto mine = [&](const Isect& it) {
                    int ic = surfaceOwner(it), ix, iy, iz;
                    L.coord(ic, ix, iy, iz);
                    int lc = localCornerOf(ix, iy, iz, vxc, vyc, vzc);
                    return lc >= 0 && comp(m, lc) == c;
                };
                F3 sum = f3(0, 0, 0);
                int nIncl = 0;
                for (int s = 0; s < cnt; s++) {
                    if (!mine(isect[off + s])) continue;
                    sum = sum + iw[off + s];
                    nIncl++;
                }
                int ccnt = accCnt[slot];
                if (ccnt > 0) sum = sum + accPos[slot] + origin * (float)ccnt;
                int total = nIncl + ccnt;
                if (total == 0) continue;  // cutFP stays zero
                F3 x = sum / (float)total;
                if (nIncl > 0) {
                    for (int it2 = 0; it2 < d.qef_iters; it2++) {
                        F3 F = f3(0, 0, 0);
                        for (int s = 0; s < cnt; s++) {
                            if (!mine(isect[off + s])) continue;
                            F3 n = ld(isect[off + s].normal);
                            F = F + n * dot(n, iw[off + s] - x);
                        }
                        float a = 0.1f * (1.0f - (float)it2 / (float)d.qef_iters);
                        F3 xn = x + F * a;
                        if (std::fabs(xn.x - bc.x) >= half.x || std::fabs(xn.y - bc.y) >= half.y ||
                            std::fabs(xn.z - bc.z) >= half.z)
                            break;
                        x = xn;
                    }
                }
                x.x = std::fmin(std::fmax(x.x, bc.x - half.x), bc.x + half.x);
                x.y = std::fmin(std::fmax(x.y, bc.y - half.y), bc.y + half.y);
                x.z = std::fmin(std::fmax(x.z, bc.z - half.z), bc.z + half.z);
                if (ccnt > 0) {
                    F3 nAcc = f3((float)(int)accNrm[3 * slot], (float)(int)accNrm[3 * slot + 1],
                                 (float)(int)accNrm[3 * slot + 2]) / POS_SCALE;
                    float nLenW = len(nAcc);
                    if (nLenW > 1e-6f && nLenW > 0.25f * (float)ccnt) {
                        F3 nbar = nAcc / nLenW;
                        F3 cutC = accPos[slot] / (float)ccnt + origin;
                        x = x - nbar * dot(nbar, x - cutC);
                    }
                }
                float* o = &cutFP[4 * slot];
                o[0] = x.x; o[1] = x.y; o[2] = x.z; o[3] = 1.f;
            }
        }
    }

    // StitchVertId (dc_recon.cu:213-225)
    int stitchVert(int x, int y, int z, int ix, int iy, int iz) const {
        int vid = L.voxel(x, y, z);
        if (!d.cut_aware) return vid;
        uint32_t m = mask[vid] & 0xFFFu;
        if (m == 0u || occupied[vid] == 0u) return vid;
        int lc = localCornerOf(ix, iy, iz, x, y, z);
        if (lc < 0) return vid;
        int slot = 8 * vid + comp(m, lc);
        return cutFP[4 * (size_t)slot + 3] > 0.5f ? (int)(0x80000000u | (uint32_t)slot) : vid;
    }

    // Candidate triangles in the fixed order: k_Stitch (dc_recon.cu:239-287), then k_BuildCutTriangles
    // (cut.cu:1023-1090).
    void candidates(std::vector<int>& tri) const {
        auto emit = [&](int a, int b, int c) {
            F3 p0 = fetch(a), p1 = fetch(b), p2 = fetch(c);
            F3 n = cross(p1 - p0, p2 - p0);
            if (dot(n, n) < 1e-16f) return;
            tri.push_back(a); tri.push_back(b); tri.push_back(c);
        };
        for (int e = 0; e < d.n_surface_edges; e++) {
            const int* ge = sedges + 8 * (size_t)e;
            int axis = ge[0], bx = ge[1], by = ge[2], bz = ge[3], insideA = ge[4];
            int ax = axis == 0, ay = axis == 1, az = axis == 2;
            int aId = L.corner(bx, by, bz), bId = L.corner(bx + ax, by + ay, bz + az);
            int es = L.slot(aId, bId);
            int ownerA = skinOwnerA(skinT[es], d.cut_aware ? cutT[es] : -1.f, insideA);
            int ix = ownerA ? bx : bx + ax, iy = ownerA ? by : by + ay, iz = ownerA ? bz : bz + az;
            int v[4];
            for (int s = 0; s < 4; s++)
                v[s] = stitchVert(bx + STENCIL[axis][s][0], by + STENCIL[axis][s][1], bz + STENCIL[axis][s][2], ix, iy,
                                  iz);
            int a = v[0], b = v[1], c = v[2], dd = v[3];
            if (insideA == 0) std::swap(b, dd);
            emit(a, b, c);
            emit(a, c, dd);
        }
        if (!d.cut_aware) return;
        for (int p = 0; p < d.n_nodes; p++) {
            int cx, cy, cz;
            L.coord(p, cx, cy, cz);
            for (int e = 0; e < 6; e++) {
                int axis = e / 2;
                int ax = axis == 0, ay = axis == 1, az = axis == 2;
                bool basedHere = (e & 1) == 0;
                int bx = basedHere ? cx : cx - ax, by = basedHere ? cy : cy - ay, bz = basedHere ? cz : cz - az;
                bool valid = true, isCut = false;
                for (int s = 0; s < 4; s++) {
                    int x = bx + STENCIL[axis][s][0], y = by + STENCIL[axis][s][1], z = bz + STENCIL[axis][s][2];
                    if (!L.voxIn(x, y, z)) { valid = false; break; }
                }
                if (!valid) continue;
                for (int s = 0; s < 4; s++) {
                    int v = L.voxel(bx + STENCIL[axis][s][0], by + STENCIL[axis][s][1], bz + STENCIL[axis][s][2]);
                    if ((mask[v] >> LOCAL_EDGE[axis][s]) & 1u) isCut = true;
                }
                if (!isCut) continue;
                int qx = basedHere ? cx + ax : cx - ax, qy = basedHere ? cy + ay : cy - ay,
                    qz = basedHere ? cz + az : cz - az;
                int other = L.corner(qx, qy, qz), es = L.slot(p, other);
                if (!cutHitsMaterial(cutT[es], skinT[es], inside[p], inside[other])) continue;
                int tag[4] = {0, 0, 0, 0}, n = 0;
                bool ok[4] = {false, false, false, false}, shared[4] = {false, false, false, false};
                for (int s = 0; s < 4; s++) {
                    int x = bx + STENCIL[axis][s][0], y = by + STENCIL[axis][s][1], z = bz + STENCIL[axis][s][2];
                    int v = L.voxel(x, y, z);
                    if (occupied[v] == 0u) continue;
                    uint32_t m = mask[v] & 0xFFFu;
                    if (m == 0u) continue;
                    int lc = localCornerOf(cx, cy, cz, x, y, z);
                    if (lc < 0) continue;
                    int c = comp(m, lc);
                    shared[s] = c == comp(m, localCornerOf(qx, qy, qz, x, y, z));
                    int slot = 8 * v + c;
                    if (cutFP[4 * (size_t)slot + 3] < 0.5f) continue;
                    tag[s] = (int)(0xC0000000u | (uint32_t)slot);
                    ok[s] = true;
                    n++;
                }
                // EmitCutQuad (cut.cu:988-1021)
                if (n < 3) continue;
                int ord[4], mm = 0;
                for (int s = 0; s < 4; s++)
                    if (ok[s]) ord[mm++] = s;
                if (mm == 4 && shared[ord[0]] && shared[ord[2]] && !shared[ord[1]] && !shared[ord[3]]) {
                    int first = ord[0];
                    ord[0] = ord[1]; ord[1] = ord[2]; ord[2] = ord[3]; ord[3] = first;
                }
                F3 q[4];
                for (int k = 0; k < mm; k++) q[k] = ld(&cutFP[4 * (size_t)((uint32_t)tag[ord[k]] & 0x3FFFFFFFu)]);
                bool flip = !basedHere;
                for (int k = 1; k < mm - 1; k++) {
                    if (shared[ord[0]] && shared[ord[k]] && shared[ord[k + 1]]) continue;
                    F3 nn = cross(q[k] - q[0], q[k + 1] - q[0]);
                    if (dot(nn, nn) < 1e-16f) continue;
                    tri.push_back(tag[ord[0]]);
                    tri.push_back(tag[ord[flip ? k + 1 : k]]);
                    tri.push_back(tag[ord[flip ? k : k + 1]]);
                }
            }
        }
    }
};

// faceKey (surface_cleanup.cuh:6-13): the three identities without the wall bit, sorted; odd = an odd permutation.
bool faceKey(const int* t, uint32_t key[3]) {
    for (int k = 0; k < 3; k++) key[k] = (uint32_t)t[k] & 0xbfffffffu;
    bool odd = false;
    if (key[0] > key[1]) { std::swap(key[0], key[1]); odd = !odd; }
    if (key[1] > key[2]) { std::swap(key[1], key[2]); odd = !odd; }
    if (key[0] > key[1]) { std::swap(key[0], key[1]); odd = !odd; }
    return odd;
}

}  // namespace

extern "C" {

struct OracleSurfaceDesc {
    int n_nodes, n_voxels, n_isect, n_surface_edges, n_grid_edges, tri_capacity;
    int vx, vy, vz;
    int qef_iters;
    int cut_aware;
    int have_rest;
    int n_cut_points;
};

// One surface rebuild. Inputs are the lattice state after a step (and its frames), the grid's surface data and
// the rest snapshot of the feature points; outputs the expanded triangle soup (k_ExpandVertices, dc_recon.cu:358-386)
// plus the triangle tags, the external feature points, the component feature points and the index count before the
// cleanup. Returns the index count.
int oracle_surface(const OracleSurfaceDesc* desc, const float* pos, const float* rest, const float* rot, const int* nbr,
                   const int* active, const int* inside, const int* corners, const int* isect_offset,
                   const int* isect_count, const void* isect, const int* surface_edges, const int* conn,
                   const int* grid_edges, const uint32_t* occupied, const uint32_t* cut_mask, const float* edge_cut_t,
                   const float* edge_skin_t, const float* cut_points, const float* rest_fp, float* out_pos,
                   float* out_nrm, float* out_aux, int* out_tags, float* out_fp, float* out_cut_fp,
                   int* out_listed) {
    Desc d;
    std::memcpy(&d, desc, sizeof(Desc));
    static_assert(sizeof(Desc) == sizeof(OracleSurfaceDesc), "descriptor layouts agree");
    Surface S(d);
    S.L = Lat{d.vx, d.vy, d.vz, d.vx + 1, d.vy + 1};
    S.pos = pos; S.rest = rest; S.rot = rot; S.nbr = nbr; S.active = active; S.inside = inside;
    S.corners = corners; S.ioff = isect_offset; S.icnt = isect_count; S.isect = (const Isect*)isect;
    S.sedges = surface_edges; S.conn = conn; S.gedges = grid_edges; S.occupied = occupied; S.mask = cut_mask;
    S.cutT = edge_cut_t; S.skinT = edge_skin_t; S.points = cut_points; S.restFP = rest_fp;

    S.isectWorld();
    S.featurePoints();
    S.componentPoints();
    std::vector<int> cand;
    S.candidates(cand);
    size_t ntri = std::min(cand.size() / 3, (size_t)d.tri_capacity);
    std::vector<int> tri(cand.begin(), cand.begin() + 3 * ntri);
    *out_listed = (int)(3 * ntri);

    if (d.cut_aware && ntri > 0) {  // cleanInternalFaces (surface_cleanup.cuh:35-53)
        struct Entry { int net = 0, minEven = 0x7fffffff, minOdd = 0x7fffffff; };
        std::map<std::array<uint32_t, 3>, Entry> faces;
        std::vector<std::array<uint32_t, 3>> keys(ntri);
        for (size_t i = 0; i < ntri; i++) {
            uint32_t k[3];
            bool odd = faceKey(&tri[3 * i], k);
            keys[i] = {k[0], k[1], k[2]};
            Entry& e = faces[keys[i]];
            e.net += odd ? -1 : 1;
            if (odd) e.minOdd = std::min(e.minOdd, (int)i);
            else e.minEven = std::min(e.minEven, (int)i);
        }
        std::vector<int> kept;
        for (size_t i = 0; i < ntri; i++) {
            const Entry& e = faces[keys[i]];
            if (e.net == 0) continue;
            if ((int)i != (e.net > 0 ? e.minEven : e.minOdd)) continue;
            for (int k = 0; k < 3; k++) kept.push_back(tri[3 * i + k]);
        }
        tri.swap(kept);
        ntri = tri.size() / 3;
    }

    // k_ClearNormals / k_NormalsScatter / k_NormalsNormalize and k_NormalizeCutFP
    const size_t V = d.n_voxels, S8 = 8 * V;
    std::vector<uint32_t> extFix(3 * V, 0u), cutFix(3 * S8, 0u);
    for (size_t t = 0; t < ntri; t++) {
        int raw[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
        F3 p0 = S.fetch(raw[0]), p1 = S.fetch(raw[1]), p2 = S.fetch(raw[2]);
        F3 fn = cross(p1 - p0, p2 - p0);
        for (int k = 0; k < 3; k++) {
            uint32_t u = (uint32_t)raw[k];
            size_t i = u & 0x3FFFFFFFu;
            uint32_t* acc = (d.cut_aware && (u & 0x80000000u)) ? &cutFix[3 * i] : &extFix[3 * i];
            acc[0] += (uint32_t)(int)(fn.x * NORMAL_SCALE);
            acc[1] += (uint32_t)(int)(fn.y * NORMAL_SCALE);
            acc[2] += (uint32_t)(int)(fn.z * NORMAL_SCALE);
        }
    }
    auto normalise = [](const uint32_t* f) {
        F3 n = f3((float)(int)f[0] / NORMAL_SCALE, (float)(int)f[1] / NORMAL_SCALE, (float)(int)f[2] / NORMAL_SCALE);
        float l = len(n);
        return l > 1e-6f ? n / l : f3(0, 0, 0);
    };

    // k_ExpandVertices (dc_recon.cu:358-386)
    const int count = (int)(3 * ntri);
    for (int i = 0; i < count; i++) {
        int raw = tri[i];
        uint32_t u = (uint32_t)raw;
        size_t idx = u & 0x3FFFFFFFu;
        F3 p, n;
        float aux[4];
        if (d.cut_aware && (u & 0x80000000u)) {
            p = ld(&S.cutFP[4 * idx]);
            n = normalise(&cutFix[3 * idx]);
            size_t v = idx >> 3;
            const float* r4 = rest_fp + 4 * v;
            F3 anchor = r4[3] > 0.f ? ld(r4) : (S.Rst(corners[8 * v]) + S.Rst(corners[8 * v + 6])) * .5f;
            aux[0] = anchor.x; aux[1] = anchor.y; aux[2] = anchor.z; aux[3] = (float)((u >> 30) & 1u);
        } else {
            p = ld(&S.fp[4 * idx]);
            n = normalise(&extFix[3 * idx]);
            const float* r4 = rest_fp + 4 * idx;
            aux[0] = r4[0]; aux[1] = r4[1]; aux[2] = r4[2]; aux[3] = 0.f;
        }
        out_pos[3 * i] = p.x; out_pos[3 * i + 1] = p.y; out_pos[3 * i + 2] = p.z;
        out_nrm[3 * i] = n.x; out_nrm[3 * i + 1] = n.y; out_nrm[3 * i + 2] = n.z;
        std::memcpy(out_aux + 4 * i, aux, sizeof(aux));
        out_tags[i] = raw;
    }
    std::memcpy(out_fp, S.fp.data(), 4 * V * sizeof(float));
    std::memcpy(out_cut_fp, S.cutFP.data(), 4 * S8 * sizeof(float));
    return count;
}

}  // extern "C"
