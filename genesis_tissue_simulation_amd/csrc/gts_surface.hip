// Surface reconstruction of the deformed, cut lattice on gfx950: the dual-contouring rebuild (recon_build_static,
// dc_recon.cu:498-551), the cut feature points and cut walls (cut.cu:761-1100) and the internal-face cleanup
// (surface_cleanup.cuh:1-53), as a chain of element-parallel kernels on one stream.
//
// Layout and order:
//   - intersections, voxels, surface edges and nodes each get one thread; everything is gathered, nothing is
//     scattered with float atomics;
//   - the cut points of a (voxel, component) are summed by the voxel's thread in cut point index order, from a
//     per-edge table of the (at most two) cut points of each cut edge: MarkAndEmit (cut.cu:211-246) emits an edge's
//     points once, on its first cut;
//   - candidate triangles have fixed slots (two per surface edge, then twelve per node) and are compacted in slot
//     order by a wave-ballot scan, so the triangle list is deterministic; the face cleanup keeps each face's first
//     listed representative of the majority orientation and compacts in index order;
//   - normals are summed in fixed point with integer atomics, as the reference does, so their order does not matter.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/gts_xpbd.h"
#include "gts_common.h"
#include "gts_contact.h"
#include "gts_lattice.h"

#define HIP_TRY(call)                          \
    do {                                       \
        hipError_t _e = (call);                \
        if (_e != hipSuccess) return -(int)_e; \
    } while (0)

namespace {

constexpr float kNormalScale = 1048576.0f;  // NORMAL_SCALE (common.cuh:45)
constexpr float kPosScale = 16384.0f;       // POS_SCALE (cut.cuh:12)
constexpr uint32_t kCutTag = 0x80000000u;   // a component feature point (dc_recon.cu:224)
constexpr uint32_t kWallTag = 0x40000000u;  // emitted by the cut walls (cut.cu:1077-1081)
constexpr uint32_t kIndexMask = 0x3FFFFFFFu;

inline unsigned blocks_for(long long n) { return (unsigned)std::max(1ll, (n + GTS_BLOCK - 1) / GTS_BLOCK); }

struct Grid {
    int vx, vy, vz, nx, ny;
    __device__ int corner(int x, int y, int z) const { return x + nx * (y + ny * z); }
    __device__ int voxel(int x, int y, int z) const { return x + vx * (y + vy * z); }
    __device__ void coord(int id, int& x, int& y, int& z) const {
        x = id % nx;
        y = (id / nx) % ny;
        z = id / (nx * ny);
    }
    __device__ bool vox_in(int x, int y, int z) const {
        return x >= 0 && y >= 0 && z >= 0 && x < vx && y < vy && z < vz;
    }
    __device__ int slot(int a, int b) const { return directed_slot(a, b, vx); }
};

__device__ __forceinline__ int ibits(float f) { return __float_as_int(f); }
__device__ __forceinline__ f3 lerp3(f3 a, f3 b, float t) {
    return mk3(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t);
}
__device__ __forceinline__ f3 div3s(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ f3 mul_rv(const float* R, f3 v) {  // mulRv (cut.cuh:70-73)
    return mk3(dot3(ld3(R), v), dot3(ld3(R + 3), v), dot3(ld3(R + 6), v));
}
__device__ __forceinline__ f3 xyz4(const float4* p, int i) {
    float4 q = p[i];
    return mk3(q.x, q.y, q.z);
}
// LocalCornerOf (cut.cuh:101-109): figure-order corner of voxel (vx, vy, vz) at node (cx, cy, cz), -1 if none.
__device__ __forceinline__ int local_corner_of(int cx, int cy, int cz, int vx, int vy, int vz) {
    int dx = cx - vx, dy = cy - vy, dz = cz - vz;
    if ((dx | dy | dz) & ~1) return -1;
    return 4 * dz + (dy ? (dx ? 2 : 3) : dx);
}
// The i-th owner of an IsectGpu record (cut.cuh:29-31 SurfaceOwner), with the cut fraction of the edge when cut-aware.
__device__ __forceinline__ int surface_owner(const float* rec, const Grid& G, const float* cutT) {
    int a = ibits(rec[6]), b = ibits(rec[7]);
    float ct = cutT ? cutT[G.slot(a, b)] : -1.f;
    return skin_owner_a(rec[8], ct, ibits(rec[9])) ? a : b;
}

// ---- k_RebuildIsectWorld (dc_recon.cu:97-142): an intact edge lerps its deformed ends; a severed one rebuilds the
// missing end from the skin owner's frame and the rest edge. nbr is null outside the cut-aware rebuild.
__device__ __forceinline__ f3 isect_world(const float* rec, const Grid& G, const float4* pos, const float* rest,
                                          const int* nbr, const float* rot, const float* cutT) {
    const int A = ibits(rec[6]), B = ibits(rec[7]);
    const float t = rec[8];
    bool severed = false;
    if (nbr) {
        int ax, ay, az, bx, by, bz;
        G.coord(A, ax, ay, az);
        G.coord(B, bx, by, bz);
        int dxi = bx - ax, dyi = by - ay, dzi = bz - az;
        int axis = dxi != 0 ? 0 : (dyi != 0 ? 1 : 2);
        int sgn = (dxi + dyi + dzi) > 0 ? 1 : -1;
        severed = nbr[6 * A + 2 * axis + (sgn > 0 ? 1 : 0)] < 0;
    }
    if (!severed) return lerp3(xyz4(pos, A), xyz4(pos, B), t);
    f3 dir = ld3(rest + 3 * (size_t)B) - ld3(rest + 3 * (size_t)A);
    if (surface_owner(rec, G, cutT) == A) {
        f3 pA = xyz4(pos, A);
        return lerp3(pA, pA + mul_rv(rot + 9 * (size_t)A, dir), t);
    }
    f3 pB = xyz4(pos, B);
    return lerp3(pB - mul_rv(rot + 9 * (size_t)B, dir), pB, t);
}

// ---- One thread per voxel: k_FeaturePoints (dc_recon.cu:149-207) over k_RebuildIsectWorld of the voxel's own
// intersections (each voxel owns a contiguous range of the list), and k_AdvectIntactSurface (dc_recon.cu:477-493):
// a voxel whose rest feature point exists and that no cut touches carries it trilinearly with its corners; any other
// voxel with intersections runs the decaying-step QEF ascent from their mean, stopping before it leaves the corners'
// box. The voxel's fixed-point normal sum is cleared for the scatter that follows. In the cut-aware rebuild, the
// threads also enter every cut point in the table of its edge's cut points (slot 2e for the owner at the edge's base
// node, 2e + 1 for the other end); the table persists between rebuilds, as the cut point list only grows.
__global__ void __launch_bounds__(GTS_BLOCK) k_voxel_pass(int V, int iters, int cut_aware, int have_rest, Grid G, const int* __restrict__ corners,
                             const int* __restrict__ ioff, const int* __restrict__ icnt,
                             const float* __restrict__ isect, float* __restrict__ iw, const float4* __restrict__ pos,
                             const float* __restrict__ rest, const int* __restrict__ nbr, const float* __restrict__ rot,
                             const float* __restrict__ cutT, const float4* __restrict__ rest_fp,
                             const uint32_t* __restrict__ mask, float4* __restrict__ fp, int* __restrict__ fixed,
                             int capacity, const uint32_t* __restrict__ counters, const int* __restrict__ gedges,
                             const float* __restrict__ points, int* __restrict__ edge_points) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (edge_points) {
        const int n = (int)min(counters[1], (uint32_t)capacity);
        for (int i = v; i < n; i += gridDim.x * blockDim.x) {
            const int owner = ibits(points[8 * (size_t)i + 3]), e = ibits(points[8 * (size_t)i + 4]);
            const int4 ge = reinterpret_cast<const int4*>(gedges)[e];
            edge_points[2 * e + (owner == G.corner(ge.y, ge.z, ge.w) ? 0 : 1)] = i;
        }
    }
    if (v >= V) return;
    fixed[3 * (size_t)v] = 0;
    fixed[3 * (size_t)v + 1] = 0;
    fixed[3 * (size_t)v + 2] = 0;
    const int* cv = corners + 8 * (size_t)v;
    if (have_rest) {
        float4 r4 = rest_fp[v];
        if (!(r4.w < .5f) && !(cut_aware && mask[v] != 0u)) {
            f3 a = ld3(rest + 3 * (size_t)cv[0]), b = ld3(rest + 3 * (size_t)cv[6]);
            f3 uv = mk3(fminf(1.f, fmaxf(0.f, (r4.x - a.x) / (b.x - a.x))),
                        fminf(1.f, fmaxf(0.f, (r4.y - a.y) / (b.y - a.y))),
                        fminf(1.f, fmaxf(0.f, (r4.z - a.z) / (b.z - a.z))));
            f3 p = mk3(0.f, 0.f, 0.f);
            for (int k = 0; k < 8; k++) {
                const int x = k == 1 || k == 2 || k == 5 || k == 6, y = k == 2 || k == 3 || k == 6 || k == 7, z = k >= 4;
                float w = (x ? uv.x : 1 - uv.x) * (y ? uv.y : 1 - uv.y) * (z ? uv.z : 1 - uv.z);
                p = p + xyz4(pos, cv[k]) * w;
            }
            fp[v] = make_float4(p.x, p.y, p.z, 1.f);
            return;
        }
    }
    const int off = ioff[v], cnt = icnt[v];
    if (cnt == 0) {
        fp[v] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    // the voxel's intersections (at most 12, one per edge) in registers: deformed point and rest normal
    f3 w[12], nn[12];
#pragma unroll
    for (int s = 0; s < 12; s++) {
        if (s >= cnt) continue;
        const float* rec = isect + 10 * (size_t)(off + s);
        w[s] = isect_world(rec, G, pos, rest, nbr, rot, cutT);
        nn[s] = ld3(rec + 3);
        float* o = iw + 3 * (size_t)(off + s);
        o[0] = w[s].x;
        o[1] = w[s].y;
        o[2] = w[s].z;
    }
    f3 x = mk3(0.f, 0.f, 0.f);
#pragma unroll
    for (int s = 0; s < 12; s++)
        if (s < cnt) x = x + w[s];
    x = div3s(x, (float)cnt);
    f3 lo = xyz4(pos, cv[0]), hi = lo;
    for (int c = 1; c < 8; c++) {
        f3 p = xyz4(pos, cv[c]);
        lo = mk3(fminf(lo.x, p.x), fminf(lo.y, p.y), fminf(lo.z, p.z));
        hi = mk3(fmaxf(hi.x, p.x), fmaxf(hi.y, p.y), fmaxf(hi.z, p.z));
    }
    for (int it = 0; it < iters; it++) {
        f3 F = mk3(0.f, 0.f, 0.f);
#pragma unroll
        for (int s = 0; s < 12; s++)
            if (s < cnt) F = F + nn[s] * dot3(nn[s], w[s] - x);
        float a = 0.1f * (1.0f - (float)it / (float)iters);
        f3 xn = x + F * a;
        if (xn.x <= lo.x || xn.y <= lo.y || xn.z <= lo.z || xn.x >= hi.x || xn.y >= hi.y || xn.z >= hi.z) break;
        x = xn;
    }
    fp[v] = make_float4(x.x, x.y, x.z, 1.f);
}

// Ascending sort of 32 ints held in registers (a bitonic network: every index is a compile-time constant).
__device__ __forceinline__ void sort32(int v[32]) {
#pragma unroll
    for (int k = 2; k <= 32; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int i = 0; i < 32; i++) {
                const int l = i ^ j;
                if (l > i) {
                    const int a = v[i], b = v[l];
                    const bool up = (i & k) == 0;
                    v[i] = up ? min(a, b) : max(a, b);
                    v[l] = up ? max(a, b) : min(a, b);
                }
            }
        }
    }
}

// ---- k_AccumulateCutFP + k_ComputeComponentFP (cut.cu:786-985), one thread per voxel: each component of a cut,
// occupied voxel gets one feature point from its intersections and its cut points (restored from the owners' frames,
// summed in cut point index order), refined like the voxel feature point inside the component's box, clamped to it
// and pulled onto the plane of its own cut points. The voxel's other slots are cleared (k_ClearCutFP, cut.cu:761-772);
// the fixed-point normal sums of its slots are cleared for the scatter. The voxel's intersections (at most 12, one
// per edge) and cut points (at most 24, two per edge) are loaded into registers once, in independent batches, so the
// ascent and the sums run without memory round trips.
__global__ void __launch_bounds__(GTS_BLOCK) k_component_points(int V, Grid G, int iters, const int* __restrict__ corners,
                                   const int* __restrict__ ioff, const int* __restrict__ icnt,
                                   const float* __restrict__ isect, const float* __restrict__ iw,
                                   const float4* __restrict__ pos, const float* __restrict__ rest,
                                   const float* __restrict__ rot, const int* __restrict__ active,
                                   const uint32_t* __restrict__ occupied, const uint32_t* __restrict__ mask,
                                   const int* __restrict__ conn, const float* __restrict__ cutT,
                                   const int* __restrict__ edge_of, const int* __restrict__ edge_points,
                                   const float* __restrict__ points, float4* __restrict__ cut_fp,
                                   int* __restrict__ fixed) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint32_t m = mask[v] & 0xFFFu;
    // A voxel's cut mask only grows and cut_fp starts zeroed, so the slots of a voxel that is not cut and occupied
    // have never been written: they are already the cleared slots k_ClearCutFP would write.
    if (m == 0u || occupied[v] == 0u) return;
    float4* out = cut_fp + 8 * (size_t)v;
    const int compCount = conn[9 * m];
    for (int c = compCount; c < 8; c++) out[c] = make_float4(0.f, 0.f, 0.f, 0.f);
    // the normal sums of the slots a triangle can name (a cut, occupied voxel's), cleared for the scatter
    int* fx = fixed + 3 * ((size_t)V + 8 * (size_t)v);
    for (int k = 0; k < 24; k++) fx[k] = 0;
    const int* row = conn + 9 * m + 1;
    int cv[8];
#pragma unroll
    for (int k = 0; k < 8; k++) cv[k] = corners[8 * (size_t)v + k];
    const int vxc = v % G.vx, vyc = (v / G.vx) % G.vy, vzc = v / (G.vx * G.vy);

    // The voxel's cut points, ascending: two table slots per edge of the voxel.
    const int8_t ec[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                              {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    int eid[12];
#pragma unroll
    for (int l = 0; l < 12; l++) {
        const int p0 = cv[ec[l][0]], p1 = cv[ec[l][1]];
        const int lo = min(p0, p1), d = max(p0, p1) - lo;
        eid[l] = edge_of[3 * lo + (d == 1 ? 0 : (d == G.nx ? 1 : 2))];
    }
    int list[32];
#pragma unroll
    for (int k = 0; k < 32; k++) list[k] = 0x7fffffff;
#pragma unroll
    for (int l = 0; l < 12; l++) {
        if (eid[l] < 0) continue;
        const int a = edge_points[2 * eid[l]], b = edge_points[2 * eid[l] + 1];
        if (a >= 0) list[2 * l] = a;
        if (b >= 0) list[2 * l + 1] = b;
    }
    sort32(list);
    // owner, then its component in this voxel (-1: not counted), for each listed cut point
    int own[24], cpc[24];
#pragma unroll
    for (int k = 0; k < 24; k++) own[k] = list[k] != 0x7fffffff ? ibits(points[8 * (size_t)list[k] + 3]) : -1;
#pragma unroll
    for (int k = 0; k < 24; k++) {
        cpc[k] = -1;
        if (own[k] < 0 || active[own[k]] == 0) continue;
        int ox, oy, oz;
        G.coord(own[k], ox, oy, oz);
        const int lc = local_corner_of(ox, oy, oz, vxc, vyc, vzc);
        if (lc >= 0) cpc[k] = row[lc];
    }

    // the intersections: deformed point, rest normal and the component of their skin owner (-1: none)
    const int off = ioff[v], cnt = icnt[v];
    f3 w[12], nn[12];
    int ic[12];
#pragma unroll
    for (int s = 0; s < 12; s++) {
        ic[s] = -1;
        if (s >= cnt) continue;
        const float* rec = isect + 10 * (size_t)(off + s);
        w[s] = ld3(iw + 3 * (size_t)(off + s));
        nn[s] = ld3(rec + 3);
        int ix, iy, iz;
        G.coord(surface_owner(rec, G, cutT), ix, iy, iz);
        const int lc = local_corner_of(ix, iy, iz, vxc, vyc, vzc);
        if (lc >= 0) ic[s] = row[lc];
    }

    const f3 ext = ld3(rest + 3 * (size_t)cv[6]) - ld3(rest + 3 * (size_t)cv[0]);
    const f3 half = mk3(fabsf(ext.x) * .5f, fabsf(ext.y) * .5f, fabsf(ext.z) * .5f);
    for (int c = 0; c < compCount; c++) {
        f3 bc = mk3(0.f, 0.f, 0.f), origin = mk3(0.f, 0.f, 0.f);
        int nCorn = 0;
#pragma unroll
        for (int lc = 0; lc < 8; lc++) {
            if (row[lc] != c) continue;
            f3 p = xyz4(pos, cv[lc]);
            if (nCorn == 0) origin = p;
            bc = bc + p;
            nCorn++;
        }
        if (nCorn > 0) bc = div3s(bc, (float)nCorn);

        f3 accPos = mk3(0.f, 0.f, 0.f);
        int ccnt = 0;
        uint32_t nacc[3] = {0u, 0u, 0u};
#pragma unroll
        for (int k = 0; k < 24; k++) {
            if (cpc[k] != c) continue;
            const float* cp = points + 8 * (size_t)list[k];
            const float* R = rot + 9 * (size_t)own[k];
            f3 nrmW = mul_rv(R, ld3(cp + 5));
            f3 local = (xyz4(pos, own[k]) - origin) + mul_rv(R, ld3(cp));
            accPos = accPos + local;
            ccnt++;
            nacc[0] += (uint32_t)(int)rintf(nrmW.x * kPosScale);
            nacc[1] += (uint32_t)(int)rintf(nrmW.y * kPosScale);
            nacc[2] += (uint32_t)(int)rintf(nrmW.z * kPosScale);
        }

        f3 sum = mk3(0.f, 0.f, 0.f);
        int nIncl = 0;
#pragma unroll
        for (int s = 0; s < 12; s++) {
            if (ic[s] != c) continue;
            sum = sum + w[s];
            nIncl++;
        }
        if (ccnt > 0) sum = sum + accPos + origin * (float)ccnt;
        const int total = nIncl + ccnt;
        if (total == 0) {
            out[c] = make_float4(0.f, 0.f, 0.f, 0.f);
            continue;
        }
        f3 x = div3s(sum, (float)total);
        if (nIncl > 0) {
            for (int it = 0; it < iters; it++) {
                f3 F = mk3(0.f, 0.f, 0.f);
#pragma unroll
                for (int s = 0; s < 12; s++)
                    if (ic[s] == c) F = F + nn[s] * dot3(nn[s], w[s] - x);
                float a = 0.1f * (1.0f - (float)it / (float)iters);
                f3 xn = x + F * a;
                if (fabsf(xn.x - bc.x) >= half.x || fabsf(xn.y - bc.y) >= half.y || fabsf(xn.z - bc.z) >= half.z) break;
                x = xn;
            }
        }
        x.x = fminf(fmaxf(x.x, bc.x - half.x), bc.x + half.x);
        x.y = fminf(fmaxf(x.y, bc.y - half.y), bc.y + half.y);
        x.z = fminf(fmaxf(x.z, bc.z - half.z), bc.z + half.z);
        if (ccnt > 0) {
            f3 nAcc = div3s(mk3((float)(int)nacc[0], (float)(int)nacc[1], (float)(int)nacc[2]), kPosScale);
            float nLenW = len3(nAcc);
            if (nLenW > 1e-6f && nLenW > 0.25f * (float)ccnt) {
                f3 nbar = div3s(nAcc, nLenW);
                f3 cutC = div3s(accPos, (float)ccnt) + origin;
                x = x - nbar * dot3(nbar, x - cutC);
            }
        }
        out[c] = make_float4(x.x, x.y, x.z, 1.f);
    }
}

// FetchPos (dc_recon.cu:228-234)
__device__ __forceinline__ f3 fetch_pos(int raw, const float4* fp, const float4* cut_fp) {
    uint32_t u = (uint32_t)raw;
    int idx = (int)(u & kIndexMask);
    float4 q = (u & kCutTag) ? cut_fp[idx] : fp[idx];
    return mk3(q.x, q.y, q.z);
}

// A candidate triangle: kept (returns 1, and its tags go to slot) unless its three points are collinear or coincide
// (k_Stitch / EmitCutQuad, dc_recon.cu:277-285, cut.cu:1013-1015).
__device__ __forceinline__ int candidate(int slot, int a, int b, int c, const float4* fp, const float4* cut_fp,
                                         int* cand) {
    f3 p0 = fetch_pos(a, fp, cut_fp), p1 = fetch_pos(b, fp, cut_fp), p2 = fetch_pos(c, fp, cut_fp);
    f3 n = cross3(p1 - p0, p2 - p0);
    if (dot3(n, n) < 1e-16f) return 0;
    cand[3 * (size_t)slot] = a;
    cand[3 * (size_t)slot + 1] = b;
    cand[3 * (size_t)slot + 2] = c;
    return 1;
}

// ---- k_Stitch (dc_recon.cu:239-287) with StitchVertId (:213-225): the quad of the four voxels around a surface edge,
// each voxel resolved to the feature point of the component holding the edge's skin owner when the voxel is cut.
__device__ __forceinline__ int stitch_item(int e, const Grid& G, int cut_aware, const int* __restrict__ sedges,
                                           const float* __restrict__ skinT, const float* __restrict__ cutT,
                                           const uint32_t* __restrict__ mask, const uint32_t* __restrict__ occupied,
                                           const int* __restrict__ conn, const float4* __restrict__ fp,
                                           const float4* __restrict__ cut_fp, int* __restrict__ cand) {
    const int* ge = sedges + 8 * (size_t)e;
    const int axis = ge[0], bx = ge[1], by = ge[2], bz = ge[3], insideA = ge[4];
    const int ax = axis == 0, ay = axis == 1, az = axis == 2;
    const int aId = G.corner(bx, by, bz), bId = G.corner(bx + ax, by + ay, bz + az);
    const int es = G.slot(aId, bId);
    const int ownerA = skin_owner_a(skinT[es], cut_aware ? cutT[es] : -1.f, insideA);
    const int ix = ownerA ? bx : bx + ax, iy = ownerA ? by : by + ay, iz = ownerA ? bz : bz + az;
    int vt[4];
    for (int s = 0; s < 4; s++) {
        int dx, dy, dz, local;
        stencil(axis, s, dx, dy, dz, local);
        const int x = bx + dx, y = by + dy, z = bz + dz;
        const int vid = G.voxel(x, y, z);
        vt[s] = vid;
        if (!cut_aware) continue;
        const uint32_t m = mask[vid] & 0xFFFu;
        if (m == 0u || occupied[vid] == 0u) continue;
        const int lc = local_corner_of(ix, iy, iz, x, y, z);
        if (lc < 0) continue;
        const int slot = 8 * vid + conn[9 * m + 1 + lc];
        if (cut_fp[slot].w > 0.5f) vt[s] = (int)(kCutTag | (uint32_t)slot);
    }
    int a = vt[0], b = vt[1], c = vt[2], d = vt[3];
    if (insideA == 0) {
        int t = b;
        b = d;
        d = t;
    }
    return candidate(2 * e, a, b, c, fp, cut_fp, cand) | candidate(2 * e + 1, a, c, d, fp, cut_fp, cand) << 1;
}

// ---- k_BuildCutTriangles with EmitCutQuad (cut.cu:988-1090): around every cut edge of a node that cuts material,
// the feature points of the node's component in the (up to four) cut voxels form the wall; a fan of at most two
// triangles, skipping faces made only of points the two ends share. Item j is edge j % 6 (-x, +x, -y, +y, -z, +z of
// the node's links, as the reference's loop visits them) of node j / 6; its slots are base + 2j + k; returns the mask
// of the slots filled.
__device__ __forceinline__ int wall_item(int j, int base, const Grid& G, const uint32_t* __restrict__ mask,
                                         const uint32_t* __restrict__ occupied, const int* __restrict__ conn,
                                         const float* __restrict__ cutT, const float* __restrict__ skinT,
                                         const int* __restrict__ inside, const float4* __restrict__ fp,
                                         const float4* __restrict__ cut_fp, int* __restrict__ cand) {
    int filled = 0;
    const int p = j / 6, e = j - 6 * p;
    int cx, cy, cz;
    G.coord(p, cx, cy, cz);
    {
        const int axis = e >> 1;
        const int ax = axis == 0, ay = axis == 1, az = axis == 2;
        const bool basedHere = (e & 1) == 0;
        const int bx = basedHere ? cx : cx - ax, by = basedHere ? cy : cy - ay, bz = basedHere ? cz : cz - az;
        bool valid = true, isCut = false;
        int vids[4];
        for (int s = 0; s < 4; s++) {
            int dx, dy, dz, local;
            stencil(axis, s, dx, dy, dz, local);
            if (!G.vox_in(bx + dx, by + dy, bz + dz)) {
                valid = false;
                break;
            }
            vids[s] = G.voxel(bx + dx, by + dy, bz + dz);
            if ((mask[vids[s]] >> local) & 1u) isCut = true;
        }
        if (!valid || !isCut) return filled;
        const int qx = basedHere ? cx + ax : cx - ax, qy = basedHere ? cy + ay : cy - ay,
                  qz = basedHere ? cz + az : cz - az;
        const int other = G.corner(qx, qy, qz), es = G.slot(p, other);
        if (!cut_hits_material(cutT[es], skinT[es], inside[p], inside[other])) return filled;
        int tag[4] = {0, 0, 0, 0}, n = 0;
        bool okq[4] = {false, false, false, false}, shared[4] = {false, false, false, false};
        for (int s = 0; s < 4; s++) {
            int dx, dy, dz, local;
            stencil(axis, s, dx, dy, dz, local);
            const int x = bx + dx, y = by + dy, z = bz + dz, v = vids[s];
            if (occupied[v] == 0u) continue;
            const uint32_t m = mask[v] & 0xFFFu;
            if (m == 0u) continue;
            const int lc = local_corner_of(cx, cy, cz, x, y, z);
            if (lc < 0) continue;
            const int c = conn[9 * m + 1 + lc];
            shared[s] = c == conn[9 * m + 1 + local_corner_of(qx, qy, qz, x, y, z)];
            const int slot = 8 * v + c;
            if (cut_fp[slot].w < 0.5f) continue;
            tag[s] = (int)(kCutTag | kWallTag | (uint32_t)slot);
            okq[s] = true;
            n++;
        }
        if (n < 3) return filled;
        int ord[4], mm = 0;
        for (int s = 0; s < 4; s++)
            if (okq[s]) ord[mm++] = s;
        if (mm == 4 && shared[ord[0]] && shared[ord[2]] && !shared[ord[1]] && !shared[ord[3]]) {
            int first = ord[0];
            ord[0] = ord[1];
            ord[1] = ord[2];
            ord[2] = ord[3];
            ord[3] = first;
        }
        const bool flip = !basedHere;
        for (int k = 1; k < mm - 1; k++) {
            if (shared[ord[0]] && shared[ord[k]] && shared[ord[k + 1]]) continue;
            filled |= candidate(base + 2 * j + (k - 1), tag[ord[0]], tag[ord[flip ? k + 1 : k]],
                                tag[ord[flip ? k : k + 1]], fp, cut_fp, cand) << (k - 1);
        }
    }
    return filled;
}

// Candidates of every item (surface edge e < Se, then node edge Se + 6p + e in the cut-aware rebuild) into their two
// fixed slots 2 * item + k, the mask of the filled ones per item, and the count per workgroup.
__global__ void k_triangles(int Se, int n_items, Grid G, int cut_aware, const int* __restrict__ sedges,
                            const float* __restrict__ skinT, const float* __restrict__ cutT,
                            const uint32_t* __restrict__ mask, const uint32_t* __restrict__ occupied,
                            const int* __restrict__ conn, const int* __restrict__ inside,
                            const float4* __restrict__ fp, const float4* __restrict__ cut_fp, int* __restrict__ cand,
                            int* __restrict__ item_mask, int* __restrict__ partial) {
    __shared__ int lds[GTS_BLOCK / 64];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int m = 0;
    if (i < Se) m = stitch_item(i, G, cut_aware, sedges, skinT, cutT, mask, occupied, conn, fp, cut_fp, cand);
    else if (i < n_items) m = wall_item(i - Se, 2 * Se, G, mask, occupied, conn, cutT, skinT, inside, fp, cut_fp, cand);
    if (i < n_items) item_mask[i] = m;
    int total;
    block_scan(__popc(m), lds, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Smallest power of two at least twice the count (at least 1): the face table the cleanup of `count` triangles uses.
__device__ __forceinline__ int face_slots(int count) {
    int fc = 1;
    while (fc < 2 * count) fc <<= 1;
    return fc;
}

// k_NormalsScatter (dc_recon.cu:306-331) of one triangle: its area-weighted face normal, in fixed point, to the sums
// of its three points (voxel points at 3 * voxel, component points at 3 * (V + slot)).
__device__ __forceinline__ void scatter_normal(const int* t, const float4* fp, const float4* cut_fp, int V,
                                               int* fixed) {
    f3 p0 = fetch_pos(t[0], fp, cut_fp), p1 = fetch_pos(t[1], fp, cut_fp), p2 = fetch_pos(t[2], fp, cut_fp);
    f3 fn = cross3(p1 - p0, p2 - p0);
    const int q[3] = {(int)(fn.x * kNormalScale), (int)(fn.y * kNormalScale), (int)(fn.z * kNormalScale)};
    for (int k = 0; k < 3; k++) {
        uint32_t u = (uint32_t)t[k];
        size_t idx = u & kIndexMask;
        int* acc = (u & kCutTag) ? fixed + 3 * ((size_t)V + idx) : fixed + 3 * idx;
        atomicAdd(acc, q[0]);
        atomicAdd(acc + 1, q[1]);
        atomicAdd(acc + 2, q[2]);
    }
}

// Ordered placement of the filled candidates: item i's triangles, in slot order, from the scanned workgroup counts;
// those past the capacity are dropped. Thread 0 records the counts (counts[1] listed, counts[2] asked for; counts[0]
// too when no cleanup follows, in which case the normals are scattered here). With a face table, the threads then
// clear the entries the cleanup of the listed triangles will use.
__global__ void k_place(int Se, int n_items, const int* __restrict__ item_mask, const int* __restrict__ partial,
                        int nblocks, const int* __restrict__ cand, int* __restrict__ out, int capacity,
                        uint32_t* __restrict__ counts, int final_list, const float4* __restrict__ fp,
                        const float4* __restrict__ cut_fp, int V, int* __restrict__ fixed,
                        int4* __restrict__ face_table) {
    __shared__ int lds[GTS_BLOCK / 64];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = i < n_items ? item_mask[i] : 0;
    int total;
    int k = partial[blockIdx.x] + block_scan(__popc(m), lds, total);
    const int base = 2 * i;
    for (int j = 0; m >> j; j++) {
        if (!((m >> j) & 1)) continue;
        if (k < capacity) {
            const int* c = cand + 3 * (size_t)(base + j);
            out[3 * k] = c[0];
            out[3 * k + 1] = c[1];
            out[3 * k + 2] = c[2];
            if (final_list) scatter_normal(c, fp, cut_fp, V, fixed);
        }
        k++;
    }
    const int asked = partial[nblocks], listed = min(asked, capacity);
    if (i == 0) {
        counts[1] = 3u * (uint32_t)listed;
        counts[2] = 3u * (uint32_t)asked;
        if (final_list) counts[0] = 3u * (uint32_t)listed;
    }
    if (face_table) {
        const int fc = face_slots(listed);
        for (int t = i; t < fc; t += gridDim.x * blockDim.x) face_table[t] = make_int4(-1, 0, 0x7fffffff, 0x7fffffff);
    }
}

// ---- cleanInternalFaces (surface_cleanup.cuh:6-53): faces keyed by their three identities without the wall bit;
// an even and an odd listing of one key cancel.
__device__ __forceinline__ bool face_key(const int* t, uint32_t key[3]) {
    for (int k = 0; k < 3; k++) key[k] = (uint32_t)t[k] & ~kWallTag;
    bool odd = false;
    uint32_t s;
    if (key[0] > key[1]) { s = key[0]; key[0] = key[1]; key[1] = s; odd = !odd; }
    if (key[1] > key[2]) { s = key[1]; key[1] = key[2]; key[2] = s; odd = !odd; }
    if (key[0] > key[1]) { s = key[0]; key[0] = key[1]; key[1] = s; odd = !odd; }
    return odd;
}
__device__ __forceinline__ uint32_t face_hash(const uint32_t key[3], int fc) {
    return ((key[0] * 73856093u) ^ (key[1] * 19349663u) ^ (key[2] * 83492791u)) & (uint32_t)(fc - 1);
}

// Only triangles with a component point (bit 31) can share a key. A triangle of three voxel points comes from one
// stitch quad, the four voxels around its surface edge; two distinct surface edges share at most two of those voxels,
// so no two such triangles have one key, and their keys never equal one that holds a component point. Each of them
// is therefore its own key's lowest listing with a net count of one, and is kept without being indexed.
__device__ __forceinline__ bool names_component(const int* t) {
    return (((uint32_t)t[0] | (uint32_t)t[1] | (uint32_t)t[2]) & kCutTag) != 0u;
}

// Entry: x = first triangle of the key, y = even minus odd listings, z / w = lowest even / odd listing. The table
// holds face_slots(listed) entries, cleared by k_place.
__global__ void k_face_index(int capacity, const uint32_t* __restrict__ counts, const int* __restrict__ tri,
                             int4* table) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = (int)(counts[1] / 3u);
    if (i >= capacity || i >= n || !names_component(tri + 3 * i)) return;
    const int fc = face_slots(n);
    uint32_t key[3];
    const bool odd = face_key(tri + 3 * i, key);
    const uint32_t h = face_hash(key, fc);
    for (int k = 0; k < fc; k++) {
        int4* entry = table + ((h + (uint32_t)k) & (uint32_t)(fc - 1));
        int old = atomicCAS(&entry->x, -1, i);
        bool same = old < 0;
        if (!same) {
            uint32_t other[3];
            face_key(tri + 3 * old, other);
            same = key[0] == other[0] && key[1] == other[1] && key[2] == other[2];
        }
        if (same) {
            atomicAdd(&entry->y, odd ? -1 : 1);
            atomicMin(odd ? &entry->w : &entry->z, i);
            return;
        }
    }
}

// Whether listed triangle i survives the cleanup (its key's net count is nonzero and it is the lowest listing of the
// majority orientation), and the survivors per workgroup.
__global__ void k_face_keep(int capacity, const uint32_t* __restrict__ counts, const int* __restrict__ tri,
                            const int4* __restrict__ table, int* __restrict__ keep, int* __restrict__ partial) {
    __shared__ int lds[GTS_BLOCK / 64];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = (int)(counts[1] / 3u);
    int kept = 0;
    if (i < capacity && i < n && !names_component(tri + 3 * i)) {
        kept = 1;
    } else if (i < capacity && i < n) {
        const int fc = face_slots(n);
        uint32_t key[3];
        face_key(tri + 3 * i, key);
        const uint32_t h = face_hash(key, fc);
        for (int k = 0; k < fc; k++) {
            int4 entry = table[(h + (uint32_t)k) & (uint32_t)(fc - 1)];
            if (entry.x < 0) break;  // not reached: every listed face was indexed
            uint32_t other[3];
            face_key(tri + 3 * entry.x, other);
            if (key[0] != other[0] || key[1] != other[1] || key[2] != other[2]) continue;
            kept = entry.y != 0 && i == (entry.y > 0 ? entry.z : entry.w);
            break;
        }
    }
    if (i < capacity) keep[i] = kept;
    int total;
    block_scan(kept, lds, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// The survivors in listing order, each scattering its face normal; thread 0 records their index count.
__global__ void k_place_kept(int capacity, const int* __restrict__ keep, const int* __restrict__ partial, int nblocks,
                             const int* __restrict__ listed, int* __restrict__ tri, uint32_t* __restrict__ counts,
                             const float4* __restrict__ fp, const float4* __restrict__ cut_fp, int V,
                             int* __restrict__ fixed) {
    __shared__ int lds[GTS_BLOCK / 64];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = i < capacity ? keep[i] : 0;
    int total;
    const int k = partial[blockIdx.x] + block_scan(f, lds, total);
    if (f) {
        tri[3 * k] = listed[3 * (size_t)i];
        tri[3 * k + 1] = listed[3 * (size_t)i + 1];
        tri[3 * k + 2] = listed[3 * (size_t)i + 2];
        scatter_normal(listed + 3 * (size_t)i, fp, cut_fp, V, fixed);
    }
    if (i == 0) counts[0] = 3u * (uint32_t)partial[nblocks];
}

// k_NormalsNormalize (dc_recon.cu:441-451) / k_NormalizeCutFP (cut.cu:1093-1100) of one fixed-point sum.
__device__ __forceinline__ f3 unit_normal(const int* f) {
    f3 v = mk3((float)f[0] / kNormalScale, (float)f[1] / kNormalScale, (float)f[2] / kNormalScale);
    float l = len3(v);
    return l > 1e-6f ? div3s(v, l) : mk3(0.f, 0.f, 0.f);
}

// ---- k_ExpandVertices (dc_recon.cu:358-386): the indexed triangles as a soup of (position, unit normal, rest anchor
// and wall flag) per index.
__global__ void k_expand(int n3, const uint32_t* __restrict__ counts, const int* __restrict__ tri, int V,
                         const float4* __restrict__ fp, const float4* __restrict__ cut_fp,
                         const int* __restrict__ fixed, const float4* __restrict__ rest_fp,
                         const float* __restrict__ rest, const int* __restrict__ corners, float* __restrict__ out_pos,
                         float* __restrict__ out_nrm, float* __restrict__ out_aux) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3 || i >= (int)counts[0]) return;
    const uint32_t u = (uint32_t)tri[i];
    const size_t idx = u & kIndexMask;
    float4 p, aux;
    f3 nrm;
    if (u & kCutTag) {
        p = cut_fp[idx];
        nrm = unit_normal(fixed + 3 * ((size_t)V + idx));
        const size_t v = idx >> 3;
        const float4 r4 = rest_fp[v];
        f3 anchor = r4.w > 0.f ? mk3(r4.x, r4.y, r4.z)
                               : (ld3(rest + 3 * (size_t)corners[8 * v]) + ld3(rest + 3 * (size_t)corners[8 * v + 6])) *
                                     .5f;
        aux = make_float4(anchor.x, anchor.y, anchor.z, (float)((u >> 30) & 1u));
    } else {
        p = fp[idx];
        nrm = unit_normal(fixed + 3 * idx);
        const float4 r4 = rest_fp[idx];
        aux = make_float4(r4.x, r4.y, r4.z, 0.f);
    }
    out_pos[3 * (size_t)i] = p.x;
    out_pos[3 * (size_t)i + 1] = p.y;
    out_pos[3 * (size_t)i + 2] = p.z;
    out_nrm[3 * (size_t)i] = nrm.x;
    out_nrm[3 * (size_t)i + 1] = nrm.y;
    out_nrm[3 * (size_t)i + 2] = nrm.z;
    reinterpret_cast<float4*>(out_aux)[i] = aux;
}

// ---------------------------------------------------------------------------------------------------------------
// Tool depth into the reconstructed surface (k_SurfaceDepth, surface_contact.cuh:662-674).
// ---------------------------------------------------------------------------------------------------------------

// shapeBounds (surface_contact.cuh:229-235) with no widening: the axis-aligned box of a capsule or an oriented box.
__device__ __forceinline__ void shape_bounds(const Shape& s, float lo[3], float hi[3]) {
    const f3 c = s.capsule ? (s.a + s.b) * .5f : s.center;
    f3 e;
    if (s.capsule) {
        e = mk3(fabsf(s.b.x - s.a.x) * .5f + s.radius, fabsf(s.b.y - s.a.y) * .5f + s.radius,
                fabsf(s.b.z - s.a.z) * .5f + s.radius);
    } else {
        e = mk3(fabsf(s.u.x) * s.half.x + fabsf(s.n.x) * s.half.y + fabsf(s.w.x) * s.half.z,
                fabsf(s.u.y) * s.half.x + fabsf(s.n.y) * s.half.y + fabsf(s.w.y) * s.half.z,
                fabsf(s.u.z) * s.half.x + fabsf(s.n.z) * s.half.y + fabsf(s.w.z) * s.half.z);
    }
    e = e + mk3(0.f, 0.f, 0.f);
    const f3 l = c - e, h = c + e;
    lo[0] = l.x; lo[1] = l.y; lo[2] = l.z;
    hi[0] = h.x; hi[1] = h.y; hi[2] = h.z;
}

// The signed gap of x to a tool shape and the shape's outward normal there.
__device__ __forceinline__ float gap_at(const Shape& sh, f3 x, f3& n) {
    const ToolHit q = tool_hit(sh, x);
    n = q.normal;
    return q.gap;
}

// corner k of a triangle by selects (a private array indexed at run time would live in scratch)
__device__ __forceinline__ f3 corner(int k, f3 a, f3 b, f3 c) { return k == 0 ? a : (k == 1 ? b : c); }

__device__ __forceinline__ f3 corner_bar(int k) { return mk3((float)(k == 0), (float)(k == 1), (float)(k == 2)); }

// triangleWitness (surface_contact.cuh:169-198): the barycentric point of a triangle deepest in a shape. The best of
// the centroid and the corners starts it; twelve Frank-Wolfe steps move towards the corner the shape's normal favours,
// each with eight golden-section reductions that reuse the surviving sample, keeping the best iterate.
__device__ __forceinline__ f3 triangle_witness(const Shape& sh, f3 a, f3 b, f3 c) {
    f3 bar = mk3(1.f / 3, 1.f / 3, 1.f / 3), best = bar, n;
    float bestD = gap_at(sh, (a + b + c) * (1.f / 3), n);
    for (int k = 0; k < 3; k++) {
        const float d = gap_at(sh, corner(k, a, b, c), n);
        if (d < bestD) { bestD = d; best = corner_bar(k); }
    }
    for (int it = 0; it < 12; it++) {
        const f3 p = a * bar.x + b * bar.y + c * bar.z;
        gap_at(sh, p, n);
        int k = dot3(n, b) < dot3(n, a) ? 1 : 0;
        if (dot3(n, c) < dot3(n, corner(k, a, b, c))) k = 2;
        const f3 target = corner_bar(k), dir = corner(k, a, b, c) - p;
        float lo = 0.f, hi = 1.f, l = .381966f, r = .618034f;
        float dl = gap_at(sh, p + dir * l, n), dr = gap_at(sh, p + dir * r, n);
        for (int j = 0; j < 8; j++) {
            if (dl < dr) {
                hi = r; r = l; dr = dl; l = lo + (hi - lo) * .381966f;
                if (j < 7) dl = gap_at(sh, p + dir * l, n);
            } else {
                lo = l; l = r; dl = dr; r = lo + (hi - lo) * .618034f;
                if (j < 7) dr = gap_at(sh, p + dir * r, n);
            }
        }
        const float t = .5f * (lo + hi);
        bar = bar * (1.f - t) + target * t;
        const float d = gap_at(sh, a * bar.x + b * bar.y + c * bar.z, n);
        if (d < bestD) { bestD = d; best = bar; }
    }
    return best;
}

struct DepthShapes {
    Shape s[4];  // 0 shaft capsule, 1 and 2 the jaw boxes, 3 the blade capsule
};

__global__ void k_depth_clear(float* __restrict__ out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0.f;
}

// One thread per (triangle, shape) of the soup, the four shapes of a triangle on neighbouring lanes, so that a
// triangle near several shapes runs their witnesses side by side. The triangle count is read from the last rebuild's
// counts[0] on the device, and a bounded grid strides over it, so a soup far below the capacity launches no empty
// workgroups. Per shape asked for: the bounds cull, the witness and the penetration there; the maximum over all of
// them is folded in with an unsigned atomic max on the bits of a non-negative float, which does not depend on the
// order.
__global__ void __launch_bounds__(GTS_BLOCK) k_surface_depth(int capacity, const uint32_t* __restrict__ counts,
                                                             const float* __restrict__ soup, DepthShapes shapes,
                                                             int kinds, float* __restrict__ out) {
    const int triangles = min(capacity, (int)(counts[0] / 3u));
    const int items = 4 * triangles;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < items; t += gridDim.x * blockDim.x) {
        const int i = t >> 2, kind = t & 3;
        if (!((kinds >> kind) & 1)) continue;
        const Shape& sh = shapes.s[kind];  // read in place from the kernel arguments
        if (sh.capsule ? sh.radius <= 0.f : (sh.half.x <= 0.f || sh.half.z <= 0.f)) continue;
        const float* v = soup + 9 * (size_t)i;
        const f3 a = mk3(v[0], v[1], v[2]), b = mk3(v[3], v[4], v[5]), c = mk3(v[6], v[7], v[8]);
        const float va[3] = {a.x, a.y, a.z}, vb[3] = {b.x, b.y, b.z}, vc[3] = {c.x, c.y, c.z};
        float lo[3], hi[3];
        shape_bounds(sh, lo, hi);
        bool outside = false;
        for (int ax = 0; ax < 3; ax++)
            outside |= fmaxf(va[ax], fmaxf(vb[ax], vc[ax])) < lo[ax] || fminf(va[ax], fminf(vb[ax], vc[ax])) > hi[ax];
        if (outside) continue;
        const f3 bar = triangle_witness(sh, a, b, c);
        f3 n;
        const float depth = fmaxf(0.f, -gap_at(sh, a * bar.x + b * bar.y + c * bar.z, n));
        if (depth > 0.f) atomicMax(reinterpret_cast<unsigned int*>(out), __float_as_uint(depth));
    }
}

}  // namespace

extern "C" int gts_surface_depth(const GtsSurfaceView* s, const GtsToolState* tool, int kinds, float thickness,
                                 float min_gap, float* depth, void* stream) {
    if (!s || !tool || !depth || kinds < 0 || kinds > 15 || s->tri_capacity <= 0 || s->tri_capacity >= (1 << 28) ||
        !s->counts || !s->out_pos)
        return GTS_ERR_BAD_ARG;
    hipStream_t st = (hipStream_t)stream;
    const ToolPose pose = pose_of(tool);
    DepthShapes shapes;
    for (int kind = 0; kind < 4; kind++) shapes.s[kind] = tool_shape(pose, kind, thickness, min_gap);
    k_depth_clear<<<1, 64, 0, st>>>(depth);
    // at most 1024 workgroups (4 per CU); they stride over the (triangle, shape) items the rebuild left
    const unsigned grid = (unsigned)std::min<long long>(blocks_for(4ll * s->tri_capacity), 1024ll);
    if (kinds) k_surface_depth<<<grid, GTS_BLOCK, 0, st>>>(s->tri_capacity, s->counts, s->out_pos, shapes, kinds, depth);
    HIP_TRY(hipGetLastError());
    return GTS_OK;
}

extern "C" int gts_surface_build(const GtsTissueView* v, const GtsSurfaceView* s, int cut_aware, int have_rest,
                                 void* stream) {
    if (!v || !s || v->n_nodes <= 0 || v->vx <= 0 || v->vy <= 0 || v->vz <= 0) return GTS_ERR_BAD_ARG;
    const long long V = (long long)v->vx * v->vy * v->vz;
    if (s->n_voxels != V || s->n_isect < 0 || s->n_surface_edges < 0 || s->tri_capacity <= 0 || s->qef_iters < 0)
        return GTS_ERR_BAD_ARG;
    if (s->face_capacity < 2 * s->tri_capacity || (s->face_capacity & (s->face_capacity - 1)) != 0)
        return GTS_ERR_BAD_ARG;
    if (!v->pos || !v->rest || !s->voxel_corners || !s->isect_offset || !s->isect_count ||
        (s->n_isect > 0 && (!s->isect || !s->isect_world)) || (s->n_surface_edges > 0 && !s->surface_edges) ||
        !v->edge_skin_t || !s->rest_fp || !s->fp || !s->candidates || !s->item_mask || !s->tri ||
        !s->normal_fixed || !s->scan || !s->counts || !s->out_pos || !s->out_nrm || !s->out_aux)
        return GTS_ERR_BAD_ARG;
    if (cut_aware && (!v->nbr || !v->rot || !v->active || !v->inside || !v->occupied || !v->cut_mask ||
                      !v->edge_cut_t || !v->grid_edges || !v->counters || !s->conn || !s->edge_of ||
                      !s->edge_points || !s->cut_fp || !s->tri_work || !s->face_table ||
                      (v->cut_point_capacity > 0 && !v->cut_points)))
        return GTS_ERR_BAD_ARG;
    const int N = v->n_nodes, Se = s->n_surface_edges, cap = s->tri_capacity;
    const long long n_items = (long long)Se + (cut_aware ? 6ll * N : 0ll);
    if (2ll * Se + 12ll * N >= (1ll << 31) / 3 || V >= (1ll << 26)) return GTS_ERR_BAD_ARG;
    hipStream_t st = (hipStream_t)stream;
    const Grid G{v->vx, v->vy, v->vz, v->vx + 1, v->vy + 1};
    const float4* pos = (const float4*)v->pos;
    const float4* fp = (const float4*)s->fp;
    const float4* cfp = (const float4*)s->cut_fp;
    const int points = cut_aware ? std::max(0, v->cut_point_capacity) : 0;

    k_voxel_pass<<<blocks_for(V), GTS_BLOCK, 0, st>>>(
        (int)V, s->qef_iters, cut_aware, have_rest, G, s->voxel_corners, s->isect_offset, s->isect_count, s->isect,
        s->isect_world, pos, v->rest, cut_aware ? v->nbr : nullptr, v->rot, cut_aware ? v->edge_cut_t : nullptr,
        (const float4*)s->rest_fp, v->cut_mask, (float4*)s->fp, s->normal_fixed, points, v->counters, v->grid_edges,
        v->cut_points, cut_aware ? s->edge_points : nullptr);
    if (cut_aware)
        k_component_points<<<blocks_for(V), GTS_BLOCK, 0, st>>>(
            (int)V, G, s->qef_iters, s->voxel_corners, s->isect_offset, s->isect_count, s->isect, s->isect_world, pos,
            v->rest, v->rot, v->active, v->occupied, v->cut_mask, s->conn, v->edge_cut_t, s->edge_of, s->edge_points,
            v->cut_points, (float4*)s->cut_fp, s->normal_fixed);

    // candidates in fixed slots, then their ordered placement
    const unsigned nb = blocks_for(n_items);
    int* listed = cut_aware ? s->tri_work : s->tri;
    k_triangles<<<nb, GTS_BLOCK, 0, st>>>(Se, (int)n_items, G, cut_aware, s->surface_edges, v->edge_skin_t,
                                          v->edge_cut_t, v->cut_mask, v->occupied, s->conn, v->inside, fp, cfp,
                                          s->candidates, s->item_mask, s->scan);
    k_scan_partials<<<1, 1024, 0, st>>>((int)nb, s->scan);
    k_place<<<nb, GTS_BLOCK, 0, st>>>(Se, (int)n_items, s->item_mask, s->scan, (int)nb, s->candidates, listed, cap,
                                      s->counts, !cut_aware, fp, cfp, (int)V, s->normal_fixed,
                                      cut_aware ? (int4*)s->face_table : nullptr);
    if (cut_aware) {  // the cleanup of internal faces, then the survivors in order
        const unsigned nt = blocks_for(cap);
        k_face_index<<<nt, GTS_BLOCK, 0, st>>>(cap, s->counts, listed, (int4*)s->face_table);
        k_face_keep<<<nt, GTS_BLOCK, 0, st>>>(cap, s->counts, listed, (const int4*)s->face_table, s->item_mask, s->scan);
        k_scan_partials<<<1, 1024, 0, st>>>((int)nt, s->scan);
        k_place_kept<<<nt, GTS_BLOCK, 0, st>>>(cap, s->item_mask, s->scan, (int)nt, listed, s->tri, s->counts, fp, cfp,
                                               (int)V, s->normal_fixed);
    }
    k_expand<<<blocks_for(3ll * cap), GTS_BLOCK, 0, st>>>(3 * cap, s->counts, s->tri, (int)V, fp, cfp, s->normal_fixed,
                                                          (const float4*)s->rest_fp, v->rest, s->voxel_corners,
                                                          s->out_pos, s->out_nrm, s->out_aux);
    HIP_TRY(hipGetLastError());
    return GTS_OK;
}
