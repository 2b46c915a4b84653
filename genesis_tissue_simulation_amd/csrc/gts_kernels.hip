// HIP kernels and C entry points of the single-instance XPBD substep and the swept-quad cut on gfx950.
//
// Layout in HBM (all device buffers belong to the caller):
//   nodes   pos / prev / vel / force as float4 rows (x, y, z, .w = inverse mass for pos): one 16-byte load per node
//           in the tet and edge gathers, coalesced rows in the node passes;
//   edges   one packed 32-bit descriptor per edge in colour order: node i << 4 | support << 3 | axis << 1 | live;
//           the other end, i + stride[axis], and the rest length (the lattice spacing) follow from the lattice;
//   tets    one packed 32-bit descriptor per tet in colour order: corner-0 node << 4 | Kuhn type << 1 | live;
//           the four nodes and Dm^-1 follow from the type, so a tet costs 4 B of topology instead of 52 B;
//   per-tet alpha~ and lambda pairs in colour order, so a colour pass streams contiguous ranges.
// Each colour of the reference's greedy colourings is one launch: no two members share a node, so the Gauss-Seidel
// pass needs no atomics and gives the same result as the serial sweep (reference: xpbd/xpbd_solver.cu:3581-3612).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/gts_xpbd.h"
#include "gts_common.h"
#include "gts_contact.h"
#include "gts_solver.h"
#include "gts_frames.h"
#include "gts_lattice.h"

#define HIP_TRY(call)                                   \
    do {                                                \
        hipError_t _e = (call);                         \
        if (_e != hipSuccess) return -(int)_e;          \
    } while (0)

namespace {

inline int grid_for(int n, int block = GTS_BLOCK) { return std::max(1, (n + block - 1) / block); }

struct Lattice {  // strides and per-type rest data passed by value to every kernel
    int nx, nxy, vx, vy, vz;
    int xcd_order;  // the colour kernels' workgroup -> item mapping (colour_item)
    int block;      // threads per workgroup of the colour kernels, an argument rather than a read of the dispatch
    float inv_rest[6][9];
    float rest_len[3];
};

// Threads per workgroup of the colour kernels. A colour of the 250k-tet kidney lattice holds about 8k tets: 64-thread
// workgroups spread it over ~125 CUs instead of ~32, which measured 5.81 ms per step against 6.32 ms at 256 threads
// and 7.14 ms at 512; 32-thread workgroups (half-empty waves on every CU) measured 6.10 ms.
#define GTS_COLOUR_BLOCK 64
inline int colour_block(const GtsTissueView* v) { return v->colour_block > 0 ? v->colour_block : GTS_COLOUR_BLOCK; }

Lattice lattice_of(const GtsTissueView* v) {
    Lattice L;
    L.block = colour_block(v);
    L.vx = v->vx; L.vy = v->vy; L.vz = v->vz;
    L.nx = v->vx + 1;
    L.nxy = (v->vx + 1) * (v->vy + 1);
    L.xcd_order = v->xcd_order;
    std::memcpy(L.inv_rest, v->inv_rest, sizeof(L.inv_rest));
    for (int a = 0; a < 3; a++) L.rest_len[a] = v->axis_rest_length[a];
    return L;
}

__device__ __forceinline__ int axis_stride(const Lattice& L, int axis) { return axis == 0 ? 1 : (axis == 1 ? L.nx : L.nxy); }

__device__ __forceinline__ f3 xyz(float4 p) { return mk3(p.x, p.y, p.z); }

// Marks kernel arguments as needed at this point: their scalar loads issue together before it, not one round at a time
// before each first use. No instruction is emitted.
__device__ __forceinline__ void asm_use(int x) { asm volatile("" ::"s"(x)); }
__device__ __forceinline__ void asm_use(float x) { asm volatile("" ::"s"(x)); }
__device__ __forceinline__ void asm_use(unsigned x) { asm volatile("" ::"s"(x)); }
template <class T>
__device__ __forceinline__ void asm_use(T* p) { asm volatile("" ::"s"(p)); }
template <class... A>
__device__ __forceinline__ void args_ready(const A&... a) {
    (asm_use(a), ...);
}
// The same for loaded values: the loads issued before this point complete here, together.
__device__ __forceinline__ void value_use(float x) { asm volatile("" ::"v"(x)); }
__device__ __forceinline__ void value_use(float4 x) {
    value_use(x.x);
    value_use(x.y);
    value_use(x.z);
    value_use(x.w);
}
template <class... A>
__device__ __forceinline__ void values_ready(const A&... a) {
    (value_use(a), ...);
}

// A bend element's three nodes, loaded ahead, as NodePair.
struct NodeTriple {
    float4* p;
    int i, j;
    float4 a, b, c;
    __device__ __forceinline__ float4 ld(int k) const { return k == i ? a : (k == j ? b : c); }
    __device__ __forceinline__ void st(int k, float4 v) const { p[k] = v; }
};

// An edge's two nodes, loaded ahead: reads return them, writes go to memory.
struct NodePair {
    float4* p;
    int i;
    float4 a, b;
    __device__ __forceinline__ float4 ld(int k) const { return k == i ? a : b; }
    __device__ __forceinline__ void st(int k, float4 v) const { p[k] = v; }
};

// The item a thread of a colour kernel takes. Workgroups are dispatched round-robin over the 8 XCDs (workgroup b on
// XCD b % 8). With L.xcd_order the workgroups of one XCD take one contiguous range of the colour; colours list their
// items in lattice order, so an XCD then meets the same slab of the lattice in every colour, and that slab's nodes can
// stay in the XCD's L2 between launches. The mapping is a bijection, so results do not depend on it. `block` is the
// launch's workgroup size, passed in (reading blockDim costs a load from the dispatch packet before the first load that
// depends on the item); the computation is branch-free, so that the kernel's argument loads issue together.
__device__ __forceinline__ int ordered_item(int xcd_order, int block) {
    const int b = blockIdx.x;
    const int G = gridDim.x, per = G >> 3, rem = G & 7, x = b & 7, slot = b >> 3;
    const int mapped = (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + slot;
    return (xcd_order ? mapped : b) * block + threadIdx.x;
}
// the colour kernels launched with colour_block threads
__device__ __forceinline__ int colour_item(const Lattice& L) { return ordered_item(L.xcd_order, L.block); }

// ------------------------------------------------------------------------------------------------ set-up kernels

__global__ void k_init_state(int N, const float* mass, const int* active, const int* pinned, const int* nbr,
                             float4* pos, uint32_t* bend_alive, uint32_t* counters) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        counters[0] = 0u;          // topology revision
        counters[1] = 0u;          // cut point counter
        counters[2] = 0xffffffffu; // revision the liveness was derived at: none yet
        counters[3] = 1u;          // stale
        counters[4] = 0u;          // edges marked by the tear law
    }
    if (i >= N) return;
    // invMass = (pinned || !active) ? 0 : 1/mass (xpbd_solver.cu:458-459); active/pinned do not change on this path.
    float4 p = pos[i];
    p.w = (pinned[i] != 0 || active[i] == 0) ? 0.f : 1.f / mass[i];
    pos[i] = p;
    // bend pair p joins slots (a, b) in lexicographic order; alive when both links exist (grid.py:257-263).
    uint32_t m = 0u;
    int pair = 0;
    for (int a = 0; a < 6; a++)
        for (int b = a + 1; b < 6; b++, pair++)
            if (nbr[6 * i + a] >= 0 && nbr[6 * i + b] >= 0) m |= 1u << pair;
    bend_alive[i] = m;
}

__global__ void k_set_gravity(int N, const float* mass, float gx, float gy, float gz, float4* force) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float m = mass[i];  // force = acceleration * mass (physics.cu:317-321)
    force[i] = make_float4(gx * m, gy * m, gz * m, 0.f);
}

// k_TopologyStale (unified_vertex.cuh:859-861): liveness is re-derived only after the revision moved.
__global__ void k_topology_stale(uint32_t* counters) {
    uint32_t rev = counters[0];
    counters[3] = rev != counters[2] ? 1u : 0u;
    counters[2] = rev;
}

__global__ void k_clear_support(int n, int* support, const uint32_t* counters) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && counters[3]) support[i] = 0;
}

// Tet domain of k_RefreshLiveness (xpbd_solver.cu:460-465) + k_MarkTetEdgeSupport (:474-479).
__global__ void k_refresh_tets(int T, Lattice L, const int* nbr, const int* active, uint32_t* tet_desc, int* support,
                               const uint32_t* counters) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T || !counters[3]) return;
    uint32_t d = tet_desc[t];
    int base = (int)(d >> 4), type = (int)((d >> 1) & 7u);
    bool live = true;
    int owner[3], na = 0;
    for (int a = 0; a < 4; a++) {
        int ca = kuhn_corner(type, a);
        live = live && active[base + corner_offset(ca, L.nx, L.nxy)] != 0;
        for (int b = a + 1; b < 4; b++) {
            int cb = kuhn_corner(type, b), diff = ca ^ cb;
            if (diff != 1 && diff != 2 && diff != 4) continue;
            int axis = diff == 1 ? 0 : (diff == 2 ? 1 : 2);
            int lower = ((ca >> axis) & 1) == 0 ? ca : cb;
            if (na < 3) owner[na++] = 6 * (base + corner_offset(lower, L.nx, L.nxy)) + 2 * axis + 1;
        }
    }
    for (int k = 0; k < 3; k++) live = live && nbr[owner[k]] >= 0;
    tet_desc[t] = (d & ~1u) | (live ? 1u : 0u);
    if (live)
        for (int k = 0; k < 3; k++) support[owner[k]] = 1;
}

// Edge domain of k_RefreshLiveness (xpbd_solver.cu:451-453) with the tet support bit for the fallback law.
__global__ void k_refresh_edges(int E, Lattice L, const int* nbr, const int* active, const int* support,
                                uint32_t* edge_desc, const uint32_t* counters) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E || !counters[3]) return;
    uint32_t d = edge_desc[e];
    int i = (int)(d >> 4), axis = (int)((d >> 1) & 3u);
    int j = i + axis_stride(L, axis), slot = 6 * i + 2 * axis + 1;
    bool live = nbr[slot] >= 0 && active[i] != 0 && active[j] != 0;
    uint32_t sup = support[slot] ? 8u : 0u;
    edge_desc[e] = (d & ~9u) | sup | (live ? 1u : 0u);
}

// k_ComputeTetAlpha (xpbd_solver.cu:578-587): alpha~ = (1/h^2) / (mu V) and (1/h^2) / (kappa V) once per frame.
__global__ void k_tet_alpha(int T, float invH2, const float* kd, const float* kh, float* alpha) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    alpha[t] = invH2 / kd[t];
    alpha[T + t] = invH2 / kh[t];
}

// ------------------------------------------------------------------------------------------------ substep kernels

// k_predict (xpbd_solver.cu:393-412): prev = x; for w > 0, v += h F w and x += h v.
__global__ void k_predict(int N, float h, float4* __restrict__ pos, float4* __restrict__ prev,
                          float4* __restrict__ vel, const float4* __restrict__ force, int xcd_order) {
    // nodes in lattice order: the XCD that owns their slab in the colour kernels
    int i = ordered_item(xcd_order, GTS_BLOCK);
    if (i >= N) return;
    float4 p = pos[i];
    prev[i] = p;
    if (p.w > 0.f) {
        float4 F = force[i];
        f3 a = mk3(F.x, F.y, F.z) * p.w;
        float4 v0 = vel[i];
        f3 v = mk3(v0.x, v0.y, v0.z) + a * h;
        vel[i] = make_float4(v.x, v.y, v.z, 0.f);
        f3 x = xyz(p) + v * h;
        pos[i] = make_float4(x.x, x.y, x.z, p.w);
    }
}

// k_solve_edges (xpbd_solver.cu:484-524) over one colour.
template <bool READ_LAMBDA, bool WRITE_LAMBDA, bool FALLBACK, bool DAMPED>
__global__ void k_edges(int off, int cnt, Lattice L, float alphaTilde, float fallbackInvEh2, float omega,
                        float betaGammaH, const uint32_t* __restrict__ edge_desc, float* __restrict__ lambda,
                        float4* __restrict__ pos, const float4* __restrict__ prev) {
    args_ready(off, cnt, edge_desc, L.xcd_order, L.block, (int)gridDim.x, L.nx, L.nxy, L.rest_len[0], L.rest_len[1],
               L.rest_len[2], pos, lambda, prev);  // as in k_tets
    const int k = colour_item(L);
    const uint32_t d = edge_desc[off + (k < cnt ? k : cnt - 1)];  // issued before the range check, as in k_tets
    if (k >= cnt) return;
    const int e = off + k;
    // a row that does not project keeps the substep's reset multiplier, 0 (xpbd_solver.cu:3534)
    const bool skip = !(d & 1u) || (FALLBACK && (d & 8u));
    if (skip) {
        if (WRITE_LAMBDA && !READ_LAMBDA) lambda[e] = 0.f;
        return;
    }
    const int axis = (int)((d >> 1) & 3u);
    const float restLen = axis == 0 ? L.rest_len[0] : (axis == 1 ? L.rest_len[1] : L.rest_len[2]);  // no indexed load
    if (FALLBACK) alphaTilde = fallbackInvEh2 / restLen;
    const int i = (int)(d >> 4), j = i + axis_stride(L, axis);
    // Both nodes and the multiplier in one round of loads: left to the compiler, the inverse masses would be read first
    // and the coordinates and the multiplier only past the tests that need them, one round each.
    const float4 pi = pos[i], pj = pos[j];
    const float lam = READ_LAMBDA ? lambda[e] : 0.f;
    values_ready(pi, pj, lam);
    float out;
    const bool ok = edge_project<DAMPED>(NodePair{pos, i, pi, pj}, prev, i, j, restLen, alphaTilde, omega, betaGammaH,
                                         lam, out);
    // a row that writes nothing in the first iteration keeps the substep's reset multiplier, 0 (xpbd_solver.cu:3534)
    if (WRITE_LAMBDA && (ok || !READ_LAMBDA)) lambda[e] = ok ? out : 0.f;
}

// k_solve_bending (xpbd_solver.cu:534-573) over one colour of bend elements. An element is live while its apex keeps
// the pair's bit in bend_alive (the refresh's bendActive, :446; SeverLink clears it); its arms are the lattice
// neighbours in the pair's two slots and its rest angle pi for opposite slots, else pi/2 (grid.py:267-268).
template <bool READ_LAMBDA, bool WRITE_LAMBDA, bool DAMPED>
__global__ void k_bends(int off, int cnt, Lattice L, float alphaB, float omega, float betaGammaH,
                        const uint32_t* __restrict__ bend_desc, const uint32_t* __restrict__ bend_alive,
                        float* __restrict__ lambda, float4* __restrict__ pos, const float4* __restrict__ prev) {
    args_ready(off, cnt, bend_desc, bend_alive, L.xcd_order, L.block, (int)gridDim.x, L.nx, L.nxy, pos, lambda, prev);
    const int q = colour_item(L);
    const uint32_t d = bend_desc[off + (q < cnt ? q : cnt - 1)];  // issued before the range check, as in k_tets
    if (q >= cnt) return;
    const int e = off + q;
    const int i = (int)(d >> 4), pair = (int)(d & 15u);
    const int sa = bend_slot_a(pair), sb = bend_slot_b(pair);
    const int da = axis_stride(L, sa >> 1), db = axis_stride(L, sb >> 1);
    const int j = i + ((sa & 1) ? da : -da), k = i + ((sb & 1) ? db : -db);
    // the liveness word, the three nodes and the multiplier in one round of loads (the nodes exist whether or not the
    // element is live: the builder keeps only pairs whose arms are lattice neighbours)
    const uint32_t alive = bend_alive[i];
    const float4 pa = pos[i], pb = pos[j], pc = pos[k];
    const float lam = READ_LAMBDA ? lambda[e] : 0.f;
    values_ready(pa, pb, pc, lam);
    if (!((alive >> pair) & 1u)) {  // a dead element keeps the substep's reset multiplier, 0 (:3535)
        if (WRITE_LAMBDA && !READ_LAMBDA) lambda[e] = 0.f;
        return;
    }
    const float theta0 = (sa >> 1) == (sb >> 1) ? 3.14159274f : 1.57079637f;
    float out;
    const bool ok = bend_project<DAMPED>(NodeTriple{pos, i, j, pa, pb, pc}, prev, i, j, k, theta0, alphaB, omega,
                                         betaGammaH, lam, out);
    // a row that writes nothing in the first iteration keeps the substep's reset multiplier, 0 (xpbd_solver.cu:3535)
    if (WRITE_LAMBDA && (ok || !READ_LAMBDA)) lambda[e] = ok ? out : 0.f;
}

// k_solve_tet (xpbd_solver.cu:749-829) over one colour. MODE 0 keeps the hydrostatic row alone; 1 and 2 are the
// coupled 2x2 block, legacy and guarded.
template <int MODE, bool READ_LAMBDA, bool WRITE_LAMBDA, bool DAMPED, bool PRECISE>
__global__ void k_tets(int off, int cnt, int T, Lattice L, float omega, float dampingOverH, float bulkDampingOverH,
                       const uint32_t* __restrict__ tet_desc, const float* __restrict__ alpha,
                       float* __restrict__ lambda, float4* __restrict__ pos, const float4* __restrict__ prev) {
    // Every argument the tet needs is requested in one round of scalar loads here, rather than each just before its
    // first use (the compiler's placement), which would put further rounds between the descriptor and the nodes.
    args_ready(off, cnt, tet_desc, L.xcd_order, L.block, (int)gridDim.x, L.nx, L.nxy, L.rest_len[0], L.rest_len[1],
               L.rest_len[2], pos, alpha, lambda, prev);
    const int k = colour_item(L);
    // the descriptor load issues before the range check (at a clamped item), with nothing but the arguments before it
    const uint32_t d = tet_desc[off + (k < cnt ? k : cnt - 1)];
    if (k >= cnt) return;
    const int e = off + k;
    if (!(d & 1u)) {  // a dead tet keeps the substep's reset multipliers, 0 (xpbd_solver.cu:3536)
        if (WRITE_LAMBDA && !READ_LAMBDA) lambda[e] = lambda[T + e] = 0.f;
        return;
    }
    int base = (int)(d >> 4), type = (int)((d >> 1) & 7u);
    int id[4];
    for (int c = 0; c < 4; c++) id[c] = base + corner_offset(kuhn_corner(type, c), L.nx, L.nxy);
    f3 rk[3];
    if (PRECISE)
        for (int c = 0; c < 3; c++) rk[c] = kuhn_rest_edge(type, c + 1, L.rest_len);
    float lamD = READ_LAMBDA ? lambda[e] : 0.f, lamH = READ_LAMBDA ? lambda[T + e] : 0.f;
    const bool ok = tet_project<MODE, DAMPED, PRECISE>(GlobalNodes{pos}, prev, id, L.inv_rest[type], alpha[e],
                                                       alpha[T + e], lamD, lamH, omega, dampingOverH,
                                                       bulkDampingOverH, rk);
    // a row that writes nothing in the first iteration keeps the substep's reset multipliers, 0 (xpbd_solver.cu:3536)
    if (WRITE_LAMBDA && (ok || !READ_LAMBDA)) {
        lambda[e] = ok ? lamD : 0.f;
        lambda[T + e] = ok ? lamH : 0.f;
    }
}

// Several consecutive tet colours in one launch over one lattice region per workgroup (the tiled sweep; the plan and
// its argument are in genesis_tissue_simulation_amd/tiling.py). The region's node box is loaded into LDS from pos_in;
// pass p projects, in LDS, the colour-(c0+p) tets within k-1-p cells of the region (the plan's entries), each pass
// after a workgroup barrier, so every tet reads its nodes as the per-colour sweep leaves them; the region then writes
// the nodes it owns to pos_out, its copies of the multipliers to tile_lambda and those of the tets it owns to
// tet_lambda. No projection reads memory another workgroup of the launch writes. Undamped forms only.

__device__ __forceinline__ int ordered_block(int xcd_order) { return ordered_item(xcd_order, 1) - (int)threadIdx.x; }

#define GTS_TILE_KMAX 4  // colours a tiled launch may hold: each thread loads its first tet of every pass up front
#define GTS_TILE_THREADS 1024  // threads of a tiled workgroup
#define GTS_TILE_BOX_LOADS 4  // nodes of the LDS box a thread loads: box^3 < 4 * 1024 (box <= 15)
#define GTS_TILE_WRITE_LOADS 1  // node writes a thread loads up front
#define GTS_TILE_MAX_REGIONS 640  // regions of a tiled launch: their dense ids fill most of the 4 KiB argument space

// Where a launch's regions find their data: every address follows from the region's index, and its box origin from
// its dense id, which travels with the launch (a scalar read of the argument segment, no memory round of its own).
struct TileLaunch {
    int k, stride, wcap, box, region, rx, ry, halo, T;
    int cap[GTS_TILE_KMAX], start[GTS_TILE_KMAX];  // each pass's rows in a region's block of `stride` entry rows
    int dense[GTS_TILE_MAX_REGIONS];  // dense id rx + rx_count (ry + ry_count rz) of each region (32-bit: a scalar load)
};

// Copies each plan entry's inputs into the entry: for a tet alpha_D, alpha_H and the descriptor's bits (liveness,
// type); for an edge (alpha null) 0, 0 and the descriptor's bits (liveness, axis).
__global__ void k_tile_gather(int n, int T, const int2* __restrict__ entries, const uint32_t* __restrict__ desc,
                              const float* __restrict__ alpha, float4* __restrict__ inputs) {
    const int q = blockIdx.x * GTS_BLOCK + threadIdx.x;
    if (q >= n) return;
    const int slot = entries[q].x & 0x3fffffff;  // padding rows name slot 0
    inputs[q] = make_float4(alpha ? alpha[slot] : 0.f, alpha ? alpha[T + slot] : 0.f,
                            __builtin_bit_cast(float, desc[slot]), 0.f);
}

// One tiled tet projection in the region's LDS box: e the plan entry, in its gathered inputs, lam its multipliers.
// Padding rows (bit 30 of e.x clear) do nothing.
template <int MODE, bool WRITE_LAMBDA, bool PRECISE>
__device__ __forceinline__ void tiled_project(int2 e, float4 in, float2 lam, float4* nodes, const float* invRest,
                                              int box, int box2, const Lattice& L, float omega, int T,
                                              float2* __restrict__ lam_copy, float* __restrict__ lam_out) {
    if (!(e.x & (1 << 30))) return;
    const bool owned = e.x < 0;
    const int slot = e.x & 0x3fffffff;
    const uint32_t d = __builtin_bit_cast(uint32_t, in.z);
    float lamD = lam.x, lamH = lam.y;
    bool ok = false;
    if (d & 1u) {  // a dead tet keeps its multipliers: 0 from the first iteration's reset
        const int type = (int)((d >> 1) & 7u);
        int id[4];
        for (int c = 0; c < 4; c++) id[c] = e.y + corner_offset(kuhn_corner(type, c), box, box2);
        f3 rk[3];
        if (PRECISE)
            for (int c = 0; c < 3; c++) rk[c] = kuhn_rest_edge(type, c + 1, L.rest_len);
        ok = tet_project<MODE, false, PRECISE>(LdsNodes{nodes}, nullptr, id, invRest + 9 * type, in.x, in.y, lamD,
                                               lamH, omega, 0.f, 0.f, rk);
    }
    // as k_tets stores in place: the update, else the multiplier read (the first iteration's reset 0)
    if (WRITE_LAMBDA) {
        const float2 out = make_float2(ok ? lamD : lam.x, ok ? lamH : lam.y);
        *lam_copy = out;
        if (owned) {
            lam_out[slot] = out.x;
            lam_out[T + slot] = out.y;
        }
    }
}

// One tiled edge projection (k_edges, undamped, without the fallback compliance): e.y is node i's index in the box,
// node j lies one node further along the edge's axis. lam.x is the multiplier; the owner writes edge_lambda.
template <bool WRITE_LAMBDA>
__device__ __forceinline__ void tiled_edge_project(int2 e, float4 in, float2 lam, float4* nodes, int box, int box2,
                                                   const Lattice& L, float alphaTilde, float omega,
                                                   float2* __restrict__ lam_copy, float* __restrict__ lam_out) {
    if (!(e.x & (1 << 30))) return;
    const uint32_t d = __builtin_bit_cast(uint32_t, in.z);
    float out = lam.x;
    bool ok = false;
    if (d & 1u) {  // a dead edge keeps its multiplier: 0 from the first iteration's reset
        const int axis = (int)((d >> 1) & 3u);
        const float restLen = axis == 0 ? L.rest_len[0] : (axis == 1 ? L.rest_len[1] : L.rest_len[2]);
        const int i = e.y, j = i + (axis == 0 ? 1 : (axis == 1 ? box : box2));
        ok = edge_project<false>(LdsNodes{nodes}, nullptr, i, j, restLen, alphaTilde, omega, 0.f, lam.x, out);
    }
    if (WRITE_LAMBDA) {
        const float v = ok ? out : lam.x;
        *lam_copy = make_float2(v, 0.f);
        if (e.x < 0) lam_out[e.x & 0x3fffffff] = v;
    }
}

template <int MODE, bool READ_LAMBDA, bool WRITE_LAMBDA, bool PRECISE, bool EDGES>
__global__ void __launch_bounds__(GTS_TILE_THREADS)
    k_tets_tiled(TileLaunch A, const int2* __restrict__ entries,
                 const float4* __restrict__ inputs, float2* __restrict__ lam_copy, const int2* __restrict__ writes,
                 Lattice L, float omega, float alphaTilde, float* __restrict__ lam_out,
                 const float4* __restrict__ pos_in, float4* __restrict__ pos_out) {
    extern __shared__ float4 nodes[];
    __shared__ float invRest[54];  // the six types' Dm^-1, staged with the box (indexed by type, not from arguments)
    args_ready(L.xcd_order, (int)gridDim.x, pos_in, A.k, A.stride, A.wcap, A.cap[0], A.cap[1], A.cap[2], A.cap[3],
               A.start[0], A.start[1], A.start[2], A.start[3], entries, inputs, lam_copy, writes);
    const int rl = ordered_block(L.xcd_order), dense = A.dense[rl];
    const int tid = threadIdx.x, bs = GTS_TILE_THREADS;  // the launch's workgroup size, not read from the dispatch
    const int k = A.k, box = A.box, box2 = box * box, box3 = box2 * box;
    // One round of loads: every address below follows from rl and tid. Each pass loads the rows of its capacity (the
    // most any region of the launch holds; the region's rows past its own count are padding, which the rows mark, so
    // no count is read), the write list those of its.
    const size_t e0 = (size_t)rl * A.stride;
    int2 en[GTS_TILE_KMAX];
    float4 in[GTS_TILE_KMAX];
    float2 lm[GTS_TILE_KMAX];
#pragma unroll
    for (int p = 0; p < GTS_TILE_KMAX; p++) {
        en[p] = make_int2(0, 0);
        in[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        lm[p] = make_float2(0.f, 0.f);
        if (tid < A.cap[p]) {  // A.cap[p] = 0 past the launch's colours
            const size_t q = e0 + A.start[p] + tid;
            en[p] = entries[q];
            in[p] = inputs[q];
            if (READ_LAMBDA) lm[p] = lam_copy[q];
        }
    }
    const int2* wr_row = writes + (size_t)rl * A.wcap;
    int2 wr[GTS_TILE_WRITE_LOADS];
#pragma unroll
    for (int j = 0; j < GTS_TILE_WRITE_LOADS; j++) {
        const int q = tid + j * bs;
        wr[j] = wr_row[q < A.wcap ? q : A.wcap - 1];  // clamped, not conditional: a conditional load waits here
    }
    // the box, at the origin of the region's dense id
    const int ox = (dense % A.rx) * A.region - A.halo, oy = ((dense / A.rx) % A.ry) * A.region - A.halo,
              oz = (dense / (A.rx * A.ry)) * A.region - A.halo;
    const int ny = L.nxy / L.nx, nz = L.vz + 1;
    const float invRestMine = reinterpret_cast<const float*>(L.inv_rest)[tid < 54 ? tid : 0];
    {
        // i / box and i / box^2 for i < 4096 as (i * ceil(2^20 / d)) >> 20, exact for d <= 256
        const uint32_t m1 = ((1u << 20) + box - 1) / box, m2 = ((1u << 20) + box2 - 1) / box2;
        float4 v[GTS_TILE_BOX_LOADS];
        bool inb[GTS_TILE_BOX_LOADS];
#pragma unroll
        for (int j = 0; j < GTS_TILE_BOX_LOADS; j++) {
            const int i = tid + j * bs;
            const int lz = (int)(((uint32_t)i * m2) >> 20), rem = i - lz * box2;
            const int ly = (int)(((uint32_t)rem * m1) >> 20), lx = rem - ly * box;
            const int x = ox + lx, y = oy + ly, z = oz + lz;
            inb[j] = i < box3 && x >= 0 && x < L.nx && y >= 0 && y < ny && z >= 0 && z < nz;
            v[j] = pos_in[inb[j] ? x + L.nx * y + L.nxy * z : 0];
        }
        // all in flight together: left alone, the compiler moves each load into its conditional store, one round each
#pragma unroll
        for (int j = 0; j < GTS_TILE_BOX_LOADS; j++) value_use(v[j]);
#pragma unroll
        for (int j = 0; j < GTS_TILE_BOX_LOADS; j++)
            if (inb[j]) nodes[tid + j * bs] = v[j];
        if (tid < 54) invRest[tid] = invRestMine;
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < GTS_TILE_KMAX; p++) {
        if (p >= k) break;
        const size_t pq = e0 + A.start[p];
        if (EDGES) {
            if (tid < A.cap[p])
                tiled_edge_project<WRITE_LAMBDA>(en[p], in[p], lm[p], nodes, box, box2, L, alphaTilde, omega,
                                                 lam_copy + pq + tid, lam_out);
            for (int q = tid + bs; q < A.cap[p]; q += bs)
                tiled_edge_project<WRITE_LAMBDA>(entries[pq + q], inputs[pq + q],
                                                 READ_LAMBDA ? lam_copy[pq + q] : make_float2(0.f, 0.f), nodes, box,
                                                 box2, L, alphaTilde, omega, lam_copy + pq + q, lam_out);
        } else {
            if (tid < A.cap[p])
                tiled_project<MODE, WRITE_LAMBDA, PRECISE>(en[p], in[p], lm[p], nodes, invRest, box, box2, L, omega,
                                                           A.T, lam_copy + pq + tid, lam_out);
            for (int q = tid + bs; q < A.cap[p]; q += bs)  // passes of more rows than threads
                tiled_project<MODE, WRITE_LAMBDA, PRECISE>(entries[pq + q], inputs[pq + q],
                                                           READ_LAMBDA ? lam_copy[pq + q] : make_float2(0.f, 0.f),
                                                           nodes, invRest, box, box2, L, omega, A.T,
                                                           lam_copy + pq + q, lam_out);
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < GTS_TILE_WRITE_LOADS; j++)  // padding rows name node -1
        if (tid + j * bs < A.wcap && wr[j].y >= 0) pos_out[wr[j].y] = nodes[wr[j].x];
    for (int q = tid + GTS_TILE_WRITE_LOADS * bs; q < A.wcap; q += bs) {
        const int2 w = wr_row[q];
        if (w.y >= 0) pos_out[w.y] = nodes[w.x];
    }
}

// Friction contact of each node with the tools (friction_contact.cuh:143-169, two solve passes as at
// xpbd_solver.cu:3663-3665), then (VEL) k_update_velocity and, unless perpendicular edge damping must run between
// them, k_apply_velscale (xpbd_solver.cu:871-919). Nodes are independent here, so one pass does all of it; with the
// tissue guard, whose safeguard runs between contact and velocity, VEL is off and k_velocity follows the safeguard.
// PREDICT (with VEL) also runs the next substep's k_predict on the node, when nothing else runs between the two.
template <bool CONTACT, bool SCALE, bool VEL = true, bool PREDICT = false>
__global__ void k_finish_nodes(int N, const StepShapes* __restrict__ shapes_p, uint32_t armed, float alpha, float muS, float muD,
                               float radius, float scaleR, float invH, float velScale, float4* __restrict__ pos,
                               float4* __restrict__ prev, float4* __restrict__ vel, float* reaction,
                               int xcd_order, float h = 0.f, const float4* __restrict__ ext = nullptr) {
    int i = ordered_item(xcd_order, GTS_BLOCK);
    f3 force = mk3(0, 0, 0);
    float contacts = 0.f;
    if (i < N) {
        float4 p4 = pos[i];
        float w = p4.w;
        f3 x = xyz(p4);
        f3 xp = xyz(prev[i]);
        if (CONTACT && w > 0.f) {
            const StepShapes& shapes = *shapes_p;
            ContactRecord rec[4];
            float lamN[4];
            f3 lamT[4];
            // every loop over the four shapes is unrolled so that the records stay in registers
#pragma unroll
            for (int kind = 0; kind < 4; kind++) {
                rec[kind].valid = 0;
                lamN[kind] = 0.f;
                lamT[kind] = mk3(0, 0, 0);
                if (((armed >> kind) & 1u) && shape_present(shapes.to[kind]))
                    rec[kind] = trace_tool_contact(shapes.from[kind], shapes.to[kind], radius, xp, x);
            }
#pragma unroll
            for (int pass = 0; pass < 2; pass++)
#pragma unroll
                for (int kind = 0; kind < 4; kind++) {
                    const ContactRecord& c = rec[kind];
                    if (!c.valid) continue;
                    float C = dot3(x - c.target, c.normal), old = lamN[kind];
                    float next = fmaxf(0.f, old + (-C - alpha * old) / (w + alpha));
                    x = x + c.normal * (w * (next - old));
                    lamN[kind] = next;
                    f3 slip = x - xp - c.motion;
                    slip = slip - c.normal * dot3(slip, c.normal);
                    f3 oldT = lamT[kind];
                    f3 trial = oldT - slip * (1.f / w);
                    float len = len3(trial);
                    f3 nextT = trial;
                    if (len > muS * next) nextT = len > 1e-15f ? trial * (muD * next / len) : mk3(0, 0, 0);
                    x = x + (nextT - oldT) * w;
                    lamT[kind] = nextT;
                }
#pragma unroll
            for (int kind = 0; kind < 4; kind++)
                if (rec[kind].valid) {
                    force = force + (rec[kind].normal * lamN[kind] + lamT[kind]) * (-scaleR);
                    if (lamN[kind] > 0.f) contacts += 1.f;
                }
            pos[i] = make_float4(x.x, x.y, x.z, w);
        }
        if (VEL) {
            f3 v = w > 0.f ? (x - xp) * invH : mk3(0, 0, 0);
            if (SCALE) v = w > 0.f ? v * velScale : mk3(0, 0, 0);
            if (PREDICT) {  // k_predict of the next substep, as k_predict below computes it from these values
                prev[i] = make_float4(x.x, x.y, x.z, w);
                if (w > 0.f) {
                    float4 F = ext[i];
                    f3 a = mk3(F.x, F.y, F.z) * w;
                    v = v + a * h;
                    x = x + v * h;
                    pos[i] = make_float4(x.x, x.y, x.z, w);
                }
            }
            vel[i] = make_float4(v.x, v.y, v.z, 0.f);
        }
    }
    if (CONTACT) {
        // reaction sum: wave reduction, one atomic per wave
        for (int o = 32; o > 0; o >>= 1) {
            force.x += __shfl_xor(force.x, o);
            force.y += __shfl_xor(force.y, o);
            force.z += __shfl_xor(force.z, o);
            contacts += __shfl_xor(contacts, o);
        }
        if ((threadIdx.x & 63) == 0 && (contacts > 0.f || force.x != 0.f || force.y != 0.f || force.z != 0.f)) {
            atomicAdd(reaction + 0, force.x);
            atomicAdd(reaction + 1, force.y);
            atomicAdd(reaction + 2, force.z);
            atomicAdd(reaction + 3, contacts);
        }
    }
}

// k_perp_damp_edges (xpbd_solver.cu:890-910) over one colour.
__global__ void k_perp_damp(int off, int cnt, Lattice L, float cs, float h, const uint32_t* edge_desc,
                            const float4* pos, float4* vel) {
    int k = ordered_item(L.xcd_order, GTS_BLOCK);
    if (k >= cnt) return;
    uint32_t d = edge_desc[off + k];
    if (!(d & 1u)) return;
    int axis = (int)((d >> 1) & 3u), i = (int)(d >> 4), j = i + axis_stride(L, axis);
    float4 pi = pos[i], pj = pos[j];
    float wi = pi.w, wj = pj.w;
    if (wi + wj <= 0.f) return;
    f3 dd = xyz(pi) - xyz(pj);
    float len = len3(dd);
    if (len < 1e-9f) return;
    f3 n = dd * (1.f / len);
    f3 vi = xyz(vel[i]), vj = xyz(vel[j]);
    f3 vrel = vi - vj;
    f3 vp = vrel - n * dot3(vrel, n);
    f3 imp = vp * (h * cs);
    if (wi > 0.f) { vi = vi - imp * wi; vel[i] = make_float4(vi.x, vi.y, vi.z, 0.f); }
    if (wj > 0.f) { vj = vj + imp * wj; vel[j] = make_float4(vj.x, vj.y, vj.z, 0.f); }
}

__global__ void k_velscale(int N, float velScale, const float4* pos, float4* vel) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    f3 v = pos[i].w > 0.f ? xyz(vel[i]) * velScale : mk3(0, 0, 0);
    vel[i] = make_float4(v.x, v.y, v.z, 0.f);
}

// k_update_velocity and, unless perpendicular edge damping runs in between, k_apply_velscale (xpbd_solver.cu:871-919),
// where the tissue guard's safeguard separates them from the node passes above.
template <bool SCALE>
__global__ void k_velocity(int N, float invH, float velScale, const float4* __restrict__ pos,
                           const float4* __restrict__ prev, float4* __restrict__ vel) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 p4 = pos[i];
    f3 v = p4.w > 0.f ? (xyz(p4) - xyz(prev[i])) * invH : mk3(0, 0, 0);
    if (SCALE) v = p4.w > 0.f ? v * velScale : mk3(0, 0, 0);
    vel[i] = make_float4(v.x, v.y, v.z, 0.f);
}

// ------------------------------------------------------------------------------------------------ tissue guard
// Tensile reinforcement after every tet sweep and the minimum-volume safeguard at the end of every substep
// (xpbd_solver.cu:588-741). Every corner of a tet lies in one lattice cell, so a tet colour and a lattice parity class
// never hold two nodes of one tet: each launch below is race-free and order-independent.

__device__ __forceinline__ int tet_corners(const Lattice& L, uint32_t d, int id[4]) {
    int base = (int)(d >> 4), type = (int)((d >> 1) & 7u);
    for (int c = 0; c < 4; c++) id[c] = base + corner_offset(kuhn_corner(type, c), L.nx, L.nxy);
    return type;
}

__device__ __forceinline__ void type_rows(const Lattice& L, int type, f3 b[3]) {
    const float* B = L.inv_rest[type];
    for (int r = 0; r < 3; r++) b[r] = mk3(B[3 * r], B[3 * r + 1], B[3 * r + 2]);
}

// projectTensile (xpbd_solver.cu:591-626) over one colour: on the tet's six edges in order, one XPBD step on
// C = s^3, s = max(L / L0 - 1, 0), of compliance alphaDev / strength; the rest edges are the columns of Dm = B^-1.
__global__ void k_tensile(int off, int cnt, Lattice L, float strength, const uint32_t* __restrict__ tet_desc,
                          const float* __restrict__ alpha, float* __restrict__ fiber, float4* __restrict__ pos) {
    int k = ordered_item(L.xcd_order, GTS_BLOCK);
    if (k >= cnt) return;
    int e = off + k;
    uint32_t d = tet_desc[e];
    if (!(d & 1u)) return;
    int id[4];
    int type = tet_corners(L, d, id);
    float4 q[4];
    f3 p[4];
    for (int c = 0; c < 4; c++) { q[c] = pos[id[c]]; p[c] = xyz(q[c]); }
    if (strength > 0.f) {
        f3 b[3];
        type_rows(L, type, b);
        float detB = dot3(b[0], cross3(b[1], b[2]));
        f3 rest[4] = {mk3(0, 0, 0), cross3(b[1], b[2]) * (1.f / detB), cross3(b[2], b[0]) * (1.f / detB),
                      cross3(b[0], b[1]) * (1.f / detB)};
        const float a = alpha[e] / fmaxf(strength, 1e-10f);
        int f = 0;
        for (int i = 0; i < 4; i++)
            for (int j = i + 1; j < 4; j++, f++) {
                f3 edge = p[j] - p[i];
                float length = len3(edge), restLength = len3(rest[j] - rest[i]);
                float strain = fmaxf(length / restLength - 1.f, 0.f), gradient = 3.f * strain * strain / restLength;
                if (strain > 0.f && length > 0.f) {
                    float wi = q[i].w, wj = q[j].w, lam = fiber[6 * e + f];
                    float dl = (-strain * strain * strain - a * lam) / ((wi + wj) * gradient * gradient + a);
                    f3 g = edge * (gradient / length);
                    fiber[6 * e + f] = lam + dl;
                    p[i] = p[i] - g * (wi * dl);
                    p[j] = p[j] + g * (wj * dl);
                }
            }
    }
    for (int c = 0; c < 4; c++) pos[id[c]] = make_float4(p[c].x, p[c].y, p[c].z, q[c].w);
}

// k_CheckCandidateVolume (:711-717): flag when a live tet's J is below the minimum or not finite.
__global__ void k_guard_check(int T, Lattice L, float minimum, const uint32_t* __restrict__ tet_desc,
                              const float4* __restrict__ pos, int* flag) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T) return;
    uint32_t d = tet_desc[e];
    if (!(d & 1u)) return;
    int id[4];
    f3 b[3];
    type_rows(L, tet_corners(L, d, id), b);
    f3 p = xyz(pos[id[0]]), u = xyz(pos[id[1]]) - p, v = xyz(pos[id[2]]) - p, w = xyz(pos[id[3]]) - p;
    float J = dot3(u, cross3(v, w)) * dot3(b[0], cross3(b[1], b[2]));
    if (!isfinite(J) || J < minimum) atomicExch(flag, 1);
}

// k_RepairMinimumVolume (:694-709) over one colour, while the flag is up: one Newton step on J toward
// minimum + 2e-4 for a live tet below minimum + 1e-4.
__global__ void k_guard_repair(int off, int cnt, Lattice L, float minimum, const uint32_t* __restrict__ tet_desc,
                               float4* __restrict__ pos, const int* __restrict__ flag) {
    if (!*flag) return;
    int k = ordered_item(L.xcd_order, GTS_BLOCK);
    if (k >= cnt) return;
    uint32_t d = tet_desc[off + k];
    if (!(d & 1u)) return;
    int id[4];
    f3 b[3];
    type_rows(L, tet_corners(L, d, id), b);
    float4 q[4];
    for (int c = 0; c < 4; c++) q[c] = pos[id[c]];
    f3 p0 = xyz(q[0]), u = xyz(q[1]) - p0, v = xyz(q[2]) - p0, x = xyz(q[3]) - p0;
    float detB = dot3(b[0], cross3(b[1], b[2]));
    float J = dot3(u, cross3(v, x)) * detB;
    if (!isfinite(J) || J >= minimum + 1e-4f) return;
    f3 g[4];
    g[1] = cross3(v, x) * detB;
    g[2] = cross3(x, u) * detB;
    g[3] = cross3(u, v) * detB;
    g[0] = (g[1] + g[2] + g[3]) * -1.f;
    float den = 0.f;
    for (int c = 0; c < 4; c++) den += q[c].w * dot3(g[c], g[c]);
    if (den < 1e-20f) return;
    float dl = (minimum + 2e-4f - J) / den;
    for (int c = 0; c < 4; c++) {
        f3 y = xyz(q[c]) + g[c] * (q[c].w * dl);
        pos[id[c]] = make_float4(y.x, y.y, y.z, q[c].w);
    }
}

// k_NeedFeasible (:663-680), while the flag is up: J is multi-affine in the corners' interpolation parameters between
// the substep start and the candidate, so a Lipschitz bound or the 16 corner configurations certify a live tet; the
// corners of an uncertified one are marked.
__global__ void k_guard_need(int T, Lattice L, float minimum, const uint32_t* __restrict__ tet_desc,
                             const float4* __restrict__ start, const float4* __restrict__ pos, int* needed,
                             const int* __restrict__ flag) {
    if (!*flag) return;
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T) return;
    uint32_t d = tet_desc[e];
    if (!(d & 1u)) return;
    int id[4];
    f3 b[3];
    type_rows(L, tet_corners(L, d, id), b);
    f3 x[4], v[4];
    for (int c = 0; c < 4; c++) { x[c] = xyz(start[id[c]]); v[c] = xyz(pos[id[c]]) - x[c]; }
    float detB = dot3(b[0], cross3(b[1], b[2]));
    f3 u = x[1] - x[0], vv = x[2] - x[0], w = x[3] - x[0];
    float r0 = len3(v[0]), r1 = r0 + len3(v[1]), r2 = r0 + len3(v[2]), r3 = r0 + len3(v[3]);
    float bound = r1 * len3(cross3(vv, w)) + r2 * len3(cross3(w, u)) + r3 * len3(cross3(u, vv)) + r1 * r2 * len3(w) +
                  r1 * r3 * len3(vv) + r2 * r3 * len3(u) + r1 * r2 * r3;
    if (dot3(u, cross3(vv, w)) * detB - bound * fabsf(detB) >= minimum + 1e-5f) return;
    bool unsafe = false;
    for (int mask = 0; mask < 16 && !unsafe; mask++) {
        f3 p[4];
        for (int c = 0; c < 4; c++) p[c] = x[c] + v[c] * (float)((mask >> c) & 1);
        float j = dot3(p[1] - p[0], cross3(p[2] - p[0], p[3] - p[0])) * detB;
        unsafe = !isfinite(j) || j < minimum + 1e-5f;
    }
    if (unsafe)
        for (int c = 0; c < 4; c++) atomicExch(needed + id[c], 1);
}

// k_PrepareFeasible (:681-683): a marked node keeps its candidate as the target and restarts from the substep start.
__global__ void k_guard_prepare(int N, const int* __restrict__ needed, const float4* __restrict__ start,
                                float4* __restrict__ target, float4* __restrict__ pos) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || !needed[i]) return;
    target[i] = pos[i];
    pos[i] = start[i];
}

// k_FeasibleVertex (:634-661) over one lattice parity class: a marked node moves toward its target, cut back so that
// no live incident tet drops below min(minimum, max(0, J)) (J is linear in this node's move), with a 0.95 margin.
__global__ void k_guard_vertex(int N, int ny, Lattice L, int colour, float minimum, const int* __restrict__ offsets,
                               const int* __restrict__ incident, const uint32_t* __restrict__ tet_desc,
                               const float4* __restrict__ target, float4* __restrict__ pos,
                               const int* __restrict__ needed) {
    const int nx = L.nx, nz = N / (nx * ny), cx = (nx + 1) / 2, cy = (ny + 1) / 2;
    int lane = blockIdx.x * blockDim.x + threadIdx.x;
    int x = 2 * (lane % cx) + (colour & 1), y = 2 * ((lane / cx) % cy) + ((colour >> 1) & 1),
        z = 2 * (lane / (cx * cy)) + ((colour >> 2) & 1);
    if (x >= nx || y >= ny || z >= nz) return;
    int i = x + nx * (y + ny * z);
    if (!needed[i]) return;
    float4 pi = pos[i];
    f3 delta = xyz(target[i]) - xyz(pi);
    float accepted = 1.f;
    if (!isfinite(delta.x + delta.y + delta.z)) return;
    for (int k = offsets[i]; k < offsets[i + 1]; k++) {
        uint32_t d = tet_desc[incident[k]];
        if (!(d & 1u)) continue;
        int id[4];
        f3 b[3];
        type_rows(L, tet_corners(L, d, id), b);
        f3 a = xyz(pos[id[0]]), u = xyz(pos[id[1]]) - a, v = xyz(pos[id[2]]) - a, w = xyz(pos[id[3]]) - a;
        float detB = dot3(b[0], cross3(b[1], b[2]));
        float j = dot3(u, cross3(v, w)) * detB;
        f3 grad[3] = {cross3(v, w), cross3(w, u), cross3(u, v)};
        f3 g = i == id[0] ? (grad[0] + grad[1] + grad[2]) * -1.f : (i == id[1] ? grad[0] : (i == id[2] ? grad[1] : grad[2]));
        float dj = dot3(g, delta) * detB;
        if (dj < 0.f) {
            float floor_ = fminf(minimum, fmaxf(0.f, j));
            if (j + dj < floor_) accepted = fminf(accepted, fmaxf(0.f, 0.95f * (j - floor_) / (-dj)));
        }
    }
    f3 moved = xyz(pi) + delta * accepted;
    pos[i] = make_float4(moved.x, moved.y, moved.z, pi.w);
}

// ------------------------------------------------------------------------------------------------ particle frames

// Frames are stored as the three rows of R per node (RotGpu, cut.cuh:67), 9 floats.
__global__ void k_identity_frames(int N, float* rot, float* rot_prev) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 9 * N) return;
    float v = (i % 9) % 4 == 0 ? 1.f : 0.f;  // entries 0, 4, 8 of a row-major 3x3
    rot[i] = v;
    rot_prev[i] = v;
}

__global__ void k_copy_floats(int n, const float* __restrict__ src, float* __restrict__ dst) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// With a gate, the frame update runs only when the cut point counter moved during the tick (counters[1] against the
// value saved in counters[5] before it).
__device__ __forceinline__ bool frames_gated_off(const uint32_t* gate) { return gate && gate[1] == gate[5]; }

__global__ void k_copy_frames(int n, const float* __restrict__ src, float* __restrict__ dst, const uint32_t* gate) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !frames_gated_off(gate)) dst[i] = src[i];
}

// k_ComputeParticleRot (physics.cu:109-205) over one tissue's nodes.
__global__ void k_particle_frames(int N, const float4* __restrict__ pos, const int* __restrict__ nbr,
                                  const int* __restrict__ active, const int* __restrict__ pinned,
                                  const float* __restrict__ prev, float* __restrict__ rot, const uint32_t* gate) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < N && !frames_gated_off(gate)) particle_frame(c, 0, pos, nbr, active, pinned, prev, rot);
}

__global__ void k_save_cut_counter(uint32_t* counters) { counters[5] = counters[1]; }

// The ribbon tick's corner snapshot: the current positions as float triples.
__global__ void k_snapshot(int N, const float4* __restrict__ pos, float* __restrict__ snap) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 p = pos[i];
    snap[3 * i] = p.x;
    snap[3 * i + 1] = p.y;
    snap[3 * i + 2] = p.z;
}

// ------------------------------------------------------------------------------------------------ legacy grasp kernels

// The jaw frame of one substep's interpolated pose as xpbd_substep derives it for the grasp kernels
// (xpbd_solver.cu:3572-3580): unit axis, the normal made orthogonal to it, w = axis x normal, half the floored gap.
struct GraspFrame {
    f3 hinge, axis, normal, w, shaftA, shaftB;
    float shaftR, jawLen, jawHalfWidth, gapHalf;
};

// The legacy grasp state and the constants of one step, passed by value.
struct GraspArgs {
    int32_t* grasped;
    int32_t* core;
    const float* off;   // [3N] (u, n, w)
    float* anchor;      // [2N] (u, w)
    float* slab;
    float* slabPrev;
    float4* coreLam;
    float* ema;
    float* jawLam;
    float* shaftLam;
    float omega, contactAlpha, mu, maxDisp, coreAlpha, breakForce, kEMA, ccdOmega, ccdVel, ccdMaxDisp;
};

// Per-substep reset of the legacy multipliers (xpbd_solver.cu:3538-3545): the CCD gate's while the tool is armed, the
// grasp's while grasping. A fill kernel, so that a captured step graph replays it exactly.
__global__ void k_grasp_reset(int N, int gate, int grasp, GraspArgs a) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    if (gate) { a.jawLam[c] = 0.f; a.shaftLam[c] = 0.f; }
    if (grasp) { a.slab[c] = 0.f; a.coreLam[c] = make_float4(0.f, 0.f, 0.f, 0.f); }
}

// k_SolveSlabCoulomb (xpbd_solver.cu:1131-1199) on the ring corners (grasped, not core) and k_SolveCorePin (:960-977)
// on the core corners. Each corner reads and writes only itself, so the two launches of the reference fuse into one.
__global__ void k_grasp_pass(int N, const GraspFrame* __restrict__ frame, float h, GraspArgs a, float4* pos) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const int ring = a.grasped[c] != 0 && a.core[c] == 0, core = a.core[c] != 0;
    if (!ring && !core) return;
    float4 p4 = pos[c];
    const float wc = p4.w;
    if (!(wc > 0.f)) return;
    const GraspFrame f = *frame;
    f3 x = xyz(p4);
    if (ring) {
        f3 rel = x - f.hinge;
        float u = dot3(rel, f.axis), v = dot3(rel, f.w), n = dot3(rel, f.normal);
        if (u >= 0.f && u <= f.jawLen && fabsf(v) <= f.jawHalfWidth) {
            float absN = fabsf(n);
            if (absN > f.gapHalf) {
                float Ctrue = f.gapHalf - absN;
                float sgn = n >= 0.f ? 1.f : -1.f;
                float denom = wc + a.contactAlpha;
                float dLambda = (-Ctrue - a.contactAlpha * a.slab[c]) / denom;
                float applied = a.omega * dLambda;
                float corrMag = wc * applied;
                if (fabsf(corrMag) > a.maxDisp) {
                    float capScale = a.maxDisp / fabsf(corrMag);
                    applied *= capScale;
                    corrMag *= capScale;
                }
                a.slab[c] += applied;
                x = x - f.normal * (sgn * corrMag);
                rel = x - f.hinge;
                n = dot3(rel, f.normal);
            }
        }
        float Fn = fabsf(a.slabPrev[c]) / (h * h);
        float aU = a.anchor[2 * c], aW = a.anchor[2 * c + 1];
        f3 anchorWorld = f.hinge + f.axis * aU + f.w * aW + f.normal * n;
        f3 tangential = x - anchorWorld;
        float dist = len3(tangential);
        float coneRadius = a.mu * Fn * (h * h * wc);
        if (dist <= coneRadius + 1e-9f) {  // stick: compliant, capped pull to the anchor
            f3 corr = tangential * a.omega;
            float corrLen = len3(corr);
            if (corrLen > a.maxDisp) corr = corr * (a.maxDisp / corrLen);
            x = x - corr;
        } else {  // slip: move by the cone radius and drag the anchor along
            f3 dir = tangential * (1.f / dist);
            x = x - dir * coneRadius;
            f3 newRel = (anchorWorld + dir * (dist - coneRadius)) - f.hinge;
            a.anchor[2 * c] = dot3(newRel, f.axis);
            a.anchor[2 * c + 1] = dot3(newRel, f.w);
        }
    } else {
        f3 target = f.hinge + f.axis * a.off[3 * c] + f.normal * a.off[3 * c + 1] + f.w * a.off[3 * c + 2];
        float4 l4 = a.coreLam[c];
        f3 lam = mk3(l4.x, l4.y, l4.z);
        f3 applied = ((x - target) + lam * a.coreAlpha) * (-a.omega / (wc + a.coreAlpha));
        float corrLen = len3(applied) * wc;
        if (corrLen > a.maxDisp) applied = applied * (a.maxDisp / corrLen);
        lam = lam + applied;
        a.coreLam[c] = make_float4(lam.x, lam.y, lam.z, 0.f);
        x = x + applied * wc;
    }
    pos[c] = make_float4(x.x, x.y, x.z, wc);
}

// After the coupling passes: k_CoreBreakCheck (xpbd_solver.cu:982-1000) while grasping, k_SolveCCDContact (:1049-1123)
// while the tool is armed, (VEL) k_update_velocity, the slab multiplier kept for the next substep (:3710-3711) and,
// unless perpendicular edge damping must run in between, k_apply_velscale (:871-919). Per corner, in that order; with
// the tissue guard VEL is off and k_velocity follows its safeguard.
template <bool SCALE, bool VEL = true>
__global__ void k_legacy_finish(int N, const GraspFrame* __restrict__ frame, int grasping, int armed, float h,
                                float invH, float velScale, GraspArgs a, float4* __restrict__ pos,
                                const float4* __restrict__ prev, float4* __restrict__ vel) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float4 p4 = pos[c];
    const float wc = p4.w;
    f3 x = xyz(p4);
    if (grasping && a.core[c] != 0) {
        float raw = len3(xyz(a.coreLam[c])) / (h * h);
        float e = a.ema[c];
        e += (raw - e) * a.kEMA;
        a.ema[c] = e;
        if (e > a.breakForce) {
            a.core[c] = 0;
            a.grasped[c] = 0;
        }
    }
    if (armed && wc > 0.f) {
        const GraspFrame f = *frame;
        f3 vp = xyz(vel[c]);
        {  // shaft capsule, every corner
            f3 ab = f.shaftB - f.shaftA;
            float ab2 = dot3(ab, ab);
            float tt = ab2 > 1e-12f ? dot3(x - f.shaftA, ab) / ab2 : 0.f;
            tt = fmaxf(0.f, fminf(1.f, tt));
            f3 dvec = x - (f.shaftA + ab * tt);
            float dist = len3(dvec);
            float C = dist - f.shaftR;
            if (C < 0.f) {
                f3 n = dist > 1e-9f ? dvec * (1.f / dist) : mk3(0.f, 1.f, 0.f);
                float vApproach = -dot3(vp, n);
                if (vApproach > a.ccdVel) {
                    float denom = wc + a.contactAlpha;
                    float dLambda = (-C - a.contactAlpha * a.shaftLam[c]) / denom;
                    float applied = a.ccdOmega * dLambda;
                    f3 corr = n * (wc * applied);
                    float corrLen = len3(corr);
                    if (corrLen > a.ccdMaxDisp) {
                        float capScale = a.ccdMaxDisp / corrLen;
                        applied *= capScale;
                        corr = corr * capScale;
                    }
                    a.shaftLam[c] += applied;
                    x = x + corr;
                }
            }
        }
        if (a.grasped[c] == 0) {  // jaw slab as a contact box, corners not grasped
            f3 rel = x - f.hinge;
            float u = dot3(rel, f.axis), v = dot3(rel, f.w), n = dot3(rel, f.normal);
            if (u >= 0.f && u <= f.jawLen && fabsf(v) <= f.jawHalfWidth) {
                float absN = fabsf(n);
                if (absN > f.gapHalf) {
                    float Ctrue = f.gapHalf - absN;
                    float sgn = n >= 0.f ? 1.f : -1.f;
                    float vApproach = -dot3(vp, f.normal * sgn);
                    if (vApproach > a.ccdVel) {
                        float denom = wc + a.contactAlpha;
                        float dLambda = (-Ctrue - a.contactAlpha * a.jawLam[c]) / denom;
                        float applied = a.ccdOmega * dLambda;
                        float corrMag = wc * applied;
                        if (fabsf(corrMag) > a.ccdMaxDisp) {
                            float capScale = a.ccdMaxDisp / fabsf(corrMag);
                            applied *= capScale;
                            corrMag *= capScale;
                        }
                        a.jawLam[c] += applied;
                        x = x - f.normal * (sgn * corrMag);
                    }
                }
            }
        }
        pos[c] = make_float4(x.x, x.y, x.z, wc);
    }
    if (VEL) {
        f3 v = wc > 0.f ? (x - xyz(prev[c])) * invH : mk3(0, 0, 0);
        if (SCALE) v = wc > 0.f ? v * velScale : mk3(0, 0, 0);
        vel[c] = make_float4(v.x, v.y, v.z, 0.f);
    }
    if (grasping) a.slabPrev[c] = a.slab[c];
}

// k_MarkGraspCandidates (xpbd_solver.cu:930-951): live, unpinned corners inside the jaw box widened by the capture skin,
// with their jaw-local (u, v, n). scratch: [N] flags, then [3N] (u, v, n).
__global__ void k_mark_candidates(int N, f3 hinge, f3 axis, f3 normal, f3 w, float jawLen, float jawHalfWidth,
                                  float band, const int* active, const int* pinned, const float4* pos, float* scratch) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    int* cand = reinterpret_cast<int*>(scratch);
    float* uvn = scratch + N;
    cand[c] = 0;
    if (active[c] == 0 || pinned[c] != 0) return;
    f3 rel = xyz(pos[c]) - hinge;
    float u = dot3(rel, axis), v = dot3(rel, w), n = dot3(rel, normal);
    if (u < 0.f || u > jawLen) return;
    if (fabsf(v) > jawHalfWidth) return;
    if (fabsf(n) > band) return;
    cand[c] = 1;
    uvn[3 * c] = u;
    uvn[3 * c + 1] = v;
    uvn[3 * c + 2] = n;
}

// Release (xpbd_solver.cu:2456-2468): k_ReleaseDampVelocity on the grasped corners, then both sets cleared.
__global__ void k_release(int N, int32_t* grasped, int32_t* core, float4* vel) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    if (grasped[c] != 0) {
        float4 v = vel[c];
        vel[c] = make_float4(v.x * 0.1f, v.y * 0.1f, v.z * 0.1f, v.w);
    }
    grasped[c] = 0;
    core[c] = 0;
}

// The tear law's grasp ring (cut_recompute_grasp_ring, cut.cu:1513-1528): one hop of the dilation over live links
// (k_DilateRingHop, :687-698), then the cut-rim carve-out (k_CarveRimBridge, :703-710).
__global__ void k_ring_hop(int N, const int* __restrict__ nbr, const int* __restrict__ ring, int* __restrict__ next) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    int v = ring[c] != 0 ? 1 : 0;
    for (int k = 0; k < 6 && !v; k++) {
        int n = nbr[6 * c + k];
        if (n >= 0 && ring[n] != 0) v = 1;
    }
    next[c] = v;
}

__global__ void k_ring_carve(int N, const int* __restrict__ nbr, const int* __restrict__ nbr_init, int* ring) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N || ring[c] == 0) return;
    for (int k = 0; k < 6; k++)
        if (nbr_init[6 * c + k] >= 0 && nbr[6 * c + k] < 0) { ring[c] = 0; return; }
}

__global__ void k_copy_ints(int n, const int* __restrict__ src, int* __restrict__ dst) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ------------------------------------------------------------------------------------------------ cut kernels

struct CutToolD {
    f3 t1v0, t1v1, t1v2, t2v0, t2v1, t2v2, ncut, lo, hi;
    float d;
    int valid;
};


__device__ __forceinline__ bool moller_trumbore(f3 O, f3 D, f3 V0, f3 V1, f3 V2, float& t, float& u, float& v) {
    f3 L1 = V1 - V0, L2 = V2 - V0, Tv = V0 - O;
    float det = dot3(cross3(L2, D), L1);
    t = 0.f; u = 0.f; v = 0.f;
    if (fabsf(det) < 1e-12f) return false;
    float inv = 1.f / det;
    t = dot3(cross3(L2, Tv), L1) * inv;
    u = -dot3(cross3(L2, D), Tv) * inv;
    v = dot3(cross3(D, Tv), L1) * inv;
    const float EPS = 1e-5f;
    return u >= -EPS && v >= -EPS && u + v <= 1.f + EPS && t > 0.f && t <= 1.f;
}

// MarkAndEmit (cut.cu:218-254) without the emission: sets the cut bits and the directed fractions, and records
// (t, hit) for the ordered emission pass when the edge was cut for the first time inside the material.
__device__ void mark_edge(int e, int axis, int bx, int by, int bz, int idA, int idB, float t, f3 hit, const Lattice& L,
                          const uint32_t* occupied, uint32_t* cut_mask, float* edge_cut_t, const float* skin,
                          const int* inside, float4* records) {
    bool firstCut = false, sawAny = false, occ = false;
    for (int s = 0; s < 4; s++) {
        int dx, dy, dz, local;
        stencil(axis, s, dx, dy, dz, local);
        int x = bx + dx, y = by + dy, z = bz + dz;
        if (x < 0 || y < 0 || z < 0 || x >= L.vx || y >= L.vy || z >= L.vz) continue;
        int vid = x + L.vx * (y + L.vy * z);
        if (occupied[vid] == 0u) continue;
        occ = true;
        uint32_t bit = 1u << local;
        uint32_t old = atomicOr(cut_mask + vid, bit);
        if (!sawAny) { sawAny = true; firstCut = (old & bit) == 0u; }
    }
    if (!(firstCut && occ)) return;
    edge_cut_t[directed_slot(idA, idB, L.vx)] = t;
    edge_cut_t[directed_slot(idB, idA, L.vx)] = 1.f - t;
    if (!cut_hits_material(t, skin[directed_slot(idA, idB, L.vx)], inside[idA], inside[idB])) return;
    records[e] = make_float4(t, hit.x, hit.y, hit.z);
}

// k_DetectCutQuad (cut.cu:421-474), rod frozen: Möller–Trumbore of each deformed lattice edge against the swept quad.
__global__ void k_cut_quad(int Eg, Lattice L, CutToolD tool, float pad, const int* __restrict__ grid_edges,
                           const float4* __restrict__ pos, const uint32_t* occupied, uint32_t* cut_mask,
                           float* edge_cut_t, const float* skin, const int* inside, float4* records) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Eg) return;
    records[e] = make_float4(-1.f, 0.f, 0.f, 0.f);
    if (tool.valid == 0) return;
    int4 ge = reinterpret_cast<const int4*>(grid_edges)[e];
    int axis = ge.x;
    int idA = ge.y + L.nx * ge.z + L.nxy * ge.w;
    int idB = idA + axis_stride(L, axis);
    f3 O = xyz(pos[idA]);
    f3 D = xyz(pos[idB]) - O;
    f3 eMin = mk3(fminf(O.x, O.x + D.x), fminf(O.y, O.y + D.y), fminf(O.z, O.z + D.z));
    f3 eMax = mk3(fmaxf(O.x, O.x + D.x), fmaxf(O.y, O.y + D.y), fmaxf(O.z, O.z + D.z));
    if (eMax.x < tool.lo.x - pad || eMin.x > tool.hi.x + pad || eMax.y < tool.lo.y - pad || eMin.y > tool.hi.y + pad ||
        eMax.z < tool.lo.z - pad || eMin.z > tool.hi.z + pad)
        return;
    float ta, ua, va, tb, ub, vb;
    bool h1 = moller_trumbore(O, D, tool.t1v0, tool.t1v1, tool.t1v2, ta, ua, va);
    bool h2 = moller_trumbore(O, D, tool.t2v0, tool.t2v1, tool.t2v2, tb, ub, vb);
    if (!h1 && !h2) return;
    float t = (h1 && h2) ? fminf(ta, tb) : (h1 ? ta : tb);
    t = fminf(fmaxf(t, 0.f), 1.f);
    f3 hit = O + D * t;
    mark_edge(e, axis, ge.y, ge.z, ge.w, idA, idB, t, hit, L, occupied, cut_mask, edge_cut_t, skin, inside, records);
}

// k_DetectCutCCD (cut.cu:486-592), tissue moving from prev_positions to pos against the static blade S->E.
__global__ void k_cut_ccd(int Eg, Lattice L, f3 S, f3 E, f3 emitN, float pad, const int* __restrict__ grid_edges,
                          const int* nbr, const float* __restrict__ prevp, const float4* __restrict__ pos,
                          const uint32_t* occupied, uint32_t* cut_mask, float* edge_cut_t, const float* skin,
                          const int* inside, float4* records, float4* normals) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Eg) return;
    records[e] = make_float4(-1.f, 0.f, 0.f, 0.f);
    int4 ge = reinterpret_cast<const int4*>(grid_edges)[e];
    int axis = ge.x;
    int idA = ge.y + L.nx * ge.z + L.nxy * ge.w;
    int idB = idA + axis_stride(L, axis);
    if (nbr[6 * idA + 2 * axis + 1] < 0) return;
    f3 a0 = ld3(prevp + 3 * idA), a1 = xyz(pos[idA]), b0 = ld3(prevp + 3 * idB), b1 = xyz(pos[idB]);
    f3 sMin = mk3(fminf(S.x, E.x) - pad, fminf(S.y, E.y) - pad, fminf(S.z, E.z) - pad);
    f3 sMax = mk3(fmaxf(S.x, E.x) + pad, fmaxf(S.y, E.y) + pad, fmaxf(S.z, E.z) + pad);
    f3 eMin = mk3(fminf(fminf(a0.x, a1.x), fminf(b0.x, b1.x)), fminf(fminf(a0.y, a1.y), fminf(b0.y, b1.y)),
                  fminf(fminf(a0.z, a1.z), fminf(b0.z, b1.z)));
    f3 eMax = mk3(fmaxf(fmaxf(a0.x, a1.x), fmaxf(b0.x, b1.x)), fmaxf(fmaxf(a0.y, a1.y), fmaxf(b0.y, b1.y)),
                  fmaxf(fmaxf(a0.z, a1.z), fmaxf(b0.z, b1.z)));
    if (eMax.x < sMin.x || eMin.x > sMax.x || eMax.y < sMin.y || eMin.y > sMax.y || eMax.z < sMin.z ||
        eMin.z > sMax.z)
        return;
    f3 D = E - S, P0 = a0 - S, Pv = a1 - a0, Q0 = b0 - S, Qv = b1 - b0;
    float c0 = dot3(cross3(P0, Q0), D);
    float c1 = dot3(cross3(P0, Qv) + cross3(Pv, Q0), D);
    float c2 = dot3(cross3(Pv, Qv), D);
    float roots[2];
    int nr = 0;
    if (fabsf(c2) < 1e-14f) {
        if (fabsf(c1) > 1e-14f) {
            float r = -c0 / c1;
            if (r >= 0.f && r <= 1.f) roots[nr++] = r;
        }
    } else {
        float disc = c1 * c1 - 4.f * c2 * c0;
        if (disc >= 0.f) {
            float sq = sqrtf(disc);
            float q = -0.5f * (c1 + (c1 >= 0.f ? sq : -sq));
            float r0 = q / c2, r1 = fabsf(q) > 1e-20f ? c0 / q : r0;
            if (r0 > r1) { float tmp = r0; r0 = r1; r1 = tmp; }
            if (r0 >= 0.f && r0 <= 1.f) roots[nr++] = r0;
            if (r1 >= 0.f && r1 <= 1.f && r1 != r0) roots[nr++] = r1;
        }
    }
    for (int ri = 0; ri < nr; ri++) {
        float tc = roots[ri];
        f3 a = a0 + Pv * tc, b = b0 + Qv * tc, e1 = b - a, n2 = cross3(e1, D);
        float n2l = dot3(n2, n2);
        if (n2l < 1e-16f) continue;
        f3 w = S - a;
        float u = dot3(cross3(w, D), n2) / n2l;
        float v = dot3(cross3(w, e1), n2) / n2l;
        const float EPS = 2e-2f;
        if (u <= 0.f || u > 1.f || v < -EPS || v > 1.f + EPS) continue;
        f3 hit = a1 + (b1 - a1) * u;
        f3 relV = (Pv + Qv) * 0.5f;
        f3 nHit = cross3(D, relV);
        float nLen = len3(nHit);
        f3 nEmit = nLen > 1e-8f ? nHit * (1.f / nLen) : emitN;
        normals[e] = make_float4(nEmit.x, nEmit.y, nEmit.z, 0.f);
        mark_edge(e, axis, ge.y, ge.z, ge.w, idA, idB, u, hit, L, occupied, cut_mask, edge_cut_t, skin, inside,
                  records);
        break;
    }
}

// k_TearOverstretch (cut.cu:607-681): an intact link between active corners, not yet marked and with an occupied
// stencil voxel, longer than tau times its rest length (tearLen2 scaled by |rest edge|^2 / L^2) is marked at its
// midpoint with the plane normal to it. A link with a grasp-ring corner is exempt (and restarts its release decay);
// out of the ring it stays exempt while longer than 0.8 tau L, for up to GTS_TEAR_DECAY_TICKS ticks. Each edge reads
// and sets only its own mask bits and exemption state, so the edges are independent. counters[4] counts the marked
// edges (the reference's dbg[15]).
__global__ void k_tear(int Eg, Lattice L, float tearLen2, float voxelL, const int* __restrict__ grid_edges,
                       const int* nbr, const int* active, const float4* __restrict__ pos, const float* __restrict__ rest,
                       const uint32_t* occupied, uint32_t* cut_mask, float* edge_cut_t, const float* skin,
                       const int* inside, float4* records, float4* normals, uint32_t* counters,
                       const int* __restrict__ ring, int* __restrict__ exempt) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Eg) return;
    records[e] = make_float4(-1.f, 0.f, 0.f, 0.f);
    int4 ge = reinterpret_cast<const int4*>(grid_edges)[e];
    int axis = ge.x;
    int idA = ge.y + L.nx * ge.z + L.nxy * ge.w;
    int idB = idA + axis_stride(L, axis);
    if (nbr[6 * idA + 2 * axis + 1] < 0) return;
    if (active[idA] == 0 || active[idB] == 0) return;
    bool cut = false, occ = false;
    for (int s = 0; s < 4; s++) {
        int dx, dy, dz, local;
        stencil(axis, s, dx, dy, dz, local);
        int x = ge.y + dx, y = ge.z + dy, z = ge.w + dz;
        if (x < 0 || y < 0 || z < 0 || x >= L.vx || y >= L.vy || z >= L.vz) continue;
        int vid = x + L.vx * (y + L.vy * z);
        if ((cut_mask[vid] >> local) & 1u) cut = true;
        if (occupied[vid] != 0u) occ = true;
    }
    if (cut || !occ) return;
    f3 A = xyz(pos[idA]), d = xyz(pos[idB]) - A;
    float len2 = dot3(d, d);
    f3 r = ld3(rest + 3 * idB) - ld3(rest + 3 * idA);
    float limit = tearLen2 * (dot3(r, r) / (voxelL * voxelL));
    if (ring && exempt) {  // the grasp-ring exemption and its release decay (cut.cu:647-673); exempt: flags, ticks
        if (ring[idA] != 0 || ring[idB] != 0) {
            exempt[e] = 1;
            exempt[Eg + e] = 0;
            return;
        }
        if (exempt[e] != 0) {
            float decayLen2 = 0.64f * limit;
            if (len2 > decayLen2) {
                int ticks = exempt[Eg + e] + 1;
                exempt[Eg + e] = ticks;
                if (ticks < GTS_TEAR_DECAY_TICKS) return;
            }
            exempt[e] = 0;
            exempt[Eg + e] = 0;
        }
    }
    if (len2 <= limit) return;
    float len = sqrtf(len2);
    f3 ncut = mk3(d.x / len, d.y / len, d.z / len);
    f3 hit = A + d * 0.5f;
    normals[e] = make_float4(ncut.x, ncut.y, ncut.z, 0.f);
    atomicAdd(counters + 4, 1u);
    mark_edge(e, axis, ge.y, ge.z, ge.w, idA, idB, 0.5f, hit, L, occupied, cut_mask, edge_cut_t, skin, inside, records);
}


struct FlagRecords {
    const float4* r;
    __device__ bool operator()(int i) const { return r[i].x >= 0.f; }
};
struct FlagDescLive {
    const uint32_t* d;
    __device__ bool operator()(int i) const { return (d[i] & 1u) != 0u; }
};
struct FlagNodeUsed {
    const int* used;
    const int* active;
    __device__ bool operator()(int i) const { return used[i] != 0 && active[i] != 0; }
};




// EmitCutPoint (cut.cu:186-215) for the edges marked by the detection, in lattice edge order: slot = counter +
// 2 * (rank of the edge among the emitting ones) for the owner A and + 1 for B.
__global__ void k_emit_points(int Eg, Lattice L, f3 ncutQuad, int useNormals, float toolD, float voxelL,
                              const int* __restrict__ grid_edges, const float4* __restrict__ pos,
                              const float* __restrict__ rest, const float* __restrict__ rot, const float4* records,
                              const float4* normals, const int* partial, uint32_t* counters, float* points,
                              int capacity) {
    __shared__ int lds[GTS_BLOCK / 64];
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    bool f = e < Eg && records[e].x >= 0.f;
    int total;
    int r = block_prefix(f, lds, total);
    uint32_t base = counters[1];
    if (f) {
        float4 rec = records[e];
        int4 ge = reinterpret_cast<const int4*>(grid_edges)[e];
        int axis = ge.x;
        int idA = ge.y + L.nx * ge.z + L.nxy * ge.w;
        int idB = idA + axis_stride(L, axis);
        float t = rec.x;
        f3 hit = mk3(rec.y, rec.z, rec.w);
        f3 ncut = ncutQuad;
        if (useNormals) { float4 nn = normals[e]; ncut = mk3(nn.x, nn.y, nn.z); }
        // rest edge from the corner rest positions when the view has them (materialRestDevice, cut.cu:248)
        f3 axisF = mk3(axis == 0 ? 1.f : 0.f, axis == 1 ? 1.f : 0.f, axis == 2 ? 1.f : 0.f);
        f3 restEdge = rest ? ld3(rest + 3 * idB) - ld3(rest + 3 * idA) : axisF * voxelL;
        f3 along[2] = {restEdge * t, restEdge * (t - 1.f)};
        int owners[2] = {idA, idB};
        uint32_t slot0 = base + 2u * (uint32_t)(partial[blockIdx.x] + r);
        for (int k = 0; k < 2; k++) {
            int p = owners[k];
            float side = dot3(xyz(pos[p]) - hit, ncut);
            side = side > 0.f ? 1.f : (side < 0.f ? -1.f : 0.f);
            if (side == 0.f) side = 1.f;
            uint32_t slot = slot0 + (uint32_t)k;
            if (slot >= (uint32_t)capacity) continue;
            // the gap and the normal in the owner's material frame: R^T v (mulRtv, cut.cuh:75-78)
            const float* Rp = rot + 9 * (size_t)p;
            const f3 r0 = ld3(Rp), r1 = ld3(Rp + 3), r2 = ld3(Rp + 6);
            f3 g = ncut * (side * (toolD * 0.5f)), sn = ncut * side;
            f3 gap = r0 * g.x + r1 * g.y + r2 * g.z;
            f3 off = along[k] + gap;
            f3 nrm = r0 * sn.x + r1 * sn.y + r2 * sn.z;
            float* cp = points + 8 * (size_t)slot;
            cp[0] = off.x; cp[1] = off.y; cp[2] = off.z;
            cp[3] = __int_as_float(p);
            cp[4] = __int_as_float(e);
            cp[5] = nrm.x; cp[6] = nrm.y; cp[7] = nrm.z;
        }
    }
}

__global__ void k_bump_counter(const int* partial, int nblocks, uint32_t* counters) {
    counters[1] += 2u * (uint32_t)partial[nblocks];
}

// k_SeverLinks / SeverLink (cut.cu:719-746): part the links whose edge bit is set in any stencil voxel, kill the bend
// pairs that cross them, count the intact links severed.
__global__ void k_sever(int Eg, Lattice L, const int* __restrict__ grid_edges, const uint32_t* cut_mask, int* nbr,
                        uint32_t* bend_alive, uint32_t* counters) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Eg) return;
    int4 ge = reinterpret_cast<const int4*>(grid_edges)[e];
    int axis = ge.x;
    bool valid = true, cut = false;
    for (int s = 0; s < 4; s++) {
        int dx, dy, dz, local;
        stencil(axis, s, dx, dy, dz, local);
        int x = ge.y + dx, y = ge.z + dy, z = ge.w + dz;
        if (x < 0 || y < 0 || z < 0 || x >= L.vx || y >= L.vy || z >= L.vz) { valid = false; continue; }
        if ((cut_mask[x + L.vx * (y + L.vy * z)] >> local) & 1u) cut = true;
    }
    if (!valid || !cut) return;
    int a = ge.y + L.nx * ge.z + L.nxy * ge.w;
    int b = a + axis_stride(L, axis);
    int sa = 2 * axis + 1, sb = 2 * axis;
    if (nbr[6 * a + sa] >= 0) atomicAdd(counters, 1u);
    nbr[6 * a + sa] = -1;
    nbr[6 * b + sb] = -1;
    // pair mask of all bend pairs of a node that use slot s
    uint32_t ma = 0u, mb = 0u;
    int pair = 0;
    for (int p = 0; p < 6; p++)
        for (int q = p + 1; q < 6; q++, pair++) {
            if (p == sa || q == sa) ma |= 1u << pair;
            if (p == sb || q == sb) mb |= 1u << pair;
        }
    atomicAnd(bend_alive + a, ~ma);
    atomicAnd(bend_alive + b, ~mb);
}

__global__ void k_fill_int(int n, int* x, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}
__global__ void k_fill_float(int n, float* x, float v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ void k_mark_used(int T, int E, Lattice L, const uint32_t* tet_desc, const uint32_t* edge_desc, int* used) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < T) {
        uint32_t d = tet_desc[i];
        if (d & 1u) {
            int base = (int)(d >> 4), type = (int)((d >> 1) & 7u);
            for (int c = 0; c < 4; c++) used[base + corner_offset(kuhn_corner(type, c), L.nx, L.nxy)] = 1;
        }
    }
    if (i < E) {
        uint32_t d = edge_desc[i];
        if (d & 1u) {
            int a = (int)(d >> 4), axis = (int)((d >> 1) & 3u);
            used[a] = 1;
            used[a + axis_stride(L, axis)] = 1;
        }
    }
}

__global__ void k_copy_counts(const int* pt, int nt, const int* pe, int ne, const int* pn, int nn, int* counts) {
    counts[0] = pt[nt];
    counts[1] = pe[ne];
    counts[2] = pn[nn];
}

// ------------------------------------------------------------------------------------------------ host helpers

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Optional per-launch timing of the dominant kernel (the tet colour pass) with HIP events on the launch stream,
// in the spirit of the reference's xpbd_profile sections (xpbd/xpbd_solver.cu:15-60).
struct Profile {
    bool on = false;
    std::vector<hipEvent_t> events;  // start, stop pairs
    std::vector<long long> units;
    size_t used = 0;
    hipError_t error = hipSuccess;
};
Profile g_profile;

// The first failing event call of a profiling window is kept and reported by gts_profile_end.
inline void profile_note(hipError_t e) {
    if (e != hipSuccess && g_profile.error == hipSuccess) g_profile.error = e;
}

template <class T> struct same_type { using type = T; };

// Launches `kern` on stream s. While profiling, the launch goes through hipExtLaunchKernel with a start and a stop
// event that the runtime stamps when the kernel itself starts and ends, so the recorded duration is the kernel's own
// (no enqueue or dispatch latency), and `units` (tets) are credited to the launch.
template <class... A>
void profiled_launch_lds(void (*kern)(A...), dim3 g, dim3 b, size_t lds, hipStream_t s, long long units,
                         typename same_type<A>::type... args) {
    if (!g_profile.on) {
        kern<<<g, b, lds, s>>>(args...);
        return;
    }
    if (2 * g_profile.used + 2 > g_profile.events.size()) {
        size_t old = g_profile.events.size();
        g_profile.events.resize(old + 4096);
        for (size_t k = old; k < g_profile.events.size(); k++) profile_note(hipEventCreate(&g_profile.events[k]));
    }
    void* ptrs[] = {(void*)&args...};
    profile_note(hipExtLaunchKernel(reinterpret_cast<const void*>(kern), g, b, ptrs, lds, s,
                                    g_profile.events[2 * g_profile.used], g_profile.events[2 * g_profile.used + 1],
                                    0));
    if (g_profile.units.size() <= g_profile.used) g_profile.units.resize(g_profile.used + 1024);
    g_profile.units[g_profile.used] = units;
    g_profile.used++;
}
template <class... A>
void profiled_launch(void (*kern)(A...), dim3 g, dim3 b, hipStream_t s, long long units,
                     typename same_type<A>::type... args) {
    profiled_launch_lds(kern, g, b, 0, s, units, args...);
}

int scan_blocks(int n) { return grid_for(n); }

// scratch layout in floats: [4*Eg) detection records, [4*Eg, 8*Eg) CCD normals, then int partials
int* partial_area(const GtsTissueView* v) { return reinterpret_cast<int*>(v->scratch + 8 * (size_t)v->n_grid_edges); }

// physics_compute_rotations (physics.cu:399-411): the current frames become the previous ones, then new frames.
void launch_frames(const GtsTissueView* v, hipStream_t s, const uint32_t* gate = nullptr) {
    const int N = v->n_nodes;
    k_copy_frames<<<grid_for(9 * N), GTS_BLOCK, 0, s>>>(9 * N, v->rot, v->rot_prev, gate);
    k_particle_frames<<<grid_for(N), GTS_BLOCK, 0, s>>>(N, (const float4*)v->pos, v->nbr, v->active, v->pinned,
                                                        v->rot_prev, v->rot, gate);
}

int launch_refresh(const GtsTissueView* v, const Lattice& L, hipStream_t s) {
    k_topology_stale<<<1, 1, 0, s>>>(v->counters);
    k_clear_support<<<grid_for(6 * v->n_nodes), GTS_BLOCK, 0, s>>>(6 * v->n_nodes, v->slot_support, v->counters);
    if (v->n_tets > 0)
        k_refresh_tets<<<grid_for(v->n_tets), GTS_BLOCK, 0, s>>>(v->n_tets, L, v->nbr, v->active, v->tet_desc,
                                                                  v->slot_support, v->counters);
    if (v->n_edges > 0)
        k_refresh_edges<<<grid_for(v->n_edges), GTS_BLOCK, 0, s>>>(v->n_edges, L, v->nbr, v->active, v->slot_support,
                                                                    v->edge_desc, v->counters);
    return -(int)hipGetLastError();
}


template <int MODE, bool DAMPED, bool PRECISE>
void launch_tet_colour_d(const GtsTissueView* v, const Lattice& L, hipStream_t s, int off, int cnt, int it, int iters,
                         float omega, float dOverH, float bOverH) {
    const int cb = colour_block(v);
    dim3 g(grid_for(cnt, cb)), b(cb);
    const bool read = it > 0, write = iters > 1;
#define GTS_TET(R, W)                                                                                             \
    profiled_launch(k_tets<MODE, R, W, DAMPED, PRECISE>, g, b, s, cnt, off, cnt, v->n_tets, L, omega, dOverH,     \
                    bOverH, v->tet_desc, v->tet_alpha, v->tet_lambda, (float4*)v->pos, (const float4*)v->prev)
    if (read && write) GTS_TET(true, true);
    else if (read) GTS_TET(true, false);
    else if (write) GTS_TET(false, true);
    else GTS_TET(false, false);
#undef GTS_TET
}

template <int MODE>
void launch_tet_colour(const GtsTissueView* v, const Lattice& L, hipStream_t s, int off, int cnt, int it, int iters,
                       float omega, float dOverH, float bOverH, bool precise) {
    const bool damped = dOverH != 0.f || bOverH != 0.f;
    if (damped && precise) launch_tet_colour_d<MODE, true, true>(v, L, s, off, cnt, it, iters, omega, dOverH, bOverH);
    else if (damped) launch_tet_colour_d<MODE, true, false>(v, L, s, off, cnt, it, iters, omega, dOverH, bOverH);
    else if (precise) launch_tet_colour_d<MODE, false, true>(v, L, s, off, cnt, it, iters, omega, dOverH, bOverH);
    else launch_tet_colour_d<MODE, false, false>(v, L, s, off, cnt, it, iters, omega, dOverH, bOverH);
}

template <bool DAMPED>
void launch_edge_colour_d(const GtsTissueView* v, const Lattice& L, hipStream_t s, int off, int cnt, int it, int iters,
                          bool fallback, float at, float fb, float omega, float bgh) {
    const int cb = colour_block(v);
    dim3 g(grid_for(cnt, cb)), b(cb);
    const bool read = it > 0, write = iters > 1;
#define GTS_EDGE(R, W, FB)                                                                                     \
    k_edges<R, W, FB, DAMPED><<<g, b, 0, s>>>(off, cnt, L, at, fb, omega, bgh, v->edge_desc, v->edge_lambda,  \
                                              (float4*)v->pos, (const float4*)v->prev)
    if (fallback) {
        if (read && write) GTS_EDGE(true, true, true);
        else if (read) GTS_EDGE(true, false, true);
        else if (write) GTS_EDGE(false, true, true);
        else GTS_EDGE(false, false, true);
    } else {
        if (read && write) GTS_EDGE(true, true, false);
        else if (read) GTS_EDGE(true, false, false);
        else if (write) GTS_EDGE(false, true, false);
        else GTS_EDGE(false, false, false);
    }
#undef GTS_EDGE
}

void launch_edge_colour(const GtsTissueView* v, const Lattice& L, hipStream_t s, int off, int cnt, int it, int iters,
                        bool fallback, float at, float fb, float omega, float bgh) {
    if (bgh != 0.f) launch_edge_colour_d<true>(v, L, s, off, cnt, it, iters, fallback, at, fb, omega, bgh);
    else launch_edge_colour_d<false>(v, L, s, off, cnt, it, iters, fallback, at, fb, omega, bgh);
}

template <bool DAMPED>
void launch_bend_colour_d(const GtsTissueView* v, const Lattice& L, hipStream_t s, int off, int cnt, int it, int iters,
                          float alphaB, float omega, float bgh) {
    const int cb = colour_block(v);
    dim3 g(grid_for(cnt, cb)), b(cb);
    const bool read = it > 0, write = iters > 1;
#define GTS_BEND(R, W)                                                                                          \
    k_bends<R, W, DAMPED><<<g, b, 0, s>>>(off, cnt, L, alphaB, omega, bgh, v->bend_desc, v->bend_alive,       \
                                          v->bend_lambda, (float4*)v->pos, (const float4*)v->prev)
    if (read && write) GTS_BEND(true, true);
    else if (read) GTS_BEND(true, false);
    else if (write) GTS_BEND(false, true);
    else GTS_BEND(false, false);
#undef GTS_BEND
}

// One iteration's tet sweep (EDGES: edge sweep) as the tiled launches of the plan's groups of that kind: positions
// alternate between pos and pos_alt from launch to launch (an even number of them, so the sweep ends in pos); each
// entry's multipliers are read from its copy, which the previous iteration's projection of that entry wrote.
template <int MODE, bool PRECISE, bool EDGES>
void launch_tiled_sweep(const GtsTissueView* v, const Lattice& L, hipStream_t s, int it, int iters, float omega,
                        float alphaTilde) {
    const bool read = it > 0, write = iters > 1;
    TileLaunch A;
    A.box = v->tile_box;
    A.region = v->tile_region;
    A.rx = v->tile_rx;
    A.ry = v->tile_ry;
    A.T = v->n_tets;
    const int nreg = v->tile_group_info[3];  // every launch takes the same regions
    for (int r = 0; r < nreg; r++) A.dense[r] = v->tile_dense_ids[r];
    const size_t lds = (size_t)v->tile_box * v->tile_box * v->tile_box * sizeof(float4);
    const int* offsets = EDGES ? v->edge_colour_offsets : v->tet_colour_offsets;
    float* lam_out = EDGES ? v->edge_lambda : v->tet_lambda;
    for (int g = 0, n = 0; g < v->tile_groups; g++) {
        const int32_t* gi = v->tile_group_info + 16 * g;
        if (gi[14] != (EDGES ? 1 : 0)) continue;
        const int regions = gi[3];
        A.k = gi[1];
        A.halo = gi[5];
        A.stride = gi[6];
        A.wcap = gi[8];
        for (int p = 0, start = 0; p < GTS_TILE_KMAX; p++) {
            A.cap[p] = p < A.k ? gi[10 + p] : 0;
            A.start[p] = start;
            start += A.cap[p];
        }
        const float4* in = reinterpret_cast<const float4*>((n & 1) ? v->pos_alt : v->pos);
        float4* out = reinterpret_cast<float4*>((n & 1) ? v->pos : v->pos_alt);
        n++;
        if (regions <= 0) continue;
        // updates of the launch (its colours' elements), credited while profiling
        const long long units = (long long)offsets[gi[0] + gi[1]] - offsets[gi[0]];
// the profiled launches are the tet sweep's (their units are tet updates); edge launches go out directly
#define GTS_TILED_ARGS                                                                                            \
    A, reinterpret_cast<const int2*>(v->tile_entries) + gi[7], reinterpret_cast<const float4*>(v->tile_inputs) + \
        gi[7], reinterpret_cast<float2*>(v->tile_lambda) + gi[7],                                                  \
        reinterpret_cast<const int2*>(v->tile_writes) + gi[9], L, omega, alphaTilde, lam_out, in, out
#define GTS_TILED(R, W)                                                                                              \
    do {                                                                                                             \
        if (EDGES)                                                                                                   \
            k_tets_tiled<MODE, R, W, PRECISE, EDGES><<<dim3(regions), dim3(GTS_TILE_THREADS), lds, s>>>(              \
                GTS_TILED_ARGS);                                                                                     \
        else                                                                                                         \
            profiled_launch_lds(k_tets_tiled<MODE, R, W, PRECISE, EDGES>, dim3(regions), dim3(GTS_TILE_THREADS), lds, \
                                s, units, GTS_TILED_ARGS);                                                           \
    } while (0)
        if (read && write) GTS_TILED(true, true);
        else if (read) GTS_TILED(true, false);
        else if (write) GTS_TILED(false, true);
        else GTS_TILED(false, false);
#undef GTS_TILED
#undef GTS_TILED_ARGS
    }
}

void launch_bend_colour(const GtsTissueView* v, const Lattice& L, hipStream_t s, int off, int cnt, int it, int iters,
                        float alphaB, float omega, float bgh) {
    if (bgh != 0.f) launch_bend_colour_d<true>(v, L, s, off, cnt, it, iters, alphaB, omega, bgh);
    else launch_bend_colour_d<false>(v, L, s, off, cnt, it, iters, alphaB, omega, bgh);
}

// The ordered emission of the cut points of the edges the last detection kernel recorded (records at v->scratch).
void emit_records(const GtsTissueView* v, const Lattice& L, hipStream_t s, f3 ncut, int useNormals, float toolD) {
    int Eg = v->n_grid_edges;
    int nb = scan_blocks(Eg);
    int* partial = partial_area(v);
    float4* records = reinterpret_cast<float4*>(v->scratch);
    float4* normals = records + Eg;
    k_count_flags<<<nb, GTS_BLOCK, 0, s>>>(Eg, FlagRecords{records}, partial);
    k_scan_partials<<<1, GTS_BLOCK, 0, s>>>(nb, partial);
    k_emit_points<<<nb, GTS_BLOCK, 0, s>>>(Eg, L, ncut, useNormals, toolD, v->voxel_length, v->grid_edges,
                                           (const float4*)v->pos, v->rest, v->rot, records, normals, partial,
                                           v->counters, v->cut_points, v->cut_point_capacity);
    k_bump_counter<<<1, 1, 0, s>>>(partial, nb, v->counters);
}

// k_SeverLinks, then the liveness refresh.
int sever_and_refresh(const GtsTissueView* v, const Lattice& L, hipStream_t s) {
    int Eg = v->n_grid_edges;
    k_sever<<<grid_for(Eg), GTS_BLOCK, 0, s>>>(Eg, L, v->grid_edges, v->cut_mask, v->nbr, v->bend_alive, v->counters);
    return launch_refresh(v, L, s);
}

int finish_cut(const GtsTissueView* v, const Lattice& L, hipStream_t s, f3 ncut, int useNormals, float toolD) {
    emit_records(v, L, s, ncut, useNormals, toolD);
    return sever_and_refresh(v, L, s);
}

}  // namespace

// ------------------------------------------------------------------------------------------------ C entry points

extern "C" int gts_init_state(const GtsTissueView* v, void* stream) {
    if (!v || v->n_nodes <= 0) return GTS_ERR_BAD_ARG;
    hipStream_t s = S(stream);
    if (!v->rot || !v->rot_prev || !v->rest) return GTS_ERR_BAD_ARG;
    k_init_state<<<grid_for(v->n_nodes), GTS_BLOCK, 0, s>>>(v->n_nodes, v->mass, v->active, v->pinned, v->nbr,
                                                            (float4*)v->pos, v->bend_alive, v->counters);
    k_identity_frames<<<grid_for(9 * v->n_nodes), GTS_BLOCK, 0, s>>>(v->n_nodes, v->rot, v->rot_prev);
    HIP_TRY(hipGetLastError());
    Lattice L = lattice_of(v);
    return launch_refresh(v, L, s);
}

extern "C" int gts_set_gravity(const GtsTissueView* v, float gx, float gy, float gz, void* stream) {
    if (!v || !std::isfinite(gx) || !std::isfinite(gy) || !std::isfinite(gz)) return GTS_ERR_BAD_ARG;
    k_set_gravity<<<grid_for(v->n_nodes), GTS_BLOCK, 0, S(stream)>>>(v->n_nodes, v->mass, gx, gy, gz,
                                                                       (float4*)v->force);
    return -(int)hipGetLastError();
}

extern "C" int gts_refresh_liveness(const GtsTissueView* v, void* stream) {
    if (!v) return GTS_ERR_BAD_ARG;
    Lattice L = lattice_of(v);
    return launch_refresh(v, L, S(stream));
}

extern "C" int gts_compute_rotations(const GtsTissueView* v, void* stream) {
    if (!v || !v->rot || !v->rot_prev || v->n_nodes <= 0) return GTS_ERR_BAD_ARG;
    launch_frames(v, S(stream));
    return -(int)hipGetLastError();
}

namespace {
// The {min J, count} accumulator starts at {1, 0} (xpbd_get_material_quality's seed, xpbd_solver.cu:2136); a fill
// kernel rather than a copy or memset, so the call can sit in a captured graph.
__global__ void k_quality_seed(float* __restrict__ out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out[0] = 1.f;
        out[1] = 0.f;
    }
}

// k_MaterialQuality (xpbd_solver.cu:2122-2132) per live tet: F = [x1 - x0, x2 - x0, x3 - x0] Dm^-1, J = f0 . (f1 x f2);
// a non-finite J is counted and goes no further; otherwise J lowers the minimum (a compare-and-swap loop on the
// float, exact in any order) and is counted when J <= 0. Counts are float adds of 1, exact below 2^24 tets.
__global__ void k_material_quality(int T, Lattice L, const uint32_t* __restrict__ tet_desc,
                                   const float4* __restrict__ pos, float* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T) return;
    const uint32_t d = tet_desc[e];
    if (!(d & 1u)) return;
    const int base = (int)(d >> 4), type = (int)((d >> 1) & 7u);
    int id[4];
    for (int c = 0; c < 4; c++) id[c] = base + corner_offset(kuhn_corner(type, c), L.nx, L.nxy);
    const f3 a = xyz(pos[id[0]]);
    const f3 e1 = xyz(pos[id[1]]) - a, e2 = xyz(pos[id[2]]) - a, e3 = xyz(pos[id[3]]) - a;
    const float* B = L.inv_rest[type];
    f3 f[3];
#pragma unroll
    for (int k = 0; k < 3; k++) f[k] = e1 * B[k] + e2 * B[3 + k] + e3 * B[6 + k];
    const float j = dot3(f[0], cross3(f[1], f[2]));
    if (!isfinite(j)) {
        atomicAdd(out + 1, 1.f);
        return;
    }
    unsigned int* address = reinterpret_cast<unsigned int*>(out);
    unsigned int old = __hip_atomic_load(address, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (j < __uint_as_float(old)) {
        const unsigned int assumed = old;
        old = atomicCAS(address, assumed, __float_as_uint(j));
        if (old == assumed) break;
    }
    if (j <= 0.f) atomicAdd(out + 1, 1.f);
}
}  // namespace

extern "C" int gts_material_quality(const GtsTissueView* v, float* out, void* stream) {
    if (!v || !out || v->n_tets < 0 || (v->n_tets > 0 && (!v->tet_desc || !v->pos))) return GTS_ERR_BAD_ARG;
    hipStream_t s = S(stream);
    k_quality_seed<<<1, 64, 0, s>>>(out);
    if (v->n_tets > 0)
        k_material_quality<<<grid_for(v->n_tets), GTS_BLOCK, 0, s>>>(v->n_tets, lattice_of(v), v->tet_desc,
                                                                     reinterpret_cast<const float4*>(v->pos), out);
    return -(int)hipGetLastError();
}

namespace {

static_assert(sizeof(StepShapes) == GTS_STEP_CONST_FLOATS * sizeof(float), "step constant layout");

struct StepPlan {  // host-side constants of one outer step (xpbd_begin_frame, xpbd_solver.cu:2481-2514)
    int Nsub, iters, mode;
    float h, at, velScale, betaGammaH, invH2, fb, dOverH, bOverH, alphaC, scaleR, invH, omega, beta_s;
    float mu_s, mu_d, radius;
    bool strain, fallback, edges, contact, perp, precise;
    uint32_t armed;
    // legacy contact model: K material passes (2 while grasping), the grasp and CCD gate kernels
    bool legacy, grasping, gate;
    int K;
    GraspArgs grasp;
    // the CCD ribbon tick after every substep while a cut tool is set: rod S -> E, fallback normal, gap, tear limit
    bool tick;
    f3 tickS, tickE, tickN;
    float tickD, tearLen2;
    // the tissue guard: tensile sweeps per tet sweep, their strength, the safeguard's minimum J
    int guardPasses;
    float strength, guardMin;
    // bending: compliance 1/(kb h^2) and its Kelvin-Voigt factor (xpbd_begin_frame, xpbd_solver.cu:2496-2499)
    bool bend;
    float alphaB, betaGammaHB;
    // the tet and edge sweeps as tiled launches (GtsTissueView.tile_*)
    bool tiled, tiledEdges;
};

// cut_ribbon_tick without the pressure blade (cut.cu:1303-1431) after substep n: CCD of the edges moving from the
// snapshot to the current positions, the tear law, one sever and liveness refresh, the snapshot rolled; a tick that
// emitted cut points, but the last of the step, updates the frames (physics.cu:366-380).
void launch_tick(const GtsTissueView* v, const Lattice& L, const StepPlan& P, hipStream_t s, int n) {
    const int Eg = v->n_grid_edges, N = v->n_nodes;
    if (Eg <= 0) return;
    float4* records = reinterpret_cast<float4*>(v->scratch);
    float4* normals = records + Eg;
    k_save_cut_counter<<<1, 1, 0, s>>>(v->counters);
    k_cut_ccd<<<grid_for(Eg), GTS_BLOCK, 0, s>>>(Eg, L, P.tickS, P.tickE, P.tickN, v->voxel_length, v->grid_edges,
                                                 v->nbr, v->prev_corner, (const float4*)v->pos, v->occupied,
                                                 v->cut_mask, v->edge_cut_t, v->edge_skin_t, v->inside, records,
                                                 normals);
    emit_records(v, L, s, P.tickN, 1, P.tickD);
    if (P.tearLen2 > 0.f) {
        k_tear<<<grid_for(Eg), GTS_BLOCK, 0, s>>>(Eg, L, P.tearLen2, v->voxel_length, v->grid_edges, v->nbr, v->active,
                                                  (const float4*)v->pos, v->rest, v->occupied, v->cut_mask,
                                                  v->edge_cut_t, v->edge_skin_t, v->inside, records, normals,
                                                  v->counters, v->grasp_ring, v->edge_exempt);
        emit_records(v, L, s, P.tickN, 1, P.tickD);
    }
    sever_and_refresh(v, L, s);
    k_snapshot<<<grid_for(N), GTS_BLOCK, 0, s>>>(N, (const float4*)v->pos, v->prev_corner);
    if (n + 1 < P.Nsub) launch_frames(v, s, v->counters);
}

// The constants gts_step and gts_step_legacy share (xpbd_begin_frame, xpbd_solver.cu:2481-2514).
int make_plan(const GtsTissueView* v, const GtsStepParams* p, StepPlan& P) {
    if (v->n_nodes <= 0 || v->n_edge_colours < 0 || v->n_edge_colours > GTS_MAX_COLOURS || v->n_tet_colours < 0 ||
        v->n_tet_colours > GTS_MAX_COLOURS || !v->edge_colour_offsets || !v->tet_colour_offsets || !v->rot ||
        !v->rot_prev)
        return GTS_ERR_BAD_ARG;
    std::memset(&P, 0, sizeof(P));
    P.Nsub = p->substeps > 0 ? p->substeps : 1;
    if (!v->step_consts || v->step_const_capacity < P.Nsub) return GTS_ERR_CAPACITY;
    P.h = p->dt / (float)P.Nsub;
    if (!(P.h > 0.f) || !std::isfinite(P.h)) return GTS_ERR_BAD_DT;
    const float ks = p->edge_ks, h = P.h;
    P.at = ks > 0.f ? 1.f / (ks * h * h) : 0.f;
    P.velScale = (1.f - p->damping) / (1.f + p->rayleigh * h);
    P.betaGammaH = P.at * p->beta_s * h;
    P.invH2 = 1.f / (h * h);
    P.strain = v->n_tets > 0;
    P.fallback = P.strain && ks <= 0.f;
    P.edges = ks > 0.f || P.fallback;
    P.fb = P.fallback ? 1.f / (p->young * h * h) : 0.f;
    P.dOverH = fmaxf(0.f, p->strain_damping) / h;
    P.bOverH = h > 0.f ? p->bulk_damping / h : 0.f;
    P.iters = p->iterations > 0 ? p->iterations : 1;
    P.mode = p->tet_mode;
    P.precise = p->precise_strain != 0;
    P.omega = p->omega;
    P.beta_s = p->beta_s;
    P.perp = p->beta_s > 0.f;
    P.invH = 1.f / h;
    P.K = 1;
    // xpbd_set_tissue_guard / xpbd_set_tissue_material's argument checks (xpbd_solver.cu:1895, :2165-2167)
    if (p->guard_passes < 0 || p->guard_passes > 4) return GTS_ERR_BAD_ARG;
    if (p->guard_passes > 0) {
        if (!std::isfinite(p->guard_min_volume) || p->guard_min_volume < 0.f || p->guard_min_volume > .95f ||
            !std::isfinite(p->tensile_strength) || p->tensile_strength < 0.f || p->tensile_strength > 1000.f)
            return GTS_ERR_BAD_ARG;
        if (P.strain && (!v->fiber_lambda || !v->feasible_target || !v->feasible_needed || !v->incident_offsets ||
                         !v->incident_tets))
            return GTS_ERR_BAD_ARG;
        P.guardPasses = P.strain ? p->guard_passes : 0;
        P.strength = p->tensile_strength;
        P.guardMin = p->guard_min_volume;
    }
    // g_doBendF = B > 0 && kb > 0 (xpbd_solver.cu:2496)
    if (!(p->bend_kb >= 0.f) || !(p->beta_b >= 0.f)) return GTS_ERR_BAD_ARG;
    P.bend = p->bend_kb > 0.f && v->n_bends > 0;
    if (P.bend) {
        if (!v->bend_desc || !v->bend_lambda || !v->bend_colour_offsets || v->n_bend_colours < 0 ||
            v->n_bend_colours > GTS_MAX_COLOURS)
            return GTS_ERR_BAD_ARG;
        P.alphaB = 1.f / (p->bend_kb * h * h);
        P.betaGammaHB = P.alphaB * p->beta_b * h;
    }
    // the tiled sweeps: a plan for the current lists; tets undamped and unguarded, edges undamped without the fallback
    // compliance (the legacy step and the ribbon tick turn both off)
    P.tiled = P.tiledEdges = false;
    if (v->tile_groups > 0 && (v->tile_groups % 2) == 0 && v->tile_group_info && v->tile_entries && v->tile_writes &&
        v->pos_alt && v->tile_inputs && v->tile_lambda && v->tile_kmax >= 1 && v->tile_kmax <= GTS_TILE_KMAX &&
        v->tile_region >= 1 && v->tile_rx >= 1 && v->tile_ry >= 1 && v->tile_box >= 2 &&
        (long long)v->tile_box * v->tile_box * v->tile_box * 16 + 256 <= 65536) {
        int next[2] = {0, 0}, count[2] = {0, 0};
        const int marked = v->tile_group_info[3];
        bool ok = v->tile_dense_ids && marked >= 1 && marked <= GTS_TILE_MAX_REGIONS;
        for (int r = 0; r < marked && ok; r++) ok = v->tile_dense_ids[r] >= 0;
        for (int g = 0; g < v->tile_groups && ok; g++) {
            const int32_t* gi = v->tile_group_info + 16 * g;
            const int kind = gi[14];
            if (kind != 0 && kind != 1) {
                ok = false;
                break;
            }
            int rows = 0;
            for (int p = 0; p < gi[1] && p < GTS_TILE_KMAX; p++) rows += gi[10 + p];
            // an edge group's colours fit the box's halo: at most tile_kmax - 1 of them
            ok = gi[0] == next[kind] && gi[1] >= 1 && gi[1] <= v->tile_kmax - kind && gi[3] == marked &&
                 gi[4] == v->tile_box && gi[5] == v->tile_kmax - 1 && gi[6] == rows && gi[7] >= 0 && gi[8] >= 1 &&
                 gi[9] >= 0;
            for (int p = 0; p < gi[1] && p < GTS_TILE_KMAX && ok; p++) ok = gi[10 + p] >= 1;
            next[kind] += gi[1];
            count[kind]++;
        }
        ok = ok && count[0] % 2 == 0 && count[1] % 2 == 0;
        P.tiled = ok && count[0] > 0 && next[0] == v->n_tet_colours && v->tile_tets == v->n_tets && P.strain &&
                  P.dOverH == 0.f && P.bOverH == 0.f && P.guardPasses == 0;
        P.tiledEdges = ok && count[1] > 0 && next[1] == v->n_edge_colours && v->tile_edges == v->n_edges &&
                       P.edges && !P.fallback && P.betaGammaH == 0.f;
    }
    return GTS_OK;
}

// finishFeasible (xpbd_solver.cu:718-741) after substep n (the reference's substep index n + 1 picks the phase):
// flags cleared, candidate check, two colour sweeps of the repair, a second check, the uncertified tets' corners sent
// back to the substep start (prev), then eight lattice parity classes of the feasible filter.
void launch_safeguard(const GtsTissueView* v, const Lattice& L, const StepPlan& P, hipStream_t s, int n) {
    const int T = v->n_tets, N = v->n_nodes, ny = L.nxy / L.nx;
    int* needed = v->feasible_needed;
    int* flag = needed + N;
    k_fill_int<<<grid_for(N + 1), GTS_BLOCK, 0, s>>>(N + 1, needed, 0);
    k_guard_check<<<grid_for(T), GTS_BLOCK, 0, s>>>(T, L, P.guardMin, v->tet_desc, (const float4*)v->pos, flag);
    for (int pass = 0; pass < 2; pass++)
        for (int c = 0; c < v->n_tet_colours; c++) {
            int off = v->tet_colour_offsets[c], cnt = v->tet_colour_offsets[c + 1] - off;
            if (cnt > 0)
                k_guard_repair<<<grid_for(cnt), GTS_BLOCK, 0, s>>>(off, cnt, L, P.guardMin, v->tet_desc,
                                                                   (float4*)v->pos, flag);
        }
    k_fill_int<<<1, 64, 0, s>>>(1, flag, 0);
    k_guard_check<<<grid_for(T), GTS_BLOCK, 0, s>>>(T, L, P.guardMin, v->tet_desc, (const float4*)v->pos, flag);
    k_guard_need<<<grid_for(T), GTS_BLOCK, 0, s>>>(T, L, P.guardMin, v->tet_desc, (const float4*)v->prev,
                                                   (const float4*)v->pos, needed, flag);
    k_guard_prepare<<<grid_for(N), GTS_BLOCK, 0, s>>>(N, needed, (const float4*)v->prev,
                                                      (float4*)v->feasible_target, (float4*)v->pos);
    const int sites = ((L.nx + 1) / 2) * ((ny + 1) / 2) * ((v->vz + 2) / 2);
    const int phase = (n + 1) & 7;
    for (int k = 0; k < 8; k++)
        k_guard_vertex<<<grid_for(sites), GTS_BLOCK, 0, s>>>(N, ny, L, (k + phase) & 7, P.guardMin,
                                                             v->incident_offsets, v->incident_tets, v->tet_desc,
                                                             (const float4*)v->feasible_target, (float4*)v->pos,
                                                             needed);
}

// Every launch of one outer step, in the order of the reference's substep (xpbd_solver.cu:3460-3721).
void enqueue_step(const GtsTissueView* v, const Lattice& L, const StepPlan& P, hipStream_t s) {
    const int T = v->n_tets, N = v->n_nodes;
    launch_refresh(v, L, s);
    if (T > 0) k_tet_alpha<<<grid_for(T), GTS_BLOCK, 0, s>>>(T, P.invH2, v->tet_k_dev, v->tet_k_hyd, v->tet_alpha);
    // each plan entry's copy of its element's inputs, after this step's liveness and compliances
    for (int kind = 0; kind < 2; kind++) {
        if (!(kind ? P.tiledEdges : P.tiled)) continue;
        int first = -1, end = 0;  // a kind's entries are contiguous
        for (int g = 0; g < v->tile_groups; g++) {
            const int32_t* gi = v->tile_group_info + 16 * g;
            if (gi[14] != kind) continue;
            if (first < 0) first = gi[7];
            end = gi[7] + gi[3] * gi[6];
        }
        const int n = end - first;
        if (first < 0 || n <= 0) continue;
        k_tile_gather<<<grid_for(n), GTS_BLOCK, 0, s>>>(
            n, T, reinterpret_cast<const int2*>(v->tile_entries) + first, kind ? v->edge_desc : v->tet_desc,
            kind ? nullptr : v->tet_alpha, reinterpret_cast<float4*>(v->tile_inputs) + first);
    }
    if (P.contact || P.legacy) k_fill_float<<<1, 64, 0, s>>>(4, v->reaction, 0.f);
    const StepShapes* shapes = reinterpret_cast<const StepShapes*>(v->step_consts);
    dim3 g(grid_for(N)), b(GTS_BLOCK);
    const GraspFrame* frames = reinterpret_cast<const GraspFrame*>(v->step_consts);
    // the multipliers are reset once per substep, so the K passes of a substep count as K * iters iterations
    const int passes = P.K * P.iters;
    // Where nothing runs between a substep's node pass and the next substep's k_predict (no legacy grasp, guard,
    // perpendicular damping or ribbon tick), the node pass also predicts the next substep: one launch fewer per substep.
    const bool fuse = !P.legacy && !P.guardPasses && !P.perp && !P.tick;
    for (int n = 0; n < P.Nsub; n++) {
        if (!fuse || n == 0)
            k_predict<<<g, b, 0, s>>>(N, P.h, (float4*)v->pos, (float4*)v->prev, (float4*)v->vel,
                                      (const float4*)v->force, L.xcd_order);
        if (P.legacy && (P.gate || P.grasping)) k_grasp_reset<<<g, b, 0, s>>>(N, P.gate, P.grasping, P.grasp);
        if (P.guardPasses) k_fill_float<<<grid_for(6 * T), GTS_BLOCK, 0, s>>>(6 * T, v->fiber_lambda, 0.f);
        for (int k = 0; k < P.K; k++) {
            for (int i = 0; i < P.iters; i++) {
                const int it = k * P.iters + i;
                if (P.edges && P.tiledEdges)
                    launch_tiled_sweep<0, false, true>(v, L, s, it, passes, P.omega, P.at);
                else if (P.edges)
                    for (int c = 0; c < v->n_edge_colours; c++) {
                        int off = v->edge_colour_offsets[c], cnt = v->edge_colour_offsets[c + 1] - off;
                        if (cnt > 0) launch_edge_colour(v, L, s, off, cnt, it, passes, P.fallback, P.at, P.fb, P.omega,
                                                        P.betaGammaH);
                    }
                if (P.bend)  // then the bending angle constraints (xpbd_solver.cu:3590-3596)
                    for (int c = 0; c < v->n_bend_colours; c++) {
                        int off = v->bend_colour_offsets[c], cnt = v->bend_colour_offsets[c + 1] - off;
                        if (cnt > 0) launch_bend_colour(v, L, s, off, cnt, it, passes, P.alphaB, P.omega,
                                                        P.betaGammaHB);
                    }
                if (P.strain && P.tiled) {
                    if (P.mode == 0 && P.precise) launch_tiled_sweep<0, true, false>(v, L, s, it, passes, P.omega, 0.f);
                    else if (P.mode == 0) launch_tiled_sweep<0, false, false>(v, L, s, it, passes, P.omega, 0.f);
                    else if (P.mode == 1 && P.precise)
                        launch_tiled_sweep<1, true, false>(v, L, s, it, passes, P.omega, 0.f);
                    else if (P.mode == 1) launch_tiled_sweep<1, false, false>(v, L, s, it, passes, P.omega, 0.f);
                    else if (P.precise) launch_tiled_sweep<2, true, false>(v, L, s, it, passes, P.omega, 0.f);
                    else launch_tiled_sweep<2, false, false>(v, L, s, it, passes, P.omega, 0.f);
                } else if (P.strain)
                    for (int c = 0; c < v->n_tet_colours; c++) {
                        int off = v->tet_colour_offsets[c], cnt = v->tet_colour_offsets[c + 1] - off;
                        if (cnt <= 0) continue;
                        if (P.mode == 0)
                            launch_tet_colour<0>(v, L, s, off, cnt, it, passes, P.omega, P.dOverH, P.bOverH, P.precise);
                        else if (P.mode == 1)
                            launch_tet_colour<1>(v, L, s, off, cnt, it, passes, P.omega, P.dOverH, P.bOverH, P.precise);
                        else
                            launch_tet_colour<2>(v, L, s, off, cnt, it, passes, P.omega, P.dOverH, P.bOverH, P.precise);
                    }
                // k_TensileReinforcement after the tet sweep, inside solveTetGraph (xpbd_solver.cu:2531-2533)
                for (int pass = 0; pass < P.guardPasses; pass++)
                    for (int c = 0; c < v->n_tet_colours; c++) {
                        int off = v->tet_colour_offsets[c], cnt = v->tet_colour_offsets[c + 1] - off;
                        if (cnt > 0)
                            k_tensile<<<grid_for(cnt), GTS_BLOCK, 0, s>>>(off, cnt, L, P.strength, v->tet_desc,
                                                                          v->tet_alpha, v->fiber_lambda,
                                                                          (float4*)v->pos);
                    }
            }
            if (P.legacy && P.grasping) k_grasp_pass<<<g, b, 0, s>>>(N, frames + n, P.h, P.grasp, (float4*)v->pos);
        }
#define GTS_LEGACY(SC, VL)                                                                                       \
    k_legacy_finish<SC, VL><<<g, b, 0, s>>>(N, frames + n, P.grasping, P.gate, P.h, P.invH, P.velScale, P.grasp, \
                                            (float4*)v->pos, (const float4*)v->prev, (float4*)v->vel)
        if (P.legacy) {
            if (P.guardPasses) GTS_LEGACY(false, false);
            else if (P.perp) GTS_LEGACY(false, true);
            else GTS_LEGACY(true, true);
        }
#undef GTS_LEGACY
        const StepShapes* sh = shapes + n;
#define GTS_FINISH(C, SC, VL, PR)                                                                                \
    k_finish_nodes<C, SC, VL, PR><<<g, b, 0, s>>>(N, sh, P.armed, P.alphaC, P.mu_s, P.mu_d, P.radius, P.scaleR,   \
                                                  P.invH, P.velScale, (float4*)v->pos, (float4*)v->prev,         \
                                                  (float4*)v->vel, v->reaction, L.xcd_order, P.h,                \
                                                  (const float4*)v->force)
        const bool predictNext = fuse && n + 1 < P.Nsub;
        if (P.legacy) {}
        else if (P.guardPasses) { if (P.contact) GTS_FINISH(true, false, false, false); }
        else if (P.contact) {
            if (P.perp) GTS_FINISH(true, false, true, false);
            else if (predictNext) GTS_FINISH(true, true, true, true);
            else GTS_FINISH(true, true, true, false);
        } else {
            if (P.perp) GTS_FINISH(false, false, true, false);
            else if (predictNext) GTS_FINISH(false, true, true, true);
            else GTS_FINISH(false, true, true, false);
        }
#undef GTS_FINISH
        if (P.guardPasses) {  // the safeguard between the node passes and the velocity update (:3678)
            launch_safeguard(v, L, P, s, n);
            if (P.perp) k_velocity<false><<<g, b, 0, s>>>(N, P.invH, P.velScale, (const float4*)v->pos,
                                                          (const float4*)v->prev, (float4*)v->vel);
            else k_velocity<true><<<g, b, 0, s>>>(N, P.invH, P.velScale, (const float4*)v->pos,
                                                  (const float4*)v->prev, (float4*)v->vel);
        }
        if (P.perp) {
            for (int c = 0; c < v->n_edge_colours; c++) {
                int off = v->edge_colour_offsets[c], cnt = v->edge_colour_offsets[c + 1] - off;
                if (cnt > 0)
                    k_perp_damp<<<grid_for(cnt), GTS_BLOCK, 0, s>>>(off, cnt, L, P.beta_s, P.h, v->edge_desc,
                                                                    (const float4*)v->pos, (float4*)v->vel);
            }
            k_velscale<<<g, b, 0, s>>>(N, P.velScale, (const float4*)v->pos, (float4*)v->vel);
        }
        if (P.tick) launch_tick(v, L, P, s, n);
    }
    launch_frames(v, s);  // LCS_Step ends with physics_compute_rotations (plugin_api.cpp:133)
}

// Captured steps, one per (view, plan) pair: replaying the graph issues the same launches with the same arguments;
// everything that changes between steps (tool shapes, positions, liveness) lives in device memory.
// Pinned host staging for the per-step constants: a ring of buffers, each reused only after the copy that last read
// it has completed (its event), so the host never overwrites data still in flight.
struct Staging {
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
};
Staging g_staging[8];
int g_staging_next = 0, g_staging_last = -1;

int staging_slot(size_t bytes, void** out) {
    Staging& st = g_staging[g_staging_next];
    if (st.pending) {
        HIP_TRY(hipEventSynchronize(st.done));
        st.pending = false;
    }
    if (st.cap < bytes) {
        if (st.host) HIP_TRY(hipHostFree(st.host));
        st.host = nullptr;
        HIP_TRY(hipHostMalloc(&st.host, bytes, hipHostMallocDefault));
        st.cap = bytes;
    }
    if (!st.done) HIP_TRY(hipEventCreateWithFlags(&st.done, hipEventDisableTiming));
    g_staging_last = g_staging_next;
    g_staging_next = (g_staging_next + 1) % 8;
    *out = st.host;
    return GTS_OK;
}

int staging_done(hipStream_t s) {
    Staging& st = g_staging[g_staging_last];
    HIP_TRY(hipEventRecord(st.done, s));
    st.pending = true;
    return GTS_OK;
}

struct GraphEntry {
    std::vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;
};
std::vector<GraphEntry> g_graphs;
hipStream_t g_capture = nullptr;

// The key holds the view, the plan and the colour offsets the launches were sized from (the view only points at them,
// and a host array freed and reallocated at the same address must not replay a graph built for other contents).
std::vector<unsigned char> graph_key(const GtsTissueView* v, const StepPlan& P) {
    const size_t ne = (size_t)v->n_edge_colours + 1, nt = (size_t)v->n_tet_colours + 1;
    const size_t nb = P.bend ? (size_t)v->n_bend_colours + 1 : 0;
    const bool tiled = P.tiled || P.tiledEdges;
    const size_t ng = tiled ? (size_t)16 * v->tile_groups : 0;
    const size_t nd = tiled ? (size_t)v->tile_group_info[3] : 0;  // the dense ids the launches carry
    std::vector<unsigned char> k(sizeof(GtsTissueView) + sizeof(StepPlan) +
                                 (ne + nt + nb + ng + nd) * sizeof(int32_t));
    unsigned char* d = k.data();
    std::memcpy(d, v, sizeof(GtsTissueView));
    d += sizeof(GtsTissueView);
    std::memcpy(d, &P, sizeof(StepPlan));
    d += sizeof(StepPlan);
    std::memcpy(d, v->edge_colour_offsets, ne * sizeof(int32_t));
    d += ne * sizeof(int32_t);
    std::memcpy(d, v->tet_colour_offsets, nt * sizeof(int32_t));
    d += nt * sizeof(int32_t);
    if (nb) std::memcpy(d, v->bend_colour_offsets, nb * sizeof(int32_t));
    d += nb * sizeof(int32_t);
    if (ng) std::memcpy(d, v->tile_group_info, ng * sizeof(int32_t));
    d += ng * sizeof(int32_t);
    if (nd) std::memcpy(d, v->tile_dense_ids, nd * sizeof(int32_t));
    return k;
}

// Views seen once and not captured yet.
std::vector<std::vector<unsigned char>> g_seen;

// A (view, plan) pair is captured on its second use: the first step of it runs as direct launches. Capturing pays
// only for views that recur; a tissue whose lists are compacted after every step presents a new view every step.
int launch_step_graph(const GtsTissueView* v, const Lattice& L, const StepPlan& P, hipStream_t s) {
    std::vector<unsigned char> key = graph_key(v, P);
    for (auto& e : g_graphs)
        if (e.key == key) return -(int)hipGraphLaunch(e.exec, s);
    auto seen = std::find(g_seen.begin(), g_seen.end(), key);
    if (seen == g_seen.end()) {
        if (g_seen.size() >= 16) g_seen.erase(g_seen.begin());
        g_seen.push_back(key);
        enqueue_step(v, L, P, s);
        return -(int)hipGetLastError();
    }
    g_seen.erase(seen);
    if (!g_capture) HIP_TRY(hipStreamCreateWithFlags(&g_capture, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(g_capture, hipStreamCaptureModeThreadLocal));
    enqueue_step(v, L, P, g_capture);
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamEndCapture(g_capture, &graph));
    GraphEntry e;
    e.key = key;
    hipError_t err = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(err);
    if (g_graphs.size() >= 16) {
        (void)hipGraphExecDestroy(g_graphs.front().exec);
        g_graphs.erase(g_graphs.begin());
    }
    g_graphs.push_back(e);
    return -(int)hipGraphLaunch(e.exec, s);
}

}  // namespace

namespace {

// The CCD ribbon tick after every substep while a cut tool is set (cut_ribbon_tick, cut.cu:1303-1431).
int plan_tick(const GtsTissueView* v, const GtsCutTool* cut, const float* emit_ncut, float tear_tau, StepPlan& P) {
    if (!emit_ncut || !v->prev_corner) return GTS_ERR_BAD_ARG;
    P.tick = true;
    P.tiled = P.tiledEdges = false;  // the tick changes liveness within the step, after the step's gather of the tiled
                                     // inputs
    P.tickS = ld3(cut->t2v2);  // the current rod: S = t2v2, E = t1v2 (cut.cu:1382-1383)
    P.tickE = ld3(cut->t1v2);
    P.tickN = ld3(emit_ncut);
    P.tickD = cut->d;
    const float tl = tear_tau * v->voxel_length;
    P.tearLen2 = tear_tau >= 1.f ? tl * tl : 0.f;  // tau < 1 or NaN: the tear law is off (cut.cu:1199-1202)
    return GTS_OK;
}

// gts_step, with the ribbon tick of a set cut tool when `cut` is given.
int step_impl(const GtsTissueView* v, const GtsStepParams* p, const GtsToolState* from, const GtsToolState* to,
              int tool_armed, int blade_armed, const GtsCutTool* cut, const float* emit_ncut, float tear_tau,
              void* stream) {
    if (!v || !p || !from || !to) return GTS_ERR_BAD_ARG;
    if (p->tet_mode < 0 || p->tet_mode > 2) return GTS_ERR_BAD_ARG;
    if (v->colour_block < 0 || v->colour_block > 1024 || v->colour_block % 32) return GTS_ERR_BAD_ARG;
    hipStream_t s = S(stream);
    Lattice L = lattice_of(v);
    StepPlan P;
    { int rc_ = make_plan(v, p, P); if (rc_) return rc_; }
    const float h = P.h;
    if (cut) { int rc_ = plan_tick(v, cut, emit_ncut, tear_tau, P); if (rc_) return rc_; }
    P.contact = p->contact_enabled && (tool_armed || blade_armed);
    P.armed = (tool_armed ? 7u : 0u) | (blade_armed ? 8u : 0u);
    P.alphaC = p->contact_compliance / (h * h);
    P.scaleR = 1.f / (h * h * fmaxf(1.f, (float)P.Nsub));
    P.mu_s = p->mu_s;
    P.mu_d = p->mu_d;
    P.radius = p->contact_radius;
    // tool pose of each substep and of the one before it (xpbd_solver.cu:3467-3483), uploaded for the kernels
    if (P.contact) {
        StepShapes* host = nullptr;
        { int rc_ = staging_slot(sizeof(StepShapes) * P.Nsub, reinterpret_cast<void**>(&host)); if (rc_) return rc_; }
        const ToolPose pf = pose_of(from), pt = pose_of(to);
        for (int n = 0; n < P.Nsub; n++) {
            float t = fminf(1.f, (float)(n + 1) / (float)P.Nsub);
            float previousT = fmaxf(0.f, t - 1.f / (float)P.Nsub);
            ToolPose prevTool = lerp_tool(pf, pt, previousT), stepTool = lerp_tool(pf, pt, t);
            for (int kind = 0; kind < 4; kind++) {
                host[n].from[kind] = tool_shape(prevTool, kind, p->jaw_thickness, p->jaw_gap_closed);
                host[n].to[kind] = tool_shape(stepTool, kind, p->jaw_thickness, p->jaw_gap_closed);
            }
        }
        HIP_TRY(hipMemcpyAsync(v->step_consts, host, sizeof(StepShapes) * P.Nsub, hipMemcpyHostToDevice, s));
        { int rc_ = staging_done(s); if (rc_) return rc_; }
    }
    if (v->use_graph && !g_profile.on) return launch_step_graph(v, L, P, s);
    enqueue_step(v, L, P, s);
    return -(int)hipGetLastError();
}

}  // namespace

extern "C" int gts_step(const GtsTissueView* v, const GtsStepParams* p, const GtsToolState* from,
                        const GtsToolState* to, int tool_armed, int blade_armed, void* stream) {
    return step_impl(v, p, from, to, tool_armed, blade_armed, nullptr, nullptr, 0.f, stream);
}

extern "C" int gts_step_cutting(const GtsTissueView* v, const GtsStepParams* p, const GtsToolState* from,
                                const GtsToolState* to, int tool_armed, int blade_armed, const GtsCutTool* tool,
                                const float* emit_ncut, float tear_tau, void* stream) {
    if (!tool) return GTS_ERR_BAD_ARG;
    return step_impl(v, p, from, to, tool_armed, blade_armed, tool, emit_ncut, tear_tau, stream);
}

extern "C" int gts_snapshot_corners(const GtsTissueView* v, void* stream) {
    if (!v || !v->prev_corner || v->n_nodes <= 0) return GTS_ERR_BAD_ARG;
    k_snapshot<<<grid_for(v->n_nodes), GTS_BLOCK, 0, S(stream)>>>(v->n_nodes, (const float4*)v->pos, v->prev_corner);
    return -(int)hipGetLastError();
}

namespace {

// The jaw frame xpbd_substep derives from a substep's pose (xpbd_solver.cu:3572-3580).
GraspFrame grasp_frame(const ToolPose& t, float closedGap) {
    GraspFrame f;
    f.axis = len3(t.jawAxis) > 1e-9f ? t.jawAxis * (1.f / len3(t.jawAxis)) : mk3(0.f, 0.f, 1.f);
    f3 n = len3(t.jawNormal) > 1e-9f ? t.jawNormal * (1.f / len3(t.jawNormal)) : mk3(0.f, 1.f, 0.f);
    n = n - f.axis * dot3(n, f.axis);
    if (len3(n) < 1e-7f) n = cross3(f.axis, fabsf(f.axis.x) < 0.9f ? mk3(1.f, 0.f, 0.f) : mk3(0.f, 1.f, 0.f));
    f.normal = n * (1.f / len3(n));
    f.w = cross3(f.axis, f.normal);
    f.gapHalf = 0.5f * fmaxf(t.jawGap, closedGap);
    f.hinge = t.jawHinge;
    f.shaftA = t.shaftA;
    f.shaftB = t.shaftB;
    f.shaftR = t.shaftR;
    f.jawLen = t.jawLen;
    f.jawHalfWidth = t.jawHalfWidth;
    return f;
}

bool grasp_state_ok(const GtsGraspState* g) {
    return g && g->grasped && g->core && g->core_offset && g->anchor && g->slab_lambda && g->core_lambda &&
           g->core_force_ema && g->gate_lambda && g->capture_scratch;
}

}  // namespace

extern "C" int gts_grasp_set_tool(const GtsTissueView* v, GtsGraspState* g, const GtsGraspParams* gp,
                                  const GtsToolState* tool, int* event, void* stream) {
    if (!v || !gp || !tool || !event || !grasp_state_ok(g)) return GTS_ERR_BAD_ARG;
    *event = 0;
    hipStream_t s = S(stream);
    const int N = v->n_nodes;
    // the jaw frame of the stored pose (xpbd_solver.cu:2326-2336)
    f3 axis = ld3(tool->jawAxis), normal = ld3(tool->jawNormal);
    float alen = len3(axis);
    axis = alen > 1e-9f ? axis * (1.f / alen) : mk3(0.f, 0.f, 1.f);
    float nlen = len3(normal);
    normal = nlen > 1e-9f ? normal * (1.f / nlen) : mk3(0.f, 1.f, 0.f);
    f3 w = cross3(axis, normal);
    float wlen = len3(w);
    w = wlen > 1e-9f ? w * (1.f / wlen) : mk3(1.f, 0.f, 0.f);
    const float gapHalf = 0.5f * fmaxf(tool->jawGap, gp->closed_gap);
    const bool nowAbove = tool->grasp01 > gp->capture, nowBelow = tool->grasp01 < gp->release;
    if (!g->grasping && nowAbove) {
        k_mark_candidates<<<grid_for(N), GTS_BLOCK, 0, s>>>(N, ld3(tool->jawHinge), axis, normal, w, tool->jawLen,
                                                            tool->jawHalfWidth, gapHalf + gp->skin, v->active,
                                                            v->pinned, (const float4*)v->pos, g->capture_scratch);
        HIP_TRY(hipGetLastError());
        std::vector<float> scratch(4 * (size_t)N);
        HIP_TRY(hipMemcpyAsync(scratch.data(), g->capture_scratch, scratch.size() * sizeof(float),
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        const int* cand = reinterpret_cast<const int*>(scratch.data());
        const float* uvn = scratch.data() + N;
        std::vector<int> idxs;
        for (int c = 0; c < N; c++)
            if (cand[c]) idxs.push_back(c);
        std::vector<int32_t> grasped(N, 0), core(N, 0);
        std::vector<float> off(3 * (size_t)N, 0.f), anchor(2 * (size_t)N, 0.f);
        if (!idxs.empty()) {
            // normalised-ellipse core over the captured set's own u / v extents (xpbd_solver.cu:2405-2431)
            float uMin = 1e30f, uMax = -1e30f, vMin = 1e30f, vMax = -1e30f;
            for (int c : idxs) {
                uMin = fminf(uMin, uvn[3 * c]); uMax = fmaxf(uMax, uvn[3 * c]);
                vMin = fminf(vMin, uvn[3 * c + 1]); vMax = fmaxf(vMax, uvn[3 * c + 1]);
            }
            float uMid = 0.5f * (uMin + uMax), uHalf = fmaxf(0.5f * (uMax - uMin), 1e-6f);
            float vMid = 0.5f * (vMin + vMax), vHalf = fmaxf(0.5f * (vMax - vMin), 1e-6f);
            float frac = gp->core_fraction;
            int bestIdx = -1;
            float bestD2 = 1e30f;
            bool anyCore = false;
            for (int c : idxs) {
                grasped[c] = 1;
                float nu = (uvn[3 * c] - uMid) / uHalf, nv = (uvn[3 * c + 1] - vMid) / vHalf;
                float d2 = nu * nu + nv * nv;
                if (d2 < frac * frac) { core[c] = 1; anyCore = true; }
                if (d2 < bestD2) { bestD2 = d2; bestIdx = c; }
            }
            if (!anyCore && bestIdx >= 0) core[bestIdx] = 1;  // a grasp is never core-less
            for (int c : idxs) {
                off[3 * c] = uvn[3 * c];          // u
                off[3 * c + 1] = uvn[3 * c + 2];  // n
                off[3 * c + 2] = uvn[3 * c + 1];  // w
                anchor[2 * c] = uvn[3 * c];
                anchor[2 * c + 1] = uvn[3 * c + 1];
            }
        }
        HIP_TRY(hipMemcpyAsync(g->grasped, grasped.data(), N * sizeof(int32_t), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(g->core, core.data(), N * sizeof(int32_t), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(g->core_offset, off.data(), off.size() * sizeof(float), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(g->anchor, anchor.data(), anchor.size() * sizeof(float), hipMemcpyHostToDevice, s));
        k_fill_float<<<grid_for(2 * N), GTS_BLOCK, 0, s>>>(2 * N, g->slab_lambda, 0.f);  // slab and its previous
        k_fill_float<<<grid_for(N), GTS_BLOCK, 0, s>>>(N, g->core_force_ema, 0.f);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));  // the host arrays above are released on return
        g->grasping = 1;
        *event = 1;
    } else if (g->grasping && nowBelow) {
        k_release<<<grid_for(N), GTS_BLOCK, 0, s>>>(N, g->grasped, g->core, (float4*)v->vel);
        HIP_TRY(hipGetLastError());
        g->grasping = 0;
        *event = 2;
    }
    // physics_set_tool then recomputes the tear law's ring from the current captured set (physics.cu:436-447)
    if (v->grasp_ring) {
        if (!v->nbr_init) return GTS_ERR_BAD_ARG;
        int* a = v->grasp_ring;
        int* b = a + N;
        k_copy_ints<<<grid_for(N), GTS_BLOCK, 0, s>>>(N, g->grasped, a);
        for (int hop = 0; hop < gp->exempt_rings; hop++) {
            k_ring_hop<<<grid_for(N), GTS_BLOCK, 0, s>>>(N, v->nbr, a, b);
            std::swap(a, b);
        }
        if (a != v->grasp_ring) k_copy_ints<<<grid_for(N), GTS_BLOCK, 0, s>>>(N, a, v->grasp_ring);
        k_ring_carve<<<grid_for(N), GTS_BLOCK, 0, s>>>(N, v->nbr, v->nbr_init, v->grasp_ring);
        HIP_TRY(hipGetLastError());
    }
    return GTS_OK;
}

namespace {

int legacy_impl(const GtsTissueView* v, const GtsStepParams* p, const GtsGraspParams* gp, const GtsGraspState* g,
                const GtsToolState* from, const GtsToolState* to, int tool_armed, const GtsCutTool* cut,
                const float* emit_ncut, float tear_tau, void* stream) {
    if (!v || !p || !gp || !from || !to || !grasp_state_ok(g)) return GTS_ERR_BAD_ARG;
    if (p->tet_mode < 0 || p->tet_mode > 2) return GTS_ERR_BAD_ARG;
    if (v->colour_block < 0 || v->colour_block > 1024 || v->colour_block % 32) return GTS_ERR_BAD_ARG;
    hipStream_t s = S(stream);
    Lattice L = lattice_of(v);
    StepPlan P;
    { int rc_ = make_plan(v, p, P); if (rc_) return rc_; }
    if ((size_t)P.Nsub * sizeof(GraspFrame) > (size_t)v->step_const_capacity * GTS_STEP_CONST_FLOATS * sizeof(float))
        return GTS_ERR_CAPACITY;
    const float h = P.h;
    const int N = v->n_nodes;
    if (cut) { int rc_ = plan_tick(v, cut, emit_ncut, tear_tau, P); if (rc_) return rc_; }
    P.legacy = true;
    P.tiled = P.tiledEdges = false;  // the coupling passes of the legacy grasp keep the per-colour launches
    P.grasping = g->grasping != 0;
    P.gate = tool_armed != 0;
    P.K = P.grasping ? 2 : 1;  // a second material pass only while genuinely grasping (xpbd_solver.cu:3571)
    GraspArgs& a = P.grasp;
    a.grasped = g->grasped;
    a.core = g->core;
    a.off = g->core_offset;
    a.anchor = g->anchor;
    a.slab = g->slab_lambda;
    a.slabPrev = g->slab_lambda + N;
    a.coreLam = reinterpret_cast<float4*>(g->core_lambda);
    a.ema = g->core_force_ema;
    a.jawLam = g->gate_lambda;
    a.shaftLam = g->gate_lambda + N;
    a.omega = gp->grasp_omega;
    a.contactAlpha = gp->contact_compliance / (h * h);
    a.mu = gp->friction;
    a.maxDisp = gp->grasp_displacement;
    a.coreAlpha = gp->core_compliance / (h * h);
    a.breakForce = gp->break_force;
    a.kEMA = -expm1f(-h);  // one-second time constant, independent of the substep count (xpbd_solver.cu:992)
    a.ccdOmega = gp->ccd_omega;
    a.ccdVel = gp->ccd_velocity;
    a.ccdMaxDisp = gp->ccd_displacement;
    // the jaw frame of every substep's interpolated pose (xpbd_solver.cu:3467-3483, :3572-3580)
    if (P.grasping || P.gate) {
        GraspFrame* host = nullptr;
        { int rc_ = staging_slot(sizeof(GraspFrame) * P.Nsub, reinterpret_cast<void**>(&host)); if (rc_) return rc_; }
        const ToolPose pf = pose_of(from), pt = pose_of(to);
        for (int n = 0; n < P.Nsub; n++) {
            float t = fminf(1.f, (float)(n + 1) / (float)P.Nsub);
            host[n] = grasp_frame(lerp_tool(pf, pt, t), gp->closed_gap);
        }
        HIP_TRY(hipMemcpyAsync(v->step_consts, host, sizeof(GraspFrame) * P.Nsub, hipMemcpyHostToDevice, s));
        { int rc_ = staging_done(s); if (rc_) return rc_; }
    }
    if (v->use_graph && !g_profile.on) return launch_step_graph(v, L, P, s);
    enqueue_step(v, L, P, s);
    return -(int)hipGetLastError();
}

}  // namespace

extern "C" int gts_step_legacy(const GtsTissueView* v, const GtsStepParams* p, const GtsGraspParams* gp,
                               const GtsGraspState* g, const GtsToolState* from, const GtsToolState* to, int tool_armed,
                               void* stream) {
    return legacy_impl(v, p, gp, g, from, to, tool_armed, nullptr, nullptr, 0.f, stream);
}

extern "C" int gts_step_legacy_cutting(const GtsTissueView* v, const GtsStepParams* p, const GtsGraspParams* gp,
                                       const GtsGraspState* g, const GtsToolState* from, const GtsToolState* to,
                                       int tool_armed, const GtsCutTool* tool, const float* emit_ncut, float tear_tau,
                                       void* stream) {
    if (!tool) return GTS_ERR_BAD_ARG;
    return legacy_impl(v, p, gp, g, from, to, tool_armed, tool, emit_ncut, tear_tau, stream);
}

extern "C" int gts_detect_cut(const GtsTissueView* v, const GtsCutTool* tool, void* stream) {
    if (!v || !tool) return GTS_ERR_BAD_ARG;
    if (v->n_grid_edges <= 0) return GTS_OK;
    hipStream_t s = S(stream);
    Lattice L = lattice_of(v);
    CutToolD t;
    t.t1v0 = ld3(tool->t1v0); t.t1v1 = ld3(tool->t1v1); t.t1v2 = ld3(tool->t1v2);
    t.t2v0 = ld3(tool->t2v0); t.t2v1 = ld3(tool->t2v1); t.t2v2 = ld3(tool->t2v2);
    t.ncut = ld3(tool->ncut); t.lo = ld3(tool->aabbMin); t.hi = ld3(tool->aabbMax);
    t.d = tool->d; t.valid = tool->valid;
    int Eg = v->n_grid_edges;
    float4* records = reinterpret_cast<float4*>(v->scratch);
    k_cut_quad<<<grid_for(Eg), GTS_BLOCK, 0, s>>>(Eg, L, t, v->voxel_length, v->grid_edges, (const float4*)v->pos,
                                                  v->occupied, v->cut_mask, v->edge_cut_t, v->edge_skin_t, v->inside,
                                                  records);
    HIP_TRY(hipGetLastError());
    return finish_cut(v, L, s, t.ncut, 0, t.d);
}

// ------------------------------------------------------------------------------------------------ allocation-free cut
// gts_detect_cut with its workspace (the per-edge detection records and the emission scan's partials) in a file-scope
// device array instead of the view's scratch buffer, for callers that cannot allocate device memory. The array is
// GTS_STATIC_SCRATCH_FLOATS floats (32 MiB) on every device the library runs on, held until the process ends; each
// device's copy sits at its own address, which hipGetSymbolAddress gives on the entry point's first call on that
// device. Host code never passes the array's name to a kernel, only that address. Every entry the cut reads is first
// written by the same call, so nothing is cleared between calls; calls that overlap on one device race on the array.
static __device__ __align__(256) float scratch[GTS_STATIC_SCRATCH_FLOATS];

namespace {
std::mutex g_scratch_mutex;
float* g_scratch_ptr[64] = {};  // per device ordinal

int static_scratch(float** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return GTS_ERR_BAD_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return GTS_ERR_NOT_READY;
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return GTS_OK;
}
}  // namespace

extern "C" long long gts_detect_cut_static_floats(int n_grid_edges) {
    return n_grid_edges <= 0 ? 0 : 8ll * n_grid_edges + scan_blocks(n_grid_edges) + 1;
}

extern "C" int gts_detect_cut_static(const GtsTissueView* v, const GtsCutTool* tool, void* stream) {
    if (!v || !tool) return GTS_ERR_BAD_ARG;
    if (v->n_grid_edges <= 0) return GTS_OK;
    if (gts_detect_cut_static_floats(v->n_grid_edges) > (long long)GTS_STATIC_SCRATCH_FLOATS) return GTS_ERR_CAPACITY;
    float* ptr = nullptr;
    { int rc_ = static_scratch(&ptr); if (rc_) return rc_; }
    GtsTissueView w = *v;  // a host copy of the descriptor: the kernels' workspace is the device array
    w.scratch = ptr;
    return gts_detect_cut(&w, tool, stream);
}

extern "C" int gts_detect_cut_ccd(const GtsTissueView* v, const float* Sb, const float* Eb, const float* emit_ncut,
                                  float tool_d, const float* prev_positions, void* stream) {
    if (!v || !Sb || !Eb || !emit_ncut || !prev_positions) return GTS_ERR_BAD_ARG;
    if (v->n_grid_edges <= 0) return GTS_OK;
    hipStream_t s = S(stream);
    Lattice L = lattice_of(v);
    int Eg = v->n_grid_edges;
    float4* records = reinterpret_cast<float4*>(v->scratch);
    float4* normals = records + Eg;
    k_cut_ccd<<<grid_for(Eg), GTS_BLOCK, 0, s>>>(Eg, L, ld3(Sb), ld3(Eb), ld3(emit_ncut), v->voxel_length,
                                                 v->grid_edges, v->nbr, prev_positions, (const float4*)v->pos,
                                                 v->occupied, v->cut_mask, v->edge_cut_t, v->edge_skin_t, v->inside,
                                                 records, normals);
    HIP_TRY(hipGetLastError());
    return finish_cut(v, L, s, ld3(emit_ncut), 1, tool_d);
}

extern "C" int gts_tear_overstretch(const GtsTissueView* v, float tau, float tool_d, void* stream) {
    if (!v) return GTS_ERR_BAD_ARG;
    if (!(tau >= 1.f) || v->n_grid_edges <= 0) return GTS_OK;  // tau < 1 or NaN: the law is off (cut.cu:1199-1202)
    if (!v->rest) return GTS_ERR_BAD_ARG;
    hipStream_t s = S(stream);
    Lattice L = lattice_of(v);
    int Eg = v->n_grid_edges;
    float tl = tau * v->voxel_length;
    float4* records = reinterpret_cast<float4*>(v->scratch);
    float4* normals = records + Eg;
    k_tear<<<grid_for(Eg), GTS_BLOCK, 0, s>>>(Eg, L, tl * tl, v->voxel_length, v->grid_edges, v->nbr, v->active,
                                              (const float4*)v->pos, v->rest, v->occupied, v->cut_mask,
                                              v->edge_cut_t, v->edge_skin_t, v->inside, records, normals,
                                              v->counters, v->grasp_ring, v->edge_exempt);
    HIP_TRY(hipGetLastError());
    return finish_cut(v, L, s, mk3(0.f, 0.f, 0.f), 1, tool_d);
}

extern "C" int gts_compact_topology(const GtsTissueView* v, int32_t* tet_remap, int32_t* edge_remap,
                                    int32_t* node_remap, int32_t* counts, int32_t* node_used, void* stream) {
    if (!v || !tet_remap || !edge_remap || !node_remap || !counts || !node_used) return GTS_ERR_BAD_ARG;
    hipStream_t s = S(stream);
    Lattice L = lattice_of(v);
    { int rc_ = launch_refresh(v, L, s); if (rc_) return rc_; }
    int T = v->n_tets, E = v->n_edges, N = v->n_nodes;
    int* partial = partial_area(v);
    int nbT = scan_blocks(T), nbE = scan_blocks(E), nbN = scan_blocks(N);
    int* pT = partial;
    int* pE = pT + nbT + 1;
    int* pN = pE + nbE + 1;
    k_fill_int<<<grid_for(N), GTS_BLOCK, 0, s>>>(N, node_used, 0);
    k_mark_used<<<grid_for(std::max(T, E)), GTS_BLOCK, 0, s>>>(T, E, L, v->tet_desc, v->edge_desc, node_used);
    k_count_flags<<<nbT, GTS_BLOCK, 0, s>>>(T, FlagDescLive{v->tet_desc}, pT);
    k_scan_partials<<<1, GTS_BLOCK, 0, s>>>(nbT, pT);
    k_write_remap<<<nbT, GTS_BLOCK, 0, s>>>(T, FlagDescLive{v->tet_desc}, pT, tet_remap);
    k_count_flags<<<nbE, GTS_BLOCK, 0, s>>>(E, FlagDescLive{v->edge_desc}, pE);
    k_scan_partials<<<1, GTS_BLOCK, 0, s>>>(nbE, pE);
    k_write_remap<<<nbE, GTS_BLOCK, 0, s>>>(E, FlagDescLive{v->edge_desc}, pE, edge_remap);
    k_count_flags<<<nbN, GTS_BLOCK, 0, s>>>(N, FlagNodeUsed{node_used, v->active}, pN);
    k_scan_partials<<<1, GTS_BLOCK, 0, s>>>(nbN, pN);
    k_write_remap<<<nbN, GTS_BLOCK, 0, s>>>(N, FlagNodeUsed{node_used, v->active}, pN, node_remap);
    k_copy_counts<<<1, 1, 0, s>>>(pT, nbT, pE, nbE, pN, nbN, counts);
    return -(int)hipGetLastError();
}

namespace {
// out[remap[k]] = in[k] for every entry the remap keeps (stable: remap is an exclusive scan of the live flags).
template <class V>
__global__ void k_gather_live(int n, const int* __restrict__ remap, const V* __restrict__ in, V* __restrict__ out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        int r = remap[k];
        if (r >= 0) out[r] = in[k];
    }
}
}  // namespace

// Moves the live tets and edges of the view into the output arrays at the positions gts_compact_topology assigned.
// The outputs hold counts[0] tets and counts[1] edges; they must not alias the inputs.
extern "C" int gts_apply_compaction(const GtsTissueView* v, const int32_t* tet_remap, const int32_t* edge_remap,
                                    uint32_t* tet_desc_out, float* tet_k_dev_out, float* tet_k_hyd_out,
                                    uint32_t* edge_desc_out, void* stream) {
    if (!v || !tet_remap || !edge_remap || !tet_desc_out || !tet_k_dev_out || !tet_k_hyd_out || !edge_desc_out)
        return GTS_ERR_BAD_ARG;
    hipStream_t s = S(stream);
    const int T = v->n_tets, E = v->n_edges;
    if (T > 0) {
        k_gather_live<<<grid_for(T), GTS_BLOCK, 0, s>>>(T, tet_remap, v->tet_desc, tet_desc_out);
        k_gather_live<<<grid_for(T), GTS_BLOCK, 0, s>>>(T, tet_remap, v->tet_k_dev, tet_k_dev_out);
        k_gather_live<<<grid_for(T), GTS_BLOCK, 0, s>>>(T, tet_remap, v->tet_k_hyd, tet_k_hyd_out);
    }
    if (E > 0) k_gather_live<<<grid_for(E), GTS_BLOCK, 0, s>>>(E, edge_remap, v->edge_desc, edge_desc_out);
    return -(int)hipGetLastError();
}

extern "C" float gts_iso_pow(float J) { return iso_pow(J); }
extern "C" float gts_log(float x) { return own_log(x); }
extern "C" float gts_acos(float c) { return own_acos(c); }

namespace {
// Streaming copy of n float4: the known-size reference for HBM byte counters and the achievable copy bandwidth.
__global__ void __launch_bounds__(256) k_stream_copy(long long n, const float4* __restrict__ src,
                                                     float4* __restrict__ dst) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256LL) dst[i] = src[i];
}
}  // namespace

extern "C" int gts_stream_copy(void* dst, const void* src, long long n_float4, void* stream) {
    if (!dst || !src || n_float4 <= 0) return GTS_ERR_BAD_ARG;
    long long blocks = (n_float4 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    k_stream_copy<<<(unsigned)blocks, 256, 0, reinterpret_cast<hipStream_t>(stream)>>>(
        n_float4, reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst));
    return -(int)hipGetLastError();
}

extern "C" int gts_profile_begin(void) {
    g_profile.on = true;
    g_profile.used = 0;
    g_profile.error = hipSuccess;
    return GTS_OK;
}

// Stops recording and returns the summed duration (ms), the launch count and the summed units (tets) of the recorded
// tet colour launches. Synchronizes on the recorded events.
extern "C" int gts_profile_end(double* total_ms, long long* launches, long long* units) {
    g_profile.on = false;
    if (g_profile.error != hipSuccess) {
        hipError_t e = g_profile.error;
        g_profile.error = hipSuccess;
        g_profile.used = 0;
        return -(int)e;
    }
    double total = 0.0;
    long long u = 0;
    for (size_t k = 0; k < g_profile.used; k++) {
        HIP_TRY(hipEventSynchronize(g_profile.events[2 * k + 1]));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, g_profile.events[2 * k], g_profile.events[2 * k + 1]));
        total += ms;
        u += g_profile.units[k];
    }
    if (total_ms) *total_ms = total;
    if (launches) *launches = (long long)g_profile.used;
    if (units) *units = u;
    g_profile.used = 0;
    return GTS_OK;
}
