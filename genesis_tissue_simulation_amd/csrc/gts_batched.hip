// Batched XPBD stepping of many independent tissue instances of uneven size in one persistent launch.
//
// The reference runs one session per process (native_backend.py:213-218); here a whole population of lattice
// instances advances together. Each workgroup takes a list of instances (longest-processing-time packing on the
// host) and runs every substep of a step for one instance at a time:
//   * the instance's node positions (float4: x, y, z, inverse mass) are staged in LDS when they fit, so all
//     Gauss-Seidel colour passes of all substeps gather and scatter on-chip; a larger instance stages as many of its
//     nodes as fit and keeps the rest in its own slice of global memory, which only this workgroup touches;
//   * colours are separated by workgroup barriers instead of kernel launches: the colour order is the reference's
//     (xpbd/xpbd_solver.cu:3581-3612), so each instance computes exactly what the single-instance path computes;
//   * HBM traffic per substep is the packed edge / tet descriptors and per-tet stiffness (8 B per tet, 4 B per edge);
//     positions, velocities and prev stay on chip or in the instance's L2-resident slice.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <queue>
#include <vector>

#include "../../include/gts_xpbd.h"
#include "gts_common.h"
#include "gts_solver.h"
#include "gts_frames.h"

#define HIP_TRY(call)                                   \
    do {                                                \
        hipError_t _e = (call);                         \
        if (_e != hipSuccess) return -(int)_e;          \
    } while (0)

namespace {

// Per-tet inputs of a projection, loaded ahead of use.
struct TetInputs {
    uint32_t d;
    float ad, ah, lamD, lamH;
};

struct BatchStep {
    int substeps, iters, mode, n_items, precise;
    int sub_begin, sub_end;  // the substeps this launch runs
    float h, invH, invH2, omega, velScale, gx, gy, gz, dOverH, bOverH;
    int lds_nodes;
};

template <int MODE, bool DAMPED, bool PRECISE, class Store>
__device__ void run_instance(const Store& S, const GtsBatchInstance& I, const BatchStep& P, const int* colours,
                             const uint32_t* __restrict__ edge_desc, const uint32_t* __restrict__ tet_desc,
                             const float* __restrict__ k_dev, const float* __restrict__ k_hyd,
                             float* __restrict__ lam_edge, float* __restrict__ lam_tet, float* __restrict__ alpha,
                             const float* __restrict__ mass, float4* __restrict__ prev, float4* __restrict__ vel) {
    const int tid = threadIdx.x, bs = blockDim.x;
    const int N = I.n_nodes, T = I.n_tets;
    const int nx = I.nx, nxy = I.nxy;
    const int* ecol = colours;
    const int* tcol = colours + I.n_edge_colours + 1;
    const float at = I.edge_ks > 0.f ? 1.f / (I.edge_ks * P.h * P.h) : 0.f;
    const float betaGammaH = 0.f;
    // k_ComputeTetAlpha (xpbd_solver.cu:578-587) once per step, as the single-instance path does per frame
    if (P.sub_begin == 0) {
        for (int k = tid; k < T; k += bs) {
            alpha[k] = P.invH2 / k_dev[k];
            alpha[T + k] = P.invH2 / k_hyd[k];
        }
        __syncthreads();
    }
    for (int n = P.sub_begin; n < P.sub_end; n++) {
        // k_predict (xpbd_solver.cu:393-412) with the gravity force g * mass of k_SetGravity
        for (int i = tid; i < N; i += bs) {
            float4 p = S.ld(i);
            prev[i] = p;
            if (p.w > 0.f) {
                float m = mass[i];
                f3 F = mk3(P.gx * m, P.gy * m, P.gz * m);
                f3 a = F * p.w;
                float4 v0 = vel[i];
                f3 v = mk3(v0.x, v0.y, v0.z) + a * P.h;
                vel[i] = make_float4(v.x, v.y, v.z, 0.f);
                f3 x = xyz4(p) + v * P.h;
                S.st(i, make_float4(x.x, x.y, x.z, p.w));
            }
        }
        __syncthreads();
        for (int it = 0; it < P.iters; it++) {
            const bool read = it > 0, write = P.iters > 1;
            if (at > 0.f)
                for (int c = 0; c < I.n_edge_colours; c++) {
                    for (int k = ecol[c] + tid; k < ecol[c + 1]; k += bs) {
                        uint32_t d = edge_desc[k];
                        if (!(d & 1u)) {  // a dead row keeps the substep's reset multiplier, 0
                            if (write && !read) lam_edge[k] = 0.f;
                            continue;
                        }
                        int axis = (int)((d >> 1) & 3u);
                        int i = (int)(d >> 4), j = i + (axis == 0 ? 1 : (axis == 1 ? nx : nxy));
                        float lam = read ? lam_edge[k] : 0.f, out;
                        const bool ok =
                            edge_project<DAMPED>(S, prev, i, j, I.rest_len[axis], at, P.omega, betaGammaH, lam, out);
                        if (write && (ok || !read)) lam_edge[k] = ok ? out : 0.f;  // 0: the substep's reset
                    }
                    __syncthreads();
                }
            // The per-tet inputs of this thread's first tet of the next colour are loaded while the current colour
            // computes, hiding their latency behind it: no projection of one colour writes another colour's
            // descriptor, alpha or multipliers.
            auto fetch = [&](int k, int end) {
                TetInputs t;
                t.d = 0u;
                t.ad = t.ah = t.lamD = t.lamH = 0.f;
                if (k < end) {
                    t.d = tet_desc[k];
                    t.ad = alpha[k];
                    t.ah = alpha[T + k];
                    if (read) {
                        t.lamD = lam_tet[k];
                        t.lamH = lam_tet[T + k];
                    }
                }
                return t;
            };
            TetInputs next = fetch(tcol[0] + tid, tcol[1]);
            for (int c = 0; c < I.n_tet_colours; c++) {
                const int k0 = tcol[c] + tid, end = tcol[c + 1];
                const TetInputs first = next;
                if (c + 1 < I.n_tet_colours) next = fetch(tcol[c + 1] + tid, tcol[c + 2]);
                for (int k = k0; k < end; k += bs) {
                    const TetInputs t = k == k0 ? first : fetch(k, end);
                    if (!(t.d & 1u)) {  // a dead tet keeps the substep's reset multipliers, 0
                        if (write && !read) lam_tet[k] = lam_tet[T + k] = 0.f;
                        continue;
                    }
                    int base = (int)(t.d >> 4), type = (int)((t.d >> 1) & 7u);
                    int id[4];
                    for (int q = 0; q < 4; q++) id[q] = base + corner_offset(kuhn_corner(type, q), nx, nxy);
                    f3 rk[3];
                    if (PRECISE)
                        for (int q = 0; q < 3; q++) rk[q] = kuhn_rest_edge(type, q + 1, I.rest_len);
                    float lamD = t.lamD, lamH = t.lamH;
                    const bool ok = tet_project<MODE, DAMPED, PRECISE>(S, prev, id, I.inv_rest[type], t.ad, t.ah,
                                                                       lamD, lamH, P.omega, P.dOverH, P.bOverH, rk);
                    if (write && (ok || !read)) {  // 0: the substep's reset
                        lam_tet[k] = ok ? lamD : 0.f;
                        lam_tet[T + k] = ok ? lamH : 0.f;
                    }
                }
                __syncthreads();
            }
        }
        // k_update_velocity + k_apply_velscale (xpbd_solver.cu:871-919)
        for (int i = tid; i < N; i += bs) {
            float4 p = S.ld(i);
            f3 v = p.w > 0.f ? (xyz4(p) - xyz4(prev[i])) * P.invH : mk3(0, 0, 0);
            v = p.w > 0.f ? v * P.velScale : mk3(0, 0, 0);
            vel[i] = make_float4(v.x, v.y, v.z, 0.f);
        }
        __syncthreads();
    }
}

template <int MODE, bool DAMPED, bool PRECISE>
__global__ void __launch_bounds__(1024) k_batched_step(BatchStep P, const GtsBatchInstance* __restrict__ instances,
                                                       const int* __restrict__ colour_table,
                                                       const int* __restrict__ work_offsets,
                                                       const int* __restrict__ work_items, float4* __restrict__ pos,
                                                       float4* __restrict__ prev, float4* __restrict__ vel,
                                                       const float* __restrict__ mass,
                                                       const uint32_t* __restrict__ edge_desc,
                                                       const uint32_t* __restrict__ tet_desc,
                                                       const float* __restrict__ k_dev, const float* __restrict__ k_hyd,
                                                       float* __restrict__ lam_edge, float* __restrict__ lam_tet,
                                                       float* __restrict__ alpha) {
    extern __shared__ float4 lds_nodes[];
    const int wg = blockIdx.x;
    for (int w = work_offsets[wg]; w < work_offsets[wg + 1]; w++) {
        // read the record in place: a private copy would index its Dm^-1 table out of scratch memory
        const GtsBatchInstance& I = instances[work_items[w]];
        float4* gpos = pos + I.node_off;
        float4* gprev = prev + I.node_off;
        float4* gvel = vel + I.node_off;
        const float* gmass = mass + I.node_off;
        const int* colours = colour_table + I.colour_off;
        const uint32_t* ed = edge_desc + I.edge_off;
        const uint32_t* td = tet_desc + I.tet_off;
        const float* kd = k_dev + I.tet_off;
        const float* kh = k_hyd + I.tet_off;
        float* le = lam_edge + I.edge_off;
        float* lt = lam_tet + 2 * (size_t)I.tet_off;
        float* al = alpha + 2 * (size_t)I.tet_off;
        const int staged = min(I.n_nodes, P.lds_nodes);
        for (int i = threadIdx.x; i < staged; i += blockDim.x) lds_nodes[i] = gpos[i];
        __syncthreads();
        if (I.n_nodes <= P.lds_nodes)
            run_instance<MODE, DAMPED, PRECISE>(LdsNodes{lds_nodes}, I, P, colours, ed, td, kd, kh, le, lt, al, gmass,
                                                gprev, gvel);
        else if (staged > 0)
            run_instance<MODE, DAMPED, PRECISE>(SplitNodes{lds_nodes, gpos, staged}, I, P, colours, ed, td, kd, kh, le,
                                                lt, al, gmass, gprev, gvel);
        else
            run_instance<MODE, DAMPED, PRECISE>(GlobalNodes{gpos}, I, P, colours, ed, td, kd, kh, le, lt, al, gmass,
                                                gprev, gvel);
        for (int i = threadIdx.x; i < staged; i += blockDim.x) gpos[i] = lds_nodes[i];
        __syncthreads();
    }
}

struct BatchGraph {
    std::vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;
};
std::vector<BatchGraph> g_batch_graphs;
hipStream_t g_batch_capture = nullptr;

template <int MODE, bool DAMPED, bool PRECISE>
int launch(const GtsBatchView* v, const BatchStep& P, hipStream_t s) {
    size_t lds = (size_t)v->lds_nodes * sizeof(float4);
    auto kern = k_batched_step<MODE, DAMPED, PRECISE>;
    hipError_t e =
        hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return -(int)e;
    kern<<<v->n_workgroups, v->block_size, lds, s>>>(
        P, reinterpret_cast<const GtsBatchInstance*>(v->instances), v->colour_table, v->work_offsets, v->work_items,
        (float4*)v->pos, (float4*)v->prev, (float4*)v->vel, v->mass, v->edge_desc, v->tet_desc, v->k_dev, v->k_hyd,
        v->lambda_edge, v->lambda_tet, v->alpha);
    return -(int)hipGetLastError();
}

template <bool PRECISE>
int launch_mode(const GtsBatchView* v, const BatchStep& P, bool damped, hipStream_t s) {
    if (damped) {
        if (P.mode == 0) return launch<0, true, PRECISE>(v, P, s);
        if (P.mode == 1) return launch<1, true, PRECISE>(v, P, s);
        return launch<2, true, PRECISE>(v, P, s);
    }
    if (P.mode == 0) return launch<0, false, PRECISE>(v, P, s);
    if (P.mode == 1) return launch<1, false, PRECISE>(v, P, s);
    return launch<2, false, PRECISE>(v, P, s);
}

int launch_range(const GtsBatchView* v, const BatchStep& P, bool damped, hipStream_t s) {
    return P.precise ? launch_mode<true>(v, P, damped, s) : launch_mode<false>(v, P, damped, s);
}

// The step as persistent launches of substeps_per_launch substeps each (one launch when 0).
// dst[0, n) = src[0, n); with both pointers 16-byte aligned, 16 bytes a thread and the n % 4 tail after them.
__global__ void k_batch_copy(long long n, const float* __restrict__ src, float* __restrict__ dst, int vec) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!vec) {
        if (i < n) dst[i] = src[i];
        return;
    }
    const long long n4 = n >> 2;
    if (i < n4) reinterpret_cast<float4*>(dst)[i] = reinterpret_cast<const float4*>(src)[i];
    else if (i < n4 + (n & 3)) dst[4 * n4 + (i - n4)] = src[4 * n4 + (i - n4)];
}

// k_ComputeParticleRot (physics.cu:109-205) over the pooled nodes of every instance.
__global__ void k_batched_frames(long long n, const float4* __restrict__ pos, const int* __restrict__ nbr,
                                 const int* __restrict__ active, const int* __restrict__ pinned,
                                 const int* __restrict__ base, const float* __restrict__ prev,
                                 float* __restrict__ rot) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) particle_frame((int)g, base[g], pos, nbr, active, pinned, prev, rot);
}

int enqueue(const GtsBatchView* v, BatchStep P, bool damped, hipStream_t s) {
    const int per = v->substeps_per_launch > 0 ? v->substeps_per_launch : P.substeps;
    for (int b = 0; b < P.substeps; b += per) {
        P.sub_begin = b;
        P.sub_end = b + per < P.substeps ? b + per : P.substeps;
        int rc = launch_range(v, P, damped, s);
        if (rc) return rc;
    }
    if (v->rot) {  // the frame update that ends LCS_Step (plugin_api.cpp:133), for every instance at once
        const long long n = v->total_nodes;
        const int vec = ((reinterpret_cast<uintptr_t>(v->rot) | reinterpret_cast<uintptr_t>(v->rot_prev)) & 15) == 0;
        const long long threads = vec ? (9 * n >> 2) + 3 : 9 * n;
        k_batch_copy<<<(unsigned)((threads + 255) / 256), 256, 0, s>>>(9 * n, v->rot, v->rot_prev, vec);
        k_batched_frames<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, (const float4*)v->pos, v->nbr, v->active,
                                                                     v->pinned, v->node_base, v->rot_prev, v->rot);
        HIP_TRY(hipGetLastError());
    }
    return GTS_OK;
}

}  // namespace

extern "C" int gts_batched_lds_bytes(void) {
    int dev = 0, bytes = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
    return bytes;
}

// Longest-processing-time packing: instances by descending cost, each to the currently least-loaded workgroup.
// offsets [n_workgroups + 1] and items [n] are host arrays.
extern "C" int gts_batched_plan(const double* cost, int n, int n_workgroups, int32_t* offsets_items) {
    if (!cost || n < 0 || n_workgroups <= 0 || !offsets_items) return GTS_ERR_BAD_ARG;
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    typedef std::pair<double, int> Load;
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    for (int w = 0; w < n_workgroups; w++) heap.push(Load(0.0, w));
    std::vector<std::vector<int>> lists(n_workgroups);
    for (int i : order) {
        Load l = heap.top();
        heap.pop();
        lists[l.second].push_back(i);
        heap.push(Load(l.first + cost[i], l.second));
    }
    int32_t* offsets = offsets_items;
    int32_t* items = offsets_items + n_workgroups + 1;
    int k = 0;
    for (int w = 0; w < n_workgroups; w++) {
        offsets[w] = k;
        for (int i : lists[w]) items[k++] = i;
    }
    offsets[n_workgroups] = k;
    return GTS_OK;
}

extern "C" int gts_batched_step(const GtsBatchView* v, const GtsStepParams* p, void* stream) {
    if (!v || !p || v->n_workgroups <= 0 || v->block_size <= 0 || v->block_size > 1024) return GTS_ERR_BAD_ARG;
    if (p->tet_mode < 0 || p->tet_mode > 2) return GTS_ERR_BAD_ARG;
    // the tissue guard and bending run on the single-tissue step only
    if (p->guard_passes != 0 || p->bend_kb != 0.f) return GTS_ERR_BAD_ARG;
    if ((long long)v->lds_nodes * 16 > (long long)gts_batched_lds_bytes()) return GTS_ERR_BAD_ARG;
    if (v->substeps_per_launch < 0) return GTS_ERR_BAD_ARG;
    if (v->rot && (!v->rot_prev || !v->nbr || !v->active || !v->pinned || !v->node_base || v->total_nodes <= 0 ||
                   v->total_nodes > (long long)INT32_MAX))
        return GTS_ERR_BAD_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    BatchStep P;
    std::memset(&P, 0, sizeof(P));
    P.substeps = p->substeps > 0 ? p->substeps : 1;
    P.h = p->dt / (float)P.substeps;
    if (!(P.h > 0.f) || !std::isfinite(P.h)) return GTS_ERR_BAD_DT;
    P.iters = p->iterations > 0 ? p->iterations : 1;
    P.mode = p->tet_mode;
    P.precise = p->precise_strain != 0;
    P.invH = 1.f / P.h;
    P.invH2 = 1.f / (P.h * P.h);
    P.omega = p->omega;
    P.velScale = (1.f - p->damping) / (1.f + p->rayleigh * P.h);
    P.gx = v->gravity[0];
    P.gy = v->gravity[1];
    P.gz = v->gravity[2];
    P.dOverH = fmaxf(0.f, p->strain_damping) / P.h;
    P.bOverH = p->bulk_damping / P.h;
    P.lds_nodes = v->lds_nodes;
    P.n_items = v->n_instances;
    const bool damped = P.dOverH != 0.f || P.bOverH != 0.f;
    if (!v->use_graph) return enqueue(v, P, damped, s);
    std::vector<unsigned char> key(sizeof(GtsBatchView) + sizeof(BatchStep));
    std::memcpy(key.data(), v, sizeof(GtsBatchView));
    std::memcpy(key.data() + sizeof(GtsBatchView), &P, sizeof(BatchStep));
    for (auto& g : g_batch_graphs)
        if (g.key == key) return -(int)hipGraphLaunch(g.exec, s);
    if (!g_batch_capture) HIP_TRY(hipStreamCreateWithFlags(&g_batch_capture, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(g_batch_capture, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(v, P, damped, g_batch_capture);
    hipGraph_t graph = nullptr;
    hipError_t end = hipStreamEndCapture(g_batch_capture, &graph);
    if (rc || end != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc ? rc : -(int)end;
    }
    BatchGraph g;
    g.key = key;
    hipError_t err = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(err);
    if (g_batch_graphs.size() >= 16) {
        (void)hipGraphExecDestroy(g_batch_graphs.front().exec);
        g_batch_graphs.erase(g_batch_graphs.begin());
    }
    g_batch_graphs.push_back(g);
    return -(int)hipGraphLaunch(g.exec, s);
}
