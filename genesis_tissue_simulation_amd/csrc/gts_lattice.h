// Lattice-edge conventions of the cut and surface kernels (cut.cuh), and the ordered stable compaction (wave ballot
// plus LDS prefix inside a workgroup, exclusive scan of the workgroup totals) that both use.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef GTS_BLOCK
#define GTS_BLOCK 256
#endif

namespace {

// EdgeStencil (cut.cuh:126-137): the four voxels around a lattice edge and the edge's local id in each.
__device__ __forceinline__ void stencil(int axis, int slot, int& dx, int& dy, int& dz, int& local) {
    const int8_t off[3][4][3] = {{{0, 0, 0}, {0, -1, 0}, {0, -1, -1}, {0, 0, -1}},
                                 {{0, 0, 0}, {0, 0, -1}, {-1, 0, -1}, {-1, 0, 0}},
                                 {{0, 0, 0}, {-1, 0, 0}, {-1, -1, 0}, {0, -1, 0}}};
    const int8_t loc[3][4] = {{0, 2, 6, 4}, {3, 7, 5, 1}, {8, 9, 10, 11}};
    dx = off[axis][slot][0]; dy = off[axis][slot][1]; dz = off[axis][slot][2];
    local = loc[axis][slot];
}

__device__ __forceinline__ int directed_slot(int a, int b, int vx) {  // DirectedEdgeSlot (cut.cuh:25-28)
    int dl = b - a, ad = dl < 0 ? -dl : dl;
    int axis = ad == 1 ? 0 : (ad == vx + 1 ? 1 : 2);
    return 6 * a + 2 * axis + (dl > 0 ? 1 : 0);
}

__device__ __forceinline__ bool cut_hits_material(float t, float skinT, int insideA, int insideB) {
    if (insideA && insideB) return true;
    if (!insideA && !insideB) return false;
    if (!(skinT >= 0.f && skinT <= 1.f)) return false;
    return insideA ? t <= skinT : t >= skinT;
}

// SkinOwnerA (cut.cuh:16-18): whether endpoint A owns the skin crossing of a lattice edge once a cut at cutT has split
// it (the original inside endpoint otherwise).
__device__ __forceinline__ int skin_owner_a(float skinT, float cutT, int insideA) {
    return cutT >= 0.f && cutT <= 1.f && fabsf(skinT - cutT) > 1e-6f ? (skinT < cutT ? 1 : 0) : insideA;
}

// ---- ordered stable compaction: wave ballot + LDS prefix inside a block, exclusive scan of the block totals

__device__ __forceinline__ int block_prefix(bool flag, int* lds, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint64_t m = __ballot(flag);
    int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) lds[wave] = __popcll(m);
    __syncthreads();
    int waveBase = 0;
    total = 0;
    for (int k = 0; k < nw; k++) {
        int c = lds[k];
        if (k < wave) waveBase += c;
        total += c;
    }
    __syncthreads();
    return waveBase + before;
}

// Exclusive prefix of `value` over the workgroup (wave scan with lane shuffles, then the wave totals in LDS); total
// receives the workgroup's sum. Every thread of the workgroup must call it.
__device__ __forceinline__ int block_scan(int value, int* lds, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int x = value;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    int waveBase = 0;
    total = 0;
    for (int k = 0; k < nw; k++) {
        int c = lds[k];
        if (k < wave) waveBase += c;
        total += c;
    }
    __syncthreads();
    return waveBase + x - value;
}

template <class Flag>
__global__ void k_count_flags(int n, Flag flag, int* partial) {
    __shared__ int lds[GTS_BLOCK / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int total;
    block_prefix(i < n && flag(i), lds, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Exclusive scan of the per-block totals in one workgroup (up to 1024 threads); partial[nblocks] receives the grand
// total.
__global__ void __launch_bounds__(1024) k_scan_partials(int nblocks, int* partial) {
    __shared__ int lds[16];
    int carry = 0;
    for (int start = 0; start < nblocks; start += blockDim.x) {
        int i = start + threadIdx.x;
        int v = i < nblocks ? partial[i] : 0;
        // inclusive scan in the wave
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        int x = v;
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) lds[wave] = x;
        __syncthreads();
        int waveBase = 0, total = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
            if (k < wave) waveBase += lds[k];
            total += lds[k];
        }
        if (i < nblocks) partial[i] = carry + waveBase + x - v;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[nblocks] = carry;
}

template <class Flag>
__global__ void k_write_remap(int n, Flag flag, const int* partial, int* remap) {
    __shared__ int lds[GTS_BLOCK / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool f = i < n && flag(i);
    int total;
    int r = block_prefix(f, lds, total);
    if (i < n) remap[i] = f ? partial[blockIdx.x] + r : -1;
}

}  // namespace
