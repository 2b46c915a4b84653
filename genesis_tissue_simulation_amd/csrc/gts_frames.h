// Particle frames of the cut points' material coordinates (k_ComputeParticleRot, physics.cu:109-205), shared by the
// single-tissue and the batched step. Frames are the three rows of R per node (RotGpu, cut.cuh:67), 9 floats.
#pragma once
#include <hip/hip_runtime.h>

#include "gts_common.h"

__device__ __forceinline__ f3 frame_xyz(float4 p) { return mk3(p.x, p.y, p.z); }

struct Rows3 {
    f3 r0, r1, r2;
};

__device__ __forceinline__ float det3(const Rows3& m) {
    return m.r0.x * (m.r1.y * m.r2.z - m.r1.z * m.r2.y) - m.r0.y * (m.r1.x * m.r2.z - m.r1.z * m.r2.x) +
           m.r0.z * (m.r1.x * m.r2.y - m.r1.y * m.r2.x);
}

// transpose(inverse(m)) by the adjugate (physics.cu:37-64)
__device__ __forceinline__ Rows3 inverse_transpose3(const Rows3& m) {
    float invDet = 1.f / det3(m);
    f3 a0 = mk3((m.r1.y * m.r2.z - m.r1.z * m.r2.y), -(m.r0.y * m.r2.z - m.r0.z * m.r2.y),
                (m.r0.y * m.r1.z - m.r0.z * m.r1.y));
    f3 a1 = mk3(-(m.r1.x * m.r2.z - m.r1.z * m.r2.x), (m.r0.x * m.r2.z - m.r0.z * m.r2.x),
                -(m.r0.x * m.r1.z - m.r0.z * m.r1.x));
    f3 a2 = mk3((m.r1.x * m.r2.y - m.r1.y * m.r2.x), -(m.r0.x * m.r2.y - m.r0.y * m.r2.x),
                (m.r0.x * m.r1.y - m.r0.y * m.r1.x));
    a0 = a0 * invDet;
    a1 = a1 * invDet;
    a2 = a2 * invDet;
    Rows3 t;
    t.r0 = mk3(a0.x, a1.x, a2.x);
    t.r1 = mk3(a0.y, a1.y, a2.y);
    t.r2 = mk3(a0.z, a1.z, a2.z);
    return t;
}

__device__ __forceinline__ f3 div3(f3 a, float d) { return mk3(a.x / d, a.y / d, a.z / d); }

// k_ComputeParticleRot (physics.cu:109-205): per axis the through-particle edge direction (one-sided where a neighbour
// is gone) blended with the mean previous axis of the live neighbours; a single lost axis rebuilt as the cross product
// of the other two; R = polar(M) by 16 Higham steps R <- (R + R^-T) / 2. Pinned or inactive nodes, and degenerate,
// reflected or unconverged frames, get the identity. Arrays are indexed by node c; link targets are relative to base.
__device__ inline void particle_frame(int c, int base, const float4* __restrict__ pos, const int* __restrict__ nbr,
                                      const int* __restrict__ active, const int* __restrict__ pinned,
                                      const float* __restrict__ prev, float* __restrict__ rot) {
    float* out = rot + 9 * (size_t)c;
    Rows3 R;
    R.r0 = mk3(1.f, 0.f, 0.f);
    R.r1 = mk3(0.f, 1.f, 0.f);
    R.r2 = mk3(0.f, 0.f, 1.f);
    const int pin = pinned[c], act = active[c];
    bool ok = pin == 0 && act != 0;
    f3 axis[3];
    bool have[3] = {false, false, false};
    if (ok) {
        // Every gather of the node is issued up front and unconditionally (a missing link reads node c itself), so
        // the loads are in flight together instead of one dependent round trip per neighbour; the masks then apply
        // the reference's per-axis choices, and the previous axes of the live neighbours are summed in link order.
        int safe[6];
        bool live[6];
        float4 np4[6];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int n = nbr[6 * c + q];
            safe[q] = n >= 0 ? base + n : c;
            live[q] = n >= 0;
        }
        int act_q[6];
        for (int q = 0; q < 6; q++) act_q[q] = active[safe[q]];
        for (int q = 0; q < 6; q++) np4[q] = pos[safe[q]];
        float r[6][9];
        for (int q = 0; q < 6; q++)
            for (int j = 0; j < 9; j++) r[q][j] = prev[9 * (size_t)safe[q] + j];
        for (int q = 0; q < 6; q++) live[q] = live[q] & (act_q[q] != 0);
        // the three sums are separate variables, not an array: selects over an array element were lowered to
        // scratch memory accesses
        f3 ps0 = mk3(0.f, 0.f, 0.f), ps1 = ps0, ps2 = ps0;
        int pc = 0;
#pragma unroll
        for (int q = 0; q < 6; q++) {
            if (live[q]) {
                ps0 = ps0 + mk3(r[q][0], r[q][3], r[q][6]);
                ps1 = ps1 + mk3(r[q][1], r[q][4], r[q][7]);
                ps2 = ps2 + mk3(r[q][2], r[q][5], r[q][8]);
                pc++;
            }
        }
        const f3 cPos = frame_xyz(pos[c]);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const bool okp = live[2 * k + 1], okm = live[2 * k];
            const f3 xp = frame_xyz(np4[2 * k + 1]), xm = frame_xyz(np4[2 * k]);
            f3 e;
            if (okp && okm) e = xp - xm;
            else if (okp) e = xp - cPos;
            else if (okm) e = cPos - xm;
            else continue;
            float len = len3(e);
            if (len < 1e-12f) continue;
            f3 dir = div3(e, len);
            if (pc > 0) {
                const f3 psum = k == 0 ? ps0 : (k == 1 ? ps1 : ps2);
                float pl = len3(psum);
                if (pl > 1e-6f) {
                    f3 blended = dir + div3(psum, pl);
                    float dl = len3(blended);
                    if (dl > 0.25f) dir = div3(blended, dl);
                }
            }
            axis[k] = dir;
            have[k] = true;
        }
        // a single lost axis is the cross product of the next two, cyclically; each case is spelled out so that the
        // axes stay in registers (an axis array indexed at run time would live in scratch memory)
        const int mc = (have[0] ? 0 : 1) + (have[1] ? 0 : 1) + (have[2] ? 0 : 1);
        if (mc == 1) {
            const f3 cx = !have[0] ? cross3(axis[1], axis[2]) : (!have[1] ? cross3(axis[2], axis[0])
                                                                          : cross3(axis[0], axis[1]));
            const float l = len3(cx);
            if (l > 1e-9f) {
                const f3 u = div3(cx, l);
                if (!have[0]) axis[0] = u;
                else if (!have[1]) axis[1] = u;
                else axis[2] = u;
                have[0] = have[1] = have[2] = true;
            }
        }
        ok = have[0] && have[1] && have[2];
    }
    if (ok) {
        Rows3 M;
        M.r0 = mk3(axis[0].x, axis[1].x, axis[2].x);
        M.r1 = mk3(axis[0].y, axis[1].y, axis[2].y);
        M.r2 = mk3(axis[0].z, axis[1].z, axis[2].z);
        ok = !(det3(M) <= 1e-4f);  // the reference's rejection test, so that a NaN frame is kept as it does
        if (ok) {
            Rows3 H = M;
            for (int it = 0; it < 16; it++) {
                Rows3 t = inverse_transpose3(H);
                H.r0 = (H.r0 + t.r0) * 0.5f;
                H.r1 = (H.r1 + t.r1) * 0.5f;
                H.r2 = (H.r2 + t.r2) * 0.5f;
            }
            f3 c0 = mk3(H.r0.x, H.r1.x, H.r2.x), c1 = mk3(H.r0.y, H.r1.y, H.r2.y), c2 = mk3(H.r0.z, H.r1.z, H.r2.z);
            float ortho = fmaxf(fmaxf(fabsf(dot3(c0, c0) - 1.f), fabsf(dot3(c1, c1) - 1.f)), fabsf(dot3(c2, c2) - 1.f));
            ortho = fmaxf(ortho, fmaxf(fabsf(dot3(c0, c1)), fmaxf(fabsf(dot3(c0, c2)), fabsf(dot3(c1, c2)))));
            if (!(ortho > 1e-3f)) R = H;
        }
    }
    out[0] = R.r0.x; out[1] = R.r0.y; out[2] = R.r0.z;
    out[3] = R.r1.x; out[4] = R.r1.y; out[5] = R.r1.z;
    out[6] = R.r2.x; out[7] = R.r2.y; out[8] = R.r2.z;
}

