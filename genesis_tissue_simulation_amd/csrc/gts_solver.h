// Constraint projections of the XPBD substep as device functions, shared by the per-colour kernels of the single
// instance and the persistent batched and adjoint kernels, so that every path rounds identically.
//
// Node storage is abstracted by a store type with ld(i) / st(i, v) on float4 (x, y, z, inverse mass): LDS-staged for
// instances that fit in a workgroup's LDS, split between LDS and global memory for larger ones, or global memory.
#pragma once
#include "gts_common.h"

struct GlobalNodes {
    float4* p;
    __device__ __forceinline__ float4 ld(int i) const { return p[i]; }
    __device__ __forceinline__ void st(int i, float4 v) const { p[i] = v; }
};
struct LdsNodes {
    float4* p;  // points into __shared__ memory
    __device__ __forceinline__ float4 ld(int i) const { return p[i]; }
    __device__ __forceinline__ void st(int i, float4 v) const { p[i] = v; }
};
// An instance larger than the LDS budget: its first `cap` nodes are staged in LDS, the rest stay in its global slice.
// Nodes are numbered along the lattice (x fastest, z slowest), so the tets a wave handles mostly lie on one side.
struct SplitNodes {
    float4* lds;
    float4* g;
    int cap;
    __device__ __forceinline__ float4 ld(int i) const { return i < cap ? lds[i] : g[i]; }
    __device__ __forceinline__ void st(int i, float4 v) const {
        if (i < cap) lds[i] = v;
        else g[i] = v;
    }
};

__device__ __forceinline__ f3 xyz4(float4 p) { return mk3(p.x, p.y, p.z); }

// k_solve_edges (xpbd/xpbd_solver.cu:484-524) on one edge between nodes i and j. `prev` is read only when the
// Kelvin-Voigt term is on. Returns false where the reference returns without writing.
template <bool DAMPED, class Store>
__device__ __forceinline__ bool edge_project(const Store& S, const float4* __restrict__ prev, int i, int j,
                                             float restLen, float alphaTilde, float omega, float betaGammaH,
                                             float lam, float& lamOut) {
    float4 pi = S.ld(i), pj = S.ld(j);
    float wi = pi.w, wj = pj.w, wsum = wi + wj;
    if (wsum <= 0.f) return false;
    f3 dd = xyz4(pi) - xyz4(pj);
    float len = len3(dd);
    if (len < 1e-9f) return false;
    f3 n = dd * (1.f / len);
    float C = len - restLen;
    float gamma = betaGammaH;
    float gradDx = DAMPED ? dot3(n, (xyz4(pi) - xyz4(prev[i])) - (xyz4(pj) - xyz4(prev[j]))) : 0.f;
    float dl = (-C - alphaTilde * lam - gamma * gradDx) / ((1.f + gamma) * wsum + alphaTilde);
    float applied = omega * dl;
    lamOut = lam + applied;
    f3 xi = xyz4(pi) + n * (wi * applied);
    f3 xj = xyz4(pj) - n * (wj * applied);
    S.st(i, make_float4(xi.x, xi.y, xi.z, wi));
    S.st(j, make_float4(xj.x, xj.y, xj.z, wj));
    return true;
}

// k_solve_bending (xpbd/xpbd_solver.cu:534-573) on one bend element: apex i, arms j and k, rest angle theta0. The
// constraint is C = theta - theta0 with theta the apex angle between x_i - x_j and x_i - x_k; the corner directions
// are the unit arms' mutual perpendicular parts g_j = n_v - c n_u, g_k = n_u - c n_v, g_i = -(g_j + g_k) (|g_j| =
// |g_k| = sin theta), the reference's choice of the angle gradient scaled by arm length times sin theta. Returns
// false where the reference returns without writing (all three pinned, a degenerate arm, arms within 1e-4 of
// collinear, a vanishing denominator).
template <bool DAMPED, class Store>
__device__ __forceinline__ bool bend_project(const Store& S, const float4* __restrict__ prev, int i, int j, int k,
                                             float theta0, float alphaB, float omega, float betaGammaH, float lam,
                                             float& lamOut) {
    const float4 apex = S.ld(i), endA = S.ld(j), endB = S.ld(k);
    const float wApex = apex.w, wA = endA.w, wB = endB.w;
    if (wApex + wA + wB <= 0.f) return false;
    const f3 armA = xyz4(apex) - xyz4(endA), armB = xyz4(apex) - xyz4(endB);
    const float lenA = len3(armA), lenB = len3(armB);
    if (lenA < 1e-9f || lenB < 1e-9f) return false;
    const f3 unitA = armA * (1.f / lenA), unitB = armB * (1.f / lenB);
    const float cosT = fmaxf(-1.f, fminf(1.f, dot3(unitA, unitB)));
    if (sqrtf(fmaxf(1.f - cosT * cosT, 0.f)) < 1e-4f) return false;  // arms (anti)parallel: no defined plane
    const float err = own_acos(cosT) - theta0;
    // each arm end moves across its arm, within the plane of the two arms; the apex takes the opposite sum
    const f3 dirA = unitB - unitA * cosT, dirB = unitA - unitB * cosT;
    const f3 dirApex = (dirA + dirB) * -1.f;
    const float weighted = wApex * dot3(dirApex, dirApex) + wA * dot3(dirA, dirA) + wB * dot3(dirB, dirB);
    float rate = 0.f;  // the angle's rate along the directions over the substep (Kelvin-Voigt term)
    if (DAMPED)
        rate = dot3(dirApex, xyz4(apex) - xyz4(prev[i])) + dot3(dirA, xyz4(endA) - xyz4(prev[j])) +
               dot3(dirB, xyz4(endB) - xyz4(prev[k]));
    const float den = (1.f + betaGammaH) * weighted + alphaB;
    if (den < 1e-12f) return false;
    const float step = omega * ((-err - alphaB * lam - betaGammaH * rate) / den);
    lamOut = lam + step;
    const f3 newApex = xyz4(apex) + dirApex * (wApex * step), newA = xyz4(endA) + dirA * (wA * step),
             newB = xyz4(endB) + dirB * (wB * step);
    S.st(i, make_float4(newApex.x, newApex.y, newApex.z, wApex));
    S.st(j, make_float4(newA.x, newA.y, newA.z, wA));
    S.st(k, make_float4(newB.x, newB.y, newB.z, wB));
    return true;
}

// k_solve_tet (xpbd/xpbd_solver.cu:749-829) on one tet with nodes id[4]. MODE 0 keeps the hydrostatic row alone;
// 1 and 2 are the coupled 2x2 block, legacy and guarded. lamD / lamH are read and updated. PRECISE evaluates J - 1,
// log J, J^(-2/3) I1 - 3 and the deviatoric gradient from the displacement gradient (precise_strain, precise_dev)
// with rk the tet type's rest edges; otherwise the reference's direct forms, and rk is not read.
template <int MODE, bool DAMPED, bool PRECISE = false, class Store>
__device__ __forceinline__ bool tet_project(const Store& S, const float4* __restrict__ prev, const int id[4],
                                            const float* B, float ad, float ah, float& lamD, float& lamH,
                                            float omega, float dampingOverH, float bulkDampingOverH,
                                            const f3* rk = nullptr) {
    float4 q[4];
    for (int c = 0; c < 4; c++) q[c] = S.ld(id[c]);
    f3 p[4] = {xyz4(q[0]), xyz4(q[1]), xyz4(q[2]), xyz4(q[3])};
    float w[4] = {q[0].w, q[1].w, q[2].w, q[3].w};
    f3 e1 = p[1] - p[0], e2 = p[2] - p[0], e3 = p[3] - p[0];
    f3 br[3] = {mk3(B[0], B[1], B[2]), mk3(B[3], B[4], B[5]), mk3(B[6], B[7], B[8])};
    f3 F[3], cof[3];
    float J;
    PreciseStrain st;
    if (PRECISE) {
        f3 d[3] = {e1 - rk[0], e2 - rk[1], e3 - rk[2]};
        st = precise_strain(d, br);
        precise_frame(st, F, cof, J);
    } else {
        for (int c = 0; c < 3; c++) F[c] = e1 * comp3(br[0], c) + e2 * comp3(br[1], c) + e3 * comp3(br[2], c);
        cof[0] = cross3(F[1], F[2]);
        cof[1] = cross3(F[2], F[0]);
        cof[2] = cross3(F[0], F[1]);
        J = dot3(F[0], cof[0]);
    }
    f3 gh[4], gd[4];
    for (int c = 1; c < 4; c++) {
        f3 b = br[c - 1];
        gh[c] = cof[0] * b.x + cof[1] * b.y + cof[2] * b.z;
    }
    // Undamped instantiations run only with both damping terms zero: the scales are then exactly 1 and the divisions
    // by them and the damping terms drop out without changing a bit.
    float scale = DAMPED ? 1.f + dampingOverH : 1.f;
    if (MODE == 0) {
        float CH = PRECISE ? st.a : J - 1.f;
        gh[0] = (gh[1] + gh[2] + gh[3]) * -1.f;
        float bulkScale = scale;
        float effectiveAh = ah / bulkScale;
        float vh = 0.f;
        float schur = effectiveAh;
        if (DAMPED)
            for (int c = 0; c < 4; c++) vh += dot3(gh[c], p[c] - xyz4(prev[id[c]]));
        for (int c = 0; c < 4; c++) schur += w[c] * dot3(gh[c], gh[c]);
        float rh = DAMPED ? (-CH - ah * lamH - (bulkScale - 1.f) * vh) / bulkScale : -CH - ah * lamH;
        if (!(schur > 0.f)) return false;
        float dlH = omega * rh / schur;
        if (!isfinite(dlH)) return false;
        lamH = lamH + dlH;
        for (int c = 0; c < 4; c++) {
            f3 x = p[c] + gh[c] * dlH * w[c];
            S.st(id[c], make_float4(x.x, x.y, x.z, w[c]));
        }
        return true;
    }
    const bool safeguarded = MODE == 2;
    float I1 = PRECISE ? 0.f : dot3(F[0], F[0]) + dot3(F[1], F[1]) + dot3(F[2], F[2]);  // the direct forms' only
    float positiveJ = fmaxf(J, 0.05f);
    float iso = iso_pow(positiveJ);
    const bool jlive = J > 0.05f;
    f3 Gh[3];
    float qq = fmaxf(PRECISE ? precise_dev(st, F, cof, jlive, positiveJ, iso, Gh) : iso * I1 - 3.f, 0.f);
    float CD = sqrtf(fmaxf(qq, 1e-12f));
    float devScale = 1.f / CD;
    float volumeGradient = jlive ? iso * I1 / (3.f * positiveJ) : 0.f;
    float CH;
    if (safeguarded)
        CH = (PRECISE && jlive ? own_log1p(st.a) : own_log(positiveJ)) + fminf(0.f, (J - 0.05f) / 0.05f);
    else
        CH = PRECISE ? st.a : J - 1.f;
    for (int c = 1; c < 4; c++) {
        f3 b = br[c - 1];
        if (PRECISE) gd[c] = (Gh[0] * b.x + Gh[1] * b.y + Gh[2] * b.z) * devScale;
        else gd[c] = ((F[0] * b.x + F[1] * b.y + F[2] * b.z) * iso - gh[c] * volumeGradient) * devScale;
        if (safeguarded) gh[c] = gh[c] * (1.f / positiveJ);
    }
    gd[0] = (gd[1] + gd[2] + gd[3]) * -1.f;
    gh[0] = (gh[1] + gh[2] + gh[3]) * -1.f;
    float dd = 0.f, dh = 0.f, vd = 0.f, vh = 0.f;
    for (int c = 0; c < 4; c++) {
        dd += w[c] * dot3(gd[c], gd[c]);
        dh += w[c] * dot3(gd[c], gh[c]);
        if (DAMPED) {
            f3 dx = p[c] - xyz4(prev[id[c]]);
            vd += dot3(gd[c], dx);
            vh += dot3(gh[c], dx);
        }
    }
    float bulkScale = DAMPED ? (safeguarded ? 1.f + bulkDampingOverH : scale) : 1.f;
    float effectiveAd = ad / scale, effectiveAh = ah / bulkScale;
    float A = dd + effectiveAd, Bc = dh, gamma = Bc / A;
    float schur = effectiveAh + effectiveAd * gamma * gamma;
    for (int c = 0; c < 4; c++) {
        f3 residual = gh[c] - gd[c] * gamma;
        schur += w[c] * dot3(residual, residual);
    }
    float rd = DAMPED ? (-CD - ad * lamD - dampingOverH * vd) / scale : -CD - ad * lamD;
    float rh = DAMPED ? (-CH - ah * lamH - (bulkScale - 1.f) * vh) / bulkScale : -CH - ah * lamH;
    if (!(A > 0.f) || !(schur > 0.f)) return false;
    float dlH = omega * (rh - gamma * rd) / schur, dlD = (omega * rd - Bc * dlH) / A;
    if (!isfinite(dlD) || !isfinite(dlH)) return false;
    if (safeguarded) {
        float largest = 0.f;
        for (int c = 0; c < 4; c++) largest = fmaxf(largest, len3((gd[c] * dlD + gh[c] * dlH) * w[c]));
        float limit = 0.15f / fmaxf(len3(br[0]), fmaxf(len3(br[1]), len3(br[2])));
        float accepted = largest > limit ? limit / largest : 1.f;
        dlD *= accepted;
        dlH *= accepted;
    }
    lamD = lamD + dlD;
    lamH = lamH + dlH;
    for (int c = 0; c < 4; c++) {
        f3 x = p[c] + (gd[c] * dlD + gh[c] * dlH) * w[c];
        S.st(id[c], make_float4(x.x, x.y, x.z, w[c]));
    }
    return true;
}
