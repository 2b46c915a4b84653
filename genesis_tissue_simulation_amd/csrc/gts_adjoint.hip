// Adjoint of the XPBD substep for per-tet compliance gradients, batched over instances (one workgroup each).
//
// No reference counterpart: the reference's solver has no gradient. The forward is the legacy substep with
// `iterations` Gauss-Seidel iterations (multipliers reset every substep), no damping and no contact (coupled
// Neo-Hookean tets, edges with the instance's compliance), driven by a per-node force field. Checkpoints of (x, v)
// are kept every `ckpt` substeps. The backward pass walks the segments last to first: it re-runs a segment forward
// from its checkpoint, taping the inputs (positions and multipliers) of every projection of every substep and
// iteration, then reverses the substeps last to first, iterations last to first and colours last to first. Within a
// colour no two projections share a node, so the adjoint scatter needs no atomics, exactly as the forward sweep.
// The operation sequence mirrors the CPU checker's restatement, so the two agree bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>

#include "../../include/gts_xpbd.h"
#include "gts_common.h"
#include "gts_solver.h"

namespace {

// The scalar fields of a GtsBatchInstance record.
struct GtsBatchInstanceHead {
    int node_off, n_nodes, edge_off, n_edges, tet_off, n_tets, nx, nxy, n_edge_colours, n_tet_colours, colour_off;
    float edge_ks;
};

// Per-tet inputs of a forward projection, loaded ahead of use.
struct TetInputs {
    uint32_t d;
    float ad, ah, lamD, lamH;
};

struct AdjStep {
    int nsub, ckpt, nckpt, iters, lds_nodes;
    float h, invH, invH2, velScale;
};

// Per-instance pointers into the pooled buffers.
struct Inst {
    const GtsBatchInstance* rec;  // read in place: a local copy of the Dm^-1 table would be indexed from scratch
    GtsBatchInstanceHead I;
    float4 *pos, *prev, *vel, *xb, *vb, *pb, *ckx, *ckv;
    float4 *ttape, *etape;  // slot 0 of this instance; slots are tslot / eslot float4 apart
    float *ltape_t, *ltape_e;
    const float4* force;
    const uint32_t *ed, *td;
    const float* comp;
    float *alpha, *abar, *lamb_t, *lamb_e, *lam_t, *lam_e, *grad;
    const int *ecol, *tcol;
    size_t tslot, eslot, lslot;
};

__device__ __forceinline__ Inst bind(const GtsAdjointView& A, const AdjStep& P, int item) {
    Inst s;
    const GtsBatchView& B = A.batch;
    s.rec = reinterpret_cast<const GtsBatchInstance*>(B.instances) + item;
    s.I.node_off = s.rec->node_off;
    s.I.n_nodes = s.rec->n_nodes;
    s.I.edge_off = s.rec->edge_off;
    s.I.n_edges = s.rec->n_edges;
    s.I.tet_off = s.rec->tet_off;
    s.I.n_tets = s.rec->n_tets;
    s.I.nx = s.rec->nx;
    s.I.nxy = s.rec->nxy;
    s.I.n_edge_colours = s.rec->n_edge_colours;
    s.I.n_tet_colours = s.rec->n_tet_colours;
    s.I.colour_off = s.rec->colour_off;
    s.I.edge_ks = s.rec->edge_ks;
    const size_t n0 = s.I.node_off, sN = A.total_nodes, sT = A.total_tets, sE = A.total_edges;
    s.pos = (float4*)B.pos + n0;
    s.prev = (float4*)B.prev + n0;
    s.vel = (float4*)B.vel + n0;
    s.xb = (float4*)A.xbar + n0;
    s.vb = (float4*)A.vbar + n0;
    s.pb = (float4*)A.prevbar + n0;
    s.force = (const float4*)A.force + n0;
    // checkpoints: [nckpt][sum N] rows of x, then [nckpt][sum N] rows of v
    s.ckx = (float4*)A.checkpoints + n0;
    s.ckv = (float4*)A.checkpoints + (size_t)P.nckpt * sN + n0;
    // tapes: slot j * iters + it for substep j of the segment and iteration it
    s.tslot = 4 * sT;
    s.eslot = 2 * sE;
    s.lslot = 2 * sT + sE;
    s.ttape = (float4*)A.tet_tape + 4 * (size_t)s.I.tet_off;
    s.etape = (float4*)A.edge_tape + 2 * (size_t)s.I.edge_off;
    s.ltape_t = A.lambda_tape + 2 * (size_t)s.I.tet_off;
    s.ltape_e = A.lambda_tape + 2 * sT + s.I.edge_off;
    s.ed = B.edge_desc + s.I.edge_off;
    s.td = B.tet_desc + s.I.tet_off;
    s.comp = A.compliance + s.I.tet_off;
    s.alpha = A.alpha + 2 * (size_t)s.I.tet_off;
    s.abar = A.abar + 2 * (size_t)s.I.tet_off;
    s.lamb_t = A.lambar + 2 * (size_t)s.I.tet_off;
    s.lamb_e = A.lambar + 2 * sT + s.I.edge_off;
    s.lam_t = B.lambda_tet + 2 * (size_t)s.I.tet_off;
    s.lam_e = B.lambda_edge + s.I.edge_off;
    s.grad = A.grad + s.I.tet_off;
    s.ecol = B.colour_table + s.I.colour_off;
    s.tcol = s.ecol + s.I.n_edge_colours + 1;
    return s;
}

__device__ __forceinline__ void tet_ids(const Inst& s, uint32_t d, int id[4], int& type) {
    int base = (int)(d >> 4);
    type = (int)((d >> 1) & 7u);
    for (int q = 0; q < 4; q++) id[q] = base + corner_offset(kuhn_corner(type, q), s.I.nx, s.I.nxy);
}

__device__ __forceinline__ int edge_partner(const Inst& s, uint32_t d, int& axis) {
    axis = (int)((d >> 1) & 3u);
    return (int)(d >> 4) + (axis == 0 ? 1 : (axis == 1 ? s.I.nx : s.I.nxy));
}

// alpha~_D = (1/h^2) / (1/c) and alpha~_H = (1/h^2) / (10/c) of every tet, once per rollout.
__device__ void compute_alphas(const Inst& s, const AdjStep& P) {
    for (int k = threadIdx.x; k < s.I.n_tets; k += blockDim.x) {
        float c = s.comp[k];
        float kd = (float)(1.0 / (double)c), kh = (float)(1.0 / ((double)c / 10.0));
        s.alpha[k] = P.invH2 / kd;
        s.alpha[s.I.n_tets + k] = P.invH2 / kh;
    }
    __syncthreads();
}

// One forward substep of the fit model on node positions S (LDS-staged or the instance's global slice); with TAPE,
// records the inputs of every projection of every iteration in the tape slots of substep `slot`.
template <bool TAPE, class Store>
__device__ void substep_forward(const Inst& s, const Store& S, const AdjStep& P, int slot) {
    const int tid = threadIdx.x, bs = blockDim.x, N = s.I.n_nodes, T = s.I.n_tets;
    for (int i = tid; i < N; i += bs) {
        float4 p = S.ld(i);
        s.prev[i] = p;
        if (p.w > 0.f) {
            float4 F4 = s.force[i];
            f3 a = mk3(F4.x, F4.y, F4.z) * p.w;
            float4 v0 = s.vel[i];
            f3 v = mk3(v0.x, v0.y, v0.z) + a * P.h;
            s.vel[i] = make_float4(v.x, v.y, v.z, 0.f);
            f3 x = xyz4(p) + v * P.h;
            S.st(i, make_float4(x.x, x.y, x.z, p.w));
        }
    }
    __syncthreads();
    const float at = s.I.edge_ks > 0.f ? 1.f / (s.I.edge_ks * P.h * P.h) : 0.f;
    const bool write = P.iters > 1;
    for (int it = 0; it < P.iters; it++) {
        const bool read = it > 0;
        const size_t sl = (size_t)slot * P.iters + it;
        float4* et = s.etape + sl * s.eslot;
        float4* tt = s.ttape + sl * s.tslot;
        float* lte = s.ltape_e + sl * s.lslot;
        float* ltt = s.ltape_t + sl * s.lslot;
        for (int c = 0; c < s.I.n_edge_colours; c++) {
            for (int k = s.ecol[c] + tid; k < s.ecol[c + 1]; k += bs) {
                uint32_t d = s.ed[k];
                if (!(d & 1u)) continue;
                int axis, i = (int)(d >> 4), j = edge_partner(s, d, axis);
                float lam = read ? s.lam_e[k] : 0.f, out;
                if (TAPE) {
                    et[2 * k] = S.ld(i);
                    et[2 * k + 1] = S.ld(j);
                    if (write) lte[k] = lam;
                }
                const bool ok = edge_project<false>(S, s.prev, i, j, s.rec->rest_len[axis], at, 1.f, 0.f, lam, out);
                if (write && (ok || !read)) s.lam_e[k] = ok ? out : 0.f;  // 0: the substep's reset
            }
            __syncthreads();
        }
        // This thread's first tet of the next colour is loaded while the current colour computes: no projection of
        // one colour writes another colour's descriptor, alpha or multipliers.
        auto fetch = [&](int k, int end) {
            TetInputs t;
            t.d = 0u;
            t.ad = t.ah = t.lamD = t.lamH = 0.f;
            if (k < end) {
                t.d = s.td[k];
                t.ad = s.alpha[k];
                t.ah = s.alpha[T + k];
                if (read) {
                    t.lamD = s.lam_t[k];
                    t.lamH = s.lam_t[T + k];
                }
            }
            return t;
        };
        TetInputs next = fetch(s.tcol[0] + tid, s.tcol[1]);
        for (int c = 0; c < s.I.n_tet_colours; c++) {
            const int k0 = s.tcol[c] + tid, end = s.tcol[c + 1];
            const TetInputs first = next;
            if (c + 1 < s.I.n_tet_colours) next = fetch(s.tcol[c + 1] + tid, s.tcol[c + 2]);
            for (int k = k0; k < end; k += bs) {
                const TetInputs in = k == k0 ? first : fetch(k, end);
                uint32_t d = in.d;
                if (!(d & 1u)) continue;
                int id[4], type;
                tet_ids(s, d, id, type);
                float lamD = in.lamD, lamH = in.lamH;
                if (TAPE) {
                    for (int q = 0; q < 4; q++) tt[4 * k + q] = S.ld(id[q]);
                    if (write) {
                        ltt[k] = lamD;
                        ltt[T + k] = lamH;
                    }
                }
                f3 rk[3];
                for (int q = 0; q < 3; q++) rk[q] = kuhn_rest_edge(type, q + 1, s.rec->rest_len);
                const bool ok = tet_project<1, false, true>(S, s.prev, id, s.rec->inv_rest[type], in.ad, in.ah, lamD,
                                                            lamH, 1.f, 0.f, 0.f, rk);
                if (write && (ok || !read)) {  // 0: the substep's reset
                    s.lam_t[k] = ok ? lamD : 0.f;
                    s.lam_t[T + k] = ok ? lamH : 0.f;
                }
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < N; i += bs) {
        float4 p = S.ld(i);
        f3 v = p.w > 0.f ? (xyz4(p) - xyz4(s.prev[i])) * P.invH : mk3(0, 0, 0);
        v = p.w > 0.f ? v * P.velScale : mk3(0, 0, 0);
        s.vel[i] = make_float4(v.x, v.y, v.z, 0.f);
    }
    __syncthreads();
}

// Reverse of the tet projection (mode 1, no damping, omega = 1, precise invariants with rest edges rk) at input
// positions p and multipliers lamD / lamH. lDb / lHb hold the adjoints of the output multipliers on entry and of the
// input multipliers on return.
//
// The deviatoric row at its clamp: where q = J^(-2/3) I1 - 3 is at most the 1e-12 floor of C_D = sqrt(max(q, 1e-12)),
// the forward divides the direction Ghat Dm^-T by the constant 1e-6, so an O(1) change of F moves the output by
// O(1e6) times the (tiny) step. That derivative is exact but meaningless, and in FP32 it overflows within a few
// substeps. The subgradient taken there holds the row's value and direction constant: no adjoint flows from C_D or
// from the direction into F. Above the floor, the q path enters F as 2 qb Ghat, Ghat evaluated precisely.
__device__ void tet_reverse(const f3 p[4], const float w[4], const float* Bm, const f3 rk[3], float ad, float ah,
                            float lamD, float lamH, const f3 ob[4], f3 pb[4], float& adb, float& ahb, float& lDb,
                            float& lHb) {
    f3 br[3] = {mk3(Bm[0], Bm[1], Bm[2]), mk3(Bm[3], Bm[4], Bm[5]), mk3(Bm[6], Bm[7], Bm[8])};
    f3 e1 = p[1] - p[0], e2 = p[2] - p[0], e3 = p[3] - p[0];
    f3 d[3] = {e1 - rk[0], e2 - rk[1], e3 - rk[2]};
    PreciseStrain st = precise_strain(d, br);
    f3 F[3], cof[3];
    float J;
    precise_frame(st, F, cof, J);
    f3 gh[4], gd[4], Fb[4], G[4];
    for (int k = 1; k < 4; k++) {
        f3 b = br[k - 1];
        gh[k] = cof[0] * b.x + cof[1] * b.y + cof[2] * b.z;
    }
    float I1 = 3.f + st.b;
    float pJ = fmaxf(J, 0.05f);
    float iso = iso_pow(pJ);
    bool jlive = J > 0.05f;
    f3 Gh[3];
    float qraw = precise_dev(st, F, cof, jlive, pJ, iso, Gh);
    float q = fmaxf(qraw, 0.f);
    float CD = sqrtf(fmaxf(q, 1e-12f));
    float ds = 1.f / CD;
    const bool frozen = !(q > 1e-12f);
    float vg = jlive ? iso * I1 / (3.f * pJ) : 0.f;
    float CH = st.a;
    for (int k = 1; k < 4; k++) {
        f3 b = br[k - 1];
        Fb[k] = F[0] * b.x + F[1] * b.y + F[2] * b.z;
        G[k] = Gh[0] * b.x + Gh[1] * b.y + Gh[2] * b.z;
        gd[k] = G[k] * ds;
    }
    gd[0] = (gd[1] + gd[2] + gd[3]) * -1.f;
    gh[0] = (gh[1] + gh[2] + gh[3]) * -1.f;
    float dd = 0.f, dh = 0.f;
    for (int k = 0; k < 4; k++) {
        dd += w[k] * dot3(gd[k], gd[k]);
        dh += w[k] * dot3(gd[k], gh[k]);
    }
    float A = dd + ad, gamma = dh / A;
    f3 Rk[4];
    float schur = ah + ad * gamma * gamma;
    for (int k = 0; k < 4; k++) {
        Rk[k] = gh[k] - gd[k] * gamma;
        schur += w[k] * dot3(Rk[k], Rk[k]);
    }
    float rd = -CD - ad * lamD, rh = -CH - ah * lamH;
    for (int k = 0; k < 4; k++) pb[k] = ob[k];
    if (!(A > 0.f) || !(schur > 0.f)) return;
    float dlH = (rh - gamma * rd) / schur, dlD = (rd - dh * dlH) / A;
    if (!isfinite(dlD) || !isfinite(dlH)) return;
    f3 gdb[4], ghb[4];
    float dlDb = lDb, dlHb = lHb;
    for (int k = 0; k < 4; k++) {
        f3 u = ob[k] * w[k];
        gdb[k] = u * dlD;
        ghb[k] = u * dlH;
        dlDb += dot3(u, gd[k]);
        dlHb += dot3(u, gh[k]);
    }
    float rdb = dlDb / A, dhb = -dlDb * dlH / A, Ab = -dlDb * dlD / A;
    dlHb += -dlDb * dh / A;
    float rhb = dlHb / schur, gammab = -dlHb * rd / schur, schurb = -dlHb * dlH / schur;
    rdb += -dlHb * gamma / schur;
    float CDb = -rdb, CHb = -rhb;
    adb += -lamD * rdb;
    ahb += -lamH * rhb;
    lDb += -ad * rdb;
    lHb += -ah * rhb;
    ahb += schurb;
    adb += schurb * gamma * gamma;
    gammab += schurb * 2.f * ad * gamma;
    for (int k = 0; k < 4; k++) {
        f3 Rb = Rk[k] * (schurb * 2.f * w[k]);
        ghb[k] = ghb[k] + Rb;
        gdb[k] = gdb[k] - Rb * gamma;
        gammab += -dot3(Rb, gd[k]);
    }
    dhb += gammab / A;
    Ab += -gammab * gamma / A;
    float ddb = Ab;
    adb += Ab;
    for (int k = 0; k < 4; k++) {
        gdb[k] = gdb[k] + gd[k] * (2.f * w[k] * ddb) + gh[k] * (w[k] * dhb);
        ghb[k] = ghb[k] + gd[k] * (w[k] * dhb);
    }
    for (int k = 1; k < 4; k++) {
        gdb[k] = gdb[k] - gdb[0];
        ghb[k] = ghb[k] - ghb[0];
    }
    float dsb = 0.f, isob = 0.f, vgb = 0.f;
    f3 Fbb[4] = {mk3(0, 0, 0), mk3(0, 0, 0), mk3(0, 0, 0), mk3(0, 0, 0)};
    if (!frozen)
        for (int k = 1; k < 4; k++) {
            f3 Gb = gdb[k] * ds;
            dsb += dot3(gdb[k], G[k]);
            Fbb[k] = Gb * iso;
            isob += dot3(Gb, Fb[k]);
            ghb[k] = ghb[k] - Gb * vg;
            vgb += -dot3(Gb, gh[k]);
        }
    f3 Fbar[3] = {mk3(0, 0, 0), mk3(0, 0, 0), mk3(0, 0, 0)};
    for (int k = 1; k < 4; k++) {
        f3 b = br[k - 1];
        Fbar[0] = Fbar[0] + Fbb[k] * b.x;
        Fbar[1] = Fbar[1] + Fbb[k] * b.y;
        Fbar[2] = Fbar[2] + Fbb[k] * b.z;
    }
    CDb += -dsb / (CD * CD);
    float qb = frozen ? 0.f : CDb / (2.f * CD);
    if (!frozen)
        for (int c = 0; c < 3; c++) Fbar[c] = Fbar[c] + Gh[c] * (2.f * qb);
    float I1b = 0.f;
    float pJb = 0.f;
    if (jlive) {
        isob += vgb * I1 / (3.f * pJ);
        I1b += vgb * iso / (3.f * pJ);
        pJb += -vgb * iso * I1 / (3.f * pJ * pJ);
    }
    pJb += isob * ((float)(-2.0 / 3.0)) * iso / pJ;
    float Jb = CHb + (jlive ? pJb : 0.f);
    for (int c = 0; c < 3; c++) Fbar[c] = Fbar[c] + F[c] * (2.f * I1b);
    f3 cofb[3] = {mk3(0, 0, 0), mk3(0, 0, 0), mk3(0, 0, 0)};
    for (int k = 1; k < 4; k++) {
        f3 b = br[k - 1];
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
        f3 eb = Fbar[0] * comp3(br[k], 0) + Fbar[1] * comp3(br[k], 1) + Fbar[2] * comp3(br[k], 2);
        pb[k + 1] = pb[k + 1] + eb;
        pb[0] = pb[0] - eb;
    }
}

// Reverse of the edge projection (no damping, omega = 1) at inputs xi, xj and multiplier lam; lb holds the adjoint
// of the output multiplier on entry and of the input multiplier on return.
__device__ __forceinline__ void edge_reverse(f3 xi, f3 xj, float wi, float wj, float L0, float at, float lam, f3 oib,
                                             f3 ojb, f3& xib, f3& xjb, float& lb) {
    xib = oib;
    xjb = ojb;
    float wsum = wi + wj;
    if (wsum <= 0.f) return;
    f3 d = xi - xj;
    float L = len3(d);
    if (L < 1e-9f) return;
    f3 n = d * (1.f / L);
    float C = L - L0;
    float kap = 1.f / (wsum + at);
    float dl = (-C - at * lam) * kap;
    f3 m = oib * wi - ojb * wj;
    f3 nb = m * dl;
    float dlb = dot3(m, n) + lb;
    float Lb = -dlb * kap;
    lb += -at * kap * dlb;
    f3 db = (nb - n * dot3(nb, n)) * (1.f / L) + n * Lb;
    xib = xib + db;
    xjb = xjb - db;
}

// Reverse of the substep whose tapes are in the slots of `slot`. wsrc holds float4 node rows whose .w is the inverse
// mass (the segment's checkpoint), X the position adjoints; the velocity adjoints and the adjoint of prev are in the
// instance's global slices.
template <class XStore>
__device__ void substep_reverse(const Inst& s, const float4* __restrict__ wsrc, const XStore& X, const AdjStep& P,
                                int slot) {
    const int tid = threadIdx.x, bs = blockDim.x, N = s.I.n_nodes, T = s.I.n_tets;
    // velocity: v' = (x_end - prev) / h * velScale for w > 0
    for (int i = tid; i < N; i += bs) {
        float w = wsrc[i].w;
        if (w > 0.f) {
            f3 g = xyz4(s.vb[i]) * (P.invH * P.velScale);
            f3 x = xyz4(X.ld(i)) + g;
            X.st(i, make_float4(x.x, x.y, x.z, 0.f));
            f3 m = g * -1.f;
            s.pb[i] = make_float4(m.x, m.y, m.z, 0.f);
        } else {
            s.pb[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    // the multipliers of the last iteration reach nothing after the substep
    for (int k = tid; k < 2 * T; k += bs) s.lamb_t[k] = 0.f;
    for (int k = tid; k < s.I.n_edges; k += bs) s.lamb_e[k] = 0.f;
    __syncthreads();
    const float at = s.I.edge_ks > 0.f ? 1.f / (s.I.edge_ks * P.h * P.h) : 0.f;
    const bool tapedLam = P.iters > 1;
    for (int it = P.iters - 1; it >= 0; it--) {
        const size_t sl = (size_t)slot * P.iters + it;
        const float4* et = s.etape + sl * s.eslot;
        const float4* tt = s.ttape + sl * s.tslot;
        const float* lte = s.ltape_e + sl * s.lslot;
        const float* ltt = s.ltape_t + sl * s.lslot;
        // The descriptor and alphas of this thread's first tet of the next (lower) colour are loaded ahead, as in
        // the forward sweep.
        auto fetch = [&](int k, int end) {
            TetInputs t;
            t.d = 0u;
            t.ad = t.ah = t.lamD = t.lamH = 0.f;
            if (k < end) {
                t.d = s.td[k];
                t.ad = s.alpha[k];
                t.ah = s.alpha[T + k];
            }
            return t;
        };
        const int top = s.I.n_tet_colours - 1;
        TetInputs next = top >= 0 ? fetch(s.tcol[top] + tid, s.tcol[top + 1]) : TetInputs{0u, 0.f, 0.f, 0.f, 0.f};
        for (int c = top; c >= 0; c--) {
            const int k0 = s.tcol[c] + tid, end = s.tcol[c + 1];
            const TetInputs first = next;
            if (c > 0) next = fetch(s.tcol[c - 1] + tid, s.tcol[c]);
            for (int k = k0; k < end; k += bs) {
                const TetInputs in = k == k0 ? first : fetch(k, end);
                uint32_t d = in.d;
                if (!(d & 1u)) continue;
                int id[4], type;
                tet_ids(s, d, id, type);
                f3 p[4], ob[4], pb[4];
                float w[4];
                for (int q = 0; q < 4; q++) {
                    float4 t = tt[4 * k + q];
                    p[q] = xyz4(t);
                    w[q] = t.w;
                    ob[q] = xyz4(X.ld(id[q]));
                }
                float lamD = tapedLam ? ltt[k] : 0.f, lamH = tapedLam ? ltt[T + k] : 0.f;
                float adb = s.abar[k], ahb = s.abar[T + k], lDb = s.lamb_t[k], lHb = s.lamb_t[T + k];
                f3 rk[3];
                for (int q = 0; q < 3; q++) rk[q] = kuhn_rest_edge(type, q + 1, s.rec->rest_len);
                tet_reverse(p, w, s.rec->inv_rest[type], rk, in.ad, in.ah, lamD, lamH, ob, pb, adb, ahb, lDb, lHb);
                for (int q = 0; q < 4; q++) X.st(id[q], make_float4(pb[q].x, pb[q].y, pb[q].z, 0.f));
                s.abar[k] = adb;
                s.abar[T + k] = ahb;
                s.lamb_t[k] = lDb;
                s.lamb_t[T + k] = lHb;
            }
            __syncthreads();
        }
        for (int c = s.I.n_edge_colours - 1; c >= 0; c--) {
            for (int k = s.ecol[c] + tid; k < s.ecol[c + 1]; k += bs) {
                uint32_t d = s.ed[k];
                if (!(d & 1u)) continue;
                int axis, i = (int)(d >> 4), j = edge_partner(s, d, axis);
                float4 ti = et[2 * k], tj = et[2 * k + 1];
                float lam = tapedLam ? lte[k] : 0.f, lb = s.lamb_e[k];
                f3 xib, xjb;
                edge_reverse(xyz4(ti), xyz4(tj), ti.w, tj.w, s.rec->rest_len[axis], at, lam, xyz4(X.ld(i)),
                             xyz4(X.ld(j)), xib, xjb, lb);
                X.st(i, make_float4(xib.x, xib.y, xib.z, 0.f));
                X.st(j, make_float4(xjb.x, xjb.y, xjb.z, 0.f));
                s.lamb_e[k] = lb;
            }
            __syncthreads();
        }
    }
    // predict: xp = x + h (v + h a) for w > 0, prev = x
    for (int i = tid; i < N; i += bs) {
        float w = wsrc[i].w;
        f3 g = xyz4(X.ld(i));
        f3 x = g + xyz4(s.pb[i]);
        X.st(i, make_float4(x.x, x.y, x.z, 0.f));
        f3 v = w > 0.f ? g * P.h : mk3(0, 0, 0);
        s.vb[i] = make_float4(v.x, v.y, v.z, 0.f);
    }
    __syncthreads();
}

template <class Store>
__device__ void save_state(const Inst& s, const Store& S, float4* xs, float4* vs) {
    for (int i = threadIdx.x; i < s.I.n_nodes; i += blockDim.x) {
        xs[i] = S.ld(i);
        vs[i] = s.vel[i];
    }
    __syncthreads();
}
template <class Store>
__device__ void load_state(const Inst& s, const Store& S, const float4* xs, const float4* vs) {
    for (int i = threadIdx.x; i < s.I.n_nodes; i += blockDim.x) {
        S.st(i, xs[i]);
        s.vel[i] = vs[i];
    }
    __syncthreads();
}

// The rollout, keeping a checkpoint before every ckpt-th substep. The last segment is taped on the way (into the
// tape half the backward reverses first), so the backward starts with its reverse.
template <class Store>
__device__ void forward_run(const Inst& s, const Store& S, const AdjStep& P, size_t sN) {
    const int last = P.nckpt - 1, n0 = last * P.ckpt, half = last & 1;
    for (int n = 0; n < P.nsub; n++) {
        if (n % P.ckpt == 0) save_state(s, S, s.ckx + (size_t)(n / P.ckpt) * sN, s.ckv + (size_t)(n / P.ckpt) * sN);
        if (n >= n0) substep_forward<true>(s, S, P, half * P.ckpt + (n - n0));
        else substep_forward<false>(s, S, P, 0);
    }
}

// Taped re-run of segment `seg` from its checkpoint into tape half `half`.
template <class Store>
__device__ void tape_segment(const Inst& s, const Store& S, const AdjStep& P, size_t sN, int seg, int half) {
    const int n0 = seg * P.ckpt, len = min(P.nsub, n0 + P.ckpt) - n0;
    load_state(s, S, s.ckx + (size_t)seg * sN, s.ckv + (size_t)seg * sN);
    for (int j = 0; j < len; j++) substep_forward<true>(s, S, P, half * P.ckpt + j);
}

// Reverse of segment `seg` from tape half `half`.
template <class XStore>
__device__ void reverse_segment(const Inst& s, const XStore& X, const AdjStep& P, size_t sN, int seg, int half) {
    const int n0 = seg * P.ckpt, len = min(P.nsub, n0 + P.ckpt) - n0;
    const float4* w = s.ckx + (size_t)seg * sN;
    for (int j = len - 1; j >= 0; j--) substep_reverse(s, w, X, P, half * P.ckpt + j);
}

// One workgroup per instance at a time. An instance whose nodes fit in the dynamic LDS (lds_nodes float4) keeps its
// positions there for the whole rollout; a larger one works on its global slice.
__global__ void __launch_bounds__(1024) k_adjoint_forward(GtsAdjointView A, AdjStep P) {
    extern __shared__ float4 lds[];
    const GtsBatchView& B = A.batch;
    const size_t sN = A.total_nodes;
    for (int w = B.work_offsets[blockIdx.x]; w < B.work_offsets[blockIdx.x + 1]; w++) {
        Inst s = bind(A, P, B.work_items[w]);
        const int N = s.I.n_nodes;
        compute_alphas(s, P);
        if (N <= P.lds_nodes) {
            for (int i = threadIdx.x; i < N; i += blockDim.x) lds[i] = s.pos[i];
            __syncthreads();
            forward_run(s, LdsNodes{lds}, P, sN);
            for (int i = threadIdx.x; i < N; i += blockDim.x) s.pos[i] = lds[i];
        } else {
            forward_run(s, GlobalNodes{s.pos}, P, sN);
        }
        __syncthreads();
    }
}

// Taped re-run of one segment of every instance (positions in LDS when they fit). It touches the forward state
// (positions, prev, velocities, multipliers) and the tapes of one half only, so it can run beside the reverse of the
// following segment, which reads the other half.
__global__ void __launch_bounds__(1024) k_adjoint_tape(GtsAdjointView A, AdjStep P, int seg, int half) {
    extern __shared__ float4 lds[];
    const GtsBatchView& B = A.batch;
    const size_t sN = A.total_nodes;
    for (int w = B.work_offsets[blockIdx.x]; w < B.work_offsets[blockIdx.x + 1]; w++) {
        Inst s = bind(A, P, B.work_items[w]);
        if (s.I.n_nodes <= P.lds_nodes) tape_segment(s, LdsNodes{lds}, P, sN, seg, half);
        else tape_segment(s, GlobalNodes{s.pos}, P, sN, seg, half);
        __syncthreads();
    }
}

// Reverse of one segment of every instance (position adjoints in LDS when they fit). The last segment starts the
// alpha adjoints at zero; segment 0 turns them into the compliance gradient.
__global__ void __launch_bounds__(GTS_ADJOINT_BACKWARD_MAX_THREADS) k_adjoint_reverse(GtsAdjointView A, AdjStep P,
                                                                                       int seg, int half) {
    extern __shared__ float4 lds[];
    const GtsBatchView& B = A.batch;
    const size_t sN = A.total_nodes;
    const bool first = seg == P.nckpt - 1, last = seg == 0;
    for (int w = B.work_offsets[blockIdx.x]; w < B.work_offsets[blockIdx.x + 1]; w++) {
        Inst s = bind(A, P, B.work_items[w]);
        const int N = s.I.n_nodes;
        if (first)
            for (int k = threadIdx.x; k < 2 * s.I.n_tets; k += blockDim.x) s.abar[k] = 0.f;
        if (N <= P.lds_nodes) {
            for (int i = threadIdx.x; i < N; i += blockDim.x) lds[i] = s.xb[i];
            __syncthreads();
            reverse_segment(s, LdsNodes{lds}, P, sN, seg, half);
            for (int i = threadIdx.x; i < N; i += blockDim.x) s.xb[i] = lds[i];
        } else {
            __syncthreads();
            reverse_segment(s, GlobalNodes{s.xb}, P, sN, seg, half);
        }
        if (last)
            for (int k = threadIdx.x; k < s.I.n_tets; k += blockDim.x)
                s.grad[k] = (s.abar[k] + s.abar[s.I.n_tets + k] / 10.f) * P.invH2;
        __syncthreads();
    }
}

int plan_step(const GtsStepParams* p, int nsub, int ckpt, AdjStep& P) {
    if (!p || nsub <= 0 || ckpt <= 0) return GTS_ERR_BAD_ARG;
    if (p->tet_mode != GTS_TET_NEOHOOKEAN || p->iterations < 1 || p->strain_damping != 0.f ||
        p->bulk_damping != 0.f || p->beta_s != 0.f || p->omega != 1.f || p->damping != 0.f || p->rayleigh != 0.f ||
        p->guard_passes != 0 || p->bend_kb != 0.f)
        return GTS_ERR_BAD_ARG;  // the adjoint covers the undamped, unguarded Neo-Hookean substep without bending
    int N = p->substeps > 0 ? p->substeps : 1;
    P.h = p->dt / (float)N;
    if (!(P.h > 0.f) || !std::isfinite(P.h)) return GTS_ERR_BAD_DT;
    P.invH = 1.f / P.h;
    P.invH2 = 1.f / (P.h * P.h);
    P.velScale = 1.f;
    P.nsub = nsub;
    P.ckpt = ckpt;
    P.nckpt = (nsub + ckpt - 1) / ckpt;
    P.iters = p->iterations;
    P.lds_nodes = 0;
    return GTS_OK;
}

int check_view(const GtsAdjointView* A) {
    if (!A || A->batch.n_workgroups <= 0 || A->batch.block_size <= 0 || A->batch.block_size > 1024 ||
        A->total_nodes <= 0 || A->total_tets < 0 || A->total_edges < 0 || A->batch.lds_nodes < 0)
        return GTS_ERR_BAD_ARG;
    return GTS_OK;
}

int lds_bytes_max() {
    int dev = 0, bytes = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
    return bytes;
}

template <class Kernel, class... Extra>
int launch(Kernel kern, const GtsAdjointView* A, AdjStep& P, int threads, hipStream_t stream, Extra... extra) {
    size_t lds = (size_t)A->batch.lds_nodes * sizeof(float4);
    if (lds > (size_t)lds_bytes_max()) return GTS_ERR_BAD_ARG;
    P.lds_nodes = A->batch.lds_nodes;
    hipError_t e =
        hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return -(int)e;
    kern<<<A->batch.n_workgroups, threads, lds, stream>>>(*A, P, extra...);
    return -(int)hipGetLastError();
}

// The side stream and events of the backward pipeline, created on first use (default priority).
struct Pipeline {
    hipStream_t side = nullptr;
    hipEvent_t start = nullptr, taped[2] = {nullptr, nullptr}, reversed[2] = {nullptr, nullptr};
};
Pipeline g_pipe;

int pipeline_ready() {
    if (g_pipe.side) return GTS_OK;
    hipError_t e = hipStreamCreateWithFlags(&g_pipe.side, hipStreamNonBlocking);
    if (e != hipSuccess) return -(int)e;
    hipEvent_t* all[5] = {&g_pipe.start, &g_pipe.taped[0], &g_pipe.taped[1], &g_pipe.reversed[0],
                          &g_pipe.reversed[1]};
    for (hipEvent_t* ev : all) {
        e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
        if (e != hipSuccess) return -(int)e;
    }
    return GTS_OK;
}

#define ADJ_TRY(call)                       \
    do {                                    \
        hipError_t e_ = (call);             \
        if (e_ != hipSuccess) return -(int)e_; \
    } while (0)

}  // namespace

extern "C" int gts_adjoint_forward(const GtsAdjointView* A, const GtsStepParams* p, int nsub, int ckpt, void* stream) {
    if (check_view(A)) return GTS_ERR_BAD_ARG;
    AdjStep P;
    int rc = plan_step(p, nsub, ckpt, P);
    if (rc) return rc;
    return launch(k_adjoint_forward, A, P, A->batch.block_size, reinterpret_cast<hipStream_t>(stream));
}

// The segments are reversed last to first on the caller's stream; the last one was taped by the forward. The taped
// re-run of segment seg - 1 (into the other tape half) runs on a side stream while segment seg is reversed; events
// order each re-run after the reverse that last read its half and each reverse after its re-run.
extern "C" int gts_adjoint_backward(const GtsAdjointView* A, const GtsStepParams* p, int nsub, int ckpt, void* stream) {
    if (check_view(A)) return GTS_ERR_BAD_ARG;
    AdjStep P;
    int rc = plan_step(p, nsub, ckpt, P);
    if (rc) return rc;
    rc = pipeline_ready();
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream), side = g_pipe.side;
    const int threads = A->batch.block_size < GTS_ADJOINT_BACKWARD_MAX_THREADS ? A->batch.block_size
                                                                                : GTS_ADJOINT_BACKWARD_MAX_THREADS;
    const int last = P.nckpt - 1;
    // the forward taped the last segment; every later re-run follows the work already on `stream`
    ADJ_TRY(hipEventRecord(g_pipe.start, s));
    ADJ_TRY(hipStreamWaitEvent(side, g_pipe.start, 0));
    for (int seg = last; seg >= 0; seg--) {
        const int half = seg & 1;
        if (seg < last) ADJ_TRY(hipStreamWaitEvent(s, g_pipe.taped[half], 0));
        if (seg > 0) {
            const int other = (seg - 1) & 1;
            if (seg + 1 <= last) ADJ_TRY(hipStreamWaitEvent(side, g_pipe.reversed[other], 0));
            rc = launch(k_adjoint_tape, A, P, A->batch.block_size, side, seg - 1, other);
            if (rc) return rc;
            ADJ_TRY(hipEventRecord(g_pipe.taped[other], side));
        }
        rc = launch(k_adjoint_reverse, A, P, threads, s, seg, half);
        if (rc) return rc;
        ADJ_TRY(hipEventRecord(g_pipe.reversed[half], s));
    }
    return GTS_OK;
}
