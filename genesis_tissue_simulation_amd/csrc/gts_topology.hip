// Host-side topology of one lattice tissue: edges, Kuhn tets, both greedy colourings, per-type rest data.
//
// Restates the reference set-up path that turns the lattice into solver topology:
//   xpbd_bind      one edge per positive neighbour slot, p ascending, axis 0..2, L0 = |restNbr|
//                  (xpbd/xpbd_solver.cu:1445-1463);
//   xpbd_init      greedy first-free edge colouring on a 64-bit per-node mask (:1213-1230);
//   xpbd_build_tets Kuhn split of every all-active cell, Dm^-1 per type from the rest spacing in double, each rest
//                  tet checked against its type, greedy 4-node tet colouring (:1600-1768);
//   xpbd_bind + xpbd_set_bending  one bend element per live bend pair (apex ascending, pair 0..14; :1466-1478),
//                  greedy 3-node colouring (:1344-1356).
// The device order is colour-major with generation order inside a colour; the descriptors are packed for the
// kernels (see gts_kernels.hip).
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/gts_xpbd.h"
#include "gts_common.h"

namespace {

const int KUHN[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7}, {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};

bool invert3(const double m[3][3], double out[3][3]) {
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

// Greedy colouring: each element takes the lowest colour none of its K nodes already carries.
template <int K>
int colour(int count, int nodes, const std::vector<int>& ids, std::vector<int>& out, int& numColours) {
    std::vector<unsigned long long> used(nodes, 0ull);
    out.assign(count, 0);
    numColours = 0;
    for (int e = 0; e < count; e++) {
        unsigned long long taken = 0ull;
        for (int k = 0; k < K; k++) taken |= used[ids[K * e + k]];
        if (taken == ~0ull) return GTS_ERR_COLOR_OVERFLOW;
        int c = __builtin_ctzll(~taken);
        out[e] = c;
        for (int k = 0; k < K; k++) used[ids[K * e + k]] |= 1ull << c;
        if (c + 1 > numColours) numColours = c + 1;
    }
    return GTS_OK;
}

// Colour-major order, generation order inside each colour.
void colour_major(const std::vector<int>& col, int numColours, std::vector<int>& order, int32_t* offsets) {
    std::vector<int> count(numColours + 1, 0);
    for (int c : col) count[c + 1]++;
    for (int c = 0; c < numColours; c++) count[c + 1] += count[c];
    for (int c = 0; c <= numColours; c++) offsets[c] = count[c];
    order.assign(col.size(), 0);
    for (size_t e = 0; e < col.size(); e++) order[count[col[e]]++] = (int)e;
}

}  // namespace

extern "C" int gts_abi_version(void) { return GTS_ABI_VERSION; }

extern "C" int gts_struct_sizes(int* out) {
    if (!out) return GTS_ERR_BAD_ARG;
    out[0] = (int)sizeof(GtsLattice);
    out[1] = (int)sizeof(GtsTopology);
    out[2] = (int)sizeof(GtsTissueView);
    out[3] = (int)sizeof(GtsStepParams);
    out[4] = (int)sizeof(GtsToolState);
    out[5] = (int)sizeof(GtsCutTool);
    out[6] = (int)sizeof(GtsBatchInstance);
    out[7] = (int)sizeof(GtsBatchView);
    out[8] = (int)sizeof(GtsAdjointView);
    out[9] = (int)sizeof(GtsGraspParams);
    out[10] = (int)sizeof(GtsGraspState);
    out[11] = (int)sizeof(GtsSurfaceView);
    return GTS_OK;
}

extern "C" int gts_build_topology(const GtsLattice* lat, GtsTopology* out) {
    if (!lat || !out || !lat->rest || !lat->active || !lat->nbr || !lat->rest_nbr) return GTS_ERR_BAD_ARG;
    if (lat->vx <= 0 || lat->vy <= 0 || lat->vz <= 0) return GTS_ERR_BAD_ARG;
    const int nx = lat->vx + 1, ny = lat->vy + 1, nz = lat->vz + 1;
    const long long Nl = (long long)nx * ny * nz;
    if (Nl >= (1ll << 28)) return GTS_ERR_BAD_ARG;  // node ids share 32 bits with 4 flag bits
    const int N = (int)Nl;
    auto cid = [&](int x, int y, int z) { return x + nx * (y + ny * z); };

    // ---- edges
    std::vector<int> ids, axisOf;
    ids.reserve((size_t)6 * N);
    float axisLen[3] = {-1.f, -1.f, -1.f};
    for (int p = 0; p < N; p++)
        for (int axis = 0; axis < 3; axis++) {
            int n = lat->nbr[6 * p + 2 * axis + 1];
            if (n < 0) continue;
            if (n != p + (axis == 0 ? 1 : (axis == 1 ? nx : nx * ny))) return GTS_ERR_BAD_ARG;
            const float* r = lat->rest_nbr + 3 * (6 * (size_t)p + 2 * axis + 1);
            float L0 = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            if (axisLen[axis] < 0.f) axisLen[axis] = L0;
            else if (L0 != axisLen[axis]) return GTS_ERR_IRREGULAR_REST;  // one rest length per axis on a lattice
            ids.push_back(p);
            ids.push_back(n);
            axisOf.push_back(axis);
        }
    const int E = (int)axisOf.size();
    std::vector<int> ecol;
    int enc = 0;
    int rc = colour<2>(E, N, ids, ecol, enc);
    if (rc) return rc;
    std::vector<int> eorder;
    colour_major(ecol, enc, eorder, out->edge_colour_offsets);
    for (int k = 0; k < E; k++) {
        int e = eorder[k];
        out->edge_desc[k] = ((uint32_t)ids[2 * e] << 4) | ((uint32_t)axisOf[e] << 1);
        out->edge_gen_index[k] = e;
    }
    // spacing along each axis from the rest lattice, in double (xpbd_solver.cu:1646-1648)
    double spacing[3] = {((double)lat->rest[3 * cid(lat->vx, 0, 0)] - (double)lat->rest[0]) / lat->vx,
                         ((double)lat->rest[3 * cid(0, lat->vy, 0) + 1] - (double)lat->rest[1]) / lat->vy,
                         ((double)lat->rest[3 * cid(0, 0, lat->vz) + 2] - (double)lat->rest[2]) / lat->vz};
    for (int a = 0; a < 3; a++) out->axis_rest_length[a] = axisLen[a] >= 0.f ? axisLen[a] : (float)spacing[a];

    // ---- Kuhn tets
    double tB[6][3][3];
    for (int ti = 0; ti < 6; ti++) {
        double Dm[3][3];
        for (int k = 1; k < 4; k++)
            for (int a = 0; a < 3; a++) Dm[a][k - 1] = ((KUHN[ti][k] >> a) & 1) * spacing[a];
        if (!invert3(Dm, tB[ti])) return GTS_ERR_IRREGULAR_REST;
        double det = Dm[0][0] * (Dm[1][1] * Dm[2][2] - Dm[1][2] * Dm[2][1]) -
                     Dm[0][1] * (Dm[1][0] * Dm[2][2] - Dm[1][2] * Dm[2][0]) +
                     Dm[0][2] * (Dm[1][0] * Dm[2][1] - Dm[1][1] * Dm[2][0]);
        out->type_volume[ti] = std::fabs(det) / 6.0;
        std::memcpy(out->type_inv_rest[ti], tB[ti], sizeof(tB[ti]));
    }
    std::vector<int> tids, bases, types;
    for (int z = 0; z < lat->vz; z++)
        for (int y = 0; y < lat->vy; y++)
            for (int x = 0; x < lat->vx; x++) {
                bool all = true;
                for (int c = 0; c < 8 && all; c++)
                    all = lat->active[cid(x + (c & 1), y + ((c >> 1) & 1), z + ((c >> 2) & 1))] != 0;
                if (!all) continue;
                for (int ti = 0; ti < 6; ti++) {
                    int g[4];
                    for (int k = 0; k < 4; k++) {
                        int c = KUHN[ti][k];
                        g[k] = cid(x + (c & 1), y + ((c >> 1) & 1), z + ((c >> 2) & 1));
                    }
                    double Dm[3][3], B[3][3];
                    for (int k = 0; k < 3; k++)
                        for (int a = 0; a < 3; a++)
                            Dm[a][k] = (double)lat->rest[3 * g[k + 1] + a] - (double)lat->rest[3 * g[0] + a];
                    if (!invert3(Dm, B)) continue;
                    double scale = 0, dev = 0;
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) {
                            scale = std::fmax(scale, std::fabs(tB[ti][a][b]));
                            dev = std::fmax(dev, std::fabs(B[a][b] - tB[ti][a][b]));
                        }
                    if (dev > 1e-4 * scale) return GTS_ERR_IRREGULAR_REST;
                    for (int k = 0; k < 4; k++) tids.push_back(g[k]);
                    bases.push_back(g[0]);
                    types.push_back(ti);
                }
            }
    const int T = (int)types.size();
    std::vector<int> tcol;
    int tnc = 0;
    rc = colour<4>(T, N, tids, tcol, tnc);
    if (rc) return rc;
    std::vector<int> torder;
    colour_major(tcol, tnc, torder, out->tet_colour_offsets);
    for (int k = 0; k < T; k++) {
        int t = torder[k];
        out->tet_desc[k] = ((uint32_t)bases[t] << 4) | ((uint32_t)types[t] << 1);
        out->tet_gen_index[k] = t;
    }
    out->n_edges = E;
    out->n_tets = T;
    out->n_edge_colours = enc;
    out->n_tet_colours = tnc;

    // ---- bend elements: the pairs of a node's live links, in the order of the lattice's bend records
    out->n_bends = 0;
    out->n_bend_colours = 0;
    if (!out->bend_desc) return GTS_OK;
    if (!out->bend_gen_index || !out->bend_colour_offsets) return GTS_ERR_BAD_ARG;
    std::vector<int> bids, bdesc;
    for (int p = 0; p < N; p++)
        for (int q = 0; q < 15; q++) {
            const int sa = bend_slot_a(q), sb = bend_slot_b(q);
            const int j = lat->nbr[6 * p + sa], k = lat->nbr[6 * p + sb];
            if (j < 0 || k < 0) continue;
            // the kernels find the arms on the lattice: slot s links p to p +- the stride of axis s / 2
            const int strideA = sa < 2 ? 1 : (sa < 4 ? nx : nx * ny), strideB = sb < 2 ? 1 : (sb < 4 ? nx : nx * ny);
            if (j != p + ((sa & 1) ? strideA : -strideA) || k != p + ((sb & 1) ? strideB : -strideB))
                return GTS_ERR_BAD_ARG;
            bids.push_back(p);
            bids.push_back(j);
            bids.push_back(k);
            bdesc.push_back((p << 4) | q);
        }
    const int B = (int)bdesc.size();
    std::vector<int> bcol;
    int bnc = 0;
    rc = colour<3>(B, N, bids, bcol, bnc);
    if (rc) return rc;
    std::vector<int> border;
    colour_major(bcol, bnc, border, out->bend_colour_offsets);
    for (int k = 0; k < B; k++) {
        out->bend_desc[k] = (uint32_t)bdesc[border[k]];
        out->bend_gen_index[k] = border[k];
    }
    out->n_bends = B;
    out->n_bend_colours = bnc;
    return GTS_OK;
}

extern "C" long long gts_scratch_floats(int n_nodes, int n_edges, int n_tets, int n_grid_edges) {
    auto blocks = [](long long n) { return (n + 255) / 256 + 1; };
    return 8ll * n_grid_edges + blocks(n_tets) + blocks(n_edges) + blocks(n_nodes) + blocks(n_grid_edges) + 64;
}
