/* gts_xpbd.h — C ABI of the MI355X (gfx950) XPBD tissue substep library, libgts_xpbd.so.
 *
 * The library computes; it allocates no device memory for a tissue. The caller owns every device buffer (the Python
 * host side allocates them as PyTorch-ROCm tensors) and passes raw pointers and sizes in the descriptor structs
 * below. All functions return int status: 0 = success, the reference's bands otherwise
 * (xpbd/xpbd_solver.h:25-33 for -1000.., physics.cu:219-228 for -2000.., cut.cu:1109-1117 for -3000..), and
 * -(int)hipError_t for a HIP runtime error. Every launch goes to the stream the caller passes.
 *
 * Reference interfaces replaced (paths relative to endoscopic_tissue/, native = native/vendor):
 *   gts_build_topology       xpbd_bind edge derivation + xpbd_init colouring + xpbd_build_tets
 *                            (native/cuda/xpbd/xpbd_solver.cu:1204-1279, :1412-1577, :1600-1848), reached from
 *                            LCS_InitPhysics (native/src/plugin_api.cpp:105-115)
 *   gts_set_gravity          LCS_SetGravity (plugin_api.cpp:103) -> k_SetGravity (physics.cu:317-329)
 *   gts_step                 LCS_Step (plugin_api.cpp:125-134) -> physics_step (physics.cu:331-394)
 *   gts_detect_cut           LCS_SetTool + LCS_DetectCut (plugin_api.cpp:145-155) -> cut_detect (cut.cu:1235-1254)
 *   gts_detect_cut_ccd       the CCD detection of cut_ribbon_tick (cut.cu:486-592, :1303-1431)
 *   gts_refresh_liveness     k_TopologyStale + k_RefreshLiveness (solver/material/unified_vertex.cuh:859-861,
 *                            xpbd_solver.cu:429-479)
 *   gts_compact_topology     no reference counterpart: stable remap of the live tets / edges / nodes
 *   gts_batched_*            no reference counterpart (the reference is single-session): many instances, one launch
 *   gts_adjoint_*            no reference counterpart: reverse of the substep for compliance gradients
 * Readbacks (LCS_GetCornerPos, LCS_GetNbrIdx, LCS_GetTopologyRevision, LCS_GetCutEdgeFractions) are plain copies of
 * the caller's own buffers and need no entry point.
 */
#ifndef GTS_XPBD_H
#define GTS_XPBD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GTS_ABI_VERSION 2

#define GTS_OK 0
#define GTS_ERR_NOT_READY (-1000)
#define GTS_ERR_BAD_DT (-1002)
#define GTS_ERR_COLOR_OVERFLOW (-1003)
#define GTS_ERR_BAD_ARG (-1004)
#define GTS_ERR_IRREGULAR_REST (-1005)
#define GTS_ERR_CAPACITY (-3004)

#define GTS_MAX_COLOURS 64

/* Tet constraint forms of k_solve_tet (xpbd_solver.cu:749-829). */
#define GTS_TET_VOLUME 0     /* hydrostatic row only, C = J - 1 */
#define GTS_TET_NEOHOOKEAN 1 /* coupled deviatoric / hydrostatic block, C_H = J - 1 */
#define GTS_TET_GUARDED 2    /* coupled block, C_H = log J + min(0,(J-.05)/.05), per-update trust region */

/* Host lattice of one tissue (TissueGrid arrays, endoscopic_tissue/grid.py:41-61). */
typedef struct GtsLattice {
    int vx, vy, vz;         /* voxel dims; nodes (vx+1)(vy+1)(vz+1), id x + nx*(y + ny*z) */
    const float* rest;      /* [3N] rest positions */
    const int* active;      /* [N] */
    const int* nbr;         /* [6N] neighbour slots -x,+x,-y,+y,-z,+z; -1 = no link */
    const float* rest_nbr;  /* [18N] rest neighbour vectors */
} GtsLattice;

/* Topology in the device's storage order, written into caller-allocated host arrays.
 * Edge descriptor: node i << 4 | axis << 1 (bits 0 and 3, liveness and tet support, are set on the device).
 * Tet descriptor: corner-0 node << 4 | Kuhn type << 1 (bit 0, liveness, is set on the device).
 * Both lists are colour-major; within a colour, in generation order. */
typedef struct GtsTopology {
    uint32_t* edge_desc;        /* capacity 3N */
    int32_t* edge_gen_index;    /* colour slot -> generation index (p ascending, axis 0..2) */
    int32_t* edge_colour_offsets; /* capacity GTS_MAX_COLOURS + 1 */
    uint32_t* tet_desc;         /* capacity 6V */
    int32_t* tet_gen_index;     /* colour slot -> generation index (cells z,y,x major, Kuhn type minor) */
    int32_t* tet_colour_offsets; /* capacity GTS_MAX_COLOURS + 1 */
    double type_inv_rest[6][3][3]; /* Dm^-1 of each Kuhn type, from the rest lattice spacing */
    double type_volume[6];
    float axis_rest_length[3];  /* L0 of the lattice edges along x, y, z */
    int n_edges, n_tets, n_edge_colours, n_tet_colours;
    /* Bend elements (xpbd_bind's live bend pairs, xpbd_solver.cu:1466-1478, coloured by xpbd_set_bending's greedy
     * 3-node colouring, :1344-1356): descriptor apex node i << 4 | pair p (0..14, the slot pair of grid.py:257-260);
     * the arms are the lattice neighbours in the pair's two slots. Null bend_desc: no bend elements are built. */
    uint32_t* bend_desc;          /* capacity 15N */
    int32_t* bend_gen_index;      /* colour slot -> generation index (apex ascending, pair 0..14) */
    int32_t* bend_colour_offsets; /* capacity GTS_MAX_COLOURS + 1 */
    int n_bends, n_bend_colours;
} GtsTopology;

/* Device view of one tissue. Pointers are device pointers unless marked host. */
typedef struct GtsTissueView {
    int n_nodes, n_edges, n_tets, n_grid_edges;
    int vx, vy, vz;
    int n_edge_colours, n_tet_colours;
    const int32_t* edge_colour_offsets; /* host, [n_edge_colours + 1] */
    const int32_t* tet_colour_offsets;  /* host, [n_tet_colours + 1] */
    float inv_rest[6][9];               /* Dm^-1 rows per Kuhn type, float */
    float axis_rest_length[3];
    float voxel_length;                 /* cut AABB pad and rest edge length of cut points */
    /* node state: xyz + inverse mass in .w for pos; xyz for prev, vel, force */
    float* pos;
    float* prev;
    float* vel;
    float* force;
    const float* mass;
    const int32_t* active;
    const int32_t* pinned;
    const int32_t* inside;
    int32_t* nbr;          /* [6N] live neighbour table */
    uint32_t* bend_alive;  /* [N] bit p = bend pair p of the node is alive (grid.py:257-268) */
    uint32_t* edge_desc;   /* [E] */
    float* edge_lambda;    /* [E] */
    int32_t* slot_support; /* [6N] live-tet support of each owner slot (fallback edges) */
    uint32_t* tet_desc;    /* [T] */
    const float* tet_k_dev;   /* [T] mu*V (or 1/compliance) */
    const float* tet_k_hyd;   /* [T] kappa*V (or 1/compliance) */
    float* tet_alpha;         /* [2T] per-frame alpha~: dev then hyd */
    float* tet_lambda;        /* [2T] dev then hyd */
    /* cut */
    const int32_t* grid_edges;   /* [4*Eg] axis, x, y, z */
    const uint32_t* occupied;    /* [V] */
    uint32_t* cut_mask;          /* [V] */
    float* edge_cut_t;           /* [6N] */
    const float* edge_skin_t;    /* [6N] */
    const float* rest;           /* [3N] corner rest positions (recon_corner_rest, dc_recon.cu:671): rest edge of cut
                                  * points and tear-law scale */
    float* rot;                  /* [9N] particle frames, rows of R per node (RotGpu, cut.cuh:67): the material frame of
                                  * the cut points' gap and normal */
    float* rot_prev;             /* [9N] the frames before the last update (the update's temporal blend) */
    float* prev_corner;          /* [3N] the CCD ribbon tick's corner snapshot (d_prevCorner, cut.cu:1303-1431) */
    float* cut_points;           /* [8*capacity] CutPointGpu records (cut.cuh:37-46) */
    int cut_point_capacity;
    uint32_t* counters;          /* [8] 0 revision, 1 cut point counter, 2 refreshed revision, 3 stale flag,
                                  * 4 edges marked by the tear law, 5 the cut point counter before a ribbon tick */
    float* scratch;              /* [4*Eg + 2*nblocks + 64] detection records and scan partials */
    float* reaction;             /* [4] force of the tissue on the tools, contacts */
    float* step_consts;          /* [step_const_capacity * GTS_STEP_CONST_FLOATS] per-substep tool shapes */
    int step_const_capacity;     /* substeps the buffer holds */
    int use_graph;               /* 1: capture the step into a hipGraph once and replay it */
    int colour_block;            /* threads per workgroup of the edge / tet colour kernels: 32..1024, multiple of 32;
                                  * 0 = 64 */
    /* tissue guard (GtsStepParams.guard_passes > 0; xpbd_set_tissue_guard, xpbd_solver.cu:1895-1915); null
     * otherwise. The substep's start positions are `prev` (k_predict writes them; nothing on this path rewrites
     * them before the safeguard reads them, so they equal the reference's d_feasibleStart copy). */
    float* fiber_lambda;             /* [6T] tensile reinforcement multipliers, per tet slot */
    float* feasible_target;          /* [4N] candidate positions of the nodes the feasible filter moves */
    int32_t* feasible_needed;        /* [N + 1] nodes needing the filter, then the candidate-unsafe flag */
    const int32_t* incident_offsets; /* [N + 1] CSR offsets of each node's tets */
    const int32_t* incident_tets;    /* [4T] tet slots incident to each node (any order) */
    /* the tear law's grasp-ring exemption (cut.cu:68-83, :647-673); null: no exemption */
    int32_t* grasp_ring;             /* [2N] ring flags (the first N), then the dilation's ping-pong half; zeroed */
    int32_t* edge_exempt;            /* [2 Eg] per cut edge: was-exempt flag, then release-decay ticks; zeroed */
    const int32_t* nbr_init;         /* [6N] the neighbour table as built (cut-rim carve-out of the ring) */
    int xcd_order;                   /* 1: the colour kernels give each XCD one contiguous range of a colour (the
                                      * workgroups of an XCD meet the same lattice slab in every colour); 0: the
                                      * plain workgroup order. Results are the same either way. */
    /* bend elements (GtsTopology.bend_desc), solved when GtsStepParams.bend_kb > 0; their liveness is the apex's
     * bend_alive bit of the pair (the refresh's bendActive, xpbd_solver.cu:446) */
    int n_bends, n_bend_colours;
    const int32_t* bend_colour_offsets; /* host, [n_bend_colours + 1] */
    const uint32_t* bend_desc;          /* [n_bends] */
    float* bend_lambda;                 /* [n_bends] */
    /* Tiled tet and edge sweeps (no reference counterpart; the results are the per-colour sweeps' bit for bit): with
     * tile_groups > 0 an undamped, unguarded step without the legacy grasp or a ribbon tick runs its tet colours in
     * launches of up to tile_kmax consecutive colours each (the groups of kind 0), one workgroup per lattice region,
     * positions in an LDS box of tile_box^3 nodes around the region (the plan: genesis_tissue_simulation_amd/tiling.py);
     * where the plan also holds groups of kind 1, each undamped edge sweep without the fallback compliance runs as
     * launches of up to tile_kmax - 1 edge colours over the same regions. Each launch reads the positions from one
     * buffer and writes them to the other (pos, pos_alt alternate; each kind has an even number of groups). Every input
     * of a region sits at an address it computes from its index: its entries and writes at fixed strides, its box
     * origin from its dense id (a kernel argument), and per entry the element's inputs (tile_inputs, gathered at the
     * start of each step) and its multipliers (tile_lambda: each region keeps the copies of the elements it projects;
     * the owners also write tet_lambda or edge_lambda). The plan holds for tile_tets tets and tile_edges edges in the
     * current colour lists; a view whose lists differ (a compaction) takes the per-colour launches. */
    int tile_groups;
    int tile_kmax, tile_box, tile_tets;
    const int32_t* tile_group_info;     /* host [tile_groups * 16]: first colour, colours, first region, regions, box
                                         * width, box halo, entry rows a region, first entry, write capacity, first
                                         * write, the capacity of each of 4 passes, kind (0 tets, 1 edges), 0 */
    const int32_t* tile_entries;        /* [2 * entries]: tet slot | 1 << 30 | owned << 31, corner-0 index in the LDS
                                         * box (padding rows: 0, 0); a region's rows start at first entry + region *
                                         * (entry rows), pass p's after the capacities of the passes before it */
    const int32_t* tile_writes;         /* [2 * writes]: LDS box index, node id (padding rows: node -1); write
                                         * capacity rows a region */
    float* pos_alt;                     /* [4N] the second position buffer */
    float* tile_inputs;                 /* [4 * entries] per entry: alpha_D, alpha_H, the descriptor's bits, 0 */
    float* tile_lambda;                 /* [2 * entries] per entry: its copy of the tet's multipliers */
    int tile_region, tile_rx, tile_ry;  /* cells along a region's edge; regions along x and y of the dense numbering */
    const int32_t* tile_dense_ids;      /* host [regions per launch, at most 640]: dense id rx + tile_rx (ry +
                                         * tile_ry rz) of each region taking part, in header order */
    int tile_edges;                     /* edges the plan's edge groups (kind 1, after the tet groups) hold; a view
                                         * whose edge lists differ takes the per-colour edge launches */
} GtsTissueView;

/* Ribbon ticks a released grasp-ring edge stays exempt from the tear law while still longer than 0.8 tau L
 * (TEAR_DECAY_TIMEOUT_TICKS, cut.cu:83). */
#define GTS_TEAR_DECAY_TICKS 2000

/* Floats of one substep's tool shapes in GtsTissueView.step_consts (8 shapes of 23 floats). */
#define GTS_STEP_CONST_FLOATS 184

typedef struct GtsStepParams {
    float dt;
    int substeps, iterations, tet_mode;
    float omega, edge_ks, young, beta_s, strain_damping, bulk_damping, rayleigh, damping;
    int contact_enabled;
    float mu_s, mu_d, jaw_thickness, contact_radius, jaw_gap_closed, contact_compliance;
    /* 0: the tet constraints are evaluated in the reference's direct forms (k_solve_tet, xpbd_solver.cu:749-829), so
     * results match an FP32 restatement of the reference bit for bit. 1: J - 1, log J, J^(-2/3) I1 - 3 and the
     * deviatoric gradient are formed from the displacement gradient F - I (no reference counterpart), which removes
     * their FP32 cancellation at small strain: the same constraints, closer to FP64. */
    int precise_strain;
    /* Tissue guard and material (LCS_SetTissueGuard / LCS_SetTissueMaterial, plugin_api.cpp:212-214 ->
     * xpbd_set_tissue_guard / xpbd_set_tissue_material, xpbd_solver.cu:1895-1915, :2165-2171). guard_passes > 0
     * (at most 4) runs that many tensile reinforcement sweeps (fibre energy of strength tensile_strength, 0..1000)
     * after every tet sweep and, at the end of every substep, the safeguard that keeps every live tet's J at or
     * above guard_min_volume (0..0.95); 0 turns both off. Needs the view's guard buffers. */
    float guard_min_volume;
    int guard_passes;
    float tensile_strength;
    /* Bending (PhysicsDesc kb / betaB, xpbd_begin_frame, xpbd_solver.cu:2491-2499): with bend_kb > 0 and bend
     * elements in the view, every iteration solves the bend colours after the edge colours (k_solve_bending,
     * :534-573, :3590-3596) with compliance 1 / (kb h^2) and Kelvin-Voigt factor alpha~ beta_b h. */
    float bend_kb, beta_b;
} GtsStepParams;

/* XpbdToolState (xpbd_solver.h:223-229) plus a blade capsule. */
typedef struct GtsToolState {
    float shaftA[3], shaftB[3], shaftR;
    float jawHinge[3], jawAxis[3], jawMid[3], jawNormal[3];
    float jawHalfWidth, jawLen, jawGap, grasp01;
    float bladeA[3], bladeB[3], bladeR;
} GtsToolState;

/* XpbdGraspParams (xpbd_solver.h:250-289) in the field order of native_backend.py's GraspParams: the legacy contact
 * model's capture / release thresholds, captured-ring friction, core pins, break force and CCD gate. */
typedef struct GtsGraspParams {
    float capture, release;     /* grasp01 above which the jaws capture, below which they release */
    float closed_gap;           /* floor of the jaw gap */
    float core_fraction;        /* normalised-ellipse core fraction */
    float friction;             /* Coulomb mu of the captured ring */
    float contact_compliance;   /* slab and CCD gate compliance, scaled by 1/h^2 */
    float skin;                 /* capture band widening */
    float ccd_velocity;         /* approach speed above which the CCD gate fires */
    float ccd_displacement;     /* per-substep cap of a CCD correction */
    float ccd_omega;            /* CCD gate under-relaxation */
    float grasp_omega;          /* under-relaxation of the ring and the core pins */
    float core_compliance;      /* core pin compliance, scaled by 1/h^2 */
    float break_force;          /* core force (one-second EMA of |lambda| / h^2) above which a core corner lets go */
    float grasp_displacement;   /* per-substep cap of a grasp correction */
    int exempt_rings;           /* tear-law ring dilation; no ring on this path */
} GtsGraspParams;

/* Legacy grasp state of one tissue (xpbd_solver.cu:250-275). Device arrays of N entries unless marked; the caller
 * allocates them and zeroes them before the first use. */
typedef struct GtsGraspState {
    int32_t* grasped;           /* captured corners */
    int32_t* core;              /* core corners, pinned to their jaw-local anchor */
    float* core_offset;         /* [3N] jaw-local (u, n, w) of each captured corner at capture */
    float* anchor;              /* [2N] Coulomb anchor (u, w) of each ring corner */
    float* slab_lambda;         /* [2N] slab multiplier of this substep, then of the previous substep */
    float* core_lambda;         /* [4N] vector core multiplier, float4 rows */
    float* core_force_ema;      /* break-check EMA */
    float* gate_lambda;         /* [2N] CCD gate multipliers: jaw slab, then shaft */
    float* capture_scratch;     /* [4N] candidate flags, then their (u, v, n) */
    int grasping;               /* host latch, 1 while grasping; gts_grasp_set_tool maintains it */
} GtsGraspState;

/* CutToolDesc (cut.h:30-38): swept quad T1 = (Sprev, Eprev, E), T2 = (Sprev, E, S). */
typedef struct GtsCutTool {
    float t1v0[3], t1v1[3], t1v2[3];
    float t2v0[3], t2v1[3], t2v2[3];
    float ncut[3];
    float d;
    int valid;
    float aabbMin[3], aabbMax[3];
} GtsCutTool;

int gts_abi_version(void);
/* Sizes of the structs above as this library sees them (ABI guard like LCS_PhysicsAbiVersion,
 * plugin_api.cpp:102). out[0..11] = GtsLattice, GtsTopology, GtsTissueView, GtsStepParams, GtsToolState, GtsCutTool,
 * GtsBatchInstance, GtsBatchView, GtsAdjointView, GtsGraspParams, GtsGraspState, GtsSurfaceView. */
int gts_struct_sizes(int* out);

/* Host: derive edges, Kuhn tets, bend elements and their greedy colourings from the lattice. Needs no GPU. */
int gts_build_topology(const GtsLattice* lattice, GtsTopology* out);

/* Host: J^(-2/3) as the tet kernel evaluates it (own cube-root sequence), for checks. */
float gts_iso_pow(float J);
/* Host: log x as the guarded tet kernel evaluates it (own atanh-series sequence), for checks. */
float gts_log(float x);
/* Host: arccos as the bend kernel evaluates it (own arcsin-series sequence), for checks. */
float gts_acos(float c);
/* Host: floats of device scratch a tissue view needs (detection records + scan partials). */
long long gts_scratch_floats(int n_nodes, int n_edges, int n_tets, int n_grid_edges);

/* Device: initialise the per-node inverse mass (.w of pos) and the bend masks; mark all liveness stale. */
int gts_init_state(const GtsTissueView* v, void* stream);
int gts_set_gravity(const GtsTissueView* v, float gx, float gy, float gz, void* stream);
/* Device: re-derive edge / tet liveness and the fallback tet support when the topology revision moved. */
int gts_refresh_liveness(const GtsTissueView* v, void* stream);
/* Device: one outer step of p->substeps substeps, then the particle frames (LCS_Step, plugin_api.cpp:125-134). The
 * tool pose moves from `from` to `to` over the step. */
int gts_step(const GtsTissueView* v, const GtsStepParams* p, const GtsToolState* from, const GtsToolState* to,
             int tool_armed, int blade_armed, void* stream);
/* Host + device: the legacy contact model's part of xpbd_set_tool (xpbd_solver.cu:2318-2341, :2372-2470) for the
 * pose `tool`. grasp01 rising past gp->capture captures: live unpinned corners inside the jaw box widened by the skin
 * are marked on the device (k_MarkGraspCandidates), read back (the stream is synchronised), the normalised-ellipse core
 * is chosen on the host and the grasp state is uploaded. grasp01 falling below gp->release releases: the grasped
 * corners' velocity is scaled by 0.1 and the sets are cleared. *event = 1 on capture, 2 on release, else 0. After a
 * capture the next step must start at `tool` (the reference sets g_toolFrom = g_tool). Then, with v->grasp_ring,
 * the tear law's ring is recomputed from the grasped set as physics_set_tool does (physics.cu:436-447 ->
 * cut_recompute_grasp_ring, cut.cu:1513-1528): gp->exempt_rings hops of dilation over live links, minus the corners
 * that have lost a link since v->nbr_init. */
int gts_grasp_set_tool(const GtsTissueView* v, GtsGraspState* g, const GtsGraspParams* gp, const GtsToolState* tool,
                       int* event, void* stream);
/* Device: one outer step under the legacy contact model (xpbd_substep with g_frictionContact false,
 * xpbd_solver.cu:3460-3721): while grasping, K = 2 material passes, each followed by the slab ring with Coulomb
 * anchors (k_SolveSlabCoulomb) and the core pins (k_SolveCorePin), then the core break check (k_CoreBreakCheck);
 * with the tool armed, the CCD gate against the shaft capsule and the jaw slab (k_SolveCCDContact). The jaw frame
 * follows the pose from `from` to `to` over the substeps. The contact fields of *p are not used, nor the reaction. */
int gts_step_legacy(const GtsTissueView* v, const GtsStepParams* p, const GtsGraspParams* gp, const GtsGraspState* g,
                    const GtsToolState* from, const GtsToolState* to, int tool_armed, void* stream);
/* Device: gts_step_legacy while LCS_SetTool's cut tool is set: the CCD ribbon tick after every substep as in
 * gts_step_cutting (physics_step runs it whatever the contact model, physics.cu:352-377), with the tear law's
 * grasp-ring exemption when tear_tau >= 1. */
int gts_step_legacy_cutting(const GtsTissueView* v, const GtsStepParams* p, const GtsGraspParams* gp,
                            const GtsGraspState* g, const GtsToolState* from, const GtsToolState* to, int tool_armed,
                            const GtsCutTool* tool, const float* emit_ncut, float tear_tau, void* stream);
/* Device: LCS_Step while LCS_SetTool's cut tool is set (physics_step, physics.cu:331-394): gts_step with the CCD
 * ribbon tick after every substep (cut_ribbon_tick without the pressure blade, cut.cu:1303-1431): k_DetectCutCCD of the
 * edges moving from v->prev_corner to the current positions against the tool's current rod S = t2v2, E = t1v2 (emit_ncut,
 * the last valid tool normal, where the hit has no normal of its own), then the tear law when tear_tau >= 1, one
 * k_SeverLinks and the snapshot rolled to the current positions; a tick that emitted cut points, but the last, is
 * followed by a frame update (physics.cu:366-380). Seed v->prev_corner with gts_snapshot_corners when the tool is set. */
int gts_step_cutting(const GtsTissueView* v, const GtsStepParams* p, const GtsToolState* from, const GtsToolState* to,
                     int tool_armed, int blade_armed, const GtsCutTool* tool, const float* emit_ncut, float tear_tau,
                     void* stream);
/* Device: the corner snapshot cut_set_tool takes when a cut tool is first set (cut.cu:1213-1220). */
int gts_snapshot_corners(const GtsTissueView* v, void* stream);
/* Device: physics_compute_rotations (physics.cu:399-411) as LCS_Finalize runs it after a cut (plugin_api.cpp:159-166):
 * the current frames become the previous ones and k_ComputeParticleRot (physics.cu:109-205) derives new ones from the
 * current positions and links. gts_step and gts_step_legacy end with the same update. */
int gts_compute_rotations(const GtsTissueView* v, void* stream);
/* Device: swept-quad cut of the current positions, then sever the cut links and refresh liveness. */
int gts_detect_cut(const GtsTissueView* v, const GtsCutTool* tool, void* stream);
/* Device: gts_detect_cut without the caller's workspace (no reference counterpart: for callers that cannot allocate
 * device memory -- code under graph capture, tight serving loops, harnesses that forbid hipMalloc). The detection
 * records and scan partials live in a library-owned static device array of GTS_STATIC_SCRATCH_FLOATS floats (32 MiB,
 * held on every device the library runs on until the process ends); v->scratch is not used. The call allocates
 * nothing. It needs gts_detect_cut_static_floats(n_grid_edges) <= GTS_STATIC_SCRATCH_FLOATS (about 1.04M cut edges)
 * and returns GTS_ERR_CAPACITY otherwise. Calls on one device share the array: calls that overlap on a device
 * race. Results equal gts_detect_cut's. */
#define GTS_STATIC_SCRATCH_FLOATS (8u << 20)
long long gts_detect_cut_static_floats(int n_grid_edges);
int gts_detect_cut_static(const GtsTissueView* v, const GtsCutTool* tool, void* stream);
/* Device: CCD of the edges moving from prev_positions ([3N] device) to the current positions against blade S->E. */
int gts_detect_cut_ccd(const GtsTissueView* v, const float* S, const float* E, const float* emit_ncut, float tool_d,
                       const float* prev_positions, void* stream);
/* Device: tear law (k_TearOverstretch, cut.cu:607-681, as cut_ribbon_tick runs it before k_SeverLinks,
 * cut.cu:1397-1405): every intact link between active corners that is not yet cut, has an occupied stencil voxel and
 * is longer than tau times its rest length is cut at its midpoint (cut points with tool gap tool_d); then the links are
 * severed and liveness refreshed. tau < 1 or NaN: no-op (cut.cu:1199-1202). Needs v->rest. Call it after the tick's
 * detection. With v->grasp_ring and v->edge_exempt, links with a ring corner are exempt, and released ones decay
 * (cut.cu:647-673). */
int gts_tear_overstretch(const GtsTissueView* v, float tau, float tool_d, void* stream);
/* Device: stable remap of the live topology. tet_remap [T] (colour order), edge_remap [E], node_remap [N],
 * counts [3] (live tets, edges, nodes); all device. scratch as in the view. */
int gts_compact_topology(const GtsTissueView* v, int32_t* tet_remap, int32_t* edge_remap, int32_t* node_remap,
                         int32_t* counts, int32_t* node_used, void* stream);
/* Device: the removal itself (no reference counterpart; the reference keeps dead tets in its static colouring):
 * gather the live tets (descriptor, k_dev, k_hyd) and edges (descriptor) of the view into output arrays of
 * counts[0] tets and counts[1] edges, in the order of gts_compact_topology's remaps. The caller then points the view
 * at the outputs and updates n_tets, n_edges and the colour offsets (the live count of each colour range). Node
 * storage keeps its lattice addressing, because descriptors locate their nodes on the lattice; node_remap is the
 * readback numbering of the used nodes. */
int gts_apply_compaction(const GtsTissueView* v, const int32_t* tet_remap, const int32_t* edge_remap,
                         uint32_t* tet_desc_out, float* tet_k_dev_out, float* tet_k_hyd_out, uint32_t* edge_desc_out,
                         void* stream);

/* ---- batched stepping: many independent lattice instances in pooled buffers, one persistent launch per step ---- */

/* Per-instance record (device array). Offsets index the pooled arrays; descriptors, stiffness and lambdas of an
 * instance are colour-major and instance-local (node ids relative to node_off). */
typedef struct GtsBatchInstance {
    int node_off, n_nodes, edge_off, n_edges, tet_off, n_tets;
    int nx, nxy, n_edge_colours, n_tet_colours, colour_off, reserved;
    float edge_ks;               /* edge stiffness 1/compliance of this instance */
    float rest_len[3];
    float inv_rest[6][9];
} GtsBatchInstance;

typedef struct GtsBatchView {
    int n_instances;
    const void* instances;        /* device GtsBatchInstance[n_instances] */
    const int32_t* colour_table;  /* device: per instance, n_edge_colours+1 edge then n_tet_colours+1 tet offsets */
    const int32_t* work_offsets;  /* device [n_workgroups + 1]: instances of each workgroup (gts_batched_plan) */
    const int32_t* work_items;    /* device [n_instances] */
    int n_workgroups, block_size, lds_nodes; /* launch shape; the first lds_nodes nodes of an instance are staged in LDS
                                              * (all of them when it has no more) */
    int use_graph;
    int substeps_per_launch;      /* substeps each persistent launch covers (0 = the whole step in one launch): the
                                   * colours per launch are this times the colours of a substep */
    float gravity[3];
    float* pos;                   /* [sum N] float4 rows as in GtsTissueView */
    float* prev;
    float* vel;
    const float* mass;
    const uint32_t* edge_desc;
    const uint32_t* tet_desc;
    const float* k_dev;
    const float* k_hyd;
    float* lambda_edge;           /* [sum E] */
    float* lambda_tet;            /* [2 sum T], dev then hyd per instance slice */
    float* alpha;                 /* [2 sum T] work: alpha~ of the step, dev then hyd per instance slice */
    /* particle frames: when rot is set, every step ends with the frame update of every instance
     * (physics_compute_rotations, as LCS_Step does) */
    long long total_nodes;        /* sum N */
    const int32_t* nbr;           /* [6 sum N] each instance's link table, node ids relative to its node_off */
    const int32_t* active;        /* [sum N] */
    const int32_t* pinned;        /* [sum N] */
    const int32_t* node_base;     /* [sum N] node_off of the instance each node belongs to */
    float* rot;                   /* [9 sum N] frames, rows of R */
    float* rot_prev;              /* [9 sum N] */
} GtsBatchView;

/* Host: LPT packing of instances (by cost) onto workgroups. out: offsets[n_workgroups+1] then items[n]. */
int gts_batched_plan(const double* cost, int n, int n_workgroups, int32_t* out);
/* Host: largest dynamic LDS a workgroup may allocate on the current device. */
int gts_batched_lds_bytes(void);
/* Device: one outer step of every instance (p->substeps substeps each) in persistent launches of
 * substeps_per_launch substeps (one launch when 0). */
int gts_batched_step(const GtsBatchView* v, const GtsStepParams* p, void* stream);

/* ---- adjoint of the substep for per-tet compliance gradients ----
 * No reference counterpart (the reference solver has no gradient; its differentiable PBD path is a stub,
 * genesis/engine/solvers/pbd_solver.py:784-815). The forward is the undamped Neo-Hookean substep with p->iterations
 * Gauss-Seidel iterations, a per-node force field and no contact or cutting; it runs on the pooled instances of a
 * GtsBatchView (one workgroup per instance, LPT-packed), whose lambda_edge / lambda_tet hold the multipliers within a
 * substep. Tape slots: 2 * ckpt * iterations (two halves of ckpt * iterations). */
typedef struct GtsAdjointView {
    GtsBatchView batch;           /* pooled instances and launch shape; k_dev and k_hyd are not used */
    long long total_nodes;        /* sum of n_nodes: row stride of the checkpoints */
    long long total_tets;         /* sum of n_tets: stride of the tet tape slots */
    long long total_edges;        /* sum of n_edges: stride of the edge tape slots */
    const float* force;           /* [sum N] float4 rows: node force (x, y, z, unused) */
    const float* compliance;      /* [sum T] per-tet deviatoric compliance (volume compliance / 10), colour-major */
    float* alpha;                 /* [2 sum T] alpha~_D then alpha~_H per instance slice, written by the forward */
    float* checkpoints;           /* [2 * n_checkpoints * sum N] float4: positions, then velocities, every ckpt-th */
    float* tet_tape;              /* [slots * 4 * sum T] float4: positions (inverse mass in .w) entering each tet
                                   * projection */
    float* edge_tape;             /* [slots * 2 * sum E] float4: positions entering each edge projection */
    float* lambda_tape;           /* [slots * (2 sum T + sum E)]: multipliers entering each projection (tets, then
                                   * edges); used when iterations > 1 */
    float* xbar;                  /* [sum N] float4: in dL/dx after the last substep, out dL/dx at the start */
    float* vbar;                  /* [sum N] float4: in dL/dv after the last substep, out dL/dv at the start */
    float* prevbar;               /* [sum N] float4 work */
    float* abar;                  /* [2 sum T] work: dL/d alpha~_D then dL/d alpha~_H per instance slice */
    float* lambar;                /* [2 sum T + sum E] work: multiplier adjoints (tets, then edges) */
    float* grad;                  /* [sum T] out: dL/d compliance, colour-major */
} GtsAdjointView;

/* Device: `nsub` substeps of every instance from the state in batch.pos / batch.vel, storing (x, v) before every
 * ckpt-th substep and taping the last segment for the backward; the end state is left in batch.pos / batch.vel. */
int gts_adjoint_forward(const GtsAdjointView* a, const GtsStepParams* p, int nsub, int ckpt, void* stream);
/* Device: reverse pass over the checkpoints of the last gts_adjoint_forward with the same compliance, nsub, ckpt and
 * iterations; consumes xbar / vbar and writes grad. Segments are reversed last to first on `stream`; the taped re-run
 * of each earlier segment runs on a library-owned side stream (default priority) meanwhile, ordered by events, and
 * all of it is complete when the work on `stream` is. Leaves batch.pos / batch.vel at the end of the first segment.
 * The reverse runs with min(batch.block_size, GTS_ADJOINT_BACKWARD_MAX_THREADS) threads per workgroup; a block is
 * staged in LDS when it has at most batch.lds_nodes nodes. Backward calls must not overlap one another. */
#define GTS_ADJOINT_BACKWARD_MAX_THREADS 512
int gts_adjoint_backward(const GtsAdjointView* a, const GtsStepParams* p, int nsub, int ckpt, void* stream);

/* Surface reconstruction of the deformed, cut lattice: the dual-contouring rebuild (recon_build_static,
 * dc_recon.cu:498-551) with the cut feature points and cut walls (cut_fp_chain / cut_build_triangles /
 * cut_normalize_fp, cut.cu:1434-1505) and the internal-face cleanup (surface_cleanup.cuh:35-53). What LCS_Finalize
 * rebuilds (plugin_api.cpp:161-169) and LCS_GetSurfaceCounts / LCS_GetSurface / LCS_GetSurfaceAux read
 * (plugin_api.cpp:82-99). Grid inputs are the ones LCS_Init and LCS_InitCut upload (dc_recon.h:29-36). */
typedef struct GtsSurfaceView {
    int n_voxels, n_isect, n_surface_edges;
    int tri_capacity;               /* triangles the surface holds (ReconInitDesc.triCapacity) */
    int qef_iters;                  /* feature point iterations (ReconInitDesc.qefIters) */
    int face_capacity;              /* entries of face_table: a power of two >= 2 * tri_capacity */
    const int32_t* voxel_corners;   /* [8V] node of each voxel corner, figure order (cut.cuh:90-99) */
    const int32_t* isect_offset;    /* [V] first intersection of the voxel */
    const int32_t* isect_count;     /* [V] at most 12: one crossing per voxel edge */
    const float* isect;             /* [10I] IsectGpu records (common.cuh:23-33): rest xyz, normal xyz, a, b, t, insideA */
    const int32_t* surface_edges;   /* [8Se] GridEdgeGpu records (common.cuh:36-42): axis, base xyz, insideA, pad */
    const int32_t* conn;            /* [9 * 4096] Conn4096Gpu (cut.cuh:51-56): component count, component of each corner */
    const int32_t* edge_of;         /* [3N] cut edge id of the lattice edge (axis, base node), -1 if it is none */
    const float* rest_fp;           /* [4V] feature points of the rest surface (the snapshot recon_init takes) */
    float* isect_world;             /* [3I] work: deformed intersections */
    float* fp;                      /* [4V] out: voxel feature points, w = 1 where the voxel has one */
    float* cut_fp;                  /* [32V] out: component feature points, slot 8 * voxel + component; zero-filled
                                     * before the first rebuild (only cut, occupied voxels' slots are rewritten) */
    int32_t* edge_points;           /* [2Eg] state: the cut points of each cut edge, owner A then B; fill with -1
                                     * before the first cut-aware rebuild, later rebuilds extend it */
    int32_t* candidates;            /* [3 * (2Se + 12N)] work: triangles in fixed slots before their placement */
    int32_t* item_mask;             /* [max(Se + 6N, tri_capacity)] work */
    int32_t* tri;                   /* [3 * tri_capacity] out: triangle vertex tags (bit 31 component point, bit 30 wall) */
    int32_t* tri_work;              /* [3 * tri_capacity] work */
    int32_t* face_table;            /* [4 * face_capacity] work */
    int32_t* normal_fixed;          /* [27V] work: fixed-point normal sums, voxels then component slots */
    int32_t* scan;                  /* [ceil(max(Se + 6N, tri_capacity) / 256) + 1] work */
    uint32_t* counts;               /* [4] out: 0 triangle indices, 1 indices before the cleanup, 2 indices asked for
                                     * (above 3 * tri_capacity when the triangles overflowed it) */
    float* out_pos;                 /* [9 * tri_capacity] out: the triangle soup, xyz per index */
    float* out_nrm;                 /* [9 * tri_capacity] out */
    float* out_aux;                 /* [12 * tri_capacity] out: rest anchor xyz, wall flag */
} GtsSurfaceView;

/* Device: rebuild the surface of the tissue in `v` (positions, frames, live links and cut state) into `s`.
 * cut_aware = 1 is the rebuild after LCS_Finalize (the cut buffers are live, dc_recon.cu:662-673); 0 is recon_init's
 * build before any cut state exists. have_rest = 1 advects the intact voxels' rest feature points
 * (k_AdvectIntactSurface, dc_recon.cu:477-493). Triangles are ordered deterministically: the stitch's two per surface
 * edge in edge order, then the cut walls in node order; the cleanup keeps each face's first listed representative. */
int gts_surface_build(const GtsTissueView* v, const GtsSurfaceView* s, int cut_aware, int have_rest, void* stream);

/* Device: LCS_SurfaceDepth (plugin_api.cpp:32, xpbd_surface_depth in xpbd_solver.cu:2046-2059, k_SurfaceDepth in
 * contact/surface_contact.cuh:662-674): the deepest penetration of the tool's shapes into the surface that `s` last
 * rebuilt, written to the device float `depth` (0 when none). `kinds` selects the shapes by bit: 1 the shaft capsule,
 * 2 and 4 the jaw boxes (thickness, min_gap: the step's jaw_thickness and jaw_gap_closed), 8 the blade capsule;
 * LCS_SurfaceDepth's grasper part is 7. The triangle count is read on the device; nothing is synchronised. */
/* Device: LCS_GetMaterialQuality (plugin_api.cpp:207, xpbd_get_material_quality and k_MaterialQuality in
 * xpbd_solver.cu:2122-2138): out[0] = the smallest det F over the live tets, seeded at 1 (so at most 1); out[1] = how
 * many live tets have det F <= 0 or a non-finite det F. `out` is 2 device floats; nothing is synchronised. */
int gts_material_quality(const GtsTissueView* v, float* out, void* stream);

int gts_surface_depth(const GtsSurfaceView* s, const GtsToolState* tool, int kinds, float thickness, float min_gap,
                      float* depth, void* stream);

/* Device: copy n_float4 16-byte elements (no reference counterpart): the known-size kernel against which HBM byte
 * counters are scaled, and the achievable copy bandwidth. */
int gts_stream_copy(void* dst, const void* src, long long n_float4, void* stream);

/* Timing of the dominant kernel (tet colour pass) with HIP events on its launch stream, per launch
 * (in the spirit of LCS_ProfileEnable/LCS_ProfileRead, plugin_api.cpp:21-26). */
int gts_profile_begin(void);
int gts_profile_end(double* total_ms, long long* launches, long long* units);

#ifdef __cplusplus
}
#endif

#endif /* GTS_XPBD_H */
