// Device batch of worlds (definition of the opaque rp3dg_batch handle).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/rp3d_mi355x.h"
#include "api.h"
#include "events.h"
#include "fused.h"
#include "islands.h"
#include "joints.h"
#include "narrowphase.h"
#include "rp3d_internal.h"

struct rp3dg_world;

struct rp3dg_batch {
    struct Pending { const char* name; hipEvent_t a, e; };
    struct KernelStat { double total_ms = 0.0; int64_t launches = 0; };

    std::vector<rp3dg_world*> worlds;
    std::vector<int> body_base, collider_base;
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;   // size of every allocation in allocs
    size_t bytes = 0;
    uint64_t uid = 0;                  // unique over the process (snapshots name their batch by it)
    hipStream_t stream = nullptr;
    int nb = 0, nc = 0, nworlds = 0, world_bits = 0;
    rp3dg::Settings settings{};
    rp3dg::DevBodies B{};
    rp3dg::DevColliders C{};
    rp3dg::DevPairs Pa{}, Pb{};
    rp3dg::DevSolver S{};
    rp3dg::DevJoints J{};
    rp3dg::DevIslands isl{};           // allocated when the worlds have sleeping enabled
    void* ray_buf = nullptr;           // raycast staging (grown on demand, freed with the batch)
    size_t ray_cap = 0;
    int4* ev_buf = nullptr;            // contact events of the last step (allocated when first enabled)
    int ev_cap = 0;
    bool events_on = false;
    int* jt_ranges = nullptr;          // [2*slot] colour ranges of the joint rows, slots in solve order
    int2* jt_wr = nullptr;             // fused solver: [w * jt_slots + slot] joint rows of world w
    int jt_slots = 0;
    std::vector<int> joint_order;      // global joint index of each joint row
    bool cur_is_a = true;
    int pair_cap = 0;
    int* pair_count = nullptr;         // device: sizes of the two pair tables (a, b)
    unsigned long long* total_points = nullptr;   // device: contact points summed over all steps
    int64_t steps_done = 0;

    // solver layout: fused = one workgroup per world for the whole solver phase, (world, colour)-major
    // rows; otherwise one launch per colour over colour-major rows.
    bool fused = false;
    int fused_block = 64;              // threads per world of the position kernel (and of the single-class solve)
    int color_block = 256;             // per-colour layout: threads per block of the colour kernels
    // fused layout: worlds grouped into size classes (bodies in (L/2, L], L = 64, 128, ... 4096), largest
    // first; each class is one launch with its own workgroup size and LDS, the classes run concurrently
    struct WorldClass { int first, count, max_bodies, block, lds_bodies; };
    std::vector<WorldClass> wclasses;
    int block_div = 2;                 // class workgroup = clamp(L / block_div, 64, 256) (autotuned)
    std::vector<hipStream_t> side_streams;
    std::vector<hipEvent_t> fork_events;
    std::vector<std::pair<int, double>> tune_table;   // autotuner trials: (shape, ms per step)
    int user_block = 0;                // threads per world requested by the caller (0 = per class, autotuned)
    int nseg = 0;                      // colour segments: kMaxColors (grid) or nworlds * kMaxColors (fused)
    int* seg_cnt = nullptr;
    int* seg_start = nullptr;
    int *d_wbody = nullptr, *d_wcoll = nullptr, *d_block_world = nullptr;
    // world groups of the fused solver: consecutive worlds sharing one workgroup (d_block_world holds
    // the group of each workgroup, largest groups first)
    int ngroups = 0, group_bodies = 256;
    std::vector<int> group_world;        // [ngroups + 1] first world of each group
    int* d_block_static = nullptr;       // the size-sorted group order (d_block_world is re-ranked per step in
                                         // residual mode: longest expected group first)
    int2* d_class_ranges = nullptr;      // (first, count) of every world class in that order
    bool order_by_sweeps = true;
    std::vector<int> world_group;        // [nworlds]
    int *d_gworld = nullptr, *d_world_group = nullptr;
    int* d_wmstart = nullptr;            // [nworlds + 1] first manifold of each world (fused colouring)
    int max_group_bodies = 0;
    bool world_scan = false;             // broadphase by per-world box scans in LDS instead of the tree
    int max_group_colliders = 0;
    float4 *res_pv = nullptr, *res_pw = nullptr;
    int* world_iters = nullptr;
    float* world_res = nullptr;
    // learned warm start and (feature, converged impulse) sample log
    rp3dg_mlp* mlp = nullptr;
    int warm_mode = 0;                 // RP3DG_WARM_START_*
    float* feat_buf = nullptr;
    int* row_slot = nullptr;
    int feat_cap = 0;
    float *log_feats = nullptr, *log_targets = nullptr;
    unsigned long long* log_count = nullptr;
    long long log_cap = 0;
    float res_thr = 0.0f;              // residual mode threshold (0: fixed iterations)
    int res_max_iters = 0;
    int max_world_bodies = 0;
    int max_lds_per_block = 65536;     // hipDeviceAttributeMaxSharedMemoryPerBlock of the batch's device
    int max_lds_per_cu = 65536;        // hipDeviceAttributeMaxSharedMemoryPerMultiprocessor

    // RigidBody calls between steps (setTransform, setLinearVelocity, apply*Force, setIsSleeping, ...),
    // queued in call order and applied by one device thread before the next step or read-back
    std::vector<rp3dg::ApiOp> api_ops;
    rp3dg::ApiOp* d_api_ops = nullptr;
    size_t api_cap = 0;
    bool api_sleep_op = false;         // a queued call puts a body to sleep (its pairs restart their records)
    int *d_coll_offs = nullptr, *d_coll_list = nullptr;   // colliders of each body (CSR)
    std::vector<int> body_type;        // RP3DG_STATIC / KINEMATIC / DYNAMIC of every body
    std::vector<int> coll_kind;        // CK_* bits of every collider
    int flush_api();

    // world queries (testOverlap / testCollision): their contacts go to these arrays, not to the pair
    // table, whose contact caches keep the last step's; exits of the pairs their broadphase removes wait in
    // ev_pend for the next step
    int* q_ncp = nullptr;
    float4 *q_pn = nullptr, *q_pl1 = nullptr, *q_pl2 = nullptr, *q_fv1 = nullptr, *q_fv2 = nullptr, *q_fimp = nullptr;
    int4* ev_pend = nullptr;
    int* ev_pend_count = nullptr;
    bool ev_pending = false;
    bool query_points = false;         // the last query was testCollision (its contact points are kept)
    bool contacts_is_a = true;         // pair table that holds the last step's contacts (read by get_contacts)
    int contacts_np = 0;               // its pair count
    int enqueue_broadphase(bool capturing, const rp3dg::DevEvents& E, int pg);
    rp3dg::NpArgs np_args(const rp3dg::DevPairs& P, const rp3dg::DevPairs& O, int query, int report, int body1, int body2);
    int enqueue_narrowphase(const rp3dg::NpArgs& A, bool capturing, int pg);
    int query(int mode, int report, int body1, int body2);

    // launch hints from the last completed step (grids are grid-stride, any size is correct)
    int hint_pairs = 0, hint_manifolds = 0, hint_rounds = 8;
    rp3dg::DevCounters last{};         // counters of the last completed step

    // captured step graphs, one per pair-table parity
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    float graph_dt = 0.0f;
    int graph_rounds = 0;

    // hull arena
    rp3dg::HullHeader* d_hulls = nullptr;
    // concave shapes (ConcaveInst per concave collider, mesh / height field arenas), the triangle tests
    // of the step and of the last step (alternating with the pair tables: CTa goes with Pa), their
    // contact points, and the tail entries of the pair tables for concave pairs' extra manifolds
    bool has_concave = false;
    rp3dg::ConcaveInst* d_cinst = nullptr;
    float4 *d_cverts = nullptr, *d_cnorm = nullptr, *d_cnodes = nullptr;
    int4* d_ctris = nullptr;
    float* d_cheights = nullptr;
    rp3dg::ConcaveTests CTa{}, CTb{};
    float4 *cp_nd = nullptr, *cp_l1 = nullptr, *cp_l2 = nullptr;
    rp3dg::TriArena tri_arena{};   // the step's triangles as hulls, one slot per concave test
    int cp_cap = 0;
    int tail_cap = 0;
    int man_cap = 0;           // contact manifolds of a step: pair_cap + tail_cap
    float4 *d_hverts = nullptr, *d_hnormals = nullptr, *d_hbounds = nullptr;
    int4* d_hedges = nullptr;
    int *d_hfedge = nullptr, *d_hfverts = nullptr, *d_hfvidx = nullptr;
    rp3dg::NoColLists nocol_lists{};          // per-body partners that must not collide
    int *d_nocol_offs = nullptr, *d_nocol_list = nullptr;

    // broadphase scratch
    uint64_t *bp_keys = nullptr, *bp_ktmp = nullptr, *bp_out = nullptr;
    uint64_t* bp_slab = nullptr;       // first kTravSlab partners of every collider (first traversal pass)
    int* bp_walk = nullptr;            // colliders walked again after the scan (more partners than the slab)
    int *bp_vals = nullptr, *bp_vtmp = nullptr;
    int2 *bp_child = nullptr, *bp_range = nullptr;
    int *bp_parent = nullptr, *bp_root = nullptr, *bp_count = nullptr;
    int* bp_offs_ab[2] = {nullptr, nullptr};   // per-collider segment starts of pair table A / B
    float4 *fit_min = nullptr, *fit_max = nullptr;   // unions of aligned 64 / 4096 / 262144-leaf blocks
    float4 *bp_nmin = nullptr, *bp_nmax = nullptr, *bp_partial = nullptr;
    // narrowphase
    int *np_lists = nullptr, *np_counts = nullptr;
    // manifolds / colouring
    int *m_flag = nullptr, *m_offs = nullptr, *m_list = nullptr;
    int *adj_cnt = nullptr, *adj_offs = nullptr, *adj_cursor = nullptr, *adj = nullptr;
    int2* mb = nullptr;
    int* color = nullptr;
    uint64_t *ckeys = nullptr, *cktmp = nullptr;
    int *cvals = nullptr, *cvtmp = nullptr;
    uint64_t* sorted_keys = nullptr;   // segment key of each solver row after the colour sort
    // scan / sort
    int* scan_sums_buf = nullptr;
    int scan_sums_cap = 0;
    int *sort_hist = nullptr, *sort_offs = nullptr;
    int sort_hist_cap = 0;

    rp3dg::DevCounters* counters = nullptr;
    rp3dg::DevCounters* hcounters = nullptr;

    bool profiling = false;
    std::string profile_only;          // time only this kernel (empty: all)
    bool timer_open = false;
    int capture_parity = -1;           // >= 0 while a graph is being captured
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, KernelStat> stats;

    ~rp3dg_batch();
    int upload(rp3dg_world* const* worlds, int n, const rp3dg_batch_options* opt);
    void set_class_blocks();
    int launch_world_classes(const char* name, const rp3dg::FusedArgs& F, bool lds);
    int upload_joints(rp3dg_world* const* worlds, int n);
    void enqueue_sleep(float dt, const rp3dg::DevPairs& N, int pair_grid);
    int enqueue_step(float dt, bool capturing);
    int finish_steps(int n);
    int step(float dt);
    int step_graph(float dt, int n);
    void drop_graphs();
    int scan(const int* in, int n, const int* nptr, int* out, hipStream_t s);
    int sort(uint64_t** keys, int** vals, uint64_t* ktmp, int* vtmp, int n, const int* nptr, int bits, hipStream_t s);
    void timer_begin(const char* name);
    void timer_end();
    void timer_flush();
};
