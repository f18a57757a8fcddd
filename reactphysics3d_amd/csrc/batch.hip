// Device batch: uploads one or many host worlds into SoA arenas in HBM and
// runs the step in the phase order of PhysicsWorld::update()
// (src/engine/PhysicsWorld.cpp:239-307).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/rp3d_mi355x.h"
#include "batch.h"
#include "raycast.h"
#include "fused.h"
#include "mlp.h"
#include "kcommon.h"
#include "narrowphase.h"
#include "solver.h"
#include "world.h"

namespace rp3dg {

// broadphase.hip
__global__ void bp_refit(DevBodies B, DevColliders C, const float4* hull_bounds, int init);
__global__ void bp_bounds(DevColliders C, float4* partial);
__global__ void bp_morton(DevColliders C, const float4* partial, int nparts, uint64_t* keys, int* vals);
__global__ void rs_hist(const uint64_t* keys, int n, const int* nptr, int shift, int* hist, int nblocks);
__global__ void rs_scatter(const uint64_t* kin, const int* vin, uint64_t* kout, int* vout, int n, const int* nptr,
                           int shift, const int* offs, int nblocks);
__global__ void bp_karras(const uint64_t* k, int n, int2* child, int* parent, int2* range);
__global__ void bp_leaves(DevColliders C, const int* sorted, int n, float4* nmin, float4* nmax, float4* bmin,
                          float4* bmax);
__global__ void bp_blocks(const float4* imin, const float4* imax, int nin, float4* omin, float4* omax);
__global__ void bp_internal(int n, const int2* range, float4* nmin, float4* nmax, const float4* b1min,
                            const float4* b1max, const float4* b2min, const float4* b2max, const float4* b3min,
                            const float4* b3max);
__global__ void bp_world_roots(const uint64_t* k, int n, const int2* range, int* root);
__global__ void bp_traverse(DevColliders C, const int* sorted, int n, const int2* child, const float4* nmin,
                            const float4* nmax, const int* root, NoColLists nc, int* count,
                            const int* offset, uint64_t* out_keys, uint64_t* slab, int capacity, int pass,
                            int* walk_list, int* walk_count, int* status);
__global__ void bp_walk_wave(DevColliders C, const int* sorted, int n, const int2* child, const float4* nmin,
                             const float4* nmax, const int* root, NoColLists ncl, const int* offset, uint64_t* out_keys,
                             int capacity, const int* walk_list, const int* walk_count, int* status);
__global__ void bp_copy_slabs(DevColliders C, int n, const int* count, const int* offset, const uint64_t* slab,
                              uint64_t* out_keys, int capacity, int* status);
__global__ void bp_world_scan(DevColliders C, const int* gworld, const int* wcoll, const int* block_group, NoColLists L,
                              int* count, const int* offset, uint64_t* out_keys, int capacity, int pass, int* status);
__global__ void bp_merge(DevColliders C, const uint64_t* keys, const int* nkeys_ptr, int capacity, DevPairs O,
                         const int* nold_ptr, const int* old_offs, DevPairs N, int* nnew_ptr, int* nb_pairs_ptr);

// ---------------------------------------------------------------------------------------------
// exclusive scan (out has n+1 entries; out[n] = total).  With nptr the element count is
// min(n, *nptr), read on the device, and n is only the capacity the grid was sized for.
// ---------------------------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ __launch_bounds__(kScanThreads) void scan_tiles(const int* in, int n, const int* nptr, int* out, int* sums) {
    if (nptr) n = min(n, *nptr);
    __shared__ int sm[17];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
    for (int k = 0; k < kScanItems; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int total;
    int run = block_excl_scan(s, sm, &total);
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void scan_sums(int* sums, int nb, int* out, int n, const int* nptr) {
    if (nptr) n = min(n, *nptr);
    __shared__ int sm[17];
    int carry = 0;
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < nb ? sums[i] : 0;
        int total;
        const int e = block_excl_scan(v, sm, &total);
        if (i < nb) sums[i] = e + carry;
        carry += total;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

__global__ __launch_bounds__(kScanThreads) void scan_add(int* out, int n, const int* nptr, const int* sums) {
    if (nptr) n = min(n, *nptr);
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    const int add = sums[blockIdx.x];
    for (int k = 0; k < kScanItems; k++)
        if (base + k < n) out[base + k] += add;
}

// Clears device memory inside the step (a kernel rather than a memset node, so the step can be
// captured into a graph).
__global__ void fill_int(int* p, int n, int v) { RP3DG_GRID_STRIDE(i, n) p[i] = v; }
// centre of mass (and its constrained copy) from the body origin and orientation:
// x = p + q * com_local (RigidBody::setTransform, RigidBody.cpp)
__global__ void set_body_com(DevBodies B) {
    RP3DG_GRID_STRIDE(i, B.n) {
        const Q4 q = ldq(B.q, i);
        const V3 x = ld3(B.p, i) + qrot(q, ld3(B.com, i));
        B.x[i] = f4(x, 0.f);
        B.cx[i] = f4(x, 0.f);
        B.cq[i] = f4q(q);
    }
}

// After a world query's broadphase: every surviving pair of the new table takes the contact cache of its
// entry in the previous table (the last step's contacts, which the next step warm-starts from: the
// reference's snapshot queries leave mPreviousContactManifolds alone); new pairs have none.
__global__ void qr_carry(DevPairs N, DevPairs O, const int* n_new) {
    const int n = *n_new;
    RP3DG_GRID_STRIDE(k, n) {
        const int p = N.prev[k];
        if (p < 0) {
            N.ncp[k] = 0;
            continue;
        }
        N.ncp[k] = O.ncp[p];
        for (int j = 0; j < 4; j++) {
            N.pn[4 * k + j] = O.pn[4 * p + j];
            N.pl1[4 * k + j] = O.pl1[4 * p + j];
            N.pl2[4 * k + j] = O.pl2[4 * p + j];
        }
        N.fv1[k] = O.fv1[p];
        N.fv2[k] = O.fv2[p];
        N.fimp[k] = O.fimp[p];
    }
}

// Dispatch order of the fused solver's workgroups inside one size class: longest expected group first
// (list scheduling), the expected length being the largest sweep count of the group's worlds in the
// previous step, ties kept in the static (size) order.  One workgroup per class; n <= kOrderMax.
constexpr int kOrderMax = 4096;
__global__ void sv_order_groups(const int* static_order, const int* gworld, const int* world_iters, int* block_world,
                                const int2* ranges, int nclass) {
    __shared__ int s_key[kOrderMax];
    const int c = blockIdx.x;
    if (c >= nclass) return;
    const int first = ranges[c].x, n = ranges[c].y;
    if (n > kOrderMax) {   // too many to rank here: the static order
        for (int i = threadIdx.x; i < n; i += blockDim.x) block_world[first + i] = static_order[first + i];
        return;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int g = static_order[first + i];
        int m = 0;
        for (int w = gworld[g]; w < gworld[g + 1]; w++) m = max(m, world_iters[w]);
        s_key[i] = m;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int ki = s_key[i];
        int rank = 0;
        for (int j = 0; j < n; j++) {
            const int kj = s_key[j];
            rank += (kj > ki) || (kj == ki && j < i);
        }
        block_world[first + rank] = static_order[first + i];
    }
}

__global__ void fill_int_n(int* p, int n, const int* nptr, int v) {
    n = min(n, *nptr);
    RP3DG_GRID_STRIDE(i, n) p[i] = v;
}

}  // namespace rp3dg

using namespace rp3dg;

#define HIPCHECK(x)                                                                              \
    do {                                                                                         \
        hipError_t e__ = (x);                                                                    \
        if (e__ != hipSuccess) {                                                                 \
            set_error(std::string("HIP error ") + hipGetErrorString(e__) + " at " #x);           \
            return RP3DG_EHIP;                                                                   \
        }                                                                                        \
    } while (0)

static int grid_for(int n, int threads = 256, int max_blocks = 4096) {
    int g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > max_blocks) g = max_blocks;
    return g;
}

template <typename T>
static T* dalloc(rp3dg_batch* b, size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
    b->allocs.push_back(p);
    b->alloc_bytes.push_back(n * sizeof(T));
    b->bytes += n * sizeof(T);
    return (T*)p;
}

// Kernel timing with HIP events on the batch's stream (eager steps; graphs are not timed).
void rp3dg_batch::timer_begin(const char* name) {
    timer_open = false;
    if (!profiling) return;
    if (!profile_only.empty() && profile_only != name) return;
    if (capture_parity >= 0) return;   // graphs are not timed
    const size_t k = pending.size();
    while (event_pool.size() < 2 * (k + 1)) {
        hipEvent_t ev;
        if (hipEventCreate(&ev) != hipSuccess) return;
        event_pool.push_back(ev);
    }
    hipEvent_t a = event_pool[2 * k], e = event_pool[2 * k + 1];
    pending.push_back({name, a, e});
    hipEventRecord(a, stream);
    timer_open = true;
}
void rp3dg_batch::timer_end() {
    if (!timer_open) return;
    timer_open = false;
    hipEventRecord(pending.back().e, stream);
}
static void accumulate(std::map<std::string, rp3dg_batch::KernelStat>& stats, const rp3dg_batch::Pending& p) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.e) != hipSuccess) return;
    rp3dg_batch::KernelStat& s = stats[p.name];
    s.total_ms += ms;
    s.launches += 1;
}
void rp3dg_batch::timer_flush() {
    if (pending.empty()) return;
    hipStreamSynchronize(stream);
    for (auto& p : pending) accumulate(stats, p);
    pending.clear();
}

int rp3dg_batch::scan(const int* in, int n, const int* nptr, int* out, hipStream_t s) {
    const int tiles = std::max(1, (n + kScanTile - 1) / kScanTile);
    if (tiles > scan_sums_cap) return RP3DG_EOVERFLOW;
    scan_tiles<<<tiles, kScanThreads, 0, s>>>(in, n, nptr, out, scan_sums_buf);
    scan_sums<<<1, 1024, 0, s>>>(scan_sums_buf, tiles, out, n, nptr);
    scan_add<<<tiles, kScanThreads, 0, s>>>(out, n, nptr, scan_sums_buf);
    return RP3DG_OK;
}

// Stable LSD radix sort of (keys, vals) over the low `bits` bits; n keys, or min(n, *nptr) when the
// count lives on the device.  The result is in (*keys, *vals), which may be the tmp buffers.
int rp3dg_batch::sort(uint64_t** keys, int** vals, uint64_t* ktmp, int* vtmp, int n, const int* nptr, int bits,
                      hipStream_t s) {
    if (n <= 1) return RP3DG_OK;
    const int nblocks = (n + 4095) / 4096;
    if (nblocks * 256 > sort_hist_cap) return RP3DG_EOVERFLOW;
    uint64_t* ki = *keys;
    int* vi = *vals;
    uint64_t* ko = ktmp;
    int* vo = vtmp;
    for (int shift = 0; shift < bits; shift += 8) {
        rs_hist<<<nblocks, 256, 0, s>>>(ki, n, nptr, shift, sort_hist, nblocks);
        int r = scan(sort_hist, nblocks * 256, nullptr, sort_offs, s);
        if (r) return r;
        rs_scatter<<<nblocks, 256, 0, s>>>(ki, vi, ko, vo, n, nptr, shift, sort_offs, nblocks);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    *keys = ki;
    *vals = vi;
    return RP3DG_OK;
}

// tail: entries past the pairs for the extra contact manifolds of concave pairs (their col and contact
// cache arrays only)
static void alloc_pairs(rp3dg_batch* b, DevPairs& P, int cap, int tail) {
    const size_t ct = (size_t)cap + tail;
    P.key = dalloc<uint64_t>(b, cap);
    P.col = dalloc<int2>(b, ct);
    P.algo = dalloc<int>(b, cap);
    P.lf = dalloc<uint32_t>(b, cap);
    P.lfaxis = dalloc<float4>(b, cap);
    P.ncp = dalloc<int>(b, ct);
    P.prev = dalloc<int>(b, cap);
    P.fv1 = dalloc<float4>(b, ct);
    P.fv2 = dalloc<float4>(b, ct);
    P.fimp = dalloc<float4>(b, ct);
    P.pn = dalloc<float4>(b, 4 * ct);
    P.pl1 = dalloc<float4>(b, 4 * ct);
    P.pl2 = dalloc<float4>(b, 4 * ct);
    P.evf = dalloc<int>(b, cap);
    P.kept = dalloc<int>(b, cap);
    P.ct0 = P.ctn = P.xm = P.xs = nullptr;
    if (b->has_concave) {
        P.ct0 = dalloc<int>(b, cap); P.ctn = dalloc<int>(b, cap);
        P.xm = dalloc<int>(b, cap); P.xs = dalloc<int>(b, cap);
        if (P.xs) {
            hipMemset(P.ctn, 0, sizeof(int) * cap);
            hipMemset(P.xm, 0, sizeof(int) * cap);
        }
    }
}

static void alloc_tests(rp3dg_batch* b, ConcaveTests& T, int cap) {
    T.cap = cap;
    T.pair = dalloc<int>(b, cap); T.tri = dalloc<int>(b, cap);
    T.lf = dalloc<uint32_t>(b, cap); T.lfaxis = dalloc<float4>(b, cap);
    T.hit = dalloc<int>(b, cap); T.pts = dalloc<int2>(b, cap);
    T.count = nullptr;   // the step's counter (np_args), the counters block is allocated later
}

// Box half-edge structure (PhysicsCommon.cpp:73-109 face lists), numbered by the reference rule.
static bool box_topology(HalfEdgeMesh& m) {
    m.faces = {{0, 1, 2, 3}, {1, 5, 6, 2}, {4, 7, 6, 5}, {4, 0, 3, 7}, {4, 5, 1, 0}, {2, 6, 7, 3}};
    std::string err;
    return m.build(8, &err);
}

// Position of every collider of a world in the iteration order of BroadPhaseSystem::mMovedShapes
// (Set<int> with identity hash, Set.h:356-400): ascending bucket (broadphase id & (H-1)), and in a
// bucket the most recently inserted first; colliders enter in creation order.  Broadphase ids of
// sequentially created colliders are 0, 1, 3, 5, ... (DynamicAABBTree.cpp:93-131: leaf and
// internal nodes share the node pool).  H is the hash size reached when all colliders were added.
static void moved_ranks(int n, std::vector<int>& rank) {
    uint64_t H = 16;
    auto alloc_entries = [](uint64_t h) -> uint64_t {
        uint64_t e = (uint64_t)(h * 0.75);
        return (uint64_t)(std::ceil(e / 16.0f) * 16);
    };
    while (alloc_entries(H) < (uint64_t)n) H *= 2;
    std::vector<std::pair<uint64_t, int>> keyed(n);
    for (int k = 0; k < n; k++) {
        const uint64_t bp = k == 0 ? 0 : (uint64_t)(2 * k - 1);
        keyed[k] = {((bp & (H - 1)) << 32) | (uint64_t)(0x7fffffff - k), k};
    }
    std::sort(keyed.begin(), keyed.end());
    rank.assign(n, 0);
    for (int r = 0; r < n; r++) rank[keyed[r].second] = r;
}

// Joint rows.  Constants follow the joint constructors with world-space anchors
// (BallAndSocketJoint.cpp:38-53, FixedJoint.cpp:36-56, HingeJoint.cpp:36-80, SliderJoint.cpp:38-91); the
// rows are sorted by (kind in solve order, colour, creation index) with a greedy colouring over all bodies.
int rp3dg_batch::upload_joints(rp3dg_world* const* ws, int nw) {
    struct Row { int kind, color, order, b1, b2, group, slot, world; float4 la1, la2, ax1, ax2, qinit, lim; int flags; };
    std::vector<Row> rows;
    int jbase = 0;
    for (int wi = 0; wi < nw; wi++) {
        const HostWorld& W = *ws[wi];
        for (size_t ji = 0; ji < W.joints.size(); ji++) {
            const HostJoint& hj = W.joints[ji];
            Row r{};
            if (hj.type == RP3DG_JOINT_BALLSOCKET) r.kind = JT_BALLSOCKET;
            else if (hj.type == RP3DG_JOINT_FIXED) r.kind = JT_FIXED;
            else if (hj.type == RP3DG_JOINT_HINGE) r.kind = JT_HINGE;
            else r.kind = JT_SLIDER;
            r.order = jbase + (int)ji;
            r.group = world_group[wi];   // fused layout: rows are ordered by world group
            r.b1 = body_base[wi] + hj.body1;
            r.b2 = body_base[wi] + hj.body2;
            const Tf& t1 = W.bodies[hj.body1].transform;
            const Tf& t2 = W.bodies[hj.body2].transform;
            const V3 a1 = tapply(tinv(t1), hj.anchor_world);
            const V3 a2 = tapply(tinv(t2), hj.anchor_world);
            r.la1 = make_float4(a1.x, a1.y, a1.z, 0.f);
            r.la2 = make_float4(a2.x, a2.y, a2.z, 0.f);
            Q4 qi = q4(0, 0, 0, 1);
            if (r.kind == JT_FIXED) {
                qi = qmul(qconj(t2.q), t1.q);
            } else if (r.kind == JT_SLIDER) {
                const V3 h1 = normalized(qrot(qconj(t1.q), hj.axis_world));
                r.ax1 = make_float4(h1.x, h1.y, h1.z, 0.f);
                qi = qmul(qconj(t2.q), t1.q);
            } else if (r.kind == JT_HINGE) {
                const V3 h1 = normalized(qrot(qconj(t1.q), hj.axis_world));
                const V3 h2 = normalized(qrot(qconj(t2.q), hj.axis_world));
                r.ax1 = make_float4(h1.x, h1.y, h1.z, 0.f);
                r.ax2 = make_float4(h2.x, h2.y, h2.z, 0.f);
                Q4 d = qmul(t2.q, qconj(t1.q));
                const float l = qlen(d);
                d = q4(d.x / l, d.y / l, d.z / l, d.w / l);
                qi = q4(-d.x, -d.y, -d.z, d.w);
            }
            r.qinit = make_float4(qi.x, qi.y, qi.z, qi.w);
            r.flags = ((hj.flags & 1) ? JF_LIMIT : 0) | ((hj.flags & 2) ? JF_MOTOR : 0) |
                      ((hj.flags & RP3DG_JOINT_BAUMGARTE) ? JF_BAUMGARTE : 0);
            r.world = wi;
            // HingeJointInfo / SliderJointInfo defaults when the limit is off (HingeJoint.h:97-105,
            // SliderJoint.h:93-101): [-1, 1]; ball-and-socket: cone half angle, pi when the cone limit is
            // off (PhysicsWorld.cpp:500)
            if (r.kind == JT_BALLSOCKET)
                r.lim = make_float4((hj.flags & 1) ? hj.limits[0] : kPi, 0.f, 0.f, 0.f);
            else
                r.lim = make_float4((hj.flags & 1) ? hj.limits[0] : -1.0f, (hj.flags & 1) ? hj.limits[1] : 1.0f,
                                    hj.limits[2], hj.limits[3]);
            if (r.kind == JT_BALLSOCKET) r.flags &= JF_LIMIT | JF_BAUMGARTE;   // cone limit, no motor
            else if (r.kind == JT_FIXED) r.flags &= JF_BAUMGARTE;
            rows.push_back(r);
        }
        jbase += (int)W.joints.size();
    }
    // greedy colouring per kind in creation order
    std::vector<uint64_t> used(nb, 0);
    for (int kind = 0; kind < JT_KINDS; kind++) {
        std::fill(used.begin(), used.end(), 0ull);
        for (Row& r : rows) {
            if (r.kind != kind) continue;
            const uint64_t u = used[r.b1] | used[r.b2];
            int c = 0;
            while (c < 63 && (u >> c) & 1ull) c++;
            r.color = c;
            used[r.b1] |= 1ull << c;
            used[r.b2] |= 1ull << c;
        }
    }
    // slots = the (kind, colour) classes present, in solve order
    std::vector<int> slot_of(JT_KINDS * 64, -1);
    for (const Row& r : rows) slot_of[r.kind * 64 + r.color] = 0;
    jt_slots = 0;
    for (int k = 0; k < JT_KINDS * 64; k++)
        if (slot_of[k] >= 0) slot_of[k] = jt_slots++;
    for (Row& r : rows) r.slot = slot_of[r.kind * 64 + r.color];
    const bool wmajor = fused;
    std::stable_sort(rows.begin(), rows.end(), [wmajor](const Row& a, const Row& b) {
        if (wmajor && a.group != b.group) return a.group < b.group;
        if (a.slot != b.slot) return a.slot < b.slot;
        return a.order < b.order;
    });
    const int n = (int)rows.size();
    std::vector<int> ranges;   // per-colour kernels: [2*slot] = [first, end) over all worlds
    std::vector<int2> wr((size_t)ngroups * std::max(jt_slots, 1), make_int2(0, 0));   // fused: per group and slot
    if (!wmajor) {
        for (int i = 0; i < n; i++) {
            if (i == 0 || rows[i].slot != rows[i - 1].slot) ranges.push_back(i);
            if (i == n - 1 || rows[i + 1].slot != rows[i].slot) ranges.push_back(i + 1);
        }
    } else {
        std::vector<int> cnt((size_t)ngroups * std::max(jt_slots, 1), 0);
        for (const Row& r : rows) cnt[(size_t)r.group * jt_slots + r.slot]++;
        int run = 0;
        for (size_t k = 0; k < cnt.size(); k++) {
            wr[k] = make_int2(run, run + cnt[k]);
            run += cnt[k];
        }
    }
    joint_order.resize(n);
    J.n = n;
    const size_t m = std::max(1, n);
    J.type = dalloc<int>(this, m); J.bodies = dalloc<int2>(this, m); J.flags = dalloc<int>(this, m);
    for (float4** a : {&J.la1, &J.la2, &J.ax1, &J.ax2, &J.qinit, &J.lim, &J.impT, &J.impR, &J.impL, &J.i1r0, &J.i1r1,
                       &J.i1r2, &J.i2r0, &J.i2r1, &J.i2r2, &J.r1, &J.r2, &J.kt0, &J.kt1, &J.kt2, &J.kr0, &J.kr1,
                       &J.kr2, &J.a1, &J.b2xa1, &J.c2xa1, &J.un1, &J.un2, &J.ua, &J.rn1, &J.rn2, &J.ra})
        *a = dalloc<float4>(this, m);
    J.bias0 = dalloc<float4>(this, m);
    J.bias1 = dalloc<float4>(this, m);
    J.srank = dalloc<int>(this, m);
    if (!J.bias0 || !J.bias1 || !J.srank) { set_error("device allocation failed (joints)"); return RP3DG_EHIP; }
    // slider ranks in each world's solve order (rows are sorted by slot, then creation index, within a world)
    std::vector<int> srank(m, -1), nsl(nw, 0);
    bool any_bslider = false;
    for (int i = 0; i < n; i++) {
        if (rows[i].kind != JT_SLIDER) continue;
        srank[i] = nsl[rows[i].world]++;
        any_bslider |= (rows[i].flags & JF_BAUMGARTE) != 0;
    }
    J.bfirst = nullptr;
    if (any_bslider) {
        J.bfirst = dalloc<int>(this, std::max(nw, 1));
        if (!J.bfirst) { set_error("device allocation failed (joints)"); return RP3DG_EHIP; }
    }
    J.imlm = dalloc<float>(this, m);
    J.imm = dalloc<float>(this, m);
    jt_ranges = dalloc<int>(this, std::max<size_t>(2, ranges.size()));
    jt_wr = dalloc<int2>(this, wr.size());
    if (!J.imlm || !J.imm || !jt_ranges || !jt_wr) { set_error("device allocation failed (joints)"); return RP3DG_EHIP; }
    if (n == 0) return RP3DG_OK;
    std::vector<int> type(n), flags(n);
    std::vector<int2> bodies(n);
    std::vector<float4> la1(n), la2(n), ax1(n), ax2(n), qinit(n), lim(n), zero(n, make_float4(0.f, 0.f, 0.f, 0.f));
    for (int i = 0; i < n; i++) {
        const Row& r = rows[i];
        type[i] = r.kind; flags[i] = r.flags; bodies[i] = make_int2(r.b1, r.b2);
        la1[i] = r.la1; la2[i] = r.la2; ax1[i] = r.ax1; ax2[i] = r.ax2; qinit[i] = r.qinit; lim[i] = r.lim;
        joint_order[i] = r.order;
    }
    HIPCHECK(hipMemcpy(J.type, type.data(), 4 * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(J.flags, flags.data(), 4 * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(J.bodies, bodies.data(), 8 * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(J.la1, la1.data(), 16 * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(J.la2, la2.data(), 16 * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(J.ax1, ax1.data(), 16 * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(J.ax2, ax2.data(), 16 * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(J.qinit, qinit.data(), 16 * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(J.lim, lim.data(), 16 * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(J.srank, srank.data(), 4 * n, hipMemcpyHostToDevice));
    for (float4* a : {J.bias0, J.bias1}) HIPCHECK(hipMemcpy(a, zero.data(), 16 * n, hipMemcpyHostToDevice));
    for (float4* a : {J.impT, J.impR, J.impL, J.a1, J.b2xa1, J.c2xa1, J.ra})
        HIPCHECK(hipMemcpy(a, zero.data(), 16 * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(J.imlm, 0, sizeof(float) * n));
    HIPCHECK(hipMemset(J.imm, 0, sizeof(float) * n));
    if (!ranges.empty()) HIPCHECK(hipMemcpy(jt_ranges, ranges.data(), sizeof(int) * ranges.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(jt_wr, wr.data(), sizeof(int2) * wr.size(), hipMemcpyHostToDevice));
    return RP3DG_OK;
}

int rp3dg_batch::upload(rp3dg_world* const* ws, int nw, const rp3dg_batch_options* opt) {
    worlds.assign(ws, ws + nw);
    rp3dg_batch_options o{};
    if (opt) o = *opt;
    if (o.solver < RP3DG_SOLVER_AUTO || o.solver > RP3DG_SOLVER_FUSED) { set_error("invalid solver layout"); return RP3DG_EINVAL; }
    if (o.fused_block && (o.fused_block < 64 || o.fused_block > kFusedMaxBlock || o.fused_block % 64)) {
        set_error("fused_block must be a multiple of 64 in [64, 256]");
        return RP3DG_EINVAL;
    }
    if (o.pairs_per_collider < 0) { set_error("pairs_per_collider must be >= 0"); return RP3DG_EINVAL; }
    if (o.group_bodies < 0) { set_error("group_bodies must be >= 0"); return RP3DG_EINVAL; }
    if (o.broadphase < RP3DG_BROADPHASE_AUTO || o.broadphase > RP3DG_BROADPHASE_WORLD_SCAN) {
        set_error("invalid broadphase");
        return RP3DG_EINVAL;
    }
    // host SoA
    std::vector<float4> x, q, p, v, w, force, torque, ii, com, llock, alock;
    std::vector<int> flags, bworld;
    std::vector<int> cbody, ctype, cworld, crank;
    std::vector<float4> shape, lpos, lq;
    std::vector<float2> mat;
    std::vector<uint32_t> filter;
    std::vector<int> ckind;
    std::vector<HullHeader> hh;
    std::vector<float4> hverts, hnormals, hbounds;
    std::vector<int4> hedges;
    std::vector<int> hfedge, hfverts, hfvidx;
    std::vector<uint64_t> nocol;
    // concave shapes: one instance per concave collider (its scale), the meshes and height fields of its
    // world once per batch; the collider's shape.x is fixed up to its bounds slot after the hulls
    std::vector<ConcaveInst> cinst;
    std::vector<float4> cverts, cnorm, cnodes, cbounds;
    std::vector<int4> ctris;
    std::vector<float> cheights;
    std::vector<int> cinst_collider;
    body_base.clear();
    collider_base.clear();
    for (int wi = 0; wi < nw; wi++) {
        const HostWorld& W = *ws[wi];
        const int b0 = (int)x.size();
        const int c0 = (int)cbody.size();
        const int h0 = (int)hh.size();
        body_base.push_back(b0);
        collider_base.push_back(c0);
        for (const HostBody& b : W.bodies) {
            x.push_back(make_float4(b.com_world.x, b.com_world.y, b.com_world.z, 0.f));
            q.push_back(make_float4(b.transform.q.x, b.transform.q.y, b.transform.q.z, b.transform.q.w));
            p.push_back(make_float4(b.transform.p.x, b.transform.p.y, b.transform.p.z, 0.f));
            v.push_back(make_float4(b.v.x, b.v.y, b.v.z, 0.f));
            w.push_back(make_float4(b.w.x, b.w.y, b.w.z, 0.f));
            force.push_back(make_float4(b.force.x, b.force.y, b.force.z, 0.f));
            torque.push_back(make_float4(b.torque.x, b.torque.y, b.torque.z, 0.f));
            ii.push_back(make_float4(b.inv_inertia.x, b.inv_inertia.y, b.inv_inertia.z, b.inv_mass));
            com.push_back(make_float4(b.com_local.x, b.com_local.y, b.com_local.z, b.mass));
            llock.push_back(make_float4(b.llock.x, b.llock.y, b.llock.z, b.ldamp));
            alock.push_back(make_float4(b.alock.x, b.alock.y, b.alock.z, b.adamp));
            int f = 0;
            if (b.type == RP3DG_DYNAMIC) f |= BF_DYNAMIC;
            if (b.type == RP3DG_KINEMATIC) f |= BF_KINEMATIC;
            if (b.type != RP3DG_STATIC) f |= BF_ENABLED;
            if (b.gravity) f |= BF_GRAVITY;
            if (b.allowed_to_sleep) f |= BF_ALLOW_SLEEP;
            if (b.llock.x == 1.f && b.llock.y == 1.f && b.llock.z == 1.f && b.alock.x == 1.f && b.alock.y == 1.f &&
                b.alock.z == 1.f)
                f |= BF_UNIT_LOCKS;
            flags.push_back(f);
            bworld.push_back(wi);
            body_type.push_back(b.type);
        }
        std::vector<int> rank;
        moved_ranks((int)W.colliders.size(), rank);
        for (size_t ci = 0; ci < W.colliders.size(); ci++) {
            const HostCollider& c = W.colliders[ci];
            cbody.push_back(b0 + c.body);
            ctype.push_back(c.type);
            cworld.push_back(wi);
            crank.push_back(rank[ci]);
            float4 s = make_float4(c.params[0], c.params[1], c.params[2], c.params[3]);
            if (c.type == SHAPE_CAPSULE) s = make_float4(c.params[0], c.params[1] * 0.5f, 0.f, 0.f);
            if (c.type == SHAPE_HULL) s = make_float4((float)(h0 + (int)c.params[0]), 0.f, 0.f, 0.f);
            if (shape_is_concave(c.type)) {
                if (W.bodies[c.body].type == RP3DG_DYNAMIC) {
                    set_error("concave shapes (triangle mesh, height field) go on static or kinematic bodies only");
                    return RP3DG_EUNSUPPORTED;
                }
                ConcaveInst I{};
                I.kind = c.type;
                I.sx = c.params[1]; I.sy = c.params[2]; I.sz = c.params[3];
                I.isx = 1.0f / I.sx; I.isy = 1.0f / I.sy; I.isz = 1.0f / I.sz;
                if (c.type == SHAPE_CONCAVE_MESH) {
                    const HostTriMesh& M = W.meshes[(int)c.params[0]];
                    I.vbase = (int)cverts.size();
                    I.tbase = (int)ctris.size();
                    I.nbase = (int)cnorm.size();
                    I.node_base = (int)cnodes.size() / 2;
                    I.root = M.root;
                    for (const V3& vv : M.vertices) cverts.push_back(make_float4(vv.x, vv.y, vv.z, 0.f));
                    for (size_t t3 = 0; t3 + 2 < M.tris.size(); t3 += 3)
                        ctris.push_back(make_int4(M.tris[t3], M.tris[t3 + 1], M.tris[t3 + 2], 0));
                    for (const V3& nn : concave_scaled_normals(M, v3(I.sx, I.sy, I.sz)))
                        cnorm.push_back(make_float4(nn.x, nn.y, nn.z, 0.f));
                    for (const HostTriMesh::Node& nd : M.nodes) {
                        auto bits = [](int v) { float f; memcpy(&f, &v, sizeof f); return f; };
                        cnodes.push_back(make_float4(nd.mn.x, nd.mn.y, nd.mn.z, bits(nd.child0)));
                        cnodes.push_back(make_float4(nd.mx.x, nd.mx.y, nd.mx.z, bits(nd.child0 < 0 ? nd.data : nd.child1)));
                    }
                } else {
                    const HostHeightField& H = W.heightfields[(int)c.params[0]];
                    I.columns = H.columns; I.rows = H.rows;
                    I.hbase = (int)cheights.size();
                    I.width = H.width; I.length = H.length; I.origin = H.origin;
                    I.bminx = H.bmin.x; I.bminy = H.bmin.y; I.bminz = H.bmin.z;
                    I.bmaxx = H.bmax.x; I.bmaxy = H.bmax.y; I.bmaxz = H.bmax.z;
                    cheights.insert(cheights.end(), H.heights.begin(), H.heights.end());
                }
                V3 lo, hi;
                concave_local_bounds(W, c, lo, hi);
                cbounds.push_back(make_float4(lo.x, lo.y, lo.z, 0.f));
                cbounds.push_back(make_float4(hi.x, hi.y, hi.z, 0.f));
                s = make_float4(0.f, (float)cinst.size(), 0.f, 0.f);
                cinst_collider.push_back((int)cbody.size() - 1);
                cinst.push_back(I);
            }
            shape.push_back(s);
            lpos.push_back(make_float4(c.local.p.x, c.local.p.y, c.local.p.z, 0.f));
            lq.push_back(make_float4(c.local.q.x, c.local.q.y, c.local.q.z, c.local.q.w));
            mat.push_back(make_float2(c.friction_sqrt, c.bounciness));
            filter.push_back((uint32_t)c.category | ((uint32_t)c.mask << 16));
            ckind.push_back(c.kind);
        }
        for (const HostHull& H : W.hulls) {
            HullHeader h;
            h.vbase = (int)hverts.size();
            h.nv = (int)H.vertices.size();
            h.fbase = (int)hnormals.size();
            h.nf = (int)H.face_normals.size();
            h.ebase = (int)hedges.size();
            h.ne = (int)H.mesh.edges.size();
            h.fvbase = (int)hfvidx.size();
            h.nfv = 0;
            h.cx = H.centroid.x; h.cy = H.centroid.y; h.cz = H.centroid.z; h.pad = 0.f;
            for (const V3& vv : H.vertices) hverts.push_back(make_float4(vv.x, vv.y, vv.z, 0.f));
            for (size_t f = 0; f < H.face_normals.size(); f++) {
                const V3& nn = H.face_normals[f];
                hnormals.push_back(make_float4(nn.x, nn.y, nn.z, 0.f));
                hfedge.push_back(H.mesh.face_edge[f]);
                hfverts.push_back((int)hfvidx.size());
                hfverts.push_back((int)H.mesh.faces[f].size());
                for (int vi : H.mesh.faces[f]) hfvidx.push_back(vi);
                h.nfv += (int)H.mesh.faces[f].size();
            }
            for (const HalfEdgeMesh::Edge& e : H.mesh.edges) hedges.push_back(make_int4(e.vertex, e.twin, e.face, e.next));
            hbounds.push_back(make_float4(H.local_min.x, H.local_min.y, H.local_min.z, 0.f));
            hbounds.push_back(make_float4(H.local_max.x, H.local_max.y, H.local_max.z, 0.f));
            hh.push_back(h);
        }
        for (const HostJoint& j : W.joints) {
            if (!j.collide_connected) {
                const int a = b0 + j.body1, c = b0 + j.body2;
                nocol.push_back(((uint64_t)(uint32_t)std::min(a, c) << 32) | (uint32_t)std::max(a, c));
            }
        }
        const rp3dg_world_settings& cfg = W.cfg;
        if (wi == 0) {
            settings.gravity = v3(cfg.gravity[0], cfg.gravity[1], cfg.gravity[2]);
            settings.persistent_threshold = cfg.persistent_contact_distance_threshold;
            settings.restitution_threshold = cfg.restitution_velocity_threshold;
            settings.cos_similar = cfg.cos_angle_similar_contact_manifold;
            settings.vel_iters = (int)cfg.default_velocity_solver_iterations;
            settings.pos_iters = (int)cfg.default_position_solver_iterations;
            settings.gravity_enabled = cfg.is_gravity_enabled;
            settings.sleeping_enabled = cfg.is_sleeping_enabled ? 1 : 0;
            settings.sleep_lin_sq = cfg.sleep_linear_velocity * cfg.sleep_linear_velocity;
            settings.sleep_ang_sq = cfg.sleep_angular_velocity * cfg.sleep_angular_velocity;
            settings.time_before_sleep = cfg.time_before_sleep;
            settings.split_impulse = W.split_impulse ? 1 : 0;
        } else if (memcmp(&cfg, &ws[0]->cfg, sizeof(cfg)) != 0 || W.split_impulse != ws[0]->split_impulse) {
            set_error("all worlds of a batch must share the same WorldSettings");
            return RP3DG_EINVAL;
        }
    }
    std::sort(nocol.begin(), nocol.end());
    nocol.erase(std::unique(nocol.begin(), nocol.end()), nocol.end());
    // concave local bounds follow the hulls' in the bounds table
    const int nhull_slots = (int)hh.size();
    for (size_t i = 0; i < cinst.size(); i++) shape[cinst_collider[i]].x = (float)(nhull_slots + (int)i);
    has_concave = !cinst.empty();
    nb = (int)x.size();
    nc = (int)cbody.size();
    nworlds = nw;
    world_bits = 0;
    while ((1 << world_bits) < nw) world_bits++;
    body_base.push_back(nb);
    collider_base.push_back(nc);
    max_world_bodies = 0;
    for (int wi = 0; wi < nw; wi++) max_world_bodies = std::max(max_world_bodies, body_base[wi + 1] - body_base[wi]);
    fused = o.solver == RP3DG_SOLVER_FUSED || (o.solver == RP3DG_SOLVER_AUTO && max_world_bodies <= 4096);
    if (fused && max_world_bodies > 4096) {
        set_error("the fused solver layout takes worlds of up to 4096 bodies (use the per-colour layout)");
        return RP3DG_EINVAL;
    }
    // (solver velocities and inverse inertia of worlds up to 512 bodies live in LDS during the sweeps,
    // 112 B per body: see the world classes below)
    // world groups: consecutive worlds share a workgroup up to group_bodies bodies (64 worlds), so the
    // colour phases of small worlds fill the wavefronts; a larger world is a group of its own
    group_bodies = std::min(o.group_bodies ? o.group_bodies : 256, 512);
    group_world.clear();
    world_group.assign(nw, 0);
    for (int wi = 0; wi < nw;) {
        int wj = wi + 1, sum = body_base[wi + 1] - body_base[wi];
        while (fused && wj < nw && wj - wi < 64 && sum + body_base[wj + 1] - body_base[wj] <= group_bodies) {
            sum += body_base[wj + 1] - body_base[wj];
            wj++;
        }
        for (int k = wi; k < wj; k++) world_group[k] = (int)group_world.size();
        group_world.push_back(wi);
        wi = wj;
    }
    group_world.push_back(nw);
    ngroups = (int)group_world.size() - 1;
    max_group_bodies = 0;
    max_group_colliders = 0;
    for (int g = 0; g < ngroups; g++) {
        max_group_bodies = std::max(max_group_bodies, body_base[group_world[g + 1]] - body_base[group_world[g]]);
        max_group_colliders =
            std::max(max_group_colliders, collider_base[group_world[g + 1]] - collider_base[group_world[g]]);
    }
    {
        int max_world_colliders = 0;
        for (int wi = 0; wi < nw; wi++)
            max_world_colliders = std::max(max_world_colliders, collider_base[wi + 1] - collider_base[wi]);
        // 32 B of LDS per collider of a group
        const bool fits = fused && max_world_colliders <= 1024 && max_group_colliders <= 2048;
        if (o.broadphase == RP3DG_BROADPHASE_WORLD_SCAN && !fits) {
            set_error("the world-scan broadphase needs the fused layout and worlds of at most 1024 colliders");
            return RP3DG_EINVAL;
        }
        // automatic choice: one workgroup scans a whole world (its colliders squared box tests), which
        // pays for itself with many worlds sharing the GPU or small worlds, not for a few large ones
        const bool worth = max_world_colliders <= 256 || nw >= 64;
        world_scan = fits && (o.broadphase == RP3DG_BROADPHASE_WORLD_SCAN || (o.broadphase == RP3DG_BROADPHASE_AUTO && worth));
    }
    user_block = o.fused_block;
    if (o.fused_block) fused_block = o.fused_block;
    else fused_block = max_world_bodies <= 128 ? 64 : (max_world_bodies <= 1024 ? 128 : 256);

    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    {
        int dev = 0, lds = 0;
        HIPCHECK(hipGetDevice(&dev));
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
            max_lds_per_block = lds;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess && lds > 0)
            max_lds_per_cu = lds;
    }
    B.n = nb;
    B.x = dalloc<float4>(this, nb); B.q = dalloc<float4>(this, nb); B.p = dalloc<float4>(this, nb);
    B.v = dalloc<float4>(this, nb); B.w = dalloc<float4>(this, nb);
    B.cv = dalloc<float4>(this, nb); B.cw = dalloc<float4>(this, nb);
    B.sv = dalloc<float4>(this, nb); B.sw = dalloc<float4>(this, nb);
    B.cx = dalloc<float4>(this, nb); B.cq = dalloc<float4>(this, nb);
    B.force = dalloc<float4>(this, nb); B.torque = dalloc<float4>(this, nb);
    B.ii = dalloc<float4>(this, nb);
    B.iw0 = dalloc<float4>(this, nb); B.iw1 = dalloc<float4>(this, nb); B.iw2 = dalloc<float4>(this, nb);
    B.com = dalloc<float4>(this, nb);
    B.llock = dalloc<float4>(this, nb); B.alock = dalloc<float4>(this, nb);
    B.flags = dalloc<int>(this, nb); B.world = dalloc<int>(this, nb);
    C.n = nc;
    C.body = dalloc<int>(this, nc); C.type = dalloc<int>(this, nc);
    C.shape = dalloc<float4>(this, nc);
    C.lpos = dalloc<float4>(this, nc); C.lq = dalloc<float4>(this, nc);
    C.wpos = dalloc<float4>(this, nc); C.wq = dalloc<float4>(this, nc);
    C.amin = dalloc<float4>(this, nc); C.amax = dalloc<float4>(this, nc);
    C.mat = dalloc<float2>(this, nc);
    C.filter = dalloc<uint32_t>(this, nc);
    C.kind = dalloc<int>(this, nc);
    C.moved = dalloc<int>(this, nc); C.rank = dalloc<int>(this, nc); C.world = dalloc<int>(this, nc);
    if (!B.world || !C.world) { set_error("device allocation failed"); return RP3DG_EHIP; }
    auto up = [&](void* d, const void* h, size_t bytes) { return hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream); };
    HIPCHECK(up(B.x, x.data(), nb * 16)); HIPCHECK(up(B.q, q.data(), nb * 16)); HIPCHECK(up(B.p, p.data(), nb * 16));
    HIPCHECK(up(B.v, v.data(), nb * 16)); HIPCHECK(up(B.w, w.data(), nb * 16));
    HIPCHECK(up(B.force, force.data(), nb * 16)); HIPCHECK(up(B.torque, torque.data(), nb * 16));
    HIPCHECK(up(B.ii, ii.data(), nb * 16)); HIPCHECK(up(B.com, com.data(), nb * 16));
    HIPCHECK(up(B.llock, llock.data(), nb * 16)); HIPCHECK(up(B.alock, alock.data(), nb * 16));
    HIPCHECK(up(B.flags, flags.data(), nb * 4)); HIPCHECK(up(B.world, bworld.data(), nb * 4));
    for (float4* a : {B.cv, B.cw, B.sv, B.sw, B.iw0, B.iw1, B.iw2}) HIPCHECK(hipMemsetAsync(a, 0, nb * 16, stream));
    HIPCHECK(up(B.cx, x.data(), nb * 16)); HIPCHECK(up(B.cq, q.data(), nb * 16));
    HIPCHECK(up(C.body, cbody.data(), nc * 4)); HIPCHECK(up(C.type, ctype.data(), nc * 4));
    HIPCHECK(up(C.shape, shape.data(), nc * 16)); HIPCHECK(up(C.lpos, lpos.data(), nc * 16));
    HIPCHECK(up(C.lq, lq.data(), nc * 16)); HIPCHECK(up(C.mat, mat.data(), nc * 8));
    HIPCHECK(up(C.filter, filter.data(), nc * 4)); HIPCHECK(up(C.rank, crank.data(), nc * 4));
    HIPCHECK(up(C.kind, ckind.data(), nc * 4));
    coll_kind = ckind;
    HIPCHECK(up(C.world, cworld.data(), nc * 4));
    // hulls (at least one dummy record so pointers are valid), then the concave shapes' bounds
    if (hh.empty()) { hh.push_back(HullHeader{}); hbounds.resize(2); }
    if (has_concave) {
        hbounds.resize(2 * (size_t)nhull_slots);
        hbounds.insert(hbounds.end(), cbounds.begin(), cbounds.end());
        auto upv = [&](auto*& d, const auto& v) {
            using T = typename std::decay<decltype(v)>::type::value_type;
            d = dalloc<T>(this, std::max<size_t>(1, v.size()));
            if (d && !v.empty()) return hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
            return d ? hipSuccess : hipErrorOutOfMemory;
        };
        HIPCHECK(upv(d_cinst, cinst)); HIPCHECK(upv(d_cverts, cverts)); HIPCHECK(upv(d_ctris, ctris));
        HIPCHECK(upv(d_cnorm, cnorm)); HIPCHECK(upv(d_cnodes, cnodes)); HIPCHECK(upv(d_cheights, cheights));
    }
    d_hulls = dalloc<HullHeader>(this, hh.size());
    d_hverts = dalloc<float4>(this, hverts.size());
    d_hnormals = dalloc<float4>(this, hnormals.size());
    d_hedges = dalloc<int4>(this, hedges.size());
    d_hfedge = dalloc<int>(this, hfedge.size());
    d_hfverts = dalloc<int>(this, hfverts.size());
    d_hfvidx = dalloc<int>(this, hfvidx.size());
    d_hbounds = dalloc<float4>(this, hbounds.size());
    HIPCHECK(up(d_hulls, hh.data(), hh.size() * sizeof(HullHeader)));
    if (!hverts.empty()) HIPCHECK(up(d_hverts, hverts.data(), hverts.size() * 16));
    if (!hnormals.empty()) HIPCHECK(up(d_hnormals, hnormals.data(), hnormals.size() * 16));
    if (!hedges.empty()) HIPCHECK(up(d_hedges, hedges.data(), hedges.size() * 16));
    if (!hfedge.empty()) HIPCHECK(up(d_hfedge, hfedge.data(), hfedge.size() * 4));
    if (!hfverts.empty()) HIPCHECK(up(d_hfverts, hfverts.data(), hfverts.size() * 4));
    if (!hfvidx.empty()) HIPCHECK(up(d_hfvidx, hfvidx.data(), hfvidx.size() * 4));
    HIPCHECK(up(d_hbounds, hbounds.data(), hbounds.size() * 16));
    {
        // colliders of every body (the between-step RigidBody calls refit them)
        std::vector<int> offs(nb + 1, 0), list(nc);
        for (int i = 0; i < nc; i++) offs[cbody[i] + 1]++;
        for (int i = 0; i < nb; i++) offs[i + 1] += offs[i];
        std::vector<int> cur(offs.begin(), offs.end() - 1);
        for (int i = 0; i < nc; i++) list[cur[cbody[i]]++] = i;
        d_coll_offs = dalloc<int>(this, nb + 1);
        d_coll_list = dalloc<int>(this, std::max(1, nc));
        if (!d_coll_list) { set_error("device allocation failed (body colliders)"); return RP3DG_EHIP; }
        HIPCHECK(hipMemcpy(d_coll_offs, offs.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice));
        if (nc) HIPCHECK(hipMemcpy(d_coll_list, list.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
    }
    {
        // per-body lists of the body pairs that must not collide, both directions
        std::vector<int> offs(nb + 1, 0), list(2 * nocol.size());
        for (uint64_t k : nocol) {
            offs[(int)(k >> 32) + 1]++;
            offs[(int)(uint32_t)k + 1]++;
        }
        for (int i = 0; i < nb; i++) offs[i + 1] += offs[i];
        std::vector<int> cur(offs.begin(), offs.end() - 1);
        for (uint64_t k : nocol) {
            const int a = (int)(k >> 32), c = (int)(uint32_t)k;
            list[cur[a]++] = c;
            list[cur[c]++] = a;
        }
        d_nocol_offs = dalloc<int>(this, nb + 1);
        d_nocol_list = dalloc<int>(this, std::max<size_t>(1, list.size()));
        if (!d_nocol_list) { set_error("device allocation failed"); return RP3DG_EHIP; }
        HIPCHECK(up(d_nocol_offs, offs.data(), offs.size() * 4));
        if (!list.empty()) HIPCHECK(up(d_nocol_list, list.data(), list.size() * 4));
        nocol_lists.offs = d_nocol_offs;
        nocol_lists.list = d_nocol_list;
    }

    // box topology
    HalfEdgeMesh box;
    if (!box_topology(box)) { set_error("box topology"); return RP3DG_EINVAL; }
    std::vector<int> e4, fe(6), fv;
    for (auto& e : box.edges) { e4.push_back(e.vertex); e4.push_back(e.twin); e4.push_back(e.face); e4.push_back(e.next); }
    for (int f = 0; f < 6; f++) fe[f] = box.face_edge[f];
    for (auto& f : box.faces) for (int vv : f) fv.push_back(vv);
    upload_box_topology(e4.data(), fe.data(), fv.data());

    // broadphase / pair / solver arenas
    pair_cap = std::max(4096, (o.pairs_per_collider ? o.pairs_per_collider : 16) * nc);
    tail_cap = has_concave ? pair_cap : 0;
    man_cap = pair_cap + tail_cap;
    alloc_pairs(this, Pa, pair_cap, tail_cap);
    alloc_pairs(this, Pb, pair_cap, tail_cap);
    if (has_concave) {
        const int ct_cap = std::max(1 << 16, 16 * nc);
        alloc_tests(this, CTa, ct_cap);
        alloc_tests(this, CTb, ct_cap);
        cp_cap = 2 * ct_cap;
        cp_nd = dalloc<float4>(this, cp_cap); cp_l1 = dalloc<float4>(this, cp_cap); cp_l2 = dalloc<float4>(this, cp_cap);
        if (!cp_l2 || !CTb.pts || !Pb.xs) { set_error("device allocation failed (concave shapes)"); return RP3DG_EHIP; }
        // TriangleShape's half-edge structure (PhysicsCommon.cpp initTriangleShapeHalfEdgeStructure)
        HalfEdgeMesh tri;
        tri.faces = {{0, 1, 2}, {0, 2, 1}};
        std::string err;
        if (!tri.build(3, &err)) { set_error("triangle topology"); return RP3DG_EINVAL; }
        std::vector<int> t4, tfe(2), tfv;
        for (auto& e : tri.edges) { t4.push_back(e.vertex); t4.push_back(e.twin); t4.push_back(e.face); t4.push_back(e.next); }
        for (int f = 0; f < 2; f++) tfe[f] = tri.face_edge[f];
        for (auto& f : tri.faces) for (int vv : f) tfv.push_back(vv);
        TriArena& TA = tri_arena;
        TA.verts = dalloc<float4>(this, 3 * ct_cap); TA.normals = dalloc<float4>(this, 2 * ct_cap);
        TA.face_edge = dalloc<int>(this, 2 * ct_cap); TA.face_verts = dalloc<int>(this, 4 * ct_cap);
        int4* d_edges = dalloc<int4>(this, 6);
        int* d_vidx = dalloc<int>(this, 6);
        if (!TA.face_verts || !d_vidx) { set_error("device allocation failed (concave shapes)"); return RP3DG_EHIP; }
        hipMemcpy(d_edges, t4.data(), sizeof(int4) * 6, hipMemcpyHostToDevice);
        hipMemcpy(d_vidx, tfv.data(), sizeof(int) * 6, hipMemcpyHostToDevice);
        TA.edges = d_edges; TA.vidx = d_vidx;
        TA.face_edge0[0] = tfe[0]; TA.face_edge0[1] = tfe[1];
    }
    const int nn = std::max(1, 2 * nc);
    bp_keys = dalloc<uint64_t>(this, nc); bp_vals = dalloc<int>(this, nc);
    bp_ktmp = dalloc<uint64_t>(this, nc); bp_vtmp = dalloc<int>(this, nc);
    bp_child = dalloc<int2>(this, nn); bp_parent = dalloc<int>(this, nn); bp_range = dalloc<int2>(this, nn);
    bp_nmin = dalloc<float4>(this, nn); bp_nmax = dalloc<float4>(this, nn);
    {
        const int nb1 = (nc + 63) / 64, nb2 = (nb1 + 63) / 64, nb3 = (nb2 + 63) / 64;
        fit_min = dalloc<float4>(this, nb1 + nb2 + nb3);
        fit_max = dalloc<float4>(this, nb1 + nb2 + nb3);
    }
    bp_root = dalloc<int>(this, (size_t)1 << world_bits);
    bp_partial = dalloc<float4>(this, 2 * 256);
    bp_count = dalloc<int>(this, nc + 1);
    for (int t = 0; t < 2; t++) {
        bp_offs_ab[t] = dalloc<int>(this, nc + 1);
        if (bp_offs_ab[t]) HIPCHECK(hipMemset(bp_offs_ab[t], 0, sizeof(int) * (nc + 1)));
    }
    bp_out = dalloc<uint64_t>(this, pair_cap);
    bp_slab = dalloc<uint64_t>(this, (size_t)nc * kTravSlab);
    bp_walk = dalloc<int>(this, nc);
    np_lists = dalloc<int>(this, (size_t)kNpListSlots * pair_cap);
    np_counts = dalloc<int>(this, ALGO_COUNT);
    m_flag = dalloc<int>(this, pair_cap); m_offs = dalloc<int>(this, pair_cap + 1);
    m_list = dalloc<int>(this, man_cap);
    adj_cnt = dalloc<int>(this, nb + 1); adj_offs = dalloc<int>(this, nb + 1); adj_cursor = dalloc<int>(this, nb);
    adj = dalloc<int>(this, 2 * (size_t)man_cap); mb = dalloc<int2>(this, man_cap);
    color = dalloc<int>(this, man_cap);
    ckeys = dalloc<uint64_t>(this, man_cap); cvals = dalloc<int>(this, man_cap);
    cktmp = dalloc<uint64_t>(this, man_cap); cvtmp = dalloc<int>(this, man_cap);
    nseg = fused ? ngroups * kMaxColors : kMaxColors;
    seg_cnt = dalloc<int>(this, nseg + 1);
    seg_start = dalloc<int>(this, nseg + 1);
    pair_count = dalloc<int>(this, 2);
    total_points = dalloc<unsigned long long>(this, 1);
    d_wbody = dalloc<int>(this, nw + 1);
    d_wcoll = dalloc<int>(this, nw + 1);
    d_block_world = dalloc<int>(this, ngroups);
    d_gworld = dalloc<int>(this, ngroups + 1);
    d_world_group = dalloc<int>(this, nw);
    d_wmstart = dalloc<int>(this, nw + 1);
    res_pv = dalloc<float4>(this, nb);
    res_pw = dalloc<float4>(this, nb);
    world_iters = dalloc<int>(this, nw);
    world_res = dalloc<float>(this, nw);
    if (!world_res) { set_error("device allocation failed (world tables)"); return RP3DG_EHIP; }
    {
        const int ng = ngroups;
        auto gsize = [&](int g) { return body_base[group_world[g + 1]] - body_base[group_world[g]]; };
        std::vector<int> order(ng);
        for (int g = 0; g < ng; g++) order[g] = g;
        // largest groups first, so the longest workgroups start first
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return gsize(a) > gsize(c); });
        HIPCHECK(up(d_wbody, body_base.data(), sizeof(int) * (nw + 1)));
        HIPCHECK(up(d_wcoll, collider_base.data(), sizeof(int) * (nw + 1)));
        HIPCHECK(hipMemcpy(d_block_world, order.data(), sizeof(int) * ng, hipMemcpyHostToDevice));
        d_block_static = dalloc<int>(this, ng);
        if (!d_block_static) { set_error("device allocation failed (group order)"); return RP3DG_EHIP; }
        HIPCHECK(hipMemcpy(d_block_static, order.data(), sizeof(int) * ng, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(d_gworld, group_world.data(), sizeof(int) * (ng + 1), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(d_world_group, world_group.data(), sizeof(int) * nw, hipMemcpyHostToDevice));
        // size classes over the size-sorted group order (largest class first)
        wclasses.clear();
        if (fused) {
            for (int i = 0; i < ng;) {
                const int sz = gsize(order[i]);
                int L = 64;
                while (L < sz) L *= 2;
                int j = i;
                while (j < ng && gsize(order[j]) > L / 2) j++;
                if (L == 64) j = ng;   // everything up to 64 bodies is one class
                wclasses.push_back({i, j - i, L, 64, L <= kFusedMaxLds ? L : 0});   // LDS: 64, 128, 256 or 512
                i = j;
            }
            set_class_blocks();
            std::vector<int2> cr;
            for (const WorldClass& wc : wclasses) cr.push_back(make_int2(wc.first, wc.count));
            d_class_ranges = dalloc<int2>(this, std::max<size_t>(1, cr.size()));
            if (!d_class_ranges) { set_error("device allocation failed (class ranges)"); return RP3DG_EHIP; }
            HIPCHECK(hipMemcpy(d_class_ranges, cr.data(), sizeof(int2) * cr.size(), hipMemcpyHostToDevice));
            for (size_t c = 1; c < wclasses.size(); c++) {
                hipStream_t s;
                HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
                side_streams.push_back(s);
            }
            for (size_t c = 0; c < 2 * wclasses.size(); c++) {
                hipEvent_t e;
                HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                fork_events.push_back(e);
            }
        }
        HIPCHECK(hipMemsetAsync(pair_count, 0, sizeof(int) * 2, stream));
        HIPCHECK(hipMemsetAsync(total_points, 0, sizeof(unsigned long long), stream));
        HIPCHECK(hipMemsetAsync(world_iters, 0, sizeof(int) * nw, stream));
        HIPCHECK(hipMemsetAsync(world_res, 0, sizeof(float) * nw, stream));
    }
    S.pair = dalloc<int>(this, man_cap); S.bodies = dalloc<int2>(this, man_cap); S.npts = dalloc<int>(this, man_cap);
    for (float4** a : {&S.n, &S.r1f, &S.r2f, &S.t1, &S.t2, &S.imp, &S.oldt1, &S.oldt2})
        *a = dalloc<float4>(this, man_cap);
    for (float4** a : {&S.pn, &S.r1, &S.r2, &S.lam, &S.pred}) *a = dalloc<float4>(this, 4 * (size_t)man_cap);
    scan_sums_cap = 1 << 16;
    scan_sums_buf = dalloc<int>(this, scan_sums_cap);
    sort_hist_cap = 256 * (std::max(man_cap, nc) / 4096 + 2);
    if ((sort_hist_cap + kScanTile - 1) / kScanTile > scan_sums_cap || (man_cap + kScanTile - 1) / kScanTile > scan_sums_cap) {
        set_error("batch too large for the scan workspace");
        return RP3DG_EOVERFLOW;
    }
    sort_hist = dalloc<int>(this, sort_hist_cap);
    sort_offs = dalloc<int>(this, sort_hist_cap + 1);
    HIPCHECK(hipMalloc(&counters, sizeof(DevCounters)));
    allocs.push_back(counters);
    alloc_bytes.push_back(sizeof(DevCounters));
    HIPCHECK(hipHostMalloc(&hcounters, sizeof(DevCounters), hipHostMallocDefault));
    HIPCHECK(hipMemsetAsync(counters, 0, sizeof(DevCounters), stream));
    if (!sort_offs || !S.lam) { set_error("device allocation failed (pair arenas)"); return RP3DG_EHIP; }
    // collider creation: world transforms, initial fat AABBs, all colliders in the moved set
    bp_refit<<<grid_for(nc), 256, 0, stream>>>(B, C, d_hbounds, 1);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(stream));

    // Static colliders never move, so their mutual overlaps are fixed at creation: list them once
    // (same closed AABB test and pair filter as the device traversal).
    std::vector<int> is_static(nc, 0), ss_offs(nc + 1, 0), ss_list;
    for (int i = 0; i < nc; i++) is_static[i] = (flags[cbody[i]] & (BF_DYNAMIC | BF_KINEMATIC)) == 0;
    {
        std::vector<float4> amn(nc), amx(nc);
        HIPCHECK(hipMemcpy(amn.data(), C.amin, sizeof(float4) * nc, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(amx.data(), C.amax, sizeof(float4) * nc, hipMemcpyDeviceToHost));
        std::vector<std::vector<int>> per_world(nw);
        for (int i = 0; i < nc; i++)
            if (is_static[i]) per_world[cworld[i]].push_back(i);
        std::vector<std::vector<int>> partners(nc);
        for (const auto& sl : per_world) {
            for (size_t a = 0; a < sl.size(); a++) {
                for (size_t b2 = a + 1; b2 < sl.size(); b2++) {
                    const int i = sl[a], j = sl[b2];
                    const float4 a0 = amn[i], a1 = amx[i], b0 = amn[j], b1 = amx[j];
                    if (a1.x < b0.x || b1.x < a0.x || a1.y < b0.y || b1.y < a0.y || a1.z < b0.z || b1.z < a0.z) continue;
                    if (cbody[i] == cbody[j]) continue;
                    if (shape_is_concave(ctype[i]) && shape_is_concave(ctype[j])) continue;
                    const uint32_t fi = filter[i], fj = filter[j];
                    if (((fi >> 16) & (fj & 0xffffu)) == 0 || ((fi & 0xffffu) & (fj >> 16)) == 0) continue;
                    const uint64_t bk = ((uint64_t)(uint32_t)std::min(cbody[i], cbody[j]) << 32) |
                                        (uint32_t)std::max(cbody[i], cbody[j]);
                    if (std::binary_search(nocol.begin(), nocol.end(), bk)) continue;
                    partners[std::min(i, j)].push_back(std::max(i, j));
                }
            }
        }
        for (int i = 0; i < nc; i++) {
            std::sort(partners[i].begin(), partners[i].end());
            ss_offs[i + 1] = ss_offs[i] + (int)partners[i].size();
            ss_list.insert(ss_list.end(), partners[i].begin(), partners[i].end());
        }
    }
    {
        const int r = upload_joints(ws, nw);
        if (r) return r;
    }
    C.is_static = dalloc<int>(this, nc);
    C.ss_offs = dalloc<int>(this, nc + 1);
    C.ss_list = dalloc<int>(this, std::max<size_t>(1, ss_list.size()));
    HIPCHECK(hipMemcpy(C.is_static, is_static.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(C.ss_offs, ss_offs.data(), sizeof(int) * (nc + 1), hipMemcpyHostToDevice));
    if (!ss_list.empty()) HIPCHECK(hipMemcpy(C.ss_list, ss_list.data(), sizeof(int) * ss_list.size(), hipMemcpyHostToDevice));
    if (settings.sleeping_enabled) {
        isl.parent = dalloc<int>(this, nb);
        isl.minst = dalloc<unsigned int>(this, nb);
        isl.active = dalloc<int>(this, nb);
        isl.sleep_time = dalloc<float>(this, nb);
        if (!isl.sleep_time) { set_error("device allocation failed (islands)"); return RP3DG_EHIP; }
        HIPCHECK(hipMemset(isl.sleep_time, 0, sizeof(float) * nb));
    }
    return RP3DG_OK;
}

#define LAUNCH_T(name, grid, block, kern, ...)               \
    do {                                                     \
        timer_begin(name);                                   \
        kern<<<(grid), (block), 0, stream>>>(__VA_ARGS__);   \
        timer_end();                                         \
    } while (0)
#define LAUNCH_TS(name, grid, block, shmem, kern, ...)           \
    do {                                                         \
        timer_begin(name);                                       \
        kern<<<(grid), (block), (shmem), stream>>>(__VA_ARGS__); \
        timer_end();                                             \
    } while (0)

// Enqueues one full step on the batch's stream with no host synchronisation (the per-colour solver
// layout reads the colour count back once, so only the fused layout can be captured).  Every
// data-dependent size lives in device memory; kernels loop grid-stride over it, so the grids below
// are performance hints only.
// CollisionDetectionSystem::computeBroadPhase (CD.cpp:103-121): the pair table of the current state from
// the previous one, for a step or a world query; E records the pairs it removes while colliding.
int rp3dg_batch::enqueue_broadphase(bool capturing, const DevEvents& E, int pg) {
    DevPairs& O = cur_is_a ? Pa : Pb;
    DevPairs& N = cur_is_a ? Pb : Pa;
    int* old_count = pair_count + (cur_is_a ? 0 : 1);
    int* new_count = pair_count + (cur_is_a ? 1 : 0);
    int* bp_offs = bp_offs_ab[cur_is_a ? 1 : 0];          // segments of the table written this step
    const int* old_offs = bp_offs_ab[cur_is_a ? 0 : 1];   // segments of the previous table
    int* dstatus = &counters->status;
    int r;
    uint64_t* k = bp_keys;
    int* vls = bp_vals;
    const size_t scan_lds = 2 * sizeof(float4) * max_group_colliders;
    if (world_scan) {
        // small worlds: every dynamic collider tests the boxes of its world in LDS
        LAUNCH_TS("bp_world_scan", ngroups, 256, scan_lds, bp_world_scan, C, d_gworld, d_wcoll, d_block_world,
                  nocol_lists, bp_count, bp_offs, bp_out, pair_cap, 0, dstatus);
    } else {
        const int nparts = std::min(256, grid_for(nc));
        LAUNCH_T("bp_bounds", nparts, 256, bp_bounds, C, bp_partial);
        LAUNCH_T("bp_morton", grid_for(nc), 256, bp_morton, C, bp_partial, nparts, bp_keys, bp_vals);
        timer_begin("bp_sort");
        r = sort(&k, &vls, bp_ktmp, bp_vtmp, nc, nullptr, 30 + world_bits, stream);
        timer_end();
        if (r) { set_error("radix sort overflow"); return r; }
        if (nc > 1) LAUNCH_T("bp_karras", grid_for(nc - 1), 256, bp_karras, k, nc, bp_child, bp_parent, bp_range);
        {
            const int nb1 = (nc + 63) / 64, nb2 = (nb1 + 63) / 64;
            timer_begin("bp_fit");
            bp_leaves<<<grid_for(nc), 256, 0, stream>>>(C, vls, nc, bp_nmin, bp_nmax, fit_min, fit_max);
            bp_blocks<<<grid_for(nb1), 256, 0, stream>>>(fit_min, fit_max, nb1, fit_min + nb1, fit_max + nb1);
            bp_blocks<<<grid_for(nb2), 256, 0, stream>>>(fit_min + nb1, fit_max + nb1, nb2, fit_min + nb1 + nb2,
                                                          fit_max + nb1 + nb2);
            if (nc > 1)
                bp_internal<<<grid_for(nc - 1), 256, 0, stream>>>(nc, bp_range, bp_nmin, bp_nmax, fit_min, fit_max,
                                                                  fit_min + nb1, fit_max + nb1, fit_min + nb1 + nb2,
                                                                  fit_max + nb1 + nb2);
            timer_end();
        }
        LAUNCH_T("bp_world_roots", grid_for(2 * nc - 1), 256, bp_world_roots, k, nc, bp_range, bp_root);
        LAUNCH_T("bp_traverse", grid_for(nc, kTravBlock, 1 << 20), kTravBlock, bp_traverse, C, vls, nc, bp_child,
                 bp_nmin, bp_nmax, bp_root, nocol_lists, bp_count, bp_offs, bp_out, bp_slab, pair_cap, 0, bp_walk,
                 &counters->nb_walk_again, dstatus);
    }
    r = scan(bp_count, nc, nullptr, bp_offs, stream);
    if (r) { set_error("scan overflow"); return r; }
    if (world_scan) {
        LAUNCH_TS("bp_world_scan", ngroups, 256, scan_lds, bp_world_scan, C, d_gworld, d_wcoll, d_block_world,
                  nocol_lists, bp_count, bp_offs, bp_out, pair_cap, 1, dstatus);
    } else {
        LAUNCH_T("bp_copy_pairs", grid_for(nc * kTravSlab, 256, 1 << 20), 256, bp_copy_slabs, C, nc, bp_count, bp_offs,
                 bp_slab, bp_out, pair_cap, dstatus);
        const int hw = capturing ? nc : std::min(nc, last.nb_walk_again + last.nb_walk_again / 4 + 256);
        LAUNCH_T("bp_walk_again", std::max(1, std::min(hw, 1 << 16)), kWave, bp_walk_wave, C, vls, nc, bp_child,
                 bp_nmin, bp_nmax, bp_root, nocol_lists, bp_offs, bp_out, pair_cap, bp_walk, &counters->nb_walk_again,
                 dstatus);
    }
    if (events_on) fill_int_n<<<pg, 256, 0, stream>>>(O.kept, pair_cap, old_count, 0);
    LAUNCH_T("bp_merge", pg, 256, bp_merge, C, bp_out, bp_offs + nc, pair_cap, O, old_count, old_offs, N, new_count,
             &counters->nb_pairs);
    if (events_on) LAUNCH_T("ev_removed", pg, 256, ev_removed, C, O, old_count, E);
    cur_is_a = !cur_is_a;
    return RP3DG_OK;
}

NpArgs rp3dg_batch::np_args(const DevPairs& P, const DevPairs& O, int query, int report, int body1, int body2) {
    NpArgs A;
    A.B = B; A.C = C; A.P = P; A.O = O;
    A.nb_pairs = &counters->nb_pairs;
    A.hulls = d_hulls; A.hull_verts = d_hverts; A.hull_normals = d_hnormals; A.hull_edges = d_hedges;
    A.hull_face_edge = d_hfedge; A.hull_face_verts = d_hfverts; A.hull_face_vidx = d_hfvidx;
    A.persistent_threshold = settings.persistent_threshold;
    A.cos_similar = settings.cos_similar;
    A.status = &counters->status;
    A.query = query;
    A.qbody1 = body1;
    A.qbody2 = body2;
    A.report = report;
    A.clip_prev = query == NP_STEP ? 1 : 0;           // testNarrowPhaseCollision(..., false) in snapshots
    A.lf_update = report == NP_REPORT_NONE ? 0 : 1;   // overlap snapshots skip processAllPotentialContacts
    A.cinst = d_cinst; A.cverts = d_cverts; A.ctris = d_ctris; A.cnorm = d_cnorm; A.cnodes = d_cnodes;
    A.cheights = d_cheights;
    A.hull_bounds = d_hbounds;
    const bool na = P.key == Pa.key;   // the test lists alternate with the pair tables
    A.CT = na ? CTa : CTb;
    A.CTo = na ? CTb : CTa;
    A.CT.count = &counters->nb_concave_tests;
    A.CTo.count = nullptr;   // the last step's tests are read through the previous table's ranges
    A.tri = tri_arena;
    A.cp_nd = cp_nd; A.cp_l1 = cp_l1; A.cp_l2 = cp_l2;
    A.cp_cap = cp_cap;
    A.cp_count = &counters->nb_concave_points;
    A.tail_base = pair_cap;
    A.tail_cap = tail_cap;
    A.tail_count = &counters->nb_tails;
    return A;
}

// CollisionDetectionSystem::testNarrowPhaseCollision (CD.cpp:618-665): the pairs are binned by algorithm
// and each bin is one launch.
int rp3dg_batch::enqueue_narrowphase(const NpArgs& A, bool capturing, int pg) {
    LAUNCH_T("np_classify", std::min(pg, 1024), 256, np_classify, A, np_lists, pair_cap, counters->nb_tests,
             &counters->nb_box_tests, &counters->nb_box_sep_tests, &counters->nb_hull_tests);
    int* L = np_lists;
    int* T = counters->nb_tests;
    auto hint = [&](int n, int block) {
        const int h = capturing ? pair_cap : std::min(pair_cap, n + n / 4 + 512);
        return grid_for(h, block);
    };
    auto tg = [&](int algo, int block) { return hint(last.nb_tests[algo], block); };
    LAUNCH_T("np_sphere_sphere", tg(ALGO_SPHERE_SPHERE, 256), 256, np_sphere_sphere, A, L + ALGO_SPHERE_SPHERE * pair_cap, T + ALGO_SPHERE_SPHERE);
    LAUNCH_T("np_sphere_capsule", tg(ALGO_SPHERE_CAPSULE, 256), 256, np_sphere_capsule, A, L + ALGO_SPHERE_CAPSULE * pair_cap, T + ALGO_SPHERE_CAPSULE);
    LAUNCH_T("np_capsule_capsule", tg(ALGO_CAPSULE_CAPSULE, 256), 256, np_capsule_capsule, A, L + ALGO_CAPSULE_CAPSULE * pair_cap, T + ALGO_CAPSULE_CAPSULE);
    LAUNCH_T("np_sphere_poly", tg(ALGO_SPHERE_POLY, 256), 256, np_sphere_poly, A, L + ALGO_SPHERE_POLY * pair_cap, T + ALGO_SPHERE_POLY);
    LAUNCH_T("np_capsule_poly", tg(ALGO_CAPSULE_POLY, 256), 256, np_capsule_poly, A, L + ALGO_CAPSULE_POLY * pair_cap, T + ALGO_CAPSULE_POLY);
    LAUNCH_T("np_box_box", hint(last.nb_box_tests, kNpBoxThreads), kNpBoxThreads, np_box_box, A, L + np_list_slot(ALGO_COUNT) * pair_cap,
             &counters->nb_box_tests);
    LAUNCH_T("np_box_box", hint(last.nb_box_sep_tests, kNpBoxThreads), kNpBoxThreads, np_box_box, A, L + np_list_slot(ALGO_COUNT + 1) * pair_cap,
             &counters->nb_box_sep_tests);
    LAUNCH_T("np_poly_poly", hint(last.nb_hull_tests, 128), 128, np_poly_poly, A, L + ALGO_POLY_POLY * pair_cap,
             &counters->nb_hull_tests);
    if (has_concave) {   // convex vs concave pairs: middle phase, triangle tests, manifolds
        const int* lc = L + np_list_slot(ALGO_CONCAVE) * pair_cap;
        LAUNCH_T("np_concave_mid", tg(ALGO_CONCAVE, 128), 128, np_concave_mid, A, lc, T + ALGO_CONCAVE);
        const int ht = capturing ? A.CT.cap : std::min(A.CT.cap, last.nb_concave_tests + last.nb_concave_tests / 4 + 512);
        LAUNCH_T("np_concave_tri", grid_for(ht, 128), 128, np_concave_tri, A);
        LAUNCH_T("np_concave_reduce", tg(ALGO_CONCAVE, 64), 64, np_concave_reduce, A, lc, T + ALGO_CONCAVE);
    }

    return RP3DG_OK;
}

int rp3dg_batch::enqueue_step(float dt, bool capturing) {
    DevPairs& O = cur_is_a ? Pa : Pb;
    DevPairs& N = cur_is_a ? Pb : Pa;
    int* old_count = pair_count + (cur_is_a ? 0 : 1);
    int* new_count = pair_count + (cur_is_a ? 1 : 0);
    int* bp_offs = bp_offs_ab[cur_is_a ? 1 : 0];          // segments of the table written this step
    const int* old_offs = bp_offs_ab[cur_is_a ? 0 : 1];   // segments of the previous table
    int* dstatus = &counters->status;
    int r;
    fill_int<<<1, 256, 0, stream>>>((int*)counters, (int)(sizeof(DevCounters) / sizeof(int)), 0);

    const int hp = capturing ? pair_cap : std::min(pair_cap, hint_pairs + hint_pairs / 4 + 1024);
    const int pg = grid_for(hp);
    DevEvents E{ev_buf, ev_cap, &counters->nb_events, events_on ? 1 : 0};
    if (ev_pending) {   // pairs that a world query's broadphase removed while colliding report ContactExit now
        ev_take_pending<<<1, 256, 0, stream>>>(ev_pend, ev_pend_count, ev_buf, ev_cap, &counters->nb_events);
        ev_pending = false;
    }
    r = enqueue_broadphase(capturing, E, pg);
    if (r) return r;
    contacts_is_a = cur_is_a;

    // ---------------- narrowphase (CD.cpp:618-718)
    NpArgs A = np_args(N, O, NP_STEP, NP_REPORT_NO_TRIGGER, -1, -1);
    r = enqueue_narrowphase(A, capturing, pg);
    if (r) return r;

    // ---------------- contact flags and events (CD.cpp:975-1011, :1708-1741)
    LAUNCH_T("ev_update", pg, 256, ev_update, B, C, N, &counters->nb_pairs, E, settings.sleeping_enabled ? 0 : 1);

    // ---------------- manifold list + colouring (replaces createIslands, PhysicsWorld.cpp:695-866)
    int* nm_ptr = &counters->nb_manifolds;
    LAUNCH_T("sv_flag", pg, 256, sv_flag_manifolds, N.ncp, N.xm, &counters->nb_pairs, m_flag);
    r = scan(m_flag, pair_cap, &counters->nb_pairs, m_offs, stream);
    if (r) { set_error("scan overflow"); return r; }
    LAUNCH_T("sv_compact", std::min(pg, 1024), 256, sv_compact_manifolds, m_flag, m_offs, &counters->nb_pairs, N.ncp, N.xm,
             N.xs, m_list, nm_ptr, &counters->nb_points, total_points);
    const int hm = capturing ? man_cap : std::min(man_cap, hint_manifolds + hint_manifolds / 4 + 1024);
    const int mg = grid_for(hm);
    if (settings.sleeping_enabled) {
        // ---------------- islands and wake-up (PhysicsWorld::createIslands, PhysicsWorld.cpp:695-866)
        LAUNCH_T("isl_init", grid_for(nb), 256, isl_init, B, isl);
        LAUNCH_T("isl_link", mg, 256, isl_link_manifolds, B, C, N, m_list, nm_ptr, isl);
        if (J.n > 0) LAUNCH_T("isl_link", grid_for(J.n), 256, isl_link_joints, B, J, isl);
        LAUNCH_T("isl_mark", grid_for(nb), 256, isl_mark, B, isl);
        LAUNCH_T("isl_wake", grid_for(nb), 256, isl_wake, B, isl);
        LAUNCH_T("ev_shift", pg, 256, ev_shift, B, C, N, &counters->nb_pairs);
    }
    const bool learned = fused && warm_mode == RP3DG_WARM_START_LEARNED && mlp;
    const bool features = learned || log_cap > 0;
    // manifold counts per body (the learned warm start's features read them; Jones-Plassmann needs them)
    if (!fused || features) {
        fill_int<<<grid_for(nb + 1), 256, 0, stream>>>(adj_cnt, nb + 1, 0);
        LAUNCH_T("sv_adj_count", mg, 256, sv_adj_count, B, C, N, m_list, nm_ptr, adj_cnt);
        r = scan(adj_cnt, nb, nullptr, adj_offs, stream);
        if (r) { set_error("scan overflow"); return r; }
    }
    uint64_t* ck = ckeys;
    int* cv = cvals;
    if (fused) {
        // greedy colouring per world group in LDS, rows placed colour by colour within each group
        LAUNCH_T("sv_world_ranges", mg, 256, sv_world_ranges, B, C, N, m_list, nm_ptr, nworlds, mb, d_wmstart);
        LAUNCH_TS("sv_group_color", ngroups, kWave, sizeof(unsigned long long) * max_group_bodies, sv_group_color, mb,
                  d_wmstart, d_gworld, d_wbody, ngroups, color, cvals, ckeys, seg_start, &counters->nb_colors, dstatus);
    } else {
        // Jones-Plassmann over the whole batch, then a radix sort of the rows by colour
        fill_int<<<grid_for(nb), 256, 0, stream>>>(adj_cursor, nb, 0);
        LAUNCH_T("sv_adj_fill", mg, 256, sv_adj_fill, B, C, N, m_list, nm_ptr, adj_offs, adj_cursor, adj, mb);
        fill_int_n<<<mg, 256, 0, stream>>>(color, man_cap, nm_ptr, -1);
        const int rounds = capturing ? graph_rounds : std::min(kColorRoundsMax, std::max(4, hint_rounds));
        for (int q = 0; q < rounds; q++)
            LAUNCH_T("sv_color_round", mg, 256, sv_color_round, mb, nm_ptr, adj_offs, adj, color,
                     q ? &counters->color_left[q - 1] : nullptr, &counters->color_left[q], dstatus);
        LAUNCH_T("sv_color_finish", 1, 1024, sv_color_finish, mb, nm_ptr, adj_offs, adj, color,
                 rounds ? &counters->color_left[rounds - 1] : nullptr, &counters->finish_rounds, dstatus);
        fill_int<<<grid_for(nseg + 1), 256, 0, stream>>>(seg_cnt, nseg + 1, 0);
        LAUNCH_T("sv_color_keys", mg, 256, sv_color_keys, color, nm_ptr, C, N, m_list, nullptr, ckeys, cvals, seg_cnt,
                 &counters->nb_colors);
        int sbits = 0;
        while ((1 << sbits) < nseg) sbits++;
        timer_begin("sv_color_sort");
        r = sort(&ck, &cv, cktmp, cvtmp, man_cap, nm_ptr, sbits, stream);
        timer_end();
        if (r) { set_error("sort overflow"); return r; }
        r = scan(seg_cnt, nseg, nullptr, seg_start, stream);
        if (r) { set_error("scan overflow"); return r; }
    }
    sorted_keys = ck;

    SvArgs SA;
    SA.B = B; SA.C = C; SA.P = N; SA.S = S;
    SA.nm = nm_ptr;
    SA.mlist = m_list;
    SA.order = cv;
    SA.cstart = seg_start;
    SA.ncolors = &counters->nb_colors;
    SA.restitution_threshold = settings.restitution_threshold;
    SA.learned = learned ? 1 : 0;
    SA.adj_offs = adj_offs;
    SA.split = settings.split_impulse;
    const float beta_dt = 0.2f / dt;

    if (fused) {
        // ---------------- one workgroup per world: integration, contact + joint solver, positions, refit
        FusedArgs F;
        F.A = SA;
        F.J = J;
        F.settings = settings;
        F.seg_start = seg_start;
        F.jt_wr = jt_wr;
        F.jt_slots = jt_slots;
        F.wbody = d_wbody;
        F.wcoll = d_wcoll;
        F.block_world = d_block_world;
        F.gworld = d_gworld;
        F.hull_bounds = d_hbounds;
        F.dt = dt;
        F.beta_dt = beta_dt;
        F.res_thr = res_thr;
        F.max_iters = res_thr > 0.0f ? res_max_iters : settings.vel_iters;
        F.pv = res_pv;
        F.pw = res_pw;
        F.world_iters = world_iters;
        F.world_res = world_res;
        F.nrows = &counters->nb_new_points;
        F.feat_cap = feat_cap;
        F.row_slot = row_slot;
        F.lds_bodies = 0;
        F.lds_residual = 0;
        F.feats = nullptr;
        F.phase = 2;
        if (res_thr > 0.0f && order_by_sweeps && !wclasses.empty())   // residual mode: groups differ in length
            LAUNCH_TS("sv_order_groups", (int)wclasses.size(), 256, 0, sv_order_groups, d_block_static, d_gworld,
                      world_iters, d_block_world, d_class_ranges, (int)wclasses.size());
        if (!features) {
            r = launch_world_classes("sv_world_solve", F, true);
            if (r) return r;
        } else {
            F.feats = feat_buf;
            F.phase = 0;
            if (settings.split_impulse) LAUNCH_TS("sv_world_prepare", ngroups, fused_block, 0, sv_world_solve<true>, F);
            else LAUNCH_TS("sv_world_prepare", ngroups, fused_block, 0, sv_world_solve<false>, F);
            if (SA.learned) {
                const int rows = capturing ? feat_cap : std::min(feat_cap, last.nb_new_points + last.nb_new_points / 4 + 1024);
                const int grid = std::max(1, std::min(2048, (rows + 63) / 64));
                LAUNCH_T("mlp_infer", grid, 256, mlp_infer, mlp_view(mlp), feat_buf, &counters->nb_new_points, feat_cap,
                         nullptr, row_slot, S.pred);
            }
            F.phase = 1;
            r = launch_world_classes("sv_world_solve", F, true);
            if (r) return r;
            if (log_cap > 0) {
                LAUNCH_T("sv_log_samples", grid_for(feat_cap), 256, sv_log_samples, F, log_feats, log_targets, log_count,
                         log_cap);
                sv_log_advance<<<1, 1, 0, stream>>>(&counters->nb_new_points, feat_cap, log_count, log_cap);
            }
        }
        LAUNCH_T("world_position", ngroups, fused_block, world_position, F);
        enqueue_sleep(dt, N, pg);
        return RP3DG_OK;
    }

    // ---------------- per-colour layout: the colour count decides the launches
    if (capturing) { set_error("graph capture needs the fused solver layout"); return RP3DG_EUNSUPPORTED; }
    HIPCHECK(hipMemcpyAsync(&hcounters->nb_colors, &counters->nb_colors, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    const int ncol = hcounters->nb_colors;
    const int bg = grid_for(nb);
    // inverse world inertia + velocity integration (PhysicsWorld.cpp:266-272)
    LAUNCH_T("dyn_integrate_velocities", bg, 256, dyn_integrate_velocities, B, settings, dt);
    // contact solver (PhysicsWorld.cpp:321-345)
    const int jg = grid_for(std::max(J.n, 1));
    LAUNCH_T("sv_prepare", mg, 256, sv_prepare, SA);
    const int cg = grid_for(hm / std::max(1, ncol) + 1, color_block);
    for (int c = 0; c < ncol; c++) LAUNCH_T("sv_warmstart", cg, color_block, sv_warmstart, SA, c);
    // ConstraintSolverSystem::initialize: initBeforeSolve + warm start of every joint kind
    if (J.n > 0) {
        if (J.bfirst) LAUNCH_T("fill_int", grid_for(nworlds), 256, fill_int, J.bfirst, nworlds, INT_MAX);
        LAUNCH_T("jt_init", jg, 256, jt_init, B, J, dt);
        for (int s = 0; s < jt_slots; s++) LAUNCH_T("jt_warmstart", jg, 256, jt_warmstart, B, J, jt_ranges, s);
    }
    for (int it = 0; it < settings.vel_iters; it++) {
        for (int s = 0; s < jt_slots; s++)
            LAUNCH_T("jt_solve_velocity", jg, 256, jt_solve_velocity, B, J, jt_ranges, s, dt);
        for (int c = 0; c < ncol; c++) LAUNCH_T("sv_solve_color", cg, color_block, sv_solve_color, SA, c, beta_dt);
    }
    LAUNCH_T("sv_store", mg, 256, sv_store, SA);
    // position integration, joint position correction, state update, refit (PhysicsWorld.cpp:275-288, :348-360)
    LAUNCH_T("dyn_integrate_positions", bg, 256, dyn_integrate_positions, B, dt, settings.split_impulse ? 1.0f : 0.0f);
    for (int it = 0; it < settings.pos_iters; it++)
        for (int s = 0; s < jt_slots; s++)
            LAUNCH_T("jt_solve_position", jg, 256, jt_solve_position, B, J, jt_ranges, s);
    LAUNCH_T("dyn_update_state", bg, 256, dyn_update_state, B);
    LAUNCH_T("bp_refit", grid_for(nc), 256, bp_refit, B, C, d_hbounds, 0);
    enqueue_sleep(dt, N, pg);
    return RP3DG_OK;
}

// PhysicsWorld::updateSleepingBodies (PhysicsWorld.cpp:871-926), after the colliders are refit
void rp3dg_batch::enqueue_sleep(float dt, const DevPairs& N, int pair_grid) {
    if (!settings.sleeping_enabled) return;
    LAUNCH_T("isl_sleep", grid_for(nb), 256, isl_sleep_time, B, isl, settings, dt);
    LAUNCH_T("isl_sleep", grid_for(nb), 256, isl_sleep, B, isl, settings);
    LAUNCH_T("isl_sleep", pair_grid, 256, isl_reset_disabled_pairs, B, C, N, &counters->nb_pairs);
}

// Workgroup size of every world class: clamp(L / block_div, 64, 256).
void rp3dg_batch::set_class_blocks() {
    for (WorldClass& c : wclasses)
        c.block = user_block ? user_block : std::max(64, std::min(kFusedMaxBlock, c.max_bodies / block_div));
}

// One launch per world class of the fused solver, each on its own stream between a fork and a join
// event (captured as parallel graph branches), so small and large worlds share the CUs.
int rp3dg_batch::launch_world_classes(const char* name, const FusedArgs& F0, bool lds) {
    timer_begin(name);
    const int nclass = (int)wclasses.size();
    if (nclass > 1) {
        HIPCHECK(hipEventRecord(fork_events[0], stream));
        for (int c = 1; c < nclass; c++) HIPCHECK(hipStreamWaitEvent(side_streams[c - 1], fork_events[0], 0));
    }
    for (int c = 0; c < nclass; c++) {
        const WorldClass& wc = wclasses[c];
        if (wc.count == 0) continue;
        FusedArgs F = F0;
        F.block_world = d_block_world + wc.first;
        F.lds_bodies = lds ? wc.lds_bodies : 0;
        // the residual's previous velocities join the staged state when that costs no resident workgroup
        // per CU: LDS bounds how many groups a CU holds, and fewer groups in flight cost more than the HBM
        // round trips saved (measured: 9.0 -> 9.6 ms per launch on the 4096-world batch when the
        // 256-body class dropped from five to four groups per CU)
        {
            const size_t base = fused_lds_bytes(F.lds_bodies, false) + 1024, with = fused_lds_bytes(F.lds_bodies, true) + 1024;
            F.lds_residual = F.lds_bodies > 0 && F.res_thr > 0.0f && with <= (size_t)max_lds_per_block &&
                             (size_t)max_lds_per_cu / with == (size_t)max_lds_per_cu / base;
        }
        const size_t shmem = fused_lds_bytes(F.lds_bodies, F.lds_residual != 0);
        hipStream_t s = c == 0 ? stream : side_streams[c - 1];
        if (settings.split_impulse) sv_world_solve<true><<<wc.count, wc.block, shmem, s>>>(F);
        else sv_world_solve<false><<<wc.count, wc.block, shmem, s>>>(F);
    }
    if (nclass > 1) {
        for (int c = 1; c < nclass; c++) {
            HIPCHECK(hipEventRecord(fork_events[c], side_streams[c - 1]));
            HIPCHECK(hipStreamWaitEvent(stream, fork_events[c], 0));
        }
    }
    timer_end();
    return RP3DG_OK;
}

// Reads the counters of the last enqueued step back and checks the device status word.
int rp3dg_batch::finish_steps(int n) {
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(hcounters, counters, sizeof(DevCounters), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    timer_flush();
    last = *hcounters;
    steps_done += n;
    if (last.status) {
        std::string what;
        if (last.status & STATUS_PAIR_OVERFLOW) what += " pair table overflow (raise pairs_per_collider);";
        if (last.status & STATUS_COLOR_OVERFLOW) what += " more than 64 solver colours;";
        if (last.status & STATUS_STACK_OVERFLOW) what += " broadphase traversal stack overflow;";
        if (last.status & STATUS_COLOR_STALL) what += " colouring did not finish;";
        if (last.status & STATUS_CLIP_OVERFLOW) what += " face clipping exceeded its buffer;";
        set_error("device status " + std::to_string(last.status) + ":" + what);
        return RP3DG_EOVERFLOW;
    }
    hint_pairs = last.nb_pairs;
    contacts_np = last.nb_pairs;
    hint_manifolds = last.nb_manifolds;
    int used = 0;
    while (used < kColorRoundsMax && last.color_left[used] != 0) used++;
    hint_rounds = used + 2;
    return RP3DG_OK;
}

int rp3dg_batch::step(float dt) {
    if (nb == 0) return RP3DG_OK;
    int r = flush_api();
    if (r) return r;
    r = enqueue_step(dt, false);
    if (r) return r;
    return finish_steps(1);
}

void rp3dg_batch::drop_graphs() {
    for (int p = 0; p < 2; p++) {
        if (graph_exec[p]) hipGraphExecDestroy(graph_exec[p]);
        if (graph[p]) hipGraphDestroy(graph[p]);
        graph_exec[p] = nullptr;
        graph[p] = nullptr;
    }
}

// The pair tables alternate between steps, so two graphs are captured (one per parity) and replayed
// alternately.  Capture happens once per time-step value; the colouring round count is frozen at
// capture (the finisher completes any remainder).
int rp3dg_batch::step_graph(float dt, int n) {
    if (nb == 0 || n <= 0) return RP3DG_OK;
    if (!fused) { set_error("graph capture needs the fused solver layout"); return RP3DG_EUNSUPPORTED; }
    if (profiling) { set_error("kernel timing is eager-only: disable it before graph replay"); return RP3DG_ESTATE; }
    const int fr = flush_api();
    if (fr) return fr;
    if (ev_pending) {   // the exits a world query recorded open the next step's events: one eager step
        const int r0 = step(dt);
        if (r0 || --n == 0) return r0;
    }
    if (graph_exec[0] && graph_dt != dt) drop_graphs();
    if (!graph_exec[0]) {
        graph_rounds = std::min(kColorRoundsMax, std::max(8, hint_rounds + 4));
        const bool start = cur_is_a;
        for (int p = 0; p < 2; p++) {
            const int parity = start == (p == 0) ? 0 : 1;   // parity 0: a -> b
            HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            capture_parity = parity;
            const int r = enqueue_step(dt, true);
            capture_parity = -1;
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(stream, &g);
            if (r || e != hipSuccess) {
                if (g) hipGraphDestroy(g);
                cur_is_a = start;
                drop_graphs();
                if (r) return r;
                HIPCHECK(e);
            }
            graph[parity] = g;
            HIPCHECK(hipGraphInstantiate(&graph_exec[parity], g, nullptr, nullptr, 0));
        }
        cur_is_a = start;   // capture does not run anything; the flips above are undone
        graph_dt = dt;
    }
    for (int i = 0; i < n; i++) {
        HIPCHECK(hipGraphLaunch(graph_exec[cur_is_a ? 0 : 1], stream));
        cur_is_a = !cur_is_a;
    }
    contacts_is_a = cur_is_a;
    return finish_steps(n);
}

// A world query (PhysicsWorld::testOverlap / testCollision, CD.cpp:1775-1896): the broadphase of the
// current state, as the reference's computeBroadPhase runs it (the pair table changes and the moved set
// empties for the next step too), then the narrowphase of the pairs the query selects, its contacts kept
// apart in the q_* arrays: q_ncp[k] > 0 points, -1 colliding without points, 0 not colliding.
int rp3dg_batch::query(int mode, int report, int body1, int body2) {
    if (nb == 0) return RP3DG_OK;
    if (has_concave) {
        set_error("world queries (testOverlap / testCollision) do not take batches with concave shapes yet");
        return RP3DG_EUNSUPPORTED;
    }
    int r = flush_api();
    if (r) return r;
    if (!q_ncp) {
        q_ncp = dalloc<int>(this, pair_cap);
        q_pn = dalloc<float4>(this, 4 * (size_t)pair_cap);
        q_pl1 = dalloc<float4>(this, 4 * (size_t)pair_cap);
        q_pl2 = dalloc<float4>(this, 4 * (size_t)pair_cap);
        q_fv1 = dalloc<float4>(this, pair_cap);
        q_fv2 = dalloc<float4>(this, pair_cap);
        q_fimp = dalloc<float4>(this, pair_cap);
        if (!q_fimp) { set_error("device allocation failed (query contacts)"); return RP3DG_EHIP; }
    }
    if (events_on && !ev_pend) {
        ev_pend = dalloc<int4>(this, (size_t)ev_cap);
        ev_pend_count = dalloc<int>(this, 1);
        if (!ev_pend_count) { set_error("device allocation failed (query events)"); return RP3DG_EHIP; }
        HIPCHECK(hipMemset(ev_pend_count, 0, sizeof(int)));
    }
    fill_int<<<1, 256, 0, stream>>>((int*)counters, (int)(sizeof(DevCounters) / sizeof(int)), 0);
    const int pg = grid_for(std::min(pair_cap, hint_pairs + hint_pairs / 4 + 1024));
    DevEvents E{ev_pend, ev_cap, ev_pend_count, events_on ? 1 : 0};
    r = enqueue_broadphase(false, E, pg);
    if (r) return r;
    if (events_on) ev_pending = true;
    fill_int<<<grid_for(nc), 256, 0, stream>>>(C.moved, nc, 0);   // BroadPhaseSystem: mMovedShapes.clear()
    DevPairs& N = cur_is_a ? Pa : Pb;
    DevPairs& O = cur_is_a ? Pb : Pa;
    DevPairs Q = N;   // the new table, with the query's contacts written apart
    Q.ncp = q_ncp; Q.pn = q_pn; Q.pl1 = q_pl1; Q.pl2 = q_pl2; Q.fv1 = q_fv1; Q.fv2 = q_fv2; Q.fimp = q_fimp;
    r = enqueue_narrowphase(np_args(Q, O, mode, report, body1, body2), false, pg);
    if (r) return r;
    qr_carry<<<pg, 256, 0, stream>>>(N, O, &counters->nb_pairs);
    HIPCHECK(hipGetLastError());
    DevCounters c{};
    HIPCHECK(hipMemcpyAsync(&c, counters, sizeof(DevCounters), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (c.status) {
        std::string what;
        if (c.status & STATUS_PAIR_OVERFLOW) what += " pair table overflow (raise pairs_per_collider);";
        if (c.status & STATUS_STACK_OVERFLOW) what += " broadphase traversal stack overflow;";
        if (c.status & STATUS_CLIP_OVERFLOW) what += " face clipping exceeded its buffer;";
        set_error("world query: device status " + std::to_string(c.status) + ":" + what);
        return RP3DG_EOVERFLOW;
    }
    last.nb_pairs = c.nb_pairs;
    hint_pairs = std::max(hint_pairs, c.nb_pairs);
    return RP3DG_OK;
}

// Applies the queued RigidBody calls on the device, in call order (api.hip), on the batch's stream.
int rp3dg_batch::flush_api() {
    if (api_ops.empty()) return RP3DG_OK;
    const size_t n = api_ops.size();
    if (n > api_cap) {
        if (d_api_ops) HIPCHECK(hipFree(d_api_ops));
        d_api_ops = nullptr;
        api_cap = std::max<size_t>(256, 2 * n);
        HIPCHECK(hipMalloc(&d_api_ops, sizeof(ApiOp) * api_cap));
    }
    HIPCHECK(hipMemcpyAsync(d_api_ops, api_ops.data(), sizeof(ApiOp) * n, hipMemcpyHostToDevice, stream));
    ApiArgs A{B, C, d_hbounds, d_coll_offs, d_coll_list, settings.sleeping_enabled ? isl.sleep_time : nullptr};
    api_apply_ops<<<1, 1, 0, stream>>>(A, d_api_ops, (int)n);
    HIPCHECK(hipGetLastError());
    if (api_sleep_op) {   // pairs of two disabled bodies restart their LastFrameCollisionInfo
        const DevPairs& P = cur_is_a ? Pa : Pb;
        isl_reset_disabled_pairs<<<grid_for(std::max(1, last.nb_pairs)), 256, 0, stream>>>(B, C, P, pair_count + (cur_is_a ? 0 : 1));
        HIPCHECK(hipGetLastError());
    }
    // the host copy stays alive until the copy has been consumed
    HIPCHECK(hipStreamSynchronize(stream));
    api_ops.clear();
    api_sleep_op = false;
    return RP3DG_OK;
}

rp3dg_batch::~rp3dg_batch() {
    drop_graphs();
    for (hipStream_t s : side_streams) hipStreamDestroy(s);
    for (hipEvent_t e : fork_events) hipEventDestroy(e);
    if (log_feats) hipFree(log_feats);
    if (log_targets) hipFree(log_targets);
    if (ray_buf) hipFree(ray_buf);
    if (d_api_ops) hipFree(d_api_ops);
    for (hipEvent_t ev : event_pool) hipEventDestroy(ev);
    for (void* p : allocs) hipFree(p);
    if (hcounters) hipHostFree(hcounters);
    if (stream) hipStreamDestroy(stream);
}

extern "C" {

rp3dg_batch* rp3dg_batch_create_ex(rp3dg_world* const* worlds, int32_t n, const rp3dg_batch_options* opt) {
    if (!worlds || n <= 0) { set_error("no worlds"); return nullptr; }
    for (int i = 0; i < n; i++) {
        if (!worlds[i]) { set_error("null world"); return nullptr; }
        if (worlds[i]->uploaded) { set_error("a world belongs to a batch already"); return nullptr; }
    }
    static std::atomic<uint64_t> next_uid{1};
    rp3dg_batch* b = new rp3dg_batch();
    b->uid = next_uid++;
    const int r = b->upload(worlds, n, opt);
    if (r) { delete b; return nullptr; }
    for (int i = 0; i < n; i++) worlds[i]->uploaded = true;
    return b;
}

rp3dg_batch* rp3dg_batch_create(rp3dg_world* const* worlds, int32_t n) { return rp3dg_batch_create_ex(worlds, n, nullptr); }

void rp3dg_batch_destroy(rp3dg_batch* b) { delete b; }

int rp3dg_batch_step(rp3dg_batch* b, float dt, int32_t nsteps) {
    if (!b) return RP3DG_EINVAL;
    if (!(dt > 0.f)) { set_error("time step must be > 0"); return RP3DG_EINVAL; }
    for (int i = 0; i < nsteps; i++) {
        const int r = b->step(dt);
        if (r) return r;
    }
    return RP3DG_OK;
}

int rp3dg_batch_synchronize(rp3dg_batch* b) {
    if (!b) return RP3DG_EINVAL;
    const int r = b->flush_api();
    if (r) return r;
    HIPCHECK(hipStreamSynchronize(b->stream));
    return RP3DG_OK;
}

// Snapshot: a device copy of every array the batch allocated (its whole simulation state lives in
// them, the step counters included), and the host-side state of the step sequence.  Arrays allocated after the snapshot (contact
// events or sample logging enabled later) are left as they are by a restore.
struct rp3dg_snapshot {
    uint64_t batch_uid = 0;
    size_t nallocs = 0;
    std::vector<size_t> offs, sizes;
    char* dev = nullptr;
    size_t bytes = 0;
    DevCounters last{};
    bool cur_is_a = true, contacts_is_a = true, ev_pending = false, query_points = false;
    int contacts_np = 0, hint_pairs = 0, hint_manifolds = 0, hint_rounds = 0;
    int64_t steps_done = 0;
    std::vector<int> body_type, coll_kind;
};

static int sync_all_streams(rp3dg_batch* b) {
    HIPCHECK(hipStreamSynchronize(b->stream));
    for (hipStream_t st : b->side_streams) HIPCHECK(hipStreamSynchronize(st));
    return RP3DG_OK;
}

int rp3dg_batch_snapshot(rp3dg_batch* b, rp3dg_snapshot** out) {
    if (!b || !out) { set_error("null batch or output"); return RP3DG_EINVAL; }
    *out = nullptr;
    if (b->capture_parity >= 0) { set_error("snapshot while a step graph is being captured"); return RP3DG_ESTATE; }
    int r = b->flush_api();
    if (r) return r;
    r = sync_all_streams(b);
    if (r) return r;
    if (b->alloc_bytes.size() != b->allocs.size()) { set_error("allocation sizes not tracked"); return RP3DG_EINVAL; }
    rp3dg_snapshot* s = new rp3dg_snapshot();
    s->batch_uid = b->uid;
    s->nallocs = b->allocs.size();
    for (size_t i = 0; i < s->nallocs; i++) {
        s->offs.push_back(s->bytes);
        s->sizes.push_back(b->alloc_bytes[i]);
        s->bytes += (b->alloc_bytes[i] + 255) & ~(size_t)255;
    }
    if (hipMalloc((void**)&s->dev, std::max<size_t>(s->bytes, 1)) != hipSuccess) {
        (void)hipGetLastError();
        delete s;
        set_error("device allocation failed (snapshot of " + std::to_string(b->bytes) + " bytes)");
        return RP3DG_EHIP;
    }
    for (size_t i = 0; i < s->nallocs; i++)
        if (hipMemcpyAsync(s->dev + s->offs[i], b->allocs[i], s->sizes[i], hipMemcpyDeviceToDevice, b->stream) != hipSuccess) {
            rp3dg_snapshot_destroy(s);
            set_error("snapshot copy failed");
            return RP3DG_EHIP;
        }
    if (hipStreamSynchronize(b->stream) != hipSuccess) {
        rp3dg_snapshot_destroy(s);
        set_error("snapshot copy failed");
        return RP3DG_EHIP;
    }
    s->last = b->last;
    s->cur_is_a = b->cur_is_a; s->contacts_is_a = b->contacts_is_a;
    s->ev_pending = b->ev_pending; s->query_points = b->query_points;
    s->contacts_np = b->contacts_np;
    s->hint_pairs = b->hint_pairs; s->hint_manifolds = b->hint_manifolds; s->hint_rounds = b->hint_rounds;
    s->steps_done = b->steps_done;
    s->body_type = b->body_type; s->coll_kind = b->coll_kind;
    *out = s;
    return RP3DG_OK;
}

int rp3dg_batch_restore(rp3dg_batch* b, const rp3dg_snapshot* s) {
    if (!b || !s) { set_error("null batch or snapshot"); return RP3DG_EINVAL; }
    if (s->batch_uid != b->uid || s->nallocs > b->allocs.size()) {
        set_error("the snapshot was taken from another batch");
        return RP3DG_EINVAL;
    }
    for (size_t i = 0; i < s->nallocs; i++)
        if (s->sizes[i] != b->alloc_bytes[i]) { set_error("snapshot layout differs from the batch"); return RP3DG_EINVAL; }
    if (b->capture_parity >= 0) { set_error("restore while a step graph is being captured"); return RP3DG_ESTATE; }
    int r = sync_all_streams(b);
    if (r) return r;
    b->api_ops.clear();   // calls queued since the snapshot belong to the state it replaces
    b->api_sleep_op = false;
    for (size_t i = 0; i < s->nallocs; i++) {
        if (b->allocs[i] == (void*)b->log_count) continue;   // the sample log is not simulation state
        HIPCHECK(hipMemcpyAsync(b->allocs[i], s->dev + s->offs[i], s->sizes[i], hipMemcpyDeviceToDevice, b->stream));
    }
    HIPCHECK(hipStreamSynchronize(b->stream));
    b->last = s->last;
    b->cur_is_a = s->cur_is_a; b->contacts_is_a = s->contacts_is_a;
    b->ev_pending = s->ev_pending; b->query_points = s->query_points;
    b->contacts_np = s->contacts_np;
    b->hint_pairs = s->hint_pairs; b->hint_manifolds = s->hint_manifolds; b->hint_rounds = s->hint_rounds;
    b->steps_done = s->steps_done;
    b->body_type = s->body_type; b->coll_kind = s->coll_kind;
    for (rp3dg_world* w : b->worlds) w->mirror_dirty = true;
    return RP3DG_OK;
}

int64_t rp3dg_snapshot_bytes(const rp3dg_snapshot* s) { return s ? (int64_t)s->bytes : 0; }

void rp3dg_snapshot_destroy(rp3dg_snapshot* s) {
    if (!s) return;
    if (s->dev) (void)hipFree(s->dev);
    delete s;
}

rp3dg_batch* rp3dg_world_upload(rp3dg_world* w) {
    if (!w) { set_error("null world"); return nullptr; }
    if (!w->own) {
        if (w->uploaded) { set_error("world belongs to a batch; step the batch"); return nullptr; }
        rp3dg_world* ws[1] = {w};
        w->own = rp3dg_batch_create(ws, 1);
        if (!w->own) return nullptr;
        if (w->events && rp3dg_batch_set_contact_events(w->own, 1)) return nullptr;
        w->mirror_dirty = true;
    }
    return w->own;
}

int rp3dg_world_update(rp3dg_world* w, float dt) {
    if (!w) return RP3DG_EINVAL;
    if (!w->own && !rp3dg_world_upload(w)) return w->uploaded ? RP3DG_ESTATE : RP3DG_EHIP;
    const int r = rp3dg_batch_step(w->own, dt, 1);
    w->mirror_dirty = true;
    return r;
}

int rp3dg_world_get_body_state(rp3dg_world* w, int32_t body, float* pos, float* ori, float* lin, float* ang) {
    if (!w || body < 0 || body >= (int32_t)w->bodies.size()) { set_error("invalid body index"); return RP3DG_EINVAL; }
    if (!w->own) {   // not stepped yet: the creation state
        const HostBody& b = w->bodies[body];
        if (pos) { pos[0] = b.transform.p.x; pos[1] = b.transform.p.y; pos[2] = b.transform.p.z; }
        if (ori) { ori[0] = b.transform.q.x; ori[1] = b.transform.q.y; ori[2] = b.transform.q.z; ori[3] = b.transform.q.w; }
        if (lin) { lin[0] = b.v.x; lin[1] = b.v.y; lin[2] = b.v.z; }
        if (ang) { ang[0] = b.w.x; ang[1] = b.w.y; ang[2] = b.w.z; }
        return RP3DG_OK;
    }
    if (w->mirror_dirty || w->mirror_pos.empty()) {
        const size_t n = w->bodies.size();
        w->mirror_pos.resize(3 * n); w->mirror_q.resize(4 * n); w->mirror_v.resize(3 * n); w->mirror_w.resize(3 * n);
        int r = rp3dg_batch_get_bodies(w->own, w->mirror_pos.data(), w->mirror_q.data(), w->mirror_v.data(),
                                       w->mirror_w.data());
        if (r) return r;
        w->mirror_sleeping.resize(n);
        r = rp3dg_batch_get_sleeping(w->own, w->mirror_sleeping.data(), (int64_t)n);
        if (r) return r;
        w->mirror_dirty = false;
    }
    for (int k = 0; k < 3; k++) {
        if (pos) pos[k] = w->mirror_pos[3 * body + k];
        if (lin) lin[k] = w->mirror_v[3 * body + k];
        if (ang) ang[k] = w->mirror_w[3 * body + k];
    }
    for (int k = 0; k < 4; k++)
        if (ori) ori[k] = w->mirror_q[4 * body + k];
    return RP3DG_OK;
}

int rp3dg_world_set_contact_events(rp3dg_world* w, int32_t enable) {
    if (!w) return RP3DG_EINVAL;
    w->events = enable != 0;
    return w->own ? rp3dg_batch_set_contact_events(w->own, enable) : RP3DG_OK;
}

rp3dg_batch* rp3dg_world_get_batch(rp3dg_world* w) { return w ? w->own : nullptr; }

int rp3dg_world_get_body_sleeping(rp3dg_world* w, int32_t body) {
    if (!w || body < 0 || body >= (int32_t)w->bodies.size()) { set_error("invalid body index"); return RP3DG_EINVAL; }
    if (!w->own) return 0;   // bodies are created awake
    const int r = rp3dg_world_get_body_state(w, body, nullptr, nullptr, nullptr, nullptr);
    if (r) return r;
    return w->mirror_sleeping[body];
}

void rp3dg_world_destroy(rp3dg_world* w) {
    if (!w) return;
    if (w->own) rp3dg_batch_destroy(w->own);
    delete w;
}

int rp3dg_batch_get_counts(rp3dg_batch* b, rp3dg_counts* o) {
    if (!b || !o) return RP3DG_EINVAL;
    o->nb_bodies = b->nb;
    o->nb_colliders = b->nc;
    o->nb_joints = b->J.n;
    o->nb_worlds = b->nworlds;
    o->nb_pairs = b->last.nb_pairs;
    o->nb_manifolds = b->last.nb_manifolds;
    o->nb_contact_points = b->last.nb_points;
    o->nb_colors = b->last.nb_colors;
    o->status = b->last.status;
    for (int a = 0; a < 7; a++) o->nb_tests[a] = b->last.nb_tests[a];
    unsigned long long tp = 0;
    HIPCHECK(hipMemcpyAsync(&tp, b->total_points, sizeof(tp), hipMemcpyDeviceToHost, b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
    o->total_contact_points = (int64_t)tp;
    o->steps = b->steps_done;
    o->nb_new_points = b->last.nb_new_points;
    o->nb_box_tests = b->last.nb_box_tests + b->last.nb_box_sep_tests;
    return RP3DG_OK;
}

int rp3dg_batch_set_contact_events(rp3dg_batch* b, int32_t enable) {
    if (!b) return RP3DG_EINVAL;
    if (enable && !b->ev_buf) {
        b->ev_cap = 2 * b->pair_cap;   // <= one event per pair of the new table + one per removed pair
        b->ev_buf = dalloc<int4>(b, (size_t)b->ev_cap);
        if (!b->ev_buf) { set_error("device allocation failed (events)"); return RP3DG_EHIP; }
    }
    if ((enable != 0) != b->events_on) b->drop_graphs();
    b->events_on = enable != 0;
    return RP3DG_OK;
}

int64_t rp3dg_batch_get_contact_events(rp3dg_batch* b, int32_t* c1, int32_t* c2, int32_t* type, int32_t* nb_points,
                                       int64_t capacity) {
    if (!b) return RP3DG_EINVAL;
    if (!b->events_on) { set_error("contact events are off: rp3dg_batch_set_contact_events(batch, 1)"); return RP3DG_ESTATE; }
    const int n = b->last.nb_events;
    if (n > b->ev_cap) { set_error("contact event buffer overflow"); return RP3DG_EOVERFLOW; }
    if (!c1 && !c2 && !type && !nb_points) return n;   // count query
    if (capacity < n) { set_error("capacity below the event count"); return RP3DG_EINVAL; }
    std::vector<int4> e(std::max(n, 1));
    HIPCHECK(hipStreamSynchronize(b->stream));
    if (n) HIPCHECK(hipMemcpy(e.data(), b->ev_buf, sizeof(int4) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
        if (c1) c1[i] = e[i].x;
        if (c2) c2[i] = e[i].y;
        if (type) type[i] = e[i].z;
        if (nb_points) nb_points[i] = e[i].w;
    }
    return n;
}

int rp3dg_batch_apply_forces(rp3dg_batch* b, const float* force, const float* torque) {
    if (!b) return RP3DG_EINVAL;
    if (!force && !torque) return RP3DG_OK;
    const int fr = b->flush_api();
    if (fr) return fr;
    const int n = b->nb;
    bool woke = false;
    std::vector<float4> f(n), t(n);
    std::vector<int> fl(n);
    std::vector<float> st;
    HIPCHECK(hipStreamSynchronize(b->stream));
    HIPCHECK(hipMemcpy(f.data(), b->B.force, sizeof(float4) * n, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(t.data(), b->B.torque, sizeof(float4) * n, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(fl.data(), b->B.flags, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (b->settings.sleeping_enabled) {
        st.resize(n);
        HIPCHECK(hipMemcpy(st.data(), b->isl.sleep_time, sizeof(float) * n, hipMemcpyDeviceToHost));
    }
    auto nonzero = [](const float* a, int i) { return a && (a[3 * i] != 0.f || a[3 * i + 1] != 0.f || a[3 * i + 2] != 0.f); };
    for (int i = 0; i < n; i++) {
        if (!(fl[i] & BF_DYNAMIC)) continue;   // only dynamic bodies take forces (RigidBody.cpp:298-311)
        const bool nf = nonzero(force, i), nt = nonzero(torque, i);
        if (!nf && !nt) continue;
        if (!(fl[i] & BF_ENABLED)) {   // asleep: setIsSleeping(false)
            fl[i] |= BF_ENABLED | BF_WOKEN;
            st[i] = 0.0f;
            woke = true;
        }
        if (nf) { f[i].x += force[3 * i]; f[i].y += force[3 * i + 1]; f[i].z += force[3 * i + 2]; }
        if (nt) { t[i].x += torque[3 * i]; t[i].y += torque[3 * i + 1]; t[i].z += torque[3 * i + 2]; }
    }
    HIPCHECK(hipMemcpy(b->B.force, f.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(b->B.torque, t.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(b->B.flags, fl.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    if (!st.empty()) HIPCHECK(hipMemcpy(b->isl.sleep_time, st.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    if (woke) {   // askForBroadPhaseCollisionCheck (RigidBody.cpp:1045-1051): the moved set takes their colliders
        bp_refit<<<grid_for(b->nc), 256, 0, b->stream>>>(b->B, b->C, b->d_hbounds, 2);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(b->stream));
    }
    return RP3DG_OK;
}

int rp3dg_batch_raycast(rp3dg_batch* b, int64_t n, const int32_t* world, const float* point1, const float* point2,
                        const float* max_fraction, uint16_t category_mask, int32_t* collider, float* point,
                        float* normal, float* fraction) {
    if (!b || n < 0 || !point1 || !point2) return RP3DG_EINVAL;
    if (n == 0) return RP3DG_OK;
    if (n > (1 << 26)) { set_error("at most 2^26 rays per call"); return RP3DG_EINVAL; }
    if (b->has_concave) { set_error("raycasts do not take batches with concave shapes yet"); return RP3DG_EUNSUPPORTED; }
    if (const int fr = b->flush_api()) return fr;
    std::vector<int> wv((size_t)n);
    std::vector<float4> a((size_t)n), c((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        const int w = world ? world[i] : 0;
        if (w < 0 || w >= b->nworlds) { set_error("raycast: invalid world index"); return RP3DG_EINVAL; }
        const float mf = max_fraction ? max_fraction[i] : 1.0f;
        wv[i] = w;
        a[i] = make_float4(point1[3 * i], point1[3 * i + 1], point1[3 * i + 2], 0.f);
        c[i] = make_float4(point2[3 * i], point2[3 * i + 1], point2[3 * i + 2], mf);
    }
    // staging: p1, p2, hit point, hit normal (float4 each), world, hit collider (int each)
    const size_t need = (size_t)n * (4 * sizeof(float4) + 2 * sizeof(int));
    if (need > b->ray_cap) {
        if (b->ray_buf) HIPCHECK(hipFree(b->ray_buf));
        b->ray_buf = nullptr;
        b->ray_cap = 0;
        HIPCHECK(hipMalloc(&b->ray_buf, need));
        b->ray_cap = need;
    }
    float4* dp1 = (float4*)b->ray_buf;
    float4* dp2 = dp1 + n;
    float4* dhp = dp2 + n;
    float4* dhn = dhp + n;
    int* dw = (int*)(dhn + n);
    int* dhc = dw + n;
    HIPCHECK(hipMemcpyAsync(dp1, a.data(), sizeof(float4) * n, hipMemcpyHostToDevice, b->stream));
    HIPCHECK(hipMemcpyAsync(dp2, c.data(), sizeof(float4) * n, hipMemcpyHostToDevice, b->stream));
    HIPCHECK(hipMemcpyAsync(dw, wv.data(), sizeof(int) * n, hipMemcpyHostToDevice, b->stream));
    RayArgs A{b->C, b->d_hulls, b->d_hverts, b->d_hnormals, b->d_hfverts, b->d_hfvidx, b->d_wcoll};
    rc_closest<<<grid_for((int)n), 256, 0, b->stream>>>(A, dw, dp1, dp2, (int)n, category_mask, dhc, dhp, dhn);
    HIPCHECK(hipGetLastError());
    std::vector<int> hc((size_t)n);
    HIPCHECK(hipMemcpyAsync(hc.data(), dhc, sizeof(int) * n, hipMemcpyDeviceToHost, b->stream));
    HIPCHECK(hipMemcpyAsync(a.data(), dhp, sizeof(float4) * n, hipMemcpyDeviceToHost, b->stream));
    HIPCHECK(hipMemcpyAsync(c.data(), dhn, sizeof(float4) * n, hipMemcpyDeviceToHost, b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
    for (int64_t i = 0; i < n; i++) {
        if (collider) collider[i] = hc[i];
        if (point) { point[3 * i] = a[i].x; point[3 * i + 1] = a[i].y; point[3 * i + 2] = a[i].z; }
        if (normal) { normal[3 * i] = c[i].x; normal[3 * i + 1] = c[i].y; normal[3 * i + 2] = c[i].z; }
        if (fraction) fraction[i] = a[i].w;
    }
    return RP3DG_OK;
}

// PhysicsWorld::raycast with a RaycastCallback whose notifyRaycastHit returns 1 (every hit, the ray is
// not clipped): the hits of all rays, sorted by (ray, collider).
int64_t rp3dg_batch_raycast_all(rp3dg_batch* b, int64_t n, const int32_t* world, const float* point1,
                                const float* point2, const float* max_fraction, uint16_t category_mask, int32_t* ray,
                                int32_t* collider, float* point, float* normal, float* fraction, int64_t capacity) {
    if (!b || n < 0 || !point1 || !point2 || capacity < 0) { set_error("raycast_all: invalid arguments"); return RP3DG_EINVAL; }
    if (n == 0) return 0;
    if (b->has_concave) { set_error("raycasts do not take batches with concave shapes yet"); return RP3DG_EUNSUPPORTED; }
    if (n > (1 << 26)) { set_error("at most 2^26 rays per call"); return RP3DG_EINVAL; }
    if (const int fr = b->flush_api()) return fr;
    std::vector<int> wv((size_t)n);
    std::vector<float4> a((size_t)n), c((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        const int w = world ? world[i] : 0;
        if (w < 0 || w >= b->nworlds) { set_error("raycast: invalid world index"); return RP3DG_EINVAL; }
        wv[i] = w;
        a[i] = make_float4(point1[3 * i], point1[3 * i + 1], point1[3 * i + 2], 0.f);
        c[i] = make_float4(point2[3 * i], point2[3 * i + 1], point2[3 * i + 2], max_fraction ? max_fraction[i] : 1.0f);
    }
    // staging: rays (2 float4 + world) and up to hcap hits (int2 + 2 float4) + the hit count
    const int64_t hcap = std::min<int64_t>(std::max<int64_t>(capacity, 1), 1 << 26);
    const size_t need = (size_t)n * (2 * sizeof(float4) + sizeof(int)) + (size_t)hcap * (sizeof(int2) + 2 * sizeof(float4)) + 64;
    if (need > b->ray_cap) {
        if (b->ray_buf) HIPCHECK(hipFree(b->ray_buf));
        b->ray_buf = nullptr;
        b->ray_cap = 0;
        HIPCHECK(hipMalloc(&b->ray_buf, need));
        b->ray_cap = need;
    }
    float4* dp1 = (float4*)b->ray_buf;
    float4* dp2 = dp1 + n;
    float4* dhp = dp2 + n;
    float4* dhn = dhp + hcap;
    int2* drc = (int2*)(dhn + hcap);
    int* dw = (int*)(drc + hcap);
    int* dcount = dw + n;
    HIPCHECK(hipMemcpyAsync(dp1, a.data(), sizeof(float4) * n, hipMemcpyHostToDevice, b->stream));
    HIPCHECK(hipMemcpyAsync(dp2, c.data(), sizeof(float4) * n, hipMemcpyHostToDevice, b->stream));
    HIPCHECK(hipMemcpyAsync(dw, wv.data(), sizeof(int) * n, hipMemcpyHostToDevice, b->stream));
    fill_int<<<1, 64, 0, b->stream>>>(dcount, 1, 0);
    RayArgs A{b->C, b->d_hulls, b->d_hverts, b->d_hnormals, b->d_hfverts, b->d_hfvidx, b->d_wcoll};
    rc_all<<<grid_for((int)n), 256, 0, b->stream>>>(A, dw, dp1, dp2, (int)n, category_mask, drc, dhp, dhn, (int)hcap,
                                                     dcount);
    HIPCHECK(hipGetLastError());
    int total = 0;
    HIPCHECK(hipMemcpyAsync(&total, dcount, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
    if (total > capacity) {
        set_error("raycast_all: " + std::to_string(total) + " hits, capacity " + std::to_string(capacity));
        return capacity == 0 && !ray && !collider ? total : RP3DG_EINVAL;
    }
    std::vector<int2> rc((size_t)total);
    std::vector<float4> hp((size_t)total), hn((size_t)total);
    if (total) {
        HIPCHECK(hipMemcpy(rc.data(), drc, sizeof(int2) * total, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(hp.data(), dhp, sizeof(float4) * total, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(hn.data(), dhn, sizeof(float4) * total, hipMemcpyDeviceToHost));
    }
    std::vector<int> order((size_t)total);
    for (int i = 0; i < total; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        return rc[x].x != rc[y].x ? rc[x].x < rc[y].x : rc[x].y < rc[y].y;
    });
    for (int k = 0; k < total; k++) {
        const int i = order[k];
        if (ray) ray[k] = rc[i].x;
        if (collider) collider[k] = rc[i].y;
        if (point) { point[3 * k] = hp[i].x; point[3 * k + 1] = hp[i].y; point[3 * k + 2] = hp[i].z; }
        if (normal) { normal[3 * k] = hn[i].x; normal[3 * k + 1] = hn[i].y; normal[3 * k + 2] = hn[i].z; }
        if (fraction) fraction[k] = hp[i].w;
    }
    return total;
}

// Results of the last world query: the colliding pairs (q_ncp != 0) of the current table with their
// contact points.
static int query_results(rp3dg_batch* b, std::vector<int2>& col, std::vector<int>& ncp, std::vector<int>& hits) {
    const DevPairs& P = b->cur_is_a ? b->Pa : b->Pb;
    const int np = b->last.nb_pairs;
    col.resize(std::max(np, 1));
    ncp.resize(std::max(np, 1));
    if (np) {
        HIPCHECK(hipMemcpy(col.data(), P.col, sizeof(int2) * np, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(ncp.data(), b->q_ncp, sizeof(int) * np, hipMemcpyDeviceToHost));
    }
    hits.clear();
    for (int k = 0; k < np; k++) {
        if (ncp[k] == 0) continue;
        // CollisionCallback::CallbackData keeps the pairs without a trigger (CollisionCallback.cpp:60-70)
        if (b->query_points && ((b->coll_kind[col[k].x] | b->coll_kind[col[k].y]) & CK_TRIGGER)) continue;
        hits.push_back(k);
    }
    return RP3DG_OK;
}

static bool valid_query_body(rp3dg_batch* b, int32_t body, bool allow_none) {
    if ((allow_none && body < 0) || (body >= 0 && body < b->nb)) return true;
    set_error("world query: invalid body index " + std::to_string(body));
    return false;
}

int rp3dg_batch_test_overlap_bodies(rp3dg_batch* b, int32_t body1, int32_t body2) {
    if (!b) return RP3DG_EINVAL;
    if (!valid_query_body(b, body1, false) || !valid_query_body(b, body2, false)) return RP3DG_EINVAL;
    int r = b->query(NP_QUERY_BODIES, NP_REPORT_NONE, body1, body2);
    if (r) return r;
    b->query_points = false;
    std::vector<int2> col;
    std::vector<int> ncp, hits;
    if ((r = query_results(b, col, ncp, hits))) return r;
    return hits.empty() ? 0 : 1;
}

int64_t rp3dg_batch_test_overlap(rp3dg_batch* b, int32_t body) {
    if (!b) return RP3DG_EINVAL;
    if (!valid_query_body(b, body, true)) return RP3DG_EINVAL;
    int r = body < 0 ? b->query(NP_QUERY_ALL, NP_REPORT_NONE, -1, -1) : b->query(NP_QUERY_BODIES, NP_REPORT_NONE, body, -1);
    if (r) return r;
    b->query_points = false;
    std::vector<int2> col;
    std::vector<int> ncp, hits;
    if ((r = query_results(b, col, ncp, hits))) return r;
    return (int64_t)hits.size();
}

int64_t rp3dg_batch_test_collision(rp3dg_batch* b, int32_t body1, int32_t body2) {
    if (!b) return RP3DG_EINVAL;
    if (!valid_query_body(b, body1, true) || !valid_query_body(b, body2, true)) return RP3DG_EINVAL;
    if (body1 < 0 && body2 >= 0) { set_error("test_collision: give body1 to filter by body2"); return RP3DG_EINVAL; }
    int r = body1 < 0 ? b->query(NP_QUERY_ALL, NP_REPORT_NO_TRIGGER, -1, -1)
                      : b->query(NP_QUERY_BODIES, NP_REPORT_ALL, body1, body2);
    if (r) return r;
    b->query_points = true;
    std::vector<int2> col;
    std::vector<int> ncp, hits;
    if ((r = query_results(b, col, ncp, hits))) return r;
    return (int64_t)hits.size();
}

int64_t rp3dg_batch_get_query_results(rp3dg_batch* b, rp3dg_contact_view* o, int64_t capacity) {
    if (!b || !o) return RP3DG_EINVAL;
    if (!b->q_ncp) { set_error("no world query has run on this batch"); return RP3DG_ESTATE; }
    std::vector<int2> col;
    std::vector<int> ncp, hits;
    int r = query_results(b, col, ncp, hits);
    if (r) return r;
    const int64_t n = (int64_t)hits.size();
    if (capacity < n) { set_error("query results: capacity below the pair count"); return RP3DG_EINVAL; }
    const int np = b->last.nb_pairs;
    std::vector<int> evf((size_t)std::max(np, 1));
    if (np) HIPCHECK(hipMemcpy(evf.data(), (b->cur_is_a ? b->Pa : b->Pb).evf, sizeof(int) * np, hipMemcpyDeviceToHost));
    std::vector<float4> pn(4 * (size_t)std::max(np, 1)), pl1(pn.size()), pl2(pn.size());
    if (np && b->query_points) {
        HIPCHECK(hipMemcpy(pn.data(), b->q_pn, sizeof(float4) * 4 * np, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(pl1.data(), b->q_pl1, sizeof(float4) * 4 * np, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(pl2.data(), b->q_pl2, sizeof(float4) * 4 * np, hipMemcpyDeviceToHost));
    }
    for (int64_t m = 0; m < n; m++) {
        const int k = hits[m];
        const int cnt = b->query_points ? std::max(ncp[k], 0) : 0;
        if (o->colliders) { o->colliders[2 * m] = col[k].x; o->colliders[2 * m + 1] = col[k].y; }
        if (o->nb_points) o->nb_points[m] = cnt;
        // ContactPair::collidingInPreviousFrame: the pair's flag in testCollision (CD.cpp:1232), false in
        // testOverlap's pairs (:831), which are therefore all OverlapStart
        if (o->color) o->color[m] = (b->query_points && (evf[k] & EV_PREV)) ? 1 : 0;
        for (int p = 0; p < 4; p++) {
            const bool in = p < cnt;
            const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
            const float4 a = in ? pn[4 * k + p] : z, c = in ? pl1[4 * k + p] : z, d = in ? pl2[4 * k + p] : z;
            if (o->normal_depth) { float* q = o->normal_depth + 16 * m + 4 * p; q[0] = a.x; q[1] = a.y; q[2] = a.z; q[3] = a.w; }
            if (o->local1) { float* q = o->local1 + 12 * m + 3 * p; q[0] = c.x; q[1] = c.y; q[2] = c.z; }
            if (o->local2) { float* q = o->local2 + 12 * m + 3 * p; q[0] = d.x; q[1] = d.y; q[2] = d.z; }
            if (o->impulse) o->impulse[4 * m + p] = 0.f;
        }
        if (o->friction) { o->friction[3 * m] = 0.f; o->friction[3 * m + 1] = 0.f; o->friction[3 * m + 2] = 0.f; }
    }
    return n;
}

int rp3dg_batch_get_sleeping(rp3dg_batch* b, int32_t* sleeping, int64_t capacity) {
    if (!b || !sleeping || capacity < b->nb) { set_error("sleeping: capacity below the body count"); return RP3DG_EINVAL; }
    if (const int fr = b->flush_api()) return fr;
    std::vector<int> f(b->nb);
    HIPCHECK(hipStreamSynchronize(b->stream));
    HIPCHECK(hipMemcpy(f.data(), b->B.flags, sizeof(int) * b->nb, hipMemcpyDeviceToHost));
    // a non-static body outside the enabled partition is asleep (RigidBody::isSleeping)
    for (int i = 0; i < b->nb; i++)
        sleeping[i] = ((f[i] & (BF_DYNAMIC | BF_KINEMATIC)) && !(f[i] & BF_ENABLED)) ? 1 : 0;
    return RP3DG_OK;
}

int rp3dg_batch_get_bodies(rp3dg_batch* b, float* pos, float* ori, float* lin, float* ang) {
    if (!b) return RP3DG_EINVAL;
    if (const int fr = b->flush_api()) return fr;
    std::vector<float4> t(b->nb);
    auto grab = [&](const float4* d, float* out, int comps) -> int {
        if (!out) return RP3DG_OK;
        HIPCHECK(hipMemcpy(t.data(), d, sizeof(float4) * b->nb, hipMemcpyDeviceToHost));
        for (int i = 0; i < b->nb; i++) {
            out[comps * i + 0] = t[i].x;
            out[comps * i + 1] = t[i].y;
            out[comps * i + 2] = t[i].z;
            if (comps == 4) out[comps * i + 3] = t[i].w;
        }
        return RP3DG_OK;
    };
    HIPCHECK(hipStreamSynchronize(b->stream));
    int r;
    if ((r = grab(b->B.p, pos, 3))) return r;
    if ((r = grab(b->B.q, ori, 4))) return r;
    if ((r = grab(b->B.v, lin, 3))) return r;
    if ((r = grab(b->B.w, ang, 3))) return r;
    return RP3DG_OK;
}

// RigidBody::setTransform / setLinearVelocity / setAngularVelocity for every body of the batch
// (src/body/RigidBody.cpp): the new origin/orientation, the centre of mass that follows from them,
// and the collider transforms and fat AABBs (refit, as the next step would).  NULL leaves a part as is.
int rp3dg_batch_set_bodies(rp3dg_batch* b, const float* pos, const float* ori, const float* lin, const float* ang) {
    if (!b) return RP3DG_EINVAL;
    int r = b->flush_api();
    if (r) return r;
    HIPCHECK(hipStreamSynchronize(b->stream));
    const int n = b->nb;
    std::vector<float4> t(n);
    // a static body keeps its transform: its colliders' static-static pairs are listed once at upload
    auto check_static = [&](const float4* d, const float* in, int comps) -> int {
        if (!in) return RP3DG_OK;
        HIPCHECK(hipMemcpy(t.data(), d, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++) {
            if (b->body_type[i] != RP3DG_STATIC) continue;
            const float cur[4] = {t[i].x, t[i].y, t[i].z, t[i].w};
            for (int k = 0; k < comps; k++)
                if (cur[k] != in[comps * i + k]) {
                    set_error("set_bodies: body " + std::to_string(i) + " is static; its transform is fixed after the first update");
                    return RP3DG_EUNSUPPORTED;
                }
        }
        return RP3DG_OK;
    };
    if ((r = check_static(b->B.p, pos, 3)) || (r = check_static(b->B.q, ori, 4))) return r;
    auto put = [&](float4* d, const float* in, int comps) -> int {
        if (!in) return RP3DG_OK;
        HIPCHECK(hipMemcpy(t.data(), d, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++) {
            t[i].x = in[comps * i]; t[i].y = in[comps * i + 1]; t[i].z = in[comps * i + 2];
            if (comps == 4) t[i].w = in[comps * i + 3];
        }
        HIPCHECK(hipMemcpy(d, t.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        return RP3DG_OK;
    };
    if ((r = put(b->B.p, pos, 3)) || (r = put(b->B.q, ori, 4)) || (r = put(b->B.v, lin, 3)) || (r = put(b->B.w, ang, 3)))
        return r;
    bool woke = false;
    if (b->settings.sleeping_enabled && (pos || ori || lin || ang)) {
        // setTransform wakes the body, setLinear/AngularVelocity do when the new velocity is not zero
        // (RigidBody.cpp); waking = setIsSleeping(false) (RigidBody.cpp:1006-1057)
        std::vector<int> f(n);
        std::vector<float> st(n);
        HIPCHECK(hipMemcpy(f.data(), b->B.flags, sizeof(int) * n, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(st.data(), b->isl.sleep_time, sizeof(float) * n, hipMemcpyDeviceToHost));
        auto nonzero = [](const float* a, int i) { return a && (a[3 * i] != 0.f || a[3 * i + 1] != 0.f || a[3 * i + 2] != 0.f); };
        for (int i = 0; i < n; i++) {
            const bool asleep = (f[i] & (BF_DYNAMIC | BF_KINEMATIC)) && !(f[i] & BF_ENABLED);
            if (asleep && (pos || ori || nonzero(lin, i) || nonzero(ang, i))) {
                f[i] |= BF_ENABLED | BF_WOKEN;
                st[i] = 0.0f;
                woke = true;
            }
        }
        HIPCHECK(hipMemcpy(b->B.flags, f.data(), sizeof(int) * n, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(b->isl.sleep_time, st.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    }
    if (pos || ori) set_body_com<<<grid_for(n), 256, 0, b->stream>>>(b->B);
    if (pos || ori || woke) {   // moved and woken bodies put their colliders into the moved set
        bp_refit<<<grid_for(b->nc), 256, 0, b->stream>>>(b->B, b->C, b->d_hbounds, 2);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(b->stream));
    }
    return RP3DG_OK;
}

// RigidBody calls on one body between steps: queued in call order, applied on the device (api.hip)
// before the next step, graph replay, read-back or raycast.
static int queue_body_op(rp3dg_batch* b, int32_t body, int op, int flags, const float* a, int na, const float* c, int nc) {
    if (!b) return RP3DG_EINVAL;
    if (body < 0 || body >= b->nb) { set_error("invalid body index " + std::to_string(body)); return RP3DG_EINVAL; }
    ApiOp o{};
    o.op = op;
    o.body = body;
    o.flags = flags;
    float va[4] = {0.f, 0.f, 0.f, 0.f}, vc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < na; k++) va[k] = a[k];
    for (int k = 0; k < nc; k++) vc[k] = c[k];
    o.a = make_float4(va[0], va[1], va[2], va[3]);
    o.b = make_float4(vc[0], vc[1], vc[2], vc[3]);
    b->api_ops.push_back(o);
    if (op == API_SET_SLEEPING && flags) b->api_sleep_op = true;
    return RP3DG_OK;
}

int rp3dg_batch_body_set_transform(rp3dg_batch* b, int32_t body, const float position[3], const float orientation[4]) {
    if (!b || !position || !orientation) { set_error("null argument"); return RP3DG_EINVAL; }
    if (body >= 0 && body < b->nb && b->body_type[body] == RP3DG_STATIC) {
        set_error("body " + std::to_string(body) + " is static; its transform is fixed after the first update");
        return RP3DG_EUNSUPPORTED;
    }
    return queue_body_op(b, body, API_SET_TRANSFORM, 0, position, 3, orientation, 4);
}

int rp3dg_batch_body_set_velocity(rp3dg_batch* b, int32_t body, const float linear[3], const float angular[3]) {
    int r = RP3DG_OK;
    if (linear) r = queue_body_op(b, body, API_SET_LINEAR_VELOCITY, 0, linear, 3, nullptr, 0);
    if (!r && angular) r = queue_body_op(b, body, API_SET_ANGULAR_VELOCITY, 0, angular, 3, nullptr, 0);
    return r;
}

int rp3dg_batch_body_apply_force(rp3dg_batch* b, int32_t body, const float force[3], const float point[3], int32_t flags) {
    if (!force) { set_error("null force"); return RP3DG_EINVAL; }
    if (flags & ~(RP3DG_FORCE_LOCAL | RP3DG_FORCE_AT_POINT | RP3DG_FORCE_POINT_LOCAL)) { set_error("unknown force flags"); return RP3DG_EINVAL; }
    if ((flags & RP3DG_FORCE_AT_POINT) && !point) { set_error("null point"); return RP3DG_EINVAL; }
    return queue_body_op(b, body, API_APPLY_FORCE, flags, force, 3, (flags & RP3DG_FORCE_AT_POINT) ? point : nullptr,
                         (flags & RP3DG_FORCE_AT_POINT) ? 3 : 0);
}

int rp3dg_batch_body_apply_torque(rp3dg_batch* b, int32_t body, const float torque[3], int32_t flags) {
    if (!torque) { set_error("null torque"); return RP3DG_EINVAL; }
    if (flags & ~RP3DG_FORCE_LOCAL) { set_error("unknown torque flags"); return RP3DG_EINVAL; }
    return queue_body_op(b, body, API_APPLY_TORQUE, flags, torque, 3, nullptr, 0);
}

int rp3dg_batch_body_reset_force_torque(rp3dg_batch* b, int32_t body, int32_t force, int32_t torque) {
    int r = RP3DG_OK;
    if (force) r = queue_body_op(b, body, API_RESET_FORCE, 0, nullptr, 0, nullptr, 0);
    if (!r && torque) r = queue_body_op(b, body, API_RESET_TORQUE, 0, nullptr, 0, nullptr, 0);
    return r;
}

int rp3dg_batch_body_set_sleeping(rp3dg_batch* b, int32_t body, int32_t sleeping) {
    if (b && !b->settings.sleeping_enabled && sleeping) {
        set_error("sleeping is disabled in the worlds of this batch");
        return RP3DG_ESTATE;
    }
    return queue_body_op(b, body, API_SET_SLEEPING, sleeping ? 1 : 0, nullptr, 0, nullptr, 0);
}

int rp3dg_batch_get_force_torque(rp3dg_batch* b, float* force, float* torque) {
    if (!b) return RP3DG_EINVAL;
    if (const int fr = b->flush_api()) return fr;
    std::vector<float4> t(b->nb);
    HIPCHECK(hipStreamSynchronize(b->stream));
    for (int part = 0; part < 2; part++) {
        float* out = part ? torque : force;
        if (!out) continue;
        HIPCHECK(hipMemcpy(t.data(), part ? b->B.torque : b->B.force, sizeof(float4) * b->nb, hipMemcpyDeviceToHost));
        for (int i = 0; i < b->nb; i++) { out[3 * i] = t[i].x; out[3 * i + 1] = t[i].y; out[3 * i + 2] = t[i].z; }
    }
    return RP3DG_OK;
}

int rp3dg_batch_get_pairs(rp3dg_batch* b, int32_t* c1, int32_t* c2, int64_t cap) {
    if (!b) return RP3DG_EINVAL;
    if (const int fr = b->flush_api()) return fr;
    const DevPairs& P = b->cur_is_a ? b->Pa : b->Pb;
    const int n = b->last.nb_pairs;
    if (cap < n) { set_error("capacity too small"); return RP3DG_EINVAL; }
    std::vector<int2> col(std::max(n, 1));
    HIPCHECK(hipStreamSynchronize(b->stream));
    if (n) HIPCHECK(hipMemcpy(col.data(), P.col, sizeof(int2) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) { c1[i] = col[i].x; c2[i] = col[i].y; }
    return n;
}

int rp3dg_batch_get_contacts(rp3dg_batch* b, rp3dg_contact_view* o, int64_t cap) {
    if (!b || !o) return RP3DG_EINVAL;
    const int nm = b->last.nb_manifolds;
    if (cap < nm) { set_error("capacity too small"); return RP3DG_EINVAL; }
    if (nm == 0) return 0;
    const DevPairs& P = b->contacts_is_a ? b->Pa : b->Pb;   // a world query since the step moved to the other table
    HIPCHECK(hipStreamSynchronize(b->stream));
    std::vector<int> order(nm), mlist(nm), pairs(nm), colors(nm);
    std::vector<uint64_t> ck(nm);
    HIPCHECK(hipMemcpy(pairs.data(), b->S.pair, sizeof(int) * nm, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(ck.data(), b->sorted_keys, sizeof(uint64_t) * nm, hipMemcpyDeviceToHost));
    // the pair entries, then the tail entries of concave pairs' extra manifolds packed after them
    const int np = b->contacts_np;
    const int nt = b->has_concave ? std::min(b->last.nb_tails, b->tail_cap) : 0;
    const size_t ne = (size_t)np + nt;
    std::vector<int2> col(ne);
    std::vector<int> ncp(ne), sub(ne, 0);
    std::vector<float4> pn(4 * ne), pl1(4 * ne), pl2(4 * ne), fimp(ne);
    auto fetch = [&](size_t dst, size_t src, size_t n) -> hipError_t {
        hipError_t e;
        if (!n) return hipSuccess;
        if ((e = hipMemcpy(col.data() + dst, P.col + src, sizeof(int2) * n, hipMemcpyDeviceToHost))) return e;
        if ((e = hipMemcpy(ncp.data() + dst, P.ncp + src, sizeof(int) * n, hipMemcpyDeviceToHost))) return e;
        if ((e = hipMemcpy(pn.data() + 4 * dst, P.pn + 4 * src, sizeof(float4) * 4 * n, hipMemcpyDeviceToHost))) return e;
        if ((e = hipMemcpy(pl1.data() + 4 * dst, P.pl1 + 4 * src, sizeof(float4) * 4 * n, hipMemcpyDeviceToHost))) return e;
        if ((e = hipMemcpy(pl2.data() + 4 * dst, P.pl2 + 4 * src, sizeof(float4) * 4 * n, hipMemcpyDeviceToHost))) return e;
        return hipMemcpy(fimp.data() + dst, P.fimp + src, sizeof(float4) * n, hipMemcpyDeviceToHost);
    };
    HIPCHECK(fetch(0, 0, np));
    HIPCHECK(fetch(np, b->pair_cap, nt));
    if (nt) {   // index of each tail manifold among its pair's manifolds
        std::vector<int> xm(np), xs(np);
        HIPCHECK(hipMemcpy(xm.data(), P.xm, sizeof(int) * np, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(xs.data(), P.xs, sizeof(int) * np, hipMemcpyDeviceToHost));
        for (int k = 0; k < np; k++)
            for (int j = 0; j < xm[k]; j++) {
                const int t = xs[k] + j - b->pair_cap;
                if (t >= 0 && t < nt) sub[np + t] = j + 1;
            }
    }
    for (int m = 0; m < nm; m++) {
        const int k = pairs[m] >= b->pair_cap ? np + pairs[m] - b->pair_cap : pairs[m];
        if (o->manifold) o->manifold[m] = sub[k];
        if (o->colliders) { o->colliders[2 * m] = col[k].x; o->colliders[2 * m + 1] = col[k].y; }
        if (o->nb_points) o->nb_points[m] = ncp[k];
        if (o->color) o->color[m] = (int)(ck[m] % kMaxColors);
        for (int p = 0; p < 4; p++) {
            const float4 a = pn[4 * k + p], c = pl1[4 * k + p], d = pl2[4 * k + p];
            if (o->normal_depth) { float* q = o->normal_depth + 16 * m + 4 * p; q[0] = a.x; q[1] = a.y; q[2] = a.z; q[3] = a.w; }
            if (o->local1) { float* q = o->local1 + 12 * m + 3 * p; q[0] = c.x; q[1] = c.y; q[2] = c.z; }
            if (o->local2) { float* q = o->local2 + 12 * m + 3 * p; q[0] = d.x; q[1] = d.y; q[2] = d.z; }
            if (o->impulse) o->impulse[4 * m + p] = c.w;
        }
        if (o->friction) { o->friction[3 * m] = fimp[k].x; o->friction[3 * m + 1] = fimp[k].y; o->friction[3 * m + 2] = fimp[k].z; }
    }
    return nm;
}

int rp3dg_batch_get_world_iterations(rp3dg_batch* b, int32_t* iterations, float* residuals, int64_t cap) {
    if (!b) return RP3DG_EINVAL;
    if (cap < b->nworlds) { set_error("capacity too small"); return RP3DG_EINVAL; }
    HIPCHECK(hipStreamSynchronize(b->stream));
    if (!b->fused) {
        for (int w = 0; w < b->nworlds; w++) {
            if (iterations) iterations[w] = b->settings.vel_iters;
            if (residuals) residuals[w] = 0.0f;
        }
        return b->nworlds;
    }
    if (iterations) HIPCHECK(hipMemcpy(iterations, b->world_iters, sizeof(int) * b->nworlds, hipMemcpyDeviceToHost));
    if (residuals) HIPCHECK(hipMemcpy(residuals, b->world_res, sizeof(float) * b->nworlds, hipMemcpyDeviceToHost));
    return b->nworlds;
}

int64_t rp3dg_batch_get_warm_start_predictions(rp3dg_batch* b, float* pred, int64_t cap) {
    if (!b) return RP3DG_EINVAL;
    if (!b->fused) { set_error("the learned warm start runs in the fused solver layout only"); return RP3DG_ESTATE; }
    const int nm = b->last.nb_manifolds;
    if (!pred) return nm;
    if (cap < nm) { set_error("capacity too small"); return RP3DG_EINVAL; }
    if (nm == 0) return 0;
    HIPCHECK(hipStreamSynchronize(b->stream));
    std::vector<float4> pr(4 * (size_t)nm);
    std::vector<float4> lam(4 * (size_t)nm);
    std::vector<int> npts(nm);
    HIPCHECK(hipMemcpy(pr.data(), b->S.pred, sizeof(float4) * 4 * nm, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(npts.data(), b->S.npts, sizeof(int) * nm, hipMemcpyDeviceToHost));
    for (int m = 0; m < nm; m++)
        for (int p = 0; p < 4; p++) {
            const bool live = b->warm_mode == RP3DG_WARM_START_LEARNED && p < npts[m];
            const float4 v = pr[4 * (size_t)m + p];
            float* o = pred + 12 * (size_t)m + 3 * p;
            o[0] = live ? v.x : 0.0f;
            o[1] = live ? v.y : 0.0f;
            o[2] = live ? v.z : 0.0f;
        }
    return nm;
}

int rp3dg_batch_get_joint_order(rp3dg_batch* b, int32_t* order, int64_t cap) {
    if (!b) return RP3DG_EINVAL;
    const int n = (int)b->joint_order.size();
    if (cap < n || (n > 0 && !order)) { set_error("capacity too small"); return RP3DG_EINVAL; }
    for (int i = 0; i < n; i++) order[i] = b->joint_order[i];
    return n;
}

int rp3dg_batch_profile(rp3dg_batch* b, int32_t enable, const char* only_kernel) {
    if (!b) return RP3DG_EINVAL;
    b->profiling = enable != 0;
    b->profile_only = (enable && only_kernel) ? only_kernel : "";
    b->stats.clear();
    return RP3DG_OK;
}

int rp3dg_batch_get_kernel_times(rp3dg_batch* b, const char* name, double* mean_ms, int64_t* launches) {
    if (!b || !name) return RP3DG_EINVAL;
    auto it = b->stats.find(name);
    if (it == b->stats.end()) { if (mean_ms) *mean_ms = 0; if (launches) *launches = 0; return RP3DG_OK; }
    if (mean_ms) *mean_ms = it->second.launches ? it->second.total_ms / it->second.launches : 0.0;
    if (launches) *launches = it->second.launches;
    return RP3DG_OK;
}

int rp3dg_batch_step_graph(rp3dg_batch* b, float dt, int32_t n) {
    if (!b) return RP3DG_EINVAL;
    if (!(dt > 0.f)) { set_error("time step must be > 0"); return RP3DG_EINVAL; }
    return b->step_graph(dt, n);
}

int rp3dg_batch_set_residual_mode(rp3dg_batch* b, float thr, int32_t maxit) {
    if (!b) return RP3DG_EINVAL;
    if (thr > 0.0f) {
        if (!b->fused) { set_error("residual mode needs the fused solver layout"); return RP3DG_EUNSUPPORTED; }
        if (maxit < 1 || maxit > 10000) { set_error("max_iterations must be in [1, 10000]"); return RP3DG_EINVAL; }
        b->res_thr = thr;
        b->res_max_iters = maxit;
    } else {
        b->res_thr = 0.0f;
        b->res_max_iters = 0;
    }
    b->drop_graphs();
    return RP3DG_OK;
}

static int ensure_features(rp3dg_batch* b) {
    if (b->feat_buf) return RP3DG_OK;
    b->feat_cap = (int)std::min<int64_t>(4 * (int64_t)b->pair_cap, std::max<int64_t>(1 << 16, 4 * (int64_t)b->nc));
    b->feat_buf = dalloc<float>(b, (size_t)b->feat_cap * kFeatures);
    b->row_slot = dalloc<int>(b, b->feat_cap);
    if (!b->feat_buf || !b->row_slot) { set_error("device allocation failed (feature rows)"); return RP3DG_EHIP; }
    return RP3DG_OK;
}

int rp3dg_batch_set_warm_start(rp3dg_batch* b, int32_t mode, rp3dg_mlp* mlp) {
    if (!b) return RP3DG_EINVAL;
    if (mode == RP3DG_WARM_START_LEARNED) {
        if (!mlp) { set_error("learned warm start needs a network"); return RP3DG_EINVAL; }
        if (!b->fused) { set_error("learned warm start needs the fused solver layout"); return RP3DG_EUNSUPPORTED; }
        const int r = ensure_features(b);
        if (r) return r;
    } else if (mode != RP3DG_WARM_START_PLAIN) {
        set_error("unknown warm start mode");
        return RP3DG_EINVAL;
    }
    b->warm_mode = mode;
    b->mlp = mode == RP3DG_WARM_START_LEARNED ? mlp : nullptr;
    b->drop_graphs();
    return RP3DG_OK;
}

int rp3dg_batch_set_logging(rp3dg_batch* b, int64_t capacity) {
    if (!b || capacity < 0) return RP3DG_EINVAL;
    if (b->log_feats) { hipFree(b->log_feats); b->log_feats = nullptr; }
    if (b->log_targets) { hipFree(b->log_targets); b->log_targets = nullptr; }
    b->log_cap = 0;
    b->drop_graphs();
    if (capacity == 0) return RP3DG_OK;
    if (!b->fused) { set_error("sample logging needs the fused solver layout"); return RP3DG_EUNSUPPORTED; }
    int r = ensure_features(b);
    if (r) return r;
    if (!b->log_count) {
        b->log_count = dalloc<unsigned long long>(b, 1);
        if (!b->log_count) { set_error("device allocation failed (log)"); return RP3DG_EHIP; }
    }
    HIPCHECK(hipMalloc(&b->log_feats, sizeof(float) * kFeatures * capacity));
    HIPCHECK(hipMalloc(&b->log_targets, sizeof(float) * 3 * capacity));
    HIPCHECK(hipMemsetAsync(b->log_count, 0, sizeof(unsigned long long), b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
    b->log_cap = capacity;
    return RP3DG_OK;
}

int64_t rp3dg_batch_get_log(rp3dg_batch* b, float* feats, float* targets, int64_t cap) {
    if (!b || cap < 0) return RP3DG_EINVAL;
    if (!b->log_cap) return 0;
    unsigned long long n = 0;
    HIPCHECK(hipMemcpyAsync(&n, b->log_count, sizeof(n), hipMemcpyDeviceToHost, b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
    const int64_t k = std::min<int64_t>((int64_t)n, cap);
    if (k > 0 && feats) HIPCHECK(hipMemcpy(feats, b->log_feats, sizeof(float) * kFeatures * k, hipMemcpyDeviceToHost));
    if (k > 0 && targets) HIPCHECK(hipMemcpy(targets, b->log_targets, sizeof(float) * 3 * k, hipMemcpyDeviceToHost));
    return k;
}

// Autotuner: times every candidate launch shape of the solver over a few real steps of this batch
// (the steps are not undone) and keeps the fastest.  Fused layout: workgroup size of each world
// class = clamp(bodies bound / d, 64, 256) for d = 1, 2, 4, 8, 16 (the choice returned is d); per-colour
// layout: threads per block of the colour kernels.
int rp3dg_batch_autotune(rp3dg_batch* b, float dt, int32_t steps_per_trial, int32_t* chosen, double* best_ms) {
    if (!b || !(dt > 0.f) || steps_per_trial < 1 || steps_per_trial > 1000) return RP3DG_EINVAL;
    std::vector<int> cand = b->fused ? std::vector<int>{1, 2, 4, 8, 16} : std::vector<int>{64, 128, 256, 512};
    const bool prof = b->profiling;
    b->profiling = false;
    b->drop_graphs();
    b->tune_table.clear();
    int best = -1;
    double best_t = 1e300;
    for (int c : cand) {
        if (b->fused) { b->block_div = c; b->set_class_blocks(); }
        else b->color_block = c;
        int r = b->step(dt);   // one untimed step with the new shape
        if (r) { b->profiling = prof; return r; }
        HIPCHECK(hipStreamSynchronize(b->stream));
        const auto t0 = std::chrono::steady_clock::now();
        for (int s = 0; s < steps_per_trial; s++) {
            r = b->step(dt);
            if (r) { b->profiling = prof; return r; }
        }
        HIPCHECK(hipStreamSynchronize(b->stream));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
                          steps_per_trial;
        b->tune_table.push_back({c, ms});
        if (ms < best_t) { best_t = ms; best = c; }
    }
    if (b->fused) { b->block_div = best; b->set_class_blocks(); }
    else b->color_block = best;
    b->profiling = prof;
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = best_t;
    return RP3DG_OK;
}

int rp3dg_batch_set_contacts_position_correction(rp3dg_batch* b, int32_t technique) {
    if (!b) return RP3DG_EINVAL;
    if (technique != RP3DG_CONTACTS_BAUMGARTE && technique != RP3DG_CONTACTS_SPLIT_IMPULSES) {
        set_error("unknown contacts position correction technique");
        return RP3DG_EINVAL;
    }
    const int split = technique == RP3DG_CONTACTS_SPLIT_IMPULSES ? 1 : 0;
    if (split != b->settings.split_impulse) b->drop_graphs();
    b->settings.split_impulse = split;
    return RP3DG_OK;
}

int rp3dg_batch_set_group_order(rp3dg_batch* b, int32_t by_sweeps) {
    if (!b) return RP3DG_EINVAL;
    if ((by_sweeps != 0) != b->order_by_sweeps) b->drop_graphs();
    b->order_by_sweeps = by_sweeps != 0;
    if (!b->order_by_sweeps && b->d_block_static && b->ngroups)   // back to the static order
        HIPCHECK(hipMemcpyAsync(b->d_block_world, b->d_block_static, sizeof(int) * b->ngroups, hipMemcpyDeviceToDevice,
                                b->stream));
    return RP3DG_OK;
}

int rp3dg_batch_set_solver_shape(rp3dg_batch* b, int32_t choice) {
    if (!b) return RP3DG_EINVAL;
    if (b->fused) {
        if (choice != 1 && choice != 2 && choice != 4 && choice != 8 && choice != 16) {
            set_error("fused layout: the shape is a workgroup divisor 1, 2, 4, 8 or 16");
            return RP3DG_EINVAL;
        }
        b->block_div = choice;
        b->set_class_blocks();
    } else {
        if (choice < 64 || choice > 512 || choice % 64) { set_error("per-colour layout: threads per block 64..512"); return RP3DG_EINVAL; }
        b->color_block = choice;
    }
    b->drop_graphs();
    return RP3DG_OK;
}

int32_t rp3dg_batch_get_tune_table(rp3dg_batch* b, int32_t* choices, double* ms, int32_t cap) {
    if (!b || cap < 0) return RP3DG_EINVAL;
    const int32_t n = (int32_t)b->tune_table.size();
    for (int32_t i = 0; i < n && i < cap; i++) {
        if (choices) choices[i] = b->tune_table[i].first;
        if (ms) ms[i] = b->tune_table[i].second;
    }
    return n;
}

int rp3dg_batch_get_iteration_stats(rp3dg_batch* b, double* mean_it, double* max_it, double* max_res) {
    if (!b) return RP3DG_EINVAL;
    double sum = 0.0, mx = 0.0, mr = 0.0;
    if (b->fused) {
        std::vector<int> it(b->nworlds);
        std::vector<float> res(b->nworlds);
        HIPCHECK(hipStreamSynchronize(b->stream));
        HIPCHECK(hipMemcpy(it.data(), b->world_iters, sizeof(int) * b->nworlds, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(res.data(), b->world_res, sizeof(float) * b->nworlds, hipMemcpyDeviceToHost));
        for (int w = 0; w < b->nworlds; w++) {
            sum += it[w];
            mx = std::max(mx, (double)it[w]);
            mr = std::max(mr, (double)res[w]);
        }
        sum /= std::max(1, b->nworlds);
    } else {
        sum = mx = b->settings.vel_iters;
    }
    if (mean_it) *mean_it = sum;
    if (max_it) *max_it = mx;
    if (max_res) *max_res = mr;
    return RP3DG_OK;
}

}  // extern "C"
