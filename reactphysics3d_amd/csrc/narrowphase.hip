// Narrowphase of the MI355X step path: one kernel per shape-pair type over a
// compacted list of pair slots (CollisionDetectionSystem::computeMiddlePhase
// buckets, :346-413), each thread running one test and then the contact
// epilogue for its pair:
//   raw contacts (<= 16, NarrowPhaseInfoBatch.h:170-185)
//   -> potential manifold (convex pair: exactly one, CD.cpp:1247-1290)
//   -> reduceContactPoints (CD.cpp:1483-1662) + removeDuplicated (:1673-1705)
//   -> warm-start transfer from the pair's cached contacts (CD.cpp:1069-1160)
//   -> cache write (the next step reads it back).
#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

#include "kcommon.h"
#include "narrowphase.h"
#include "refit_dev.h"
#include "rp3d_internal.h"

namespace rp3dg {

__constant__ int4 c_box_edges[24];   // vertex, twin, face, next (HalfEdgeStructure of the box, PhysicsCommon.cpp:73-109)
__constant__ int c_box_face_edge[6];
__constant__ int c_box_face_verts[24];

// ------------------------------------------------------------------------------------------------
// Polyhedron view: BoxShape (implicit) or ConvexMeshShape (hull arena).  A concave shape's triangle
// (TriangleShape.h) is a hull of the arena's triangle slots, written by its test (header pad = 1).
// ------------------------------------------------------------------------------------------------
struct Poly {
    int hull;          // -1 for a box
    V3 he;             // box half-extents
    HullHeader h;
    const float4* hv;  // hull vertices
    const float4* hn;  // hull face normals
    const int4* hedges;
    const int* hfedge;
    const int* hfverts;   // [2*f] = start, [2*f+1] = count into hfvidx
    const int* hfvidx;

    __device__ bool tri() const { return hull >= 0 && h.pad != 0.0f; }
    __device__ int nfaces() const { return hull < 0 ? 6 : h.nf; }
    __device__ int nhalfedges() const { return hull < 0 ? 24 : h.ne; }
    // BoxShape.h:186-205 (vertices 0-3 on +z, 4-7 on -z; +x for 1, 2, 5, 6; +y for 2, 3, 6, 7), as
    // selects so that lanes holding different indices do not branch
    __device__ V3 vertex(int i) const {
        if (hull >= 0) return xyz(hv[h.vbase + i]);
        return v3(((i + 1) & 2) ? he.x : -he.x, (i & 2) ? he.y : -he.y, i < 4 ? he.z : -he.z);
    }
    // BoxShape.h:208-221 (faces +z, +x, -z, -x, -y, +y)
    __device__ V3 face_normal(int f) const {
        if (hull >= 0) return xyz(hn[h.fbase + f]);
        return v3(f == 1 ? 1.0f : f == 3 ? -1.0f : 0.0f, f == 5 ? 1.0f : f == 4 ? -1.0f : 0.0f,
                  f == 0 ? 1.0f : f == 2 ? -1.0f : 0.0f);
    }
    __device__ int4 edge(int e) const { return hull < 0 ? c_box_edges[e] : hedges[h.ebase + e]; }
    __device__ int face_edge(int f) const { return hull < 0 ? c_box_face_edge[f] : hfedge[h.fbase + f]; }
    __device__ int face_nverts(int f) const { return hull < 0 ? 4 : hfverts[2 * (h.fbase + f) + 1]; }
    __device__ int face_vertex(int f, int k) const {
        return hull < 0 ? c_box_face_verts[4 * f + k] : hfvidx[hfverts[2 * (h.fbase + f)] + k];
    }
    __device__ V3 centroid() const { return hull < 0 ? v3zero() : v3(h.cx, h.cy, h.cz); }
    // BoxShape.h:167-172, ConvexMeshShape.cpp:80-102
    // (a triangle's getMaxAxis of its three dot products picks the first maximum, as this walk does)
    __device__ V3 support(V3 d) const {
        if (hull < 0)
            return v3(d.x < 0.0f ? -he.x : he.x, d.y < 0.0f ? -he.y : he.y, d.z < 0.0f ? -he.z : he.z);
        float best = kDecimalSmallest;
        int bi = 0;
        for (int i = 0; i < h.nv; i++) {
            const float dp = dot(d, xyz(hv[h.vbase + i]));
            if (dp > best) { bi = i; best = dp; }
        }
        return xyz(hv[h.vbase + bi]);
    }
    // ConvexPolyhedronShape.h:105-123
    __device__ int most_anti_parallel_face(V3 dir) const {
        float mn = kDecimalLargest;
        int best = 0;
        const int nf = nfaces();
        for (int i = 0; i < nf; i++) {
            const float dp = dot(face_normal(i), dir);
            if (dp < mn) { mn = dp; best = i; }
        }
        return best;
    }
};

__device__ inline Poly make_poly(const NpArgs& A, int collider) {
    Poly P;
    const float4 s = A.C.shape[collider];
    if (A.C.type[collider] == SHAPE_BOX) {
        P.hull = -1;
        P.he = v3(s.x, s.y, s.z);
    } else {
        P.hull = (int)s.x;
        P.h = A.hulls[P.hull];
        P.h.pad = 0.0f;   // a collider's hull is never a triangle: lets the triangle branches fold away
        P.hv = A.hull_verts;
        P.hn = A.hull_normals;
        P.hedges = A.hull_edges;
        P.hfedge = A.hull_face_edge;
        P.hfverts = A.hull_face_verts;
        P.hfvidx = A.hull_face_vidx;
        P.he = v3zero();
    }
    return P;
}

// ------------------------------------------------------------------------------------------------
// Raw contact buffer of one narrowphase test
// ------------------------------------------------------------------------------------------------
// N = kMaxRawContacts in general; box-box tests produce at most 8 points (a quad clipped by four
// planes), so their kernel uses a buffer of 8 and the reference's cap of 16 is never reached.
// TriangleShape::computeSmoothTriangleMeshContact (TriangleShape.h:330-350, TriangleShape.cpp:85-114): a
// contact with a concave shape's triangle takes the mesh's interpolated vertex normal at the point on the
// triangle, and the point on the other shape moves along it by the penetration depth.  Applied to every
// contact of a triangle test right before it is added, as the reference does at each contact site.
struct TriSmooth {
    bool tri1;                 // the triangle is shape 1
    V3 p0, p1, p2, n;          // triangle (concave shape space) and its normal
    V3 vn0, vn1, vn2;          // vertex normals
    Tf triToWorld, worldToOther;
    __device__ void apply(V3& normal, float depth, V3& l1, V3& l2) const {
        const V3 lp = tri1 ? l1 : l2;
        // computeBarycentricCoordinatesInTriangle (mathematics_functions.h:203-219)
        const V3 e0 = p1 - p0, e1 = p2 - p0, e2 = lp - p0;
        const float d00 = dot(e0, e0), d01 = dot(e0, e1), d11 = dot(e1, e1), d20 = dot(e2, e0), d21 = dot(e2, e1);
        const float denom = d00 * d11 - d01 * d01;
        const float v = (d11 * d20 - d01 * d21) / denom;
        const float w = (d00 * d21 - d01 * d20) / denom;
        const float u = 1.0f - v - w;
        // computeSmoothLocalContactNormalForTriangle (TriangleShape.h:310-327)
        V3 nl = n;
        if (!(u > kMachineEps && v > kMachineEps && w > kMachineEps)) {
            const V3 in = u * vn0 + v * vn1 + w * vn2;
            if (!(len2(in) < kMachineEps)) nl = unit(in);
        }
        V3 nw = qrot(triToWorld.q, nl);
        const V3 pen = tri1 ? normal : -normal;
        if (dot(nw, pen) < 0.0f) {
            nw = -nw;
            nl = -nl;
        }
        normal = tri1 ? nw : -nw;
        const V3 other = tapply(tmul(worldToOther, triToWorld), lp - nl * depth);
        if (tri1) l2 = other;
        else l1 = other;
    }
};

template <int N>
struct RawBuf {
    static constexpr int kN = N;
    int n;
    bool clip_full = false;   // a clipped polygon had more vertices than the clip buffer holds
    const TriSmooth* smooth = nullptr;   // set for the tests of a concave shape's triangle
    V3 nrm[N];
    float depth[N];
    V3 l1[N];
    V3 l2[N];
    __device__ void add(V3 normal, float d, V3 a, V3 b) { add_carry(normal, d, a, b); }
    // a contact site whose normal variable the smoothing rewrites for the next contact of the same loop
    // (SATAlgorithm::computeCapsulePolyhedronFaceContactPoints takes it by reference)
    __device__ void add_carry(V3& normal, float d, V3 a, V3 b) {  // NarrowPhaseInfoBatch::addContactPoint
        if (smooth) smooth->apply(normal, d, a, b);
        if (n < N) {
            nrm[n] = normal; depth[n] = d; l1[n] = a; l2[n] = b;
            n++;
        }
    }
    __device__ V3 normal(int r) const { return nrm[r]; }
    __device__ V3 p1(int r) const { return l1[r]; }
    __device__ void point(int r, V3& a, V3& b, float& d) const { a = l1[r]; b = l2[r]; d = depth[r]; }
};
using Raw = RawBuf<kMaxRawContacts>;

// Up to 16 indices below 16, four bits each, in one register (dynamic indexing without scratch)
struct IdxList {
    uint64_t v = 0;
    __device__ int get(int i) const { return (int)((v >> (4 * i)) & 15u); }
    __device__ void set(int i, int x) { v = (v & ~(15ull << (4 * i))) | ((uint64_t)x << (4 * i)); }
};

// Clip buffers of one thread.  General polyhedra keep them in private arrays; the box-box kernel keeps
// them in a strided slice of its workgroup's LDS, so that it needs no scratch memory.
template <int CM>
struct PrivClip {
    V3 v[CM];
    __device__ V3 get(int i) const { return v[i]; }
    __device__ void put(int i, V3 x) { v[i] = x; }
};
struct LdsClip {
    float* p;   // component c of vertex i at p[(3 * i + c) * kNpBoxThreads]
    __device__ V3 get(int i) const {
        return v3(p[(3 * i) * kNpBoxThreads], p[(3 * i + 1) * kNpBoxThreads], p[(3 * i + 2) * kNpBoxThreads]);
    }
    __device__ void put(int i, V3 x) {
        p[(3 * i) * kNpBoxThreads] = x.x;
        p[(3 * i + 1) * kNpBoxThreads] = x.y;
        p[(3 * i + 2) * kNpBoxThreads] = x.z;
    }
};

// Raw contacts of a box-box test, held without private arrays.  Every contact point of one test carries
// the same normal (the reference face normal of the face contacts, or the single edge-edge contact).
// Face contacts are kept as indices of clipped vertices in the LDS clip buffer, and each read recomputes
// the point pair and depth from the vertex with the operations that the general buffer stores
// (SATAlgorithm.cpp:1010-1025), so the values are identical; the one edge-edge contact is kept as is.
struct RawBoxView {
    static constexpr int kN = 8;
    int n;
    bool clip_full = false;
    V3 nrm0;
    IdxList ci;       // clip-buffer index of each face contact
    LdsClip clip;     // the buffer that holds the clipped incident face
    bool face;        // contacts are face contacts (else the one edge-edge contact below)
    bool refIs1;
    Tf refToInc;
    V3 axisRef, refV;
    V3 e1, e2;
    float ed;
    __device__ void add(V3 normal, float d, V3 a, V3 b) {   // edge-edge contact
        if (n < kN) {
            nrm0 = normal; face = false; e1 = a; e2 = b; ed = d;
            n++;
        }
    }
    __device__ void add_clipped(V3 normal, int i) {
        if (n < kN) {
            nrm0 = normal; face = true; ci.set(n, i);
            n++;
        }
    }
    __device__ V3 normal(int) const { return nrm0; }
    __device__ void point(int r, V3& a, V3& b, float& d) const {
        if (!face) { a = e1; b = e2; d = ed; return; }
        const V3 c = clip.get(ci.get(r));
        d = dot(refV - c, axisRef);
        const V3 cInc = tapply(refToInc, c);
        const V3 cRef = project_point_onto_plane(c, axisRef, refV);
        a = refIs1 ? cRef : cInc;
        b = refIs1 ? cInc : cRef;
    }
    __device__ V3 p1(int r) const {
        if (!face) return e1;
        const V3 c = clip.get(ci.get(r));
        return refIs1 ? project_point_onto_plane(c, axisRef, refV) : tapply(refToInc, c);
    }
};

constexpr int kBoxClip = 8;     // vertices of a box face clipped by the four side planes at most

// ------------------------------------------------------------------------------------------------
// SAT, polyhedron vs polyhedron (SATAlgorithm.cpp:485-1191)
// ------------------------------------------------------------------------------------------------
constexpr float kSatRel = 1.002f;    // SATAlgorithm.cpp:43
constexpr float kSatAbs = 0.0005f;   // SATAlgorithm.cpp:44
constexpr int kClipMax = 2 * kMaxHullFaceVertices;

// :1095-1116
__device__ inline float sat_single_face(const Poly& P1, const Poly& P2, const Tf& t12, int f) {
    const V3 n = P1.face_normal(f);
    const V3 n2 = qrot(t12.q, n);
    const V3 sp = P2.support(-n2);
    const V3 fv = tapply(t12, P1.vertex(P1.face_vertex(f, 0)));
    return dot(fv - sp, n2);
}

// :1119-1143
__device__ inline float sat_faces(const Poly& P1, const Poly& P2, const Tf& t12, int& minFace) {
    float mn = kDecimalLargest;
    const int nf = P1.nfaces();
    for (int f = 0; f < nf; f++) {
        const float pen = sat_single_face(P1, P2, t12, f);
        if (pen <= 0.0f) { minFace = f; return pen; }
        if (pen < mn) { mn = pen; minFace = f; }
    }
    return mn;
}

// :1146-1191
__device__ inline bool sat_minkowski_face(const Poly& P1, int4 e1, const Poly& P2, int4 e2, const Tf& t12) {
    const V3 a = qrot(t12.q, P1.face_normal(e1.z));
    const V3 b = qrot(t12.q, P1.face_normal(P1.edge(e1.y).z));
    const V3 c = P2.face_normal(e2.z);
    const V3 d = P2.face_normal(P2.edge(e2.y).z);
    const V3 ev1 = P1.vertex(e1.x);
    const V3 ev2 = P1.vertex(P1.edge(e1.y).x);
    const V3 bxa = qrot(t12.q, ev1 - ev2);
    const V3 fv1 = P2.vertex(e2.x);
    const V3 fv2 = P2.vertex(P2.edge(e2.y).x);
    const V3 dxc = fv1 - fv2;
    const V3 nc = -c, nd = -d;
    const float cba = dot(nc, bxa);
    const float dba = dot(nd, bxa);
    const float adc = dot(a, dxc);
    const float bdc = dot(b, dxc);
    return cba * dba < 0.0f && adc * bdc < 0.0f && cba * bdc > 0.0f;
}

// :1040-1083; when shape 1 is a triangle (tri1) the axis is oriented with shape 2's centroid c2
__device__ inline float sat_edge_distance(V3 e1a, V3 e2a, V3 c1, V3 e1d, V3 e2d, V3& axis_out, bool tri1 = false,
                                          V3 c2 = V3{0.f, 0.f, 0.f}) {
    if (parallel_vectors(e1d, e2d)) return kDecimalLargest;
    V3 axis = unit(cross(e1d, e2d));
    const float dp = tri1 ? dot(axis, e2a - c2) : dot(axis, c1 - e1a);
    if (dp > 0.0f) axis = -axis;
    axis_out = axis;
    return -dot(axis, e2a - e1a);
}

// mathematics_functions.h:332-392
// A vertex past the buffer's CM is dropped and `full` is set (the caller reports it in the status word).
template <int CM, class CB>
__device__ inline int clip_polygon(const CB& in, int nin, V3 pp, V3 pn, CB& out, bool& full) {
    int nout = 0;
    int s = nin - 1;
    const float pnd = dot(pn, pp);
    float sdn = dot(in.get(s) - pp, pn);
    for (int e = 0; e < nin; e++) {
        const V3 vs = in.get(s), ve = in.get(e);
        const float edn = dot(ve - pp, pn);
        if (edn >= 0.0f) {
            if (sdn < 0.0f) {
                const float t = plane_segment_intersection(vs, ve, pnd, pn);
                if (t >= 0.0f && t <= 1.0f) { if (nout < CM) out.put(nout++, vs + t * (ve - vs)); else full = true; }
                else { if (nout < CM) out.put(nout++, ve); else full = true; }
            }
            if (nout < CM) out.put(nout++, ve);
            else full = true;
        } else {
            if (sdn >= 0.0f) {
                const float t = plane_segment_intersection(vs, ve, -pnd, -pn);
                if (t >= 0.0f && t <= 1.0f) { if (nout < CM) out.put(nout++, vs + t * (ve - vs)); else full = true; }
                else { if (nout < CM) out.put(nout++, vs); else full = true; }
            }
        }
        s = e;
        sdn = edn;
    }
    return nout;
}

// :893-1036
template <int CM, class RB, class CB>
__device__ __forceinline__ bool sat_face_contacts_t(bool refIs1, const Poly& P1, const Poly& P2, const Tf& t12,
                                                    const Tf& t21, int minFace, const Tf& T1, const Tf& T2, RB& raw,
                                                    CB& bufA, CB& bufB) {
    const Tf& refToInc = refIs1 ? t12 : t21;
    const Tf& incToRef = refIs1 ? t21 : t12;
    const Poly& R = refIs1 ? P1 : P2;
    const Poly& I = refIs1 ? P2 : P1;
    const V3 axisRef = R.face_normal(minFace);
    const V3 axisInc = qrot(refToInc.q, axisRef);
    const V3 normalWorld = refIs1 ? qrot(T1.q, axisRef) : -(qrot(T2.q, axisRef));
    const int incFace = I.most_anti_parallel_face(axisInc);
    int na = 0, nb = 0;
    const int niv = I.face_nverts(incFace);
    if (niv > CM) raw.clip_full = true;
    for (int i = 0; i < niv && na < CM; i++) bufA.put(na++, tapply(incToRef, I.vertex(I.face_vertex(incFace, i))));
    const int firstEdge = R.face_edge(minFace);
    bool v1Input = false;
    int nOut = 0;
    int curIdx = firstEdge;
    int4 cur = R.edge(firstEdge);
    V3 ev1 = R.vertex(cur.x);
    int guard = R.nhalfedges();   // a face has fewer edges than the mesh has half-edges
    do {
        if (--guard < 0) break;
        v1Input = !v1Input;
        const int4 nxt = R.edge(cur.w);
        const V3 ev2 = R.vertex(nxt.x);
        const V3 planeNormal = cross(axisRef, ev2 - ev1);
        // one inlined copy of the clipping, alternating its input and output buffers
        const int nn = clip_polygon<CM>(v1Input ? bufA : bufB, v1Input ? na : nb, ev1, planeNormal,
                                        v1Input ? bufB : bufA, raw.clip_full);
        if (v1Input) nb = nn;
        else na = nn;
        curIdx = cur.w;
        cur = nxt;
        ev1 = ev2;
        if (v1Input) { na = 0; nOut = nb; }
        else { nb = 0; nOut = na; }
    } while (curIdx != firstEdge && nOut > 0);
    const CB& clipped = v1Input ? bufB : bufA;
    const int nclip = v1Input ? nb : na;
    const V3 refV = R.vertex(R.edge(firstEdge).x);
    bool found = false;
    if constexpr (std::is_same<RB, RawBoxView>::value) {
        raw.clip = clipped;
        raw.refIs1 = refIs1;
        raw.refToInc = refToInc;
        raw.axisRef = axisRef;
        raw.refV = refV;
    }
    for (int i = 0; i < nclip; i++) {
        const V3 ci = clipped.get(i);
        const float pen = dot(refV - ci, axisRef);
        if (pen > 0.0f) {
            found = true;
            if constexpr (std::is_same<RB, RawBoxView>::value) {
                raw.add_clipped(normalWorld, i);
            } else {
                const V3 cInc = tapply(refToInc, ci);
                const V3 cRef = project_point_onto_plane(ci, axisRef, refV);
                raw.add(normalWorld, pen, refIs1 ? cRef : cInc, refIs1 ? cInc : cRef);
            }
        }
    }
    return found;
}

// General polyhedra (hulls): out of line, large clip buffers
__device__ __noinline__ bool sat_face_contacts(bool refIs1, const Poly& P1, const Poly& P2, const Tf& t12,
                                               const Tf& t21, int minFace, const Tf& T1, const Tf& T2, Raw& raw) {
    PrivClip<kClipMax> bufA, bufB;
    return sat_face_contacts_t<kClipMax>(refIs1, P1, P2, t12, t21, minFace, T1, T2, raw, bufA, bufB);
}

// Box-box tests inline the face clipping with buffers sized for a box; hull tests call it out of line.
template <bool BOX, class R>
__device__ __forceinline__ bool face_contacts(bool refIs1, const Poly& P1, const Poly& P2, const Tf& t12, const Tf& t21,
                                              int minFace, const Tf& T1, const Tf& T2, R& raw) {
    if constexpr (BOX) {
        // one slice per thread of the np_box_box workgroup (kNpBoxThreads threads), shared by every
        // inlined call site of the test
        __shared__ float s_box_clip[2 * 3 * kBoxClip * kNpBoxThreads];
        LdsClip bufA{s_box_clip + threadIdx.x};
        LdsClip bufB{s_box_clip + 3 * kBoxClip * kNpBoxThreads + threadIdx.x};
        return sat_face_contacts_t<kBoxClip>(refIs1, P1, P2, t12, t21, minFace, T1, T2, raw, bufA, bufB);
    } else {
        return sat_face_contacts(refIs1, P1, P2, t12, t21, minFace, T1, T2, raw);
    }
}

__device__ inline uint32_t lf_set_face(uint32_t lf, bool face1, int face) {
    lf &= ~(LF_SAT_FACE1 | LF_SAT_FACE2 | 0xff00u);
    lf |= face1 ? LF_SAT_FACE1 : LF_SAT_FACE2;
    lf |= ((uint32_t)face & 0xffu) << 8;
    return lf;
}
__device__ inline uint32_t lf_set_edges(uint32_t lf, int e1, int e2) {
    lf &= ~(LF_SAT_FACE1 | LF_SAT_FACE2 | 0xffff0000u);
    lf |= ((uint32_t)e1 & 0xffu) << 16;
    lf |= ((uint32_t)e2 & 0xffu) << 24;
    return lf;
}

// Returns isColliding; updates the SAT part of the temporal-coherence word.

// ------------------------------------------------------------------------------------------------
// Sphere / capsule closed forms
// ------------------------------------------------------------------------------------------------

// SphereVsSphereAlgorithm.cpp:34-113
__device__ bool sphere_sphere(float r1, float r2, const Tf& T1, const Tf& T2, Raw& raw) {
    const V3 d = T2.p - T1.p;
    const float d2 = len2(d);
    const float sum = r1 + r2;
    const float sum2 = sum * sum;
    if (!(d2 < sum2)) return false;
    const float pen = sum - sqrtf(d2);
    if (!(pen > 0.0f)) return false;
    const Tf i1 = tinv(T1), i2 = tinv(T2);
    V3 p1, p2, n;
    if (d2 > kMachineEps) {
        const V3 c2in1 = tapply(i1, T2.p);
        const V3 c1in2 = tapply(i2, T1.p);
        p1 = r1 * unit(c2in1);
        p2 = r2 * unit(c1in2);
        n = unit(d);
    } else {
        n = v3(0, 1, 0);
        p1 = r1 * qrot(i1.q, n);
        p2 = r2 * qrot(i2.q, n);
    }
    raw.add(n, pen, p1, p2);
    return true;
}

// SphereVsCapsuleAlgorithm.cpp:38-152
__device__ bool sphere_capsule(bool sphereIs1, float rs, float rc, float capHeight, const Tf& Ts, const Tf& Tc, Raw& raw) {
    const Tf worldToCap = tinv(Tc);
    const Tf s2c = tmul(worldToCap, Ts);
    const V3 sc = s2c.p;
    const float hh = capHeight * 0.5f;
    const V3 sa = v3(0, -hh, 0), sb = v3(0, hh, 0);
    const V3 cps = closest_point_on_segment(sa, sb, sc);
    V3 sc2seg = cps - sc;
    const float dist2 = len2(sc2seg);
    const float sumR = rs + rc;
    if (!(dist2 < sumR * sumR)) return false;
    float pen;
    V3 nw, cpS, cpC;
    if (dist2 > kMachineEps) {
        const float dist = sqrtf(dist2);
        sc2seg = sc2seg / dist;
        cpS = tapply(tinv(s2c), sc + sc2seg * rs);
        cpC = cps - sc2seg * rc;
        nw = qrot(Tc.q, sc2seg);
        pen = sumR - dist;
        if (!sphereIs1) nw = -nw;
    } else {
        const V3 seg = unit(sb - sa);
        const float cosA1 = fabsf(seg.x);
        const float cosA2 = fabsf(seg.y);
        pen = sumR;
        const V3 ncs = cosA1 < cosA2 ? cross(seg, v3(1, 0, 0)) : cross(seg, v3(0, 1, 0));
        nw = qrot(Tc.q, ncs);
        cpS = tapply(tinv(s2c), sc + ncs * rs);
        cpC = sc - ncs * rc;
    }
    if (pen <= 0.0f) return false;
    raw.add(nw, pen, sphereIs1 ? cpS : cpC, sphereIs1 ? cpC : cpS);
    return true;
}

// mathematics_functions.h:216-224
__device__ inline float point_to_line_distance(V3 a, V3 b, V3 p) {
    const float dab = len(b - a);
    if (dab < kMachineEps) return len(p - a);
    return len(cross(p - a, p - b)) / dab;
}

// CapsuleVsCapsuleAlgorithm.cpp:37-243
__device__ bool capsule_capsule(float r1, float h1, float r2, float h2, const Tf& T1, const Tf& T2, Raw& raw) {
    const Tf c1to2 = tmul(tinv(T2), T1);
    const float hh1 = h1 * 0.5f;
    V3 a1 = tapply(c1to2, v3(0, -hh1, 0));
    V3 b1 = tapply(c1to2, v3(0, hh1, 0));
    const float hh2 = h2 * 0.5f;
    const V3 a2 = v3(0, -hh2, 0), b2 = v3(0, hh2, 0);
    const V3 seg1 = b1 - a1;
    const V3 seg2 = b2 - a2;
    const float sumR = r1 + r2;
    const bool par = parallel_vectors(seg1, seg2);
    if (par) {
        const float perp = point_to_line_distance(a1, b1, a2);
        if (perp >= sumR) return false;
        const float d1 = dot(seg1, a1);
        const float d2 = -dot(seg1, b1);
        float t1 = plane_segment_intersection(b2, a2, d1, seg1);
        float t2 = plane_segment_intersection(a2, b2, d2, -seg1);
        if (t1 > 0.0f && t2 > 0.0f) {
            if (t1 > 1.0f) t1 = 1.0f;
            const V3 clipA = b2 - t1 * seg2;
            if (t2 > 1.0f) t2 = 1.0f;
            const V3 clipB = a2 + t2 * seg2;
            const V3 s1n = unit(seg1);
            const V3 pOn1 = a1 + dot(s1n, a2 - a1) * s1n;
            V3 n2, s1to2;
            if (perp > kMachineEps) {
                s1to2 = a2 - pOn1;
                n2 = unit(s1to2);
            } else {
                const V3 s2n = unit(seg2);
                const float cA1 = fabsf(s2n.x), cA2 = fabsf(s2n.y);
                s1to2 = v3zero();
                n2 = cA1 < cA2 ? cross(s2n, v3(1, 0, 0)) : cross(s2n, v3(0, 1, 0));
            }
            const Tf c2to1 = tinv(c1to2);
            const V3 pa1 = tapply(c2to1, clipA - s1to2 + n2 * r1);
            const V3 pb1 = tapply(c2to1, clipB - s1to2 + n2 * r1);
            const V3 pa2 = clipA - n2 * r2;
            const V3 pb2 = clipB - n2 * r2;
            const float pen = sumR - perp;
            const V3 nw = qrot(T2.q, n2);
            raw.add(nw, pen, pa1, pa2);
            raw.add(nw, pen, pb1, pb2);
            return true;
        }
    }
    V3 c1, c2;
    closest_points_between_segments(a1, b1, a2, b2, c1, c2);
    V3 d = c2 - c1;
    const float dd = len2(d);
    if (!(dd < sumR * sumR)) return false;
    if (dd > kMachineEps) {
        const float dist = sqrtf(dd);
        d = d / dist;
        const V3 p1 = tapply(tinv(c1to2), c1 + d * r1);
        const V3 p2 = c2 - d * r2;
        const V3 nw = qrot(T2.q, d);
        const float pen = fmax_s(sumR - dist, kMachineEps);
        raw.add(nw, pen, p1, p2);
    } else if (par) {
        const float sq = len2(a1 - c2);
        const V3 ext = sq > kMachineEps ? a1 : b1;
        const V3 n2 = normalized(c2 - ext);
        const V3 p1 = tapply(tinv(c1to2), c1 + n2 * r1);
        const V3 p2 = c2 - n2 * r2;
        raw.add(qrot(T2.q, n2), sumR, p1, p2);
    } else {
        const V3 n2 = normalized(cross(seg1, seg2));
        const V3 p1 = tapply(tinv(c1to2), c1 + n2 * r1);
        const V3 p2 = c2 - n2 * r2;
        raw.add(qrot(T2.q, n2), sumR, p1, p2);
    }
    return true;
}

// ------------------------------------------------------------------------------------------------
// Contact epilogue: manifold reduction, duplicate removal, warm-start transfer, cache write.
// ------------------------------------------------------------------------------------------------

// CollisionDetectionSystem::reduceContactPoints (:1483-1662): picks 4 of n > 4 raw points.
template <class R, class L = IdxList>
__device__ void reduce_points(const R& raw, const Tf& T1, int* keep) {
    L cand;
    int nc = raw.n;
    for (int i = 0; i < nc; i++) cand.set(i, i);
    const Tf w2s = tinv(T1);
    const V3 nS1 = qrot(w2s.q, raw.normal(cand.get(0)));
    const V3 dir = v3(1, 1, 1);
    float maxDot = kDecimalSmallest;
    int idx = 0;
    for (int i = 0; i < nc; i++) {
        const float dp = dot(dir, raw.p1(cand.get(i)));
        if (dp > maxDot) { maxDot = dp; idx = i; }
    }
    keep[0] = cand.get(idx);
    cand.set(idx, cand.get(nc - 1)); nc--;
    float maxDist = 0.0f;
    idx = 0;
    for (int i = 0; i < nc; i++) {
        const float d = len2(raw.p1(keep[0]) - raw.p1(cand.get(i)));
        if (d >= maxDist) { maxDist = d; idx = i; }
    }
    keep[1] = cand.get(idx);
    cand.set(idx, cand.get(nc - 1)); nc--;
    int maxI = 0, minI = 0;
    float minArea = 0.0f, maxArea = 0.0f;
    for (int i = 0; i < nc; i++) {
        const V3 nf = raw.p1(keep[0]) - raw.p1(cand.get(i));
        const V3 ns = raw.p1(keep[1]) - raw.p1(cand.get(i));
        const float area = dot(cross(nf, ns), nS1);
        if (area >= maxArea) { maxArea = area; maxI = i; }
        if (area <= minArea) { minArea = area; minI = i; }
    }
    bool prevPositive;
    if (maxArea > (-minArea)) {
        prevPositive = true;
        keep[2] = cand.get(maxI);
        cand.set(maxI, cand.get(nc - 1)); nc--;
    } else {
        prevPositive = false;
        keep[2] = cand.get(minI);
        cand.set(minI, cand.get(nc - 1)); nc--;
    }
    float largest = 0.0f;
    idx = 0;
    for (int i = 0; i < nc; i++) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int va = j, vb = j < 2 ? j + 1 : 0;
            const V3 nf = raw.p1(keep[va]) - raw.p1(cand.get(i));
            const V3 ns = raw.p1(keep[vb]) - raw.p1(cand.get(i));
            const float area = dot(cross(nf, ns), nS1);
            if (prevPositive && area <= largest) { largest = area; idx = i; }
            else if (!prevPositive && area >= largest) { largest = area; idx = i; }
        }
    }
    keep[3] = cand.get(idx);
}

// A pair with a trigger collider is tested with reportContacts = false (CD.cpp:402): it can collide, but it
// has no contact points, no manifold and no solver rows.
__device__ __forceinline__ bool trigger_pair(const NpArgs& A, int slot) {
    const int2 c = A.P.col[slot];
    return ((A.C.kind[c.x] | A.C.kind[c.y]) & CK_TRIGGER) != 0;
}

// NarrowPhaseInfo::reportContacts of a tested pair: steps and testCollision(callback) leave trigger pairs
// without contacts (CD.cpp:402), testCollision(body[, body]) reports every pair (CD.cpp:462-511),
// testOverlap reports none.
__device__ __forceinline__ bool pair_reports(const NpArgs& A, int slot) {
    return A.report == NP_REPORT_ALL || (A.report == NP_REPORT_NO_TRIGGER && !trigger_pair(A, slot));
}

template <class R>
__device__ void contact_epilogue(const NpArgs& A, int slot, bool colliding, const R& raw, int collider1) {
    DevPairs P = A.P;
    const DevPairs& O = A.O;
    const int prev = P.prev[slot];   // this pair's entry in the previous table, -1 when new
    if (raw.clip_full) atomicOr(A.status, STATUS_CLIP_OVERFLOW);
    if (!colliding || raw.n == 0 || !pair_reports(A, slot)) {
        // a query marks a colliding pair without points with -1 (its overlap is reported)
        P.ncp[slot] = (A.query != NP_STEP && colliding) ? -1 : 0;
        return;
    }
    const int oldN = prev >= 0 ? O.ncp[prev] : 0;
    IdxList idx;
    int n = raw.n;
    for (int i = 0; i < n; i++) idx.set(i, i);
    if (n > kMaxManifoldPoints) {
        const Tf T1 = tf(ld3(A.C.wpos, collider1), ldq(A.C.wq, collider1));
        int keep[4];
        reduce_points(raw, T1, keep);
        for (int i = 0; i < 4; i++) idx.set(i, keep[i]);
        n = 4;
    }
    // removeDuplicatedContactPointsInManifold (:1673-1705)
    const float thr2 = 0.01f * 0.01f;
    for (int i = 0; i < n; i++) {
        for (int j = i + 1; j < n; j++) {
            const float d2 = len2(raw.p1(idx.get(j)) - raw.p1(idx.get(i)));
            if (d2 < thr2) {
                idx.set(j, idx.get(n - 1));
                n--;
                j--;
            }
        }
    }
    // warm start transfer (initContactsWithPreviousOnes, :1069-1160); the old points are held in
    // registers (fixed-trip loops over the four cache slots)
    float4 fv1 = make_float4(0.f, 0.f, 0.f, 0.f), fv2 = fv1, fimp = fv1;
    float4 oldl1[kMaxManifoldPoints];
    float oldrest[kMaxManifoldPoints];
#pragma unroll
    for (int k = 0; k < kMaxManifoldPoints; k++) {
        oldl1[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        oldrest[k] = 0.0f;
    }
    if (oldN > 0) {
        const float4 oldn0 = O.pn[4 * prev];
        if (dot(v3(oldn0.x, oldn0.y, oldn0.z), raw.normal(idx.get(0))) >= A.cos_similar) {
            fv1 = O.fv1[prev];
            fv2 = O.fv2[prev];
            fimp = O.fimp[prev];
        }
#pragma unroll
        for (int k = 0; k < kMaxManifoldPoints; k++) {
            if (k < oldN) {
                oldl1[k] = O.pl1[4 * prev + k];
                oldrest[k] = O.pl2[4 * prev + k].w;
            }
        }
    }
    const float pt2 = A.persistent_threshold * A.persistent_threshold;
    for (int k = 0; k < n; k++) {
        const int r = idx.get(k);
        V3 pl1, pl2;
        float depth;
        raw.point(r, pl1, pl2, depth);
        float imp = 0.0f, resting = 0.0f;
        bool matched = false;
#pragma unroll
        for (int o = 0; o < kMaxManifoldPoints; o++) {
            if (matched || o >= oldN) continue;
            const float d2 = len2(pl1 - xyz(oldl1[o]));
            if (d2 <= pt2) {
                imp = oldl1[o].w;
                resting = oldrest[o];
                matched = true;
            }
        }
        P.pn[4 * slot + k] = f4(raw.normal(r), depth);
        P.pl1[4 * slot + k] = f4(pl1, imp);
        P.pl2[4 * slot + k] = f4(pl2, resting);
    }
    P.fv1[slot] = fv1;
    P.fv2[slot] = fv2;
    P.fimp[slot] = fimp;
    P.ncp[slot] = n;
}

// processPotentialContacts (:1210-1219): the test updates wasColliding / isValid of every tested pair.
__device__ inline uint32_t lf_after_test(uint32_t lf, bool colliding) {
    lf |= LF_VALID;
    if (colliding) lf |= LF_WAS_COLLIDING;
    else lf &= ~LF_WAS_COLLIDING;
    return lf;
}

// ------------------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------------------

// Middle phase: bucket the tested pairs by algorithm; a pair is tested only if at least one of its
// bodies is enabled (CD.cpp:391-400).  Untested pairs have no contact pair this step.  The list order
// is irrelevant because every test writes only its own pair slot.
// Polyhedron pairs are split further (np_list_slot): box-box pairs that collided in the last step go
// to list slot 0 (ALGO_NONE is never listed) counted in box_count, the other box-box pairs to slot
// ALGO_COUNT counted in box_sep_count (a wave of either kind follows one path through the test), pairs
// with a mesh stay in the ALGO_POLY_POLY slot counted in hull_count; counts[ALGO_POLY_POLY] keeps the
// total.  Each workgroup counts its pairs per bin first and reserves one contiguous range per bin with
// a single global atomic (same-address atomics from every wave serialise in the L2), then lists its
// pairs into those ranges.
__device__ __forceinline__ int np_bin(const NpArgs& A, int k, bool count_pass) {
    const int2 c = A.P.col[k];
    if (count_pass && A.P.xm) {   // no triangle tests and no extra manifolds unless the concave kernels list them
        A.P.ctn[k] = 0;
        A.P.xm[k] = 0;
    }
    const int f1 = A.B.flags[A.C.body[c.x]], f2 = A.B.flags[A.C.body[c.y]];
    // tested only if a body is enabled and both colliders are simulation colliders or triggers (CD.cpp:386-400);
    // a world query tests the pairs of world-query colliders (:386), a query about bodies the pairs of those
    // bodies (filterOverlappingPairs, :1898-1940), among the enabled pairs in both cases
    const int k1 = A.C.kind[c.x], k2 = A.C.kind[c.y];
    bool tested = (f1 & BF_ENABLED) || (f2 & BF_ENABLED);
    if (A.query == NP_STEP) {
        tested = tested && (k1 & (CK_SIMULATION | CK_TRIGGER)) && (k2 & (CK_SIMULATION | CK_TRIGGER));
    } else if (A.query == NP_QUERY_ALL) {
        tested = tested && (k1 & CK_WORLD_QUERY) && (k2 & CK_WORLD_QUERY);
    } else {
        const int b1 = A.C.body[c.x], b2 = A.C.body[c.y];
        tested = tested && (A.qbody2 < 0 ? (b1 == A.qbody1 || b2 == A.qbody1)
                                         : ((b1 == A.qbody1 && b2 == A.qbody2) || (b1 == A.qbody2 && b2 == A.qbody1)));
    }
    if (!tested) {
        if (count_pass) A.P.ncp[k] = 0;
        return 0;
    }
    const int a = A.P.algo[k];
    if (a == ALGO_POLY_POLY && A.C.type[c.x] == SHAPE_BOX && A.C.type[c.y] == SHAPE_BOX)   // box-box
        return (A.P.lf[k] & LF_WAS_COLLIDING) ? ALGO_COUNT : ALGO_COUNT + 1;
    return a;
}

__global__ __launch_bounds__(256) void np_classify(NpArgs A, int* lists, int list_stride, int* counts, int* box_count,
                                                   int* box_sep_count, int* hull_count) {
    constexpr int kBins = kNpBins;
    __shared__ int s_cnt[kBins];
    __shared__ int s_base[kBins];
    __shared__ int s_cur[kBins];
    const int np = *A.nb_pairs;
    const int stride = gridDim.x * blockDim.x;
    const int lane = lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    if (threadIdx.x < kBins) {
        s_cnt[threadIdx.x] = 0;
        s_cur[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int base = blockIdx.x * blockDim.x; base < np; base += stride) {
        const int k = base + threadIdx.x;
        const int b = k < np ? np_bin(A, k, true) : 0;
        for (int t = 1; t < kBins; t++) {
            const uint64_t m = __ballot(b == t);
            if (m && lane == __ffsll((unsigned long long)m) - 1) atomicAdd(&s_cnt[t], __popcll(m));
        }
    }
    __syncthreads();
    if (threadIdx.x < kBins && threadIdx.x > 0) {
        const int t = threadIdx.x, n = s_cnt[t];
        int* ctr = t == ALGO_COUNT ? box_count
                   : t == ALGO_COUNT + 1 ? box_sep_count
                   : t == ALGO_POLY_POLY ? hull_count : &counts[t];
        s_base[t] = n ? atomicAdd(ctr, n) : 0;
        if (t == ALGO_POLY_POLY) {
            const int polys = n + s_cnt[ALGO_COUNT] + s_cnt[ALGO_COUNT + 1];
            if (polys) atomicAdd(&counts[ALGO_POLY_POLY], polys);
        }
    }
    __syncthreads();
    for (int base = blockIdx.x * blockDim.x; base < np; base += stride) {
        const int k = base + threadIdx.x;
        const int b = k < np ? np_bin(A, k, false) : 0;
        for (int t = 1; t < kBins; t++) {
            const uint64_t m = __ballot(b == t);
            if (m == 0) continue;
            const int leader = __ffsll((unsigned long long)m) - 1;
            int pos = 0;
            if (lane == leader) pos = atomicAdd(&s_cur[t], __popcll(m));
            pos = __shfl(pos, leader, kWave) + s_base[t] + __popcll(m & lt);
            if (b == t) lists[np_list_slot(t) * list_stride + pos] = k;
        }
    }
}

__global__ void np_sphere_sphere(NpArgs A, const int* list, const int* count) {
    RP3DG_GRID_STRIDE(t, *count) {
        const int k = list[t];
        const int2 c = A.P.col[k];
        const Tf T1 = tf(ld3(A.C.wpos, c.x), ldq(A.C.wq, c.x));
        const Tf T2 = tf(ld3(A.C.wpos, c.y), ldq(A.C.wq, c.y));
        Raw raw; raw.n = 0;
        const bool hit = sphere_sphere(A.C.shape[c.x].x, A.C.shape[c.y].x, T1, T2, raw);
        A.P.lf[k] = A.lf_update ? lf_after_test(A.P.lf[k], hit) : A.P.lf[k];
        contact_epilogue(A, k, hit, raw, c.x);
    }
}

__global__ void np_sphere_capsule(NpArgs A, const int* list, const int* count) {
    RP3DG_GRID_STRIDE(t, *count) {
        const int k = list[t];
        const int2 c = A.P.col[k];
        const bool s1 = A.C.type[c.x] == SHAPE_SPHERE;
        const int cs = s1 ? c.x : c.y, cc = s1 ? c.y : c.x;
        const Tf Ts = tf(ld3(A.C.wpos, cs), ldq(A.C.wq, cs));
        const Tf Tc = tf(ld3(A.C.wpos, cc), ldq(A.C.wq, cc));
        const float4 cap = A.C.shape[cc];
        Raw raw; raw.n = 0;
        const bool hit = sphere_capsule(s1, A.C.shape[cs].x, cap.x, cap.y + cap.y, Ts, Tc, raw);
        A.P.lf[k] = A.lf_update ? lf_after_test(A.P.lf[k], hit) : A.P.lf[k];
        contact_epilogue(A, k, hit, raw, c.x);
    }
}

__global__ void np_capsule_capsule(NpArgs A, const int* list, const int* count) {
    RP3DG_GRID_STRIDE(t, *count) {
        const int k = list[t];
        const int2 c = A.P.col[k];
        const Tf T1 = tf(ld3(A.C.wpos, c.x), ldq(A.C.wq, c.x));
        const Tf T2 = tf(ld3(A.C.wpos, c.y), ldq(A.C.wq, c.y));
        const float4 s1 = A.C.shape[c.x], s2 = A.C.shape[c.y];
        Raw raw; raw.n = 0;
        const bool hit = capsule_capsule(s1.x, s1.y + s1.y, s2.x, s2.y + s2.y, T1, T2, raw);
        A.P.lf[k] = A.lf_update ? lf_after_test(A.P.lf[k], hit) : A.P.lf[k];
        contact_epilogue(A, k, hit, raw, c.x);
    }
}

// Box-box pairs (the same SAT test with the box topology known at compile time and box-sized buffers)
__device__ __forceinline__ Poly box_poly(const NpArgs& A, int c) {
    Poly P{};
    const float4 s = A.C.shape[c];
    P.hull = -1;
    P.he = v3(s.x, s.y, s.z);
    return P;
}

// The polyhedron test in three stages so that the costly search over the edge pairs can be shared by a
// wavefront.  Stage 1 (one lane per pair): the cached-axis shortcut and both face searches of
// testCollisionConvexPolyhedronVsConvexPolyhedron.  Returns 0 / 1 when the test is decided (no contact / contact), 2 when the edge pairs
// must be searched; minPen0, minFace and isMinFace1 then hold the face result (SATAlgorithm.cpp:485-747).
// With FACES false it returns 3 instead of running the face searches (the caller shares them).
template <bool BOX, bool FACES, class R>
__device__ __forceinline__ int sat_begin(const Poly& P1, const Poly& P2, const Tf& T1, const Tf& T2, const Tf& t12,
                                         const Tf& t21, uint32_t& lf, R& raw, float& minPen0, int& minFace,
                                         bool& isMinFace1, bool clip_prev) {
    const bool wasColliding = (lf & LF_WAS_COLLIDING) != 0;
    // clipWithPreviousAxisIfStillColliding (SATAlgorithm.cpp:538, 578, 636): off in world queries
    const bool clipAgain = wasColliding && clip_prev;
    if ((lf & LF_VALID) && (lf & LF_WAS_SAT)) {
        if (lf & (LF_SAT_FACE1 | LF_SAT_FACE2)) {
            const bool ref1 = (lf & LF_SAT_FACE1) != 0;
            const int f = (int)lf_face(lf);
            const float pen = ref1 ? sat_single_face(P1, P2, t12, f) : sat_single_face(P2, P1, t21, f);
            if (!wasColliding && pen <= 0.0f) return 0;
            if (clipAgain && pen > 0.0f) {
                if (face_contacts<BOX>(ref1, P1, P2, t12, t21, f, T1, T2, raw)) {
                    lf = lf_set_face(lf, ref1, f);
                    return 1;
                }
            }
        } else {
            const int4 e1 = P1.edge((int)lf_edge1(lf));
            const int4 e2 = P2.edge((int)lf_edge2(lf));
            const V3 e1a = tapply(t12, P1.vertex(e1.x));
            const V3 e1b = tapply(t12, P1.vertex(P1.edge(e1.w).x));
            const V3 e1d = e1b - e1a;
            const V3 e2a = P2.vertex(e2.x);
            const V3 e2b = P2.vertex(P2.edge(e2.w).x);
            const V3 e2d = e2b - e2a;
            if (sat_minkowski_face(P1, e1, P2, e2, t12)) {
                V3 axis;
                const V3 c1 = tapply(t12, P1.centroid());
                const float pen = sat_edge_distance(e1a, e2a, c1, e1d, e2d, axis, P1.tri(), P1.tri() ? P2.centroid() : v3zero());
                if (!wasColliding && pen <= 0.0f) return 0;
                if (clipAgain && pen > 0.0f && pen < kDecimalLargest) {
                    V3 cp1, cp2;
                    closest_points_between_segments(e1a, e1b, e2a, e2b, cp1, cp2);
                    const V3 vec1 = cp1 - e2a;
                    const V3 vec2 = cp2 - e1a;
                    const float l1 = len2(e1d);
                    const float l2 = len2(e2d);
                    const float t1 = dot(vec1, e2d) / l2;
                    const float t2 = dot(vec2, e1d) / l1;
                    if (t1 >= 0.0f && t1 <= 1.0f && t2 >= 0.0f && t2 <= 1.0f) {
                        raw.add(qrot(T2.q, axis), pen, tapply(t21, cp1), cp2);
                        return 1;
                    }
                }
            }
        }
    }
    if constexpr (!FACES) return 3;
    int face1 = 0, face2 = 0;
    const float pen1 = sat_faces(P1, P2, t12, face1);
    if (pen1 <= 0.0f) { lf = lf_set_face(lf, true, face1); return 0; }
    const float pen2 = sat_faces(P2, P1, t21, face2);
    if (pen2 <= 0.0f) { lf = lf_set_face(lf, false, face2); return 0; }
    minPen0 = fmin_s(pen1, pen2);
    isMinFace1 = pen1 < pen2 * kSatRel + kSatAbs;
    minFace = isMinFace1 ? face1 : face2;
    return 2;
}

// Edge pair k in the order of the reference's double loop over the half-edges i = 0, 2, .. of shape 1
// and j = 0, 2, .. of shape 2 (k = (i / 2) * h2 + j / 2, h2 = half-edges of shape 2 / 2); false when the
// edges do not build a face of the Minkowski difference.
__device__ __forceinline__ int sat_edge_pairs(const Poly& P1, const Poly& P2, int& h2) {
    h2 = (P2.nhalfedges() + 1) / 2;
    return ((P1.nhalfedges() + 1) / 2) * h2;
}
__device__ __forceinline__ bool sat_edge_pair(const Poly& P1, const Poly& P2, const Tf& t12, int k, int h2, float& pen,
                                              V3* axis_out = nullptr) {
    const int i = 2 * (k / h2), j = 2 * (k % h2);
    const int4 e1 = P1.edge(i);
    const int4 e2 = P2.edge(j);
    if (!sat_minkowski_face(P1, e1, P2, e2, t12)) return false;
    const V3 e1a = tapply(t12, P1.vertex(e1.x));
    const V3 e1b = tapply(t12, P1.vertex(P1.edge(e1.w).x));
    const V3 e2a = P2.vertex(e2.x);
    const V3 e2b = P2.vertex(P2.edge(e2.w).x);
    V3 axis = v3zero();
    pen = sat_edge_distance(e1a, e2a, tapply(t12, P1.centroid()), e1b - e1a, e2b - e2a, axis, P1.tri(),
                            P1.tri() ? P2.centroid() : v3zero());
    if (axis_out) *axis_out = axis;
    return true;
}

// The reference walks the edge pairs in order and returns at the first separating pair; otherwise an
// edge pair replaces the face result when pen * kSatRel + kSatAbs < the face penetration, and later
// pairs replace it when strictly smaller (SATAlgorithm.cpp:749-815).  Since rounding is monotonic, the
// pair it ends with is the first pair of smallest penetration, provided that this one passes the face
// test; so the walk reduces to (first separating k, first k of the smallest positive penetration),
// which lanes can compute over disjoint subsets of k and combine.
struct EdgeSearch {
    int sep;     // first separating pair, INT_MAX when none
    int k;       // first pair of the smallest penetration among the others, INT_MAX when none
    float pen;   // its penetration (+inf when none)
    __device__ void init(float start = __builtin_huge_valf()) { sep = INT_MAX; k = INT_MAX; pen = start; }
    __device__ void visit(int kk, bool valid, float p) {   // in increasing kk within one lane
        if (!valid) return;
        if (p <= 0.0f) { if (kk < sep) sep = kk; return; }
        if (p < pen) { pen = p; k = kk; }
    }
    __device__ void combine(const EdgeSearch& o) {
        if (o.sep < sep) sep = o.sep;
        if (o.pen < pen || (o.pen == pen && o.k < k)) { pen = o.pen; k = o.k; }
    }
};

// Stage 3: the choice between the face and the edge result (SATAlgorithm.cpp:828-880), then the contacts.
template <bool BOX, class R>
__device__ __forceinline__ bool sat_finish(const Poly& P1, const Poly& P2, const Tf& T1, const Tf& T2, const Tf& t12,
                                           const Tf& t21, uint32_t& lf, R& raw, float minPen0, int minFace,
                                           bool isMinFace1, const EdgeSearch& es) {
    int h2;
    sat_edge_pairs(P1, P2, h2);
    if (es.sep != INT_MAX) {
        lf = lf_set_edges(lf, 2 * (es.sep / h2), 2 * (es.sep % h2));
        return false;
    }
    if (es.k != INT_MAX && es.pen * kSatRel + kSatAbs < minPen0) {
        const int i = 2 * (es.k / h2), j = 2 * (es.k % h2);
        float pen;
        V3 axis;
        sat_edge_pair(P1, P2, t12, es.k, h2, pen, &axis);
        const int4 e1 = P1.edge(i);
        const int4 e2 = P2.edge(j);
        const V3 s1a = tapply(t12, P1.vertex(e1.x));
        const V3 s1b = tapply(t12, P1.vertex(P1.edge(e1.w).x));
        const V3 s2a = P2.vertex(e2.x);
        const V3 s2b = P2.vertex(P2.edge(e2.w).x);
        V3 cp1, cp2;
        closest_points_between_segments(s1a, s1b, s2a, s2b, cp1, cp2);
        raw.add(qrot(T2.q, axis), es.pen, tapply(t21, cp1), cp2);
        lf = lf_set_edges(lf, i, j);
        return true;
    }
    const bool found = face_contacts<BOX>(isMinFace1, P1, P2, t12, t21, minFace, T1, T2, raw);
    lf = lf_set_face(lf, isMinFace1, minFace);
    return found;
}

// One search shared by the wavefront for every lane in m (wave-uniform): the lane's colliders (ca, cb)
// and transform t are broadcast, each lane evaluates every 64th candidate kk < count(B1, B2) with
// eval(B1, B2, tb, kk, pen) -> valid, and the combined result lands in es of the lane it belongs to.
template <bool BOX, class Count, class Eval>
__device__ __forceinline__ void shared_search(const NpArgs& A, uint64_t m, int lane, const Tf& t, int ca, int cb,
                                              float start, EdgeSearch& es, Count count, Eval eval) {
    while (m) {
        const int L = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        Tf tb;
        tb.p = v3(__shfl(t.p.x, L), __shfl(t.p.y, L), __shfl(t.p.z, L));
        tb.q.x = __shfl(t.q.x, L); tb.q.y = __shfl(t.q.y, L);
        tb.q.z = __shfl(t.q.z, L); tb.q.w = __shfl(t.q.w, L);
        const int cx = __shfl(ca, L), cy = __shfl(cb, L);
        const Poly B1 = BOX ? box_poly(A, cx) : make_poly(A, cx);
        const Poly B2 = BOX ? box_poly(A, cy) : make_poly(A, cy);
        const int np = count(B1, B2);
        EdgeSearch w;
        w.init(start);
        for (int kk = lane; kk < np; kk += kWave) {
            float p = 0.0f;
            const bool v = eval(B1, B2, tb, kk, p);
            w.visit(kk, v, p);
        }
        for (int off = kWave / 2; off > 0; off >>= 1) {
            EdgeSearch o;
            o.sep = __shfl_xor(w.sep, off);
            o.k = __shfl_xor(w.k, off);
            o.pen = __shfl_xor(w.pen, off);
            w.combine(o);
        }
        if (lane == L) es = w;
    }
}

// sat_faces (SATAlgorithm.cpp:1119-1143) as a shared search: the first face with pen <= 0, else the
// first face of the smallest penetration below kDecimalLargest (face 0 and kDecimalLargest when none).
__device__ __forceinline__ float face_search_result(const EdgeSearch& es, int& face) {
    if (es.sep != INT_MAX) { face = es.sep; return 0.0f; }
    if (es.k == INT_MAX) { face = 0; return kDecimalLargest; }
    face = es.k;
    return es.pen;
}

// Polyhedron pairs, one lane per pair.  The edge search of a pair that needs one is shared by the whole
// wavefront (each lane takes every 64th edge pair and the results are combined across lanes) while at
// most coop_max lanes need it; beyond that every such lane walks its own pairs.  Both give the same
// result.
template <bool BOX, class R>
__device__ __forceinline__ void sat_pair_list(const NpArgs& A, const int* list, int n, int coop_max) {
    const int lane = lane_id();
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {   // uniform in the block
        const int t = base + threadIdx.x;
        const bool act = t < n;
        int k = 0;
        int2 c = make_int2(0, 0);
        Poly P1{}, P2{};
        if constexpr (BOX) {   // inactive lanes too: every access folds to the box's implicit topology
            P1.hull = -1;
            P2.hull = -1;
        }
        Tf T1{}, T2{}, t12{}, t21{};
        R raw;
        raw.n = 0;
        uint32_t lf = 0;
        int stage = 0;
        float minPen0 = 0.0f;
        int minFace = 0;
        bool isMinFace1 = false;
        if (act) {
            k = list[t];
            c = A.P.col[k];
            T1 = tf(ld3(A.C.wpos, c.x), ldq(A.C.wq, c.x));
            T2 = tf(ld3(A.C.wpos, c.y), ldq(A.C.wq, c.y));
            P1 = BOX ? box_poly(A, c.x) : make_poly(A, c.x);
            P2 = BOX ? box_poly(A, c.y) : make_poly(A, c.y);
            lf = A.P.lf[k];
            t12 = tmul(tinv(T2), T1);
            t21 = tinv(t12);
            stage = sat_begin<BOX, BOX>(P1, P2, T1, T2, t12, t21, lf, raw, minPen0, minFace, isMinFace1, A.clip_prev != 0);
        }
        if constexpr (!BOX) {
            // mesh pairs share the two face searches as well (each face's support point walks every
            // vertex of the other mesh)
            auto nfaces = [](const Poly& B1, const Poly&) { return B1.nfaces(); };
            auto face_pen = [](const Poly& B1, const Poly& B2, const Tf& tb, int f, float& p) {
                p = sat_single_face(B1, B2, tb, f);
                return true;
            };
            EdgeSearch fs;
            fs.init(kDecimalLargest);
            shared_search<BOX>(A, __ballot(stage == 3), lane, t12, c.x, c.y, kDecimalLargest, fs, nfaces, face_pen);
            int face1 = 0, face2 = 0;
            float pen1 = 0.0f;
            if (stage == 3) {
                pen1 = face_search_result(fs, face1);
                if (fs.sep != INT_MAX) {
                    lf = lf_set_face(lf, true, face1);
                    stage = 0;
                }
            }
            fs.init(kDecimalLargest);
            shared_search<BOX>(A, __ballot(stage == 3), lane, t21, c.y, c.x, kDecimalLargest, fs, nfaces, face_pen);
            if (stage == 3) {
                const float pen2 = face_search_result(fs, face2);
                if (fs.sep != INT_MAX) {
                    lf = lf_set_face(lf, false, face2);
                    stage = 0;
                } else {
                    minPen0 = fmin_s(pen1, pen2);
                    isMinFace1 = pen1 < pen2 * kSatRel + kSatAbs;
                    minFace = isMinFace1 ? face1 : face2;
                    stage = 2;
                }
            }
        }
        EdgeSearch es;
        es.init();
        uint64_t m = __ballot(stage == 2);
        if (__popcll(m) > coop_max) {
            if (stage == 2) {
                int h2;
                const int np = sat_edge_pairs(P1, P2, h2);
                for (int kk = 0; kk < np; kk++) {
                    float p = 0.0f;
                    const bool v = sat_edge_pair(P1, P2, t12, kk, h2, p);
                    es.visit(kk, v, p);
                    if (es.sep != INT_MAX) break;
                }
            }
        } else {
            auto npairs = [](const Poly& B1, const Poly& B2) {
                int h2;
                return sat_edge_pairs(B1, B2, h2);
            };
            auto edge_pen = [](const Poly& B1, const Poly& B2, const Tf& tb, int kk, float& p) {
                int h2;
                sat_edge_pairs(B1, B2, h2);
                return sat_edge_pair(B1, B2, tb, kk, h2, p);
            };
            shared_search<BOX>(A, m, lane, t12, c.x, c.y, __builtin_huge_valf(), es, npairs, edge_pen);
        }
        if (act) {
            bool hit = stage == 1;
            if (stage == 2)
                hit = sat_finish<BOX>(P1, P2, T1, T2, t12, t21, lf, raw, minPen0, minFace, isMinFace1, es);
            // ConvexPolyhedronVsConvexPolyhedronAlgorithm.cpp:55-62
            lf |= LF_WAS_SAT;
            lf &= ~LF_WAS_GJK;
            A.P.lf[k] = A.lf_update ? lf_after_test(lf, hit) : lf;
            contact_epilogue(A, k, hit, raw, c.x);
        }
    }
}

// Pairs with a convex mesh (the general buffers); their edge searches run over hundreds of pairs, so
// they are always shared.
__global__ __launch_bounds__(128) void np_poly_poly(NpArgs A, const int* list, const int* count) {
    sat_pair_list<false, Raw>(A, list, *count, kWave);
}

// Box-box pairs: the box topology is known at compile time and the buffers are box-sized.  A lane's own
// walk over the 144 edge pairs costs about as much as 32 shared searches, hence the limit.
__global__ __launch_bounds__(kNpBoxThreads) void np_box_box(NpArgs A, const int* list, const int* count) {
    sat_pair_list<true, RawBoxView>(A, list, *count, 32);
}

void upload_box_topology(const int* edges4, const int* face_edge, const int* face_verts) {
    hipMemcpyToSymbol(HIP_SYMBOL(c_box_edges), edges4, sizeof(int4) * 24);
    hipMemcpyToSymbol(HIP_SYMBOL(c_box_face_edge), face_edge, sizeof(int) * 6);
    hipMemcpyToSymbol(HIP_SYMBOL(c_box_face_verts), face_verts, sizeof(int) * 24);
}


}  // namespace rp3dg

namespace rp3dg {

// ------------------------------------------------------------------------------------------------
// GJK on the core shapes with margins (GJKAlgorithm.cpp:50-238, VoronoiSimplex.cpp)
// ------------------------------------------------------------------------------------------------

struct Conv {
    int type;        // SHAPE_SPHERE / SHAPE_CAPSULE / polyhedron
    float margin;    // sphere/capsule radius, 0 for polyhedra (ConvexShape.h:62)
    float hh;        // capsule half height
    Poly P;
    // SphereShape.h:149-153, CapsuleShape.h:191-206, BoxShape.h:167-172, ConvexMeshShape.cpp:80-102
    __device__ V3 support(V3 d) const {
        if (type == SHAPE_SPHERE) return v3(0.0f, 0.0f, 0.0f);
        if (type == SHAPE_CAPSULE) {
            const float top = hh * d.y;
            const float bottom = -hh * d.y;
            return top > bottom ? v3(0.0f, hh, 0.0f) : v3(0.0f, -hh, 0.0f);
        }
        return P.support(d);
    }
};

__device__ inline Conv make_conv(const NpArgs& A, int collider) {
    Conv c;
    c.type = A.C.type[collider];
    const float4 s = A.C.shape[collider];
    c.hh = 0.0f;
    c.margin = 0.0f;
    c.P.hull = -1;
    c.P.he = v3zero();
    if (c.type == SHAPE_SPHERE) c.margin = s.x;
    else if (c.type == SHAPE_CAPSULE) { c.margin = s.x; c.hh = s.y; }
    else c.P = make_poly(A, collider);
    return c;
}

constexpr float kSimplexEps = 0.0001f;   // VoronoiSimplex.cpp:33

struct Simplex {
    int n;
    V3 pts[4], sa[4], sb[4];
    V3 closest, csa, csb;
    float bary[4];
    bool recompute, valid;

    __device__ void init() { n = 0; recompute = false; valid = false; }
    __device__ bool full() const { return n == 4; }
    __device__ void add(V3 p, V3 a, V3 b) { pts[n] = p; sa[n] = a; sb[n] = b; n++; recompute = true; }
    __device__ void remove(int i) { n--; pts[i] = pts[n]; sa[i] = sa[n]; sb[i] = sb[n]; }
    __device__ void reduce(int bits) {
        if (n >= 4 && (bits & 8) == 0) remove(3);
        if (n >= 3 && (bits & 4) == 0) remove(2);
        if (n >= 2 && (bits & 2) == 0) remove(1);
        if (n >= 1 && (bits & 1) == 0) remove(0);
    }
    __device__ bool contains(V3 p) const {
        for (int i = 0; i < n; i++)
            if (len2(pts[i] - p) <= kSimplexEps) return true;
        return false;
    }
    __device__ bool affinely_dependent() const {
        switch (n) {
            case 2: return len2(pts[1] - pts[0]) <= kSimplexEps;
            case 3: return len2(cross(pts[1] - pts[0], pts[2] - pts[0])) <= kSimplexEps;
            case 4: return fabsf(dot(pts[1] - pts[0], cross(pts[2] - pts[0], pts[3] - pts[0]))) <= kSimplexEps;
        }
        return false;
    }
    __device__ float max_len2() const {
        float m = 0.0f;
        for (int i = 0; i < n; i++) {
            const float l = len2(pts[i]);
            if (l > m) m = l;
        }
        return m;
    }
    __device__ void set_bary(float a, float b, float c, float d) { bary[0] = a; bary[1] = b; bary[2] = c; bary[3] = d; }
    __device__ bool bary_valid() const { return bary[0] >= 0.0f && bary[1] >= 0.0f && bary[2] >= 0.0f && bary[3] >= 0.0f; }

    // VoronoiSimplex.cpp:296-331
    __device__ static void closest_on_segment(V3 a, V3 b, int& bits, float& t) {
        const V3 ap = -a;
        const V3 ab = b - a;
        const float apab = dot(ap, ab);
        if (apab > 0.0f) {
            const float l2 = len2(ab);
            if (apab < l2) { t = apab / l2; bits = 3; }
            else { t = 1.0f; bits = 2; }
        } else { t = 0.0f; bits = 1; }
    }
    // VoronoiSimplex.cpp:339-427 (Ericson RTCD 5.1.5)
    __device__ static void closest_on_triangle(V3 a, V3 b, V3 c, int& bits, V3& bc) {
        const V3 ab = b - a, ac = c - a, ap = -a;
        const float d1 = dot(ab, ap), d2 = dot(ac, ap);
        if (d1 <= 0.0f && d2 <= 0.0f) { bc = v3(1.0f, 0, 0); bits = 1; return; }
        const V3 bp = -b;
        const float d3 = dot(ab, bp), d4 = dot(ac, bp);
        if (d3 >= 0.0f && d4 <= d3) { bc = v3(0.0f, 1.0f, 0); bits = 2; return; }
        const float vc = d1 * d4 - d3 * d2;
        if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
            const float v = d1 / (d1 - d3);
            bc = v3(1.0f - v, v, 0); bits = 3; return;
        }
        const V3 cp = -c;
        const float d5 = dot(ab, cp), d6 = dot(ac, cp);
        if (d6 >= 0.0f && d5 <= d6) { bc = v3(0.0f, 0.0f, 1.0f); bits = 4; return; }
        const float vb = d5 * d2 - d1 * d6;
        if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
            const float w = d2 / (d2 - d6);
            bc = v3(1.0f - w, 0, w); bits = 5; return;
        }
        const float va = d3 * d6 - d5 * d4;
        if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
            const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
            bc = v3(0.0f, 1.0f - w, w); bits = 6; return;
        }
        const float denom = 1.0f / (va + vb + vc);
        const float v = vb * denom;
        const float w = vc * denom;
        bc = v3(1 - v - w, v, w);
        bits = 7;
    }
    // VoronoiSimplex.cpp:601-617
    __device__ static int outside_plane(V3 a, V3 b, V3 c, V3 d) {
        const V3 nrm = cross(b - a, c - a);
        const float sp = dot(-a, nrm);
        const float sd = dot(d - a, nrm);
        if (sd * sd < kSimplexEps * kSimplexEps) return -1;
        return sp * sd < 0.0f;
    }
    __device__ static int map_tri(int tb, int f, int s, int t) {
        return (((1 & tb) != 0) << f) | (((2 & tb) != 0) << s) | (((4 & tb) != 0) << t);
    }
    // VoronoiSimplex.cpp:434-598
    __device__ static bool closest_on_tetra(V3 a, V3 b, V3 c, V3 d, int& bits, float& ab0, float& ab1, float& cd0,
                                            float& cd1, bool& degenerate) {
        degenerate = false;
        bits = 15;
        ab0 = ab1 = cd0 = cd1 = 0.0f;
        const int oABC = outside_plane(a, b, c, d);
        const int oACD = outside_plane(a, c, d, b);
        const int oADB = outside_plane(a, d, b, c);
        const int oBDC = outside_plane(b, d, c, a);
        if (oABC < 0 || oACD < 0 || oADB < 0 || oBDC < 0) { degenerate = true; return false; }
        if (oABC == 0 && oACD == 0 && oADB == 0 && oBDC == 0) return true;
        float best = kDecimalLargest;
        int tb;
        V3 bc;
        if (oABC) {
            closest_on_triangle(a, b, c, tb, bc);
            const V3 cp = bc.x * a + bc.y * b + bc.z * c;
            const float sq = len2(cp);
            if (sq < best) { best = sq; ab0 = bc.x; ab1 = bc.y; cd0 = bc.z; cd1 = 0.0f; bits = tb; }
        }
        if (oACD) {
            closest_on_triangle(a, c, d, tb, bc);
            const V3 cp = bc.x * a + bc.y * c + bc.z * d;
            const float sq = len2(cp);
            if (sq < best) { best = sq; ab0 = bc.x; ab1 = 0.0f; cd0 = bc.y; cd1 = bc.z; bits = map_tri(tb, 0, 2, 3); }
        }
        if (oADB) {
            closest_on_triangle(a, d, b, tb, bc);
            const V3 cp = bc.x * a + bc.y * d + bc.z * b;
            const float sq = len2(cp);
            if (sq < best) { best = sq; ab0 = bc.x; ab1 = bc.z; cd0 = 0.0f; cd1 = bc.y; bits = map_tri(tb, 0, 3, 1); }
        }
        if (oBDC) {
            closest_on_triangle(b, d, c, tb, bc);
            const V3 cp = bc.x * b + bc.y * d + bc.z * c;
            const float sq = len2(cp);
            if (sq < best) { ab0 = 0.0f; ab1 = bc.x; cd0 = bc.z; cd1 = bc.y; bits = map_tri(tb, 1, 3, 2); }
        }
        return true;
    }
    // VoronoiSimplex.cpp:165-289
    __device__ bool recompute_closest() {
        if (recompute) {
            recompute = false;
            switch (n) {
                case 0: valid = false; break;
                case 1:
                    closest = pts[0]; csa = sa[0]; csb = sb[0];
                    set_bary(1, 0, 0, 0);
                    valid = bary_valid();
                    break;
                case 2: {
                    int bits = 0;
                    float t;
                    closest_on_segment(pts[0], pts[1], bits, t);
                    csa = sa[0] + t * (sa[1] - sa[0]);
                    csb = sb[0] + t * (sb[1] - sb[0]);
                    closest = csa - csb;
                    set_bary(1.0f - t, t, 0, 0);
                    valid = bary_valid();
                    reduce(bits);
                } break;
                case 3: {
                    int bits = 0;
                    V3 bc;
                    closest_on_triangle(pts[0], pts[1], pts[2], bits, bc);
                    csa = bc.x * sa[0] + bc.y * sa[1] + bc.z * sa[2];
                    csb = bc.x * sb[0] + bc.y * sb[1] + bc.z * sb[2];
                    closest = csa - csb;
                    set_bary(bc.x, bc.y, bc.z, 0.0f);
                    valid = bary_valid();
                    reduce(bits);
                } break;
                case 4: {
                    int bits = 0;
                    float ab0, ab1, cd0, cd1;
                    bool degenerate;
                    const bool outside = closest_on_tetra(pts[0], pts[1], pts[2], pts[3], bits, ab0, ab1, cd0, cd1, degenerate);
                    if (outside) {
                        csa = ab0 * sa[0] + ab1 * sa[1] + cd0 * sa[2] + cd1 * sa[3];
                        csb = ab0 * sb[0] + ab1 * sb[1] + cd0 * sb[2] + cd1 * sb[3];
                        closest = csa - csb;
                        set_bary(ab0, ab1, cd0, cd1);
                        reduce(bits);
                    } else {
                        if (degenerate) valid = false;
                        else {
                            set_bary(0.0f, 0.0f, 0.0f, 0.0f);
                            csa = csb = closest = v3zero();
                            valid = true;
                        }
                        break;
                    }
                    valid = bary_valid();
                } break;
            }
        }
        return valid;
    }
};

enum GjkResult { GJK_SEPARATED = 0, GJK_COLLIDE_IN_MARGIN = 1, GJK_INTERPENETRATE = 2 };

// GJKAlgorithm::testCollision for one pair; lf_axis is LastFrameCollisionInfo::gjkSeparatingAxis.
__device__ int gjk(const Conv& S1, const Conv& S2, const Tf& T1, const Tf& T2, uint32_t lf, float4& lf_axis, Raw& raw) {
    const float relErrSq = 1.0e-3f * 1.0e-3f;   // REL_ERROR_SQUARE (GJKAlgorithm.h:46)
    const Tf t1inv = tinv(T1);
    const Tf b2to1 = tmul(t1inv, T2);
    const Q4 rotTo2 = qmul(qconj(T2.q), T1.q);
    const float margin = S1.margin + S2.margin;
    const float marginSq = margin * margin;
    Simplex sx;
    sx.init();
    V3 v;
    if ((lf & LF_VALID) && (lf & LF_WAS_GJK)) v = xyz(lf_axis);
    else v = v3(0.0f, 1.0f, 0.0f);
    float distSq = kDecimalLargest;
    bool contact = false;
    V3 suppA, suppB, w;
    int iters = 0;
    do {
        if (++iters > 64) { contact = true; break; }   // safety bound; the reference terminates far earlier
        suppA = S1.support(-v);
        suppB = tapply(b2to1, S2.support(qrot(rotTo2, v)));
        w = suppA - suppB;
        const float vw = dot(v, w);
        if (vw > 0.0f && vw * vw > distSq * marginSq) {
            lf_axis = f4(v, 0.f);
            return GJK_SEPARATED;
        }
        if (sx.contains(w) || distSq - vw <= distSq * relErrSq) { contact = true; break; }
        sx.add(w, suppA, suppB);
        if (sx.affinely_dependent()) { contact = true; break; }
        const bool ok = sx.recompute_closest();
        v = sx.closest;
        if (!ok) { contact = true; break; }
        const float prev = distSq;
        distSq = len2(v);
        if (prev - distSq <= kMachineEps * prev) {
            v = sx.closest;
            distSq = len2(v);
            contact = true;
            break;
        }
    } while (!sx.full() && distSq > kMachineEps * sx.max_len2());
    if (contact && distSq > kMachineEps) {
        V3 pA = sx.csa, pB = sx.csb;
        const float dist = sqrtf(distSq);
        pA = pA - (S1.margin / dist) * v;
        pB = tapply(tinv(b2to1), pB + (S2.margin / dist) * v);
        const V3 normal = qrot(T1.q, -unit(v));
        const float pen = margin - dist;
        if (pen <= 0.0f) return GJK_SEPARATED;
        if (len2(normal) < kMachineEps) return GJK_SEPARATED;
        raw.add(normal, pen, pA, pB);
        return GJK_COLLIDE_IN_MARGIN;
    }
    return GJK_INTERPENETRATE;
}

// SAT sphere vs polyhedron (SATAlgorithm.cpp:58-144)
__device__ bool sat_sphere_poly(bool sphere1, float r, const Poly& P, const Tf& Ts, const Tf& Tp, Raw& raw) {
    const Tf s2p = tmul(tinv(Tp), Ts);
    const V3 c = s2p.p;
    float minPen = kDecimalLargest;
    int minFace = 0;
    const int nf = P.nfaces();
    for (int f = 0; f < nf; f++) {
        const V3 fn = P.face_normal(f);
        const float pen = dot(P.vertex(P.face_vertex(f, 0)) - c, fn) + r;
        if (pen <= 0.0f) return false;
        if (pen < minPen) { minPen = pen; minFace = f; }
    }
    const V3 mfn = P.face_normal(minFace);
    const V3 mfnw = qrot(Tp.q, mfn);
    const V3 cpS = qrot(tinv(Ts).q, (-mfnw) * r);
    const V3 cpP = c + mfn * (minPen - r);
    const V3 nw = sphere1 ? -mfnw : mfnw;
    raw.add(nw, minPen, sphere1 ? cpS : cpP, sphere1 ? cpP : cpS);
    return true;
}

// ConvexShape::getLocalSupportPointWithMargin (ConvexShape.cpp:40-56) for a capsule
__device__ inline V3 capsule_support_margin(float r, float hh, V3 d) {
    const float top = hh * d.y, bottom = -hh * d.y;
    V3 s = top > bottom ? v3(0.0f, hh, 0.0f) : v3(0.0f, -hh, 0.0f);
    V3 u = v3(0.0f, -1.0f, 0.0f);
    if (len2(d) > kMachineEps * kMachineEps) u = unit(d);
    s += u * r;
    return s;
}

// clipSegmentWithPlanes (mathematics_functions.h:227-289); planes given by the reference face's edges.
__device__ int clip_segment_face(const Poly& P, int face, V3 n, V3 a, V3 b, V3* out, V3& plane0) {
    V3 in0 = a, in1 = b;
    int nin = 2;
    const int first = P.face_edge(face);
    int e = first;
    bool firstPlane = true;
    int guard = P.nhalfedges();
    do {
        if (--guard < 0) break;
        const int4 edge = P.edge(e);
        const int4 twin = P.edge(edge.y);
        const V3 ev1 = P.vertex(edge.x);
        const V3 ev2 = P.vertex(twin.x);
        const V3 pn = cross(n, ev2 - ev1);
        const V3 pp = P.vertex(edge.x);
        if (firstPlane) { plane0 = pp; firstPlane = false; }
        if (nin == 0) return 0;
        V3 o[2];
        int no = 0;
        const float d1 = dot(in0 - pp, pn);
        const float d2 = dot(in1 - pp, pn);
        if (d2 >= 0.0f) {
            if (d1 < 0.0f) {
                const float t = plane_segment_intersection(in0, in1, dot(pn, pp), pn);
                if (t >= 0.0f && t <= 1.0f) o[no++] = in0 + t * (in1 - in0);
                else o[no++] = in1;
            } else {
                o[no++] = in0;
            }
            o[no++] = in1;
        } else {
            if (d1 >= 0.0f) {
                o[no++] = in0;
                const float t = plane_segment_intersection(in0, in1, -dot(pn, pp), -pn);
                if (t >= 0.0f && t <= 1.0f) o[no++] = in0 + t * (in1 - in0);
            }
        }
        nin = no;
        if (no > 0) in0 = o[0];
        if (no > 1) in1 = o[1];
        e = edge.w;
    } while (e != first);
    if (nin > 0) out[0] = in0;
    if (nin > 1) out[1] = in1;
    return nin;
}

// SATAlgorithm::computeCapsulePolyhedronFaceContactPoints (:381-470).  Returns contactFound; on the first
// contact point the raw buffer is reset (NarrowPhaseInfoBatch::resetContactPoints).
__device__ bool capsule_face_contacts(int face, float r, const Poly& P, float pen, const Tf& p2c, V3 normalWorld,
                                      V3 sepAxisCap, V3 segA, V3 segB, bool capsule1, Raw& raw) {
    const V3 fn = P.face_normal(face);
    V3 clip[2];
    V3 plane0 = v3zero();
    const int nclip = clip_segment_face(P, face, fn, segA, segB, clip, plane0);
    const V3 delta = fn * (pen - r);
    bool found = false;
    for (int i = 0; i < nclip; i++) {
        const float cpd = dot(plane0 - clip[i], fn);
        if (cpd > pen - r - 0.001f) {
            if (!found) raw.n = 0;
            found = true;
            const V3 cpP = clip[i] + delta;
            const V3 cpC = tapply(p2c, clip[i]) - sepAxisCap * r;
            raw.add_carry(normalWorld, pen, capsule1 ? cpC : cpP, capsule1 ? cpP : cpC);
        }
    }
    return found;
}

// SATAlgorithm::testCollisionCapsuleVsConvexPolyhedron (:165-319)
__device__ bool sat_capsule_poly(bool capsule1, float r, float hh, const Poly& P, const Tf& Tc, const Tf& Tp, Raw& raw,
                                 bool report) {
    const Tf p2c = tmul(tinv(Tc), Tp);
    const float height = hh + hh;
    const V3 segA = v3(0.0f, -height * 0.5f, 0.0f);
    const V3 segB = v3(0.0f, height * 0.5f, 0.0f);
    const V3 segAxis = segB - segA;
    float minPen = kDecimalLargest;
    int minFace = 0;
    bool isFace = false;
    V3 sepAxis = v3zero(), sev1 = v3zero(), sev2 = v3zero();
    const int nf = P.nfaces();
    for (int f = 0; f < nf; f++) {
        const V3 fn = P.face_normal(f);
        const V3 fnc = qrot(p2c.q, fn);
        const V3 sp = capsule_support_margin(r, hh, -fnc);
        const V3 pf = tapply(p2c, P.vertex(P.face_vertex(f, 0)));
        const float pen = dot(pf - sp, fnc);
        if (pen <= 0.0f) return false;
        if (pen < minPen) { minPen = pen; minFace = f; isFace = true; sepAxis = fnc; }
    }
    const int ne = P.nhalfedges();
    for (int e = 0; e < ne; e += 2) {
        const int4 edge = P.edge(e);
        const V3 ev1 = P.vertex(edge.x);
        const V3 ev2 = P.vertex(P.edge(edge.w).x);
        const V3 edc = qrot(p2c.q, ev2 - ev1);
        const int4 twin = P.edge(edge.y);
        const V3 n1 = qrot(p2c.q, P.face_normal(edge.z));
        const V3 n2 = qrot(p2c.q, P.face_normal(twin.z));
        if (dot(segAxis, n1) * dot(segAxis, n2) < 0.0f) {   // isMinkowskiFaceCapsuleVsEdge (:474-482)
            // computeEdgeVsCapsuleInnerSegmentPenetrationDepth (:323-352)
            float pen = kDecimalLargest;
            V3 axis = cross(segAxis, edc);
            if (len2(axis) >= 0.00001f) {
                const V3 pc = tapply(p2c, P.centroid());
                const V3 pe = tapply(p2c, ev1);
                if (dot(axis, pe - pc) < 0) axis = -axis;
                axis = normalized(axis);
                const V3 sp = capsule_support_margin(r, hh, -axis);
                pen = dot(pe - sp, axis);
            }
            if (pen <= 0.0f) return false;
            if (pen < minPen) { minPen = pen; isFace = false; sepAxis = axis; sev1 = ev1; sev2 = ev2; }
        }
    }
    const Tf c2p = tinv(p2c);
    const V3 segAP = tapply(c2p, segA);
    const V3 segBP = tapply(c2p, segB);
    V3 nw = qrot(Tc.q, sepAxis);
    if (capsule1) nw = -nw;
    // without contact reporting the face case collides whatever the clipping finds (SAT.cpp:278-287)
    if (isFace) return report ? capsule_face_contacts(minFace, r, P, minPen, p2c, nw, sepAxis, segAP, segBP, capsule1, raw)
                              : true;
    V3 cpPoly, cpSeg;
    closest_points_between_segments(segAP, segBP, sev1, sev2, cpSeg, cpPoly);
    const V3 cpCap = tapply(p2c, cpSeg) - sepAxis * r;
    raw.add(nw, minPen, capsule1 ? cpCap : cpPoly, capsule1 ? cpPoly : cpCap);
    return true;
}

// SphereVsConvexPolyhedronAlgorithm.cpp:38-103
__device__ __forceinline__ bool sphere_poly_test(const Conv& S1, const Conv& S2, const Tf& T1, const Tf& T2, uint32_t& lf,
                                                 float4& axis, Raw& raw) {
    const int g = gjk(S1, S2, T1, T2, lf, axis, raw);
    lf |= LF_WAS_GJK;
    lf &= ~LF_WAS_SAT;
    bool hit = false;
    if (g == GJK_COLLIDE_IN_MARGIN) hit = true;
    else if (g == GJK_INTERPENETRATE) {
        const bool s1 = S1.type == SHAPE_SPHERE;
        const Conv& Ss = s1 ? S1 : S2;
        const Conv& Sp = s1 ? S2 : S1;
        hit = sat_sphere_poly(s1, Ss.margin, Sp.P, s1 ? T1 : T2, s1 ? T2 : T1, raw);
        lf &= ~LF_WAS_GJK;
        lf |= LF_WAS_SAT;
    }
    return hit;
}

__global__ void np_sphere_poly(NpArgs A, const int* list, const int* count) {
    RP3DG_GRID_STRIDE(t, *count) {
        const int k = list[t];
        const int2 c = A.P.col[k];
        const Tf T1 = tf(ld3(A.C.wpos, c.x), ldq(A.C.wq, c.x));
        const Tf T2 = tf(ld3(A.C.wpos, c.y), ldq(A.C.wq, c.y));
        const Conv S1 = make_conv(A, c.x), S2 = make_conv(A, c.y);
        uint32_t lf = A.P.lf[k];
        float4 axis = A.P.lfaxis[k];
        Raw raw; raw.n = 0;
        const bool hit = sphere_poly_test(S1, S2, T1, T2, lf, axis, raw);
        A.P.lfaxis[k] = axis;
        A.P.lf[k] = A.lf_update ? lf_after_test(lf, hit) : lf;
        contact_epilogue(A, k, hit, raw, c.x);
    }
}

// CapsuleVsConvexPolyhedronAlgorithm.cpp:42-177
__device__ __forceinline__ bool capsule_poly_test(const Conv& S1, const Conv& S2, const Tf& T1, const Tf& T2, uint32_t& lf,
                                                  float4& axis, Raw& raw, bool report) {
    const int g = gjk(S1, S2, T1, T2, lf, axis, raw);
    lf |= LF_WAS_GJK;
    lf &= ~LF_WAS_SAT;
    bool hit = false;
    const bool cap1 = S1.type == SHAPE_CAPSULE;
    const Conv& Sc = cap1 ? S1 : S2;
    const Conv& Sp = cap1 ? S2 : S1;
    const Tf& Tc = cap1 ? T1 : T2;
    const Tf& Tp = cap1 ? T2 : T1;
    if (g == GJK_COLLIDE_IN_MARGIN && !report) {
        hit = true;   // the face-contact search only adds points (CapsuleVsConvexPolyhedronAlgorithm.cpp:83)
    } else if (g == GJK_COLLIDE_IN_MARGIN) {
        const V3 cn = raw.nrm[0];
        const float cpen = raw.depth[0];
        const float height = Sc.hh + Sc.hh;
        const int nf = Sp.P.nfaces();
        for (int f = 0; f < nf; f++) {
            const V3 fn = Sp.P.face_normal(f);
            V3 fnw = qrot(Tp.q, fn);
            const V3 sa = v3(0.0f, -height * 0.5f, 0.0f);
            const V3 sb = v3(0.0f, height * 0.5f, 0.0f);
            const V3 sdir = normalized(qrot(Tc.q, sb - sa));
            const bool inDir = dot(fnw, cn) > 0.0f;
            const bool inContactDir = (cap1 && !inDir) || (!cap1 && inDir);
            if (inContactDir && fabsf(dot(fnw, sdir)) < 0.001f && parallel_vectors(fnw, cn)) {
                const Tf p2c = tmul(tinv(Tc), Tp);
                const Tf c2p = tinv(p2c);
                const V3 saP = tapply(c2p, sa);
                const V3 sbP = tapply(c2p, sb);
                const V3 sepCap = qrot(p2c.q, fn);
                if (cap1) fnw = -fnw;
                if (capsule_face_contacts(f, Sc.margin, Sp.P, cpen, p2c, fnw, sepCap, saP, sbP, cap1, raw)) break;
            }
        }
        hit = true;
    } else if (g == GJK_INTERPENETRATE) {
        hit = sat_capsule_poly(cap1, Sc.margin, Sc.hh, Sp.P, Tc, Tp, raw, report);
        lf &= ~LF_WAS_GJK;
        lf |= LF_WAS_SAT;
    }
    return hit;
}

__global__ void np_capsule_poly(NpArgs A, const int* list, const int* count) {
    RP3DG_GRID_STRIDE(t, *count) {
        const int k = list[t];
        const int2 c = A.P.col[k];
        const Tf T1 = tf(ld3(A.C.wpos, c.x), ldq(A.C.wq, c.x));
        const Tf T2 = tf(ld3(A.C.wpos, c.y), ldq(A.C.wq, c.y));
        const Conv S1 = make_conv(A, c.x), S2 = make_conv(A, c.y);
        uint32_t lf = A.P.lf[k];
        float4 axis = A.P.lfaxis[k];
        Raw raw; raw.n = 0;
        const bool hit = capsule_poly_test(S1, S2, T1, T2, lf, axis, raw, pair_reports(A, k));
        A.P.lfaxis[k] = axis;
        A.P.lf[k] = A.lf_update ? lf_after_test(lf, hit) : lf;
        contact_epilogue(A, k, hit, raw, c.x);
    }
}

}  // namespace rp3dg

// ------------------------------------------------------------------------------------------------
// Convex vs concave pairs (CollisionDetectionSystem::computeConvexVsConcaveMiddlePhase, CD.cpp:522-615;
// processPotentialContacts for non-convex pairs, :1292-1385; reducePotentialContactManifolds, :1390-1458)
// ------------------------------------------------------------------------------------------------
namespace rp3dg {

constexpr int kConcaveStack = 64;          // tree walk stack (a balanced tree of 2^31 triangles needs 45)
constexpr int kConcavePairPoints = 64;     // contact points of one pair's tests that its reduction holds
constexpr int kMaxPotentialManifolds = 64;   // NB_MAX_POTENTIAL_CONTACT_MANIFOLDS (configuration.h:140)
constexpr int kMaxPairManifolds = 4;         // NB_MAX_CONTACT_MANIFOLDS (configuration.h:137)
constexpr int kMaxPotentialManifoldPoints = 255;   // NB_MAX_CONTACT_POINTS_IN_POTENTIAL_MANIFOLD (:143)

// convex collider cv and concave collider cc of a pair; convex1: the convex one is collider 1
__device__ __forceinline__ void concave_sides(const NpArgs& A, int2 c, int& cv, int& cc, bool& convex1) {
    convex1 = !shape_is_concave(A.C.type[c.x]);
    cv = convex1 ? c.x : c.y;
    cc = convex1 ? c.y : c.x;
}

// HeightField::getVertexAt (HeightField.cpp:399-409)
__device__ __forceinline__ V3 hf_vertex(const NpArgs& A, const ConcaveInst& I, int x, int y) {
    const float h = A.cheights[I.hbase + y * I.columns + x];
    return v3(-I.width * 0.5f + (float)(uint32_t)x, I.origin + h, -I.length * 0.5f + (float)(uint32_t)y);
}

// HeightFieldShape::computeOverlappingTriangles -> HeightField::computeMinMaxGridCoordinates + clamps
// (HeightFieldShape.cpp:46-52, HeightField.cpp:133-216): the cells [i0, i1) x [j0, j1)
__device__ __forceinline__ void hf_cells(const ConcaveInst& I, V3 mn, V3 mx, int& i0, int& i1, int& j0, int& j1) {
    const V3 is = v3(I.isx, I.isy, I.isz);
    mn = cmul(mn, is);
    mx = cmul(mx, is);
    const V3 bmin = v3(I.bminx, I.bminy, I.bminz), bmax = v3(I.bmaxx, I.bmaxy, I.bmaxz);
    V3 lo = vmin(vmax(mn, bmin), bmax);
    V3 hi = vmax(vmin(mx, bmax), bmin);
    const V3 tr = (bmax - bmin) * 0.5f;
    lo = lo + tr;
    hi = hi + tr;
    const int a0 = (int)((double)lo.x + 0.5) - 1, a2 = (int)((double)lo.z + 0.5) - 1;
    const int b0 = (int)((double)hi.x + 0.5) + 1, b2 = (int)((double)hi.z + 0.5) + 1;
    i0 = min(max(a0, 0), I.columns - 1);
    i1 = min(max(b0, 0), I.columns - 1);
    j0 = min(max(a2, 0), I.rows - 1);
    j1 = min(max(b2, 0), I.rows - 1);
}

// DynamicAABBTree::reportAllShapesOverlappingWithAABB (DynamicAABBTree.cpp:641-681) on a mesh's tree:
// emit(index, triangle) for the overlapping leaves in the reference's order; false on a stack overflow
template <class Emit>
__device__ __forceinline__ bool mesh_walk(const NpArgs& A, const ConcaveInst& I, V3 mn, V3 mx, Emit emit) {
    int stack[kConcaveStack];
    int sp = 0, n = 0;
    stack[sp++] = I.root;
    while (sp > 0) {
        const int id = stack[--sp];
        if (id < 0) continue;
        const float4 a = A.cnodes[2 * (I.node_base + id)];
        const float4 b = A.cnodes[2 * (I.node_base + id) + 1];
        if (mx.x < a.x || b.x < mn.x || mx.y < a.y || b.y < mn.y || mx.z < a.z || b.z < mn.z) continue;
        const int c0 = __float_as_int(a.w), c1 = __float_as_int(b.w);
        if (c0 < 0) {
            emit(n++, c1);
        } else {
            if (sp + 2 > kConcaveStack) return false;
            stack[sp++] = c0;
            stack[sp++] = c1;
        }
    }
    return true;
}

// Middle phase: the convex shape's AABB in the concave shape's space and the triangles it overlaps
// (ConcaveShape::computeOverlappingTriangles), listed as this step's tests of the pair, contiguous and in
// order.
__global__ void np_concave_mid(NpArgs A, const int* list, const int* count) {
    RP3DG_GRID_STRIDE(t, *count) {
        const int k = list[t];
        const int2 c = A.P.col[k];
        int cv, cc;
        bool convex1;
        concave_sides(A, c, cv, cc, convex1);
        const Tf Tv = tf(ld3(A.C.wpos, cv), ldq(A.C.wq, cv));
        const Tf Tc = tf(ld3(A.C.wpos, cc), ldq(A.C.wq, cc));
        V3 mn, mx;
        transformed_aabb(A.C.type[cv], A.C.shape[cv], A.hull_bounds, tmul(tinv(Tc), Tv), mn, mx);
        const ConcaveInst I = A.cinst[(int)A.C.shape[cc].y];
        int n = 0, base = 0;
        bool ok = true, reserved = false;
        if (I.kind == SHAPE_CONCAVE_MESH) {
            const V3 is = v3(I.isx, I.isy, I.isz);   // AABB::applyScale(1 / scale)
            mn = cmul(mn, is);
            mx = cmul(mx, is);
            ok = mesh_walk(A, I, mn, mx, [&](int i, int) { n = i + 1; });
            if (ok) {
                base = n ? atomicAdd(A.CT.count, n) : 0;
                reserved = n > 0;
                if (base + n > A.CT.cap) ok = false;
                else mesh_walk(A, I, mn, mx, [&](int i, int tri) {
                    A.CT.pair[base + i] = k;
                    A.CT.tri[base + i] = tri;
                });
            }
        } else {
            int i0, i1, j0, j1;
            hf_cells(I, mn, mx, i0, i1, j0, j1);
            n = (i1 > i0 && j1 > j0) ? 2 * (i1 - i0) * (j1 - j0) : 0;
            base = n ? atomicAdd(A.CT.count, n) : 0;
            reserved = n > 0;
            if (base + n > A.CT.cap) ok = false;
            else {
                int q = base;
                for (int i = i0; i < i1; i++) {
                    for (int j = j0; j < j1; j++) {
                        const int id = (j * (I.columns - 1) + i) * 2;   // HeightField::computeTriangleShapeId
                        A.CT.pair[q] = k; A.CT.tri[q] = id; q++;
                        A.CT.pair[q] = k; A.CT.tri[q] = id + 1; q++;
                    }
                }
            }
        }
        if (!ok) {
            atomicOr(A.status, STATUS_CONCAVE_OVERFLOW);
            if (reserved)   // slots reserved but not listed
                for (int q = base; q < A.CT.cap && q < base + n; q++) A.CT.pair[q] = -1;
            n = 0;
        }
        A.P.ct0[k] = base;
        A.P.ctn[k] = n;
    }
}

// One triangle test: the TriangleShape of the middle phase (TriangleShape.cpp:40-62) against the convex
// shape with the pair's algorithm, the triangle's LastFrameCollisionInfo (kept by the pair for the
// triangles it tested in the last step, OverlappingPairs.h:231-294) and smoothed contacts.
__global__ __launch_bounds__(128) void np_concave_tri(NpArgs A) {
    const int n = min(*A.CT.count, A.CT.cap);
    RP3DG_GRID_STRIDE(t, n) {
        const int k = A.CT.pair[t];
        if (k < 0) continue;
        const int id = A.CT.tri[t];
        const int2 c = A.P.col[k];
        int cv, cc;
        bool convex1;
        concave_sides(A, c, cv, cc, convex1);
        const ConcaveInst I = A.cinst[(int)A.C.shape[cc].y];
        const V3 sc = v3(I.sx, I.sy, I.sz);
        TriSmooth sm;
        if (I.kind == SHAPE_CONCAVE_MESH) {
            const int4 ix = A.ctris[I.tbase + id];
            sm.p0 = cmul(xyz(A.cverts[I.vbase + ix.x]), sc);
            sm.p1 = cmul(xyz(A.cverts[I.vbase + ix.y]), sc);
            sm.p2 = cmul(xyz(A.cverts[I.vbase + ix.z]), sc);
            sm.vn0 = xyz(A.cnorm[I.nbase + ix.x]);
            sm.vn1 = xyz(A.cnorm[I.nbase + ix.y]);
            sm.vn2 = xyz(A.cnorm[I.nbase + ix.z]);
        } else {   // HeightField::computeOverlappingTriangles (:150-176)
            const int cell = id >> 1;
            const int i = cell % (I.columns - 1), j = cell / (I.columns - 1);
            const V3 q1 = cmul(hf_vertex(A, I, i, j), sc), q2 = cmul(hf_vertex(A, I, i, j + 1), sc);
            const V3 q3 = cmul(hf_vertex(A, I, i + 1, j), sc), q4 = cmul(hf_vertex(A, I, i + 1, j + 1), sc);
            V3 fnrm;
            if ((id & 1) == 0) {
                sm.p0 = q1; sm.p1 = q2; sm.p2 = q3;
                fnrm = unit(cross(q2 - q1, q3 - q1));
            } else {
                sm.p0 = q3; sm.p1 = q2; sm.p2 = q4;
                fnrm = unit(cross(q2 - q3, q4 - q3));
            }
            sm.vn0 = sm.vn1 = sm.vn2 = fnrm;
        }
        sm.n = normalized(cross(sm.p1 - sm.p0, sm.p2 - sm.p0));
        const Tf Tv = tf(ld3(A.C.wpos, cv), ldq(A.C.wq, cv));
        const Tf Tc = tf(ld3(A.C.wpos, cc), ldq(A.C.wq, cc));
        sm.tri1 = !convex1;
        sm.triToWorld = Tc;
        sm.worldToOther = tinv(Tv);
        Conv Sv = make_conv(A, cv);
        // the TriangleShape as a hull in the test's slots of the triangle arena: its three vertices and
        // two faces (the normal and its opposite) over the half-edge structure every triangle shares
        const TriArena& TA = A.tri;
        HullHeader h;
        h.vbase = 3 * t; h.nv = 3;
        h.fbase = 2 * t; h.nf = 2;
        h.ebase = 0; h.ne = 6;
        h.fvbase = 0; h.nfv = 6;
        const V3 cen = (sm.p0 + sm.p1 + sm.p2) / 3.0f;   // TriangleShape.h getCentroid
        h.cx = cen.x; h.cy = cen.y; h.cz = cen.z;
        h.pad = 1.0f;
        TA.verts[h.vbase] = f4(sm.p0, 0.f);
        TA.verts[h.vbase + 1] = f4(sm.p1, 0.f);
        TA.verts[h.vbase + 2] = f4(sm.p2, 0.f);
        TA.normals[h.fbase] = f4(sm.n, 0.f);
        TA.normals[h.fbase + 1] = f4(-sm.n, 0.f);
        TA.face_edge[h.fbase] = TA.face_edge0[0];
        TA.face_edge[h.fbase + 1] = TA.face_edge0[1];
        TA.face_verts[2 * h.fbase] = 0;
        TA.face_verts[2 * h.fbase + 1] = 3;
        TA.face_verts[2 * h.fbase + 2] = 3;
        TA.face_verts[2 * h.fbase + 3] = 3;
        Conv St;
        St.type = SHAPE_HULL;
        St.margin = 0.0f;
        St.hh = 0.0f;
        St.P.hull = t;
        St.P.he = v3zero();
        St.P.h = h;
        St.P.hv = TA.verts; St.P.hn = TA.normals; St.P.hedges = TA.edges;
        St.P.hfedge = TA.face_edge; St.P.hfverts = TA.face_verts; St.P.hfvidx = TA.vidx;
        const Conv& S1 = convex1 ? Sv : St;
        const Conv& S2 = convex1 ? St : Sv;
        const Tf& T1 = convex1 ? Tv : Tc;
        const Tf& T2 = convex1 ? Tc : Tv;
        // the triangle's last-frame info: from the pair's tests of the last step, else the default
        uint32_t lf = 0u;
        float4 axis = make_float4(0.f, 1.f, 0.f, 0.f);
        const int prev = A.P.prev[k];
        if (prev >= 0) {
            const int b = A.O.ct0[prev], m = A.O.ctn[prev];
            for (int q = 0; q < m; q++) {
                if (A.CTo.tri[b + q] == id) {
                    lf = A.CTo.lf[b + q];
                    axis = A.CTo.lfaxis[b + q];
                    break;
                }
            }
        }
        Raw raw;
        raw.n = 0;
        raw.smooth = &sm;
        const int vt = Sv.type;
        bool hit = false;
        if (vt == SHAPE_SPHERE) {
            hit = sphere_poly_test(S1, S2, T1, T2, lf, axis, raw);
        } else if (vt == SHAPE_CAPSULE) {
            hit = capsule_poly_test(S1, S2, T1, T2, lf, axis, raw, pair_reports(A, k));
        } else {   // ConvexPolyhedronVsConvexPolyhedronAlgorithm: SAT, one lane per test
            const Tf t12 = tmul(tinv(T2), T1), t21 = tinv(t12);
            float minPen0 = 0.0f;
            int minFace = 0;
            bool isMinFace1 = false;
            const int stage = sat_begin<false, true>(S1.P, S2.P, T1, T2, t12, t21, lf, raw, minPen0, minFace, isMinFace1,
                                                     A.clip_prev != 0);
            hit = stage == 1;
            if (stage == 2) {
                EdgeSearch es;
                es.init();
                int h2;
                const int np = sat_edge_pairs(S1.P, S2.P, h2);
                for (int kk = 0; kk < np; kk++) {
                    float p = 0.0f;
                    const bool v = sat_edge_pair(S1.P, S2.P, t12, kk, h2, p);
                    es.visit(kk, v, p);
                    if (es.sep != INT_MAX) break;
                }
                hit = sat_finish<false>(S1.P, S2.P, T1, T2, t12, t21, lf, raw, minPen0, minFace, isMinFace1, es);
            }
            lf |= LF_WAS_SAT;
            lf &= ~LF_WAS_GJK;
        }
        if (raw.clip_full) atomicOr(A.status, STATUS_CLIP_OVERFLOW);
        A.CT.lf[t] = A.lf_update ? lf_after_test(lf, hit) : lf;
        A.CT.lfaxis[t] = axis;
        A.CT.hit[t] = hit ? 1 : 0;
        int2 pts = make_int2(0, 0);
        if (hit && raw.n > 0 && pair_reports(A, k)) {
            const int base = atomicAdd(A.cp_count, raw.n);
            if (base + raw.n > A.cp_cap) {
                atomicOr(A.status, STATUS_CONCAVE_OVERFLOW);
            } else {
                for (int q = 0; q < raw.n; q++) {
                    A.cp_nd[base + q] = f4(raw.nrm[q], raw.depth[q]);
                    A.cp_l1[base + q] = f4(raw.l1[q], 0.f);
                    A.cp_l2[base + q] = f4(raw.l2[q], 0.f);
                }
                pts = make_int2(base, raw.n);
            }
        }
        A.CT.pts[t] = pts;
    }
}

// The points of one manifold of a concave pair, as reduce_points reads them
struct PoolPoints {
    const float4 *nd, *l1, *l2;
    const int* idx;   // pool index of each point
    int n;
    __device__ V3 normal(int r) const { return xyz(nd[idx[r]]); }
    __device__ V3 p1(int r) const { return xyz(l1[idx[r]]); }
};
template <int N>
struct IntList {
    int v[N];
    __device__ int get(int i) const { return v[i]; }
    __device__ void set(int i, int x) { v[i] = x; }
};

// Per pair: the potential manifolds of its triangle tests' points (contact points go to the first
// manifold of a similar normal, CD.cpp:1325-1380), at most NB_MAX_CONTACT_MANIFOLDS kept (the one of the
// smallest largest depth removed first, swapped with the last), each reduced to four points and without
// duplicates, then the warm-start transfer over the pair's previous manifolds (initContactsWithPreviousOnes,
// :1069-1160) and the cache write: manifold 0 in the pair's entry, the others in tail entries.
__global__ void np_concave_reduce(NpArgs A, const int* list, const int* count) {
    RP3DG_GRID_STRIDE(t, *count) {
        const int k = list[t];
        DevPairs P = A.P;
        const DevPairs& O = A.O;
        const int b = P.ct0[k], nt = P.ctn[k];
        bool colliding = false;
        int pidx[kConcavePairPoints];
        int np = 0;
        for (int q = 0; q < nt; q++) {
            colliding |= A.CT.hit[b + q] != 0;
            const int2 r = A.CT.pts[b + q];
            for (int p = 0; p < r.y; p++) {
                if (np < kConcavePairPoints) pidx[np++] = r.x + p;
                else atomicOr(A.status, STATUS_CONCAVE_OVERFLOW);
            }
        }
        // the pair's wasColliding (trigger pairs report their overlap from it)
        uint32_t plf = P.lf[k] | LF_VALID;
        plf = colliding ? (plf | LF_WAS_COLLIDING) : (plf & ~LF_WAS_COLLIDING);
        if (A.lf_update) P.lf[k] = plf;
        P.xm[k] = 0;
        P.xs[k] = 0;
        if (!colliding || np == 0 || !pair_reports(A, k)) {
            P.ncp[k] = 0;
            continue;
        }
        // potential manifolds in creation order
        int pm[kConcavePairPoints];   // manifold of each point (-1: dropped)
        int mfirst[kMaxPotentialManifolds], mcnt[kMaxPotentialManifolds];
        int nm = 0;
        for (int q = 0; q < np; q++) {
            const V3 nq = xyz(A.cp_nd[pidx[q]]);
            int found = -1;
            for (int m = 0; m < nm; m++) {
                if (mcnt[m] < kMaxPotentialManifoldPoints &&
                    dot(xyz(A.cp_nd[pidx[mfirst[m]]]), nq) >= A.cos_similar) {
                    found = m;
                    break;
                }
            }
            if (found < 0 && nm < kMaxPotentialManifolds) {
                found = nm++;
                mfirst[found] = q;
                mcnt[found] = 0;
            }
            pm[q] = found;
            if (found >= 0) mcnt[found]++;
        }
        // at most four manifolds (ContactPair::removePotentialManifoldAtIndex swaps with the last)
        int ml[kMaxPotentialManifolds];
        for (int m = 0; m < nm; m++) ml[m] = m;
        int L = nm;
        while (L > kMaxPairManifolds) {
            float minDepth = kDecimalLargest;
            int mi = -1;
            for (int j = 0; j < L; j++) {
                float largest = 0.0f;
                for (int q = 0; q < np; q++)
                    if (pm[q] == ml[j] && A.cp_nd[pidx[q]].w > largest) largest = A.cp_nd[pidx[q]].w;
                if (largest < minDepth) {
                    minDepth = largest;
                    mi = j;
                }
            }
            ml[mi] = ml[L - 1];
            L--;
        }
        // previous manifolds of the pair: its entry, then its tail entries
        int prevSlots[kMaxPairManifolds];
        int nprev = 0;
        const int prev = P.prev[k];
        if (prev >= 0 && O.ncp[prev] > 0) {
            prevSlots[nprev++] = prev;
            const int pxm = O.xm[prev], pxs = O.xs[prev];
            for (int j = 0; j < pxm && nprev < kMaxPairManifolds; j++) prevSlots[nprev++] = pxs + j;
        }
        const int tbase = L > 1 ? atomicAdd(A.tail_count, L - 1) : 0;
        if (L > 1 && tbase + L - 1 > A.tail_cap) {
            atomicOr(A.status, STATUS_CONCAVE_OVERFLOW);
            L = 1;
        }
        const Tf T1 = tf(ld3(A.C.wpos, P.col[k].x), ldq(A.C.wq, P.col[k].x));
        const float pt2 = A.persistent_threshold * A.persistent_threshold;
        const float thr2 = 0.01f * 0.01f;
        for (int j = 0; j < L; j++) {
            const int slot = j == 0 ? k : A.tail_base + tbase + j - 1;
            int mp[kConcavePairPoints];
            int n = 0;
            for (int q = 0; q < np; q++)
                if (pm[q] == ml[j]) mp[n++] = pidx[q];
            IntList<kConcavePairPoints> idx;
            for (int i = 0; i < n; i++) idx.set(i, mp[i]);
            if (n > kMaxManifoldPoints) {   // reduceContactPoints
                PoolPoints view{A.cp_nd, A.cp_l1, A.cp_l2, mp, n};
                int keep[4];
                reduce_points<PoolPoints, IntList<kConcavePairPoints>>(view, T1, keep);
                for (int i = 0; i < 4; i++) idx.set(i, mp[keep[i]]);
                n = 4;
            }
            // removeDuplicatedContactPointsInManifold (:1673-1705)
            for (int i = 0; i < n; i++) {
                for (int m = i + 1; m < n; m++) {
                    if (len2(xyz(A.cp_l1[idx.get(m)]) - xyz(A.cp_l1[idx.get(i)])) < thr2) {
                        idx.set(m, idx.get(n - 1));
                        n--;
                        m--;
                    }
                }
            }
            // friction of the first previous manifold of a similar normal
            float4 fv1 = make_float4(0.f, 0.f, 0.f, 0.f), fv2 = fv1, fimp = fv1;
            const V3 n0 = xyz(A.cp_nd[idx.get(0)]);
            for (int p = 0; p < nprev; p++) {
                const float4 on = O.pn[4 * prevSlots[p]];
                if (dot(v3(on.x, on.y, on.z), n0) >= A.cos_similar) {
                    fv1 = O.fv1[prevSlots[p]];
                    fv2 = O.fv2[prevSlots[p]];
                    fimp = O.fimp[prevSlots[p]];
                    break;
                }
            }
            for (int i = 0; i < n; i++) {
                const int r = idx.get(i);
                const V3 l1 = xyz(A.cp_l1[r]);
                float imp = 0.0f, resting = 0.0f;
                bool matched = false;
                for (int p = 0; p < nprev && !matched; p++) {
                    const int os = prevSlots[p];
                    const int on = O.ncp[os];
                    for (int o = 0; o < on; o++) {
                        const float4 ol1 = O.pl1[4 * os + o];
                        if (len2(l1 - xyz(ol1)) <= pt2) {
                            imp = ol1.w;
                            resting = O.pl2[4 * os + o].w;
                            matched = true;
                            break;
                        }
                    }
                }
                P.pn[4 * slot + i] = A.cp_nd[r];
                P.pl1[4 * slot + i] = f4(l1, imp);
                P.pl2[4 * slot + i] = f4(xyz(A.cp_l2[r]), resting);
            }
            P.fv1[slot] = fv1;
            P.fv2[slot] = fv2;
            P.fimp[slot] = fimp;
            P.ncp[slot] = n;
            if (j > 0) P.col[slot] = P.col[k];
        }
        P.xm[k] = L - 1;
        P.xs[k] = A.tail_base + tbase;
    }
}

}  // namespace rp3dg
