// Convex hull of a point cloud, for PhysicsCommon::createConvexMesh(const VertexArray&, messages)
// (src/engine/PhysicsCommon.cpp:638-660, which calls QuickHull::computeConvexHull,
// src/utils/quickhull/QuickHull.cpp:44-131).
//
// The result has to be the reference's hull polygon for polygon, not just the same convex set: the order
// of the faces, the first vertex of each face and the vertex numbering become the half-edge numbering of
// the convex mesh (HalfEdgeStructure.cpp:35-121), and the SAT temporal-coherence record stores face and
// edge indices of that numbering.  So the incremental construction below takes the reference's steps in
// the reference's order: the same initial tetrahedron, the farthest conflict point first, a depth-first
// horizon walk, new faces prepended to the face list, the two face-merging passes, the repair of
// vertices left with two faces, and the same FP32 arithmetic (Newell normals, centroids, plane
// distances against the 3 * sum|max coordinate| * FLT_EPSILON tolerance).  The data structure is this
// file's own: faces and half-edges live in index arrays, and the (start, end) -> half-edge table keeps the
// reference's keying (an edge is filed under the vertices it had when it was created), since the twin
// lookups of later faces depend on it.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "quickhull.h"

namespace rp3dg {
namespace {

constexpr float kFltEps = 1.1920928955078125e-07f;   // std::numeric_limits<float>::epsilon()

struct QEdge {
    int v0, v1;     // start / end point index
    int face;
    int twin, next, prev;
};

struct QFace {
    int edge;       // first half-edge of the face's cycle
    V3 normal, centroid;
    float area;
    std::vector<uint32_t> conflict;   // points outside this face, in the order they were assigned
    int prev, next;                   // face list (most recent first)
    bool alive;
};

class Builder {
public:
    explicit Builder(std::vector<V3> points) : P(std::move(points)) {}

    bool run(std::string* err);
    void output(std::vector<float>& verts, std::vector<int32_t>& counts, std::vector<int32_t>& idx) const;

private:
    std::vector<V3> P;
    float eps = 0.0f;
    std::vector<QEdge> E;
    std::vector<QFace> F;
    int head = -1;
    std::unordered_map<uint64_t, int> edge_of;   // (start, end) at creation -> half-edge
    bool broken = false;                         // a twin lookup failed (degenerate input)

    static uint64_t key(int a, int b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; }
    float plane_dist(V3 p, int f) const { return dot(F[f].normal, p - F[f].centroid); }
    bool triangle(int f) const { const int e = F[f].edge; return E[E[E[e].next].next].next == e; }

    int add_face(const int* v, int n);
    void remove_edge(int e);
    void delete_face(int f);
    void remove_face(int f);
    void recalc(int f);
    bool initial_hull(std::vector<int>& faces, std::vector<uint32_t>& orphans, std::string* err);
    void assign(const std::vector<uint32_t>& orphans, const std::vector<int>& faces, const std::vector<char>& deleted);
    bool next_candidate(uint32_t& vertex, int& face);
    void add_vertex(uint32_t vertex, int face);
    void horizon(V3 p, int face, std::vector<int>& hv, std::vector<int>& visible);
    void merge_large(std::vector<int>& nf, std::vector<char>& deleted);
    void merge_concave(std::vector<int>& nf, std::vector<char>& deleted);
    bool convex_edge(int e) const;
    void merge_at_edge(int e, std::vector<char>& deleted);
    void fix_topology(int f, std::vector<char>& deleted);
    void fix_at_edge(int f, int in, std::vector<char>& deleted);
    static void mark(std::vector<char>& deleted, int f) {
        if ((int)deleted.size() <= f) deleted.resize(f + 1, 0);
        deleted[f] = 1;
    }
    static bool is_deleted(const std::vector<char>& deleted, int f) { return f < (int)deleted.size() && deleted[f]; }
};

// QHHalfEdgeStructure::Face::recalculateFace (QHHalfEdgeStructure.h:148-179): Newell normal.
void Builder::recalc(int f) {
    V3 c = v3zero(), n = v3zero();
    uint32_t cnt = 0;
    const int first = F[f].edge;
    int e = first;
    do {
        const V3 a = P[E[e].v0], b = P[E[e].v1];
        c += a;
        n += v3((a.y - b.y) * (a.z + b.z), (a.z - b.z) * (a.x + b.x), (a.x - b.x) * (a.y + b.y));
        cnt++;
        e = E[e].next;
    } while (e != first);
    F[f].centroid = c / (float)cnt;
    const float l = len(n);
    F[f].normal = n / l;
    F[f].area = l * 0.5f;
}

// QHHalfEdgeStructure::addFace (QHHalfEdgeStructure.cpp): vertices in CCW order seen from outside.
int Builder::add_face(const int* v, int n) {
    const int f = (int)F.size();
    F.push_back(QFace{-1, v3zero(), v3zero(), 0.0f, {}, -1, -1, true});
    int prev = -1, first = -1;
    for (int i = 0; i < n; i++) {
        const int a = v[i], b = v[i == n - 1 ? 0 : i + 1];
        const int e = (int)E.size();
        E.push_back(QEdge{a, b, f, -1, -1, prev});
        auto it = edge_of.find(key(b, a));
        if (it != edge_of.end()) {
            E[it->second].twin = e;
            E[e].twin = it->second;
        }
        if (prev >= 0) E[prev].next = e;
        if (i == 0) first = e;
        edge_of[key(a, b)] = e;
        prev = e;
    }
    E[prev].next = first;
    E[first].prev = prev;
    F[f].edge = first;
    if (head >= 0) F[head].prev = f;
    F[f].next = head;
    head = f;
    recalc(f);
    return f;
}

// QHHalfEdgeStructure::removeHalfEdge: drops the table entry under the edge's current vertices.
void Builder::remove_edge(int e) {
    edge_of.erase(key(E[e].v0, E[e].v1));
    if (E[e].twin >= 0) E[E[e].twin].twin = -1;
}

void Builder::delete_face(int f) {
    if (F[f].next >= 0) F[F[f].next].prev = F[f].prev;
    if (F[f].prev >= 0) F[F[f].prev].next = F[f].next;
    else head = F[f].next;
    F[f].alive = false;
    F[f].conflict.clear();
}

void Builder::remove_face(int f) {
    const int first = F[f].edge;
    int e = first;
    do {
        const int nx = E[e].next;
        remove_edge(e);
        e = nx;
    } while (e != first);
    delete_face(f);
}

// QuickHull::computeInitialHull (QuickHull.cpp:873-1057)
bool Builder::initial_hull(std::vector<int>& faces, std::vector<uint32_t>& orphans, std::string* err) {
    uint32_t ext[6] = {0, 0, 0, 0, 0, 0};
    const uint32_t n = (uint32_t)P.size();
    for (uint32_t i = 0; i < n; i++) {
        if (P[i].x < P[ext[0]].x) ext[0] = i;
        else if (P[i].x > P[ext[1]].x) ext[1] = i;
        if (P[i].y < P[ext[2]].y) ext[2] = i;
        else if (P[i].y > P[ext[3]].y) ext[3] = i;
        if (P[i].z < P[ext[4]].z) ext[4] = i;
        else if (P[i].z > P[ext[5]].z) ext[5] = i;
    }
    const char* close = "Error during initial hull creation in QuickHull: vertices too close to each other";
    float best = 0.0f;
    uint32_t axis = 0;
    for (uint32_t i = 0; i < 3; i++) {
        const float d = len2(P[ext[2 * i]] - P[ext[2 * i + 1]]);
        if (d > best) { axis = i; best = d; }
    }
    if (best < kFltEps) { *err = close; return false; }
    const uint32_t i1 = ext[2 * axis], i2 = ext[2 * axis + 1];
    const V3 dir = normalized(P[i2] - P[i1]);
    uint32_t i3 = 0;
    best = 0.0f;
    for (uint32_t i = 0; i < n; i++) {
        const float d = len2(cross(P[i] - P[i1], dir));
        if (d > best) { i3 = i; best = d; }
    }
    if (best < kFltEps) { *err = close; return false; }
    uint32_t i4 = 0;
    best = 0.0f;
    const V3 pn = cross(P[i3] - P[i1], P[i2] - P[i1]);
    for (uint32_t i = 0; i < n; i++) {
        const float d = std::fabs(dot(P[i] - P[i1], pn));
        if (d > best) { i4 = i; best = d; }
    }
    if (best < kFltEps) { *err = close; return false; }
    const int a = (int)i1, b = (int)i2, c = (int)i3, d = (int)i4;
    if (dot(pn, P[i4] - P[i1]) < 0.0f) {
        const int t[4][3] = {{a, c, b}, {b, c, d}, {a, b, d}, {a, d, c}};
        for (const auto& tri : t) faces.push_back(add_face(tri, 3));
    } else {
        const int t[4][3] = {{a, b, c}, {b, d, c}, {a, d, b}, {a, c, d}};
        for (const auto& tri : t) faces.push_back(add_face(tri, 3));
    }
    // Array::remove(value): first occurrence, order kept (vertices added as i1, i2, i3, i4)
    for (uint32_t v : {i1, i2, i3, i4}) {
        for (size_t k = 0; k < orphans.size(); k++)
            if (orphans[k] == v) { orphans.erase(orphans.begin() + k); break; }
    }
    return true;
}

// QuickHull::associateOrphanPointsToNewFaces / findFarthestFaceForVertex (QuickHull.cpp:833-870)
void Builder::assign(const std::vector<uint32_t>& orphans, const std::vector<int>& faces, const std::vector<char>& deleted) {
    for (uint32_t p : orphans) {
        float best = eps;
        int far = -1;
        for (int f : faces) {
            if (is_deleted(deleted, f)) continue;
            const float d = dot(F[f].normal, P[p] - F[f].centroid);
            if (d > best) { best = d; far = f; }
        }
        if (far >= 0) F[far].conflict.push_back(p);
    }
}

// QuickHull::findNextVertexCandidate (QuickHull.cpp:791-830)
bool Builder::next_candidate(uint32_t& vertex, int& face) {
    float best = eps;
    int bi = -1;
    face = -1;
    for (int f = head; f >= 0; f = F[f].next) {
        const std::vector<uint32_t>& cl = F[f].conflict;
        for (size_t i = 0; i < cl.size(); i++) {
            const float d = dot(P[cl[i]] - F[f].centroid, F[f].normal);
            if (d > best) { best = d; vertex = cl[i]; bi = (int)i; face = f; }
        }
    }
    if (face < 0) return false;
    F[face].conflict.erase(F[face].conflict.begin() + bi);
    return true;
}

// QuickHull::findHorizon (QuickHull.cpp:352-413): depth-first walk over the visible faces; every edge to a
// face not visible from p is a horizon edge (start, end pairs in hv).
void Builder::horizon(V3 p, int face, std::vector<int>& hv, std::vector<int>& visible) {
    struct Cand { int face, start, cur; };
    std::vector<Cand> stack;
    std::vector<char> visited(F.size(), 0);
    stack.push_back(Cand{face, F[face].edge, F[face].edge});
    visible.push_back(face);
    while (!stack.empty()) {
        const size_t t = stack.size() - 1;
        visited[stack[t].face] = 1;
        bool descend = false;
        do {
            const int e = stack[t].cur;
            const int tw = E[e].twin;
            if (tw < 0) { broken = true; return; }
            const int nf = E[tw].face;
            if (!visited[nf]) {
                if (dot(F[nf].normal, p - F[nf].centroid) > eps) {
                    stack.push_back(Cand{nf, E[tw].next, E[tw].next});
                    visible.push_back(nf);
                    descend = true;
                    break;
                }
                hv.push_back(E[e].v0);
                hv.push_back(E[e].v1);
            }
            stack[t].cur = E[e].next;
        } while (stack[t].cur != stack[t].start);
        if (!descend) stack.pop_back();
    }
}

// QuickHull::testIsConvexEdge (QuickHull.cpp:334-349)
bool Builder::convex_edge(int e) const {
    const int f1 = E[e].face, f2 = E[E[e].twin].face;
    if (plane_dist(F[f1].centroid, f2) >= -eps) return false;
    if (plane_dist(F[f2].centroid, f1) >= -eps) return false;
    return true;
}

// QuickHull::mergeLargeConcaveFaces (QuickHull.cpp:417-482): the larger face of each pair decides.
void Builder::merge_large(std::vector<int>& nf, std::vector<char>& deleted) {
    size_t i = 0;
    while (i < nf.size()) {
        const int f1 = nf[i];
        if (!is_deleted(deleted, f1)) {
            int concave = -1;
            const int first = F[f1].edge;
            int e = first;
            do {
                if (E[e].twin < 0) { broken = true; return; }
                const int f2 = E[E[e].twin].face;
                if (F[f1].area > F[f2].area) {
                    if (plane_dist(F[f2].centroid, f1) < -eps) { e = E[e].next; continue; }
                    concave = e;
                    break;
                }
                if (plane_dist(F[f1].centroid, f2) < -eps) { e = E[e].next; continue; }
                concave = e;
                break;
            } while (e != first);
            if (concave >= 0) {
                merge_at_edge(concave, deleted);
                if (broken) return;
                continue;
            }
        }
        i++;
    }
}

// QuickHull::mergeConcaveFaces (QuickHull.cpp:485-533)
void Builder::merge_concave(std::vector<int>& nf, std::vector<char>& deleted) {
    size_t i = 0;
    while (i < nf.size()) {
        const int f = nf[i];
        if (!is_deleted(deleted, f)) {
            int concave = -1;
            const int first = F[f].edge;
            int e = first;
            do {
                if (E[e].twin < 0) { broken = true; return; }
                if (convex_edge(e)) { e = E[e].next; continue; }
                concave = e;
                break;
            } while (e != first);
            if (concave >= 0) {
                merge_at_edge(concave, deleted);
                if (broken) return;
                continue;
            }
        }
        i++;
    }
}

// QuickHull::mergeConcaveFacesAtEdge (QuickHull.cpp:536-600): the face across e joins e's face.
void Builder::merge_at_edge(int e, std::vector<char>& deleted) {
    const int tw = E[e].twin;
    const int drop = E[tw].face, keep = E[e].face;
    F[keep].edge = E[e].prev;
    for (int fe = E[tw].next; fe != tw; fe = E[fe].next) E[fe].face = keep;
    const int ep = E[e].prev, en = E[e].next, tp = E[tw].prev, tn = E[tw].next;
    E[ep].next = tn;
    E[en].prev = tp;
    E[tp].next = en;
    E[tn].prev = ep;
    F[keep].conflict.insert(F[keep].conflict.end(), F[drop].conflict.begin(), F[drop].conflict.end());
    mark(deleted, drop);
    delete_face(drop);
    remove_edge(tw);
    remove_edge(e);
    recalc(keep);
    fix_topology(keep, deleted);
}

// QuickHull::fixTopologicalIssues (QuickHull.cpp:603-640): a vertex whose two edges in the face border the
// same other face is redundant.
void Builder::fix_topology(int f, std::vector<char>& deleted) {
    int bad;
    do {
        bad = -1;
        const int first = F[f].edge;
        int in = first;
        do {
            const int out = E[in].next;
            if (E[in].twin < 0 || E[out].twin < 0) { broken = true; return; }
            if (E[E[in].twin].face == E[E[out].twin].face) { bad = in; break; }
            in = out;
        } while (in != first);
        if (bad >= 0) fix_at_edge(f, bad, deleted);
    } while (bad >= 0 && !broken);
}

// QuickHull::fixTopologicalIssueAtEdge (QuickHull.cpp:643-733)
void Builder::fix_at_edge(int f, int in, std::vector<char>& deleted) {
    const int tw = E[in].twin;
    if (triangle(E[tw].face)) {
        const int before = E[in].prev;
        const int after = E[E[in].next].next;
        const int drop = E[tw].face, keep = f;
        F[keep].edge = before;
        const int tn = E[tw].next;
        E[tn].face = keep;
        E[before].next = tn;
        E[tn].prev = before;
        E[after].prev = tn;
        E[tn].next = after;
        F[keep].conflict.insert(F[keep].conflict.end(), F[drop].conflict.begin(), F[drop].conflict.end());
        delete_face(drop);
        mark(deleted, drop);
        const int out = E[in].next;
        if (E[out].twin >= 0) remove_edge(E[out].twin);
        remove_edge(out);
        remove_edge(tw);
        remove_edge(in);
        recalc(keep);
    } else {
        const int out = E[in].next;
        const int otw = E[out].twin;
        E[tw].v0 = E[out].v1;
        E[E[otw].prev].next = tw;
        E[tw].prev = E[otw].prev;
        E[in].v1 = E[out].v1;
        E[in].next = E[out].next;
        E[E[in].next].prev = in;
        F[f].edge = in;
        F[E[tw].face].edge = tw;
        remove_edge(otw);
        remove_edge(out);
        recalc(f);
        recalc(E[tw].face);
    }
}

// QuickHull::addVertexToHull (QuickHull.cpp:197-233) with deleteVisibleFaces and buildNewFaces.
void Builder::add_vertex(uint32_t vertex, int face) {
    std::vector<int> hv, visible;
    horizon(P[vertex], face, hv, visible);
    if (broken) return;
    std::vector<uint32_t> orphans;
    for (int f : visible) {
        orphans.insert(orphans.end(), F[f].conflict.begin(), F[f].conflict.end());
        remove_face(f);
    }
    std::vector<int> nf;
    for (size_t i = 0; i < hv.size(); i += 2) {
        const int tri[3] = {hv[i], hv[i + 1], (int)vertex};
        nf.push_back(add_face(tri, 3));
    }
    std::vector<char> deleted;
    merge_large(nf, deleted);
    if (broken) return;
    merge_concave(nf, deleted);
    if (broken) return;
    assign(orphans, nf, deleted);
}

bool Builder::run(std::string* err) {
    const uint32_t n = (uint32_t)P.size();
    std::vector<uint32_t> orphans;
    orphans.reserve(n);
    float mx = 0.0f, my = 0.0f, mz = 0.0f;
    for (uint32_t i = 0; i < n; i++) {
        orphans.push_back(i);
        const float ax = std::fabs(P[i].x), ay = std::fabs(P[i].y), az = std::fabs(P[i].z);
        if (ax > mx) mx = ax;
        if (ay > my) my = ay;
        if (az > mz) mz = az;
    }
    eps = 3.0f * (mx + my + mz) * kFltEps;
    std::vector<int> init;
    if (!initial_hull(init, orphans, err)) return false;
    assign(orphans, init, std::vector<char>());
    uint32_t v = 0;
    int f = -1;
    while (next_candidate(v, f)) {
        add_vertex(v, f);
        if (broken) {
            *err = "QuickHull: degenerate point set (a hull edge lost its twin)";
            return false;
        }
    }
    return true;
}

// QuickHull::computeFinalPolygonVertexArray (QuickHull.cpp:134-194): faces from the most recent, each
// from its first half-edge; vertices numbered in order of first use.
void Builder::output(std::vector<float>& verts, std::vector<int32_t>& counts, std::vector<int32_t>& idx) const {
    std::unordered_map<int, int> renum;
    for (int f = head; f >= 0; f = F[f].next) {
        int cnt = 0;
        const int first = F[f].edge;
        int e = first;
        do {
            const int old = E[e].v0;
            auto it = renum.find(old);
            int nv;
            if (it != renum.end()) {
                nv = it->second;
            } else {
                nv = (int)(verts.size() / 3);
                renum.emplace(old, nv);
                verts.push_back(P[old].x);
                verts.push_back(P[old].y);
                verts.push_back(P[old].z);
            }
            idx.push_back(nv);
            cnt++;
            e = E[e].next;
        } while (e != first);
        counts.push_back(cnt);
    }
}

}  // namespace

// QuickHull::removeDuplicatedVertices (QuickHull.cpp:736-787): points closer than 1e-5 per axis are one
// point, the one farther from the cloud's centre kept.
static std::vector<V3> dedupe(const std::vector<V3>& pts) {
    V3 c = v3zero();
    for (const V3& p : pts) c += p;
    c = v3(c.x / (float)pts.size(), c.y / (float)pts.size(), c.z / (float)pts.size());
    std::vector<V3> keep;
    for (const V3& p : pts) {
        size_t j = 0;
        for (; j < keep.size(); j++) {
            const float dx = std::fabs(keep[j].x - p.x), dy = std::fabs(keep[j].y - p.y), dz = std::fabs(keep[j].z - p.z);
            if (dx < 0.00001f && dy < 0.00001f && dz < 0.00001f) {
                if (len2(p - c) > len2(keep[j] - c)) keep[j] = p;
                break;
            }
        }
        if (j == keep.size()) keep.push_back(p);
    }
    return keep;
}

bool quickhull_polygons(const std::vector<V3>& points, std::vector<float>& verts, std::vector<int32_t>& counts,
                        std::vector<int32_t>& idx, std::string* err) {
    verts.clear();
    counts.clear();
    idx.clear();
    if (points.empty()) {
        *err = "The VertexArray must contain at least 4 vertices to create a convex mesh";
        return false;
    }
    std::vector<V3> pts = dedupe(points);
    if (pts.size() < 4) {
        *err = "The VertexArray must contain at least 4 vertices to create a convex mesh";
        return false;
    }
    Builder b(std::move(pts));
    if (!b.run(err)) return false;
    b.output(verts, counts, idx);
    return true;
}

}  // namespace rp3dg
