// Oracle triangle-mesh narrowphase: CPU restatement of the reference's TriangleMesh (BVH
// build and query, src/TriangleMesh.cpp), the triangle distance routines of
// src/GeomUtil.cpp:61-127,469-665 and the mesh collision of src/CollisionTriangleMesh.cpp.
// TEST INFRASTRUCTURE ONLY: used by tests/ and the CPU baseline, never by the product.
//
// The edge/vertex contacts are ordered with std::sort, as the reference orders them
// (CollisionTriangleMesh.cpp:925); ties keep whatever order this standard library's
// introsort leaves, the same library the reference harness is linked against.
#include <algorithm>
#include <cfloat>
#include <limits>
#include <unordered_set>
#include "orc_collision.h"

namespace orc {

// ---------------------------------------------------------------- bounds helpers (src/Bounds.cpp)
static Bounds getUnion(const Bounds& a, const Bounds& b) { return {vmin(a.min, b.min), vmax(a.max, b.max)}; }
static float area(const Bounds& b) {                       // Bounds::area, src/Bounds.cpp:28-31
    V3 d = b.max - b.min;
    return 2.f * (d.x * d.y + d.y * d.z + d.z * d.x);
}
static bool intersects(const Bounds& a, const Bounds& b) {  // src/BoundsUtil.cpp:87-92
    for (int i = 0; i < 3; ++i)
        if (a.max[i] < b.min[i] || a.min[i] > b.max[i]) return false;
    return true;
}

// ---------------------------------------------------------------- TriangleMesh (src/TriangleMesh.cpp)
static const int NUM_BUCKETS = 6;

static void updateNodeBounds(TriMesh& m, int nodeId) {                // :7-14
    Bounds b{v3s(FLT_MAX), v3s(-FLT_MAX)};
    TriNode& node = m.bvh[nodeId];
    for (int i = 0; i < node.triCount; ++i) b = getUnion(b, m.tris[node.index + i].bounds);
    node.bounds = b;
}
static bool fillBuckets(TriMesh& m, int nodeId, int axis, Bounds* bb, int* bc) {   // :16-38
    TriNode& node = m.bvh[nodeId];
    for (int i = 0; i < NUM_BUCKETS; ++i) { bb[i] = {v3s(FLT_MAX), v3s(-FLT_MAX)}; bc[i] = 0; }
    float start = node.bounds.min[axis];
    float end = node.bounds.max[axis];
    float len = end - start;
    if (!len) return false;
    for (int i = 0; i < node.triCount; ++i) {
        const Tri& t = m.tris[node.index + i];
        int bi = std::min(NUM_BUCKETS - 1, static_cast<int>((t.centroid[axis] - start) / len * NUM_BUCKETS));
        bb[bi] = getUnion(bb[bi], t.bounds);
        bc[bi]++;
    }
    return true;
}
static bool evaluateSplit(int split, const Bounds* bb, const int* bc, float& cost, Bounds& bl, Bounds& br, int& cl,
                          int& cr) {                                   // :40-62
    bl = {v3s(FLT_MAX), v3s(-FLT_MAX)};
    br = {v3s(FLT_MAX), v3s(-FLT_MAX)};
    cl = 0; cr = 0;
    for (int i = 0; i < NUM_BUCKETS; ++i) {
        if (i < split) { bl = getUnion(bl, bb[i]); cl += bc[i]; }
        else { br = getUnion(br, bb[i]); cr += bc[i]; }
    }
    if (!cl || !cr) return false;
    cost = cl * area(bl) + cr * area(br);
    return true;
}
static bool chooseSplit(TriMesh& m, int nodeId, int& bestAxis, int& splitIndex, Bounds& bbl, Bounds& bbr, int& bcl,
                        int& bcr) {                                    // :64-94
    float bestCost = FLT_MAX;
    for (int axis = 0; axis < 3; ++axis) {
        Bounds bb[NUM_BUCKETS];
        int bc[NUM_BUCKETS];
        if (!fillBuckets(m, nodeId, axis, bb, bc)) continue;
        for (int i = 1; i < NUM_BUCKETS; ++i) {
            float cost;
            Bounds bl, br;
            int cl = 0, cr = 0;
            if (evaluateSplit(i, bb, bc, cost, bl, br, cl, cr) && cost < bestCost) {
                bestCost = cost; bestAxis = axis; splitIndex = i;
                bbl = bl; bbr = br; bcl = cl; bcr = cr;
            }
        }
    }
    return bcl && bcr;
}
static void subdivide(TriMesh& m, int nodeId) {                        // :96-137
    int bestAxis = 0, splitIndex = 0, bcl = 0, bcr = 0;
    Bounds bbl{}, bbr{};
    if (!chooseSplit(m, nodeId, bestAxis, splitIndex, bbl, bbr, bcl, bcr)) return;
    TriNode& node = m.bvh[nodeId];
    float start = node.bounds.min[bestAxis];
    float end = node.bounds.max[bestAxis];
    float len = end - start;
    int i = node.index;
    int j = i + node.triCount - 1;
    while (i <= j) {
        const Tri& t = m.tris[i];
        int bi = static_cast<int>((t.centroid[bestAxis] - start) / len * NUM_BUCKETS);
        if (bi < splitIndex) ++i;
        else std::swap(m.tris[i], m.tris[j--]);
    }
    int nodeIndex = node.index;
    m.bvh.push_back({bbl, bcl, nodeIndex});
    m.bvh.push_back({bbr, bcr, i});
    TriNode& n2 = m.bvh[nodeId];
    n2.triCount = 0;
    n2.index = (int)m.bvh.size() - 2;
    subdivide(m, n2.index);
    subdivide(m, m.bvh[nodeId].index + 1);
}

// TriangleMesh(vertices, indices), src/TriangleMesh.cpp:139-161
TriMesh buildTriMesh(const std::vector<V3>& verts, const std::vector<uint32_t>& indices) {
    TriMesh m;
    m.verts = verts;
    for (size_t i = 0; i + 2 < indices.size(); i += 3) {
        Tri t;
        t.idx[0] = indices[i]; t.idx[1] = indices[i + 1]; t.idx[2] = indices[i + 2];
        V3 a = verts[t.idx[0]], b = verts[t.idx[1]], c = verts[t.idx[2]];
        t.bounds = {vmin(vmin(a, b), c), vmax(vmax(a, b), c)};      // getBoundsTriangle (BoundsUtil.cpp:71-75)
        t.normal = normalize(cross(b - a, c - a));
        t.centroid = (a + b + c) / 3.f;
        m.tris.push_back(t);
    }
    m.bvh.reserve(2 * m.tris.size() + 1);
    m.bvh.push_back({{}, (int)m.tris.size(), 0});
    updateNodeBounds(m, 0);
    subdivide(m, 0);
    return m;
}

// TriangleMesh::overlapBvh, src/TriangleMesh.cpp:163-192 (depth first, left child first)
std::vector<int> overlapBvh(const TriMesh& m, const Bounds& b) {
    std::vector<int> out, stack;
    stack.push_back(0);
    while (!stack.empty()) {
        const TriNode& node = m.bvh[stack.back()];
        stack.pop_back();
        if (intersects(b, node.bounds)) {
            if (node.triCount) {
                for (int i = 0; i < node.triCount; ++i) out.push_back(node.index + i);
            } else {
                stack.push_back(node.index + 1);
                stack.push_back(node.index);
            }
        }
    }
    return out;
}

// ---------------------------------------------------------------- GeomUtil triangle routines
enum TriFeature { T_FACE = 0, T_EDGE = 1, T_VERTEX = 2 };

// src/GeomUtil.cpp:61-127
static void closestPointsBetweenSegments4(V3 o0, V3 v0, V3 o1, V3 v1, V3& p0, V3& p1) {
    float v0v0 = dot(v0, v0);
    if (v0v0 < 0.0001f) {
        float v1v1 = dot(v1, v1);
        if (v1v1 < 0.0001f) { p0 = o0; p1 = o1; return; }
        float t1 = -dot(o1 - o0, v1) / v1v1;
        p0 = o0; p1 = o1 + t1 * v1;
        return;
    }
    float v1v1 = dot(v1, v1);
    if (v1v1 < 0.0001f) {
        float t0 = -dot(o0 - o1, v0) / v0v0;
        p0 = o0 + t0 * v0; p1 = o1;
        return;
    }
    float v0v1 = dot(v0, v1);
    V3 r = o1 - o0;
    float rv0 = dot(r, v0), rv1 = dot(r, v1);
    float denom = v0v0 * v1v1 - v0v1 * v0v1;
    float t1 = gabs(denom) < 0.0001f ? 0 : (rv0 * v0v1 - rv1 * v0v0) / denom;
    float t0 = (rv0 + t1 * v0v1) / v0v0;
    if (t0 < 0) { t0 = 0; p0 = o0 + t0 * v0; t1 = -dot(o1 - p0, v1) / v1v1; }
    else if (t0 > 1) { t0 = 1; p0 = o0 + t0 * v0; t1 = -dot(o1 - p0, v1) / v1v1; }
    else p0 = o0 + t0 * v0;
    if (t1 < 0) {
        t1 = 0; p1 = o1 + t1 * v1;
        t0 = gclamp(-dot(o0 - p1, v0) / v0v0, 0.f, 1.f);
        p0 = o0 + t0 * v0;
    } else if (t1 > 1) {
        t1 = 1; p1 = o1 + t1 * v1;
        t0 = gclamp(-dot(o0 - p1, v0) / v0v0, 0.f, 1.f);
        p0 = o0 + t0 * v0;
    } else {
        p1 = o1 + t1 * v1;
    }
}

// src/GeomUtil.cpp:469-538
static float sqrDistPointTriangle(V3 p, V3 a, V3 b, V3 c, V3& q, int& feature, int& fi) {
    V3 ab = b - a, ac = c - a, ap = p - a;
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) { feature = T_VERTEX; fi = 0; q = a; return distance2(p, q); }
    V3 bp = p - b;
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) { feature = T_VERTEX; fi = 1; q = b; return distance2(p, q); }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0.f && d3 <= 0.f) {
        feature = T_EDGE; fi = 2;
        q = a + d1 / (d1 - d3) * ab;
        return distance2(p, q);
    }
    V3 cp = p - c;
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.f && d5 <= d6) { feature = T_VERTEX; fi = 2; q = c; return distance2(p, q); }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0 && d4 >= d3 && d5 >= d6) {
        feature = T_EDGE; fi = 0;
        q = b + (d4 - d3) / (d4 - d3 + d5 - d6) * (c - b);
        return distance2(p, q);
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0.f && d6 <= 0.f) {
        feature = T_EDGE; fi = 1;
        q = c + d2 / (d2 - d6) * ac;
        return distance2(p, q);
    }
    feature = T_FACE;
    float denom = va + vb + vc;
    float u = va / denom, v = vb / denom, w = 1 - u - v;
    q = u * a + v * b + w * c;
    return distance2(p, q);
}

// src/GeomUtil.cpp:540-581
static float sqrDistLineSegment(V3 p, V3 dir, V3 a, V3 b, float& t, V3& q, int& vflag) {
    V3 ab = b - a;
    float v1v2 = dot(dir, ab), v2v2 = dot(ab, ab);
    V3 r = p - a;
    float rv1 = dot(r, dir), rv2 = dot(r, ab);
    float denom = (v1v2 * v1v2 - v2v2);
    if (gabs(denom) <= 1e-6f) {
        q = a; t = -rv1; vflag = 0;
        V3 d = p + t * dir - a;
        return dot(d, d);
    }
    float s = (rv1 * v1v2 - rv2) / denom;
    if (s <= 0.f) {
        q = a; t = -rv1; vflag = 1;
        V3 d = p + t * dir - a;
        return dot(d, d);
    }
    if (s >= 1.f) {
        q = b; t = -dot(p - b, dir); vflag = 2;
        V3 d = p + t * dir - b;
        return dot(d, d);
    }
    q = a + s * ab;
    t = s * v1v2 - rv1;
    vflag = 0;
    V3 d = p + t * dir - q;
    return dot(d, d);
}

// glm::determinant(mat3), detail/func_matrix.inl:185-193
static float det3(V3 c0, V3 c1, V3 c2) {
    return +c0.x * (c1.y * c2.z - c2.y * c1.z) - c1.x * (c0.y * c2.z - c2.y * c0.z) + c2.x * (c0.y * c1.z - c1.y * c0.z);
}

// src/GeomUtil.cpp:583-647
static float sqrDistLineTriangle(V3 p, V3 dir, V3 a, V3 b, V3 c, float& t, V3& q, int& feature, int& fi) {
    V3 e0 = b - a, e1 = c - a;
    V3 nd = -dir;
    float det = det3(e0, e1, nd);
    if (gabs(det) > 1e-6f) {
        V3 diff = p - a;
        float invDet = 1 / det;
        float u = invDet * det3(diff, e1, nd);
        float v = invDet * det3(e0, diff, nd);
        float w = 1 - u - v;
        if (u >= 0.f && v >= 0.f && w >= 0.f) {
            t = invDet * det3(e0, e1, diff);
            q = a + u * e0 + v * e1;
            feature = T_FACE;
            return 0;
        }
    }
    float t1;
    V3 q1;
    int vf, vf1;
    fi = 2;
    float dist = sqrDistLineSegment(p, dir, a, b, t, q, vf);
    float dist1 = sqrDistLineSegment(p, dir, b, c, t1, q1, vf1);
    if (dist1 < dist) { fi = 0; t = t1; q = q1; vf = vf1; dist = dist1; }
    dist1 = sqrDistLineSegment(p, dir, c, a, t1, q1, vf1);
    if (dist1 < dist) { fi = 1; t = t1; q = q1; vf = vf1; dist = dist1; }
    if (vf) { feature = T_VERTEX; fi = (fi + vf) % 3; }
    else feature = T_EDGE;
    return dist;
}

// src/GeomUtil.cpp:649-660
static float sqrDistSegmentTriangle(V3 p, V3 dir, float mn, float mx, V3 a, V3 b, V3 c, float& t, V3& q,
                                    int& feature, int& fi) {
    float sq = sqrDistLineTriangle(p, dir, a, b, c, t, q, feature, fi);
    if (t < mn) { t = mn; sq = sqrDistPointTriangle(p + dir * t, a, b, c, q, feature, fi); }
    else if (t > mx) { t = mx; sq = sqrDistPointTriangle(p + dir * t, a, b, c, q, feature, fi); }
    return sq;
}

// src/GeomUtil.cpp:662-665
static float distanceAABBPlane(V3 he, V3 n, float d) {
    float r = he.x * gabs(n.x) + he.y * gabs(n.y) + he.z * gabs(n.z);
    return gabs(d) - r;
}

// ---------------------------------------------------------------- per-triangle tests
enum BoxFeat { B_FACE = 0, B_EDGE = 1, B_UNKNOWN = 2 };
struct TCI {                                       // TriangleContactInfo, CollisionTriangleMesh.cpp:22-36
    int triangleIndex = 0;
    V3 normal = {0, 0, 0}, cpBody = {0, 0, 0}, cpTri = {0, 0, 0};
    int feature = 0, featureIndex = 0;
    float distance = 0;
    int boxFeature = 0, boxAxis = 0;
    bool operator<(const TCI& o) const { return distance < o.distance; }
};

// :44-71
static bool sphereTriangle(V3 pos, float r, V3 a, V3 b, V3 c, V3 n, TCI& ci) {
    V3 u;
    int f, fi = 0;
    float sq = sqrDistPointTriangle(pos, a, b, c, u, f, fi);
    if (sq > r * r) return false;
    if (f == T_FACE) ci.normal = -n;
    else {
        V3 normal = u - pos;
        float len = length(normal);
        if (len) { normal = normal / len; ci.normal = normal; }
        else ci.normal = -n;
    }
    ci.cpBody = pos + ci.normal * r;
    ci.cpTri = u;
    ci.feature = f;
    ci.featureIndex = fi;
    ci.distance = sq;
    return true;
}
// :73-81
static void contactsSphereTriangle(V3 pos1, Q or1, const TCI& ci, std::vector<Manifold>& out) {
    Manifold m;
    m.normal = qrot(or1, ci.normal);
    m.numPoints = 1;
    m.points[0] = {pos1 + qrot(or1, ci.cpBody), pos1 + qrot(or1, ci.cpTri)};
    m.triangleIndex = ci.triangleIndex;
    out.push_back(m);
}
// :85-121
static bool capsuleTriangle(V3 pos, Q ori, float hh, float r, V3 a, V3 b, V3 c, V3 n, TCI& ci) {
    V3 dir = qrot(ori, v3(0, 1, 0));
    float t;
    V3 u;
    int f, fi = 0;
    float sq = sqrDistSegmentTriangle(pos, dir, -hh, hh, a, b, c, t, u, f, fi);
    if (sq > r * r) return false;
    V3 onSeg = pos + dir * t;
    if (f == T_FACE) ci.normal = -n;
    else {
        V3 normal = u - onSeg;
        float len = length(normal);
        if (len) { normal = normal / len; ci.normal = normal; }
        else ci.normal = -n;
    }
    ci.cpBody = onSeg + ci.normal * r;
    ci.cpTri = u;
    ci.feature = f;
    ci.featureIndex = fi;
    ci.distance = sq;
    return true;
}

// the shared face frame of the triangle-face generators: u0 toward vertex 0, u1 the normal
static void triFrame(const TriMesh& mesh, const Tri& tri, bool normalizeN, V3& refOrig, V3& u1, M3& meshToRef,
                     V2* clip) {
    refOrig = tri.centroid;
    V3 u0 = normalize(mesh.verts[tri.idx[0]] - refOrig);
    u1 = normalizeN ? normalize(tri.normal) : tri.normal;
    V3 u2 = cross(u0, u1);
    M3 b; b[0] = u0; b[1] = u1; b[2] = u2;
    meshToRef = transpose(b);
    for (int i = 0; i < 3; i++) {
        V3 v = mulmv(meshToRef, mesh.verts[tri.idx[i]]);
        clip[3 - i - 1] = {v.z, v.x};
    }
}

// :123-217
static void contactsCapsuleTriangleFace(V3 pos0, Q or0, float hh0, float r0, V3 pos1, Q or1, const TriMesh& mesh,
                                        const TCI& ci, std::vector<Manifold>& out) {
    V3 dir = qrot(or0, v3(0, 1, 0));
    V3 p0local = pos0 - dir * hh0;
    V3 p1local = pos0 + dir * hh0;
    const Tri& tri = mesh.tris[ci.triangleIndex];
    V3 refOrig, u1;
    M3 meshToRef;
    V2 clip[3];
    triFrame(mesh, tri, true, refOrig, u1, meshToRef, clip);
    V3 p0Ref = mulmv(meshToRef, p0local), p1Ref = mulmv(meshToRef, p1local);
    V2 line0 = {p0Ref.z, p0Ref.x}, line1 = {p1Ref.z, p1Ref.x};
    if (!clipLine(line0, line1, clip, 3)) return;
    M3 meshToWorld = mat3_cast(or1);
    M3 refToWorld = mulmm(meshToWorld, transpose(meshToRef));
    float distToPlane = dot(refOrig, u1);
    V3 p0 = v3s(0), p1 = v3s(0);
    float distX = p0Ref.z - p1Ref.z, distY = p0Ref.x - p1Ref.x;
    float adx = gabs(distX), ady = gabs(distY);
    if (adx && adx >= ady) {
        p0.y = gmix(p0Ref.y, p1Ref.y, (p0Ref.z - line0.x) / distX);
        p1.y = gmix(p1Ref.y, p0Ref.y, (p1Ref.z - line1.x) / -distX);
    } else if (ady > adx) {
        p0.y = gmix(p0Ref.y, p1Ref.y, (p0Ref.x - line0.y) / distY);
        p1.y = gmix(p1Ref.y, p0Ref.y, (p1Ref.x - line1.y) / -distY);
    } else {
        p0.y = p0Ref.y;
        p1.y = p1Ref.y;
    }
    p0.y -= r0;
    p1.y -= r0;
    int n = 0;
    Manifold res;
    res.triangleIndex = ci.triangleIndex;
    res.normal = mulmv(meshToWorld, ci.normal);
    if (p0.y < distToPlane) {
        p0.z = line0.x; p0.x = line0.y;
        V3 on = p0; on.y = distToPlane;
        res.points[n++] = {pos1 + mulmv(refToWorld, p0), pos1 + mulmv(refToWorld, on)};
    }
    if (p1.y < distToPlane) {
        p1.z = line1.x; p1.x = line1.y;
        V3 on = p1; on.y = distToPlane;
        res.points[n++] = {pos1 + mulmv(refToWorld, p1), pos1 + mulmv(refToWorld, on)};
    }
    if (!n) {
        n = 1;
        res.points[0] = {pos1 + mulmv(meshToWorld, ci.cpBody), pos1 + mulmv(meshToWorld, ci.cpTri)};
    }
    res.numPoints = n;
    out.push_back(res);
}

// :221-396
static bool boxTriangle(V3 pos, Q ori, V3 he, V3 a, V3 b, V3 c, V3 n, TCI& ci) {
    float edgeOffset = 0.f;
    const float edgeLimit = 0.999f;
    M3 u = mat3_cast(ori);
    M3 invU = transpose(u);
    V3 v0 = mulmv(invU, a - pos), v1 = mulmv(invU, b - pos), v2 = mulmv(invU, c - pos);
    V3 f0 = v1 - v0, f1 = v2 - v1, f2 = v0 - v2;
    float lim[3] = {edgeLimit * dot(f0, f0), edgeLimit * dot(f1, f1), edgeLimit * dot(f2, f2)};
    V3 fe[3] = {f0, f1, f2};
    float mx = std::numeric_limits<float>::lowest();
    // edge axes u_k x f_e in the reference's order: k = 0..2, e = 0..2; the projected pair
    // of vertices and the feature index of each edge
    const V3* pv[3][2] = {{&v0, &v2}, {&v0, &v1}, {&v1, &v2}};
    const int efi[3] = {2, 0, 1};
    for (int k = 0; k < 3; ++k)
        for (int e = 0; e < 3; ++e) {
            V3 f = fe[e];
            V3 A = *pv[e][0], B = *pv[e][1];
            float p0, p1, r;
            if (k == 0) {
                p0 = A.z * f.y - A.y * f.z; p1 = B.z * f.y - B.y * f.z;
                r = he.y * gabs(f.z) + he.z * gabs(f.y);
            } else if (k == 1) {
                p0 = A.x * f.z - A.z * f.x; p1 = B.x * f.z - B.z * f.x;
                r = he.x * gabs(f.z) + he.z * gabs(f.x);
            } else {
                p0 = A.y * f.x - A.x * f.y; p1 = B.y * f.x - B.x * f.y;
                r = he.x * gabs(f.y) + he.y * gabs(f.x);
            }
            float d = gmax(-gmax(p0, p1), gmin(p0, p1)) - r;
            if (d > 0) return false;
            if (f[k] * f[k] < lim[e] && d > mx) {
                mx = d;
                ci.feature = T_EDGE; ci.featureIndex = efi[e]; ci.distance = d;
                ci.boxFeature = B_EDGE; ci.boxAxis = k;
                edgeOffset = 0.1f;
            }
        }
    for (int i = 0; i < 3; ++i) {
        float d = gmax(-gmax(gmax(v0[i], v1[i]), v2[i]), gmin(gmin(v0[i], v1[i]), v2[i])) - he[i];
        if (d > 0) return false;
        if (d > mx - edgeOffset) {
            mx = d;
            bool i0 = gabs(v0[i]) < he[i], i1 = gabs(v1[i]) < he[i], i2 = gabs(v2[i]) < he[i];
            if (i0) {
                if (i1) {
                    if (i2) ci.feature = T_FACE;
                    else { ci.feature = T_EDGE; ci.featureIndex = 2; }
                } else {
                    if (i2) { ci.feature = T_EDGE; ci.featureIndex = 1; }
                    else { ci.feature = T_VERTEX; ci.featureIndex = 0; }
                }
            } else {
                if (i1) {
                    if (i2) { ci.feature = T_EDGE; ci.featureIndex = 0; }
                    else { ci.feature = T_VERTEX; ci.featureIndex = 1; }
                } else {
                    ci.feature = T_VERTEX; ci.featureIndex = 2;
                }
            }
            ci.boxFeature = B_FACE; ci.boxAxis = i;
            ci.distance = d;
            edgeOffset = 0.f;
        }
    }
    V3 nLocal = mulmv(invU, n);
    float d = distanceAABBPlane(he, nLocal, dot(nLocal, v0));
    if (d > 0) return false;
    if (d > mx - edgeOffset) {
        ci.feature = T_FACE;
        ci.distance = d;
        ci.boxFeature = B_UNKNOWN;
    }
    return true;
}

// :398-467
static void contactsBoxTriangleFace(V3 boxPos, Q boxOr, V3 he, V3 meshPos, Q meshOr, const TriMesh& mesh,
                                    const TCI& ci, std::vector<Manifold>& out) {
    M3 basis = mat3_cast(boxOr);
    const Tri& tri = mesh.tris[ci.triangleIndex];
    V3 refOrig, u1;
    M3 meshToRef;
    V2 clipA[3];
    triFrame(mesh, tri, true, refOrig, u1, meshToRef, clipA);
    std::vector<V2> clip(clipA, clipA + 3);
    int incAxis = 0;
    float incSign = 0;
    if (ci.boxFeature == B_FACE) {
        incAxis = ci.boxAxis;
        incSign = dot(tri.normal, basis[incAxis]) > 0 ? -1.f : 1.f;
    } else {
        float maxDot = 0;
        for (int i = 0; i < 3; ++i) {
            float d = dot(tri.normal, basis[i]);
            float ad = gabs(d);
            if (ad > maxDot) { maxDot = ad; incAxis = i; incSign = d > 0 ? -1.f : 1.f; }
        }
    }
    int ia1 = (incAxis + 1) % 3, ia2 = (incAxis + 2) % 3;
    V3 incOrig = mulmv(meshToRef, boxPos + basis[incAxis] * he[incAxis] * incSign);
    V3 A = basis[ia1] * he[ia1], B = basis[ia2] * he[ia2];
    V3 nA = -basis[ia1] * he[ia1];
    V3 poly0 = incOrig + mulmv(meshToRef, A + B);
    V3 poly1 = incOrig + mulmv(meshToRef, nA + B);
    V3 poly2 = incOrig + mulmv(meshToRef, nA - B);
    V3 poly3 = incOrig + mulmv(meshToRef, A - B);
    std::vector<V2> polygon = {{poly0.z, poly0.x}, {poly1.z, poly1.x}, {poly2.z, poly2.x}, {poly3.z, poly3.x}};
    suthHodgClip(polygon, clip);
    V3 incNormal = mulmv(meshToRef, basis[incAxis]);
    M3 meshToWorld = mat3_cast(meshOr);
    M3 refToWorld = mulmm(meshToWorld, transpose(meshToRef));
    Manifold m;
    m.normal = mulmv(mneg(meshToWorld), tri.normal);
    m.triangleIndex = ci.triangleIndex;
    contactsPolygonPolygonFace(meshPos, refToWorld, refOrig, u1, incOrig, incNormal, polygon, 2, 0, m.points,
                               m.numPoints);
    for (int i = 0; i < m.numPoints; ++i) m.points[i] = {m.points[i].p1, m.points[i].p0};
    out.push_back(m);
}

// :469-512
static void contactsBoxFaceTriangle(V3 boxPos, Q boxOr, V3 he, V3 meshPos, Q meshOr, const TriMesh& mesh,
                                    const TCI& ci, std::vector<Manifold>& out) {
    M3 basis = mat3_cast(boxOr);
    const Tri& tri = mesh.tris[ci.triangleIndex];
    M3 meshToBox = transpose(basis);
    int refAxis = ci.boxAxis;
    float refSign = dot(basis[refAxis], tri.centroid - boxPos) < 0 ? -1.f : 1.f;
    int cx = (refAxis + 1) % 3, cy = (refAxis + 2) % 3;
    std::vector<V2> polygon(3);
    for (int i = 0; i < 3; ++i) {
        V3 v = mulmv(meshToBox, mesh.verts[tri.idx[i]] - boxPos);
        polygon[i] = {v[cx], v[cy]};
    }
    std::vector<V2> clip = {{he[cx], he[cy]}, {he[cx], -he[cy]}, {-he[cx], -he[cy]}, {-he[cx], he[cy]}};
    suthHodgClip(polygon, clip);
    V3 incOrig = mulmv(meshToBox, tri.centroid - boxPos);
    V3 incNormal = mulmv(meshToBox, tri.normal);
    M3 meshToWorld = mat3_cast(meshOr);
    Manifold m;
    m.normal = mulmv(meshToWorld, refSign * basis[refAxis]);
    m.triangleIndex = ci.triangleIndex;
    generateContactsPolygonBoxFace(boxPos, basis, refAxis, (int)refSign, he, incOrig, incNormal, polygon, cx, cy,
                                   m.points, m.numPoints);
    for (int i = 0; i < m.numPoints; ++i) {
        m.points[i].p0 = meshPos + mulmv(meshToWorld, m.points[i].p0);
        m.points[i].p1 = meshPos + mulmv(meshToWorld, m.points[i].p1);
    }
    out.push_back(m);
}

// :514-542
static void contactsBoxEdgeTriangleEdge(V3 boxPos, Q boxOr, V3 he, V3 meshPos, Q meshOr, const TriMesh& mesh,
                                        const TCI& ci, std::vector<Manifold>& out) {
    M3 basis = mat3_cast(boxOr);
    const Tri& tri = mesh.tris[ci.triangleIndex];
    int boxAxis = ci.boxAxis;
    int e = ci.featureIndex;
    V3 triA = mesh.verts[tri.idx[(e + 1) % 3]];
    V3 triB = mesh.verts[tri.idx[(e + 2) % 3]];
    V3 triEdge = triB - triA;
    V3 axis = cross(basis[boxAxis], triEdge);
    M3 meshToWorld = mat3_cast(meshOr);
    Manifold m;
    m.normal = mulmv(meshToWorld, normalize(dot(axis, tri.centroid - boxPos) < 0 ? -axis : axis));
    m.numPoints = 1;
    m.triangleIndex = ci.triangleIndex;
    V3 sb = v3s(-1);
    sb[(boxAxis + 1) % 3] = dot(basis[(boxAxis + 1) % 3], m.normal) < 0 ? -1.f : 1.f;
    sb[(boxAxis + 2) % 3] = dot(basis[(boxAxis + 2) % 3], m.normal) < 0 ? -1.f : 1.f;
    V3 boxA = boxPos + sb[0] * he[0] * basis[0] + sb[1] * he[1] * basis[1] + sb[2] * he[2] * basis[2];
    V3 boxEdge = basis[boxAxis] * he[boxAxis] * 2.f;
    V3 p0, p1;
    closestPointsBetweenSegments4(boxA, boxEdge, triA, triEdge, p0, p1);
    m.points[0] = {meshPos + mulmv(meshToWorld, p0), meshPos + mulmv(meshToWorld, p1)};
    out.push_back(m);
}

// :616-697
static bool convexTriangle(V3 pos, Q ori, const ConvexMesh* mesh, V3 scale, V3 a, V3 b, V3 c, V3 centroid,
                           TCI& ci) {
    V3 cb, ct;
    V3 nrm;
    if (!gjkEpaConvexTriangle(pos, ori, mesh, scale, a, b, c, centroid - pos, nrm, cb, ct)) return false;
    ci.normal = nrm;
    ci.cpBody = cb;
    ci.cpTri = ct;
    ci.distance = dot(ci.cpTri - ci.cpBody, ci.normal);
    V3 v0 = b - a, v1 = c - a, v2 = ci.cpTri - a;
    float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    float denom = d00 * d11 - d01 * d01;
    float u, v, w;
    if (denom) {
        v = (d20 * d11 - d21 * d01) / denom;
        w = (d00 * d21 - d01 * d20) / denom;
        u = 1.0f - v - w;
    } else {
        w = 0;
        if (d00) { v = d20 / d00; u = 1.0f - v; }
        else { v = 0; u = 1.0; }
    }
    const float eps = 1e-4f;
    if (u < eps) {
        if (v < eps) { ci.feature = T_VERTEX; ci.featureIndex = 2; }
        else if (w < eps) { ci.feature = T_VERTEX; ci.featureIndex = 1; }
        else { ci.feature = T_EDGE; ci.featureIndex = 0; }
    } else {
        if (v < eps) {
            if (w < eps) { ci.feature = T_VERTEX; ci.featureIndex = 0; }
            else { ci.feature = T_EDGE; ci.featureIndex = 1; }
        } else if (w < eps) {
            if (u < eps) { ci.feature = T_VERTEX; ci.featureIndex = 0; }
            else { ci.feature = T_EDGE; ci.featureIndex = 2; }
        } else {
            ci.feature = T_FACE;
        }
    }
    return true;
}

// :699-757
static void contactsConvexTriangleFace(V3 cpos, Q cor, const ConvexMesh* cv, V3 scale, V3 meshPos, Q meshOr,
                                       const TriMesh& mesh, const TCI& ci, std::vector<Manifold>& out) {
    const Tri& tri = mesh.tris[ci.triangleIndex];
    V3 refOrig, u1;
    M3 meshToRef;
    V2 clipA[3];
    triFrame(mesh, tri, false, refOrig, u1, meshToRef, clipA);
    std::vector<V2> clip(clipA, clipA + 3);
    int fi = 0;   // the reference leaves it unset when no face has a negative dot
    float minDot = 0;
    for (int i = 0; i < (int)cv->faces.size(); ++i) {
        float d = dot(tri.normal, normalize(qrot(cor, cv->faces[i].normal / scale)));
        if (d < minDot) { minDot = d; fi = i; }
    }
    const ConvexFace& F = cv->faces[fi];
    std::vector<V2> polygon(F.idx.size());
    for (size_t i = 0; i < F.idx.size(); ++i) {
        V3 v = mulmv(meshToRef, cpos + qrot(cor, scale * cv->verts[F.idx[i]]));
        polygon[i] = {v.z, v.x};
    }
    suthHodgClip(polygon, clip);
    V3 incOrig = mulmv(meshToRef, cpos + qrot(cor, scale * F.centroid));
    V3 incNormal = mulmv(meshToRef, qrot(cor, normalize(F.normal / scale)));
    M3 meshToWorld = mat3_cast(meshOr);
    M3 refToWorld = mulmm(meshToWorld, transpose(meshToRef));
    Manifold m;
    m.normal = mulmv(mneg(meshToWorld), tri.normal);
    m.triangleIndex = ci.triangleIndex;
    contactsPolygonPolygonFace(meshPos, refToWorld, refOrig, u1, incOrig, incNormal, polygon, 2, 0, m.points,
                               m.numPoints);
    for (int i = 0; i < m.numPoints; ++i) m.points[i] = {m.points[i].p1, m.points[i].p0};
    out.push_back(m);
}

// :759-824
static void contactsConvexFaceTriangle(V3 cpos, Q cor, const ConvexMesh* cv, V3 scale, V3 meshPos, Q meshOr,
                                       const TriMesh& mesh, const TCI& ci, std::vector<Manifold>& out) {
    const Tri& tri = mesh.tris[ci.triangleIndex];
    M3 convexToMesh = mat3_cast(cor);
    M3 meshToConvex = transpose(convexToMesh);
    int fi = 0;   // unset in the reference when no face has a positive dot
    float maxDot = 0;
    for (int i = 0; i < (int)cv->faces.size(); ++i) {
        float d = dot(ci.normal, normalize(qrot(cor, cv->faces[i].normal / scale)));
        if (d > maxDot) { maxDot = d; fi = i; }
    }
    const ConvexFace& F = cv->faces[fi];
    V3 refOrig = scale * F.centroid;
    V3 u0 = normalize(scale * cv->verts[F.idx[0]] - refOrig);
    V3 u1 = normalize(F.normal / scale);
    V3 u2 = cross(u0, u1);
    M3 b; b[0] = u0; b[1] = u1; b[2] = u2;
    M3 convexToRef = transpose(b);
    M3 meshToRef = mulmm(convexToRef, meshToConvex);
    std::vector<V2> clip(F.idx.size());
    for (size_t i = 0; i < F.idx.size(); ++i) {
        V3 v = mulmv(convexToRef, scale * cv->verts[F.idx[i]]);
        clip[i] = {v.z, v.x};
    }
    std::vector<V2> polygon(3);
    for (int i = 0; i < 3; ++i) {
        V3 v = mulmv(meshToRef, mesh.verts[tri.idx[i]] - cpos);
        polygon[i] = {v.z, v.x};
    }
    suthHodgClip(polygon, clip);
    V3 incOrig = mulmv(meshToRef, tri.centroid - cpos);
    V3 incNormal = mulmv(meshToRef, tri.normal);
    M3 meshToWorld = mat3_cast(meshOr);
    M3 refToMesh = transpose(meshToRef);
    Manifold m;
    m.triangleIndex = ci.triangleIndex;
    contactsPolygonPolygonFace(cpos, refToMesh, refOrig, u1, incOrig, incNormal, polygon, 2, 0, m.points,
                               m.numPoints);
    if (m.numPoints) {
        for (int i = 0; i < m.numPoints; ++i) {
            m.points[i].p0 = meshPos + mulmv(meshToWorld, m.points[i].p0);
            m.points[i].p1 = meshPos + mulmv(meshToWorld, m.points[i].p1);
        }
        m.normal = mulmv(mulmm(meshToWorld, convexToMesh), u1);
    } else {
        m.normal = mulmv(meshToWorld, ci.normal);
        m.numPoints = 1;
        m.points[0] = {meshPos + mulmv(meshToWorld, ci.cpBody), meshPos + mulmv(meshToWorld, ci.cpTri)};
    }
    out.push_back(m);
}

// ---------------------------------------------------------------- dispatch (:828-880)
static bool collisionTriangle(V3 pos, Q ori, const Geom& g, V3 a, V3 b, V3 c, V3 n, V3 centroid, TCI& ci) {
    switch (g.type) {
        case SPHERE: return sphereTriangle(pos, g.p[0], a, b, c, n, ci);
        case CAPSULE: return capsuleTriangle(pos, ori, g.p[0], g.p[1], a, b, c, n, ci);
        case BOX: return boxTriangle(pos, ori, v3(g.p[0], g.p[1], g.p[2]), a, b, c, n, ci);
        case CONVEX_MESH: return convexTriangle(pos, ori, g.mesh, v3(g.p[0], g.p[1], g.p[2]), a, b, c, centroid, ci);
    }
    return false;
}
static void contactsFace(V3 p0, Q o0, const Geom& g, V3 p1, Q o1, const TriMesh& m, const TCI& ci,
                         std::vector<Manifold>& out) {
    switch (g.type) {
        case SPHERE: return contactsSphereTriangle(p1, o1, ci, out);
        case CAPSULE: return contactsCapsuleTriangleFace(p0, o0, g.p[0], g.p[1], p1, o1, m, ci, out);
        case BOX: return contactsBoxTriangleFace(p0, o0, v3(g.p[0], g.p[1], g.p[2]), p1, o1, m, ci, out);
        case CONVEX_MESH:
            return contactsConvexTriangleFace(p0, o0, g.mesh, v3(g.p[0], g.p[1], g.p[2]), p1, o1, m, ci, out);
    }
}
static void contactsEdge(V3 p0, Q o0, const Geom& g, V3 p1, Q o1, const TriMesh& m, const TCI& ci,
                         std::vector<Manifold>& out) {
    switch (g.type) {
        case SPHERE: case CAPSULE: return contactsSphereTriangle(p1, o1, ci, out);
        case BOX:
            if (ci.boxFeature == B_FACE) return contactsBoxFaceTriangle(p0, o0, v3(g.p[0], g.p[1], g.p[2]), p1, o1, m, ci, out);
            return contactsBoxEdgeTriangleEdge(p0, o0, v3(g.p[0], g.p[1], g.p[2]), p1, o1, m, ci, out);
        case CONVEX_MESH:
            return contactsConvexFaceTriangle(p0, o0, g.mesh, v3(g.p[0], g.p[1], g.p[2]), p1, o1, m, ci, out);
    }
}
static void contactsVertex(V3 p0, Q o0, const Geom& g, V3 p1, Q o1, const TriMesh& m, const TCI& ci,
                           std::vector<Manifold>& out) {
    switch (g.type) {
        case SPHERE: case CAPSULE: return contactsSphereTriangle(p1, o1, ci, out);
        case BOX: return contactsBoxFaceTriangle(p0, o0, v3(g.p[0], g.p[1], g.p[2]), p1, o1, m, ci, out);
        case CONVEX_MESH:
            return contactsConvexFaceTriangle(p0, o0, g.mesh, v3(g.p[0], g.p[1], g.p[2]), p1, o1, m, ci, out);
    }
}

// physecs::collisionTriangleMesh, src/CollisionTriangleMesh.cpp:882-956
bool collisionTriangleMesh(V3 pos0, Q or0, const Geom& g0, V3 pos1, Q or1, const TriMesh& mesh,
                           std::vector<Manifold>& out) {
    Q inv1 = qinverse(or1);
    V3 localPos = qrot(inv1, pos0 - pos1);
    Q localOr = qmul(inv1, or0);
    Bounds lb = getBounds(localPos, localOr, g0);
    std::vector<int> cand = overlapBvh(mesh, lb);
    if (cand.empty()) return false;
    std::vector<TCI> contacts;
    for (int ti : cand) {
        const Tri& tri = mesh.tris[ti];
        V3 a = mesh.verts[tri.idx[0]], b = mesh.verts[tri.idx[1]], c = mesh.verts[tri.idx[2]];
        TCI ci;
        if (collisionTriangle(localPos, localOr, g0, a, b, c, tri.normal, tri.centroid, ci)) {
            ci.triangleIndex = ti;
            contacts.push_back(ci);
        }
    }
    if (contacts.empty()) return false;
    std::unordered_set<uint32_t> voided;
    for (int i = (int)contacts.size() - 1; i >= 0; i--) {
        if (contacts[i].feature == T_FACE) {
            const Tri& tri = mesh.tris[contacts[i].triangleIndex];
            voided.insert(tri.idx[0]); voided.insert(tri.idx[1]); voided.insert(tri.idx[2]);
            contactsFace(localPos, localOr, g0, pos1, or1, mesh, contacts[i], out);
            std::swap(contacts[i], contacts.back());
            contacts.pop_back();
        }
    }
    std::sort(contacts.begin(), contacts.end());
    for (auto& ci : contacts) {
        const Tri& tri = mesh.tris[ci.triangleIndex];
        if (ci.feature == T_EDGE) {
            if (voided.count(tri.idx[(ci.featureIndex + 1) % 3]) && voided.count(tri.idx[(ci.featureIndex + 2) % 3]))
                continue;
            contactsEdge(localPos, localOr, g0, pos1, or1, mesh, ci, out);
        } else if (ci.feature == T_VERTEX) {
            if (voided.count(tri.idx[ci.featureIndex])) continue;
            contactsVertex(localPos, localOr, g0, pos1, or1, mesh, ci, out);
        }
        voided.insert(tri.idx[0]); voided.insert(tri.idx[1]); voided.insert(tri.idx[2]);
    }
    return true;
}

}  // namespace orc
