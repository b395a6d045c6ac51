// Oracle convex-mesh narrowphase: CPU restatement of the reference's GJK (src/GJK.h),
// EPA (src/EPA.h) and the convex contact generators (src/Collision.cpp:352-426,488-499,
// 694-887; src/ContactsUtil.cpp:109-204). TEST INFRASTRUCTURE ONLY: used by tests/ and
// the CPU baseline, never by the product.
//
// GJK and EPA keep the reference's mixed precision exactly: simplex and polytope
// vertices are doubles (GjkVertex holds glm::dvec3), every support query narrows its
// direction to float, and the places where the reference passes a double vector to a
// float helper (isZero, sqrDistPointToLine, distPointToPlane, the barycentric edge
// vectors) narrow it the same way.
#include <limits>
#include "orc_collision.h"

namespace orc {

thread_local long epaExhausted = 0, epaZeroNormal = 0;

// ---------------------------------------------------------------- support maps
// src/GJK.h:13-20
struct SupSphere {
    V3 pos; float r;
    V3 operator()(V3 d) const { return pos + d * r; }
};
// src/GJK.h:22-31
struct SupCapsule {
    V3 pos, axis; float hh, r;
    V3 operator()(V3 d) const { return pos + axis * gsign(dot(d, axis)) * hh + d * r; }
};
// src/GJK.h:33-45
struct SupBox {
    V3 pos; M3 l2w; V3 he;
    V3 operator()(V3 d) const {
        V3 s = pos;
        for (int i = 0; i < 3; ++i) s = s + gsign(dot(d, l2w[i])) * l2w[i] * he[i];
        return s;
    }
};
// src/MathUtil.h:6-8
static V3 multiplyTranspose(const M3& m, V3 v) { return {dot(v, m[0]), dot(v, m[1]), dot(v, m[2])}; }
// src/GJK.h:47-107: four interleaved running maxima (vertex i goes to lane i % 4, strict
// '>' keeps the first index per lane), then lane 0..3 compared with strict '>'.
struct SupConvex {
    V3 pos; M3 l2w; const ConvexMesh* m; V3 scale;
    V3 operator()(V3 dir) const {
        dir = scale * multiplyTranspose(l2w, dir);
        float mv[4];
        int mi[4] = {0, 1, 2, 3};
        for (int k = 0; k < 4; ++k) mv[k] = std::numeric_limits<float>::lowest();
        for (int i = 0; i < m->count; i += 4)
            for (int k = 0; k < 4; ++k) {
                const V3& v = m->verts[i + k];
                float d = (v.x * dir.x + v.y * dir.y) + v.z * dir.z;
                if (d > mv[k]) mi[k] = i + k;
                mv[k] = d > mv[k] ? d : mv[k];
            }
        int best = mi[0];
        float bv = mv[0];
        for (int k = 1; k < 4; ++k)
            if (mv[k] > bv) { bv = mv[k]; best = mi[k]; }
        return pos + mulmv(l2w, scale * m->verts[best]);
    }
};

// ---------------------------------------------------------------- GJK (src/GJK.h:131-292)
struct GjkVertex { D3 pos, sp0, sp1; };
struct GjkSimplex { GjkVertex v[4]; GjkVertex& operator[](int i) { return v[i]; } };

template <class S0, class S1>
static GjkVertex mdiff(const S0& s0, const S1& s1, V3 dir) {          // src/GJK.h:192-196
    V3 a = s0(dir);
    V3 b = s1(-dir);
    return {d3(a - b), d3(a), d3(b)};
}
static float sqrDistPointToLine(V3 p, V3 a, V3 b) {                   // src/GJK.h:142-148
    V3 r = p - a;
    V3 d = b - a;
    V3 q = a + dot(r, d) / dot(d, d) * d;
    V3 pq = p - q;
    return dot(pq, pq);
}
static float distPointToPlane(V3 p, V3 o, V3 n) { return gabs(dot(p - o, n)); }   // :150-152
static bool isZero(V3 v) { return (double)dot(v, v) < 1e-6; }                      // :158-160

template <class S0, class S1>
static bool checkDirection(const S0& s0, const S1& s1, D3 dir, GjkVertex& v) {   // :162-167
    dir = dnormalize(dir);
    v = mdiff(s0, s1, f3(dir));
    return ddot(v.pos, dir) >= 0;
}
static bool checkFace(D3 v0, D3 v1, D3 v2, D3 opposite, D3& n) {                  // :169-173
    n = dcross(v1 - v0, v2 - v0);
    n = dsign(ddot(v0 - opposite, n)) * n;
    return ddot(n, -v0) > (double)0.00001f;
}
static bool simplexContainsOrigin(GjkSimplex& s, D3& dir, int& faceIndex) {        // :175-181
    faceIndex = 2;
    if (checkFace(s[3].pos, s[0].pos, s[1].pos, s[faceIndex].pos, dir)) return false;
    if (checkFace(s[3].pos, s[2].pos, s[0].pos, s[--faceIndex].pos, dir)) return false;
    if (checkFace(s[3].pos, s[1].pos, s[2].pos, s[--faceIndex].pos, dir)) return false;
    return true;
}
// the arbitrary tetrahedron the reference builds around a degenerate start (:233-243,253-263)
template <class S0, class S1>
static void completeFromLine(const S0& s0, const S1& s1, GjkSimplex& s) {
    D3 dir = dcross(s[0].pos - s[1].pos, d3(0, 1, 0));
    if (isZero(f3(dir))) dir = dcross(s[0].pos - s[1].pos, d3(0, 0, 1));
    s[2] = mdiff(s0, s1, f3(dnormalize(dir)));
    if (sqrDistPointToLine(f3(s[2].pos), f3(s[0].pos), f3(s[1].pos)) < 0.00001f) s[2] = mdiff(s0, s1, f3(-dir));
    dir = dnormalize(dcross(s[1].pos - s[0].pos, s[2].pos - s[0].pos));
    s[3] = mdiff(s0, s1, f3(dir));
    if (distPointToPlane(f3(s[3].pos), f3(s[0].pos), f3(dir)) < 0.0001f) s[3] = mdiff(s0, s1, f3(-dir));
}
// src/GJK.h:226-292 (the overload that returns the simplex)
template <class S0, class S1>
static bool gjk(const S0& s0, const S1& s1, V3 startDir, GjkSimplex& s) {
    D3 dir = d3(normalize(startDir));
    s[0] = mdiff(s0, s1, f3(dir));
    if (isZero(f3(s[0].pos))) {
        dir = -dir;
        s[1] = mdiff(s0, s1, f3(dir));
        completeFromLine(s0, s1, s);
        return true;
    }
    dir = -s[0].pos;
    if (!checkDirection(s0, s1, dir, s[1])) return false;
    dir = dcross(dcross(s[0].pos - s[1].pos, -s[1].pos), s[0].pos - s[1].pos);
    if (isZero(f3(dir))) {
        completeFromLine(s0, s1, s);
        return true;
    }
    if (!checkDirection(s0, s1, dir, s[2])) return false;
    dir = dcross(s[1].pos - s[0].pos, s[2].pos - s[0].pos);
    double sg = dsign(ddot(-s[0].pos, dir));
    if (!sg) {
        dir = dnormalize(dir);
        s[3] = mdiff(s0, s1, f3(dir));
        if (distPointToPlane(f3(s[3].pos), f3(s[0].pos), f3(dir)) < 0.0001f) s[3] = mdiff(s0, s1, f3(-dir));
        return true;
    }
    dir = sg * dir;
    if (!checkDirection(s0, s1, dir, s[3])) return false;
    int faceIndex;
    for (int i = 0; i < 100; ++i) {
        if (simplexContainsOrigin(s, dir, faceIndex)) return true;
        if (!checkDirection(s0, s1, dir, s[faceIndex])) return false;
        GjkVertex t = s[faceIndex];
        s[faceIndex] = s[3];
        s[3] = t;
    }
    return false;
}

// ---------------------------------------------------------------- EPA (src/EPA.h:9-186)
struct EpaFace { int i[3]; D3 n; };
struct EpaEdge { int i0, i1; };
struct Polytope {
    std::vector<GjkVertex> v;
    std::vector<EpaFace> f;
    void createFace(int i0, int i1, int i2) {                            // :24-34
        D3 n = dcross(v[i0].pos - v[i1].pos, v[i2].pos - v[i1].pos);
        double len = dlength(n);
        if (len) n = n / len;
        else ++epaZeroNormal;
        f.push_back({{i0, i1, i2}, n});
    }
    double distanceToFace(const EpaFace& fc) const { return ddot(v[fc.i[0]].pos, fc.n); }       // :41-43
    bool visible(const EpaFace& fc, D3 p) const { return ddot(fc.n, p - v[fc.i[0]].pos) > 0; } // :45-47
    explicit Polytope(GjkSimplex& s) {                                   // :50-66
        v.assign(&s[0], &s[0] + 4);
        D3 n = dcross(v[0].pos - v[1].pos, v[2].pos - v[1].pos);
        if (ddot(v[3].pos - v[0].pos, n) < 0) {
            createFace(0, 1, 2); createFace(3, 1, 0); createFace(3, 2, 1); createFace(3, 0, 2);
        } else {
            createFace(0, 2, 1); createFace(3, 2, 0); createFace(3, 1, 2); createFace(3, 0, 1);
        }
    }
    // :68-79 -- the running minimum is a float compared against double distances
    void findClosestFace(float& distance, D3& normal, int& faceIndex) const {
        distance = std::numeric_limits<float>::infinity();   // (float)DBL_MAX
        for (int i = 0; i < (int)f.size(); ++i) {
            double d = distanceToFace(f[i]);
            if (d < distance) { distance = (float)d; normal = f[i].n; faceIndex = i; }
        }
    }
    void insertVertex(const GjkVertex& vx) {                              // :81-120
        std::vector<EpaEdge> loose;
        for (int i = (int)f.size() - 1; i >= 0; --i) {
            EpaFace fc = f[i];
            if (visible(fc, vx.pos)) {
                EpaEdge e[3] = {{fc.i[0], fc.i[1]}, {fc.i[1], fc.i[2]}, {fc.i[2], fc.i[0]}};
                bool found[3] = {false, false, false};
                for (int j = 0; j < 3; ++j)
                    for (int k = (int)loose.size() - 1; k >= 0; --k)
                        if (loose[k].i0 == e[j].i1 && loose[k].i1 == e[j].i0) {
                            found[j] = true;
                            loose[k] = loose.back();
                            loose.pop_back();
                            break;
                        }
                for (int j = 0; j < 3; ++j)
                    if (!found[j]) loose.push_back(e[j]);
                f[i] = f.back();
                f.pop_back();
            }
        }
        int vi = (int)v.size();
        v.push_back(vx);
        for (auto& e : loose) createFace(vi, e.i0, e.i1);
    }
    // :122-168, including the proj = normal + distance sum (:133)
    void closestPoints(int faceIndex, V3& cp0, V3& cp1) const {
        const EpaFace& fc = f[faceIndex];
        GjkVertex a = v[fc.i[0]], b = v[fc.i[1]], c = v[fc.i[2]];
        D3 proj = fc.n + distanceToFace(fc);
        V3 v0 = f3(b.pos - a.pos), v1 = f3(c.pos - a.pos), v2 = f3(proj - a.pos);
        double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
        double denom = d00 * d11 - d01 * d01;
        double u, vv, w;
        if (denom) {
            vv = (d20 * d11 - d21 * d01) / denom;
            w = (d00 * d21 - d01 * d20) / denom;
            u = 1.0f - vv - w;
        } else {
            w = 0;
            if (d00) { vv = d20 / d00; u = 1.0f - vv; }
            else { vv = 0; u = 1.0; }
        }
        cp0 = f3(u * a.sp0 + vv * b.sp0 + w * c.sp0);
        cp1 = f3(u * a.sp1 + vv * b.sp1 + w * c.sp1);
    }
};

// src/EPA.h:171-186
template <class S0, class S1>
static V3 epa(const S0& s0, const S1& s1, GjkSimplex& s, V3& cp0, V3& cp1) {
    Polytope p(s);
    float minDist;
    D3 normal = {0, 0, 0};
    int faceIndex = 0;
    int i = 0;
    for (; i < 100; ++i) {
        p.findClosestFace(minDist, normal, faceIndex);
        GjkVertex v = mdiff(s0, s1, f3(normal));
        if (ddot(v.pos, normal) - minDist < (double)0.00001f) break;
        p.insertVertex(v);
    }
    if (i == 100) ++epaExhausted;
    p.closestPoints(faceIndex, cp0, cp1);
    return f3(normal);
}

// TriangleSupportFunction, src/GJK.h:109-125
struct SupTriangle {
    V3 a, b, c;
    V3 operator()(V3 d) const {
        if (dot(d, a) > dot(d, b)) {
            if (dot(d, a) > dot(d, c)) return a;
            return c;
        }
        if (dot(d, b) > dot(d, c)) return b;
        return c;
    }
};

bool gjkEpaConvexTriangle(V3 pos, Q ori, const ConvexMesh* m, V3 scale, V3 a, V3 b, V3 c, V3 startDir, V3& normal,
                          V3& cpBody, V3& cpTri) {
    SupConvex s0{pos, mat3_cast(ori), m, scale};
    SupTriangle s1{a, b, c};
    GjkSimplex s;
    if (!gjk(s0, s1, startDir, s)) return false;
    normal = epa(s0, s1, s, cpBody, cpTri);
    return true;
}

// ---------------------------------------------------------------- contact generation
// src/ContactsUtil.cpp:109-204 (epsilon 1e-4f)
void contactsPolygonPolygonFace(V3 refPos, const M3& refToWorld, V3 refPlaneOrigin, V3 refPlaneNormal,
                                       V3 incPlaneOrigin, V3 incPlaneNormal, const std::vector<V2>& polygon,
                                       int clipX, int clipY, ContactPoint* cps, int& numPoints) {
    std::vector<V3> pen;
    float distanceToRefPlane = dot(refPlaneOrigin, refPlaneNormal);
    for (V2 point : polygon) {
        V3 p3 = v3(point.y, 0, point.x);
        V3 dir = v3(0, 1, 0);
        float distance = dot(incPlaneOrigin - p3, incPlaneNormal) / dot(dir, incPlaneNormal);
        if (distance < distanceToRefPlane + 1e-4f) {
            p3.y = distance;
            pen.push_back(p3);
        }
    }
    auto emit = [&](int slot, int idx) {
        V3 on = pen[idx];
        on.y = distanceToRefPlane;
        cps[slot] = {refPos + mulmv(refToWorld, on), refPos + mulmv(refToWorld, pen[idx])};
    };
    int n = (int)pen.size();
    if (n > 4) {
        emit(0, 0);
        float maxDistance = 0; int maxDistanceIndex = 0;
        for (int i = 0; i < n; ++i) {
            float d = distance2(pen[0], pen[i]);
            if (d > maxDistance) { maxDistance = d; maxDistanceIndex = i; }
        }
        emit(1, maxDistanceIndex);
        float maxArea = 0; int maxAreaIndex = 0;
        for (int i = 0; i < n; ++i) {
            V3 ca = pen[0] - pen[i], cb = pen[maxDistanceIndex] - pen[i];
            float area = det2({ca[clipX], ca[clipY]}, {cb[clipX], cb[clipY]});
            if (area > maxArea) { maxArea = area; maxAreaIndex = i; }
        }
        emit(2, maxAreaIndex);
        float minArea = 0; int minAreaIndex = 0;
        for (int i = 0; i < n; ++i) {
            V3 da = pen[0] - pen[i], db = pen[maxDistanceIndex] - pen[i];
            float area = det2({da[clipX], da[clipY]}, {db[clipX], db[clipY]});
            if (area < minArea) { minArea = area; minAreaIndex = i; }
        }
        emit(3, minAreaIndex);
        numPoints = 4;
    } else {
        for (int i = 0; i < n; ++i) emit(i, i);
        numPoints = n;
    }
}

static V3 worldFaceNormal(Q q, const ConvexFace& f, V3 scale) { return normalize(qrot(q, f.normal / scale)); }

// src/Collision.cpp:352-410
static void convexConvexContacts(V3 pos0, Q or0, const ConvexMesh* m0, V3 s0, V3 pos1, Q or1,
                                 const ConvexMesh* m1, V3 s1, V3 normal, ContactPoint* cps, int& numPoints) {
    int f0 = 0;
    float maxDot = 0;
    for (int i = 0; i < (int)m0->faces.size(); ++i) {
        float d = dot(normal, worldFaceNormal(or0, m0->faces[i], s0));
        if (d > maxDot) { maxDot = d; f0 = i; }
    }
    const ConvexFace& F0 = m0->faces[f0];
    int f1 = 0;
    float minDot = 0;
    for (int i = 0; i < (int)m1->faces.size(); ++i) {
        float d = dot(normal, worldFaceNormal(or1, m1->faces[i], s1));
        if (d < minDot) { minDot = d; f1 = i; }
    }
    const ConvexFace& F1 = m1->faces[f1];
    M3 c0ToWorld = mat3_cast(or0);
    M3 worldToC0 = transpose(c0ToWorld);
    V3 refPlaneOrigin = s0 * F0.centroid;
    V3 u0 = normalize(s0 * m0->verts[F0.idx[0]] - refPlaneOrigin);
    V3 u1 = normalize(F0.normal / s0);
    V3 u2 = cross(u0, u1);
    M3 b; b[0] = u0; b[1] = u1; b[2] = u2;
    M3 c0ToRef = transpose(b);
    M3 worldToRef = mulmm(c0ToRef, worldToC0);
    const int clipX = 2, clipY = 0;
    std::vector<V2> clip(F0.idx.size()), polygon(F1.idx.size());
    for (size_t i = 0; i < F0.idx.size(); ++i) {
        V3 vx = mulmv(c0ToRef, s0 * m0->verts[F0.idx[i]]);
        clip[i] = {vx[clipX], vx[clipY]};
    }
    for (size_t i = 0; i < F1.idx.size(); ++i) {
        V3 vx = mulmv(worldToRef, pos1 + qrot(or1, s1 * m1->verts[F1.idx[i]]) - pos0);
        polygon[i] = {vx[clipX], vx[clipY]};
    }
    suthHodgClip(polygon, clip);
    V3 incPlaneOrigin = mulmv(worldToRef, pos1 + qrot(or1, s1 * F1.centroid) - pos0);
    V3 incPlaneNormal = mulmv(worldToRef, qrot(or1, normalize(F1.normal / s1)));
    M3 refToWorld = transpose(worldToRef);
    contactsPolygonPolygonFace(pos0, refToWorld, refPlaneOrigin, u1, incPlaneOrigin, incPlaneNormal, polygon, clipX,
                               clipY, cps, numPoints);
}

// src/Collision.cpp:412-426
bool collisionConvexConvex(V3 pos0, Q or0, const ConvexMesh* m0, V3 s0, V3 pos1, Q or1, const ConvexMesh* m1,
                           V3 s1, Manifold& res) {
    SupConvex a{pos0, mat3_cast(or0), m0, s0};
    SupConvex b{pos1, mat3_cast(or1), m1, s1};
    GjkSimplex s;
    if (!gjk(a, b, pos1 - pos0, s)) return false;
    res.normal = epa(a, b, s, res.points[0].p0, res.points[0].p1);
    convexConvexContacts(pos0, or0, m0, s0, pos1, or1, m1, s1, res.normal, res.points, res.numPoints);
    if (!res.numPoints) res.numPoints = 1;
    return true;
}

// The GJK-decided overlaps of src/Overlap.cpp:141-200. The box-convex case builds the
// convex support function at the box's position (pos0 in both, Overlap.cpp:200), which is
// kept.
bool overlapSphereConvex(V3 pos0, float r0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1) {
    GjkSimplex s;
    return gjk(SupSphere{pos0, r0}, SupConvex{pos1, mat3_cast(or1), m1, s1}, pos1 - pos0, s);
}
bool overlapCapsuleConvex(V3 pos0, Q or0, float hh0, float r0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1) {
    GjkSimplex s;
    return gjk(SupCapsule{pos0, qrot(or0, v3(0, 1, 0)), hh0, r0}, SupConvex{pos1, mat3_cast(or1), m1, s1},
               pos1 - pos0, s);
}
bool overlapBoxConvex(V3 pos0, Q or0, V3 he0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1) {
    GjkSimplex s;
    return gjk(SupBox{pos0, mat3_cast(or0), he0}, SupConvex{pos0, mat3_cast(or1), m1, s1}, pos1 - pos0, s);
}
bool overlapConvexConvex(V3 pos0, Q or0, const ConvexMesh* m0, V3 s0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1) {
    GjkSimplex s;
    return gjk(SupConvex{pos0, mat3_cast(or0), m0, s0}, SupConvex{pos1, mat3_cast(or1), m1, s1}, pos1 - pos0, s);
}

// src/Collision.cpp:488-499
bool collisionSphereConvex(V3 pos0, float r0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1, Manifold& res) {
    SupSphere a{pos0, r0};
    SupConvex b{pos1, mat3_cast(or1), m1, s1};
    GjkSimplex s;
    if (!gjk(a, b, pos1 - pos0, s)) return false;
    res.normal = epa(a, b, s, res.points[0].p0, res.points[0].p1);
    res.numPoints = 1;
    return true;
}

// src/Collision.cpp:694-785
static void capsuleConvexContacts(V3 p0L, V3 p1L, float radius, V3 meshPos, const M3& convexToWorld,
                                  const ConvexMesh* m, V3 scale, Manifold& res) {
    int fi = 0;
    float minDot = 0;
    for (int i = 0; i < (int)m->faces.size(); ++i) {
        float d = dot(res.normal, normalize(mulmv(convexToWorld, m->faces[i].normal / scale)));
        if (d < minDot) { minDot = d; fi = i; }
    }
    const ConvexFace& F = m->faces[fi];
    V3 refPlaneOrigin = scale * F.centroid;
    V3 u0 = normalize(scale * m->verts[F.idx[0]] - refPlaneOrigin);
    V3 u1 = normalize(F.normal / scale);
    V3 u2 = cross(u0, u1);
    M3 b; b[0] = u0; b[1] = u1; b[2] = u2;
    M3 convexToRef = transpose(b);
    const int faceAxis = 1, clipX = 2, clipY = 0;
    std::vector<V2> clip(F.idx.size());
    for (size_t i = 0; i < F.idx.size(); ++i) {
        V3 vx = mulmv(convexToRef, scale * m->verts[F.idx[i]]);
        clip[i] = {vx[clipX], vx[clipY]};
    }
    V3 p0Ref = mulmv(convexToRef, p0L);
    V3 p1Ref = mulmv(convexToRef, p1L);
    V2 line0 = {p0Ref[clipX], p0Ref[clipY]};
    V2 line1 = {p1Ref[clipX], p1Ref[clipY]};
    if (clipLine(line0, line1, clip.data(), (int)clip.size())) {
        M3 refToWorld = mulmm(convexToWorld, transpose(convexToRef));
        float distToPlane = dot(refPlaneOrigin, u1);
        V3 p0 = v3s(0), p1 = v3s(0);   // only faceAxis is read before it is written
        float distX = p0Ref[clipX] - p1Ref[clipX];
        float distY = p0Ref[clipY] - p1Ref[clipY];
        float absDistX = gabs(distX), absDistY = gabs(distY);
        if (absDistX && absDistX >= absDistY) {
            p0[faceAxis] = gmix(p0Ref[faceAxis], p1Ref[faceAxis], (p0Ref[clipX] - line0.x) / distX);
            p1[faceAxis] = gmix(p1Ref[faceAxis], p0Ref[faceAxis], (p1Ref[clipX] - line1.x) / -distX);
        } else if (absDistY > absDistX) {
            p0[faceAxis] = gmix(p0Ref[faceAxis], p1Ref[faceAxis], (p0Ref[clipY] - line0.y) / distY);
            p1[faceAxis] = gmix(p1Ref[faceAxis], p0Ref[faceAxis], (p1Ref[clipY] - line1.y) / -distY);
        } else {
            p0[faceAxis] = p0Ref[faceAxis];
            p1[faceAxis] = p1Ref[faceAxis];
        }
        p0[faceAxis] -= radius;
        p1[faceAxis] -= radius;
        int n = 0;
        if (p0[faceAxis] < distToPlane) {
            p0[clipX] = line0.x;
            p0[clipY] = line0.y;
            V3 on = p0;
            on[faceAxis] = distToPlane;
            res.points[n++] = {meshPos + mulmv(refToWorld, p0), meshPos + mulmv(refToWorld, on)};
        }
        if (p1[faceAxis] < distToPlane) {
            p1[clipX] = line1.x;
            p1[clipY] = line1.y;
            V3 on = p1;
            on[faceAxis] = distToPlane;
            res.points[n++] = {meshPos + mulmv(refToWorld, p1), meshPos + mulmv(refToWorld, on)};
        }
        res.numPoints = n;
    } else {
        res.numPoints = 0;
    }
}

// src/Collision.cpp:787-806
bool collisionCapsuleConvex(V3 pos0, Q or0, float hh0, float r0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1,
                            Manifold& res) {
    M3 convexToWorld = mat3_cast(or1);
    V3 axis = qrot(or0, v3(0, 1, 0));
    SupCapsule a{pos0, axis, hh0, r0};
    SupConvex b{pos1, convexToWorld, m1, s1};
    GjkSimplex s;
    if (!gjk(a, b, pos1 - pos0, s)) return false;
    res.normal = epa(a, b, s, res.points[0].p0, res.points[0].p1);
    M3 worldToConvex = inverse(convexToWorld);
    V3 p = mulmv(worldToConvex, pos0 - pos1);
    V3 axisLc = mulmv(worldToConvex, axis);
    capsuleConvexContacts(p + axisLc * hh0, p - axisLc * hh0, r0, pos1, convexToWorld, m1, s1, res);
    if (!res.numPoints) res.numPoints = 1;
    return true;
}

// src/Collision.cpp:808-871
static void boxConvexContacts(V3 boxCenter, const M3& boxBasis, V3 he, V3 convexPos, Q convexOr,
                              const ConvexMesh* m, V3 scale, V3 normal, ContactPoint* cps, int& numPoints) {
    int boxAxis = 0;
    float boxAxisSign = 0;
    float maxDot = 0;
    for (int i = 0; i < 3; ++i) {
        float d = dot(normal, boxBasis[i]);
        float ad = gabs(d);
        if (ad > maxDot) { maxDot = ad; boxAxis = i; boxAxisSign = d < 0 ? -1.f : 1.f; }
    }
    int fi = 0;
    float minDot = 0;
    for (int i = 0; i < (int)m->faces.size(); ++i) {
        float d = dot(normal, worldFaceNormal(convexOr, m->faces[i], scale));
        if (d < minDot) { minDot = d; fi = i; }
    }
    const ConvexFace& F = m->faces[fi];
    int clipX = (boxAxis + 1) % 3, clipY = (boxAxis + 2) % 3;
    M3 worldToBox = transpose(boxBasis);
    std::vector<V2> polygon(F.idx.size());
    for (size_t i = 0; i < F.idx.size(); ++i) {
        V3 vx = mulmv(worldToBox, convexPos + qrot(convexOr, scale * m->verts[F.idx[i]]) - boxCenter);
        polygon[i] = {vx[clipX], vx[clipY]};
    }
    std::vector<V2> clip = {{he[clipX], he[clipY]}, {he[clipX], -he[clipY]},
                            {-he[clipX], -he[clipY]}, {-he[clipX], he[clipY]}};
    suthHodgClip(polygon, clip);
    V3 incPlaneOrig = mulmv(worldToBox, convexPos + qrot(convexOr, scale * F.centroid) - boxCenter);
    V3 incPlaneNormal = normalize(mulmv(worldToBox, qrot(convexOr, F.normal / scale)));
    generateContactsPolygonBoxFace(boxCenter, boxBasis, boxAxis, (int)boxAxisSign, he, incPlaneOrig, incPlaneNormal,
                                   polygon, clipX, clipY, cps, numPoints);
}

// src/Collision.cpp:873-887
bool collisionBoxConvex(V3 pos0, Q or0, V3 he0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1, Manifold& res) {
    M3 boxBasis = mat3_cast(or0);
    SupBox a{pos0, boxBasis, he0};
    SupConvex b{pos1, mat3_cast(or1), m1, s1};
    GjkSimplex s;
    if (!gjk(a, b, pos1 - pos0, s)) return false;
    res.normal = epa(a, b, s, res.points[0].p0, res.points[0].p1);
    boxConvexContacts(pos0, boxBasis, he0, pos1, or1, m1, s1, res.normal, res.points, res.numPoints);
    if (!res.numPoints) res.numPoints = 1;
    return true;
}

}  // namespace orc
