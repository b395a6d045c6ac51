// Triangle-mesh narrowphase on the device: physecs::collisionTriangleMesh
// (src/CollisionTriangleMesh.cpp:882-956) for one (body collider, static mesh) pair per
// lane -- the body's mesh-local AABB queries the mesh BVH depth first (left child first,
// src/TriangleMesh.cpp:163-192), every overlapping triangle is tested
// (sphere/capsule/box closed forms, GJK/EPA for convex meshes, :44-121,221-396,616-697),
// face contacts are emitted in reverse order with swap-removal, the rest are ordered with
// std::sort's own introsort and emitted unless their feature's vertices were already
// covered (:906-953). The triangle distance routines follow src/GeomUtil.cpp:61-127,469-665.
//
// Three passes: each pair queries the BVH and reserves a block of candidate triangles;
// every (pair, candidate) is tested by a lane of its own (contact records stored flat,
// one per candidate); each pair then emits its manifolds from its records. Per-lane state
// of the query and emission passes (candidate triangles, the contact order, covered
// vertices) lives in a lane-interleaved workspace (word w of lane l at w * lanes + l). A
// pair's manifolds go to a block of slots reserved with one atomic add, in emission
// order. Included by narrowphase.hip after np_convex.h.
#pragma once
#include "introsort.h"

namespace pgx {

struct TriMeshView {
    const float4* verts;     // mesh-local vertices
    const int4* tris;        // mesh-local vertex indices, post-build order
    const float4* triN;
    const float4* triC;
    const float4* nodes;     // 2 per node: {min, triCount bits}, {max, index bits}
    const int4* mesh;        // {first vertex, vertices, first triangle, first node}
};

enum { TF_FACE = 0, TF_EDGE = 1, TF_VERTEX = 2 };
enum { BF_FACE = 0, BF_EDGE = 1, BF_UNKNOWN = 2 };

struct TCI {                 // TriangleContactInfo (:22-36)
    int tri;
    f3 normal, cpBody, cpTri;
    float distance;
    int feature, featureIndex, boxFeature, boxAxis;
};

// per-lane workspace of the query and emission passes
struct MeshWs {
    int32_t* w;              // lane's first word
    size_t L;
    __device__ int& cand(int i) { return w[(size_t)i * L]; }
    __device__ int& order(int i) { return w[(size_t)(TM_MAXCAND + i) * L]; }
    __device__ int& voided(int i) { return w[(size_t)(2 * TM_MAXCAND + i) * L]; }
};

// contact records, one per candidate triangle: word f of record k at f * cap + k (word
// TM_CI_WORDS: 1 if the triangle test hit)
struct MeshRecs {
    int32_t* w;
    size_t cap;
    __device__ int& rec(int k, int f) const { return w[(size_t)f * cap + k]; }
    __device__ bool hit(int k) const { return rec(k, TM_CI_WORDS) != 0; }
    __device__ float dist(int k) const { return __int_as_float(rec(k, 10)); }
    __device__ int feature(int k) const { return rec(k, 11); }
    __device__ void store(int k, bool h, const TCI& t) const {
        rec(k, TM_CI_WORDS) = h ? 1 : 0;
        if (!h) return;
        rec(k, 0) = t.tri;
        rec(k, 1) = __float_as_int(t.normal.x); rec(k, 2) = __float_as_int(t.normal.y); rec(k, 3) = __float_as_int(t.normal.z);
        rec(k, 4) = __float_as_int(t.cpBody.x); rec(k, 5) = __float_as_int(t.cpBody.y); rec(k, 6) = __float_as_int(t.cpBody.z);
        rec(k, 7) = __float_as_int(t.cpTri.x); rec(k, 8) = __float_as_int(t.cpTri.y); rec(k, 9) = __float_as_int(t.cpTri.z);
        rec(k, 10) = __float_as_int(t.distance);
        rec(k, 11) = t.feature; rec(k, 12) = t.featureIndex; rec(k, 13) = t.boxFeature | (t.boxAxis << 8);
    }
    __device__ TCI load(int k) const {
        TCI t;
        t.tri = rec(k, 0);
        t.normal = mk3(__int_as_float(rec(k, 1)), __int_as_float(rec(k, 2)), __int_as_float(rec(k, 3)));
        t.cpBody = mk3(__int_as_float(rec(k, 4)), __int_as_float(rec(k, 5)), __int_as_float(rec(k, 6)));
        t.cpTri = mk3(__int_as_float(rec(k, 7)), __int_as_float(rec(k, 8)), __int_as_float(rec(k, 9)));
        t.distance = __int_as_float(rec(k, 10));
        t.feature = rec(k, 11); t.featureIndex = rec(k, 12);
        t.boxFeature = rec(k, 13) & 255; t.boxAxis = rec(k, 13) >> 8;
        return t;
    }
};

// index array over the contact order (record indices), compared by distance
// (TriangleContactInfo::operator<)
struct OrderArr {
    MeshWs* ws;
    const MeshRecs* recs;
    __device__ int get(int i) { return ws->order(i); }
    __device__ void set(int i, int v) { ws->order(i) = v; }
    __device__ bool less(int a, int b) { return recs->dist(a) < recs->dist(b); }
};

// ------------------------------------------------------------------ GeomUtil (triangles)
__device__ void closest_points_segments4(f3 o0, f3 v0, f3 o1, f3 v1, f3& p0, f3& p1) {   // :61-127
    float v0v0 = dot(v0, v0);
    if (v0v0 < 0.0001f) {
        float v1v1 = dot(v1, v1);
        if (v1v1 < 0.0001f) { p0 = o0; p1 = o1; return; }
        float t1 = -dot(sub(o1, o0), v1) / v1v1;
        p0 = o0; p1 = add(o1, mul(t1, v1));
        return;
    }
    float v1v1 = dot(v1, v1);
    if (v1v1 < 0.0001f) {
        float t0 = -dot(sub(o0, o1), v0) / v0v0;
        p0 = add(o0, mul(t0, v0)); p1 = o1;
        return;
    }
    float v0v1 = dot(v0, v1);
    f3 r = sub(o1, o0);
    float rv0 = dot(r, v0), rv1 = dot(r, v1);
    float denom = v0v0 * v1v1 - v0v1 * v0v1;
    float t1 = gabs(denom) < 0.0001f ? 0.f : (rv0 * v0v1 - rv1 * v0v0) / denom;
    float t0 = (rv0 + t1 * v0v1) / v0v0;
    if (t0 < 0) { t0 = 0; p0 = add(o0, mul(t0, v0)); t1 = -dot(sub(o1, p0), v1) / v1v1; }
    else if (t0 > 1) { t0 = 1; p0 = add(o0, mul(t0, v0)); t1 = -dot(sub(o1, p0), v1) / v1v1; }
    else p0 = add(o0, mul(t0, v0));
    if (t1 < 0) {
        t1 = 0; p1 = add(o1, mul(t1, v1));
        t0 = gclamp(-dot(sub(o0, p1), v0) / v0v0, 0.f, 1.f);
        p0 = add(o0, mul(t0, v0));
    } else if (t1 > 1) {
        t1 = 1; p1 = add(o1, mul(t1, v1));
        t0 = gclamp(-dot(sub(o0, p1), v0) / v0v0, 0.f, 1.f);
        p0 = add(o0, mul(t0, v0));
    } else {
        p1 = add(o1, mul(t1, v1));
    }
}

__device__ __forceinline__ float dist2(f3 a, f3 b) { f3 d = sub(b, a); return dot(d, d); }

__device__ float sqr_dist_point_tri(f3 p, f3 a, f3 b, f3 c, f3& q, int& feature, int& fi) {   // :469-538
    f3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) { feature = TF_VERTEX; fi = 0; q = a; return dist2(p, q); }
    f3 bp = sub(p, b);
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) { feature = TF_VERTEX; fi = 1; q = b; return dist2(p, q); }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0.f && d3 <= 0.f) {
        feature = TF_EDGE; fi = 2;
        q = add(a, mul(d1 / (d1 - d3), ab));
        return dist2(p, q);
    }
    f3 cp = sub(p, c);
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.f && d5 <= d6) { feature = TF_VERTEX; fi = 2; q = c; return dist2(p, q); }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0 && d4 >= d3 && d5 >= d6) {
        feature = TF_EDGE; fi = 0;
        q = add(b, mul((d4 - d3) / (d4 - d3 + d5 - d6), sub(c, b)));
        return dist2(p, q);
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0.f && d6 <= 0.f) {
        feature = TF_EDGE; fi = 1;
        q = add(c, mul(d2 / (d2 - d6), ac));
        return dist2(p, q);
    }
    feature = TF_FACE;
    float denom = va + vb + vc;
    float u = va / denom, v = vb / denom, w = 1 - u - v;
    q = add(add(mul(u, a), mul(v, b)), mul(w, c));
    return dist2(p, q);
}

__device__ float sqr_dist_line_segment(f3 p, f3 dir, f3 a, f3 b, float& t, f3& q, int& vflag) {   // :540-581
    f3 ab = sub(b, a);
    float v1v2 = dot(dir, ab), v2v2 = dot(ab, ab);
    f3 r = sub(p, a);
    float rv1 = dot(r, dir), rv2 = dot(r, ab);
    float denom = (v1v2 * v1v2 - v2v2);
    if (gabs(denom) <= 1e-6f) {
        q = a; t = -rv1; vflag = 0;
        f3 d = sub(add(p, mul(t, dir)), a);
        return dot(d, d);
    }
    float s = (rv1 * v1v2 - rv2) / denom;
    if (s <= 0.f) {
        q = a; t = -rv1; vflag = 1;
        f3 d = sub(add(p, mul(t, dir)), a);
        return dot(d, d);
    }
    if (s >= 1.f) {
        q = b; t = -dot(sub(p, b), dir); vflag = 2;
        f3 d = sub(add(p, mul(t, dir)), b);
        return dot(d, d);
    }
    q = add(a, mul(s, ab));
    t = s * v1v2 - rv1;
    vflag = 0;
    f3 d = sub(add(p, mul(t, dir)), q);
    return dot(d, d);
}

__device__ __forceinline__ float det3(f3 c0, f3 c1, f3 c2) {   // glm::determinant(mat3)
    return +c0.x * (c1.y * c2.z - c2.y * c1.z) - c1.x * (c0.y * c2.z - c2.y * c0.z) + c2.x * (c0.y * c1.z - c1.y * c0.z);
}

__device__ float sqr_dist_line_tri(f3 p, f3 dir, f3 a, f3 b, f3 c, float& t, f3& q, int& feature, int& fi) {   // :583-647
    f3 e0 = sub(b, a), e1 = sub(c, a), nd = neg(dir);
    float det = det3(e0, e1, nd);
    if (gabs(det) > 1e-6f) {
        f3 diff = sub(p, a);
        float invDet = 1 / det;
        float u = invDet * det3(diff, e1, nd);
        float v = invDet * det3(e0, diff, nd);
        float w = 1 - u - v;
        if (u >= 0.f && v >= 0.f && w >= 0.f) {
            t = invDet * det3(e0, e1, diff);
            q = add(add(a, mul(u, e0)), mul(v, e1));
            feature = TF_FACE;
            return 0;
        }
    }
    float t1;
    f3 q1;
    int vf, vf1;
    fi = 2;
    float dist = sqr_dist_line_segment(p, dir, a, b, t, q, vf);
    float dist1 = sqr_dist_line_segment(p, dir, b, c, t1, q1, vf1);
    if (dist1 < dist) { fi = 0; t = t1; q = q1; vf = vf1; dist = dist1; }
    dist1 = sqr_dist_line_segment(p, dir, c, a, t1, q1, vf1);
    if (dist1 < dist) { fi = 1; t = t1; q = q1; vf = vf1; dist = dist1; }
    if (vf) { feature = TF_VERTEX; fi = (fi + vf) % 3; }
    else feature = TF_EDGE;
    return dist;
}

__device__ float sqr_dist_segment_tri(f3 p, f3 dir, float mn, float mx, f3 a, f3 b, f3 c, float& t, f3& q,
                                      int& feature, int& fi) {   // :649-660
    float sq = sqr_dist_line_tri(p, dir, a, b, c, t, q, feature, fi);
    if (t < mn) { t = mn; sq = sqr_dist_point_tri(add(p, mul(dir, t)), a, b, c, q, feature, fi); }
    else if (t > mx) { t = mx; sq = sqr_dist_point_tri(add(p, mul(dir, t)), a, b, c, q, feature, fi); }
    return sq;
}

// ------------------------------------------------------------------ per-triangle tests
__device__ bool sphere_tri(f3 pos, float r, f3 a, f3 b, f3 c, f3 n, TCI& ci) {   // :44-71
    f3 u;
    int f, fi = 0;
    float sq = sqr_dist_point_tri(pos, a, b, c, u, f, fi);
    if (sq > r * r) return false;
    if (f == TF_FACE) ci.normal = neg(n);
    else {
        f3 nn = sub(u, pos);
        float len = length(nn);
        if (len) ci.normal = divs(nn, len);
        else ci.normal = neg(n);
    }
    ci.cpBody = add(pos, mul(ci.normal, r));
    ci.cpTri = u;
    ci.feature = f; ci.featureIndex = fi; ci.distance = sq;
    return true;
}

__device__ bool capsule_tri(f3 pos, qf ori, float hh, float r, f3 a, f3 b, f3 c, f3 n, TCI& ci) {   // :85-121
    f3 dir = qrot(ori, mk3(0.f, 1.f, 0.f));
    float t;
    f3 u;
    int f, fi = 0;
    float sq = sqr_dist_segment_tri(pos, dir, -hh, hh, a, b, c, t, u, f, fi);
    if (sq > r * r) return false;
    f3 onSeg = add(pos, mul(dir, t));
    if (f == TF_FACE) ci.normal = neg(n);
    else {
        f3 nn = sub(u, onSeg);
        float len = length(nn);
        if (len) ci.normal = divs(nn, len);
        else ci.normal = neg(n);
    }
    ci.cpBody = add(onSeg, mul(ci.normal, r));
    ci.cpTri = u;
    ci.feature = f; ci.featureIndex = fi; ci.distance = sq;
    return true;
}

__device__ bool box_tri(f3 pos, qf ori, f3 he, f3 a, f3 b, f3 c, f3 n, TCI& ci) {   // :221-396
    float edgeOffset = 0.f;
    const float edgeLimit = 0.999f;
    m3 invU = transpose(mat3_cast(ori));
    f3 v0 = mulmv(invU, sub(a, pos)), v1 = mulmv(invU, sub(b, pos)), v2 = mulmv(invU, sub(c, pos));
    f3 fe[3] = {sub(v1, v0), sub(v2, v1), sub(v0, v2)};
    float lim[3] = {edgeLimit * dot(fe[0], fe[0]), edgeLimit * dot(fe[1], fe[1]), edgeLimit * dot(fe[2], fe[2])};
    float mx = -3.402823466e+38f;
    for (int k = 0; k < 3; ++k)
        for (int e = 0; e < 3; ++e) {
            f3 f = fe[e];
            f3 A = e == 2 ? v1 : v0;               // the two vertices off the edge's line
            f3 B = e == 1 ? v1 : v2;
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
            float fk = at(f, k);
            if (fk * fk < lim[e] && d > mx) {
                mx = d;
                ci.feature = TF_EDGE; ci.featureIndex = e == 0 ? 2 : (e == 1 ? 0 : 1); ci.distance = d;
                ci.boxFeature = BF_EDGE; ci.boxAxis = k;
                edgeOffset = 0.1f;
            }
        }
    for (int i = 0; i < 3; ++i) {
        float x0 = at(v0, i), x1 = at(v1, i), x2 = at(v2, i), hi = at(he, i);
        float d = gmax(-gmax(gmax(x0, x1), x2), gmin(gmin(x0, x1), x2)) - hi;
        if (d > 0) return false;
        if (d > mx - edgeOffset) {
            mx = d;
            bool i0 = gabs(x0) < hi, i1 = gabs(x1) < hi, i2 = gabs(x2) < hi;
            if (i0) {
                if (i1) {
                    if (i2) ci.feature = TF_FACE;
                    else { ci.feature = TF_EDGE; ci.featureIndex = 2; }
                } else {
                    if (i2) { ci.feature = TF_EDGE; ci.featureIndex = 1; }
                    else { ci.feature = TF_VERTEX; ci.featureIndex = 0; }
                }
            } else {
                if (i1) {
                    if (i2) { ci.feature = TF_EDGE; ci.featureIndex = 0; }
                    else { ci.feature = TF_VERTEX; ci.featureIndex = 1; }
                } else {
                    ci.feature = TF_VERTEX; ci.featureIndex = 2;
                }
            }
            ci.boxFeature = BF_FACE; ci.boxAxis = i;
            ci.distance = d;
            edgeOffset = 0.f;
        }
    }
    f3 nl = mulmv(invU, n);
    float r = he.x * gabs(nl.x) + he.y * gabs(nl.y) + he.z * gabs(nl.z);   // distanceAABBPlane
    float d = gabs(dot(nl, v0)) - r;
    if (d > 0) return false;
    if (d > mx - edgeOffset) {
        ci.feature = TF_FACE;
        ci.distance = d;
        ci.boxFeature = BF_UNKNOWN;
    }
    return true;
}

// TriangleSupportFunction, src/GJK.h:109-125
struct SupTriangle {
    f3 a, b, c;
    __device__ f3 operator()(f3 d) const {
        if (dot(d, a) > dot(d, b)) {
            if (dot(d, a) > dot(d, c)) return a;
            return c;
        }
        if (dot(d, b) > dot(d, c)) return b;
        return c;
    }
};

__device__ bool convex_tri(const SupConvex& sc, f3 a, f3 b, f3 c, f3 centroid, EpaWs& ews, TCI& ci) {   // :616-697
    SupTriangle st{a, b, c};
    GjkV s[4];
    if (!gjk(sc, st, sub(centroid, sc.pos), s)) return false;
    ci.normal = epa(sc, st, s, ews, ci.cpBody, ci.cpTri);
    ci.distance = dot(sub(ci.cpTri, ci.cpBody), ci.normal);
    f3 v0 = sub(b, a), v1 = sub(c, a), v2 = sub(ci.cpTri, a);
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
        else { v = 0; u = 1.0f; }
    }
    const float eps = 1e-4f;
    if (u < eps) {
        if (v < eps) { ci.feature = TF_VERTEX; ci.featureIndex = 2; }
        else if (w < eps) { ci.feature = TF_VERTEX; ci.featureIndex = 1; }
        else { ci.feature = TF_EDGE; ci.featureIndex = 0; }
    } else {
        if (v < eps) {
            if (w < eps) { ci.feature = TF_VERTEX; ci.featureIndex = 0; }
            else { ci.feature = TF_EDGE; ci.featureIndex = 1; }
        } else if (w < eps) {
            if (u < eps) { ci.feature = TF_VERTEX; ci.featureIndex = 0; }
            else { ci.feature = TF_EDGE; ci.featureIndex = 2; }
        } else {
            ci.feature = TF_FACE;
        }
    }
    return true;
}

// ------------------------------------------------------------------ contact generation
struct MeshCtx {
    const TriMeshView* tv;
    int4 mi;
    __device__ f3 vert(int local) const { return xyz(tv->verts[mi.x + local]); }
    __device__ int4 tri(int t) const { return tv->tris[mi.z + t]; }
    __device__ f3 n(int t) const { return xyz(tv->triN[mi.z + t]); }
    __device__ f3 c(int t) const { return xyz(tv->triC[mi.z + t]); }
};

// the triangle's face frame (u0 toward vertex 0, u1 = normal) and its clip triangle in
// (z, x), listed backwards (:132-146)
__device__ void tri_frame(const MeshCtx& mc, int t, bool normalizeN, f3& refOrig, f3& u1, m3& meshToRef, f2* clip) {
    int4 ix = mc.tri(t);
    refOrig = mc.c(t);
    f3 u0 = normalize(sub(mc.vert(ix.x), refOrig));
    u1 = normalizeN ? normalize(mc.n(t)) : mc.n(t);
    f3 u2 = cross(u0, u1);
    meshToRef = transpose(mk_m3(u0, u1, u2));
    int iv[3] = {ix.x, ix.y, ix.z};
    for (int i = 0; i < 3; i++) {
        f3 v = mulmv(meshToRef, mc.vert(iv[i]));
        clip[3 - i - 1] = mk2(v.z, v.x);
    }
}

__device__ void gen_sphere(f3 pos1, qf or1, const TCI& ci, Mf& m) {   // :73-81
    m.normal = qrot(or1, ci.normal);
    m.numPoints = 1;
    m.pts[0].p0 = add(pos1, qrot(or1, ci.cpBody));
    m.pts[0].p1 = add(pos1, qrot(or1, ci.cpTri));
}

__device__ bool gen_capsule_face(f3 pos0, qf or0, float hh0, float r0, f3 pos1, qf or1, const MeshCtx& mc,
                                 const TCI& ci, Mf& m) {   // :123-217; false when nothing is pushed
    f3 dir = qrot(or0, mk3(0.f, 1.f, 0.f));
    f3 p0l = sub(pos0, mul(dir, hh0)), p1l = add(pos0, mul(dir, hh0));
    f3 refOrig, u1;
    m3 meshToRef;
    f2 clip[3];
    tri_frame(mc, ci.tri, true, refOrig, u1, meshToRef, clip);
    f3 p0Ref = mulmv(meshToRef, p0l), p1Ref = mulmv(meshToRef, p1l);
    f2 line0 = mk2(p0Ref.z, p0Ref.x), line1 = mk2(p1Ref.z, p1Ref.x);
    if (!clip_line_n(line0, line1, clip, 3)) return false;
    m3 meshToWorld = mat3_cast(or1);
    m3 refToWorld = mulmm(meshToWorld, transpose(meshToRef));
    float distToPlane = dot(refOrig, u1);
    float y0, y1;
    float distX = p0Ref.z - p1Ref.z, distY = p0Ref.x - p1Ref.x;
    float adx = gabs(distX), ady = gabs(distY);
    if (adx && adx >= ady) {
        y0 = gmix(p0Ref.y, p1Ref.y, (p0Ref.z - line0.x) / distX);
        y1 = gmix(p1Ref.y, p0Ref.y, (p1Ref.z - line1.x) / -distX);
    } else if (ady > adx) {
        y0 = gmix(p0Ref.y, p1Ref.y, (p0Ref.x - line0.y) / distY);
        y1 = gmix(p1Ref.y, p0Ref.y, (p1Ref.x - line1.y) / -distY);
    } else {
        y0 = p0Ref.y;
        y1 = p1Ref.y;
    }
    y0 -= r0;
    y1 -= r0;
    int n = 0;
    m.normal = mulmv(meshToWorld, ci.normal);
    if (y0 < distToPlane) {
        f3 p = mk3(line0.y, y0, line0.x), on = mk3(line0.y, distToPlane, line0.x);
        m.pts[n].p0 = add(pos1, mulmv(refToWorld, p));
        m.pts[n].p1 = add(pos1, mulmv(refToWorld, on));
        ++n;
    }
    if (y1 < distToPlane) {
        f3 p = mk3(line1.y, y1, line1.x), on = mk3(line1.y, distToPlane, line1.x);
        m.pts[n].p0 = add(pos1, mulmv(refToWorld, p));
        m.pts[n].p1 = add(pos1, mulmv(refToWorld, on));
        ++n;
    }
    if (!n) {
        n = 1;
        m.pts[0].p0 = add(pos1, mulmv(meshToWorld, ci.cpBody));
        m.pts[0].p1 = add(pos1, mulmv(meshToWorld, ci.cpTri));
    }
    m.numPoints = n;
    return true;
}

__device__ void gen_box_tri_face(f3 boxPos, qf boxOr, f3 he, f3 meshPos, qf meshOr, const MeshCtx& mc,
                                 const TCI& ci, Mf& m) {   // :398-467
    m3 basis = mat3_cast(boxOr);
    f3 tn = mc.n(ci.tri);
    f3 refOrig, u1;
    m3 meshToRef;
    f2 clip[3];
    tri_frame(mc, ci.tri, true, refOrig, u1, meshToRef, clip);
    int incAxis = 0;
    float incSign = 0.f;
    if (ci.boxFeature == BF_FACE) {
        incAxis = ci.boxAxis;
        incSign = dot(tn, col(basis, incAxis)) > 0 ? -1.f : 1.f;
    } else {
        float maxDot = 0.f;
        for (int i = 0; i < 3; ++i) {
            float d = dot(tn, col(basis, i));
            float ad = gabs(d);
            if (ad > maxDot) { maxDot = ad; incAxis = i; incSign = d > 0 ? -1.f : 1.f; }
        }
    }
    int ia1 = (incAxis + 1) % 3, ia2 = (incAxis + 2) % 3;
    f3 ui = col(basis, incAxis), ua = col(basis, ia1), ub = col(basis, ia2);
    f3 incOrig = mulmv(meshToRef, add(boxPos, mul(mul(ui, at(he, incAxis)), incSign)));
    f3 A = mul(ua, at(he, ia1)), B = mul(ub, at(he, ia2)), nA = mul(neg(ua), at(he, ia1));
    f3 pp[4] = {add(incOrig, mulmv(meshToRef, add(A, B))), add(incOrig, mulmv(meshToRef, add(nA, B))),
                add(incOrig, mulmv(meshToRef, sub(nA, B))), add(incOrig, mulmv(meshToRef, sub(A, B)))};
    f2 poly[CVX_MAXPOLY];
    for (int i = 0; i < 4; ++i) poly[i] = mk2(pp[i].z, pp[i].x);
    int n = 4;
    clip_polygon_n<CVX_MAXPOLY>(poly, n, clip, 3);
    f3 incNormal = mulmv(meshToRef, ui);
    m3 meshToWorld = mat3_cast(meshOr);
    m3 refToWorld = mulmm(meshToWorld, transpose(meshToRef));
    m3 negM = mk_m3(neg(meshToWorld.c0), neg(meshToWorld.c1), neg(meshToWorld.c2));
    m.normal = mulmv(negM, tn);
    polygon_polygon_face_contacts(meshPos, refToWorld, refOrig, u1, incOrig, incNormal, poly, n, 2, 0, m.pts,
                                  m.numPoints);
    for (int i = 0; i < m.numPoints; ++i) { f3 t = m.pts[i].p0; m.pts[i].p0 = m.pts[i].p1; m.pts[i].p1 = t; }
}

__device__ void gen_box_face_tri(f3 boxPos, qf boxOr, f3 he, f3 meshPos, qf meshOr, const MeshCtx& mc,
                                 const TCI& ci, Mf& m) {   // :469-512
    m3 basis = mat3_cast(boxOr);
    m3 meshToBox = transpose(basis);
    int refAxis = ci.boxAxis;
    f3 tc = mc.c(ci.tri), tn = mc.n(ci.tri);
    float refSign = dot(col(basis, refAxis), sub(tc, boxPos)) < 0 ? -1.f : 1.f;
    int cx = (refAxis + 1) % 3, cy = (refAxis + 2) % 3;
    int4 ix = mc.tri(ci.tri);
    int iv[3] = {ix.x, ix.y, ix.z};
    f2 poly[CVX_MAXPOLY];
    for (int i = 0; i < 3; ++i) {
        f3 v = mulmv(meshToBox, sub(mc.vert(iv[i]), boxPos));
        poly[i] = mk2(at(v, cx), at(v, cy));
    }
    float hx = at(he, cx), hy = at(he, cy);
    f2 clip[4] = {mk2(hx, hy), mk2(hx, -hy), mk2(-hx, -hy), mk2(-hx, hy)};
    int n = 3;
    clip_polygon_n<CVX_MAXPOLY>(poly, n, clip, 4);
    f3 incOrig = mulmv(meshToBox, sub(tc, boxPos));
    f3 incNormal = mulmv(meshToBox, tn);
    m3 meshToWorld = mat3_cast(meshOr);
    m.normal = mulmv(meshToWorld, mul(refSign, col(basis, refAxis)));
    polygon_box_face_contacts_n<CVX_MAXPOLY>(boxPos, basis, refAxis, (int)refSign, he, incOrig, incNormal, poly, n,
                                             cx, cy, m.pts, m.numPoints);
    for (int i = 0; i < m.numPoints; ++i) {
        m.pts[i].p0 = add(meshPos, mulmv(meshToWorld, m.pts[i].p0));
        m.pts[i].p1 = add(meshPos, mulmv(meshToWorld, m.pts[i].p1));
    }
}

__device__ void gen_box_edge_tri_edge(f3 boxPos, qf boxOr, f3 he, f3 meshPos, qf meshOr, const MeshCtx& mc,
                                      const TCI& ci, Mf& m) {   // :514-542
    m3 basis = mat3_cast(boxOr);
    int boxAxis = ci.boxAxis;
    int e = ci.featureIndex;
    int4 ix = mc.tri(ci.tri);
    int iv[3] = {ix.x, ix.y, ix.z};
    f3 triA = mc.vert(iv[(e + 1) % 3]), triB = mc.vert(iv[(e + 2) % 3]);
    f3 triEdge = sub(triB, triA);
    f3 axis = cross(col(basis, boxAxis), triEdge);
    m3 meshToWorld = mat3_cast(meshOr);
    m.normal = mulmv(meshToWorld, normalize(dot(axis, sub(mc.c(ci.tri), boxPos)) < 0 ? neg(axis) : axis));
    m.numPoints = 1;
    float sb[3] = {-1.f, -1.f, -1.f};
    int a1 = (boxAxis + 1) % 3, a2 = (boxAxis + 2) % 3;
    sb[a1] = dot(col(basis, a1), m.normal) < 0 ? -1.f : 1.f;
    sb[a2] = dot(col(basis, a2), m.normal) < 0 ? -1.f : 1.f;
    f3 boxA = add(add(add(boxPos, mul(sb[0] * he.x, basis.c0)), mul(sb[1] * he.y, basis.c1)), mul(sb[2] * he.z, basis.c2));
    f3 boxEdge = mul(mul(col(basis, boxAxis), at(he, boxAxis)), 2.f);
    f3 p0, p1;
    closest_points_segments4(boxA, boxEdge, triA, triEdge, p0, p1);
    m.pts[0].p0 = add(meshPos, mulmv(meshToWorld, p0));
    m.pts[0].p1 = add(meshPos, mulmv(meshToWorld, p1));
}

__device__ void gen_convex_tri_face(f3 cpos, qf cor, const MeshRef& cm, f3 scale, f3 meshPos, qf meshOr,
                                    const MeshCtx& mc, const TCI& ci, Mf& m) {   // :699-757
    f3 tn = mc.n(ci.tri);
    f3 refOrig, u1;
    m3 meshToRef;
    f2 clip[3];
    tri_frame(mc, ci.tri, false, refOrig, u1, meshToRef, clip);
    int fi = 0;   // the reference leaves the index unset when no face has a negative dot
    float minDot = 0.f;
    for (int i = 0; i < cm.nfaces(); ++i) {
        float d = dot(tn, normalize(qrot(cor, divv(cm.fnormal(i), scale))));
        if (d < minDot) { minDot = d; fi = i; }
    }
    int2 F = cm.face(fi);
    f2 poly[CVX_MAXPOLY];
    int n = F.y < CVX_MAX_FACE_VERTS ? F.y : CVX_MAX_FACE_VERTS;
    for (int i = 0; i < n; ++i) {
        f3 v = mulmv(meshToRef, add(cpos, qrot(cor, mulv(scale, cm.face_vert(F, i)))));
        poly[i] = mk2(v.z, v.x);
    }
    clip_polygon_n<CVX_MAXPOLY>(poly, n, clip, 3);
    f3 incOrig = mulmv(meshToRef, add(cpos, qrot(cor, mulv(scale, cm.fcentroid(fi)))));
    f3 incNormal = mulmv(meshToRef, qrot(cor, normalize(divv(cm.fnormal(fi), scale))));
    m3 meshToWorld = mat3_cast(meshOr);
    m3 refToWorld = mulmm(meshToWorld, transpose(meshToRef));
    m3 negM = mk_m3(neg(meshToWorld.c0), neg(meshToWorld.c1), neg(meshToWorld.c2));
    m.normal = mulmv(negM, tn);
    polygon_polygon_face_contacts(meshPos, refToWorld, refOrig, u1, incOrig, incNormal, poly, n, 2, 0, m.pts,
                                  m.numPoints);
    for (int i = 0; i < m.numPoints; ++i) { f3 t = m.pts[i].p0; m.pts[i].p0 = m.pts[i].p1; m.pts[i].p1 = t; }
}

__device__ void gen_convex_face_tri(f3 cpos, qf cor, const MeshRef& cm, f3 scale, f3 meshPos, qf meshOr,
                                    const MeshCtx& mc, const TCI& ci, Mf& m) {   // :759-824
    m3 convexToMesh = mat3_cast(cor);
    m3 meshToConvex = transpose(convexToMesh);
    int fi = 0;   // unset in the reference when no face has a positive dot
    float maxDot = 0.f;
    for (int i = 0; i < cm.nfaces(); ++i) {
        float d = dot(ci.normal, normalize(qrot(cor, divv(cm.fnormal(i), scale))));
        if (d > maxDot) { maxDot = d; fi = i; }
    }
    int2 F = cm.face(fi);
    f3 refOrig = mulv(scale, cm.fcentroid(fi));
    f3 u0 = normalize(sub(mulv(scale, cm.face_vert(F, 0)), refOrig));
    f3 u1 = normalize(divv(cm.fnormal(fi), scale));
    f3 u2 = cross(u0, u1);
    m3 convexToRef = transpose(mk_m3(u0, u1, u2));
    m3 meshToRef = mulmm(convexToRef, meshToConvex);
    f2 clip[CVX_MAX_FACE_VERTS];
    int nc = F.y < CVX_MAX_FACE_VERTS ? F.y : CVX_MAX_FACE_VERTS;
    for (int i = 0; i < nc; ++i) {
        f3 v = mulmv(convexToRef, mulv(scale, cm.face_vert(F, i)));
        clip[i] = mk2(v.z, v.x);
    }
    int4 ix = mc.tri(ci.tri);
    int iv[3] = {ix.x, ix.y, ix.z};
    f2 poly[CVX_MAXPOLY];
    for (int i = 0; i < 3; ++i) {
        f3 v = mulmv(meshToRef, sub(mc.vert(iv[i]), cpos));
        poly[i] = mk2(v.z, v.x);
    }
    int n = 3;
    clip_polygon_n<CVX_MAXPOLY>(poly, n, clip, nc);
    f3 incOrig = mulmv(meshToRef, sub(mc.c(ci.tri), cpos));
    f3 incNormal = mulmv(meshToRef, mc.n(ci.tri));
    m3 meshToWorld = mat3_cast(meshOr);
    m3 refToMesh = transpose(meshToRef);
    polygon_polygon_face_contacts(cpos, refToMesh, refOrig, u1, incOrig, incNormal, poly, n, 2, 0, m.pts,
                                  m.numPoints);
    if (m.numPoints) {
        for (int i = 0; i < m.numPoints; ++i) {
            m.pts[i].p0 = add(meshPos, mulmv(meshToWorld, m.pts[i].p0));
            m.pts[i].p1 = add(meshPos, mulmv(meshToWorld, m.pts[i].p1));
        }
        m.normal = mulmv(mulmm(meshToWorld, convexToMesh), u1);
    } else {
        m.normal = mulmv(meshToWorld, ci.normal);
        m.numPoints = 1;
        m.pts[0].p0 = add(meshPos, mulmv(meshToWorld, ci.cpBody));
        m.pts[0].p1 = add(meshPos, mulmv(meshToWorld, ci.cpTri));
    }
}

// ------------------------------------------------------------------ one (body, mesh) pair
struct MeshOut {
    int32_t* npNum; float4* npNormal; float4* npPts; int32_t* npTri;
    int base;        // first reserved slot (pairCap + block offset for the data arrays)
    int triBase;     // same block in npTri
    int n;
    bool flip;
    __device__ void emit(Mf& m, int tri) {
        if (m.numPoints <= 0) return;                  // simulate skips empty manifolds (Physecs.cpp:233)
        if (flip) {
            m.normal = neg(m.normal);
            for (int i = 0; i < m.numPoints; ++i) { f3 t = m.pts[i].p0; m.pts[i].p0 = m.pts[i].p1; m.pts[i].p1 = t; }
        }
        size_t q = (size_t)base + n;
        npNum[q] = m.numPoints;
        npNormal[q] = f4(m.normal, 0.f);
        for (int k = 0; k < m.numPoints; ++k) {
            npPts[8 * q + 2 * k] = f4(m.pts[k].p0, 0.f);
            npPts[8 * q + 2 * k + 1] = f4(m.pts[k].p1, 0.f);
        }
        npTri[triBase + n] = tri;
        ++n;
    }
};

// The body collider (type t, params g, world pose pos0/or0) in the mesh's frame (mesh at
// pos1/or1), and for a convex mesh its support function there.
__device__ void mesh_local(int t, float4 g, f3 pos0, qf or0, f3 pos1, qf or1, const ConvexView& cv, f3& localPos,
                           qf& localOr, SupConvex& sc) {
    qf inv1 = qinverse(or1);
    localPos = qrot(inv1, sub(pos0, pos1));
    localOr = qmul(inv1, or0);
    if (t == PGX_CONVEX_MESH) {
        int4 ci = cv.mesh[(int)g.w];
        sc = SupConvex{localPos, mat3_cast(localOr), mk3(g.x, g.y, g.z), cv.verts + ci.x, ci.y};
    }
}

// The body's bounds in the mesh frame (getBounds(localPos, localOr, geom0),
// src/BoundsUtil.cpp:35-69) and the mesh's triangles whose BVH leaves they overlap, depth
// first, left child first (overlapBvh, src/TriangleMesh.cpp:163-192). Returns the number
// of candidates, -1 on a capacity error.
__device__ int mesh_query(int t, float4 g, f3 localPos, qf localOr, const ConvexView& cv, const MeshCtx& mc,
                          MeshWs& ws) {
    f3 mn, mx;
    if (t == PGX_SPHERE) {
        mn = sub(localPos, s3(g.x)); mx = add(localPos, s3(g.x));
    } else if (t == PGX_CAPSULE) {
        f3 a = add(localPos, qrot(localOr, mk3(0.f, g.x, 0.f))), b = add(localPos, qrot(localOr, mk3(0.f, -g.x, 0.f)));
        mn = subs(vmin(a, b), g.y); mx = adds(vmax(a, b), g.y);
    } else if (t == PGX_BOX) {
        m3 u = mat3_cast(localOr);
        f3 w = mulmv(mk_m3(vabs(u.c0), vabs(u.c1), vabs(u.c2)), mk3(g.x, g.y, g.z));
        mn = sub(localPos, w); mx = add(localPos, w);
    } else {
        int4 ci = cv.mesh[(int)g.w];
        f3 scale = mk3(g.x, g.y, g.z);
        mn = s3(3.402823466e+38f); mx = s3(-3.402823466e+38f);
        int np = (ci.y + 3) & ~3;
        for (int i = 0; i < np; ++i) {
            f3 p = add(localPos, qrot(localOr, mulv(scale, xyz(cv.verts[ci.x + i]))));
            mn = vmin(mn, p); mx = vmax(mx, p);
        }
    }
    int stack[TM_STACK];
    int sp = 0, nc = 0;
    stack[sp++] = 0;
    const float4* nodes = mc.tv->nodes + 2 * (size_t)mc.mi.w;
    while (sp > 0) {
        int id = stack[--sp];
        float4 lo = nodes[2 * id], hi = nodes[2 * id + 1];
        bool over = !(mx.x < lo.x || mn.x > hi.x) && !(mx.y < lo.y || mn.y > hi.y) && !(mx.z < lo.z || mn.z > hi.z);
        if (!over) continue;
        int triCount = __float_as_int(lo.w), index = __float_as_int(hi.w);
        if (triCount) {
            for (int i = 0; i < triCount; ++i) {
                if (nc >= TM_MAXCAND) return -1;
                ws.cand(nc++) = index + i;
            }
        } else {
            if (sp + 2 > TM_STACK) return -1;
            stack[sp++] = index + 1;
            stack[sp++] = index;
        }
    }
    return nc;
}

// the per-shape test of one candidate triangle ti (src/CollisionTriangleMesh.cpp:886-903)
__device__ bool mesh_test(int t, float4 g, f3 localPos, qf localOr, const SupConvex& sc, const MeshCtx& mc, int ti,
                          EpaWs& ews, TCI& ci, bool& epaErr) {
    int4 ix = mc.tri(ti);
    f3 a = mc.vert(ix.x), b = mc.vert(ix.y), c = mc.vert(ix.z);
    f3 n = mc.n(ti);
    ci.featureIndex = 0; ci.boxFeature = 0; ci.boxAxis = 0;
    bool hit;
    if (t == PGX_SPHERE) hit = sphere_tri(localPos, g.x, a, b, c, n, ci);
    else if (t == PGX_CAPSULE) hit = capsule_tri(localPos, localOr, g.x, g.y, a, b, c, n, ci);
    else if (t == PGX_BOX) hit = box_tri(localPos, localOr, mk3(g.x, g.y, g.z), a, b, c, n, ci);
    else {
        hit = convex_tri(sc, a, b, c, mc.c(ti), ews, ci);
        if (ews.overflow) epaErr = true;
    }
    ci.tri = ti;
    return hit;
}

__device__ __forceinline__ bool is_voided(MeshWs& ws, int nv, int v) {
    for (int i = 0; i < nv; ++i)
        if (ws.voided(i) == v) return true;
    return false;
}

// the emission passes (:906-953) over the nct contact records ws.order(0 .. nct-1) lists
// in candidate order; out.emit(manifold, triangle) takes each manifold in order
template <class Out>
__device__ void mesh_emit(int t, float4 g, f3 localPos, qf localOr, f3 pos1, qf or1, const MeshCtx& mc,
                          const ConvexView& cv, MeshWs& ws, const MeshRecs& recs, int nct, Out& out) {
    f3 he = mk3(g.x, g.y, g.z);
    MeshRef cm{&cv, make_int4(0, 0, 0, 0)};
    if (t == PGX_CONVEX_MESH) cm.info = cv.mesh[(int)g.w];
    int nv = 0;
    int n = nct;
    // first pass: face contacts, last to first, swap-removed
    for (int i = n - 1; i >= 0; i--) {
        int c = ws.order(i);
        if (recs.feature(c) != TF_FACE) continue;
        TCI ci = recs.load(c);
        int4 ix = mc.tri(ci.tri);
        ws.voided(nv++) = ix.x; ws.voided(nv++) = ix.y; ws.voided(nv++) = ix.z;
        Mf m;
        m.numPoints = 0;
        if (t == PGX_SPHERE) { gen_sphere(pos1, or1, ci, m); out.emit(m, ci.tri); }
        else if (t == PGX_CAPSULE) { if (gen_capsule_face(localPos, localOr, g.x, g.y, pos1, or1, mc, ci, m)) out.emit(m, ci.tri); }
        else if (t == PGX_BOX) { gen_box_tri_face(localPos, localOr, he, pos1, or1, mc, ci, m); out.emit(m, ci.tri); }
        else { gen_convex_tri_face(localPos, localOr, cm, he, pos1, or1, mc, ci, m); out.emit(m, ci.tri); }
        int last = ws.order(n - 1);
        ws.order(n - 1) = ws.order(i);
        ws.order(i) = last;
        --n;
    }
    OrderArr arr{&ws, &recs};
    introsort(arr, n);
    for (int k = 0; k < n; ++k) {
        int c = ws.order(k);
        TCI ci = recs.load(c);
        int4 ix = mc.tri(ci.tri);
        int iv[3] = {ix.x, ix.y, ix.z};
        Mf m;
        m.numPoints = 0;
        if (ci.feature == TF_EDGE) {
            if (is_voided(ws, nv, iv[(ci.featureIndex + 1) % 3]) && is_voided(ws, nv, iv[(ci.featureIndex + 2) % 3]))
                continue;
            if (t == PGX_SPHERE || t == PGX_CAPSULE) gen_sphere(pos1, or1, ci, m);
            else if (t == PGX_BOX) {
                if (ci.boxFeature == BF_FACE) gen_box_face_tri(localPos, localOr, he, pos1, or1, mc, ci, m);
                else gen_box_edge_tri_edge(localPos, localOr, he, pos1, or1, mc, ci, m);
            } else gen_convex_face_tri(localPos, localOr, cm, he, pos1, or1, mc, ci, m);
            out.emit(m, ci.tri);
        } else if (ci.feature == TF_VERTEX) {
            if (is_voided(ws, nv, iv[ci.featureIndex])) continue;
            if (t == PGX_SPHERE || t == PGX_CAPSULE) gen_sphere(pos1, or1, ci, m);
            else if (t == PGX_BOX) gen_box_face_tri(localPos, localOr, he, pos1, or1, mc, ci, m);
            else gen_convex_face_tri(localPos, localOr, cm, he, pos1, or1, mc, ci, m);
            out.emit(m, ci.tri);
        }
        ws.voided(nv++) = ix.x; ws.voided(nv++) = ix.y; ws.voided(nv++) = ix.z;
    }
}

}  // namespace pgx
