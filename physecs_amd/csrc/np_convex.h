// Convex-mesh narrowphase on the device: GJK (src/GJK.h:131-292), EPA (src/EPA.h:9-186)
// and the convex contact generators (src/Collision.cpp:352-426,488-499,694-887,
// src/ContactsUtil.cpp:109-204), operation for operation with the reference, including
// its mixed precision: simplex and polytope vertices are doubles, every support query
// narrows its direction to float, and the float helpers the reference feeds with double
// vectors (isZero, sqrDistPointToLine, distPointToPlane, the barycentric edges) narrow
// them the same way.
//
// Pairs with a convex mesh on either side are taken out of the one-pair-per-lane
// primitive kernel into a list and solved by k_narrowphase_convex, a grid-stride kernel
// of a fixed number of lanes. The EPA polytope of every lane lives in a device workspace
// laid out lane-interleaved (element e of lane l at e * lanes + l), so that the lanes of a
// wave walking the same face or vertex index read one contiguous line.
// Included by narrowphase.hip (shares its clipping and box-face routines).
#pragma once

namespace pgx {

constexpr int CVX_MAXPOLY = 40;   // a clipped face polygon (faces have at most 16 vertices)


// ------------------------------------------------------------------ support maps
struct SupSphere {                                  // src/GJK.h:13-20
    f3 pos; float r;
    __device__ f3 operator()(f3 d) const { return add(pos, mul(d, r)); }
};
struct SupCapsule {                                 // src/GJK.h:22-31
    f3 pos, axis; float hh, r;
    __device__ f3 operator()(f3 d) const {
        return add(add(pos, mul(mul(axis, gsign(dot(d, axis))), hh)), mul(d, r));
    }
};
struct SupBox {                                     // src/GJK.h:33-45
    f3 pos; m3 l2w; f3 he;
    __device__ f3 operator()(f3 d) const {
        f3 s = pos;
        for (int i = 0; i < 3; ++i) {
            f3 c = col(l2w, i);
            s = add(s, mul(mul(gsign(dot(d, c)), c), at(he, i)));
        }
        return s;
    }
};
// src/GJK.h:47-107: four interleaved running maxima (vertex i -> lane i % 4, strict '>'),
// then the lanes compared in order with strict '>'
struct SupConvex {
    f3 pos; m3 l2w; f3 scale; const float4* v; int count;
    __device__ f3 operator()(f3 dir) const {
        dir = mulv(scale, mk3(dot(dir, l2w.c0), dot(dir, l2w.c1), dot(dir, l2w.c2)));
        float mv[4];
        int mi[4] = {0, 1, 2, 3};
        for (int k = 0; k < 4; ++k) mv[k] = -3.402823466e+38f;
        for (int i = 0; i < count; i += 4) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float4 p = v[i + k];
                float d = (p.x * dir.x + p.y * dir.y) + p.z * dir.z;
                if (d > mv[k]) mi[k] = i + k;
                mv[k] = d > mv[k] ? d : mv[k];
            }
        }
        int best = mi[0];
        float bv = mv[0];
        for (int k = 1; k < 4; ++k)
            if (mv[k] > bv) { bv = mv[k]; best = mi[k]; }
        return add(pos, mulmv(l2w, mulv(scale, xyz(v[best]))));
    }
};

// ------------------------------------------------------------------ GJK
struct GjkV { d3 pos, sp0, sp1; };

template <class S0, class S1>
__device__ __forceinline__ GjkV mdiff(const S0& s0, const S1& s1, f3 dir) {      // src/GJK.h:192-196
    f3 a = s0(dir);
    f3 b = s1(neg(dir));
    GjkV r;
    r.pos = wide(sub(a, b)); r.sp0 = wide(a); r.sp1 = wide(b);
    return r;
}
__device__ __forceinline__ float sqr_dist_point_line(f3 p, f3 a, f3 b) {          // src/GJK.h:142-148
    f3 r = sub(p, a), d = sub(b, a);
    f3 q = add(a, mul(dot(r, d) / dot(d, d), d));
    f3 pq = sub(p, q);
    return dot(pq, pq);
}
__device__ __forceinline__ float dist_point_plane(f3 p, f3 o, f3 n) { return gabs(dot(sub(p, o), n)); }
__device__ __forceinline__ bool is_zero(f3 v) { return (double)dot(v, v) < 1e-6; }

template <class S0, class S1>
__device__ bool check_direction(const S0& s0, const S1& s1, d3 dir, GjkV& v) {    // src/GJK.h:162-167
    dir = dnormalize(dir);
    v = mdiff(s0, s1, narrow(dir));
    return ddot(v.pos, dir) >= 0;
}
__device__ __forceinline__ bool check_face(d3 v0, d3 v1, d3 v2, d3 opp, d3& n) {  // src/GJK.h:169-173
    n = dcross(dsub(v1, v0), dsub(v2, v0));
    n = dmul(dsign(ddot(dsub(v0, opp), n)), n);
    return ddot(n, dneg(v0)) > (double)0.00001f;
}
template <class S0, class S1>
__device__ void complete_from_line(const S0& s0, const S1& s1, GjkV* s) {         // src/GJK.h:233-243
    d3 dir = dcross(dsub(s[0].pos, s[1].pos), mkd(0, 1, 0));
    if (is_zero(narrow(dir))) dir = dcross(dsub(s[0].pos, s[1].pos), mkd(0, 0, 1));
    s[2] = mdiff(s0, s1, narrow(dnormalize(dir)));
    if (sqr_dist_point_line(narrow(s[2].pos), narrow(s[0].pos), narrow(s[1].pos)) < 0.00001f)
        s[2] = mdiff(s0, s1, narrow(dneg(dir)));
    dir = dnormalize(dcross(dsub(s[1].pos, s[0].pos), dsub(s[2].pos, s[0].pos)));
    s[3] = mdiff(s0, s1, narrow(dir));
    if (dist_point_plane(narrow(s[3].pos), narrow(s[0].pos), narrow(dir)) < 0.0001f)
        s[3] = mdiff(s0, s1, narrow(dneg(dir)));
}
template <class S0, class S1>
__device__ bool gjk(const S0& s0, const S1& s1, f3 startDir, GjkV* s) {          // src/GJK.h:226-292
    d3 dir = wide(normalize(startDir));
    s[0] = mdiff(s0, s1, narrow(dir));
    if (is_zero(narrow(s[0].pos))) {
        dir = dneg(dir);
        s[1] = mdiff(s0, s1, narrow(dir));
        complete_from_line(s0, s1, s);
        return true;
    }
    dir = dneg(s[0].pos);
    if (!check_direction(s0, s1, dir, s[1])) return false;
    dir = dcross(dcross(dsub(s[0].pos, s[1].pos), dneg(s[1].pos)), dsub(s[0].pos, s[1].pos));
    if (is_zero(narrow(dir))) {
        complete_from_line(s0, s1, s);
        return true;
    }
    if (!check_direction(s0, s1, dir, s[2])) return false;
    dir = dcross(dsub(s[1].pos, s[0].pos), dsub(s[2].pos, s[0].pos));
    double sg = dsign(ddot(dneg(s[0].pos), dir));
    if (!sg) {
        dir = dnormalize(dir);
        s[3] = mdiff(s0, s1, narrow(dir));
        if (dist_point_plane(narrow(s[3].pos), narrow(s[0].pos), narrow(dir)) < 0.0001f)
            s[3] = mdiff(s0, s1, narrow(dneg(dir)));
        return true;
    }
    dir = dmul(sg, dir);
    if (!check_direction(s0, s1, dir, s[3])) return false;
    for (int i = 0; i < 100; ++i) {
        // simplexContainsOrigin (src/GJK.h:175-181)
        int fi = 2;
        bool outside = check_face(s[3].pos, s[0].pos, s[1].pos, s[2].pos, dir);
        if (!outside) { fi = 1; outside = check_face(s[3].pos, s[2].pos, s[0].pos, s[1].pos, dir); }
        if (!outside) { fi = 0; outside = check_face(s[3].pos, s[1].pos, s[2].pos, s[0].pos, dir); }
        if (!outside) return true;
        if (!check_direction(s0, s1, dir, s[fi])) return false;
        GjkV t = s[fi];
        s[fi] = s[3];
        s[3] = t;
    }
    return false;
}

// ------------------------------------------------------------------ EPA
// one lane's polytope in the lane-interleaved workspace
struct EpaWs {
    double* vd;       // 9 doubles per vertex: pos, sp0, sp1
    int* fi;          // 3 ints per face
    double* fn;       // 3 doubles per face (unit normal)
    int* ed;          // 2 ints per loose edge
    size_t L;         // lanes
    int nv, nf;
    bool overflow;
    __device__ double& V(int v, int c) { return vd[(size_t)(v * 9 + c) * L]; }
    __device__ int& FI(int f, int k) { return fi[(size_t)(f * 3 + k) * L]; }
    __device__ double& FN(int f, int c) { return fn[(size_t)(f * 3 + c) * L]; }
    __device__ int& E(int e, int k) { return ed[(size_t)(e * 2 + k) * L]; }
    __device__ d3 pos(int v) { return mkd(V(v, 0), V(v, 1), V(v, 2)); }
    __device__ d3 normal(int f) { return mkd(FN(f, 0), FN(f, 1), FN(f, 2)); }
    __device__ void putV(int v, const GjkV& g) {
        V(v, 0) = g.pos.x; V(v, 1) = g.pos.y; V(v, 2) = g.pos.z;
        V(v, 3) = g.sp0.x; V(v, 4) = g.sp0.y; V(v, 5) = g.sp0.z;
        V(v, 6) = g.sp1.x; V(v, 7) = g.sp1.y; V(v, 8) = g.sp1.z;
    }
    __device__ GjkV getV(int v) {
        GjkV g;
        g.pos = mkd(V(v, 0), V(v, 1), V(v, 2));
        g.sp0 = mkd(V(v, 3), V(v, 4), V(v, 5));
        g.sp1 = mkd(V(v, 6), V(v, 7), V(v, 8));
        return g;
    }
    __device__ void create_face(int i0, int i1, int i2) {              // src/EPA.h:24-34
        if (nf >= EPA_MAXF) { overflow = true; return; }
        d3 n = dcross(dsub(pos(i0), pos(i1)), dsub(pos(i2), pos(i1)));
        double len = dlength(n);
        if (len) n = ddivs(n, len);
        FI(nf, 0) = i0; FI(nf, 1) = i1; FI(nf, 2) = i2;
        FN(nf, 0) = n.x; FN(nf, 1) = n.y; FN(nf, 2) = n.z;
        ++nf;
    }
    __device__ void copy_face(int dst, int src) {
        for (int k = 0; k < 3; ++k) { FI(dst, k) = FI(src, k); FN(dst, k) = FN(src, k); }
    }
    __device__ double face_distance(int f) { return ddot(pos(FI(f, 0)), normal(f)); }   // src/EPA.h:41-43
};

template <class S0, class S1>
__device__ f3 epa(const S0& s0, const S1& s1, GjkV* s, EpaWs& P, f3& cp0, f3& cp1) {   // src/EPA.h:171-186
    P.nv = 4; P.nf = 0; P.overflow = false;
    for (int i = 0; i < 4; ++i) P.putV(i, s[i]);
    {   // EpaPolytope(GjkSimplex&), src/EPA.h:50-66
        d3 n = dcross(dsub(s[0].pos, s[1].pos), dsub(s[2].pos, s[1].pos));
        if (ddot(dsub(s[3].pos, s[0].pos), n) < 0) {
            P.create_face(0, 1, 2); P.create_face(3, 1, 0); P.create_face(3, 2, 1); P.create_face(3, 0, 2);
        } else {
            P.create_face(0, 2, 1); P.create_face(3, 2, 0); P.create_face(3, 1, 2); P.create_face(3, 0, 1);
        }
    }
    float minDist = 0.f;
    d3 normal = mkd(0, 0, 0);
    int faceIndex = 0;
    for (int it = 0; it < 100 && !P.overflow; ++it) {
        // findClosestFace: a float running minimum compared against double distances (:68-79)
        minDist = __builtin_inff();
        for (int f = 0; f < P.nf; ++f) {
            double d = P.face_distance(f);
            if (d < (double)minDist) { minDist = (float)d; normal = P.normal(f); faceIndex = f; }
        }
        GjkV v = mdiff(s0, s1, narrow(normal));
        if (ddot(v.pos, normal) - (double)minDist < (double)0.00001f) break;
        // insertVertex (:81-120)
        if (P.nv >= EPA_MAXV) { P.overflow = true; break; }
        int ne = 0;
        for (int f = P.nf - 1; f >= 0; --f) {
            if (ddot(P.normal(f), dsub(v.pos, P.pos(P.FI(f, 0)))) > 0) {
                int e[3][2] = {{P.FI(f, 0), P.FI(f, 1)}, {P.FI(f, 1), P.FI(f, 2)}, {P.FI(f, 2), P.FI(f, 0)}};
                bool found[3] = {false, false, false};
                for (int j = 0; j < 3; ++j)
                    for (int k = ne - 1; k >= 0; --k)
                        if (P.E(k, 0) == e[j][1] && P.E(k, 1) == e[j][0]) {
                            found[j] = true;
                            P.E(k, 0) = P.E(ne - 1, 0);
                            P.E(k, 1) = P.E(ne - 1, 1);
                            --ne;
                            break;
                        }
                for (int j = 0; j < 3; ++j)
                    if (!found[j]) {
                        if (ne >= EPA_MAXE) { P.overflow = true; break; }
                        P.E(ne, 0) = e[j][0];
                        P.E(ne, 1) = e[j][1];
                        ++ne;
                    }
                P.copy_face(f, P.nf - 1);
                --P.nf;
            }
        }
        int vi = P.nv++;
        P.putV(vi, v);
        for (int k = 0; k < ne; ++k) P.create_face(vi, P.E(k, 0), P.E(k, 1));
    }
    // getClosestPoints (:122-168), including proj = normal + distance (:133)
    int fa = P.FI(faceIndex, 0), fb = P.FI(faceIndex, 1), fc = P.FI(faceIndex, 2);
    GjkV a = P.getV(fa), b = P.getV(fb), c = P.getV(fc);
    d3 proj = dadds(P.normal(faceIndex), P.face_distance(faceIndex));
    f3 v0 = narrow(dsub(b.pos, a.pos)), v1 = narrow(dsub(c.pos, a.pos)), v2 = narrow(dsub(proj, a.pos));
    double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    double denom = d00 * d11 - d01 * d01;
    double u, vv, w;
    if (denom) {
        vv = (d20 * d11 - d21 * d01) / denom;
        w = (d00 * d21 - d01 * d20) / denom;
        u = 1.0 - vv - w;
    } else {
        w = 0;
        if (d00) { vv = d20 / d00; u = 1.0 - vv; }
        else { vv = 0; u = 1.0; }
    }
    cp0 = narrow(dadd(dadd(dmul(u, a.sp0), dmul(vv, b.sp0)), dmul(w, c.sp0)));
    cp1 = narrow(dadd(dadd(dmul(u, a.sp1), dmul(vv, b.sp1)), dmul(w, c.sp1)));
    return narrow(normal);
}

// ------------------------------------------------------------------ contact generation
template <int CAP>
__device__ void clip_polygon_n(f2* poly, int& n, const f2* clip, int nclip) {   // src/Clipping.cpp:21-53
    f2 out[CAP];
    for (int e = 0; e < nclip; ++e) {
        f2 a = clip[e], b = clip[(e + 1) % nclip];
        int m = 0;
        for (int i = 0; i < n; ++i) {
            int j = (i + 1) == n ? 0 : i + 1;
            bool ii = inside(poly[i], a, b), ji = inside(poly[j], a, b);
            if (ii && ji) { if (m < CAP) out[m++] = poly[j]; }
            else if (ii) { if (m < CAP) out[m++] = intersect2(poly[i], poly[j], a, b); }
            else if (ji) {
                if (m < CAP) out[m++] = intersect2(poly[i], poly[j], a, b);
                if (m < CAP) out[m++] = poly[j];
            }
        }
        n = m;
        for (int i = 0; i < m; ++i) poly[i] = out[i];
    }
}
__device__ bool clip_line_n(f2& p0, f2& p1, const f2* clip, int nclip) {        // src/Clipping.cpp:55-71
    for (int i = 0; i < nclip; ++i) {
        int j = (i + 1) % nclip;
        bool in0 = inside(p0, clip[i], clip[j]), in1 = inside(p1, clip[i], clip[j]);
        if (in0 && in1) continue;
        if (in0) p1 = intersect2(p0, p1, clip[i], clip[j]);
        else if (in1) p0 = intersect2(p0, p1, clip[i], clip[j]);
        else return false;
    }
    return true;
}

// src/ContactsUtil.cpp:109-204
__device__ void polygon_polygon_face_contacts(f3 refPos, const m3& refToWorld, f3 refOrig, f3 refNormal,
                                              f3 incOrig, f3 incNormal, const f2* poly, int n, int cx, int cy,
                                              CPt* cps, int& numPoints) {
    f3 pen[CVX_MAXPOLY];
    int np = 0;
    float distRef = dot(refOrig, refNormal);
    f3 dir = mk3(0.f, 1.f, 0.f);
    for (int i = 0; i < n; ++i) {
        f3 p3 = mk3(poly[i].y, 0.f, poly[i].x);
        float distance = dot(sub(incOrig, p3), incNormal) / dot(dir, incNormal);
        if (distance < distRef + 1e-4f) {
            p3.y = distance;
            pen[np++] = p3;
        }
    }
    auto emit = [&](int slot, int idx) {
        f3 on = pen[idx];
        on.y = distRef;
        cps[slot].p0 = add(refPos, mulmv(refToWorld, on));
        cps[slot].p1 = add(refPos, mulmv(refToWorld, pen[idx]));
    };
    if (np > 4) {
        emit(0, 0);
        float maxD = 0.f;
        int maxDi = 0;
        for (int i = 0; i < np; ++i) {
            f3 d = sub(pen[i], pen[0]);
            float dd = dot(d, d);
            if (dd > maxD) { maxD = dd; maxDi = i; }
        }
        emit(1, maxDi);
        float maxA = 0.f, minA = 0.f;
        int maxAi = 0, minAi = 0;
        for (int i = 0; i < np; ++i) {
            f3 ca = sub(pen[0], pen[i]), cb = sub(pen[maxDi], pen[i]);
            float area = det2(mk2(at(ca, cx), at(ca, cy)), mk2(at(cb, cx), at(cb, cy)));
            if (area > maxA) { maxA = area; maxAi = i; }
        }
        emit(2, maxAi);
        for (int i = 0; i < np; ++i) {
            f3 da = sub(pen[0], pen[i]), db = sub(pen[maxDi], pen[i]);
            float area = det2(mk2(at(da, cx), at(da, cy)), mk2(at(db, cx), at(db, cy)));
            if (area < minA) { minA = area; minAi = i; }
        }
        emit(3, minAi);
        numPoints = 4;
    } else {
        for (int i = 0; i < np; ++i) emit(i, i);
        numPoints = np;
    }
}

__device__ __forceinline__ f3 world_face_normal(qf q, f3 n, f3 scale) { return normalize(qrot(q, divv(n, scale))); }

struct MeshRef {
    const ConvexView* cv; int4 info;
    __device__ int nfaces() const { return info.w; }
    __device__ int2 face(int f) const { return cv->face[info.z + f]; }
    __device__ f3 fnormal(int f) const { return xyz(cv->faceN[info.z + f]); }
    __device__ f3 fcentroid(int f) const { return xyz(cv->faceC[info.z + f]); }
    __device__ f3 vert(int i) const { return xyz(cv->verts[info.x + i]); }
    __device__ f3 face_vert(int2 fc, int k) const { return vert(cv->faceIdx[fc.x + k]); }
};

// src/Collision.cpp:352-410
__device__ void convex_convex_contacts(f3 pos0, qf or0, const MeshRef& m0, f3 s0, f3 pos1, qf or1,
                                       const MeshRef& m1, f3 s1, f3 normal, CPt* cps, int& numPoints) {
    int f0 = 0;
    float maxDot = 0.f;
    for (int i = 0; i < m0.nfaces(); ++i) {
        float d = dot(normal, world_face_normal(or0, m0.fnormal(i), s0));
        if (d > maxDot) { maxDot = d; f0 = i; }
    }
    int f1 = 0;
    float minDot = 0.f;
    for (int i = 0; i < m1.nfaces(); ++i) {
        float d = dot(normal, world_face_normal(or1, m1.fnormal(i), s1));
        if (d < minDot) { minDot = d; f1 = i; }
    }
    int2 F0 = m0.face(f0), F1 = m1.face(f1);
    m3 worldToC0 = transpose(mat3_cast(or0));
    f3 refOrig = mulv(s0, m0.fcentroid(f0));
    f3 u0 = normalize(sub(mulv(s0, m0.face_vert(F0, 0)), refOrig));
    f3 u1 = normalize(divv(m0.fnormal(f0), s0));
    f3 u2 = cross(u0, u1);
    m3 c0ToRef = transpose(mk_m3(u0, u1, u2));
    m3 worldToRef = mulmm(c0ToRef, worldToC0);
    f2 clip[16];
    f2 poly[CVX_MAXPOLY];
    int nc = F0.y < 16 ? F0.y : 16;
    for (int i = 0; i < nc; ++i) {
        f3 v = mulmv(c0ToRef, mulv(s0, m0.face_vert(F0, i)));
        clip[i] = mk2(v.z, v.x);
    }
    int n = F1.y < 16 ? F1.y : 16;
    for (int i = 0; i < n; ++i) {
        f3 v = mulmv(worldToRef, sub(add(pos1, qrot(or1, mulv(s1, m1.face_vert(F1, i)))), pos0));
        poly[i] = mk2(v.z, v.x);
    }
    clip_polygon_n<CVX_MAXPOLY>(poly, n, clip, nc);
    f3 incOrig = mulmv(worldToRef, sub(add(pos1, qrot(or1, mulv(s1, m1.fcentroid(f1)))), pos0));
    f3 incNormal = mulmv(worldToRef, qrot(or1, normalize(divv(m1.fnormal(f1), s1))));
    m3 refToWorld = transpose(worldToRef);
    polygon_polygon_face_contacts(pos0, refToWorld, refOrig, u1, incOrig, incNormal, poly, n, 2, 0, cps, numPoints);
}

// src/Collision.cpp:694-785
__device__ void capsule_convex_contacts(f3 p0L, f3 p1L, float radius, f3 meshPos, const m3& c2w, const MeshRef& m,
                                        f3 scale, Mf& res) {
    int fi = 0;
    float minDot = 0.f;
    for (int i = 0; i < m.nfaces(); ++i) {
        float d = dot(res.normal, normalize(mulmv(c2w, divv(m.fnormal(i), scale))));
        if (d < minDot) { minDot = d; fi = i; }
    }
    int2 F = m.face(fi);
    f3 refOrig = mulv(scale, m.fcentroid(fi));
    f3 u0 = normalize(sub(mulv(scale, m.face_vert(F, 0)), refOrig));
    f3 u1 = normalize(divv(m.fnormal(fi), scale));
    f3 u2 = cross(u0, u1);
    m3 convexToRef = transpose(mk_m3(u0, u1, u2));
    f2 clip[16];
    int nc = F.y < 16 ? F.y : 16;
    for (int i = 0; i < nc; ++i) {
        f3 v = mulmv(convexToRef, mulv(scale, m.face_vert(F, i)));
        clip[i] = mk2(v.z, v.x);
    }
    f3 p0Ref = mulmv(convexToRef, p0L), p1Ref = mulmv(convexToRef, p1L);
    f2 line0 = mk2(p0Ref.z, p0Ref.x), line1 = mk2(p1Ref.z, p1Ref.x);
    if (clip_line_n(line0, line1, clip, nc)) {
        m3 refToWorld = mulmm(c2w, transpose(convexToRef));
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
        y0 -= radius;
        y1 -= radius;
        int n = 0;
        if (y0 < distToPlane) {
            f3 p = mk3(line0.y, y0, line0.x), on = mk3(line0.y, distToPlane, line0.x);
            res.pts[n].p0 = add(meshPos, mulmv(refToWorld, p));
            res.pts[n].p1 = add(meshPos, mulmv(refToWorld, on));
            ++n;
        }
        if (y1 < distToPlane) {
            f3 p = mk3(line1.y, y1, line1.x), on = mk3(line1.y, distToPlane, line1.x);
            res.pts[n].p0 = add(meshPos, mulmv(refToWorld, p));
            res.pts[n].p1 = add(meshPos, mulmv(refToWorld, on));
            ++n;
        }
        res.numPoints = n;
    } else {
        res.numPoints = 0;
    }
}

// src/Collision.cpp:808-871
__device__ void box_convex_contacts(f3 boxCenter, const m3& basis, f3 he, f3 cpos, qf cor, const MeshRef& m,
                                    f3 scale, f3 normal, CPt* cps, int& numPoints) {
    int boxAxis = 0;
    float boxAxisSign = 0.f, maxDot = 0.f;
    for (int i = 0; i < 3; ++i) {
        float d = dot(normal, col(basis, i));
        float ad = gabs(d);
        if (ad > maxDot) { maxDot = ad; boxAxis = i; boxAxisSign = d < 0 ? -1.f : 1.f; }
    }
    int fi = 0;
    float minDot = 0.f;
    for (int i = 0; i < m.nfaces(); ++i) {
        float d = dot(normal, world_face_normal(cor, m.fnormal(i), scale));
        if (d < minDot) { minDot = d; fi = i; }
    }
    int2 F = m.face(fi);
    int cx = (boxAxis + 1) % 3, cy = (boxAxis + 2) % 3;
    m3 worldToBox = transpose(basis);
    f2 poly[CVX_MAXPOLY];
    int n = F.y < 16 ? F.y : 16;
    for (int i = 0; i < n; ++i) {
        f3 v = mulmv(worldToBox, sub(add(cpos, qrot(cor, mulv(scale, m.face_vert(F, i)))), boxCenter));
        poly[i] = mk2(at(v, cx), at(v, cy));
    }
    float hx = at(he, cx), hy = at(he, cy);
    f2 clip[4] = {mk2(hx, hy), mk2(hx, -hy), mk2(-hx, -hy), mk2(-hx, hy)};
    clip_polygon_n<CVX_MAXPOLY>(poly, n, clip, 4);
    f3 incOrig = mulmv(worldToBox, sub(add(cpos, qrot(cor, mulv(scale, m.fcentroid(fi)))), boxCenter));
    f3 incNormal = normalize(mulmv(worldToBox, qrot(cor, divv(m.fnormal(fi), scale))));
    polygon_box_face_contacts_n<CVX_MAXPOLY>(boxCenter, basis, boxAxis, (int)boxAxisSign, he, incOrig, incNormal,
                                             poly, n, cx, cy, cps, numPoints);
}

// one pair with a convex mesh on either side (src/Collision.cpp:895-1016); flips as the
// reference's dispatcher does
__device__ bool collide_convex(int t0, float4 g0, f3 p0, qf q0, int t1, float4 g1, f3 p1, qf q1,
                               const ConvexView& cv, EpaWs& ws, Mf& m, bool& epaOverflow) {
    m.numPoints = 0;
    ws.overflow = false;
    bool flipRes = false;
    if (t0 == PGX_CONVEX_MESH && t1 != PGX_CONVEX_MESH) {   // mirrored: solve (other, convex) and flip
        int tt = t0; t0 = t1; t1 = tt;
        float4 gg = g0; g0 = g1; g1 = gg;
        f3 pp = p0; p0 = p1; p1 = pp;
        qf qq = q0; q0 = q1; q1 = qq;
        flipRes = true;
    }
    MeshRef m1{&cv, cv.mesh[(int)g1.w]};
    SupConvex b{p1, mat3_cast(q1), mk3(g1.x, g1.y, g1.z), cv.verts + m1.info.x, m1.info.y};
    GjkV s[4];
    bool hit = false;
    if (t0 == PGX_SPHERE) {                                  // src/Collision.cpp:488-499
        SupSphere a{p0, g0.x};
        if (gjk(a, b, sub(p1, p0), s)) {
            m.normal = epa(a, b, s, ws, m.pts[0].p0, m.pts[0].p1);
            m.numPoints = 1;
            hit = true;
        }
    } else if (t0 == PGX_CAPSULE) {                          // src/Collision.cpp:787-806
        f3 axis = qrot(q0, mk3(0.f, 1.f, 0.f));
        SupCapsule a{p0, axis, g0.x, g0.y};
        if (gjk(a, b, sub(p1, p0), s)) {
            m.normal = epa(a, b, s, ws, m.pts[0].p0, m.pts[0].p1);
            m3 w2c = inverse(b.l2w);
            f3 p = mulmv(w2c, sub(p0, p1));
            f3 axisLc = mulmv(w2c, axis);
            capsule_convex_contacts(add(p, mul(axisLc, g0.x)), sub(p, mul(axisLc, g0.x)), g0.y, p1, b.l2w, m1,
                                    b.scale, m);
            if (!m.numPoints) m.numPoints = 1;
            hit = true;
        }
    } else if (t0 == PGX_BOX) {                              // src/Collision.cpp:873-887
        SupBox a{p0, mat3_cast(q0), mk3(g0.x, g0.y, g0.z)};
        if (gjk(a, b, sub(p1, p0), s)) {
            m.normal = epa(a, b, s, ws, m.pts[0].p0, m.pts[0].p1);
            box_convex_contacts(p0, a.l2w, a.he, p1, q1, m1, b.scale, m.normal, m.pts, m.numPoints);
            if (!m.numPoints) m.numPoints = 1;
            hit = true;
        }
    } else {                                                 // src/Collision.cpp:412-426
        MeshRef m0{&cv, cv.mesh[(int)g0.w]};
        SupConvex a{p0, mat3_cast(q0), mk3(g0.x, g0.y, g0.z), cv.verts + m0.info.x, m0.info.y};
        if (gjk(a, b, sub(p1, p0), s)) {
            m.normal = epa(a, b, s, ws, m.pts[0].p0, m.pts[0].p1);
            convex_convex_contacts(p0, q0, m0, a.scale, p1, q1, m1, b.scale, m.normal, m.pts, m.numPoints);
            if (!m.numPoints) m.numPoints = 1;
            hit = true;
        }
    }
    epaOverflow = ws.overflow;
    if (flipRes) flip(m);
    return hit;
}

}  // namespace pgx
