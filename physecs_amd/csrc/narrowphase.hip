// Narrowphase: one candidate pair per lane, writing a manifold slot (SoA) per pair.
//
// The pair routines follow the reference's src/Collision.cpp (dispatcher :895-1016,
// sphere/capsule/box closed forms :109-133,428-486, box-box SAT + face clipping
// :54-107,135-350, capsule-box :501-692), src/GeomUtil.cpp (segment distances and the
// segment-AABB regions :7-59,129-467), src/Clipping.cpp (:10-71) and
// src/ContactsUtil.cpp (:9-107), operation for operation, so the manifold is bit-identical
// to the reference's for the same poses. The polygon clip buffers live in registers
// (a quad clipped by four edges has at most 8 vertices).
#include "pgx_internal.h"
#include "shape_bounds.h"
#include <iterator>

namespace pgx {

struct CPt { f3 p0, p1; };
struct Mf {
    int numPoints;
    CPt pts[4];
    f3 normal;
};

// ------------------------------------------------------------------ GeomUtil
__device__ f3 closest_point_on_segment(f3 point, f3 o, f3 d, float mn, float mx) {
    f3 r = sub(o, point);
    float t = -dot(r, d);
    t = gclamp(t, mn, mx);
    return add(o, mul(t, d));
}

__device__ void closest_points_segments(f3 o0, f3 d0, float mn0, float mx0, f3 o1, f3 d1, float mn1, float mx1,
                                        f3& p0, f3& p1) {
    float v1v2 = dot(d0, d1);
    f3 r = sub(o1, o0);
    float rv1 = dot(r, d0);
    float rv2 = dot(r, d1);
    float t1 = gabs(v1v2) > 1.f - 0.0001f ? 0.f : (rv1 * v1v2 - rv2) / (1.f - v1v2 * v1v2);
    float t0 = rv1 + t1 * v1v2;
    if (t0 < mn0 || t0 > mx0) {
        t0 = t0 < mn0 ? mn0 : mx0;
        p0 = add(o0, mul(t0, d0));
        t1 = -dot(sub(o1, p0), d1);
    } else {
        p0 = add(o0, mul(t0, d0));
    }
    if (t1 < mn1 || t1 > mx1) {
        t1 = t1 < mn1 ? mn1 : mx1;
        p1 = add(o1, mul(t1, d1));
        t0 = -dot(sub(o0, p1), d0);
        t0 = gclamp(t0, mn0, mx0);
        p0 = add(o0, mul(t0, d0));
    } else {
        p1 = add(o1, mul(t1, d1));
    }
}

__device__ float case_two_zeros(f3 p, int axis, f3 he, float& t, f3& q) {
    float sq = 0.f;
    t = at(he, axis) - at(p, axis);
    set(q, axis, at(p, axis) + t);
    for (int k = 1; k <= 2; ++k) {
        int a = (axis + k) % 3;
        float pa = at(p, a), ha = at(he, a);
        if (pa > ha) { float d = pa - ha; sq += d * d; set(q, a, ha); }
        else if (pa < -ha) { float d = pa + ha; sq += d * d; set(q, a, -ha); }
        else set(q, a, pa);
    }
    return sq;
}

__device__ float case_one_zero(f3 p, int z0, f3 dir, f3 he, float& t, f3& q) {
    float sq = 0.f;
    set(q, z0, at(p, z0));
    f3 pme = sub(p, he);
    int x = (z0 + 1) % 3, y = (z0 + 2) % 3;
    float prod0 = at(dir, y) * at(pme, x);
    float prod1 = at(dir, x) * at(pme, y);
    if (prod0 >= prod1) {
        set(q, x, at(he, x));
        float eY = at(p, y) + at(he, y);
        float delta = prod0 - at(dir, x) * eY;
        if (delta >= 0) {
            sq += delta * delta;
            set(q, y, -at(he, y));
            t = -(at(pme, x) * at(dir, x) + eY * at(dir, y));
        } else {
            set(q, y, at(p, y) - prod0 / at(dir, x));
            t = -at(pme, x) / at(dir, x);
        }
    } else {
        set(q, y, at(he, y));
        float eX = at(p, x) + at(he, x);
        float delta = prod1 - at(dir, y) * eX;
        if (delta >= 0) {
            sq += delta * delta;
            set(q, x, -at(he, x));
            t = -(eX * at(dir, x) + at(pme, y) * at(dir, y));
        } else {
            set(q, x, at(p, x) - prod1 / at(dir, y));
            t = -at(pme, y) / at(dir, y);
        }
    }
    float pz = at(p, z0), hz = at(he, z0);
    if (pz < -hz) { float d = pz + hz; sq += d * d; set(q, z0, -hz); }
    else if (pz > hz) { float d = pz - hz; sq += d * d; set(q, z0, hz); }
    return sq;
}

// Region 4/5 and 1/2/3 helpers of the face case. "Region 5" keeps the reference's
// pme[y] + pme[y] term (src/GeomUtil.cpp:273,326).
__device__ __forceinline__ float region45(f3 p, f3 dir, f3 he, int x, int y, int z, f3 pme, f3 ppe, float npY,
                                          float& t, f3& q) {
    float dx = at(dir, x), dy = at(dir, y), dz = at(dir, z);
    float pmx = at(pme, x), pmy = at(pme, y), ppz = at(ppe, z);
    if (npY <= at(he, y)) {
        float tmp = at(p, y) - npY;
        float delta = dx * pmx + dy * tmp + dz * ppz;
        t = -delta;
        set(q, x, at(he, x)); set(q, y, npY); set(q, z, -at(he, z));
        return pmx * pmx + tmp * tmp + ppz * ppz - delta * delta;
    }
    float delta = dx * pmx + dy * pmy + dz * ppz;
    t = -delta;
    set(q, x, at(he, x)); set(q, y, at(he, y)); set(q, z, -at(he, z));
    return pmx * pmx + pmy + pmy + ppz * ppz - delta * delta;
}
__device__ __forceinline__ float region12(f3 p, f3 dir, f3 he, int x, int y, int z, f3 pme, f3 ppe, float npZ,
                                          float& t, f3& q) {
    float dx = at(dir, x), dy = at(dir, y), dz = at(dir, z);
    float pmx = at(pme, x), ppy = at(ppe, y), pmz = at(pme, z);
    if (npZ <= at(he, z)) {
        float tmp = at(p, z) - npZ;
        float delta = dx * pmx + dy * ppy + dz * tmp;
        t = -delta;
        set(q, x, at(he, x)); set(q, y, -at(he, y)); set(q, z, npZ);
        return pmx * pmx + ppy * ppy + tmp * tmp - delta * delta;
    }
    float delta = dx * pmx + dy * ppy + dz * pmz;
    t = -delta;
    set(q, x, at(he, x)); set(q, y, -at(he, y)); set(q, z, at(he, z));
    return pmx * pmx + ppy * ppy + pmz * pmz - delta * delta;
}

__device__ float line_aabb_face(f3 p, f3 dir, f3 he, int x, f3 pme, float& t, f3& q) {
    q = p;
    t = 0.f;
    int y = (x + 1) % 3, z = (y + 1) % 3;
    f3 ppe = add(p, he);
    float dx = at(dir, x), dy = at(dir, y), dz = at(dir, z);
    float pmx = at(pme, x);
    if (dx * at(ppe, y) >= dy * pmx) {
        if (dx * at(ppe, z) >= dz * pmx) {
            set(q, x, at(he, x));
            set(q, y, at(q, y) - pmx * dy / dx);
            set(q, z, at(q, z) - pmx * dz / dx);
            t = -pmx / dx;
            return 0.f;
        }
        float sqrLen = dx * dx + dz * dz;
        float npY = at(p, y) - dy * (dx * pmx + dz * at(ppe, z)) / sqrLen;
        return region45(p, dir, he, x, y, z, pme, ppe, npY, t, q);
    }
    if (dx * at(ppe, z) >= dz * pmx) {
        float sqrLen = dx * dx + dy * dy;
        float npZ = at(p, z) - dz * (dx * pmx + dy * at(ppe, y)) / sqrLen;
        return region12(p, dir, he, x, y, z, pme, ppe, npZ, t, q);
    }
    float npY = at(p, y) - dy * (dx * pmx + dz * at(ppe, z)) / (dx * dx + dz * dz);
    if (npY >= -at(he, y)) return region45(p, dir, he, x, y, z, pme, ppe, npY, t, q);
    float npZ = at(p, z) - dz * (dx * pmx + dy * at(ppe, y)) / (dx * dx + dy * dy);
    if (npZ >= -at(he, z)) return region12(p, dir, he, x, y, z, pme, ppe, npZ, t, q);
    float delta = dx * pmx + dy * at(ppe, y) + dz * at(ppe, z);
    t = -delta;
    set(q, x, at(he, x)); set(q, y, -at(he, y)); set(q, z, -at(he, z));
    return pmx * pmx + at(ppe, y) * at(ppe, y) + at(ppe, z) * at(ppe, z) - delta * delta;
}

__device__ float case_no_zeroes(f3 p, f3 dir, f3 he, float& t, f3& q) {
    f3 pme = sub(p, he);
    float dxEy = dir.x * pme.y, dyEx = dir.y * pme.x;
    if (dyEx >= dxEy) {
        float dzEx = dir.z * pme.x, dxEz = dir.x * pme.z;
        return line_aabb_face(p, dir, he, dzEx >= dxEz ? 0 : 2, pme, t, q);
    }
    float dzEy = dir.z * pme.y, dyEz = dir.y * pme.z;
    return line_aabb_face(p, dir, he, dzEy >= dyEz ? 1 : 2, pme, t, q);
}

__device__ float point_aabb(f3 p, f3 he, f3& q) {
    q = vclamp(p, neg(he), he);
    f3 d = sub(p, q);
    return dot(d, d);
}

__device__ float segment_aabb(f3 p, f3 dir, float mn, float mx, f3 he, float& t, f3& q) {
    bool refl[3];
    f3 pp = p, dd = dir;
    for (int i = 0; i < 3; ++i) {
        if (at(dd, i) < 0.f) { set(pp, i, -at(pp, i)); set(dd, i, -at(dd, i)); refl[i] = true; }
        else refl[i] = false;
    }
    float sq;
    bool px = dd.x > 0, py = dd.y > 0, pz = dd.z > 0;
    if (px && py && pz) sq = case_no_zeroes(pp, dd, he, t, q);
    else if (px && py) sq = case_one_zero(pp, 2, dd, he, t, q);
    else if (px && pz) sq = case_one_zero(pp, 1, dd, he, t, q);
    else if (py && pz) sq = case_one_zero(pp, 0, dd, he, t, q);
    else if (px) sq = case_two_zeros(pp, 0, he, t, q);
    else if (py) sq = case_two_zeros(pp, 1, he, t, q);
    else if (pz) sq = case_two_zeros(pp, 2, he, t, q);
    else { sq = point_aabb(pp, he, q); t = 0.f; }
    for (int i = 0; i < 3; ++i)
        if (refl[i]) set(q, i, -at(q, i));
    if (t < mn) { t = mn; sq = point_aabb(add(p, mul(dir, t)), he, q); }
    else if (t > mx) { t = mx; sq = point_aabb(add(p, mul(dir, t)), he, q); }
    return sq;
}

// ------------------------------------------------------------------ clipping
constexpr int MAXPOLY = 12;

__device__ __forceinline__ bool inside(f2 pt, f2 a, f2 b) {
    return (b.y - a.y) * (pt.x - a.x) - (b.x - a.x) * (pt.y - a.y) > 0;
}
__device__ __forceinline__ f2 intersect2(f2 a1, f2 b1, f2 a2, f2 b2) {
    f2 r1 = sub2(b1, a1), r2 = sub2(b2, a2);
    float t = det2(sub2(a1, a2), r2) / det2(mk2(-r1.x, -r1.y), r2);
    return add2(a1, mk2(t * r1.x, t * r1.y));
}
// Sutherland-Hodgman against an axis-aligned rectangle given as 4 corners (Clipping.cpp:21-53)
__device__ void clip_polygon(f2* poly, int& n, const f2* clip) {
    f2 out[MAXPOLY];
    for (int e = 0; e < 4; ++e) {
        f2 a = clip[e], b = clip[(e + 1) & 3];
        int m = 0;
        for (int i = 0; i < n; ++i) {
            int j = (i + 1) == n ? 0 : i + 1;
            bool ii = inside(poly[i], a, b), ji = inside(poly[j], a, b);
            if (ii && ji) { if (m < MAXPOLY) out[m++] = poly[j]; }
            else if (ii) { if (m < MAXPOLY) out[m++] = intersect2(poly[i], poly[j], a, b); }
            else if (ji) {
                if (m < MAXPOLY) out[m++] = intersect2(poly[i], poly[j], a, b);
                if (m < MAXPOLY) out[m++] = poly[j];
            }
        }
        n = m;
        for (int i = 0; i < m; ++i) poly[i] = out[i];
    }
}
__device__ bool clip_line(f2& p0, f2& p1, const f2* clip) {
    for (int i = 0; i < 4; ++i) {
        int j = (i + 1) & 3;
        bool in0 = inside(p0, clip[i], clip[j]), in1 = inside(p1, clip[i], clip[j]);
        if (in0 && in1) continue;
        if (in0) p1 = intersect2(p0, p1, clip[i], clip[j]);
        else if (in1) p0 = intersect2(p0, p1, clip[i], clip[j]);
        else return false;
    }
    return true;
}

// src/ContactsUtil.cpp:9-107
template <int CAP>
__device__ void polygon_box_face_contacts_n(f3 boxCenter, const m3& basis, int axis, int axisSign, f3 he,
                                            f3 incOrig, f3 incNormal, const f2* poly, int n, int cx, int cy,
                                            CPt* cps, int& numPoints) {
    f3 pen[CAP];
    int np = 0;
    f3 dir = s3(0.f);
    set(dir, axis, (float)axisSign);
    float denom = dot(dir, incNormal);
    float lim = at(he, axis) + 1e-4f;
    for (int i = 0; i < n; ++i) {
        f3 p3 = s3(0.f);
        set(p3, cx, poly[i].x);
        set(p3, cy, poly[i].y);
        float distance = dot(sub(incOrig, p3), incNormal) / denom;
        if (distance < lim) {
            set(p3, axis, axisSign * distance);
            pen[np++] = p3;
        }
    }
    auto emit = [&](int slot, int idx) {
        f3 on = pen[idx];
        set(on, axis, axisSign * at(he, axis));
        cps[slot].p0 = add(boxCenter, mulmv(basis, on));
        cps[slot].p1 = add(boxCenter, mulmv(basis, pen[idx]));
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
            f3 ca = sub(pen[0], pen[i]);
            f3 cb = sub(pen[maxDi], pen[i]);
            float area = det2(mk2(at(ca, cx), at(ca, cy)), mk2(at(cb, cx), at(cb, cy)));
            if (area > maxA) { maxA = area; maxAi = i; }
        }
        emit(2, maxAi);
        for (int i = 0; i < np; ++i) {
            f3 da = sub(pen[0], pen[i]);
            f3 db = sub(pen[maxDi], pen[i]);
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

__device__ __forceinline__ void polygon_box_face_contacts(f3 boxCenter, const m3& basis, int axis, int axisSign,
                                                          f3 he, f3 incOrig, f3 incNormal, const f2* poly, int n,
                                                          int cx, int cy, CPt* cps, int& numPoints) {
    polygon_box_face_contacts_n<MAXPOLY>(boxCenter, basis, axis, axisSign, he, incOrig, incNormal, poly, n, cx, cy,
                                         cps, numPoints);
}

// ------------------------------------------------------------------ pair routines
// src/Collision.cpp:54-107
__device__ void box_face_contacts(const m3& uRef, f3 posRef, f3 heRef, const m3& uInc, f3 posInc, f3 heInc,
                                  int refAxis, CPt* cps, int& numPoints) {
    f3 axis = col(uRef, refAxis);
    float refSign = dot(axis, sub(posInc, posRef)) < 0 ? -1.f : 1.f;
    float maxDot = 0.f, incSign = 1.f;
    int incAxis = 0;
    for (int i = 0; i < 3; ++i) {
        float d = dot(mul(-refSign, axis), col(uInc, i));
        float ad = gabs(d);
        if (ad > maxDot) { maxDot = ad; incAxis = i; incSign = d < 0 ? -1.f : 1.f; }
    }
    int ia1 = (incAxis + 1) % 3, ia2 = (incAxis + 2) % 3;
    m3 inv = inverse(uRef);
    f3 ui = col(uInc, incAxis), u1 = col(uInc, ia1), u2 = col(uInc, ia2);
    float h1 = at(heInc, ia1), h2 = at(heInc, ia2);
    f3 incOrig = mulmv(inv, sub(add(posInc, mul(mul(ui, at(heInc, incAxis)), incSign)), posRef));
    f3 a = mul(u1, h1), b = mul(u2, h2);
    f3 na = mul(neg(u1), h1);
    f3 poly3[4] = {add(incOrig, mulmv(inv, add(a, b))), add(incOrig, mulmv(inv, add(na, b))),
                   add(incOrig, mulmv(inv, sub(na, b))), add(incOrig, mulmv(inv, sub(a, b)))};
    int cx = (refAxis + 1) % 3, cy = (refAxis + 2) % 3;
    f2 poly[MAXPOLY];
    for (int i = 0; i < 4; ++i) poly[i] = mk2(at(poly3[i], cx), at(poly3[i], cy));
    float hx = at(heRef, cx), hy = at(heRef, cy);
    f2 clip[4] = {mk2(hx, hy), mk2(hx, -hy), mk2(-hx, -hy), mk2(-hx, hy)};
    int n = 4;
    clip_polygon(poly, n, clip);
    f3 incNormal = mulmv(inv, ui);
    polygon_box_face_contacts(posRef, uRef, refAxis, (int)refSign, heRef, incOrig, incNormal, poly, n, cx, cy, cps,
                              numPoints);
}

__device__ bool sphere_sphere(f3 p0, float r0, f3 p1, float r1, Mf& m) {
    f3 d = sub(p1, p0);
    float rs = r0 + r1;
    if (dot(d, d) > rs * rs) return false;
    float dl = length(d);
    m.normal = dl != 0.f ? divs(d, dl) : mk3(0.f, 1.f, 0.f);
    m.numPoints = 1;
    m.pts[0].p0 = add(p0, mul(m.normal, r0));
    m.pts[0].p1 = sub(p1, mul(m.normal, r1));
    return true;
}

__device__ bool capsule_capsule(f3 p0, qf q0, float hh0, float r0, f3 p1, qf q1, float hh1, float r1, Mf& m) {
    f3 up = mk3(0.f, 1.f, 0.f);
    f3 v0 = qrot(q0, up), v1 = qrot(q1, up);
    f3 a, b;
    closest_points_segments(p0, v0, -hh0, hh0, p1, v1, -hh1, hh1, a, b);
    f3 d = sub(b, a);
    float rs = r0 + r1;
    if (dot(d, d) > rs * rs) return false;
    float dl = length(d);
    m.normal = dl != 0.f ? divs(d, dl) : up;
    m.numPoints = 1;
    m.pts[0].p0 = add(a, mul(m.normal, r0));
    m.pts[0].p1 = sub(b, mul(m.normal, r1));
    return true;
}

__device__ bool sphere_capsule(f3 p0, float r0, f3 p1, qf q1, float hh1, float r1, Mf& m) {
    f3 up = mk3(0.f, 1.f, 0.f);
    f3 dir = qrot(q1, up);
    f3 q = closest_point_on_segment(p0, p1, dir, -hh1, hh1);
    f3 v = sub(q, p0);
    float rs = r0 + r1;
    if (dot(v, v) > rs * rs) return false;
    float vl = length(v);
    m.normal = vl != 0.f ? divs(v, vl) : up;
    m.numPoints = 1;
    m.pts[0].p0 = add(p0, mul(m.normal, r0));
    m.pts[0].p1 = sub(q, mul(m.normal, r1));
    return true;
}

__device__ bool sphere_box(f3 p0, float r0, f3 p1, qf q1, f3 he, Mf& m) {
    m3 u = mat3_cast(q1);
    f3 d = sub(p0, p1);
    d = mk3(dot(d, u.c0), dot(d, u.c1), dot(d, u.c2));
    f3 q = add(p1, mulmv(u, vclamp(d, neg(he), he)));
    f3 v = sub(q, p0);
    if (dot(v, v) > r0 * r0) return false;
    float vl = length(v);
    m.numPoints = 1;
    if (vl != 0.f) {
        m.normal = divs(v, vl);
        m.pts[0].p0 = add(p0, mul(m.normal, r0));
        m.pts[0].p1 = q;
    } else {
        float mn = 3.402823466e+38f;
        int axis = 0;
        for (int i = 0; i < 3; ++i) {
            float depth = at(he, i) - gabs(at(d, i));
            if (depth < mn) { mn = depth; axis = i; }
        }
        float sign = at(d, axis) != 0.f ? gsign(at(d, axis)) : 1.f;
        f3 bp = d;
        set(bp, axis, sign * at(he, axis));
        m.normal = mul(-sign, col(u, axis));
        m.pts[0].p0 = p0;
        m.pts[0].p1 = add(p1, mulmv(u, bp));
    }
    return true;
}

__device__ void capsule_box_contacts(f3 p0L, f3 p1L, float radius, f3 normal, f3 boxCenter, f3 he,
                                     const m3& basis, Mf& m) {
    int axis = 0;
    float axisSign = 0.f, maxDot = 0.f;
    for (int i = 0; i < 3; ++i) {
        float d = dot(normal, col(basis, i));
        float ad = gabs(d);
        if (ad > maxDot) { maxDot = ad; axis = i; axisSign = d < 0 ? 1.f : -1.f; }
    }
    int cx = (axis + 1) % 3, cy = (axis + 2) % 3;
    f2 l0 = mk2(at(p0L, cx), at(p0L, cy)), l1 = mk2(at(p1L, cx), at(p1L, cy));
    float hx = at(he, cx), hy = at(he, cy);
    f2 clip[4] = {mk2(hx, hy), mk2(hx, -hy), mk2(-hx, -hy), mk2(-hx, hy)};
    if (!clip_line(l0, l1, clip)) { m.numPoints = 0; return; }
    float onAxis = axisSign * at(he, axis);
    f3 a = s3(0.f), b = s3(0.f);
    float distX = at(p0L, cx) - at(p1L, cx), distY = at(p0L, cy) - at(p1L, cy);
    float adx = gabs(distX), ady = gabs(distY);
    if (adx > ady) {
        set(a, axis, gmix(at(p0L, axis), at(p1L, axis), (at(p0L, cx) - l0.x) / distX));
        set(b, axis, gmix(at(p1L, axis), at(p0L, axis), (at(p1L, cx) - l1.x) / -distX));
    } else if (ady > adx) {
        set(a, axis, gmix(at(p0L, axis), at(p1L, axis), (at(p0L, cy) - l0.y) / distY));
        set(b, axis, gmix(at(p1L, axis), at(p0L, axis), (at(p1L, cy) - l1.y) / -distY));
    } else {
        set(a, axis, at(p0L, axis));
        set(b, axis, at(p1L, axis));
    }
    set(a, axis, at(a, axis) - axisSign * radius);
    set(b, axis, at(b, axis) - axisSign * radius);
    int np = 0;
    float ha = at(he, axis);
    if (gabs(at(a, axis)) < ha) {
        set(a, cx, l0.x); set(a, cy, l0.y);
        f3 ob = a; set(ob, axis, onAxis);
        m.pts[np].p0 = add(boxCenter, mulmv(basis, a));
        m.pts[np].p1 = add(boxCenter, mulmv(basis, ob));
        ++np;
    }
    if (gabs(at(b, axis)) < ha) {
        set(b, cx, l1.x); set(b, cy, l1.y);
        f3 ob = b; set(ob, axis, onAxis);
        m.pts[np].p0 = add(boxCenter, mulmv(basis, b));
        m.pts[np].p1 = add(boxCenter, mulmv(basis, ob));
        ++np;
    }
    m.numPoints = np;
}

__device__ bool capsule_box(f3 p0, qf q0, float hh0, float r0, f3 p1, qf q1, f3 he, Mf& m) {
    m3 u = mat3_cast(q1);
    f3 p = sub(p0, p1);
    p = mk3(dot(p, u.c0), dot(p, u.c1), dot(p, u.c2));
    f3 dir = qrot(q0, mk3(0.f, 1.f, 0.f));
    f3 dl = mk3(dot(dir, u.c0), dot(dir, u.c1), dot(dir, u.c2));
    float t = 0.f;
    f3 q = s3(0.f);
    float sq = segment_aabb(p, dl, -hh0, hh0, he, t, q);
    if (sq >= r0 * r0) return false;
    f3 onSeg = add(p0, mul(dir, t));
    f3 onBox = add(p1, mulmv(u, q));
    f3 nrm = sub(onBox, onSeg);
    float nl = length(nrm);
    if (nl != 0.f) {
        m.normal = divs(nrm, nl);
    } else {
        f3 r, ar;
        r = mk3(dot(dir, u.c0), dot(dir, u.c1), dot(dir, u.c2));
        ar = vabs(r);
        float mx = -3.402823466e+38f;
        bool edge = false;
        int axis = 0;
        for (int i = 0; i < 3; ++i) {
            float d = gabs(at(p, i)) - (hh0 * at(ar, i) + r0) - at(he, i);
            if (d > mx) { mx = d; axis = i; }
        }
        float edgeOffset = 0.1f;
        {
            float rb = he.y * ar.z + he.z * ar.y;
            float d = gabs(p.z * r.y - p.y * r.z) - r0 - rb;
            if (ar.x < 0.999f && d > mx + edgeOffset) { mx = d; edge = true; axis = 0; edgeOffset = 0.f; }
        }
        {
            float rb = he.x * ar.z + he.z * ar.x;
            float d = gabs(p.x * r.z - p.z * r.x) - r0 - rb;
            if (ar.y < 0.999f && d > mx + edgeOffset) { mx = d; edge = true; axis = 1; edgeOffset = 0.f; }
        }
        {
            float rb = he.x * ar.y + he.y * ar.x;
            float d = gabs(p.y * r.x - p.x * r.y) - r0 - rb;
            if (ar.z < 0.999f && d > mx + edgeOffset) { mx = d; edge = true; axis = 2; }
        }
        f3 sep = edge ? cross(dir, col(u, axis)) : col(u, axis);
        m.normal = normalize(dot(sep, sub(p1, p0)) > 0 ? sep : neg(sep));
    }
    capsule_box_contacts(sub(p, mul(dl, hh0)), add(p, mul(dl, hh0)), r0, m.normal, p1, he, u, m);
    if (!m.numPoints) {
        m.numPoints = 1;
        m.pts[0].p0 = add(onSeg, mul(m.normal, r0));
        m.pts[0].p1 = onBox;
    }
    return true;
}

// src/Collision.cpp:135-350
__device__ bool box_box(f3 p0, qf q0, f3 he0, f3 p1, qf q1, f3 he1, Mf& m) {
    float edgeOffset = 0.1f;
    const float edgeLimit = 0.999f;
    bool isEdge = false;
    int faceBox = 0, faceAxis = 0, e0 = 0, e1 = 0;
    float mx = -3.402823466e+38f;
    m3 u0 = mat3_cast(q0), u1 = mat3_cast(q1);
    f3 t = sub(p1, p0);
    t = mk3(dot(t, u0.c0), dot(t, u0.c1), dot(t, u0.c2));
    float r[3][3], ar[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            r[i][j] = dot(col(u0, i), col(u1, j));
            ar[i][j] = gabs(r[i][j]);
        }
    float h0[3] = {he0.x, he0.y, he0.z}, h1[3] = {he1.x, he1.y, he1.z}, tt[3] = {t.x, t.y, t.z};
    for (int i = 0; i < 3; ++i) {
        float rb = h1[0] * ar[i][0] + h1[1] * ar[i][1] + h1[2] * ar[i][2];
        float d = gabs(tt[i]) - h0[i] - rb;
        if (d > 0) return false;
        if (d > mx) { mx = d; faceBox = 0; faceAxis = i; }
    }
    for (int i = 0; i < 3; ++i) {
        float ra = h0[0] * ar[0][i] + h0[1] * ar[1][i] + h0[2] * ar[2][i];
        float l = gabs(tt[0] * r[0][i] + tt[1] * r[1][i] + tt[2] * r[2][i]);
        float d = l - ra - h1[i];
        if (d > 0) return false;
        if (d > mx) { mx = d; faceBox = 1; faceAxis = i; }
    }
    // A_i x B_j in the reference's order; each radius sum lists the lower axis first
    for (int i = 0; i < 3; ++i) {
        int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
        int lo = i1 < i2 ? i1 : i2, hi = i1 < i2 ? i2 : i1;
        for (int j = 0; j < 3; ++j) {
            int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            int jlo = j1 < j2 ? j1 : j2, jhi = j1 < j2 ? j2 : j1;
            float ra = h0[lo] * ar[lo == i1 ? i2 : i1][j] + h0[hi] * ar[hi == i1 ? i2 : i1][j];
            float rb = h1[jlo] * ar[i][jlo == j1 ? j2 : j1] + h1[jhi] * ar[i][jhi == j1 ? j2 : j1];
            float l = gabs(tt[i2] * r[i1][j] - tt[i1] * r[i2][j]);
            float d = l - ra - rb;
            if (d > 0) return false;
            if (ar[i][j] < edgeLimit && d > mx + edgeOffset) {
                mx = d; isEdge = true; e0 = i; e1 = j; edgeOffset = 0.f;
            }
        }
    }
    if (!isEdge) {
        CPt cps[4];
        int np = 0;
        f3 axis;
        if (faceBox) {
            axis = col(u1, faceAxis);
            box_face_contacts(u1, p1, he1, u0, p0, he0, faceAxis, cps, np);
            for (int i = 0; i < np; ++i) { m.pts[i].p0 = cps[i].p1; m.pts[i].p1 = cps[i].p0; }
        } else {
            axis = col(u0, faceAxis);
            box_face_contacts(u0, p0, he0, u1, p1, he1, faceAxis, cps, np);
            for (int i = 0; i < np; ++i) m.pts[i] = cps[i];
        }
        m.numPoints = np;
        m.normal = dot(axis, sub(p1, p0)) < 0 ? neg(axis) : axis;
    } else {
        f3 a0 = col(u0, e0), a1 = col(u1, e1);
        f3 axis = cross(a0, a1);
        m.normal = normalize(dot(axis, sub(p1, p0)) < 0 ? neg(axis) : axis);
        float s0[3] = {-1.f, -1.f, -1.f}, s1[3] = {-1.f, -1.f, -1.f};
        s0[(e0 + 1) % 3] = dot(col(u0, (e0 + 1) % 3), m.normal) < 0 ? -1.f : 1.f;
        s0[(e0 + 2) % 3] = dot(col(u0, (e0 + 2) % 3), m.normal) < 0 ? -1.f : 1.f;
        f3 c0 = add(add(add(p0, mul(s0[0] * h0[0], u0.c0)), mul(s0[1] * h0[1], u0.c1)), mul(s0[2] * h0[2], u0.c2));
        s1[(e1 + 1) % 3] = dot(col(u1, (e1 + 1) % 3), m.normal) > 0 ? -1.f : 1.f;
        s1[(e1 + 2) % 3] = dot(col(u1, (e1 + 2) % 3), m.normal) > 0 ? -1.f : 1.f;
        f3 c1 = add(add(add(p1, mul(s1[0] * h1[0], u1.c0)), mul(s1[1] * h1[1], u1.c1)), mul(s1[2] * h1[2], u1.c2));
        m.numPoints = 1;
        f3 a, b;
        closest_points_segments(c0, a0, 0.f, h0[e0] * 2, c1, a1, 0.f, h1[e1] * 2, a, b);
        m.pts[0].p0 = a;
        m.pts[0].p1 = b;
    }
    return true;
}

__device__ void flip(Mf& m) {
    m.normal = neg(m.normal);
    for (int i = 0; i < m.numPoints; ++i) { f3 a = m.pts[i].p0; m.pts[i].p0 = m.pts[i].p1; m.pts[i].p1 = a; }
}

// src/Collision.cpp:895-1016 for primitive shapes
__device__ bool collide(int t0, float4 g0, f3 p0, qf q0, int t1, float4 g1, f3 p1, qf q1, Mf& m) {
    m.numPoints = 0;
    bool hit = false;
    f3 he0 = mk3(g0.x, g0.y, g0.z), he1 = mk3(g1.x, g1.y, g1.z);
    if (t0 == PGX_SPHERE) {
        if (t1 == PGX_SPHERE) return sphere_sphere(p0, g0.x, p1, g1.x, m);
        if (t1 == PGX_CAPSULE) return sphere_capsule(p0, g0.x, p1, q1, g1.x, g1.y, m);
        return sphere_box(p0, g0.x, p1, q1, he1, m);
    }
    if (t0 == PGX_CAPSULE) {
        if (t1 == PGX_SPHERE) { hit = sphere_capsule(p1, g1.x, p0, q0, g0.x, g0.y, m); flip(m); return hit; }
        if (t1 == PGX_CAPSULE) return capsule_capsule(p0, q0, g0.x, g0.y, p1, q1, g1.x, g1.y, m);
        return capsule_box(p0, q0, g0.x, g0.y, p1, q1, he1, m);
    }
    if (t1 == PGX_SPHERE) { hit = sphere_box(p1, g1.x, p0, q0, he0, m); flip(m); return hit; }
    if (t1 == PGX_CAPSULE) { hit = capsule_box(p1, q1, g1.x, g1.y, p0, q0, he0, m); flip(m); return hit; }
    return box_box(p0, q0, he0, p1, q1, he1, m);
}

}  // namespace pgx

#include "np_convex.h"
#include "np_trimesh.h"

namespace pgx {

// ------------------------------------------------------------------ kernel
__device__ __forceinline__ bool in_sorted(const uint64_t* __restrict__ a, int n, uint64_t key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo < n && a[lo] == key;
}

// Index in a list of this lane's entry when the lanes with `want` append one entry each:
// one atomic on the list's counter per wave.
__device__ __forceinline__ int wave_append(int32_t* counter, bool want) {
    const uint64_t m = __ballot(want);
    const int lane = threadIdx.x & (WAVE - 1);
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader);
    return base + __popcll(m & ((1ull << lane) - 1ull));
}

// Shape-pair class of every pair: the unordered pair of primitive types (6 classes), or
// 15 for pairs that leave this kernel's collision routines (triangle or convex meshes,
// triggers, non-colliding pairs route or skip quickly). Sorted by class, a wave of the
// narrowphase runs one collision routine instead of every routine its pairs need.
__global__ void k_pair_class(int nPairs, const uint64_t* __restrict__ pkeys, int localBits,
                             const int32_t* __restrict__ worldColStart, const int32_t* __restrict__ colOfLocal,
                             const int32_t* __restrict__ ctype, uint64_t* __restrict__ cls,
                             int32_t* __restrict__ order) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nPairs) return;
    uint64_t key = pkeys[p];
    uint64_t lmask = (1ull << localBits) - 1;
    int base = worldColStart[(int)(key >> (2 * localBits))];
    int t0 = ctype[colOfLocal[base + (int)((key >> localBits) & lmask)]];
    int t1 = ctype[colOfLocal[base + (int)(key & lmask)]];
    bool prim = t0 <= PGX_BOX && t1 <= PGX_BOX;
    cls[p] = prim ? (uint64_t)(min(t0, t1) * 3 + max(t0, t1)) : 15ull;
    order[p] = p;
}

__global__ __launch_bounds__(256) void k_narrowphase(
    int nPairs, const int32_t* __restrict__ order, const uint64_t* __restrict__ pkeys, int localBits, const int32_t* __restrict__ worldColStart,
    const int32_t* __restrict__ colOfLocal, const int32_t* __restrict__ cbody, const int32_t* __restrict__ ctype,
    const int32_t* __restrict__ cflags, const float4* __restrict__ cparams, const float4* __restrict__ clpos,
    const float4* __restrict__ clquat, const uint32_t* __restrict__ centity, const float4* __restrict__ bpos,
    const float4* __restrict__ bquat, const uint64_t* __restrict__ nonColliding, int nNonColliding,
    int32_t* __restrict__ npNum, float4* __restrict__ npNormal, float4* __restrict__ npPts,
    int32_t* __restrict__ heavy, int32_t* __restrict__ nHeavy, int32_t* __restrict__ meshPairs,
    int32_t* __restrict__ nMesh, int32_t* __restrict__ meshCount, int32_t* __restrict__ trigList,
    int32_t* __restrict__ nTrig, const int32_t* __restrict__ filterClass, const uint8_t* __restrict__ filterTable,
    int filterK) {
    // pairs per lane = pairs / (grid size): a lane takes every grid-size'th pair of the
    // class order; results are stored at the pair's own slot
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nPairs; i += gridDim.x * blockDim.x) {
        const int p = order[i];
        meshCount[p] = 0;
        uint64_t key = pkeys[p];
        uint64_t lmask = (1ull << localBits) - 1;
        int wd = (int)(key >> (2 * localBits));
        int base = worldColStart[wd];
        int c0 = colOfLocal[base + (int)((key >> localBits) & lmask)];
        int c1 = colOfLocal[base + (int)(key & lmask)];
        int num = 0;
        // the contact filter on (collider of the lower entity, other collider): TRIGGER makes
        // a trigger pair, not a contact (src/Physecs.cpp:200-207); defaultContactFilter is a
        // trigger on either side (:20-23). Joint-connected / setCanCollide pairs are skipped (:209)
        bool skip = filterK ? filterTable[filterClass[c0] * filterK + filterClass[c1]] != 0
                            : ((cflags[c0] | cflags[c1]) & CF_TRIGGER) != 0;
        // trigger pairs go to k_triggers (before the non-colliding test, as in the reference)
        const bool toTrig = skip && trigList != nullptr;
        if (__ballot(toTrig)) {
            int i = wave_append(nTrig, toTrig);
            if (toTrig) trigList[i] = p;
        }
        if (!skip && nNonColliding) {
            uint64_t ek = ((uint64_t)centity[c0] << 32) | centity[c1];
            skip = in_sorted(nonColliding, nNonColliding, ek);
        }
        Mf m;
        m.numPoints = 0;
        // pairs with a triangle mesh go to k_narrowphase_mesh, pairs with a convex mesh to
        // k_narrowphase_convex: appended with one atomic per wave
        const bool toMesh = !skip && (ctype[c0] == PGX_TRIANGLE_MESH || ctype[c1] == PGX_TRIANGLE_MESH);
        const bool toHeavy = !skip && !toMesh && (ctype[c0] == PGX_CONVEX_MESH || ctype[c1] == PGX_CONVEX_MESH);
        if (__ballot(toMesh)) {
            int i = wave_append(nMesh, toMesh);
            if (toMesh) meshPairs[i] = p;
        }
        if (__ballot(toHeavy)) {
            int i = wave_append(nHeavy, toHeavy);
            if (toHeavy) heavy[i] = p;
        }
        skip = skip || toMesh || toHeavy;
        if (!skip) {
            int b0 = cbody[c0], b1 = cbody[c1];
            f3 bp0 = xyz(bpos[b0]), bp1 = xyz(bpos[b1]);
            qf bq0 = q4(bquat[b0]), bq1 = q4(bquat[b1]);
            f3 p0 = add(bp0, qrot(bq0, xyz(clpos[c0])));
            qf q0 = qmul(bq0, q4(clquat[c0]));
            f3 p1 = add(bp1, qrot(bq1, xyz(clpos[c1])));
            qf q1 = qmul(bq1, q4(clquat[c1]));
            bool hit = collide(ctype[c0], cparams[c0], p0, q0, ctype[c1], cparams[c1], p1, q1, m);
            num = hit ? m.numPoints : 0;
        }
        npNum[p] = num;
        if (num) {
            npNormal[p] = f4(m.normal, 0.f);
            for (int k = 0; k < 4; ++k) {
                if (k < num) {
                    npPts[8 * (size_t)p + 2 * k] = f4(m.pts[k].p0, 0.f);
                    npPts[8 * (size_t)p + 2 * k + 1] = f4(m.pts[k].p1, 0.f);
                }
            }
        }
    }
}

// Pairs with a convex mesh: a fixed grid of `lanes` threads walks the list; lane l owns
// EPA workspace l. The list length is read on the device, so no host round trip.
__global__ __launch_bounds__(64) void k_narrowphase_convex(
    const int32_t* __restrict__ nHeavyP, const int32_t* __restrict__ heavy, const uint64_t* __restrict__ pkeys,
    int localBits, const int32_t* __restrict__ worldColStart, const int32_t* __restrict__ colOfLocal,
    const int32_t* __restrict__ cbody, const int32_t* __restrict__ ctype, const float4* __restrict__ cparams,
    const float4* __restrict__ clpos, const float4* __restrict__ clquat, const float4* __restrict__ bpos,
    const float4* __restrict__ bquat, ConvexView cv, double* epaV, double* epaFN, int32_t* epaFI, int32_t* epaE,
    int lanes, int32_t* __restrict__ npNum, float4* __restrict__ npNormal, float4* __restrict__ npPts,
    int32_t* __restrict__ errors) {
    int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= lanes) return;
    int nHeavy = *nHeavyP;
    EpaWs ws;
    ws.vd = epaV + lane; ws.fn = epaFN + lane; ws.fi = epaFI + lane; ws.ed = epaE + lane;
    ws.L = (size_t)lanes;
    ws.overflow = false;
    uint64_t lmask = (1ull << localBits) - 1;
    for (int h = lane; h < nHeavy; h += lanes) {
        int p = heavy[h];
        uint64_t key = pkeys[p];
        int wd = (int)(key >> (2 * localBits));
        int base = worldColStart[wd];
        int c0 = colOfLocal[base + (int)((key >> localBits) & lmask)];
        int c1 = colOfLocal[base + (int)(key & lmask)];
        int b0 = cbody[c0], b1 = cbody[c1];
        qf bq0 = q4(bquat[b0]), bq1 = q4(bquat[b1]);
        f3 p0 = add(xyz(bpos[b0]), qrot(bq0, xyz(clpos[c0])));
        qf q0 = qmul(bq0, q4(clquat[c0]));
        f3 p1 = add(xyz(bpos[b1]), qrot(bq1, xyz(clpos[c1])));
        qf q1 = qmul(bq1, q4(clquat[c1]));
        Mf m;
        bool ovf = false;
        bool hit = collide_convex(ctype[c0], cparams[c0], p0, q0, ctype[c1], cparams[c1], p1, q1, cv, ws, m, ovf);
        if (ovf) atomicAdd(errors, 1);
        int num = hit ? m.numPoints : 0;
        npNum[p] = num;
        if (num) {
            npNormal[p] = f4(m.normal, 0.f);
            for (int k = 0; k < 4; ++k)
                if (k < num) {
                    npPts[8 * (size_t)p + 2 * k] = f4(m.pts[k].p0, 0.f);
                    npPts[8 * (size_t)p + 2 * k + 1] = f4(m.pts[k].p1, 0.f);
                }
        }
    }
}

// ------------------------------------------------------------------ overlap queries
// physecs::overlap (src/Overlap.cpp): the pair tests as yes/no answers, no manifold.
__device__ __forceinline__ bool ov_sphere_sphere(f3 p0, float r0, f3 p1, float r1) {   // Overlap.cpp:19-24
    f3 d = sub(p1, p0);
    float rs = r0 + r1;
    return !(dot(d, d) > rs * rs);
}

__device__ bool ov_capsule_capsule(f3 p0, qf q0, float hh0, float r0, f3 p1, qf q1, float hh1,   // :26-35
                                   float r1) {
    const f3 up = mk3(0.f, 1.f, 0.f);
    f3 a, b;
    closest_points_segments(p0, qrot(q0, up), -hh0, hh0, p1, qrot(q1, up), -hh1, hh1, a, b);
    f3 d = sub(b, a);
    float rs = r0 + r1;
    return !(dot(d, d) > rs * rs);
}

// Overlap.cpp:37-139: the 15 separating axes, separated when |t.L| - rA - rB > 0. For the
// edge axes Ai x Bj the radius sums take the other two axes of each box, lower index first.
__device__ bool ov_box_box(f3 p0, qf q0, f3 he0, f3 p1, qf q1, f3 he1) {
    const m3 u0 = mat3_cast(q0), u1 = mat3_cast(q1);
    f3 tv = sub(p1, p0);
    tv = mk3(dot(tv, u0.c0), dot(tv, u0.c1), dot(tv, u0.c2));
    float r[3][3], ar[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            r[i][j] = dot(col(u0, i), col(u1, j));
            ar[i][j] = gabs(r[i][j]);
        }
    const float a[3] = {he0.x, he0.y, he0.z}, b[3] = {he1.x, he1.y, he1.z}, t[3] = {tv.x, tv.y, tv.z};
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (gabs(t[i]) - a[i] - (b[0] * ar[i][0] + b[1] * ar[i][1] + b[2] * ar[i][2]) > 0) return false;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float ra = a[0] * ar[0][i] + a[1] * ar[1][i] + a[2] * ar[2][i];
        if (gabs(t[0] * r[0][i] + t[1] * r[1][i] + t[2] * r[2][i]) - ra - b[i] > 0) return false;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
        const int lo = i1 < i2 ? i1 : i2, hi = i1 < i2 ? i2 : i1;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            const int blo = j1 < j2 ? j1 : j2, bhi = j1 < j2 ? j2 : j1;
            float ra = a[lo] * ar[hi][j] + a[hi] * ar[lo][j];
            float rb = b[blo] * ar[i][bhi] + b[bhi] * ar[i][blo];
            if (gabs(t[i2] * r[i1][j] - t[i1] * r[i2][j]) - ra - rb > 0) return false;
        }
    }
    return true;
}

__device__ bool ov_sphere_capsule(f3 p0, float r0, f3 p1, qf q1, float hh1, float r1) {   // :145-153
    f3 q = closest_point_on_segment(p0, p1, qrot(q1, mk3(0.f, 1.f, 0.f)), -hh1, hh1);
    f3 v = sub(q, p0);
    float rs = r0 + r1;
    return !(dot(v, v) > rs * rs);
}

__device__ bool ov_sphere_box(f3 p0, float r0, f3 p1, qf q1, f3 he) {   // :155-169
    m3 u = mat3_cast(q1);
    f3 d = sub(p0, p1);
    d = mk3(dot(d, u.c0), dot(d, u.c1), dot(d, u.c2));
    f3 q = add(p1, mulmv(u, vclamp(d, neg(he), he)));
    f3 v = sub(q, p0);
    return !(dot(v, v) > r0 * r0);
}

// :175-193; separated when the squared distance reaches r^2 (>=, where the others use >)
__device__ bool ov_capsule_box(f3 p0, qf q0, float hh0, float r0, f3 p1, qf q1, f3 he) {
    m3 u = mat3_cast(q1);
    f3 p = sub(p0, p1);
    p = mk3(dot(p, u.c0), dot(p, u.c1), dot(p, u.c2));
    f3 dir = qrot(q0, mk3(0.f, 1.f, 0.f));
    f3 dl = mk3(dot(dir, u.c0), dot(dir, u.c1), dot(dir, u.c2));
    float t = 0.f;
    f3 q = s3(0.f);
    return !(segment_aabb(p, dl, -hh0, hh0, he, t, q) >= r0 * r0);
}

__device__ __forceinline__ SupConvex ov_convex(const ConvexView& cv, float4 g, f3 pos, qf q) {
    const int4 mi = cv.mesh[(int)g.w];
    return SupConvex{pos, mat3_cast(q), mk3(g.x, g.y, g.z), cv.verts + mi.x, mi.y};
}

// physecs::overlap (src/Overlap.cpp:203-235). A triangle mesh on either side never
// overlaps. The convex cases run GJK (Overlap.cpp:141-200); the box-convex case places the
// convex support at the box's position, as the reference does (Overlap.cpp:200).
__device__ bool overlap_shapes(int t0, float4 g0, f3 p0, qf q0, int t1, float4 g1, f3 p1, qf q1,
                               const ConvexView& cv) {
    if (t0 > PGX_CONVEX_MESH || t1 > PGX_CONVEX_MESH) return false;
    const f3 he0 = mk3(g0.x, g0.y, g0.z), he1 = mk3(g1.x, g1.y, g1.z);
    const f3 up = mk3(0.f, 1.f, 0.f);
    GjkV s[4];
    switch (t0) {
        case PGX_SPHERE:
            switch (t1) {
                case PGX_SPHERE: return ov_sphere_sphere(p0, g0.x, p1, g1.x);
                case PGX_CAPSULE: return ov_sphere_capsule(p0, g0.x, p1, q1, g1.x, g1.y);
                case PGX_BOX: return ov_sphere_box(p0, g0.x, p1, q1, he1);
                default: return gjk(SupSphere{p0, g0.x}, ov_convex(cv, g1, p1, q1), sub(p1, p0), s);
            }
        case PGX_CAPSULE:
            switch (t1) {
                case PGX_SPHERE: return ov_sphere_capsule(p1, g1.x, p0, q0, g0.x, g0.y);
                case PGX_CAPSULE: return ov_capsule_capsule(p0, q0, g0.x, g0.y, p1, q1, g1.x, g1.y);
                case PGX_BOX: return ov_capsule_box(p0, q0, g0.x, g0.y, p1, q1, he1);
                default:
                    return gjk(SupCapsule{p0, qrot(q0, up), g0.x, g0.y}, ov_convex(cv, g1, p1, q1), sub(p1, p0), s);
            }
        case PGX_BOX:
            switch (t1) {
                case PGX_SPHERE: return ov_sphere_box(p1, g1.x, p0, q0, he0);
                case PGX_CAPSULE: return ov_capsule_box(p1, q1, g1.x, g1.y, p0, q0, he0);
                case PGX_BOX: return ov_box_box(p0, q0, he0, p1, q1, he1);
                default: return gjk(SupBox{p0, mat3_cast(q0), he0}, ov_convex(cv, g1, p0, q1), sub(p1, p0), s);
            }
        default:
            switch (t1) {
                case PGX_SPHERE: return gjk(SupSphere{p1, g1.x}, ov_convex(cv, g0, p0, q0), sub(p0, p1), s);
                case PGX_CAPSULE:
                    return gjk(SupCapsule{p1, qrot(q1, up), g1.x, g1.y}, ov_convex(cv, g0, p0, q0), sub(p0, p1), s);
                case PGX_BOX: return gjk(SupBox{p1, mat3_cast(q1), he1}, ov_convex(cv, g0, p1, q0), sub(p0, p1), s);
                default: return gjk(ov_convex(cv, g0, p0, q0), ov_convex(cv, g1, p1, q1), sub(p1, p0), s);
            }
    }
}

// the primitive pairs of physecs::overlap (sphere, capsule, box on both sides), no GJK
__device__ bool overlap_prims(int t0, float4 g0, f3 p0, qf q0, int t1, float4 g1, f3 p1, qf q1) {
    const f3 he0 = mk3(g0.x, g0.y, g0.z), he1 = mk3(g1.x, g1.y, g1.z);
    if (t0 == PGX_SPHERE) {
        if (t1 == PGX_SPHERE) return ov_sphere_sphere(p0, g0.x, p1, g1.x);
        if (t1 == PGX_CAPSULE) return ov_sphere_capsule(p0, g0.x, p1, q1, g1.x, g1.y);
        return ov_sphere_box(p0, g0.x, p1, q1, he1);
    }
    if (t0 == PGX_CAPSULE) {
        if (t1 == PGX_SPHERE) return ov_sphere_capsule(p1, g1.x, p0, q0, g0.x, g0.y);
        if (t1 == PGX_CAPSULE) return ov_capsule_capsule(p0, q0, g0.x, g0.y, p1, q1, g1.x, g1.y);
        return ov_capsule_box(p0, q0, g0.x, g0.y, p1, q1, he1);
    }
    if (t1 == PGX_SPHERE) return ov_sphere_box(p1, g1.x, p0, q0, he0);
    if (t1 == PGX_CAPSULE) return ov_capsule_box(p1, q1, g1.x, g1.y, p0, q0, he0);
    return ov_box_box(p0, q0, he0, p1, q1, he1);
}

// intersects (src/BoundsUtil.cpp:87-92): inclusive on every axis
__device__ __forceinline__ bool boxes_meet(f3 amn, f3 amx, float4 bmn, float4 bmx) {
    return !(amx.x < bmn.x || amn.x > bmx.x || amx.y < bmn.y || amn.y > bmx.y || amx.z < bmn.z || amn.z > bmx.z);
}

// one collider of the query's world: the reference's leaf test (bounds, filter, shape).
// 0: no overlap; 1: overlap; 2: a pair with a convex mesh, left to k_overlap_convex (the
// walk keeps no GJK state, so it needs far fewer registers)
__device__ __forceinline__ int overlap_collider(const OverlapArgs& A, int c, int t, float4 g, f3 pos, qf q, f3 mn,
                                                f3 mx, int filter) {
    if (!boxes_meet(mn, mx, A.cbmin[c], A.cbmax[c])) return 0;
    if (filter && !(A.cdata[c] & filter)) return 0;
    const int ct = A.ctype[c];
    if (ct == PGX_TRIANGLE_MESH) return 0;   // overlaps nothing (src/Overlap.cpp:203-235)
    if (ct == PGX_CONVEX_MESH || t == PGX_CONVEX_MESH) return 2;
    const int b = A.cbody[c];
    const qf bq = q4(A.bquat[b]);
    const f3 cp = add(xyz(A.bpos[b]), qrot(bq, xyz(A.clpos[c])));
    return overlap_prims(t, g, pos, q, ct, A.cparams[c], cp, qmul(bq, q4(A.clquat[c]))) ? 1 : 0;
}

// One query per thread: its bounds, the walk of its world's LBVH (boxes tested inclusive,
// as the reference's recursion does) and its world's static colliders. Each hit appends the
// key (query << cbits | collider) with one atomic per wave; keys beyond the capacity are
// counted, not written. count[i] = the query's hits.
__global__ __launch_bounds__(64) void k_overlap(OverlapArgs A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < A.n;
    int t = PGX_TRIANGLE_MESH, filter = 0, wd = 0;
    float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
    f3 pos = s3(0.f), mn = s3(0.f), mx = s3(0.f);
    qf q = q4(make_float4(0.f, 0.f, 0.f, 1.f));
    if (live) {
        t = A.qType[i];
        g = A.qPrm[i];
        const float4 pw = A.qPos[i];
        pos = xyz(pw);
        wd = __float_as_int(pw.w);
        q = q4(A.qQuat[i]);
        filter = A.qFilter[i];
        if (t <= PGX_CONVEX_MESH)
            shape_bounds(t, g, pos, q, A.cv.verts, A.cv.mesh, nullptr, nullptr, mn, mx);
    }
    // a triangle-mesh query overlaps nothing (src/Overlap.cpp:203-235)
    const bool active = live && t <= PGX_CONVEX_MESH;
    int hits = 0;
    int node = active ? A.worldRoot[wd] : -1;
    const int stop = active && node != -1 ? A.worldStop[wd] : -1;
    int visits = 0;
    int st = active ? A.worldStaticStart[wd] : 0;
    const int stEnd = active ? A.worldStaticStart[wd + 1] : 0;
    // the walk and then the static list, in one loop so that the wave's lanes append together
    while (true) {
        int c = -1;
        bool more = false;
        if (node != -1 && node != stop && visits < 2 * A.nLeaves) {   // a walk visits at most 2n - 1 nodes
            ++visits;
            more = true;
            if (node & LEAF_BIT) {
                const int l = node & ~LEAF_BIT;
                c = A.mvals[l];
                node = __float_as_int(A.leafA[l].w);
            } else {
                const float4 na = A.nodeA[node], nb = A.nodeB[node];
                node = boxes_meet(mn, mx, na, nb) ? __float_as_int(na.w) : __float_as_int(nb.w);
            }
        } else if (st < stEnd) {
            more = true;
            c = A.staticCols[st++];
        }
        const int r = c >= 0 ? overlap_collider(A, c, t, g, pos, q, mn, mx, filter) : 0;
        const bool hit = r == 1, toCvx = r == 2;
        if (__ballot(hit)) {
            const int slot = wave_append(A.total, hit);
            if (hit && slot < A.cap) A.keys[slot] = ((uint64_t)(uint32_t)i << A.cbits) | (uint32_t)c;
        }
        if (__ballot(toCvx)) {
            const int k = wave_append(A.nCvx, toCvx);
            if (toCvx && k < A.listCap) A.cvxPairs[k] = make_int2(i, c);
        }
        hits += hit;
        if (!__ballot(more)) break;
    }
    if (live) A.count[i] = hits;
}

// Trigger pairs: physecs::overlap of the two colliders (src/Physecs.cpp:199-206); the keys
// of the overlapping pairs are appended with one atomic per wave. A fixed grid walks the
// list, whose length is read on the device.
__global__ __launch_bounds__(64) void k_triggers(
    const int32_t* __restrict__ nTrigP, const int32_t* __restrict__ trigList, const uint64_t* __restrict__ pkeys,
    int localBits, const int32_t* __restrict__ worldColStart, const int32_t* __restrict__ colOfLocal,
    const int32_t* __restrict__ cbody, const int32_t* __restrict__ ctype, const float4* __restrict__ cparams,
    const float4* __restrict__ clpos, const float4* __restrict__ clquat, const float4* __restrict__ bpos,
    const float4* __restrict__ bquat, ConvexView cv, int32_t* __restrict__ nHits, uint64_t* __restrict__ hitKeys) {
    const int nTrig = *nTrigP;
    const uint64_t lmask = (1ull << localBits) - 1;
    // blocks are one wave, so the trip count is wave-uniform and the appends see whole waves
    for (int base = blockIdx.x * blockDim.x; base < nTrig; base += gridDim.x * blockDim.x) {
        const int h = base + (int)threadIdx.x;
        bool hit = false;
        uint64_t key = 0;
        if (h < nTrig) {
            key = pkeys[trigList[h]];
            const int b = worldColStart[(int)(key >> (2 * localBits))];
            const int c0 = colOfLocal[b + (int)((key >> localBits) & lmask)];
            const int c1 = colOfLocal[b + (int)(key & lmask)];
            const int b0 = cbody[c0], b1 = cbody[c1];
            const qf bq0 = q4(bquat[b0]), bq1 = q4(bquat[b1]);
            const f3 p0 = add(xyz(bpos[b0]), qrot(bq0, xyz(clpos[c0])));
            const f3 p1 = add(xyz(bpos[b1]), qrot(bq1, xyz(clpos[c1])));
            hit = overlap_shapes(ctype[c0], cparams[c0], p0, qmul(bq0, q4(clquat[c0])), ctype[c1], cparams[c1], p1,
                                 qmul(bq1, q4(clquat[c1])), cv);
        }
        if (__ballot(hit)) {
            const int slot = wave_append(nHits, hit);
            if (hit) hitKeys[slot] = key;
        }
    }
}

// ------------------------------------------------------------------ triangle-mesh pairs
// Everything the three mesh passes read.
struct MeshArgs {
    const int32_t* nMeshP;                 // number of mesh pairs (device)
    const int32_t* meshPairs;              // their pair slots
    const uint64_t* pkeys;
    int localBits;
    const int32_t *worldColStart, *colOfLocal, *cbody, *ctype;
    const float4 *cparams, *clpos, *clquat, *bpos, *bquat;
    ConvexView cv;
    TriMeshView tv;
    int32_t* errors;                       // capacity and EPA errors
};

struct MeshPair {
    int p, tb;
    bool flip;
    float4 g;
    f3 pos0, pos1, lpos;
    qf or0, or1, lor;
    int4 mi;
};

// mesh pair h: the body collider and the mesh, their world poses, the body in the mesh's
// frame. src/Collision.cpp:897-906: a mesh on side 0 is solved as (other, mesh) and flipped.
// False for a mesh-mesh pair (no routine).
__device__ bool mesh_pair(const MeshArgs& A, int h, MeshPair& m, SupConvex& sc) {
    m.p = A.meshPairs[h];
    uint64_t key = A.pkeys[m.p];
    uint64_t lmask = (1ull << A.localBits) - 1;
    int wd = (int)(key >> (2 * A.localBits));
    int base = A.worldColStart[wd];
    int c0 = A.colOfLocal[base + (int)((key >> A.localBits) & lmask)];
    int c1 = A.colOfLocal[base + (int)(key & lmask)];
    int t0 = A.ctype[c0], t1 = A.ctype[c1];
    if (t0 == PGX_TRIANGLE_MESH && t1 == PGX_TRIANGLE_MESH) return false;
    m.flip = t0 == PGX_TRIANGLE_MESH;
    int cb = m.flip ? c1 : c0, cm = m.flip ? c0 : c1;
    int bb = A.cbody[cb], bm = A.cbody[cm];
    qf qb = q4(A.bquat[bb]), qm = q4(A.bquat[bm]);
    m.pos0 = add(xyz(A.bpos[bb]), qrot(qb, xyz(A.clpos[cb])));
    m.or0 = qmul(qb, q4(A.clquat[cb]));
    m.pos1 = add(xyz(A.bpos[bm]), qrot(qm, xyz(A.clpos[cm])));
    m.or1 = qmul(qm, q4(A.clquat[cm]));
    m.tb = A.ctype[cb];
    m.g = A.cparams[cb];
    m.mi = A.tv.mesh[(int)A.cparams[cm].w];
    mesh_local(m.tb, m.g, m.pos0, m.or0, m.pos1, m.or1, A.cv, m.lpos, m.lor, sc);
    return true;
}

// exclusive prefix of n over the wave and the wave's total
__device__ __forceinline__ int wave_scan(int n, int& total) {
    const int lane = threadIdx.x & (WAVE - 1);
    int incl = n;
    for (int o = 1; o < WAVE; o <<= 1) {
        int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    total = __shfl(incl, WAVE - 1);
    return incl - n;
}

// pass 1, one lane per mesh pair: the BVH query, and a block of candidate slots (triangle,
// pair) reserved for the pair with one atomic per wave
__global__ __launch_bounds__(64) void k_mesh_query(MeshArgs A, int32_t* tmWs, int lanes, int32_t* __restrict__ candTotal,
                                                   int candCap, int32_t* __restrict__ cand,
                                                   int32_t* __restrict__ pairOf, int32_t* __restrict__ candOff,
                                                   int32_t* __restrict__ candCount, int32_t* __restrict__ meshCount) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    const int nMesh = *A.nMeshP;
    MeshWs ws{tmWs + lane, (size_t)lanes};
    for (int h0 = lane & ~(WAVE - 1); h0 < nMesh; h0 += lanes) {   // the same trip count for a whole wave
        const int h = h0 + (lane & (WAVE - 1));
        int nc = 0;
        if (h < nMesh) {
            MeshPair m;
            SupConvex sc;
            meshCount[A.meshPairs[h]] = 0;
            if (mesh_pair(A, h, m, sc)) {
                MeshCtx mc{&A.tv, m.mi};
                nc = mesh_query(m.tb, m.g, m.lpos, m.lor, A.cv, mc, ws);
                if (nc < 0) { atomicAdd(A.errors, 1); nc = 0; }
            }
        }
        int total;
        int off = wave_scan(nc, total);
        int base = 0;
        if ((lane & (WAVE - 1)) == WAVE - 1 && total) base = atomicAdd(candTotal, total);
        base = __shfl(base, WAVE - 1) + off;
        if (h < nMesh) {
            if (base + nc > candCap) { atomicAdd(A.errors, 1); nc = 0; }
            for (int i = 0; i < nc; ++i) {
                cand[base + i] = ws.cand(i);
                pairOf[base + i] = h;
            }
            candOff[h] = base;
            candCount[h] = nc;
        }
    }
}

// pass 2, one lane per (pair, candidate triangle): the per-shape test, its contact record
__global__ __launch_bounds__(64) void k_mesh_test(MeshArgs A, double* epaV, double* epaFN, int32_t* epaFI,
                                                  int32_t* epaE, int epaLanes, const int32_t* __restrict__ candTotal,
                                                  int candCap, const int32_t* __restrict__ cand,
                                                  const int32_t* __restrict__ pairOf, MeshRecs recs) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= epaLanes) return;
    const int total = min(*candTotal, candCap);
    EpaWs ews;
    ews.vd = epaV + lane; ews.fn = epaFN + lane; ews.fi = epaFI + lane; ews.ed = epaE + lane;
    ews.L = (size_t)epaLanes;
    for (int k = lane; k < total; k += epaLanes) {
        MeshPair m;
        SupConvex sc;
        mesh_pair(A, pairOf[k], m, sc);
        MeshCtx mc{&A.tv, m.mi};
        ews.overflow = false;
        TCI ci;
        bool epaErr = false;
        bool hit = mesh_test(m.tb, m.g, m.lpos, m.lor, sc, mc, cand[k], ews, ci, epaErr);
        if (epaErr) atomicAdd(A.errors, 1);
        recs.store(k, hit, ci);
    }
}

// pass 3, one lane per mesh pair: the pair's hit records in candidate order, a block of
// manifold slots (one atomic per wave), the emission passes
__global__ __launch_bounds__(64) void k_mesh_emit(MeshArgs A, int32_t* tmWs, int lanes,
                                                  const int32_t* __restrict__ candOff,
                                                  const int32_t* __restrict__ candCount, MeshRecs recs,
                                                  int32_t* __restrict__ npNum, float4* __restrict__ npNormal,
                                                  float4* __restrict__ npPts, int32_t* __restrict__ npTri,
                                                  int32_t* __restrict__ meshFirst, int32_t* __restrict__ meshCount,
                                                  int32_t* __restrict__ slotCounter, int mmCap, int pairCap) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    const int nMesh = *A.nMeshP;
    MeshWs ws{tmWs + lane, (size_t)lanes};
    for (int h0 = lane & ~(WAVE - 1); h0 < nMesh; h0 += lanes) {
        const int h = h0 + (lane & (WAVE - 1));
        int nct = 0;
        MeshPair m;
        SupConvex sc;
        bool ok = h < nMesh && mesh_pair(A, h, m, sc);
        if (ok) {
            const int off = candOff[h], nc = candCount[h];
            for (int i = 0; i < nc; ++i)
                if (recs.hit(off + i)) ws.order(nct++) = off + i;
        }
        int total;
        int pre = wave_scan(nct, total);
        int base = 0;
        if ((lane & (WAVE - 1)) == WAVE - 1 && total) base = atomicAdd(slotCounter, total);
        const int slot = __shfl(base, WAVE - 1) + pre;
        if (!ok || nct == 0) continue;
        if (slot + nct > mmCap) { atomicAdd(A.errors, 1); continue; }
        MeshCtx mc{&A.tv, m.mi};
        MeshOut out{npNum, npNormal, npPts, npTri, pairCap + slot, slot, 0, m.flip};
        mesh_emit(m.tb, m.g, m.lpos, m.lor, m.pos1, m.or1, mc, A.cv, ws, recs, nct, out);
        meshFirst[m.p] = slot;
        meshCount[m.p] = out.n;
    }
}

// ------------------------------------------------------------------ MTD overlap queries
// Scene::overlapWithMinTranslationalDistance (src/Physecs.cpp:653-688): collision(collider,
// query) for every collider whose bounds meet the query's; each manifold with points gives
// its normal and mtd. Results are appended with keys (query, collider, manifold index).
constexpr int MTD_MBITS = 8;   // manifold index within a pair: at most TM_MAXCAND

__device__ __forceinline__ float manifold_mtd(const Mf& m) {   // src/Physecs.cpp:663-670
    float mtd = 0;
    for (int i = 0; i < m.numPoints; i++) {
        float dist = dot(sub(m.pts[i].p0, m.pts[i].p1), m.normal);
        if (dist > mtd) mtd = dist;
    }
    return mtd;
}

__device__ __forceinline__ void mtd_put(const MtdArgs& A, int slot, int q, int c, int k, const Mf& m) {
    if (slot >= A.cap) return;
    A.keys[slot] = ((uint64_t)(uint32_t)q << (A.cbits + MTD_MBITS)) | ((uint64_t)(uint32_t)c << MTD_MBITS) |
                   (uint64_t)k;
    A.vals[slot] = slot;
    A.res[slot] = f4(m.normal, manifold_mtd(m));
}

__device__ __forceinline__ void mtd_query(const MtdArgs& A, int i, int& t, float4& g, f3& pos, qf& q, int& wd) {
    t = A.qType[i];
    g = A.qPrm[i];
    const float4 pw = A.qPos[i];
    pos = xyz(pw);
    wd = __float_as_int(pw.w);
    q = q4(A.qQuat[i]);
}

__device__ __forceinline__ void mtd_collider_pose(const MtdArgs& A, int c, f3& cp, qf& cq) {
    const int b = A.cbody[c];
    const qf bq = q4(A.bquat[b]);
    cp = add(xyz(A.bpos[b]), qrot(bq, xyz(A.clpos[c])));
    cq = qmul(bq, q4(A.clquat[c]));
}

// pass 1, one lane per query: the walk of its world's LBVH and static list (as k_overlap);
// pairs of primitives are solved here, pairs with a convex or triangle mesh are listed
__global__ __launch_bounds__(64) void k_mtd_walk(MtdArgs A, ConvexView cv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < A.n;
    int t = PGX_SPHERE, wd = 0;
    float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
    f3 pos = s3(0.f), mn = s3(0.f), mx = s3(0.f);
    qf q = q4(make_float4(0.f, 0.f, 0.f, 1.f));
    if (live) {
        mtd_query(A, i, t, g, pos, q, wd);
        shape_bounds(t, g, pos, q, cv.verts, cv.mesh, nullptr, nullptr, mn, mx);
    }
    int node = live ? A.worldRoot[wd] : -1;
    const int stop = live && node != -1 ? A.worldStop[wd] : -1;
    int visits = 0;
    int st = live ? A.worldStaticStart[wd] : 0;
    const int stEnd = live ? A.worldStaticStart[wd + 1] : 0;
    while (true) {
        int c = -1;
        bool more = false;
        if (node != -1 && node != stop && visits < 2 * A.nLeaves) {   // a walk visits at most 2n - 1 nodes
            ++visits;
            more = true;
            if (node & LEAF_BIT) {
                const int l = node & ~LEAF_BIT;
                c = A.mvals[l];
                node = __float_as_int(A.leafA[l].w);
            } else {
                const float4 na = A.nodeA[node], nb = A.nodeB[node];
                node = boxes_meet(mn, mx, na, nb) ? __float_as_int(na.w) : __float_as_int(nb.w);
            }
        } else if (st < stEnd) {
            more = true;
            c = A.staticCols[st++];
        }
        bool hit = false, toCvx = false, toMesh = false;
        Mf m;
        m.numPoints = 0;
        if (c >= 0 && boxes_meet(mn, mx, A.cbmin[c], A.cbmax[c])) {
            const int ct = A.ctype[c];
            if (ct == PGX_TRIANGLE_MESH) {
                toMesh = true;
            } else if (ct == PGX_CONVEX_MESH || t == PGX_CONVEX_MESH) {
                toCvx = true;
            } else {   // collision(collider, query), src/Physecs.cpp:659
                f3 cp;
                qf cq;
                mtd_collider_pose(A, c, cp, cq);
                hit = collide(ct, A.cparams[c], cp, cq, t, g, pos, q, m) && m.numPoints > 0;
            }
        }
        if (__ballot(hit)) {
            const int slot = wave_append(A.counts, hit);
            if (hit) mtd_put(A, slot, i, c, 0, m);
        }
        if (__ballot(toCvx)) {
            const int k = wave_append(A.counts + 1, toCvx);
            if (toCvx && k < A.listCap) A.convexPairs[k] = make_int2(i, c);
        }
        if (__ballot(toMesh)) {
            const int k = wave_append(A.counts + 2, toMesh);
            if (toMesh && k < A.listCap) A.meshPairs[k] = make_int2(i, c);
        }
        if (!__ballot(more)) break;
    }
}

// pass 2: the pairs with a convex mesh, a fixed grid of `lanes` threads; lane l owns EPA
// workspace l
__global__ __launch_bounds__(64) void k_mtd_convex(MtdArgs A, ConvexView cv, double* epaV, double* epaFN,
                                                   int32_t* epaFI, int32_t* epaE, int lanes) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= lanes) return;
    const int n = min(A.counts[1], A.listCap);
    EpaWs ws;
    ws.vd = epaV + lane; ws.fn = epaFN + lane; ws.fi = epaFI + lane; ws.ed = epaE + lane;
    ws.L = (size_t)lanes;
    ws.overflow = false;
    for (int h = lane; h < n; h += lanes) {
        const int2 qc = A.convexPairs[h];
        int t, wd;
        float4 g;
        f3 pos, cp;
        qf q, cq;
        mtd_query(A, qc.x, t, g, pos, q, wd);
        mtd_collider_pose(A, qc.y, cp, cq);
        Mf m;
        bool ovf = false;
        const bool hit = collide_convex(A.ctype[qc.y], A.cparams[qc.y], cp, cq, t, g, pos, q, cv, ws, m, ovf);
        if (ovf) atomicAdd(A.counts + 3, 1);
        if (hit && m.numPoints > 0) mtd_put(A, atomicAdd(A.counts, 1), qc.x, qc.y, 0, m);
    }
}

// the manifolds of one (query, mesh) pair: the mesh is the collider, geom0, so the pair is
// solved as (query, mesh) and flipped (src/Collision.cpp:897-901)
struct MtdOut {
    const MtdArgs* A;
    int q, c, n;
    __device__ void emit(Mf& m, int) {
        if (m.numPoints <= 0) return;
        m.normal = neg(m.normal);
        for (int i = 0; i < m.numPoints; ++i) { f3 t = m.pts[i].p0; m.pts[i].p0 = m.pts[i].p1; m.pts[i].p1 = t; }
        mtd_put(*A, atomicAdd(A->counts, 1), q, c, n, m);
        ++n;
    }
};

// pass 3: the pairs with a triangle mesh, one lane per pair: the mesh BVH query, the
// per-triangle tests into the lane's own block of TM_MAXCAND contact records (a buffer of
// recLanes * TM_MAXCAND records), the emission passes (the step's three mesh passes in one
// lane)
__global__ __launch_bounds__(64) void k_mtd_mesh(MtdArgs A, ConvexView cv, TriMeshView tv, int32_t* tmWs,
                                                 int wsLanes, MeshRecs recs, double* epaV, double* epaFN,
                                                 int32_t* epaFI, int32_t* epaE, int epaLanes) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= A.recLanes) return;
    const int n = min(A.counts[2], A.listCap);
    MeshWs ws{tmWs + lane, (size_t)wsLanes};
    EpaWs ews;
    ews.vd = epaV + lane; ews.fn = epaFN + lane; ews.fi = epaFI + lane; ews.ed = epaE + lane;
    ews.L = (size_t)epaLanes;
    ews.overflow = false;
    for (int h = lane; h < n; h += A.recLanes) {
        const int2 qc = A.meshPairs[h];
        int t, wd;
        float4 g;
        f3 pos, mpos, lpos;
        qf q, mq, lor;
        mtd_query(A, qc.x, t, g, pos, q, wd);
        mtd_collider_pose(A, qc.y, mpos, mq);
        SupConvex sc;
        mesh_local(t, g, pos, q, mpos, mq, cv, lpos, lor, sc);
        const MeshCtx mc{&tv, tv.mesh[(int)A.cparams[qc.y].w]};
        const int nc = mesh_query(t, g, lpos, lor, cv, mc, ws);
        if (nc < 0) { atomicAdd(A.counts + 3, 1); continue; }
        const int base = lane * TM_MAXCAND;
        int nct = 0;
        for (int j = 0; j < nc; ++j) {
            TCI ci;
            bool epaErr = false;
            ews.overflow = false;
            const bool hit = mesh_test(t, g, lpos, lor, sc, mc, ws.cand(j), ews, ci, epaErr);
            if (epaErr) atomicAdd(A.counts + 3, 1);
            recs.store(base + j, hit, ci);
            if (hit) ws.order(nct++) = base + j;
        }
        MtdOut out{&A, qc.x, qc.y, 0};
        mesh_emit(t, g, lpos, lor, mpos, mq, mc, cv, ws, recs, nct, out);
    }
}

void np_mtd_launch(World& w, const MtdArgs& A) {
    const ConvexView cv{w.cvVerts.p, w.cvMesh.p, w.cvFace.p, w.cvFaceIdx.p, w.cvFaceN.p, w.cvFaceC.p};
    hipLaunchKernelGGL(k_mtd_walk, dim3(div_up(A.n, 64)), dim3(64), 0, w.stream, A, cv);
    PGX_LAUNCH_CHECK();
    if (w.nConvex > 0) {
        hipLaunchKernelGGL(k_mtd_convex, dim3(div_up(w.epaLanes, 64)), dim3(64), 0, w.stream, A, cv, w.epaV.p,
                           w.epaFN.p, w.epaFI.p, w.epaE.p, w.epaLanes);
        PGX_LAUNCH_CHECK();
    }
    if (w.nTriMesh > 0) {
        const TriMeshView tv{w.tmVerts.p, w.tmTris.p, w.tmTriN.p, w.tmTriC.p, w.tmNodes.p, w.tmMesh.p};
        const MeshRecs recs{A.recs, (size_t)A.recLanes * TM_MAXCAND};
        hipLaunchKernelGGL(k_mtd_mesh, dim3(div_up(A.recLanes, 64)), dim3(64), 0, w.stream, A, cv, tv, w.tmWs.p,
                           w.meshLanes, recs, w.epaV.p, w.epaFN.p, w.epaFI.p, w.epaE.p, w.epaLanes);
        PGX_LAUNCH_CHECK();
    }
}

// the overlap queries' pairs with a convex mesh (GJK), a fixed grid of `lanes` threads;
// each hit appends its key and adds to its query's count
__global__ __launch_bounds__(64) void k_overlap_convex(OverlapArgs A, int lanes) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= lanes) return;
    const int n = min(*A.nCvx, A.listCap);
    for (int h = lane; h < n; h += lanes) {
        const int2 qc = A.cvxPairs[h];
        const int i = qc.x, c = qc.y;
        const float4 pw = A.qPos[i];
        const int b = A.cbody[c];
        const qf bq = q4(A.bquat[b]);
        const f3 cp = add(xyz(A.bpos[b]), qrot(bq, xyz(A.clpos[c])));
        if (!overlap_shapes(A.qType[i], A.qPrm[i], xyz(pw), q4(A.qQuat[i]), A.ctype[c], A.cparams[c], cp,
                            qmul(bq, q4(A.clquat[c])), A.cv))
            continue;
        const int slot = atomicAdd(A.total, 1);
        if (slot < A.cap) A.keys[slot] = ((uint64_t)(uint32_t)i << A.cbits) | (uint32_t)c;
        atomicAdd(A.count + i, 1);
    }
}

void np_overlap_launch(hipStream_t s, const OverlapArgs& A, bool convex) {
    hipLaunchKernelGGL(k_overlap, dim3(div_up(A.n, 64)), dim3(64), 0, s, A);
    PGX_LAUNCH_CHECK();
    if (convex) {
        constexpr int lanes = 16384;
        hipLaunchKernelGGL(k_overlap_convex, dim3(lanes / 64), dim3(64), 0, s, A, lanes);
        PGX_LAUNCH_CHECK();
    }
}

void np_run(World& w, int nPairs) {
    if (w.hasTriggers) {   // this step's list: its pair and hit counts start at zero
        // every trigger pair and hit has a pair slot: the lists must hold pairCap entries
        if (w.trigList.n < (size_t)w.pairCap || w.trigKeys[0].n < (size_t)w.pairCap ||
            w.trigKeys[1].n < (size_t)w.pairCap || w.trigN.n < 4)
            throw PgxError{PGX_ERR_HIP, "trigger lists smaller than the pair capacity"};
        w.trigCur ^= 1;
        fill_i32(w.stream, w.trigN.p + 2 * w.trigCur, 0, 2);
    }
    if (nPairs == 0) return;
    const int ppl = std::max(1, w.tune.pairsPerLane);
    w.npClass.alloc(w.pairCap);
    w.npOrder.alloc(w.pairCap);
    hipLaunchKernelGGL(k_pair_class, dim3(div_up(nPairs, 256)), dim3(256), 0, w.stream, nPairs, w.pkeys.p, w.localBits,
                       w.worldColStart.p, w.colOfLocal.p, w.ctype.p, w.npClass.p, w.npOrder.p);
    PGX_LAUNCH_CHECK();
    radix_sort_u64(w.stream, w.npClass.p, w.npOrder.p, nPairs, 4, w.sortT);   // stable: slot order within a class
    hipLaunchKernelGGL(k_narrowphase, dim3(div_up(nPairs, 256 * ppl)), dim3(256), 0, w.stream, nPairs, w.npOrder.p,
                       w.pkeys.p,
                       w.localBits, w.worldColStart.p, w.colOfLocal.p, w.cbody.p, w.ctype.p, w.cflags.p,
                       w.cparams.p, w.clpos.p, w.clquat.p, w.centity.p, w.bpos.p, w.bquat.p, w.nonColliding.p,
                       w.nNonColliding, w.npNum.p, w.npNormal.p, w.npPts.p, w.heavy.p, w.counters.p + 8,
                       w.meshPairs.p, w.counters.p + 10, w.meshCount.p,
                       w.hasTriggers ? w.trigList.p : nullptr, w.trigN.p + 2 * w.trigCur, w.filterClass.p,
                       w.filterTable.p, w.filterK);
    PGX_LAUNCH_CHECK();
    if (w.hasTriggers) {
        ConvexView cv{w.cvVerts.p, w.cvMesh.p, w.cvFace.p, w.cvFaceIdx.p, w.cvFaceN.p, w.cvFaceC.p};
        hipLaunchKernelGGL(k_triggers, dim3(std::min(div_up(nPairs, 64), 1024)), dim3(64), 0, w.stream,
                           w.trigN.p + 2 * w.trigCur, w.trigList.p, w.pkeys.p, w.localBits, w.worldColStart.p,
                           w.colOfLocal.p, w.cbody.p, w.ctype.p, w.cparams.p, w.clpos.p, w.clquat.p, w.bpos.p,
                           w.bquat.p, cv, w.trigN.p + 2 * w.trigCur + 1, w.trigKeys[w.trigCur].p);
        PGX_LAUNCH_CHECK();
    }
    if (w.nConvex > 0) {
        ConvexView cv{w.cvVerts.p, w.cvMesh.p, w.cvFace.p, w.cvFaceIdx.p, w.cvFaceN.p, w.cvFaceC.p};
        hipLaunchKernelGGL(k_narrowphase_convex, dim3(div_up(w.epaLanes, 64)), dim3(64), 0, w.stream,
                           w.counters.p + 8, w.heavy.p, w.pkeys.p, w.localBits, w.worldColStart.p, w.colOfLocal.p,
                           w.cbody.p, w.ctype.p, w.cparams.p, w.clpos.p, w.clquat.p, w.bpos.p, w.bquat.p, cv,
                           w.epaV.p, w.epaFN.p, w.epaFI.p, w.epaE.p, w.epaLanes, w.npNum.p, w.npNormal.p, w.npPts.p,
                           w.counters.p + 9);
        PGX_LAUNCH_CHECK();
    }
    if (w.nTriMesh > 0) {
        MeshArgs A;
        A.nMeshP = w.counters.p + 10; A.meshPairs = w.meshPairs.p; A.pkeys = w.pkeys.p; A.localBits = w.localBits;
        A.worldColStart = w.worldColStart.p; A.colOfLocal = w.colOfLocal.p; A.cbody = w.cbody.p; A.ctype = w.ctype.p;
        A.cparams = w.cparams.p; A.clpos = w.clpos.p; A.clquat = w.clquat.p; A.bpos = w.bpos.p; A.bquat = w.bquat.p;
        A.cv = ConvexView{w.cvVerts.p, w.cvMesh.p, w.cvFace.p, w.cvFaceIdx.p, w.cvFaceN.p, w.cvFaceC.p};
        A.tv = TriMeshView{w.tmVerts.p, w.tmTris.p, w.tmTriN.p, w.tmTriC.p, w.tmNodes.p, w.tmMesh.p};
        A.errors = w.counters.p + 12;
        MeshRecs recs{w.tmRecs.p, (size_t)w.tmCandCap};
        hipLaunchKernelGGL(k_mesh_query, dim3(div_up(w.meshLanes, 64)), dim3(64), 0, w.stream, A, w.tmWs.p,
                           w.meshLanes, w.counters.p + 13, w.tmCandCap, w.tmCand.p, w.tmPairOf.p, w.tmCandOff.p,
                           w.tmCandCount.p, w.meshCount.p);
        hipLaunchKernelGGL(k_mesh_test, dim3(div_up(w.epaLanes, 64)), dim3(64), 0, w.stream, A, w.epaV.p, w.epaFN.p,
                           w.epaFI.p, w.epaE.p, w.epaLanes, w.counters.p + 13, w.tmCandCap, w.tmCand.p, w.tmPairOf.p,
                           recs);
        hipLaunchKernelGGL(k_mesh_emit, dim3(div_up(w.meshLanes, 64)), dim3(64), 0, w.stream, A, w.tmWs.p,
                           w.meshLanes, w.tmCandOff.p, w.tmCandCount.p, recs, w.npNum.p, w.npNormal.p, w.npPts.p,
                           w.npTri.p, w.meshFirst.p, w.meshCount.p, w.counters.p + 11, w.mmCap, w.pairCap);
        PGX_LAUNCH_CHECK();
    }
}

int np_trigger_events(World& w, int cap, int32_t* enter, int32_t* nEnter, int32_t* exit, int32_t* nExit) {
    *nEnter = 0;
    *nExit = 0;
    if (!w.hasTriggers) return PGX_OK;
    hipStream_t s = w.stream;
    int32_t n[4];
    PGX_HIP(hipMemcpyAsync(n, w.trigN.p, sizeof n, hipMemcpyDeviceToHost, s));
    PGX_HIP(hipStreamSynchronize(s));
    const int cur = w.trigCur, prev = cur ^ 1;
    std::vector<uint64_t> a((size_t)n[2 * cur + 1]), b((size_t)n[2 * prev + 1]);
    if (!a.empty()) PGX_HIP(hipMemcpyAsync(a.data(), w.trigKeys[cur].p, 8 * a.size(), hipMemcpyDeviceToHost, s));
    if (!b.empty()) PGX_HIP(hipMemcpyAsync(b.data(), w.trigKeys[prev].p, 8 * b.size(), hipMemcpyDeviceToHost, s));
    PGX_HIP(hipStreamSynchronize(s));
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    std::vector<uint64_t> en, ex;
    std::set_difference(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(en));   // src/Physecs.cpp:538-544
    std::set_difference(b.begin(), b.end(), a.begin(), a.end(), std::back_inserter(ex));   // :545-551
    const uint64_t lmask = (1ull << w.localBits) - 1;
    auto decode = [&](const std::vector<uint64_t>& keys) {
        std::vector<std::pair<int32_t, int32_t>> out;
        for (uint64_t k : keys) {
            const int base = w.hostWorldColStart[(size_t)(k >> (2 * w.localBits))];
            out.push_back({w.hostColOfLocal[base + (int)((k >> w.localBits) & lmask)],
                           w.hostColOfLocal[base + (int)(k & lmask)]});
        }
        std::sort(out.begin(), out.end());
        return out;
    };
    const auto pe = decode(en), px = decode(ex);
    *nEnter = (int32_t)pe.size();
    *nExit = (int32_t)px.size();
    if ((int)pe.size() > cap || (int)px.size() > cap) return PGX_ERR_CAPACITY;
    for (size_t i = 0; i < pe.size(); ++i) { enter[2 * i] = pe[i].first; enter[2 * i + 1] = pe[i].second; }
    for (size_t i = 0; i < px.size(); ++i) { exit[2 * i] = px[i].first; exit[2 * i + 1] = px[i].second; }
    return PGX_OK;
}

}  // namespace pgx
