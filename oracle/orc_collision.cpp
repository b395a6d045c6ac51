// Oracle narrowphase and bounds: CPU restatement of the reference's primitive collision
// routines. TEST INFRASTRUCTURE ONLY: used by tests/ and the CPU baseline, never by the
// product. Every function names the reference lines it follows; expressions keep the
// reference's association order so that results are bit-identical to a -ffp-contract=off
// build of the reference.
#include "orc_collision.h"
#include <limits>

namespace orc {

// ---------------------------------------------------------------- bounds
// src/BoundsUtil.cpp:35-59
Bounds getBounds(V3 pos, Q ori, const Geom& g) {
    switch (g.type) {
        case SPHERE: {
            float r = g.p[0];
            return {pos - v3s(r), pos + v3s(r)};
        }
        case CAPSULE: {
            float hh = g.p[0], r = g.p[1];
            V3 p0 = pos + qrot(ori, v3(0, hh, 0));
            V3 p1 = pos + qrot(ori, v3(0, -hh, 0));
            return {vmin(p0, p1) - r, vmax(p0, p1) + r};
        }
        case BOX: {
            M3 u = mat3_cast(ori);
            M3 au;
            au[0] = vabs(u[0]); au[1] = vabs(u[1]); au[2] = vabs(u[2]);
            V3 he = v3(g.p[0], g.p[1], g.p[2]);
            V3 w = mulmv(au, he);
            return {pos - w, pos + w};
        }
        case TRIANGLE_MESH: {   // src/BoundsUtil.cpp:77-85
            V3 mn = v3s(std::numeric_limits<float>::max()), mx = v3s(std::numeric_limits<float>::lowest());
            for (const V3& v : g.tmesh->verts) {
                V3 t = pos + qrot(ori, v);
                mn = vmin(mn, t);
                mx = vmax(mx, t);
            }
            return {mn, mx};
        }
        case CONVEX_MESH: {   // src/BoundsUtil.cpp:61-69, over the padded vertex buffer
            V3 mn = v3s(std::numeric_limits<float>::max()), mx = v3s(std::numeric_limits<float>::lowest());
            V3 sc = v3(g.p[0], g.p[1], g.p[2]);
            for (const V3& v : g.mesh->verts) {
                V3 t = pos + qrot(ori, sc * v);
                mn = vmin(mn, t);
                mx = vmax(mx, t);
            }
            return {mn, mx};
        }
    }
    return {v3s(0), v3s(0)};
}

// ---------------------------------------------------------------- GeomUtil
// src/GeomUtil.cpp:7-12
V3 closestPointOnSegment(V3 point, V3 lineOrig, V3 lineDir, float mn, float mx) {
    V3 r = lineOrig - point;
    float t = -dot(r, lineDir);
    t = gclamp(t, mn, mx);
    return lineOrig + t * lineDir;
}

// src/GeomUtil.cpp:14-59
void closestPointsBetweenSegments(V3 o0, V3 d0, float mn0, float mx0, V3 o1, V3 d1,
                                  float mn1, float mx1, V3& p0, V3& p1) {
    float v1v2 = dot(d0, d1);
    V3 r = o1 - o0;
    float rv1 = dot(r, d0);
    float rv2 = dot(r, d1);
    float t1 = gabs(v1v2) > 1.f - 0.0001f ? 0.f : (rv1 * v1v2 - rv2) / (1.f - v1v2 * v1v2);
    float t0 = rv1 + t1 * v1v2;
    if (t0 < mn0) {
        t0 = mn0;
        p0 = o0 + t0 * d0;
        t1 = -dot(o1 - p0, d1);
    } else if (t0 > mx0) {
        t0 = mx0;
        p0 = o0 + t0 * d0;
        t1 = -dot(o1 - p0, d1);
    } else {
        p0 = o0 + t0 * d0;
    }
    if (t1 < mn1) {
        t1 = mn1;
        p1 = o1 + t1 * d1;
        t0 = -dot(o0 - p1, d0);
        t0 = gclamp(t0, mn0, mx0);
        p0 = o0 + t0 * d0;
    } else if (t1 > mx1) {
        t1 = mx1;
        p1 = o1 + t1 * d1;
        t0 = -dot(o0 - p1, d0);
        t0 = gclamp(t0, mn0, mx0);
        p0 = o0 + t0 * d0;
    } else {
        p1 = o1 + t1 * d1;
    }
}

// src/GeomUtil.cpp:129-169
static float caseTwoZeros(V3 p, int axis, V3 he, float& t, V3& q) {
    float sqrDist = 0;
    t = he[axis] - p[axis];
    q[axis] = p[axis] + t;
    int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    if (p[a1] > he[a1]) { float d = p[a1] - he[a1]; sqrDist += d * d; q[a1] = he[a1]; }
    else if (p[a1] < -he[a1]) { float d = p[a1] + he[a1]; sqrDist += d * d; q[a1] = -he[a1]; }
    else q[a1] = p[a1];
    if (p[a2] > he[a2]) { float d = p[a2] - he[a2]; sqrDist += d * d; q[a2] = he[a2]; }
    else if (p[a2] < -he[a2]) { float d = p[a2] + he[a2]; sqrDist += d * d; q[a2] = -he[a2]; }
    else q[a2] = p[a2];
    return sqrDist;
}

// src/GeomUtil.cpp:171-235
static float caseOneZero(V3 p, int zeroAxis, V3 dir, V3 he, float& t, V3& q) {
    float sqrDist = 0;
    q[zeroAxis] = p[zeroAxis];
    V3 pme = p - he;
    int x = (zeroAxis + 1) % 3, y = (zeroAxis + 2) % 3;
    float prod0 = dir[y] * pme[x];
    float prod1 = dir[x] * pme[y];
    if (prod0 >= prod1) {
        q[x] = he[x];
        float eY = p[y] + he[y];
        float delta = prod0 - dir[x] * eY;
        if (delta >= 0) {
            sqrDist += delta * delta;
            q[y] = -he[y];
            t = -(pme[x] * dir[x] + eY * dir[y]);
        } else {
            q[y] = p[y] - prod0 / dir[x];
            t = -pme[x] / dir[x];
        }
    } else {
        q[y] = he[y];
        float eX = p[x] + he[x];
        float delta = prod1 - dir[y] * eX;
        if (delta >= 0) {
            sqrDist += delta * delta;
            q[x] = -he[x];
            t = -(eX * dir[x] + pme[y] * dir[y]);
        } else {
            q[x] = p[x] - prod1 / dir[y];
            t = -pme[y] / dir[y];
        }
    }
    if (p[zeroAxis] < -he[zeroAxis]) {
        float d = p[zeroAxis] + he[zeroAxis]; sqrDist += d * d; q[zeroAxis] = -he[zeroAxis];
    } else if (p[zeroAxis] > he[zeroAxis]) {
        float d = p[zeroAxis] - he[zeroAxis]; sqrDist += d * d; q[zeroAxis] = he[zeroAxis];
    }
    return sqrDist;
}

// src/GeomUtil.cpp:237-370. The "region 5" branches keep the reference's
// pMinusExtents[y] + pMinusExtents[y] (GeomUtil.cpp:273,326) where a square is meant.
static float sqrDistLineAABBFace(V3 p, V3 dir, V3 he, int x, V3 pme, float& t, V3& q) {
    q = p;
    t = 0;
    float sqrDist = 0;
    int y = (x + 1) % 3;
    int z = (y + 1) % 3;
    V3 ppe = p + he;
    if (dir[x] * ppe[y] >= dir[y] * pme[x]) {
        if (dir[x] * ppe[z] >= dir[z] * pme[x]) {
            q[x] = he[x];
            q[y] -= pme[x] * dir[y] / dir[x];
            q[z] -= pme[x] * dir[z] / dir[x];
            t = -pme[x] / dir[x];
        } else {
            float sqrLen = dir[x] * dir[x] + dir[z] * dir[z];
            float npY = p[y] - dir[y] * (dir[x] * pme[x] + dir[z] * ppe[z]) / sqrLen;
            if (npY <= he[y]) {
                float tmp = p[y] - npY;
                float delta = dir[x] * pme[x] + dir[y] * tmp + dir[z] * ppe[z];
                t = -delta;
                sqrDist = pme[x] * pme[x] + tmp * tmp + ppe[z] * ppe[z] - delta * delta;
                q[x] = he[x]; q[y] = npY; q[z] = -he[z];
            } else {
                float delta = dir[x] * pme[x] + dir[y] * pme[y] + dir[z] * ppe[z];
                t = -delta;
                sqrDist = pme[x] * pme[x] + pme[y] + pme[y] + ppe[z] * ppe[z] - delta * delta;
                q[x] = he[x]; q[y] = he[y]; q[z] = -he[z];
            }
        }
    } else {
        if (dir[x] * ppe[z] >= dir[z] * pme[x]) {
            float sqrLen = dir[x] * dir[x] + dir[y] * dir[y];
            float npZ = p[z] - dir[z] * (dir[x] * pme[x] + dir[y] * ppe[y]) / sqrLen;
            if (npZ <= he[z]) {
                float tmp = p[z] - npZ;
                float delta = dir[x] * pme[x] + dir[y] * ppe[y] + dir[z] * tmp;
                t = -delta;
                sqrDist = pme[x] * pme[x] + ppe[y] * ppe[y] + tmp * tmp - delta * delta;
                q[x] = he[x]; q[y] = -he[y]; q[z] = npZ;
            } else {
                float delta = dir[x] * pme[x] + dir[y] * ppe[y] + dir[z] * pme[z];
                t = -delta;
                sqrDist = pme[x] * pme[x] + ppe[y] * ppe[y] + pme[z] * pme[z] - delta * delta;
                q[x] = he[x]; q[y] = -he[y]; q[z] = he[z];
            }
        } else {
            float npY = p[y] - dir[y] * (dir[x] * pme[x] + dir[z] * ppe[z]) / (dir[x] * dir[x] + dir[z] * dir[z]);
            if (npY >= -he[y]) {
                if (npY <= he[y]) {
                    float tmp = p[y] - npY;
                    float delta = dir[x] * pme[x] + dir[y] * tmp + dir[z] * ppe[z];
                    t = -delta;
                    sqrDist = pme[x] * pme[x] + tmp * tmp + ppe[z] * ppe[z] - delta * delta;
                    q[x] = he[x]; q[y] = npY; q[z] = -he[z];
                } else {
                    float delta = dir[x] * pme[x] + dir[y] * pme[y] + dir[z] * ppe[z];
                    t = -delta;
                    sqrDist = pme[x] * pme[x] + pme[y] + pme[y] + ppe[z] * ppe[z] - delta * delta;
                    q[x] = he[x]; q[y] = he[y]; q[z] = -he[z];
                }
                return sqrDist;
            }
            float npZ = p[z] - dir[z] * (dir[x] * pme[x] + dir[y] * ppe[y]) / (dir[x] * dir[x] + dir[y] * dir[y]);
            if (npZ >= -he[z]) {
                if (npZ <= he[z]) {
                    float tmp = p[z] - npZ;
                    float delta = dir[x] * pme[x] + dir[y] * ppe[y] + dir[z] * tmp;
                    t = -delta;
                    sqrDist = pme[x] * pme[x] + ppe[y] * ppe[y] + tmp * tmp - delta * delta;
                    q[x] = he[x]; q[y] = -he[y]; q[z] = npZ;
                } else {
                    float delta = dir[x] * pme[x] + dir[y] * ppe[y] + dir[z] * pme[z];
                    t = -delta;
                    sqrDist = pme[x] * pme[x] + ppe[y] * ppe[y] + pme[z] * pme[z] - delta * delta;
                    q[x] = he[x]; q[y] = -he[y]; q[z] = he[z];
                }
                return sqrDist;
            }
            float delta = dir[x] * pme[x] + dir[y] * ppe[y] + dir[z] * ppe[z];
            t = -delta;
            sqrDist = pme[x] * pme[x] + ppe[y] * ppe[y] + ppe[z] * ppe[z] - delta * delta;
            q[x] = he[x]; q[y] = -he[y]; q[z] = -he[z];
        }
    }
    return sqrDist;
}

// src/GeomUtil.cpp:372-399
static float caseNoZeroes(V3 p, V3 dir, V3 he, float& t, V3& q) {
    V3 pme = p - he;
    float dxEy = dir.x * pme.y;
    float dyEx = dir.y * pme.x;
    if (dyEx >= dxEy) {
        float dzEx = dir.z * pme.x;
        float dxEz = dir.x * pme.z;
        if (dzEx >= dxEz) return sqrDistLineAABBFace(p, dir, he, 0, pme, t, q);
        return sqrDistLineAABBFace(p, dir, he, 2, pme, t, q);
    }
    float dzEy = dir.z * pme.y;
    float dyEz = dir.y * pme.z;
    if (dzEy >= dyEz) return sqrDistLineAABBFace(p, dir, he, 1, pme, t, q);
    return sqrDistLineAABBFace(p, dir, he, 2, pme, t, q);
}

// src/GeomUtil.cpp:401-405
static float sqrDistPointAABB(V3 p, V3 he, V3& q) {
    q = vclamp(p, -he, he);
    V3 d = p - q;
    return dot(d, d);
}

// src/GeomUtil.cpp:407-453
static float sqrDistLineAABB(V3 p, V3 dir, V3 he, float& t, V3& q) {
    bool reflect[3];
    for (int i = 0; i < 3; i++) {
        if (dir[i] < 0.0f) { p[i] = -p[i]; dir[i] = -dir[i]; reflect[i] = true; }
        else reflect[i] = false;
    }
    float sqrDist = 0;
    if (dir.x > 0) {
        if (dir.y > 0) {
            if (dir.z > 0) sqrDist = caseNoZeroes(p, dir, he, t, q);
            else sqrDist = caseOneZero(p, 2, dir, he, t, q);
        } else {
            if (dir.z > 0) sqrDist = caseOneZero(p, 1, dir, he, t, q);
            else sqrDist = caseTwoZeros(p, 0, he, t, q);
        }
    } else {
        if (dir.y > 0) {
            if (dir.z > 0) sqrDist = caseOneZero(p, 0, dir, he, t, q);
            else sqrDist = caseTwoZeros(p, 1, he, t, q);
        } else {
            if (dir.z > 0) sqrDist = caseTwoZeros(p, 2, he, t, q);
            else { sqrDist = sqrDistPointAABB(p, he, q); t = 0; }
        }
    }
    for (int i = 0; i < 3; i++)
        if (reflect[i]) q[i] = -q[i];
    return sqrDist;
}

// src/GeomUtil.cpp:455-467
float sqrDistSegmentAABB(V3 p, V3 dir, float mn, float mx, V3 he, float& t, V3& q) {
    float sqrDist = sqrDistLineAABB(p, dir, he, t, q);
    if (t < mn) {
        t = mn;
        sqrDist = sqrDistPointAABB(p + dir * t, he, q);
    } else if (t > mx) {
        t = mx;
        sqrDist = sqrDistPointAABB(p + dir * t, he, q);
    }
    return sqrDist;
}

// ---------------------------------------------------------------- Clipping
// src/Clipping.cpp:10-12
static bool isInside(V2 pt, V2 a, V2 b) {
    return (b.y - a.y) * (pt.x - a.x) - (b.x - a.x) * (pt.y - a.y) > 0;
}
// src/Clipping.cpp:14-19
static V2 intersection(V2 a1, V2 b1, V2 a2, V2 b2) {
    V2 r1 = b1 - a1;
    V2 r2 = b2 - a2;
    float t = det2(a1 - a2, r2) / det2(-r1, r2);
    return a1 + t * r1;
}
// src/Clipping.cpp:21-46 (MAX_POINTS = 128 scratch, :6-8)
static void clipEdge(std::vector<V2>& poly, V2 a, V2 b) {
    V2 np[128];
    int n = 0;
    int sz = (int)poly.size();
    for (int i = 0; i < sz; ++i) {
        int j = (i + 1) % sz;
        bool ii = isInside(poly[i], a, b);
        bool ji = isInside(poly[j], a, b);
        if (ii && ji) np[n++] = poly[j];
        else if (ii) np[n++] = intersection(poly[i], poly[j], a, b);
        else if (ji) { np[n++] = intersection(poly[i], poly[j], a, b); np[n++] = poly[j]; }
    }
    poly.assign(np, np + n);
}
// src/Clipping.cpp:48-53
void suthHodgClip(std::vector<V2>& polygon, const std::vector<V2>& clip) {
    int cs = (int)clip.size();
    for (int i = 0; i < cs; ++i) clipEdge(polygon, clip[i], clip[(i + 1) % cs]);
}
// src/Clipping.cpp:55-71
bool clipLine(V2& p0, V2& p1, const V2* clip, int clipSize) {
    for (int i = 0; i < clipSize; ++i) {
        int j = (i + 1) % clipSize;
        bool in0 = isInside(p0, clip[i], clip[j]);
        bool in1 = isInside(p1, clip[i], clip[j]);
        if (in0 && in1) continue;
        if (in0) p1 = intersection(p0, p1, clip[i], clip[j]);
        else if (in1) p0 = intersection(p0, p1, clip[i], clip[j]);
        else return false;
    }
    return true;
}

// ---------------------------------------------------------------- ContactsUtil
// src/ContactsUtil.cpp:9-107 (epsilon 1e-4f, :6)
void generateContactsPolygonBoxFace(V3 boxCenter, const M3& boxBasis, int boxAxis, int boxAxisSign,
                                           V3 he, V3 incPlaneOrig, V3 incPlaneNormal,
                                           const std::vector<V2>& polygon, int clipX, int clipY,
                                           ContactPoint* cps, int& numPoints) {
    std::vector<V3> pen;
    pen.reserve(polygon.size());
    for (V2 point : polygon) {
        V3 p3 = v3s(0);
        p3[clipX] = point.x;
        p3[clipY] = point.y;
        V3 dir = v3s(0);
        dir[boxAxis] = (float)boxAxisSign;
        float distance = dot(incPlaneOrig - p3, incPlaneNormal) / dot(dir, incPlaneNormal);
        if (distance < he[boxAxis] + 1e-4f) {
            p3[boxAxis] = boxAxisSign * distance;
            pen.push_back(p3);
        }
    }
    auto emit = [&](int slot, int idx) {
        V3 on = pen[idx];
        on[boxAxis] = boxAxisSign * he[boxAxis];
        cps[slot] = {boxCenter + mulmv(boxBasis, on), boxCenter + mulmv(boxBasis, pen[idx])};
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
            V3 ca = pen[0] - pen[i];
            V3 cb = pen[maxDistanceIndex] - pen[i];
            float area = det2({ca[clipX], ca[clipY]}, {cb[clipX], cb[clipY]});
            if (area > maxArea) { maxArea = area; maxAreaIndex = i; }
        }
        emit(2, maxAreaIndex);
        float minArea = 0; int minAreaIndex = 0;
        for (int i = 0; i < n; ++i) {
            V3 da = pen[0] - pen[i];
            V3 db = pen[maxDistanceIndex] - pen[i];
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

// ---------------------------------------------------------------- pair routines
// src/Collision.cpp:54-107
static void generateBoxBoxFaceContacts(const M3& uRef, V3 posRef, V3 heRef, const M3& uInc, V3 posInc,
                                       V3 heInc, int referenceAxis, ContactPoint* cps, int& numPoints) {
    V3 axis = uRef[referenceAxis];
    float refSign = dot(axis, posInc - posRef) < 0 ? -1.f : 1.f;
    float maxDot = 0;
    int incAxis = 0;
    float incSign = 1;
    for (int i = 0; i < 3; ++i) {
        float d = dot((-refSign) * axis, uInc[i]);
        float ad = gabs(d);
        if (ad > maxDot) { maxDot = ad; incAxis = i; incSign = d < 0 ? -1.f : 1.f; }
    }
    int ia1 = (incAxis + 1) % 3, ia2 = (incAxis + 2) % 3;
    M3 invURef = inverse(uRef);
    V3 incPlaneOrig = mulmv(invURef, posInc + uInc[incAxis] * heInc[incAxis] * incSign - posRef);
    V3 poly0 = incPlaneOrig + mulmv(invURef, uInc[ia1] * heInc[ia1] + uInc[ia2] * heInc[ia2]);
    V3 poly1 = incPlaneOrig + mulmv(invURef, -uInc[ia1] * heInc[ia1] + uInc[ia2] * heInc[ia2]);
    V3 poly2 = incPlaneOrig + mulmv(invURef, -uInc[ia1] * heInc[ia1] - uInc[ia2] * heInc[ia2]);
    V3 poly3 = incPlaneOrig + mulmv(invURef, uInc[ia1] * heInc[ia1] - uInc[ia2] * heInc[ia2]);
    int cx = (referenceAxis + 1) % 3, cy = (referenceAxis + 2) % 3;
    std::vector<V2> polygon = {{poly0[cx], poly0[cy]}, {poly1[cx], poly1[cy]},
                               {poly2[cx], poly2[cy]}, {poly3[cx], poly3[cy]}};
    std::vector<V2> clip = {{heRef[cx], heRef[cy]}, {heRef[cx], -heRef[cy]},
                            {-heRef[cx], -heRef[cy]}, {-heRef[cx], heRef[cy]}};
    suthHodgClip(polygon, clip);
    V3 incPlaneNormal = mulmv(invURef, uInc[incAxis]);
    generateContactsPolygonBoxFace(posRef, uRef, referenceAxis, (int)refSign, heRef, incPlaneOrig,
                                   incPlaneNormal, polygon, cx, cy, cps, numPoints);
}

// src/Collision.cpp:109-118
static bool sphereSphere(V3 pos0, float r0, V3 pos1, float r1, Manifold& res) {
    V3 d = pos1 - pos0;
    float rs = r0 + r1;
    if (dot(d, d) > rs * rs) return false;
    float dl = length(d);
    res.normal = dl != 0.f ? d / dl : v3(0, 1, 0);
    res.numPoints = 1;
    res.points[0] = {pos0 + res.normal * r0, pos1 - res.normal * r1};
    return true;
}

// src/Collision.cpp:120-133
static bool capsuleCapsule(V3 pos0, Q or0, float hh0, float r0, V3 pos1, Q or1, float hh1, float r1, Manifold& res) {
    V3 up = v3(0, 1, 0);
    V3 v0 = qrot(or0, up);
    V3 v1 = qrot(or1, up);
    V3 pt0, pt1;
    closestPointsBetweenSegments(pos0, v0, -hh0, hh0, pos1, v1, -hh1, hh1, pt0, pt1);
    V3 d = pt1 - pt0;
    float rs = r0 + r1;
    if (dot(d, d) > rs * rs) return false;
    float dl = length(d);
    res.normal = dl != 0.f ? d / dl : up;
    res.numPoints = 1;
    res.points[0] = {pt0 + res.normal * r0, pt1 - res.normal * r1};
    return true;
}

// src/Collision.cpp:135-350
static bool boxBox(V3 pos0, Q or0, V3 he0, V3 pos1, Q or1, V3 he1, Manifold& res) {
    float edgeOffset = 0.1f;
    const float edgeLimit = 0.999f;
    float ra, rb, l, d;
    bool isEdge = false;
    int faceBox = 0, faceAxis = 0, edge0 = 0, edge1 = 0;
    float mx = std::numeric_limits<float>::lowest();
    M3 r, ar;
    M3 u0 = mat3_cast(or0);
    M3 u1 = mat3_cast(or1);
    V3 t = pos1 - pos0;
    t = v3(dot(t, u0[0]), dot(t, u0[1]), dot(t, u0[2]));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { r[i][j] = dot(u0[i], u1[j]); ar[i][j] = gabs(r[i][j]); }
    for (int i = 0; i < 3; i++) {
        ra = he0[i];
        rb = he1[0] * ar[i][0] + he1[1] * ar[i][1] + he1[2] * ar[i][2];
        l = gabs(t[i]);
        d = l - ra - rb;
        if (d > 0) return false;
        if (d > mx) { mx = d; isEdge = false; faceBox = 0; faceAxis = i; }
    }
    for (int i = 0; i < 3; i++) {
        ra = he0[0] * ar[0][i] + he0[1] * ar[1][i] + he0[2] * ar[2][i];
        rb = he1[i];
        l = gabs(t[0] * r[0][i] + t[1] * r[1][i] + t[2] * r[2][i]);
        d = l - ra - rb;
        if (d > 0) return false;
        if (d > mx) { mx = d; isEdge = false; faceBox = 1; faceAxis = i; }
    }
    // edge axes A_i x B_j, in the reference's order (Collision.cpp:194-308)
    struct EdgeAxis { int i, j; };
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            // ra = he0[i1]*|R[i2][j]| + he0[i2]*|R[i1][j]|, reference lists the
            // lower axis index first in each sum; reproduce its operand order.
            int lo_i = i1 < i2 ? i1 : i2, hi_i = i1 < i2 ? i2 : i1;
            int lo_j = j1 < j2 ? j1 : j2, hi_j = j1 < j2 ? j2 : j1;
            ra = he0[lo_i] * ar[lo_i == i1 ? i2 : i1][j] + he0[hi_i] * ar[hi_i == i1 ? i2 : i1][j];
            rb = he1[lo_j] * ar[i][lo_j == j1 ? j2 : j1] + he1[hi_j] * ar[i][hi_j == j1 ? j2 : j1];
            // l = |t[i2]*R[i1][j] - t[i1]*R[i2][j]|  (cyclic order as in the reference)
            l = gabs(t[i2] * r[i1][j] - t[i1] * r[i2][j]);
            d = l - ra - rb;
            if (d > 0) return false;
            if (ar[i][j] < edgeLimit && d > mx + edgeOffset) {
                mx = d; isEdge = true; edge0 = i; edge1 = j;
                if (!(i == 2 && j == 2)) edgeOffset = 0;
            }
        }
    }
    if (!isEdge) {
        V3 axis;
        ContactPoint cps[4];
        int np = 0;
        if (faceBox) {
            axis = u1[faceAxis];
            generateBoxBoxFaceContacts(u1, pos1, he1, u0, pos0, he0, faceAxis, cps, np);
            for (int i = 0; i < np; i++) res.points[i] = {cps[i].p1, cps[i].p0};
        } else {
            axis = u0[faceAxis];
            generateBoxBoxFaceContacts(u0, pos0, he0, u1, pos1, he1, faceAxis, cps, np);
            for (int i = 0; i < np; i++) res.points[i] = cps[i];
        }
        res.numPoints = np;
        res.normal = dot(axis, pos1 - pos0) < 0 ? -axis : axis;
    } else {
        V3 axis = cross(u0[edge0], u1[edge1]);
        res.normal = normalize(dot(axis, pos1 - pos0) < 0 ? -axis : axis);
        V3 s0 = v3s(-1);
        s0[(edge0 + 1) % 3] = dot(u0[(edge0 + 1) % 3], res.normal) < 0 ? -1.f : 1.f;
        s0[(edge0 + 2) % 3] = dot(u0[(edge0 + 2) % 3], res.normal) < 0 ? -1.f : 1.f;
        V3 p0 = pos0 + s0[0] * he0[0] * u0[0] + s0[1] * he0[1] * u0[1] + s0[2] * he0[2] * u0[2];
        V3 s1 = v3s(-1);
        s1[(edge1 + 1) % 3] = dot(u1[(edge1 + 1) % 3], res.normal) > 0 ? -1.f : 1.f;
        s1[(edge1 + 2) % 3] = dot(u1[(edge1 + 2) % 3], res.normal) > 0 ? -1.f : 1.f;
        V3 p1 = pos1 + s1[0] * he1[0] * u1[0] + s1[1] * he1[1] * u1[1] + s1[2] * he1[2] * u1[2];
        res.numPoints = 1;
        V3 q0, q1;
        closestPointsBetweenSegments(p0, u0[edge0], 0, he0[edge0] * 2, p1, u1[edge1], 0, he1[edge1] * 2, q0, q1);
        res.points[0] = {q0, q1};
    }
    return true;
}

// src/Collision.cpp:428-440
static bool sphereCapsule(V3 pos0, float r0, V3 pos1, Q or1, float hh1, float r1, Manifold& res) {
    V3 up = v3(0, 1, 0);
    V3 dir = qrot(or1, up);
    V3 q = closestPointOnSegment(pos0, pos1, dir, -hh1, hh1);
    V3 v = q - pos0;
    float rs = r0 + r1;
    if (dot(v, v) > rs * rs) return false;
    float vl = length(v);
    res.normal = vl != 0.f ? v / vl : up;
    res.numPoints = 1;
    res.points[0] = {pos0 + res.normal * r0, q - res.normal * r1};
    return true;
}

// src/Collision.cpp:442-486
static bool sphereBox(V3 pos0, float r0, V3 pos1, Q or1, V3 he1, Manifold& res) {
    M3 u1 = mat3_cast(or1);
    V3 d = pos0 - pos1;
    d = v3(dot(d, u1[0]), dot(d, u1[1]), dot(d, u1[2]));
    V3 q = pos1 + mulmv(u1, vclamp(d, -he1, he1));
    V3 v = q - pos0;
    if (dot(v, v) > r0 * r0) return false;
    float vl = length(v);
    if (vl != 0.f) {
        res.normal = v / vl;
        res.numPoints = 1;
        res.points[0] = {pos0 + res.normal * r0, q};
    } else {
        float mn = std::numeric_limits<float>::max();
        int axis = 0;
        for (int i = 0; i < 3; i++) {
            float depth = he1[i] - gabs(d[i]);
            if (depth < mn) { mn = depth; axis = i; }
        }
        float sign = d[axis] != 0.f ? gsign(d[axis]) : 1.f;
        V3 bp = d;
        bp[axis] = sign * he1[axis];
        res.normal = (-sign) * u1[axis];
        res.numPoints = 1;
        res.points[0] = {pos0, pos1 + mulmv(u1, bp)};
    }
    return true;
}

// src/Collision.cpp:501-582
static void generateCapsuleBoxContacts(V3 p0L, V3 p1L, float radius, V3 normal, V3 boxCenter, V3 he,
                                       const M3& basis, Manifold& res) {
    int boxAxis = 0;
    float boxAxisSign = 0;
    float maxDot = 0;
    for (int i = 0; i < 3; ++i) {
        float d = dot(normal, basis[i]);
        float ad = gabs(d);
        if (ad > maxDot) { maxDot = ad; boxAxis = i; boxAxisSign = d < 0 ? 1.f : -1.f; }
    }
    int cx = (boxAxis + 1) % 3, cy = (boxAxis + 2) % 3;
    V2 line0 = {p0L[cx], p0L[cy]};
    V2 line1 = {p1L[cx], p1L[cy]};
    V2 clip[4] = {{he[cx], he[cy]}, {he[cx], -he[cy]}, {-he[cx], -he[cy]}, {-he[cx], he[cy]}};
    if (clipLine(line0, line1, clip, 4)) {
        float onBoxAxis = boxAxisSign * he[boxAxis];
        V3 p0 = v3s(0), p1 = v3s(0);
        float distX = p0L[cx] - p1L[cx];
        float distY = p0L[cy] - p1L[cy];
        float adx = gabs(distX), ady = gabs(distY);
        if (adx > ady) {
            p0[boxAxis] = gmix(p0L[boxAxis], p1L[boxAxis], (p0L[cx] - line0.x) / distX);
            p1[boxAxis] = gmix(p1L[boxAxis], p0L[boxAxis], (p1L[cx] - line1.x) / -distX);
        } else if (ady > adx) {
            p0[boxAxis] = gmix(p0L[boxAxis], p1L[boxAxis], (p0L[cy] - line0.y) / distY);
            p1[boxAxis] = gmix(p1L[boxAxis], p0L[boxAxis], (p1L[cy] - line1.y) / -distY);
        } else {
            p0[boxAxis] = p0L[boxAxis];
            p1[boxAxis] = p1L[boxAxis];
        }
        p0[boxAxis] -= boxAxisSign * radius;
        p1[boxAxis] -= boxAxisSign * radius;
        int np = 0;
        if (gabs(p0[boxAxis]) < he[boxAxis]) {
            p0[cx] = line0.x; p0[cy] = line0.y;
            V3 ob = p0; ob[boxAxis] = onBoxAxis;
            res.points[np++] = {boxCenter + mulmv(basis, p0), boxCenter + mulmv(basis, ob)};
        }
        if (gabs(p1[boxAxis]) < he[boxAxis]) {
            p1[cx] = line1.x; p1[cy] = line1.y;
            V3 ob = p1; ob[boxAxis] = onBoxAxis;
            res.points[np++] = {boxCenter + mulmv(basis, p1), boxCenter + mulmv(basis, ob)};
        }
        res.numPoints = np;
    } else {
        res.numPoints = 0;
    }
}

// src/Collision.cpp:584-692
static bool capsuleBox(V3 pos0, Q or0, float hh0, float r0, V3 pos1, Q or1, V3 he1, Manifold& res) {
    M3 u1 = mat3_cast(or1);
    V3 p = pos0 - pos1;
    p = v3(dot(p, u1[0]), dot(p, u1[1]), dot(p, u1[2]));
    V3 dir = qrot(or0, v3(0, 1, 0));
    V3 dirLc = v3(dot(dir, u1[0]), dot(dir, u1[1]), dot(dir, u1[2]));
    float t = 0;
    V3 q = v3s(0);
    float sq = sqrDistSegmentAABB(p, dirLc, -hh0, hh0, he1, t, q);
    if (sq >= r0 * r0) return false;
    V3 onSeg = pos0 + dir * t;
    V3 onBox = pos1 + mulmv(u1, q);
    V3 normal = onBox - onSeg;
    float nl = length(normal);
    if (nl != 0.f) {
        res.normal = normal / nl;
    } else {
        V3 r, ar;
        for (int i = 0; i < 3; ++i) { r[i] = dot(dir, u1[i]); ar[i] = gabs(r[i]); }
        float mx = std::numeric_limits<float>::lowest();
        bool isEdge = false;
        int axis = 0;
        float ra, rb, l, d;
        for (int i = 0; i < 3; i++) {
            ra = hh0 * ar[i] + r0;
            rb = he1[i];
            l = gabs(p[i]);
            d = l - ra - rb;
            if (d > mx) { mx = d; isEdge = false; axis = i; }
        }
        const float edgeLimit = 0.999f;
        float edgeOffset = 0.1f;
        rb = he1[1] * ar[2] + he1[2] * ar[1];
        l = gabs(p[2] * r[1] - p[1] * r[2]);
        d = l - r0 - rb;
        if (ar[0] < edgeLimit && d > mx + edgeOffset) { mx = d; isEdge = true; axis = 0; edgeOffset = 0; }
        rb = he1[0] * ar[2] + he1[2] * ar[0];
        l = gabs(p[0] * r[2] - p[2] * r[0]);
        d = l - r0 - rb;
        if (ar[1] < edgeLimit && d > mx + edgeOffset) { mx = d; isEdge = true; axis = 1; edgeOffset = 0; }
        rb = he1[0] * ar[1] + he1[1] * ar[0];
        l = gabs(p[1] * r[0] - p[0] * r[1]);
        d = l - r0 - rb;
        if (ar[2] < edgeLimit && d > mx + edgeOffset) { mx = d; isEdge = true; axis = 2; }
        V3 sep = isEdge ? cross(dir, u1[axis]) : u1[axis];
        res.normal = normalize(dot(sep, pos1 - pos0) > 0 ? sep : -sep);
    }
    generateCapsuleBoxContacts(p - dirLc * hh0, p + dirLc * hh0, r0, res.normal, pos1, he1, u1, res);
    if (!res.numPoints) {
        res.numPoints = 1;
        res.points[0] = {onSeg + res.normal * r0, onBox};
    }
    return true;
}

// src/Collision.cpp:889-893
static void flipContacts(Manifold& m) {
    m.normal = -m.normal;
    for (int i = 0; i < m.numPoints; ++i) m.points[i] = {m.points[i].p1, m.points[i].p0};
}

// src/Collision.cpp:895-1016
bool collision(V3 pos0, Q or0, const Geom& g0, V3 pos1, Q or1, const Geom& g1, std::vector<Manifold>& out) {
    if (g0.type == TRIANGLE_MESH) {
        size_t first = out.size();
        bool hit = collisionTriangleMesh(pos1, or1, g1, pos0, or0, *g0.tmesh, out);
        for (size_t i = first; i < out.size(); ++i) flipContacts(out[i]);
        return hit;
    }
    if (g1.type == TRIANGLE_MESH) return collisionTriangleMesh(pos0, or0, g0, pos1, or1, *g1.tmesh, out);
    out.push_back(Manifold());
    Manifold& m = out.back();
    auto he = [](const Geom& g) { return v3(g.p[0], g.p[1], g.p[2]); };
    bool hit = false;
    switch (g0.type) {
        case SPHERE:
            switch (g1.type) {
                case SPHERE: return sphereSphere(pos0, g0.p[0], pos1, g1.p[0], m);
                case CAPSULE: return sphereCapsule(pos0, g0.p[0], pos1, or1, g1.p[0], g1.p[1], m);
                case BOX: return sphereBox(pos0, g0.p[0], pos1, or1, he(g1), m);
                case CONVEX_MESH: return collisionSphereConvex(pos0, g0.p[0], pos1, or1, g1.mesh, he(g1), m);
            }
            break;
        case CAPSULE:
            switch (g1.type) {
                case SPHERE:
                    hit = sphereCapsule(pos1, g1.p[0], pos0, or0, g0.p[0], g0.p[1], m);
                    flipContacts(m);
                    return hit;
                case CAPSULE: return capsuleCapsule(pos0, or0, g0.p[0], g0.p[1], pos1, or1, g1.p[0], g1.p[1], m);
                case BOX: return capsuleBox(pos0, or0, g0.p[0], g0.p[1], pos1, or1, he(g1), m);
                case CONVEX_MESH:
                    return collisionCapsuleConvex(pos0, or0, g0.p[0], g0.p[1], pos1, or1, g1.mesh, he(g1), m);
            }
            break;
        case BOX:
            switch (g1.type) {
                case SPHERE:
                    hit = sphereBox(pos1, g1.p[0], pos0, or0, he(g0), m);
                    flipContacts(m);
                    return hit;
                case CAPSULE:
                    hit = capsuleBox(pos1, or1, g1.p[0], g1.p[1], pos0, or0, he(g0), m);
                    flipContacts(m);
                    return hit;
                case BOX: return boxBox(pos0, or0, he(g0), pos1, or1, he(g1), m);
                case CONVEX_MESH: return collisionBoxConvex(pos0, or0, he(g0), pos1, or1, g1.mesh, he(g1), m);
            }
            break;
        case CONVEX_MESH:
            switch (g1.type) {
                case SPHERE:
                    hit = collisionSphereConvex(pos1, g1.p[0], pos0, or0, g0.mesh, he(g0), m);
                    flipContacts(m);
                    return hit;
                case CAPSULE:
                    hit = collisionCapsuleConvex(pos1, or1, g1.p[0], g1.p[1], pos0, or0, g0.mesh, he(g0), m);
                    flipContacts(m);
                    return hit;
                case BOX:
                    hit = collisionBoxConvex(pos1, or1, he(g1), pos0, or0, g0.mesh, he(g0), m);
                    flipContacts(m);
                    return hit;
                case CONVEX_MESH:
                    return collisionConvexConvex(pos0, or0, g0.mesh, he(g0), pos1, or1, g1.mesh, he(g1), m);
            }
            break;
    }
    return false;
}

}  // namespace orc
