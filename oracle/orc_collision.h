// Oracle narrowphase for the primitive shapes (sphere, capsule, box): a CPU restatement
// of the reference's pairwise collision routines with the reference's floating-point
// operation order. TEST INFRASTRUCTURE ONLY (never linked into the product).
#pragma once
#include <vector>
#include "orc_math.h"

namespace orc {

enum GeomType { SPHERE = 0, CAPSULE = 1, BOX = 2, CONVEX_MESH = 3, TRIANGLE_MESH = 4 };

// params: sphere {r}, capsule {halfHeight, r}, box {hx, hy, hz}
struct Geom { int type; float p[4]; };

struct ContactPoint { V3 p0, p1; };
struct Manifold {                          // src/ContactManifold.h:11-16
    int numPoints = 0;
    ContactPoint points[4];
    V3 normal = {0, 0, 0};
    int triangleIndex = -1;
};

// Reference collision dispatcher restated (src/Collision.cpp:895-1016) for primitive
// pairs; returns the bool the reference returns and appends one manifold.
bool collision(V3 pos0, Q or0, const Geom& g0, V3 pos1, Q or1, const Geom& g1,
               std::vector<Manifold>& out);

struct Bounds { V3 min, max; };
Bounds getBounds(V3 pos, Q ori, const Geom& g);          // src/BoundsUtil.cpp:23-59

// geometry helpers (exported for unit pinning)
V3 closestPointOnSegment(V3 point, V3 lineOrig, V3 lineDir, float mn, float mx);
void closestPointsBetweenSegments(V3 o0, V3 d0, float mn0, float mx0, V3 o1, V3 d1,
                                  float mn1, float mx1, V3& p0, V3& p1);
float sqrDistSegmentAABB(V3 p, V3 dir, float mn, float mx, V3 he, float& t, V3& q);
void suthHodgClip(std::vector<V2>& polygon, const std::vector<V2>& clip);
bool clipLine(V2& p0, V2& p1, const V2* clip, int clipSize);

}  // namespace orc
