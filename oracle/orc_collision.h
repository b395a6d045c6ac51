// Oracle narrowphase for the primitive shapes (sphere, capsule, box): a CPU restatement
// of the reference's pairwise collision routines with the reference's floating-point
// operation order. TEST INFRASTRUCTURE ONLY (never linked into the product).
#pragma once
#include <vector>
#include "orc_math.h"

namespace orc {

enum GeomType { SPHERE = 0, CAPSULE = 1, BOX = 2, CONVEX_MESH = 3, TRIANGLE_MESH = 4 };

// include/Physecs/ConvexMesh.h:7-34. Vertices are padded to a multiple of 4 with copies
// of the last vertex (src/ConvexMesh.cpp:3-11); `count` is the unpadded number.
struct ConvexFace { std::vector<int> idx; V3 normal, centroid; };
struct ConvexMesh {
    int count = 0;
    std::vector<V3> verts;
    std::vector<ConvexFace> faces;
};

struct Bounds { V3 min, max; };

// include/Physecs/TriangleMesh.h: triangles are permuted by the BVH build; a manifold's
// triangleIndex is the post-build index
struct Tri { uint32_t idx[3]; Bounds bounds; V3 normal, centroid; };
struct TriNode { Bounds bounds; int triCount; int index; };
struct TriMesh {
    std::vector<V3> verts;
    std::vector<Tri> tris;
    std::vector<TriNode> bvh;
};
TriMesh buildTriMesh(const std::vector<V3>& verts, const std::vector<uint32_t>& indices);   // src/TriangleMesh.cpp:139-161
std::vector<int> overlapBvh(const TriMesh& m, const Bounds& b);                         // src/TriangleMesh.cpp:163-192

// params: sphere {r}, capsule {halfHeight, r}, box {hx, hy, hz}, convex {sx, sy, sz} (scale)
struct Geom { int type; float p[4]; const ConvexMesh* mesh = nullptr; const TriMesh* tmesh = nullptr; };

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

Bounds getBounds(V3 pos, Q ori, const Geom& g);          // src/BoundsUtil.cpp:23-59

// geometry helpers (exported for unit pinning)
V3 closestPointOnSegment(V3 point, V3 lineOrig, V3 lineDir, float mn, float mx);
void closestPointsBetweenSegments(V3 o0, V3 d0, float mn0, float mx0, V3 o1, V3 d1,
                                  float mn1, float mx1, V3& p0, V3& p1);
float sqrDistSegmentAABB(V3 p, V3 dir, float mn, float mx, V3 he, float& t, V3& q);
void suthHodgClip(std::vector<V2>& polygon, const std::vector<V2>& clip);
bool clipLine(V2& p0, V2& p1, const V2* clip, int clipSize);
void generateContactsPolygonBoxFace(V3 boxCenter, const M3& boxBasis, int boxAxis, int boxAxisSign, V3 he,
                                    V3 incPlaneOrig, V3 incPlaneNormal, const std::vector<V2>& polygon,
                                    int clipX, int clipY, ContactPoint* cps, int& numPoints);

// convex-mesh pairs (orc_convex.cpp): src/Collision.cpp:352-426,488-499,694-887
bool collisionSphereConvex(V3 pos0, float r0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1, Manifold& res);
bool collisionCapsuleConvex(V3 pos0, Q or0, float hh0, float r0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1,
                            Manifold& res);
bool collisionBoxConvex(V3 pos0, Q or0, V3 he0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1, Manifold& res);
bool collisionConvexConvex(V3 pos0, Q or0, const ConvexMesh* m0, V3 s0, V3 pos1, Q or1, const ConvexMesh* m1,
                           V3 s1, Manifold& res);
// src/Overlap.cpp:141-200 (GJK only)
bool overlapSphereConvex(V3 pos0, float r0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1);
bool overlapCapsuleConvex(V3 pos0, Q or0, float hh0, float r0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1);
bool overlapBoxConvex(V3 pos0, Q or0, V3 he0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1);
bool overlapConvexConvex(V3 pos0, Q or0, const ConvexMesh* m0, V3 s0, V3 pos1, Q or1, const ConvexMesh* m1, V3 s1);
// GJK + EPA of a convex mesh against one triangle (src/CollisionTriangleMesh.cpp:616-620)
bool gjkEpaConvexTriangle(V3 pos, Q ori, const ConvexMesh* m, V3 scale, V3 a, V3 b, V3 c, V3 startDir, V3& normal,
                          V3& cpBody, V3& cpTri);
void contactsPolygonPolygonFace(V3 refPos, const M3& refToWorld, V3 refPlaneOrigin, V3 refPlaneNormal,
                                V3 incPlaneOrigin, V3 incPlaneNormal, const std::vector<V2>& polygon, int clipX,
                                int clipY, ContactPoint* cps, int& numPoints);
// physecs::collisionTriangleMesh (src/CollisionTriangleMesh.cpp:882-956): appends one
// manifold per contacting triangle
bool collisionTriangleMesh(V3 pos0, Q or0, const Geom& g0, V3 pos1, Q or1, const TriMesh& mesh,
                           std::vector<Manifold>& out);
// counters of EPA runs that needed the whole 100 iterations / hit a zero face normal
extern thread_local long epaExhausted, epaZeroNormal;

}  // namespace orc
