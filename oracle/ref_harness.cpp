// Reference harness: a C ABI over functions of the reference itself, compiled from the
// reference's own sources where they lie (see oracle/Makefile), so that the tests can
// pin the CPU oracle against the real reference code. TEST INFRASTRUCTURE ONLY.
//
// Calls into the reference:
//   physecs::getBounds            src/BoundsUtil.cpp:23-33
//   physecs::collision            src/Collision.cpp:895-1016
//   sqrDistSegmentAABB            src/GeomUtil.cpp:455-467
//   ContactConstraints::preSolve / solve   src/ContactConstraints.cpp:4-124
//   physecs::setMassProps         src/MassUtil.cpp:185-192
//   solve33                       src/MathUtil.h:10-21
//   ConvexMesh(vertices, faces)   include/Physecs/ConvexMesh.h:29-34, src/ConvexMesh.cpp
//   TriangleMesh(vertices, indices) src/TriangleMesh.cpp:139-161
// The two integrate functions below are GLM-written restatements of src/Physecs.cpp
// (that file does not compile here): they pin the oracle's GLM-free arithmetic against
// GLM's own operator order.
#include <cstring>
#include <vector>
#include <glm/glm.hpp>
#include <glm/gtx/quaternion.hpp>
#include <glm/gtx/matrix_cross_product.hpp>
#include "BoundsUtil.h"
#include "Collision.h"
#include "ContactConstraints.h"
#include "ContactManifold.h"
#include "GeomUtil.h"
#include "MassUtil.h"
#include "MathUtil.h"
#include "Overlap.h"
#include "Raycast.h"
#include "SolverData.h"
#include "TriangleMesh.h"
#include "Constraint1D.h"
#include "Constraint1D.cpp"   // the row templates, compiled from the reference source
#include "ConvexMesh.cpp"     // ConvexMeshVertices<4>'s constructor template (not instantiated there)

using namespace physecs;

// convex meshes registered by ref_set_convex_mesh (include/Physecs/ConvexMesh.h:7-34)
static std::vector<ConvexMesh*> g_convex;
// triangle meshes registered by ref_set_tri_mesh (include/Physecs/TriangleMesh.h)
static std::vector<TriangleMesh*> g_tri;

static Geometry makeGeom(int type, const float* p) {
    Geometry g;
    std::memset(&g, 0, sizeof(g));
    g.type = (GeometryType)type;
    switch (type) {
        case SPHERE: g.sphere.radius = p[0]; break;
        case CAPSULE: g.capsule.halfHeight = p[0]; g.capsule.radius = p[1]; break;
        case BOX: g.box.halfExtents = glm::vec3(p[0], p[1], p[2]); break;
        case CONVEX_MESH:   // params: scale xyz, mesh index
            g.convex.mesh = g_convex.at((size_t)p[3]);
            g.convex.scale = glm::vec3(p[0], p[1], p[2]);
            break;
        case TRIANGLE_MESH:   // params: -, -, -, mesh index
            g.triangleMesh.mesh = g_tri.at((size_t)p[3]);
            break;
        default: break;
    }
    return g;
}
static glm::vec3 V(const float* p) { return glm::vec3(p[0], p[1], p[2]); }
static glm::quat Qt(const float* q) { return glm::quat(q[3], q[0], q[1], q[2]); }

// One Constraint1D<flags> row (src/Constraint1D.cpp): preSolve, then `iters` solves, the
// first with warm start. in: linear(3) angular0(3) angular1(3) tv c min max freq damp;
// masses: invMass0, invI0(9), invMass1, invI1(9); vel: v0 w0 v1 w1 (12, updated);
// pv: pv0 pw0 count0 pv1 pw1 count1 (14, updated). out: invEff, lambda.
// lambda0 is the accumulated lambda the row starts with (totalLambda, which carries over from
// the previous substep within a step); bit i of warmMask asks for warm start on solve i.
template <int F>
static void runRow(int b0, int b1, const float* in, const float* masses, float* vel, float* pv, float h, int iters,
                   float* out, float lambda0 = 0.f, unsigned warmMask = 1u) {
    Constraint1D<F> r(b0 >= 0 ? 0 : -1, b1 >= 0 ? 1 : -1);
    r.linear = V(in); r.angular0 = V(in + 3); r.angular1 = V(in + 6);
    r.targetVelocity = in[9]; r.c = in[10]; r.min = in[11]; r.max = in[12]; r.frequency = in[13]; r.dampingRatio = in[14];
    MassData md[2];
    VelocityData vd[2];
    PseudoVelocityData pd[2];
    for (int k = 0; k < 2; ++k) {
        md[k].invMass = masses[10 * k];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) md[k].invInertiaTensor[i][j] = masses[10 * k + 1 + 3 * i + j];
        vd[k].velocity = V(vel + 6 * k); vd[k].angularVelocity = V(vel + 6 * k + 3);
        pd[k].pseudoVelocity = V(pv + 7 * k); pd[k].pseudoAngularVelocity = V(pv + 7 * k + 3);
        pd[k].constraintCount = (int)pv[7 * k + 6];
    }
    r.preSolve(md, pd);
    r.totalLambda = lambda0;
    for (int i = 0; i < iters; ++i) r.solve(vd, h, ((warmMask >> i) & 1u) != 0);
    for (int k = 0; k < 2; ++k) {
        vel[6 * k] = vd[k].velocity.x; vel[6 * k + 1] = vd[k].velocity.y; vel[6 * k + 2] = vd[k].velocity.z;
        vel[6 * k + 3] = vd[k].angularVelocity.x; vel[6 * k + 4] = vd[k].angularVelocity.y; vel[6 * k + 5] = vd[k].angularVelocity.z;
        pv[7 * k] = pd[k].pseudoVelocity.x; pv[7 * k + 1] = pd[k].pseudoVelocity.y; pv[7 * k + 2] = pd[k].pseudoVelocity.z;
        pv[7 * k + 3] = pd[k].pseudoAngularVelocity.x; pv[7 * k + 4] = pd[k].pseudoAngularVelocity.y;
        pv[7 * k + 5] = pd[k].pseudoAngularVelocity.z; pv[7 * k + 6] = (float)pd[k].constraintCount;
    }
    out[0] = r.invEffMass;
    out[1] = r.totalLambda;
}

extern "C" {

int ref_bounds(int type, const float* prm, const float* pos, const float* q, float* out) {
    Geometry g = makeGeom(type, prm);
    Bounds b = getBounds(V(pos), Qt(q), g);
    out[0] = b.min.x; out[1] = b.min.y; out[2] = b.min.z;
    out[3] = b.max.x; out[4] = b.max.y; out[5] = b.max.z;
    return 0;
}

// out: numPoints, normal(3), 4 x (p0, p1); returns collision()'s bool
int ref_collision(int type0, const float* prm0, const float* pos0, const float* q0,
                  int type1, const float* prm1, const float* pos1, const float* q1, float* out) {
    Geometry g0 = makeGeom(type0, prm0), g1 = makeGeom(type1, prm1);
    std::vector<ContactManifold> res;
    bool hit = collision(V(pos0), Qt(q0), g0, V(pos1), Qt(q1), g1, res);
    std::memset(out, 0, sizeof(float) * 28);
    const ContactManifold& m = res.back();
    out[0] = (float)m.numPoints;
    if (m.numPoints) { out[1] = m.normal.x; out[2] = m.normal.y; out[3] = m.normal.z; }
    for (int k = 0; k < m.numPoints && k < 4; ++k) {
        float* o = out + 4 + 6 * k;
        o[0] = m.points[k].position0.x; o[1] = m.points[k].position0.y; o[2] = m.points[k].position0.z;
        o[3] = m.points[k].position1.x; o[4] = m.points[k].position1.y; o[5] = m.points[k].position1.z;
    }
    return hit ? 1 : 0;
}

// Registers a ConvexMesh built by the reference's own constructor; returns its index.
// faces: face_start[n_faces+1] offsets into face_idx; normals/centroids xyz per face.
int ref_set_convex_mesh(int n_verts, const float* verts, int n_faces, const int* face_start, const int* face_idx,
                        const float* normals, const float* centroids) {
    std::vector<glm::vec3> v(n_verts);
    for (int i = 0; i < n_verts; ++i) v[i] = V(verts + 3 * i);
    std::vector<ConvexMeshFace> f(n_faces);
    for (int i = 0; i < n_faces; ++i) {
        f[i].indices.assign(face_idx + face_start[i], face_idx + face_start[i + 1]);
        f[i].normal = V(normals + 3 * i);
        f[i].centroid = V(centroids + 3 * i);
    }
    g_convex.push_back(new ConvexMesh(std::move(v), std::move(f)));
    return (int)g_convex.size() - 1;
}

// Builds a TriangleMesh with the reference's constructor (BVH included); returns its index.
int ref_set_tri_mesh(int n_verts, const float* verts, int n_tris, const uint32_t* idx) {
    std::vector<glm::vec3> v(n_verts);
    for (int i = 0; i < n_verts; ++i) v[i] = V(verts + 3 * i);
    std::vector<unsigned int> ind(idx, idx + 3 * n_tris);
    g_tri.push_back(new TriangleMesh(v, ind));
    return (int)g_tri.size() - 1;
}

// triangle vertex indices after the build's permutation, and BVH nodes {min, max, triCount, index}
int ref_tri_mesh_layout(int mesh, uint32_t* triIdx, float* nodes) {
    const TriangleMesh& m = *g_tri.at((size_t)mesh);
    for (size_t t = 0; t < m.triangles.size(); ++t)
        for (int k = 0; k < 3; ++k) triIdx[3 * t + k] = m.triangles[t].indices[k];
    for (size_t n = 0; n < m.bvh.size(); ++n) {
        const TriangleMeshBVHNode& b = m.bvh[n];
        float* o = nodes + 8 * n;
        o[0] = b.bounds.min.x; o[1] = b.bounds.min.y; o[2] = b.bounds.min.z;
        o[3] = b.bounds.max.x; o[4] = b.bounds.max.y; o[5] = b.bounds.max.z;
        o[6] = (float)b.triCount; o[7] = (float)b.index;
    }
    return (int)m.bvh.size();
}

// every manifold of one collision() call (see orc_collision_multi for the layout)
int ref_collision_multi(int type0, const float* prm0, const float* pos0, const float* q0, int type1,
                        const float* prm1, const float* pos1, const float* q1, int cap, float* out) {
    Geometry g0 = makeGeom(type0, prm0), g1 = makeGeom(type1, prm1);
    std::vector<ContactManifold> res;
    bool hit = collision(V(pos0), Qt(q0), g0, V(pos1), Qt(q1), g1, res);
    int n = 0;
    for (const ContactManifold& m : res) {
        if (n >= cap) break;
        float* o = out + 30 * n++;
        std::memset(o, 0, sizeof(float) * 30);
        o[0] = (float)m.numPoints;
        o[1] = m.normal.x; o[2] = m.normal.y; o[3] = m.normal.z;
        for (int k = 0; k < m.numPoints && k < 4; ++k) {
            o[4 + 6 * k] = m.points[k].position0.x; o[5 + 6 * k] = m.points[k].position0.y;
            o[6 + 6 * k] = m.points[k].position0.z; o[7 + 6 * k] = m.points[k].position1.x;
            o[8 + 6 * k] = m.points[k].position1.y; o[9 + 6 * k] = m.points[k].position1.z;
        }
        o[28] = (float)m.triangleIndex;
        o[29] = hit ? 1.f : 0.f;
    }
    return hit ? n : -1;
}

int ref_sqr_dist_segment_aabb(const float* p, const float* dir, float mn, float mx, const float* he, float* out) {
    float t = 0;
    glm::vec3 q(0);
    float sq = sqrDistSegmentAABB(V(p), V(dir), mn, mx, V(he), t, q);
    out[0] = sq; out[1] = t; out[2] = q.x; out[3] = q.y; out[4] = q.z;
    return 0;
}

// mass properties of a one-collider body: out invMass, com(3), invInertia(9, column-major)
int ref_mass_props(int type, const float* prm, const float* lpos, const float* lq, float mass, float* out) {
    RigidBodyCollisionComponent col;
    Collider c;
    std::memset(&c, 0, sizeof(c));
    c.position = V(lpos);
    c.orientation = Qt(lq);
    c.geometry = makeGeom(type, prm);
    c.enableSimulation = true;
    col.colliders.push_back(c);
    RigidBodyDynamicComponent d{};
    setMassProps(d, col, mass);
    out[0] = d.invMass;
    out[1] = d.com.x; out[2] = d.com.y; out[3] = d.com.z;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[4 + 3 * i + j] = d.invInertiaTensor[i][j];
    return 0;
}

// One ContactConstraints (up to 4 points) run through the reference's preSolve and
// solve: `iters` biased passes with timeStep h, then `relax` unbiased passes.
// in_f layout: n(3) friction isSoft freq damping | per point x4: r0xn r1xn t r0xt r1xt (15) c tv
// masses: invMass0, invI0(9), invMass1, invI1(9); vel: v0 w0 v1 w1 (12, solver velocities);
// comp: v0 w0 v1 w1 (12, component velocities). out: vel(12) then per point lN lT (8)
int ref_contact_solve(int b0, int b1, int numPoints, const float* in_f, const float* masses,
                      const float* vel, const float* comp, float h, int iters, int relax, float* out) {
    TransformComponent tr0{}, tr1{};
    RigidBodyDynamicComponent d0{}, d1{};
    d0.invMass = masses[0]; d1.invMass = masses[10];
    d0.velocity = V(comp); d0.angularVelocity = V(comp + 3);
    d1.velocity = V(comp + 6); d1.angularVelocity = V(comp + 9);
    int ib0 = b0 >= 0 ? 0 : -1, ib1 = b1 >= 0 ? 1 : -1;
    ContactConstraints cc{tr0, tr1, &d0, &d1, ib0, ib1, V(in_f), in_f[3], in_f[4] != 0.f, in_f[5], in_f[6], numPoints, {}};
    for (int k = 0; k < 4; ++k) {
        const float* p = in_f + 7 + 17 * k;
        auto& pc = cc.contactPointConstraints[k];
        pc.r0 = glm::vec3(0); pc.r1 = glm::vec3(0);
        pc.r0xn = V(p); pc.r1xn = V(p + 3); pc.t = V(p + 6); pc.r0xt = V(p + 9); pc.r1xt = V(p + 12);
        pc.c = p[15]; pc.targetVelocity = p[16];
        pc.totalLambdaN = 0; pc.totalLambdaT = 0;
    }
    MassData md[2];
    md[0].invMass = masses[0]; md[1].invMass = masses[10];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { md[0].invInertiaTensor[i][j] = masses[1 + 3 * i + j]; md[1].invInertiaTensor[i][j] = masses[11 + 3 * i + j]; }
    VelocityData vd[2];
    vd[0].velocity = V(vel); vd[0].angularVelocity = V(vel + 3);
    vd[1].velocity = V(vel + 6); vd[1].angularVelocity = V(vel + 9);
    cc.preSolve(md);
    for (int i = 0; i < iters; ++i) cc.solve(vd, true, h);
    for (int i = 0; i < relax; ++i) cc.solve(vd, false);
    float* o = out;
    o[0] = vd[0].velocity.x; o[1] = vd[0].velocity.y; o[2] = vd[0].velocity.z;
    o[3] = vd[0].angularVelocity.x; o[4] = vd[0].angularVelocity.y; o[5] = vd[0].angularVelocity.z;
    o[6] = vd[1].velocity.x; o[7] = vd[1].velocity.y; o[8] = vd[1].velocity.z;
    o[9] = vd[1].angularVelocity.x; o[10] = vd[1].angularVelocity.y; o[11] = vd[1].angularVelocity.z;
    for (int k = 0; k < 4; ++k) { o[12 + 2 * k] = cc.contactPointConstraints[k].totalLambdaN; o[13 + 2 * k] = cc.contactPointConstraints[k].totalLambdaT; }
    return 0;
}

// GLM restatement of src/Physecs.cpp:450-467 for one body.
// io: pos(3) quat(4 xyzw) v(3) w(3); invI(9); out: v(3) w(3) worldInvI(9)
int ref_integrate_velocity(const float* quat, const float* v, const float* w, const float* invI, float h, float g, float* out) {
    glm::quat q = Qt(quat);
    glm::mat3 invInertia;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) invInertia[i][j] = invI[3 * i + j];
    glm::vec3 vel = V(v), ang = V(w);
    glm::mat3 rot = glm::toMat3(q);
    glm::mat3 invRot = glm::transpose(rot);
    vel += h * glm::vec3(0, -g, 0);
    glm::vec3 omegaLocal = invRot * ang;
    glm::mat3x3 I = glm::inverse(invInertia);
    glm::vec3 f = h * glm::cross(omegaLocal, I * omegaLocal);
    glm::mat3x3 J = I + h * (glm::matrixCross3(omegaLocal) * I - glm::matrixCross3(I * omegaLocal));
    omegaLocal = omegaLocal - solve33(J, f);
    ang = rot * omegaLocal;
    glm::mat3 wI = rot * invInertia * invRot;
    out[0] = vel.x; out[1] = vel.y; out[2] = vel.z;
    out[3] = ang.x; out[4] = ang.y; out[5] = ang.z;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) out[6 + 3 * i + j] = wI[i][j];
    return 0;
}

// GLM restatement of src/Physecs.cpp:501-510 for one body.
// in: pos(3) quat(4) v(3) w(3) pv(3) pw(3) count com(3); out: pos(3) quat(4)
int ref_integrate_position(const float* pos, const float* quat, const float* v, const float* w,
                           const float* pv, const float* pw, int count, const float* com, float h, float* out) {
    glm::vec3 position = V(pos);
    glm::quat orientation = Qt(quat);
    float s = count ? 1.f / count : 1.f;
    position += h * V(v) + s * V(pv);
    glm::vec3 prevComWorld = orientation * V(com);
    orientation += glm::quat(0, 0.5f * (h * V(w) + s * V(pw))) * orientation;
    orientation = glm::normalize(orientation);
    position += prevComWorld - orientation * V(com);
    out[0] = position.x; out[1] = position.y; out[2] = position.z;
    out[3] = orientation.x; out[4] = orientation.y; out[5] = orientation.z; out[6] = orientation.w;
    return 0;
}

int ref_row_scalar(int flags, int b0, int b1, const float* in, const float* masses, float* vel, float* pv, float h,
                   int iters, float* out) {
    switch (flags) {
        case NONE: runRow<NONE>(b0, b1, in, masses, vel, pv, h, iters, out); return 0;
        case ANGULAR: runRow<ANGULAR>(b0, b1, in, masses, vel, pv, h, iters, out); return 0;
        case SOFT: runRow<SOFT>(b0, b1, in, masses, vel, pv, h, iters, out); return 0;
        case LIMITED: runRow<LIMITED>(b0, b1, in, masses, vel, pv, h, iters, out); return 0;
        case ANGULAR | SOFT: runRow<ANGULAR | SOFT>(b0, b1, in, masses, vel, pv, h, iters, out); return 0;
        case ANGULAR | LIMITED: runRow<ANGULAR | LIMITED>(b0, b1, in, masses, vel, pv, h, iters, out); return 0;
    }
    return -1;
}

// ref_row_scalar with a starting accumulated lambda and a per-solve warm-start mask
int ref_row_scalar_ex(int flags, int b0, int b1, const float* in, const float* masses, float* vel, float* pv, float h,
                      int iters, float lambda0, int warmMask, float* out) {
    const unsigned m = (unsigned)warmMask;
    switch (flags) {
        case NONE: runRow<NONE>(b0, b1, in, masses, vel, pv, h, iters, out, lambda0, m); return 0;
        case ANGULAR: runRow<ANGULAR>(b0, b1, in, masses, vel, pv, h, iters, out, lambda0, m); return 0;
        case SOFT: runRow<SOFT>(b0, b1, in, masses, vel, pv, h, iters, out, lambda0, m); return 0;
        case LIMITED: runRow<LIMITED>(b0, b1, in, masses, vel, pv, h, iters, out, lambda0, m); return 0;
        case ANGULAR | SOFT: runRow<ANGULAR | SOFT>(b0, b1, in, masses, vel, pv, h, iters, out, lambda0, m); return 0;
        case ANGULAR | LIMITED: runRow<ANGULAR | LIMITED>(b0, b1, in, masses, vel, pv, h, iters, out, lambda0, m); return 0;
    }
    return -1;
}

// intersectRayAABB (src/Raycast.cpp:5-30): returns the hit, t when hit (else untouched)
int ref_ray_aabb(const float* o, const float* d, const float* mn, const float* mx, float* t) {
    return intersectRayAABB(V(o), V(d), V(mn), V(mx), *t) ? 1 : 0;
}

// intersectRayGeometry (src/Raycast.cpp:141-150)
int ref_ray_geometry(const float* o, const float* d, int type, const float* prm, const float* pos, const float* q,
                     float* t) {
    Geometry g = makeGeom(type, prm);
    return intersectRayGeometry(V(o), V(d), V(pos), Qt(q), g, *t) ? 1 : 0;
}

// setMassProps (src/MassUtil.cpp:183-192) for a body of n colliders; convex params[3] is a
// ref_set_convex_mesh index. out: invMass, com(3), invInertia (9, column-major)
int ref_mass_props_multi(int n, const int32_t* types, const float* prm, const float* lpos, const float* lq,
                         const uint8_t* trig, float mass, float* out) {
    RigidBodyCollisionComponent col;
    for (int i = 0; i < n; ++i) {
        Collider c;
        std::memset(&c, 0, sizeof(c));
        c.position = V(lpos + 3 * i);
        c.orientation = Qt(lq + 4 * i);
        c.geometry = makeGeom(types[i], prm + 4 * i);
        c.isTrigger = trig && trig[i];
        c.enableSimulation = true;
        col.colliders.push_back(c);
    }
    RigidBodyDynamicComponent d{};
    setMassProps(d, col, mass);
    out[0] = d.invMass;
    out[1] = d.com.x; out[2] = d.com.y; out[3] = d.com.z;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[4 + 3 * i + j] = d.invInertiaTensor[i][j];
    return 0;
}

// physecs::overlap (src/Overlap.cpp:203-235)
int ref_overlap(int type0, const float* prm0, const float* pos0, const float* q0, int type1, const float* prm1,
                const float* pos1, const float* q1) {
    Geometry g0 = makeGeom(type0, prm0), g1 = makeGeom(type1, prm1);
    return overlap(V(pos0), Qt(q0), g0, V(pos1), Qt(q1), g1) ? 1 : 0;
}

}  // extern "C"
