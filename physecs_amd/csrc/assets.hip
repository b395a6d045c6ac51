// Host-side asset pipeline: mass properties of a rigid body from its colliders, as
// physecs::setMassProps and computeCOMAndInvInertiaTensor compute them (src/MassUtil.cpp).
// Host code only. Every expression keeps GLM 0.9.9.8's association order, and the
// library is built with -ffp-contract=off. The volume terms keep the reference's
// double-precision detour: glm::pow(float, int) resolves to std::pow(double, double).
#include <cmath>
#include "pgx_internal.h"

namespace {
using namespace pgx;

constexpr float PI_F = 3.14159265358979323846264338327950288f;   // glm::pi<float>()

m3 diag(f3 v) { return mk_m3(mk3(v.x, 0.f, 0.f), mk3(0.f, v.y, 0.f), mk3(0.f, 0.f, v.z)); }

// src/MassUtil.cpp:6-8
m3 inertia_sphere(float mass, float radius) { return diag(s3(2.f * mass * radius * radius / 5.f)); }

// src/MassUtil.cpp:10-20
m3 inertia_capsule(float mass, float halfHeight, float radius) {
    float vc = 2 * radius * radius * PI_F * halfHeight;
    float vs = (float)(4 * std::pow((double)radius, 3.0) * PI_F / 3.f);
    float v = vc + vs;
    float mc = vc * mass / v;
    float ms = vs * mass / v;
    float ixx = mc * (halfHeight * halfHeight / 3.f + radius * radius / 4.f) +
                ms * (halfHeight * halfHeight + 3.f * halfHeight * radius / 4.f + 2 * radius * radius / 5.f);
    float iyy = mc * radius * radius / 2.f + ms * 2 * radius * radius / 5.f;
    return diag(mk3(ixx, iyy, ixx));
}

// src/MassUtil.cpp:22-28
m3 inertia_box(float mass, f3 he) {
    float m = mass / 12.f;
    float x = he.x * he.x, y = he.y * he.y, z = he.z * he.z;
    return diag(mk3(m * (y + z), m * (x + z), m * (x + y)));
}

// src/MassUtil.cpp:30-72, about the origin
m3 inertia_tetrahedron(float mass, const f3 (&v)[4]) {
    f3 d = s3(0.f);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j <= i; j++) d = add(d, mk3(v[j].x * v[i].x, v[j].y * v[i].y, v[j].z * v[i].z));
    float a = (d.y + d.z) / 10.f, b = (d.x + d.z) / 10.f, c = (d.x + d.y) / 10.f;
    float ap = 0, bp = 0, cp = 0;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float mult = i == j ? 2 : 1;
            ap += mult * v[i].y * v[j].z;
            bp += mult * v[i].x * v[j].z;
            cp += mult * v[i].x * v[j].y;
        }
    ap /= 20.f; bp /= 20.f; cp /= 20.f;
    m3 I = mk_m3(mk3(a, -bp, -cp), mk3(-bp, b, -ap), mk3(-cp, -ap, c));
    return mscale(mass, I);
}

// glm::mat3(length2(r)) - outerProduct(r, r)
m3 parallel_axis(f3 r) {
    float l2 = dot(r, r);
    m3 o = mk_m3(mul(r, r.x), mul(r, r.y), mul(r, r.z));
    return msub(diag(s3(l2)), o);
}

struct ConvexTables {
    const int32_t *vertStart, *faceStart, *idxStart, *idx;
    const float* verts;
};

// ConvexMeshVertices<4> pads its buffer to a multiple of 4 with the last vertex; the
// centre sums that whole buffer and divides by the unpadded count (src/MassUtil.cpp:99-103)
f3 convex_center(const ConvexTables& T, int mesh, f3 scale) {
    const int v0 = T.vertStart[mesh], nv = T.vertStart[mesh + 1] - v0;
    const int padded = (nv + 3) / 4 * 4;
    f3 c = s3(0.f);
    for (int i = 0; i < padded; ++i) {
        const int k = v0 + (i < nv ? i : nv - 1);
        c = add(c, mk3(scale.x * T.verts[3 * k], scale.y * T.verts[3 * k + 1], scale.z * T.verts[3 * k + 2]));
    }
    const float n = (float)nv;
    return mk3(c.x / n, c.y / n, c.z / n);
}

f3 convex_vertex(const ConvexTables& T, int mesh, f3 scale, int local) {
    const int k = T.vertStart[mesh] + local;
    return mk3(scale.x * T.verts[3 * k], scale.y * T.verts[3 * k + 1], scale.z * T.verts[3 * k + 2]);
}

// the fan tetrahedra of a convex collider: f(volume, v0, v1, v2) per triangle
template <class F>
void convex_fan(const ConvexTables& T, int mesh, f3 scale, f3 center, F&& f) {
    for (int fc = T.faceStart[mesh]; fc < T.faceStart[mesh + 1]; ++fc) {
        const int i0 = T.idxStart[fc], ni = T.idxStart[fc + 1] - i0;
        const f3 v0 = convex_vertex(T, mesh, scale, T.idx[i0]);
        for (int i = 1; i < ni - 1; i++) {
            const f3 v1 = convex_vertex(T, mesh, scale, T.idx[i0 + i]);
            const f3 v2 = convex_vertex(T, mesh, scale, T.idx[i0 + i + 1]);
            const float volume = std::fabs(dot(cross(sub(v0, center), sub(v1, center)), sub(v2, center))) / 6.f;
            f(volume, v0, v1, v2);
        }
    }
}

float sphere_volume(float r) { return (float)(4 * PI_F * std::pow((double)r, 3.0) / 3.f); }
float capsule_volume(float hh, float r) {
    return (float)(4 * PI_F * std::pow((double)r, 3.0) / 3.f + r * r * PI_F * hh * 2.f);
}
}  // namespace

extern "C" int pgx_mass_properties(int n, const int32_t* types, const float* params, const float* local_pos,
                                   const float* local_quat, const uint8_t* is_trigger, int n_convex,
                                   const int32_t* convex_vert_start, const float* convex_verts,
                                   const int32_t* convex_face_start, const int32_t* face_idx_start,
                                   const int32_t* face_idx, float mass, float* inv_mass, float* com_out,
                                   float* inv_inertia) {
    if (n < 0 || (n > 0 && (!types || !params || !local_pos || !local_quat))) return PGX_ERR_INVALID;
    if (!inv_mass || !com_out || !inv_inertia) return PGX_ERR_INVALID;
    for (int c = 0; c < n; ++c) {
        if (types[c] < PGX_SPHERE || types[c] > PGX_TRIANGLE_MESH) return PGX_ERR_INVALID;
        if (types[c] == PGX_CONVEX_MESH) {
            const float m = params[4 * c + 3];
            if (!convex_vert_start || !convex_verts || !convex_face_start || !face_idx_start || !face_idx ||
                !(m >= 0.f && m < (float)n_convex) || m != (float)(int)m)
                return PGX_ERR_INVALID;
        }
    }
    const ConvexTables T{convex_vert_start, convex_face_start, face_idx_start, face_idx, convex_verts};
    auto pos = [&](int c) { return mk3(local_pos[3 * c], local_pos[3 * c + 1], local_pos[3 * c + 2]); };
    auto prm = [&](int c, int k) { return params[4 * c + k]; };
    // src/MassUtil.cpp:74-125: the centre of mass, volume-weighted; triangle meshes add nothing
    f3 com = s3(0.f);
    float total = 0;
    for (int c = 0; c < n; ++c) {
        if (is_trigger && is_trigger[c]) continue;
        switch (types[c]) {
            case PGX_SPHERE: {
                float v = sphere_volume(prm(c, 0));
                com = add(com, mul(pos(c), v));
                total += v;
            } break;
            case PGX_CAPSULE: {
                float v = capsule_volume(prm(c, 0), prm(c, 1));
                com = add(com, mul(pos(c), v));
                total += v;
            } break;
            case PGX_BOX: {
                float v = prm(c, 0) * prm(c, 1) * prm(c, 2);   // the reference's: half extents, no factor 8
                com = add(com, mul(pos(c), v));
                total += v;
            } break;
            case PGX_CONVEX_MESH: {
                const int mesh = (int)prm(c, 3);
                const f3 scale = mk3(prm(c, 0), prm(c, 1), prm(c, 2));
                const f3 center = convex_center(T, mesh, scale);
                convex_fan(T, mesh, scale, center, [&](float v, f3 v0, f3 v1, f3 v2) {
                    const f3 centroid = divs(add(add(add(center, v0), v1), v2), 4.f);
                    com = add(com, mul(add(pos(c), centroid), v));
                    total += v;
                });
            } break;
            default: break;
        }
    }
    com = divs(com, total);
    // src/MassUtil.cpp:127-181: inertia about the centre of mass (parallel axis theorem)
    m3 I = mzero();
    for (int c = 0; c < n; ++c) {
        if (is_trigger && is_trigger[c]) continue;
        const qf q{local_quat[4 * c], local_quat[4 * c + 1], local_quat[4 * c + 2], local_quat[4 * c + 3]};
        switch (types[c]) {
            case PGX_SPHERE: {
                float v = sphere_volume(prm(c, 0));
                float m = mass * v / total;
                const f3 r = sub(com, pos(c));
                I = madd(I, madd(inertia_sphere(m, prm(c, 0)), mscale(m, parallel_axis(r))));
            } break;
            case PGX_CAPSULE:
            case PGX_BOX: {
                const bool cap = types[c] == PGX_CAPSULE;
                float v = cap ? capsule_volume(prm(c, 0), prm(c, 1)) : prm(c, 0) * prm(c, 1) * prm(c, 2);
                float m = mass * v / total;
                const m3 local = cap ? inertia_capsule(m, prm(c, 0), prm(c, 1))
                                     : inertia_box(m, mk3(prm(c, 0), prm(c, 1), prm(c, 2)));
                const m3 rot = mat3_cast(q);
                const f3 r = sub(com, pos(c));
                I = madd(I, madd(mulmm(mulmm(rot, local), transpose(rot)), mscale(m, parallel_axis(r))));
            } break;
            case PGX_CONVEX_MESH: {
                const int mesh = (int)prm(c, 3);
                const f3 scale = mk3(prm(c, 0), prm(c, 1), prm(c, 2));
                const f3 center = convex_center(T, mesh, scale);
                const f3 d = sub(pos(c), com);
                convex_fan(T, mesh, scale, center, [&](float v, f3 v0, f3 v1, f3 v2) {
                    float m = mass * v / total;
                    const f3 tet[4] = {add(center, d), add(v0, d), add(v1, d), add(v2, d)};
                    I = madd(I, inertia_tetrahedron(m, tet));
                });
            } break;
            default: break;
        }
    }
    const m3 inv = inverse(I);
    *inv_mass = 1.f / mass;   // src/MassUtil.cpp:185
    com_out[0] = com.x; com_out[1] = com.y; com_out[2] = com.z;
    const f3 cols[3] = {inv.c0, inv.c1, inv.c2};
    for (int k = 0; k < 3; ++k) {
        inv_inertia[3 * k] = cols[k].x; inv_inertia[3 * k + 1] = cols[k].y; inv_inertia[3 * k + 2] = cols[k].z;
    }
    return PGX_OK;
}
