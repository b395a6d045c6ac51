// Host-side asset pipeline: mass properties of a rigid body from its colliders, as
// physecs::setMassProps and computeCOMAndInvInertiaTensor compute them (src/MassUtil.cpp).
// Host code only. Every expression keeps GLM 0.9.9.8's association order, and the
// library is built with -ffp-contract=off. The volume terms keep the reference's
// double-precision detour: glm::pow(float, int) resolves to std::pow(double, double).
#include <cmath>
#include <map>
#include <numeric>
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

// ------------------------------------------------------------------ convex hulls
// ConvexMesh data (include/Physecs/ConvexMesh.h:7-34) for the convex hull of a point cloud;
// the reference builds none and takes vertices and faces as given. Incremental hull in
// double precision: a tetrahedron of extreme points, then every other point, farthest from
// the hull first, replaces the triangles it sees by a fan to their horizon. Coplanar
// neighbouring triangles are merged into polygons.
namespace {
struct D3 { double x, y, z; };
D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dlen(D3 a) { return std::sqrt(ddot(a, a)); }

struct Tri {
    int v[3];
    D3 n;          // outward unit normal
    double d;      // plane offset: dot(n, x) = d
    bool alive;
};

Tri make_tri(const std::vector<D3>& P, int a, int b, int c, D3 inside) {
    Tri t{{a, b, c}, {0, 0, 0}, 0, true};
    D3 n = dcross(dsub(P[b], P[a]), dsub(P[c], P[a]));
    double l = dlen(n);
    n = {n.x / l, n.y / l, n.z / l};
    if (ddot(n, dsub(inside, P[a])) > 0) {   // keep the normal outward: swap the winding
        std::swap(t.v[1], t.v[2]);
        n = {-n.x, -n.y, -n.z};
    }
    t.n = n;
    t.d = ddot(n, P[a]);
    return t;
}
}  // namespace

extern "C" int pgx_convex_hull(const float* points, int n, float* verts, int* n_verts, int* face_start,
                               int* face_idx, float* normals, float* centroids, int* n_faces) {
    if (n < 4 || !points || !verts || !n_verts || !face_start || !face_idx || !normals || !centroids || !n_faces)
        return PGX_ERR_INVALID;
    std::vector<D3> P((size_t)n);
    for (int i = 0; i < n; ++i) P[i] = {points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    double ext = 0;
    for (int i = 0; i < n; ++i) ext = std::max({ext, std::fabs(P[i].x), std::fabs(P[i].y), std::fabs(P[i].z)});
    const double eps = 1e-10 * std::max(ext, 1e-30);
    // the initial tetrahedron: extreme x, farthest from their line, farthest from their plane
    int i0 = 0, i1 = 0;
    for (int i = 1; i < n; ++i) {
        if (P[i].x < P[i0].x) i0 = i;
        if (P[i].x > P[i1].x) i1 = i;
    }
    if (i0 == i1) return PGX_ERR_INVALID;
    int i2 = -1, i3 = -1;
    double best = eps;
    for (int i = 0; i < n; ++i) {
        double d = dlen(dcross(dsub(P[i], P[i0]), dsub(P[i1], P[i0])));
        if (d > best) { best = d; i2 = i; }
    }
    if (i2 < 0) return PGX_ERR_INVALID;   // collinear
    const D3 pn = dcross(dsub(P[i1], P[i0]), dsub(P[i2], P[i0]));
    best = eps * dlen(pn);
    for (int i = 0; i < n; ++i) {
        double d = std::fabs(ddot(pn, dsub(P[i], P[i0])));
        if (d > best) { best = d; i3 = i; }
    }
    if (i3 < 0) return PGX_ERR_INVALID;   // coplanar
    const D3 inside = {(P[i0].x + P[i1].x + P[i2].x + P[i3].x) / 4, (P[i0].y + P[i1].y + P[i2].y + P[i3].y) / 4,
                       (P[i0].z + P[i1].z + P[i2].z + P[i3].z) / 4};
    std::vector<Tri> T = {make_tri(P, i0, i1, i2, inside), make_tri(P, i0, i1, i3, inside),
                          make_tri(P, i0, i2, i3, inside), make_tri(P, i1, i2, i3, inside)};
    // the other points, farthest from the tetrahedron's centre first
    std::vector<int> order;
    for (int i = 0; i < n; ++i)
        if (i != i0 && i != i1 && i != i2 && i != i3) order.push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return ddot(dsub(P[a], inside), dsub(P[a], inside)) > ddot(dsub(P[b], inside), dsub(P[b], inside));
    });
    for (int p : order) {
        std::vector<int> vis;
        for (int t = 0; t < (int)T.size(); ++t)
            if (T[t].alive && ddot(T[t].n, P[p]) - T[t].d > eps) vis.push_back(t);
        if (vis.empty()) continue;   // inside the hull
        // horizon: directed edges of visible triangles whose reverse is on no visible one
        std::map<std::pair<int, int>, int> edges;
        for (int t : vis)
            for (int k = 0; k < 3; ++k) edges[{T[t].v[k], T[t].v[(k + 1) % 3]}] = t;
        for (int t : vis) T[t].alive = false;
        for (const auto& e : edges) {
            if (edges.count({e.first.second, e.first.first})) continue;
            Tri nt{{e.first.first, e.first.second, p}, {0, 0, 0}, 0, true};
            D3 nn = dcross(dsub(P[nt.v[1]], P[nt.v[0]]), dsub(P[nt.v[2]], P[nt.v[0]]));
            double l = dlen(nn);
            if (l <= eps * eps) continue;
            nt.n = {nn.x / l, nn.y / l, nn.z / l};
            nt.d = ddot(nt.n, P[nt.v[0]]);
            T.push_back(nt);
        }
    }
    // merge coplanar triangles that share an edge (union-find over the live triangles)
    std::vector<int> live;
    for (int t = 0; t < (int)T.size(); ++t)
        if (T[t].alive) live.push_back(t);
    std::vector<int> parent(T.size());
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    std::map<std::pair<int, int>, int> owner;
    for (int t : live)
        for (int k = 0; k < 3; ++k) owner[{T[t].v[k], T[t].v[(k + 1) % 3]}] = t;
    for (int t : live)
        for (int k = 0; k < 3; ++k) {
            auto it = owner.find({T[t].v[(k + 1) % 3], T[t].v[k]});
            if (it == owner.end()) continue;
            const Tri& u = T[it->second];
            if (dlen(dsub(T[t].n, u.n)) < 1e-6 && std::fabs(T[t].d - u.d) < 1e-6 * std::max(ext, 1e-30))
                parent[find(t)] = find(it->second);
        }
    // hull vertices in ascending input order
    std::vector<int> used((size_t)n, 0), remap((size_t)n, -1);
    for (int t : live)
        for (int k = 0; k < 3; ++k) used[T[t].v[k]] = 1;
    int nv = 0;
    for (int i = 0; i < n; ++i)
        if (used[i]) {
            remap[i] = nv;
            verts[3 * nv] = points[3 * i]; verts[3 * nv + 1] = points[3 * i + 1]; verts[3 * nv + 2] = points[3 * i + 2];
            ++nv;
        }
    // one polygon per group: its vertices ordered clockwise seen from outside
    std::map<int, std::vector<int>> groups;
    for (int t : live) groups[find(t)].push_back(t);
    int nf = 0, ni = 0;
    face_start[0] = 0;
    for (const auto& g : groups) {
        const Tri& t0 = T[g.first];
        std::vector<int> vs;
        for (int t : g.second)
            for (int k = 0; k < 3; ++k)
                if (std::find(vs.begin(), vs.end(), T[t].v[k]) == vs.end()) vs.push_back(T[t].v[k]);
        D3 c = {0, 0, 0};
        for (int v : vs) c = {c.x + P[v].x, c.y + P[v].y, c.z + P[v].z};
        c = {c.x / vs.size(), c.y / vs.size(), c.z / vs.size()};
        D3 u = dsub(P[vs[0]], c);
        const double ul = dlen(u);
        u = {u.x / ul, u.y / ul, u.z / ul};
        const D3 w = dcross(t0.n, u);
        std::vector<std::pair<double, int>> ang;
        for (int v : vs) ang.push_back({std::atan2(ddot(dsub(P[v], c), w), ddot(dsub(P[v], c), u)), v});
        std::sort(ang.begin(), ang.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        for (const auto& a : ang) face_idx[ni++] = remap[a.second];
        face_start[++nf] = ni;
        normals[3 * (nf - 1)] = (float)t0.n.x; normals[3 * (nf - 1) + 1] = (float)t0.n.y;
        normals[3 * (nf - 1) + 2] = (float)t0.n.z;
        centroids[3 * (nf - 1)] = (float)c.x; centroids[3 * (nf - 1) + 1] = (float)c.y;
        centroids[3 * (nf - 1) + 2] = (float)c.z;
    }
    *n_verts = nv;
    *n_faces = nf;
    return PGX_OK;
}
