// Host-side build of the static triangle meshes: the triangles and the BVH that the
// reference's TriangleMesh constructor builds (src/TriangleMesh.cpp:7-161), built once at
// world creation and uploaded. The narrowphase queries the BVH on the device.
//
// The build follows the reference step for step -- 6 SAH buckets per axis over the node's
// bounds, the in-place two-sided partition, depth-first child creation -- because the
// triangle order it leaves is observable: a contact manifold carries the post-build
// triangle index (it keys the restitution cache and the solve order), so a different
// build would give different manifolds.
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <vector>
#include "pgx_internal.h"

namespace pgx {

namespace {
struct HBounds { f3 mn, mx; };
struct HTri { int32_t idx[3]; HBounds b; f3 n, c; };
struct HNode { HBounds b; int triCount; int index; };

HBounds hunion(const HBounds& a, const HBounds& b) { return {vmin(a.mn, b.mn), vmax(a.mx, b.mx)}; }
HBounds hempty() { return {s3(FLT_MAX), s3(-FLT_MAX)}; }
float harea(const HBounds& b) {                                  // Bounds::area, src/Bounds.cpp:28-31
    f3 d = sub(b.mx, b.mn);
    return 2.f * (d.x * d.y + d.y * d.z + d.z * d.x);
}

struct Builder {
    std::vector<HTri>& tris;
    std::vector<HNode>& bvh;
    static constexpr int NB = 6;

    void node_bounds(int id) {                                   // :7-14
        HBounds b = hempty();
        for (int i = 0; i < bvh[id].triCount; ++i) b = hunion(b, tris[bvh[id].index + i].b);
        bvh[id].b = b;
    }
    bool fill(int id, int axis, HBounds* bb, int* bc) {          // :16-38
        for (int i = 0; i < NB; ++i) { bb[i] = hempty(); bc[i] = 0; }
        float start = at(bvh[id].b.mn, axis), end = at(bvh[id].b.mx, axis);
        float len = end - start;
        if (!len) return false;
        for (int i = 0; i < bvh[id].triCount; ++i) {
            const HTri& t = tris[bvh[id].index + i];
            int k = std::min(NB - 1, static_cast<int>((at(t.c, axis) - start) / len * NB));
            bb[k] = hunion(bb[k], t.b);
            bc[k]++;
        }
        return true;
    }
    bool choose(int id, int& axisOut, int& split, HBounds& bl, HBounds& br, int& cl, int& cr) {   // :40-94
        float best = FLT_MAX;
        for (int axis = 0; axis < 3; ++axis) {
            HBounds bb[NB];
            int bc[NB];
            if (!fill(id, axis, bb, bc)) continue;
            for (int s = 1; s < NB; ++s) {
                HBounds l = hempty(), r = hempty();
                int nl = 0, nr = 0;
                for (int i = 0; i < NB; ++i) {
                    if (i < s) { l = hunion(l, bb[i]); nl += bc[i]; }
                    else { r = hunion(r, bb[i]); nr += bc[i]; }
                }
                if (!nl || !nr) continue;
                float cost = nl * harea(l) + nr * harea(r);
                if (cost < best) { best = cost; axisOut = axis; split = s; bl = l; br = r; cl = nl; cr = nr; }
            }
        }
        return cl && cr;
    }
    void subdivide(int root) {                                   // :96-137, with an explicit stack
        std::vector<int> todo{root};
        while (!todo.empty()) {
            int id = todo.back();
            todo.pop_back();
            int axis = 0, split = 0, cl = 0, cr = 0;
            HBounds bl{}, br{};
            if (!choose(id, axis, split, bl, br, cl, cr)) continue;
            float start = at(bvh[id].b.mn, axis), end = at(bvh[id].b.mx, axis);
            float len = end - start;
            int i = bvh[id].index, j = i + bvh[id].triCount - 1;
            while (i <= j) {
                int k = static_cast<int>((at(tris[i].c, axis) - start) / len * NB);
                if (k < split) ++i;
                else std::swap(tris[i], tris[j--]);
            }
            int first = bvh[id].index;
            bvh.push_back({bl, cl, first});
            bvh.push_back({br, cr, i});
            bvh[id].triCount = 0;
            bvh[id].index = (int)bvh.size() - 2;
            // the reference recurses into the left child first, then the right: the left
            // subtree's nodes are created before the right child's
            todo.push_back(bvh[id].index + 1);
            todo.push_back(bvh[id].index);
        }
    }
};
}  // namespace

// Builds every mesh of the descriptor and uploads vertices, triangles (post-build order),
// triangle normals / centroids and BVH nodes. Returns the deepest BVH path (stack bound).
// One mesh as the constructor leaves it: triangles in post-build order with their normals
// and centroids, the BVH nodes ({min, triCount bits}, {max, index bits}), the tree depth.
struct BuiltMesh {
    std::vector<int4> tris;
    std::vector<float4> triN, triC, nodes;
    int depth = 0;
};

BuiltMesh build_mesh(const float* vp, int nv, const int32_t* ip, int nt) {
    std::vector<f3> v(nv);
    for (int i = 0; i < nv; ++i) v[i] = mk3(vp[3 * i], vp[3 * i + 1], vp[3 * i + 2]);
    std::vector<HTri> ht(nt);
    for (int t = 0; t < nt; ++t) {
        HTri& h = ht[t];
        for (int k = 0; k < 3; ++k) h.idx[k] = ip[3 * t + k];
        f3 a = v[h.idx[0]], b = v[h.idx[1]], c = v[h.idx[2]];
        h.b = {vmin(vmin(a, b), c), vmax(vmax(a, b), c)};        // getBoundsTriangle
        h.n = normalize(cross(sub(b, a), sub(c, a)));
        h.c = divs(add(add(a, b), c), 3.f);
    }
    std::vector<HNode> bvh;
    bvh.reserve(2 * (size_t)nt + 1);
    bvh.push_back({hempty(), nt, 0});
    Builder bld{ht, bvh};
    bld.node_bounds(0);
    bld.subdivide(0);
    BuiltMesh out;
    // depth of the tree (bounds the device traversal stack)
    std::vector<int> depth(bvh.size(), 0);
    for (size_t n = 0; n < bvh.size(); ++n)
        if (!bvh[n].triCount) {
            depth[bvh[n].index] = depth[bvh[n].index + 1] = depth[n] + 1;
            out.depth = std::max(out.depth, depth[n] + 1);
        }
    for (const HTri& h : ht) {
        out.tris.push_back(make_int4(h.idx[0], h.idx[1], h.idx[2], 0));
        out.triN.push_back(f4(h.n, 0.f));
        out.triC.push_back(f4(h.c, 0.f));
    }
    auto asf = [](int i) { float f; std::memcpy(&f, &i, sizeof f); return f; };
    for (const HNode& n : bvh) {
        out.nodes.push_back(make_float4(n.b.mn.x, n.b.mn.y, n.b.mn.z, asf(n.triCount)));
        out.nodes.push_back(make_float4(n.b.mx.x, n.b.mx.y, n.b.mx.z, asf(n.index)));
    }
    return out;
}

// the cache's layout: header, then tris, triN, triC (nt each), nodes (2 per node)
constexpr uint32_t MESH_MAGIC = 0x4d584750u;   // "PGXM"
constexpr uint32_t MESH_VERSION = 2;
struct MeshCacheHeader {
    uint32_t magic, version;
    int32_t nv, nt, nNodes, depth;
    uint64_t bytes;         // the whole cache, header included
    uint64_t hash;          // FNV-1a of the vertex and index bytes the cache was built from
    uint64_t payloadHash;   // FNV-1a of everything after the header
};

static uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}
constexpr uint64_t FNV_BASIS = 1469598103934665603ull;

uint64_t mesh_hash(const float* vp, int nv, const int32_t* ip, int nt) {
    return fnv(fnv(FNV_BASIS, vp, 12 * (size_t)nv), ip, 12 * (size_t)nt);
}

size_t cache_bytes(int nt, int nNodes) {
    return sizeof(MeshCacheHeader) + 16 * (3 * (size_t)nt + 2 * (size_t)nNodes);
}

// The mesh from a cache of `len` bytes written for these vertices and indices. Throws
// PGX_ERR_INVALID unless the cache is one: its length, the hashes of the mesh it was built
// from and of its payload, and the structure the device walks -- every internal node's
// children inside the node list and after it, every leaf's triangle range inside the
// triangles, every triangle's vertices inside the mesh, the depth recomputed from the nodes.
BuiltMesh load_mesh(const void* cache, size_t len, const float* vp, int nv, const int32_t* ip, int nt) {
    const auto bad = [](const char* why) {
        return PgxError{PGX_ERR_INVALID, std::string("triangle-mesh cache rejected: ") + why};
    };
    MeshCacheHeader h;
    if (len < sizeof h) throw bad("shorter than its header");
    std::memcpy(&h, cache, sizeof h);
    if (h.magic != MESH_MAGIC || h.version != MESH_VERSION) throw bad("not a cache of this version");
    if (h.nv != nv || h.nt != nt || h.hash != mesh_hash(vp, nv, ip, nt)) throw bad("built from another mesh");
    if (h.nNodes < 1 || h.nNodes > 2 * nt) throw bad("node count out of range");
    if (h.bytes != cache_bytes(nt, h.nNodes) || len != h.bytes) throw bad("length does not match its contents");
    const char* p = static_cast<const char*>(cache) + sizeof h;
    if (fnv(FNV_BASIS, p, len - sizeof h) != h.payloadHash) throw bad("payload hash differs (corrupted)");
    BuiltMesh m;
    auto take = [&](auto& vec, size_t n) {
        vec.resize(n);
        std::memcpy(vec.data(), p, sizeof(vec[0]) * n);
        p += sizeof(vec[0]) * n;
    };
    take(m.tris, nt); take(m.triN, nt); take(m.triC, nt); take(m.nodes, 2 * (size_t)h.nNodes);
    for (const int4& t : m.tris)
        if (t.x < 0 || t.x >= nv || t.y < 0 || t.y >= nv || t.z < 0 || t.z >= nv) throw bad("vertex index out of range");
    auto asi = [](float f) { int i; std::memcpy(&i, &f, sizeof i); return i; };
    std::vector<int> depth(h.nNodes, 0);
    int maxDepth = 0;
    for (int n = 0; n < h.nNodes; ++n) {
        const int count = asi(m.nodes[2 * n].w), index = asi(m.nodes[2 * n + 1].w);
        if (count < 0) throw bad("negative triangle count");
        if (count == 0) {       // internal: children index and index + 1, stored after the node
            if (index <= n || index + 1 >= h.nNodes) throw bad("child index out of range");
            depth[index] = depth[index + 1] = depth[n] + 1;
            maxDepth = std::max(maxDepth, depth[n] + 1);
        } else if (index < 0 || (long)index + count > nt) {
            throw bad("leaf triangle range out of range");
        }
    }
    if (maxDepth != h.depth) throw bad("depth differs from its nodes");
    m.depth = h.depth;
    return m;
}

int upload_triangle_meshes(World& w, const pgx_scene_desc* sc, hipStream_t s) {
    std::vector<float4> verts, triN, triC, nodes;
    std::vector<int4> tris, meshes;
    int maxDepth = 0;
    for (int m = 0; m < sc->n_trimesh; ++m) {
        int v0 = sc->tri_vert_start[m], nv = sc->tri_vert_start[m + 1] - v0;
        int t0 = sc->tri_start[m], nt = sc->tri_start[m + 1] - t0;
        const float* vp = sc->tri_verts + 3 * (size_t)v0;
        const int32_t* ip = sc->tri_idx + 3 * (size_t)t0;
        const void* cache = sc->tri_bvh_cache ? sc->tri_bvh_cache[m] : nullptr;
        if (cache && !sc->tri_bvh_cache_bytes)
            throw PgxError{PGX_ERR_INVALID, "tri_bvh_cache given without tri_bvh_cache_bytes"};
        const BuiltMesh bm = cache ? load_mesh(cache, sc->tri_bvh_cache_bytes[m], vp, nv, ip, nt)
                                   : build_mesh(vp, nv, ip, nt);
        maxDepth = std::max(maxDepth, bm.depth);
        int vbase = (int)verts.size(), tbase = (int)tris.size(), nbase = (int)nodes.size() / 2;
        meshes.push_back(make_int4(vbase, nv, tbase, nbase));
        for (int i = 0; i < nv; ++i) verts.push_back(make_float4(vp[3 * i], vp[3 * i + 1], vp[3 * i + 2], 0.f));
        tris.insert(tris.end(), bm.tris.begin(), bm.tris.end());
        triN.insert(triN.end(), bm.triN.begin(), bm.triN.end());
        triC.insert(triC.end(), bm.triC.begin(), bm.triC.end());
        nodes.insert(nodes.end(), bm.nodes.begin(), bm.nodes.end());
    }
    auto up = [&](auto& buf, const auto& vec) {
        buf.alloc(vec.size());
        if (!vec.empty())
            PGX_HIP(hipMemcpyAsync(buf.p, vec.data(), sizeof(vec[0]) * vec.size(), hipMemcpyHostToDevice, s));
    };
    up(w.tmVerts, verts); up(w.tmTris, tris); up(w.tmTriN, triN); up(w.tmTriC, triC); up(w.tmNodes, nodes);
    up(w.tmMesh, meshes);
    PGX_HIP(hipStreamSynchronize(s));   // the host vectors above end with this function
    w.nTriMesh = sc->n_trimesh;
    return maxDepth;
}

}  // namespace pgx

extern "C" int pgx_trimesh_cache(const float* verts, int n_verts, const int32_t* tris, int n_tris, void* out,
                                 size_t cap, size_t* bytes) {
    using namespace pgx;
    if (!verts || !tris || !bytes || n_verts < 3 || n_tris < 1) return PGX_ERR_INVALID;
    for (int t = 0; t < 3 * n_tris; ++t)
        if (tris[t] < 0 || tris[t] >= n_verts) return PGX_ERR_INVALID;
    try {
        const BuiltMesh m = build_mesh(verts, n_verts, tris, n_tris);
        const int nNodes = (int)m.nodes.size() / 2;
        *bytes = cache_bytes(n_tris, nNodes);
        if (!out) return PGX_OK;
        if (cap < *bytes) return PGX_ERR_CAPACITY;
        MeshCacheHeader h{MESH_MAGIC, MESH_VERSION, n_verts, n_tris, nNodes, m.depth, *bytes,
                          mesh_hash(verts, n_verts, tris, n_tris), 0};
        char* p = static_cast<char*>(out) + sizeof h;
        char* const payload = p;
        auto put = [&](const auto& vec) {
            std::memcpy(p, vec.data(), sizeof(vec[0]) * vec.size());
            p += sizeof(vec[0]) * vec.size();
        };
        put(m.tris); put(m.triN); put(m.triC); put(m.nodes);
        h.payloadHash = fnv(FNV_BASIS, payload, (size_t)(p - payload));
        std::memcpy(out, &h, sizeof h);
        return PGX_OK;
    } catch (...) {
        return PGX_ERR_INVALID;
    }
}

extern "C" int pgx_trimesh_cache_validate(const float* verts, int n_verts, const int32_t* tris, int n_tris,
                                          const void* cache, size_t bytes) {
    using namespace pgx;
    if (!verts || !tris || !cache || n_verts < 3 || n_tris < 1) return PGX_ERR_INVALID;
    try {
        (void)load_mesh(cache, bytes, verts, n_verts, tris, n_tris);
        return PGX_OK;
    } catch (...) {
        return PGX_ERR_INVALID;
    }
}

extern "C" int pgx_trimesh_cache_pack(const float* verts, int n_verts, const int32_t* tris, int n_tris,
                                      const float* tri_normals, const float* tri_centroids, const float* node_bounds,
                                      const int32_t* node_tri_count, const int32_t* node_index, int n_nodes, void* out,
                                      size_t cap, size_t* bytes) {
    using namespace pgx;
    if (!verts || !tris || !tri_normals || !tri_centroids || !node_bounds || !node_tri_count || !node_index || !bytes ||
        n_verts < 3 || n_tris < 1 || n_nodes < 1 || n_nodes > 2 * n_tris)
        return PGX_ERR_INVALID;
    *bytes = cache_bytes(n_tris, n_nodes);
    if (!out) return PGX_OK;
    if (cap < *bytes) return PGX_ERR_CAPACITY;
    try {
        BuiltMesh m;
        for (int t = 0; t < n_tris; ++t) {
            m.tris.push_back(make_int4(tris[3 * t], tris[3 * t + 1], tris[3 * t + 2], 0));
            m.triN.push_back(make_float4(tri_normals[3 * t], tri_normals[3 * t + 1], tri_normals[3 * t + 2], 0.f));
            m.triC.push_back(make_float4(tri_centroids[3 * t], tri_centroids[3 * t + 1], tri_centroids[3 * t + 2], 0.f));
        }
        auto asf = [](int i) { float f; std::memcpy(&f, &i, sizeof f); return f; };
        std::vector<int> depth(n_nodes, 0);
        for (int n = 0; n < n_nodes; ++n) {
            const float* b = node_bounds + 6 * (size_t)n;
            m.nodes.push_back(make_float4(b[0], b[1], b[2], asf(node_tri_count[n])));
            m.nodes.push_back(make_float4(b[3], b[4], b[5], asf(node_index[n])));
            if (node_tri_count[n] == 0) {
                const int c = node_index[n];
                if (c <= n || c + 1 >= n_nodes) return PGX_ERR_INVALID;
                depth[c] = depth[c + 1] = depth[n] + 1;
                m.depth = std::max(m.depth, depth[n] + 1);
            }
        }
        MeshCacheHeader h{MESH_MAGIC, MESH_VERSION, n_verts, n_tris, n_nodes, m.depth, *bytes,
                          mesh_hash(verts, n_verts, tris, n_tris), 0};
        char* p = static_cast<char*>(out) + sizeof h;
        char* const payload = p;
        auto put = [&](const auto& vec) {
            std::memcpy(p, vec.data(), sizeof(vec[0]) * vec.size());
            p += sizeof(vec[0]) * vec.size();
        };
        put(m.tris); put(m.triN); put(m.triC); put(m.nodes);
        h.payloadHash = fnv(FNV_BASIS, payload, (size_t)(p - payload));
        std::memcpy(out, &h, sizeof h);
        (void)load_mesh(out, *bytes, verts, n_verts, tris, n_tris);   // the structure checks world creation runs
        return PGX_OK;
    } catch (...) {
        return PGX_ERR_INVALID;
    }
}
