// Scene raycasts on the device: Scene::raycastClosest (src/Physecs.cpp:571-621) for a batch
// of rays, each in one world. A ray walks the LBVH of its world's dynamic colliders (built
// from the current bounds for the query) and then its world's non-dynamic colliders. The
// reference walks its dynamic AABB tree; a leaf is reached when the ray enters the leaf's
// own bounds within maxDistance, and every ancestor's box holds those bounds, so pruning
// with the same ray-box test on the LBVH's boxes accepts the same hits. Of the accepted
// hits the closest is kept; equal distances go to the lower collider index (the
// reference settles them by its tree's shape). The filter rejects one entity per ray.
#include <cstring>
#include "pgx_internal.h"

using namespace pgx;

namespace {

__device__ __forceinline__ float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// intersectRayAABB (src/Raycast.cpp:5-30)
__device__ bool ray_aabb(f3 o, f3 d, f3 mn, f3 mx, float& t) {
    float tMin = 0.f;
    float tMax = 3.402823466e+38f;   // std::numeric_limits<float>::max()
    for (int i = 0; i < 3; ++i) {
        const float oi = comp(o, i), di = comp(d, i), lo = comp(mn, i), hi = comp(mx, i);
        if (gabs(di) < 0.001f) {
            if (oi < lo || oi > hi) return false;   // parallel to the slab, outside it
        } else {
            const float ood = 1.f / di;
            float t1 = (lo - oi) * ood;
            float t2 = (hi - oi) * ood;
            if (t1 > t2) { float x = t1; t1 = t2; t2 = x; }
            if (t1 > tMin) tMin = t1;
            if (t2 < tMax) tMax = t2;
            if (tMin > tMax) return false;
        }
    }
    t = tMin;
    return true;
}

// intersectRaySphere (src/Raycast.cpp:32-46)
__device__ bool ray_sphere(f3 o, f3 d, f3 pos, float radius, float& t) {
    f3 m = sub(o, pos);
    float b = dot(m, d);
    float c = dot(m, m) - radius * radius;
    if (c > 0.f && b > 0.f) return false;
    float discr = b * b - c;
    if (discr < 0.f) return false;
    t = -b - sqrtf(discr);
    if (t < 0.f) t = 0.f;
    return true;
}

// intersectRayCapsule (src/Raycast.cpp:48-103): top hemisphere, cylinder, bottom hemisphere
__device__ bool ray_capsule(f3 o, f3 d0, f3 pos, qf ori, float hh, float radius, float& t) {
    qf inv = qinverse(ori);
    f3 p = qrot(inv, sub(o, pos));
    f3 d = qrot(inv, d0);
    float pd = dot(p, d);
    float hdy = hh * d.y;
    float thpy = 2.f * hh * p.y;
    float phr = dot(p, p) + hh * hh - radius * radius;
    float tMin = 3.402823466e+38f;
    bool hit = false;
    float b = pd - hdy, c = phr - thpy, discr = b * b - c;
    if (discr >= 0.f) {
        t = -b - sqrtf(discr);
        if (t < 0.f) t = 0.f;
        if (add(p, mul(t, d)).y >= hh) { hit = true; tMin = t; }
    }
    float a = d.x * d.x + d.z * d.z;
    b = p.x * d.x + p.z * d.z;
    c = p.x * p.x + p.z * p.z - radius * radius;
    discr = b * b - a * c;
    if (discr >= 0.f) {
        t = (-b - sqrtf(discr)) / a;
        if (t < 0.f) t = 0.f;
        float y = add(p, mul(t, d)).y;
        if (-hh <= y && y < hh && t < tMin) { hit = true; tMin = t; }
    }
    b = pd + hdy;
    c = phr + thpy;
    discr = b * b - c;
    if (discr >= 0.f) {
        t = -b - sqrtf(discr);
        if (t < 0.f) t = 0.f;
        if (add(p, mul(t, d)).y < -hh && t < tMin) { hit = true; tMin = t; }
    }
    t = tMin;
    return hit;
}

// intersectRayBox (src/Raycast.cpp:105-110)
__device__ bool ray_box(f3 o, f3 d, f3 pos, qf ori, f3 he, float& t) {
    qf inv = qinverse(ori);
    return ray_aabb(qrot(inv, sub(o, pos)), qrot(inv, d), neg(he), he, t);
}

// intersectRayConvexMesh (src/Raycast.cpp:112-139): the faces' half-spaces, in face order
// convex meshes: per mesh {first vertex, vertex count, first face, face count}; per face
// outward normal and centroid
struct ConvexFaces { const int4* mesh; const float4 *faceN, *faceC; };

__device__ bool ray_convex(f3 o, f3 d, f3 pos, qf ori, const ConvexFaces& cv, int mesh, f3 scale, float& t) {
    const int4 mi = cv.mesh[mesh];
    float tMin = 0.f;
    float tMax = 3.402823466e+38f;
    for (int f = 0; f < mi.w; ++f) {
        f3 planeOrig = add(pos, qrot(ori, mulv(scale, xyz(cv.faceC[mi.z + f]))));
        f3 planeNormal = qrot(ori, mulv(divv(s3(1.f), scale), xyz(cv.faceN[mi.z + f])));
        float denom = dot(d, planeNormal);
        float dist = dot(sub(planeOrig, o), planeNormal);
        if (denom == 0.f) {
            if (dist > 0.f) return false;
        } else {
            t = dist / denom;
            if (denom < 0.f) { if (t > tMin) tMin = t; }
            else { if (t < tMax) tMax = t; }
            if (tMin > tMax) return false;
        }
    }
    t = tMin;
    return true;
}

struct RayScene {
    int W, nLeaves;
    const int32_t *worldRoot, *worldStop, *worldStaticStart, *staticCols, *mvals;
    const float4 *leafA, *nodeA, *nodeB;
    const float4 *cbmin, *cbmax, *cparams, *clpos, *clquat, *bpos, *bquat;
    const int32_t *cbody, *ctype;
    const uint32_t* centity;
    ConvexFaces cv;
};

// one collider: the reference's leaf test (its bounds, then its geometry, both within
// maxDistance, and the filter), kept when closer than the best so far
__device__ __forceinline__ void test_collider(const RayScene& S, int c, f3 o, f3 d, float maxDist, uint32_t ignore,
                                              float& bestT, int& best) {
    if (S.centity[c] == ignore) return;
    float t;
    if (!ray_aabb(o, d, xyz(S.cbmin[c]), xyz(S.cbmax[c]), t) || t > maxDist) return;
    const int b = S.cbody[c];
    const qf bq = q4(S.bquat[b]);
    const f3 pos = add(xyz(S.bpos[b]), qrot(bq, xyz(S.clpos[c])));
    const qf q = qmul(bq, q4(S.clquat[c]));
    const float4 prm = S.cparams[c];
    bool hit;
    switch (S.ctype[c]) {
        case PGX_SPHERE: hit = ray_sphere(o, d, pos, prm.x, t); break;
        case PGX_CAPSULE: hit = ray_capsule(o, d, pos, q, prm.x, prm.y, t); break;
        case PGX_BOX: hit = ray_box(o, d, pos, q, xyz(prm), t); break;
        case PGX_CONVEX_MESH: hit = ray_convex(o, d, pos, q, S.cv, (int)prm.w, xyz(prm), t); break;
        default: hit = false;   // triangle meshes are never hit (src/Raycast.cpp:148)
    }
    if (!hit || t > maxDist) return;
    if (best < 0 || t < bestT || (t == bestT && c < best)) { best = c; bestT = t; }
}

// ray i: origin (.w = maxDistance), direction (.w = world id bits), entity to reject
__global__ __launch_bounds__(256) void k_raycast(int n, const float4* __restrict__ ro, const float4* __restrict__ rd,
                                                 const uint32_t* __restrict__ ignore, RayScene S,
                                                 int32_t* __restrict__ hitCol, float* __restrict__ hitT) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = ro[i], b = rd[i];
    const f3 o = xyz(a), d = xyz(b);
    const float maxDist = a.w;
    const int wd = __float_as_int(b.w);
    const uint32_t ig = ignore[i];
    float bestT = 0.f;
    int best = -1;
    int node = S.worldRoot[wd];
    if (node != -1) {
        const int stop = S.worldStop[wd];
        int visits = 0;
        while (node != stop && ++visits <= 2 * S.nLeaves) {   // a walk visits at most 2n - 1 nodes
            if (node & LEAF_BIT) {
                const int l = node & ~LEAF_BIT;
                test_collider(S, S.mvals[l], o, d, maxDist, ig, bestT, best);
                node = __float_as_int(S.leafA[l].w);
            } else {
                const float4 na = S.nodeA[node], nb = S.nodeB[node];
                float tn;
                node = (ray_aabb(o, d, xyz(na), xyz(nb), tn) && tn <= maxDist) ? __float_as_int(na.w)
                                                                                  : __float_as_int(nb.w);
            }
        }
    }
    for (int st = S.worldStaticStart[wd]; st < S.worldStaticStart[wd + 1]; ++st)
        test_collider(S, S.staticCols[st], o, d, maxDist, ig, bestT, best);
    hitCol[i] = best;
    hitT[i] = best >= 0 ? bestT : 0.f;
}

}  // namespace

namespace pgx {

void qr_raycast(World& w, int n, const float* origins, const float* dirs, const float* maxDist,
                const int32_t* rayWorld, const uint32_t* ignoreEntity, int32_t* hitCollider, float* hitDistance,
                double* deviceUs) {
    hipStream_t s = w.stream;
    std::vector<float4> hr(2 * (size_t)n);
    std::vector<uint32_t> hig((size_t)n);
    for (int i = 0; i < n; ++i) {
        int32_t wd = rayWorld ? rayWorld[i] : 0;
        float wbits;
        std::memcpy(&wbits, &wd, sizeof wbits);
        hr[i] = make_float4(origins[3 * i], origins[3 * i + 1], origins[3 * i + 2], maxDist[i]);
        hr[n + i] = make_float4(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2], wbits);
        hig[i] = ignoreEntity ? ignoreEntity[i] : 0xffffffffu;
    }
    w.qRays.alloc(2 * (size_t)n); w.qIgnore.alloc(n); w.qHit.alloc(n); w.qT.alloc(n);
    PGX_HIP(hipMemcpyAsync(w.qRays.p, hr.data(), sizeof(float4) * 2 * (size_t)n, hipMemcpyHostToDevice, s));
    PGX_HIP(hipMemcpyAsync(w.qIgnore.p, hig.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    if (!w.qEv[0])
        for (auto& e : w.qEv) PGX_HIP(hipEventCreate(&e));
    PGX_HIP(hipEventRecord(w.qEv[0], s));
    fill_i32(s, w.counters.p + 7, 0, 1);   // the refit's work counter
    bp_build(w);
    RayScene S;
    S.W = w.W;
    S.nLeaves = w.nDynCol;
    S.worldRoot = w.worldRoot.p; S.worldStop = w.worldStop.p;
    S.worldStaticStart = w.worldStaticStart.p; S.staticCols = w.staticCols.p; S.mvals = w.mvals.p;
    S.leafA = w.leafA.p; S.nodeA = w.nodeA.p; S.nodeB = w.nodeB.p;
    S.cbmin = w.cbmin.p; S.cbmax = w.cbmax.p; S.cparams = w.cparams.p; S.clpos = w.clpos.p; S.clquat = w.clquat.p;
    S.bpos = w.bpos.p; S.bquat = w.bquat.p; S.cbody = w.cbody.p; S.ctype = w.ctype.p; S.centity = w.centity.p;
    S.cv = ConvexFaces{w.cvMesh.p, w.cvFaceN.p, w.cvFaceC.p};
    if (w.nDynCol == 0) fill_i32(s, w.worldRoot.p, -1, w.W);   // no tree: static lists only
    hipLaunchKernelGGL(k_raycast, dim3(div_up(n, 256)), dim3(256), 0, s, n, w.qRays.p, w.qRays.p + n, w.qIgnore.p,
                       S, w.qHit.p, w.qT.p);
    PGX_LAUNCH_CHECK();
    PGX_HIP(hipEventRecord(w.qEv[1], s));
    PGX_HIP(hipMemcpyAsync(hitCollider, w.qHit.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    PGX_HIP(hipMemcpyAsync(hitDistance, w.qT.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    PGX_HIP(hipStreamSynchronize(s));
    if (deviceUs) {
        float ms = 0.f;
        PGX_HIP(hipEventElapsedTime(&ms, w.qEv[0], w.qEv[1]));
        *deviceUs = 1000.0 * ms;
    }
}

// query shapes to the device: params, position (.w = world id bits), orientation; types
static void upload_query_shapes(World& w, int n, const int32_t* types, const float* params, const float* positions,
                                const float* orientations, const int32_t* queryWorld) {
    std::vector<float4> hs(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        int32_t wd = queryWorld ? queryWorld[i] : 0;
        float wbits;
        std::memcpy(&wbits, &wd, sizeof wbits);
        hs[i] = make_float4(params[4 * i], params[4 * i + 1], params[4 * i + 2], params[4 * i + 3]);
        hs[n + i] = make_float4(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2], wbits);
        hs[2 * (size_t)n + i] = make_float4(orientations[4 * i], orientations[4 * i + 1], orientations[4 * i + 2],
                                            orientations[4 * i + 3]);
    }
    w.qShape.alloc(3 * (size_t)n); w.qType.alloc(n);
    PGX_HIP(hipMemcpyAsync(w.qShape.p, hs.data(), sizeof(float4) * 3 * (size_t)n, hipMemcpyHostToDevice, w.stream));
    PGX_HIP(hipMemcpyAsync(w.qType.p, types, sizeof(int32_t) * n, hipMemcpyHostToDevice, w.stream));
    PGX_HIP(hipStreamSynchronize(w.stream));   // hs is released on return
}

int qr_overlap_mtd(World& w, int n, const int32_t* types, const float* params, const float* positions,
                   const float* orientations, const int32_t* queryWorld, int capacity, int32_t* offsets,
                   int32_t* colliders, float* normalMtd, int32_t* total, double* deviceUs) {
    constexpr int MBITS = 8;   // manifold index within a pair (the kernels' MTD_MBITS)
    hipStream_t s = w.stream;
    int cbits = 1;
    while ((1L << cbits) < (long)w.C) ++cbits;
    int qbits = 1;
    while ((1L << qbits) < (long)n) ++qbits;
    if (qbits + cbits + MBITS > 64) throw PgxError{PGX_ERR_UNSUPPORTED, "too many queries and colliders for one call"};
    upload_query_shapes(w, n, types, params, positions, orientations, queryWorld);
    w.qMtdN.alloc(4);
    w.qKeys.alloc(std::max(capacity, 1)); w.qVals.alloc(std::max(capacity, 1)); w.qRes.alloc(std::max(capacity, 1));
    if (!w.qEv[0])
        for (auto& e : w.qEv) PGX_HIP(hipEventCreate(&e));
    MtdArgs A;
    A.n = n; A.nLeaves = w.nDynCol; A.cap = capacity; A.cbits = cbits;
    // triangle-mesh pairs: one lane each, with TM_MAXCAND contact records of its own
    // (a buffer of the query's own, allocated on first use: 8192 lanes, 126 MB)
    A.recLanes = 1;
    A.recs = nullptr;
    if (w.nTriMesh > 0) {
        A.recLanes = std::max(1, std::min({w.meshLanes, w.epaLanes, 8192}));
        w.qRecs.alloc((size_t)(TM_CI_WORDS + 1) * TM_MAXCAND * A.recLanes);
        A.recs = w.qRecs.p;
    }
    A.qPrm = w.qShape.p; A.qPos = w.qShape.p + n; A.qQuat = w.qShape.p + 2 * (size_t)n; A.qType = w.qType.p;
    A.worldRoot = w.worldRoot.p; A.worldStop = w.worldStop.p;
    A.worldStaticStart = w.worldStaticStart.p; A.staticCols = w.staticCols.p; A.mvals = w.mvals.p;
    A.leafA = w.leafA.p; A.nodeA = w.nodeA.p; A.nodeB = w.nodeB.p;
    A.cbmin = w.cbmin.p; A.cbmax = w.cbmax.p; A.cparams = w.cparams.p; A.clpos = w.clpos.p; A.clquat = w.clquat.p;
    A.bpos = w.bpos.p; A.bquat = w.bquat.p; A.cbody = w.cbody.p; A.ctype = w.ctype.p;
    A.counts = w.qMtdN.p; A.keys = w.qKeys.p; A.vals = w.qVals.p; A.res = w.qRes.p;
    int cnt[4] = {0, 0, 0, 0};
    int listCap = (int)std::max<size_t>(std::max<size_t>(w.qCvxPairs.n, 4096), 4 * (size_t)n);
    float ms0 = 0.f;
    for (int attempt = 0; attempt < 2; ++attempt) {   // the second with pair lists as long as the first needed
        w.qCvxPairs.alloc(listCap); w.qMeshPairs.alloc(listCap);
        A.listCap = listCap; A.convexPairs = w.qCvxPairs.p; A.meshPairs = w.qMeshPairs.p;
        PGX_HIP(hipEventRecord(w.qEv[0], s));
        fill_i32(s, w.qMtdN.p, 0, 4);
        fill_i32(s, w.counters.p + 7, 0, 1);   // the refit's work counter
        bp_build(w);
        if (w.nDynCol == 0) fill_i32(s, w.worldRoot.p, -1, w.W);   // no tree: static lists only
        np_mtd_launch(w, A);
        PGX_HIP(hipEventRecord(w.qEv[1], s));
        PGX_HIP(hipMemcpyAsync(cnt, w.qMtdN.p, sizeof cnt, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipStreamSynchronize(s));
        PGX_HIP(hipEventElapsedTime(&ms0, w.qEv[0], w.qEv[1]));
        if (std::max(cnt[1], cnt[2]) <= listCap) break;
        if (attempt == 1) throw PgxError{PGX_ERR_HIP, "MTD query: pair lists overflowed twice"};
        listCap = std::max(cnt[1], cnt[2]);
    }
    if (cnt[3]) throw PgxError{PGX_ERR_CAPACITY, "MTD query: EPA or triangle-candidate workspace overflow in " +
                                                     std::to_string(cnt[3]) + " pairs"};
    *total = cnt[0];
    if (deviceUs) *deviceUs = 1000.0 * ms0;
    if (cnt[0] > capacity) return PGX_ERR_CAPACITY;
    const int m = cnt[0];
    float ms1 = 0.f;
    if (m > 1) {
        PGX_HIP(hipEventRecord(w.qEv[2], s));
        radix_sort_u64(s, w.qKeys.p, w.qVals.p, m, qbits + cbits + MBITS, w.sortT);
        PGX_HIP(hipEventRecord(w.qEv[3], s));
    }
    std::vector<uint64_t> keys((size_t)m);
    std::vector<int32_t> vals((size_t)m);
    std::vector<float4> res((size_t)m);
    if (m) {
        PGX_HIP(hipMemcpyAsync(keys.data(), w.qKeys.p, 8 * (size_t)m, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipMemcpyAsync(vals.data(), w.qVals.p, 4 * (size_t)m, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipMemcpyAsync(res.data(), w.qRes.p, 16 * (size_t)m, hipMemcpyDeviceToHost, s));
    }
    PGX_HIP(hipStreamSynchronize(s));
    if (m > 1 && deviceUs) {
        PGX_HIP(hipEventElapsedTime(&ms1, w.qEv[2], w.qEv[3]));
        *deviceUs = 1000.0 * (ms0 + ms1);
    }
    const uint64_t cmask = (1ull << cbits) - 1;
    int k = 0;
    for (int i = 0; i < n; ++i) {
        offsets[i] = k;
        while (k < m && (int)(keys[k] >> (cbits + MBITS)) == i) {
            colliders[k] = (int32_t)((keys[k] >> MBITS) & cmask);
            const float4 r = res[vals[k]];
            normalMtd[4 * k] = r.x; normalMtd[4 * k + 1] = r.y; normalMtd[4 * k + 2] = r.z; normalMtd[4 * k + 3] = r.w;
            ++k;
        }
    }
    offsets[n] = k;
    if (k != m) throw PgxError{PGX_ERR_HIP, "MTD query: result keys out of order"};
    return PGX_OK;
}

// CharacterController::move (src/CharacterController.cpp:5-45) for every character: the
// rounds of MTD queries run on the device for all characters still moving, and each
// character's hits are applied on the host in order, as the reference's loop does.
void qr_character_move(World& w, int n, pgx_character* chars, const float* translations, uint8_t* flags) {
    std::vector<f3> pos((size_t)n), up((size_t)n);
    int rounds = 0;
    for (int i = 0; i < n; ++i) {
        const pgx_character& c = chars[i];
        const qf q{c.orientation[0], c.orientation[1], c.orientation[2], c.orientation[3]};
        up[i] = qrot(q, mk3(0.f, 1.f, 0.f));
        pos[i] = add(mk3(c.position[0], c.position[1], c.position[2]),
                     mk3(translations[3 * i], translations[3 * i + 1], translations[3 * i + 2]));
        flags[i] = 0;
        rounds = std::max(rounds, (int)c.num_iterations);
    }
    std::vector<int32_t> idx, types, qworld, off, cols;
    std::vector<float> prm, qpos, qori, nm;
    int cap = 8 * n + 1024;
    for (int r = 0; r < rounds; ++r) {
        idx.clear(); types.clear(); qworld.clear(); prm.clear(); qpos.clear(); qori.clear();
        for (int i = 0; i < n; ++i) {
            const pgx_character& c = chars[i];
            if (r >= c.num_iterations) continue;
            idx.push_back(i);
            types.push_back(PGX_CAPSULE);
            qworld.push_back(c.world);
            prm.insert(prm.end(), {c.height * 0.5f, c.radius, 0.f, 0.f});
            const f3 centre = add(pos[i], mul(up[i], c.height * 0.5f + c.radius));
            qpos.insert(qpos.end(), {centre.x, centre.y, centre.z});
            qori.insert(qori.end(), {c.orientation[0], c.orientation[1], c.orientation[2], c.orientation[3]});
        }
        const int m = (int)idx.size();
        if (m == 0) continue;
        off.assign((size_t)m + 1, 0);
        int32_t total = 0;
        while (true) {
            cols.resize((size_t)cap);
            nm.resize(4 * (size_t)cap);
            if (qr_overlap_mtd(w, m, types.data(), prm.data(), qpos.data(), qori.data(), qworld.data(), cap,
                               off.data(), cols.data(), nm.data(), &total, nullptr) == PGX_OK)
                break;
            cap = total;   // PGX_ERR_CAPACITY: room for every hit
        }
        for (int j = 0; j < m; ++j) {
            const int i = idx[j];
            const float slope = chars[i].slope_limit;
            for (int k = off[j]; k < off[j + 1]; ++k) {
                const f3 nk = mk3(nm[4 * (size_t)k], nm[4 * (size_t)k + 1], nm[4 * (size_t)k + 2]);
                const float mtd = nm[4 * (size_t)k + 3];
                const float d = dot(nk, up[i]);
                if (d > 0.001f) flags[i] |= 1;          // DOWN
                else if (d < -0.001f) flags[i] |= 4;    // UP
                else flags[i] |= 2;                     // SIDES
                pos[i] = d > slope ? add(pos[i], divs(mul(up[i], mtd), d)) : add(pos[i], mul(nk, mtd));
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        chars[i].position[0] = pos[i].x; chars[i].position[1] = pos[i].y; chars[i].position[2] = pos[i].z;
    }
}

int qr_overlap(World& w, int n, const int32_t* types, const float* params, const float* positions,
               const float* orientations, const int32_t* queryWorld, const int32_t* filters, int capacity,
               int32_t* offsets, int32_t* colliders, int32_t* total, double* deviceUs) {
    hipStream_t s = w.stream;
    int cbits = 1;
    while ((1L << cbits) < (long)w.C) ++cbits;
    int qbits = 1;
    while ((1L << qbits) < (long)n) ++qbits;
    std::vector<float4> hs(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        int32_t wd = queryWorld ? queryWorld[i] : 0;
        float wbits;
        std::memcpy(&wbits, &wd, sizeof wbits);
        hs[i] = make_float4(params[4 * i], params[4 * i + 1], params[4 * i + 2], params[4 * i + 3]);
        hs[n + i] = make_float4(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2], wbits);
        hs[2 * (size_t)n + i] = make_float4(orientations[4 * i], orientations[4 * i + 1], orientations[4 * i + 2],
                                            orientations[4 * i + 3]);
    }
    std::vector<int32_t> hf(filters ? filters : nullptr, filters ? filters + n : nullptr);
    if (!filters) hf.assign((size_t)n, 0);
    w.qShape.alloc(3 * (size_t)n); w.qType.alloc(n); w.qFilter.alloc(n); w.qCount.alloc((size_t)n + 2);
    w.qKeys.alloc(std::max(capacity, 1)); w.qVals.alloc(std::max(capacity, 1));
    PGX_HIP(hipMemcpyAsync(w.qShape.p, hs.data(), sizeof(float4) * 3 * (size_t)n, hipMemcpyHostToDevice, s));
    PGX_HIP(hipMemcpyAsync(w.qType.p, types, sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    PGX_HIP(hipMemcpyAsync(w.qFilter.p, hf.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    if (!w.qEv[0])
        for (auto& e : w.qEv) PGX_HIP(hipEventCreate(&e));
    OverlapArgs A;
    A.n = n; A.nLeaves = w.nDynCol; A.cap = capacity; A.cbits = cbits;
    A.qPrm = w.qShape.p; A.qPos = w.qShape.p + n; A.qQuat = w.qShape.p + 2 * (size_t)n;
    A.qType = w.qType.p; A.qFilter = w.qFilter.p;
    A.worldRoot = w.worldRoot.p; A.worldStop = w.worldStop.p;
    A.worldStaticStart = w.worldStaticStart.p; A.staticCols = w.staticCols.p; A.mvals = w.mvals.p;
    A.leafA = w.leafA.p; A.nodeA = w.nodeA.p; A.nodeB = w.nodeB.p;
    A.cbmin = w.cbmin.p; A.cbmax = w.cbmax.p; A.cparams = w.cparams.p; A.clpos = w.clpos.p; A.clquat = w.clquat.p;
    A.bpos = w.bpos.p; A.bquat = w.bquat.p; A.cbody = w.cbody.p; A.ctype = w.ctype.p; A.cdata = w.cdata.p;
    A.cv = ConvexView{w.cvVerts.p, w.cvMesh.p, w.cvFace.p, w.cvFaceIdx.p, w.cvFaceN.p, w.cvFaceC.p};
    A.count = w.qCount.p; A.total = w.qCount.p + n; A.keys = w.qKeys.p; A.nCvx = w.qCount.p + n + 1;
    std::vector<int32_t> cnt((size_t)n + 2);
    int listCap = (int)std::max<size_t>(std::max<size_t>(w.qCvxPairs.n, 4096), (size_t)n);
    float ms0 = 0.f;
    for (int attempt = 0; attempt < 2; ++attempt) {   // the second with a convex-pair list as long as needed
        w.qCvxPairs.alloc(listCap);
        A.cvxPairs = w.qCvxPairs.p; A.listCap = listCap;
        PGX_HIP(hipEventRecord(w.qEv[0], s));
        fill_i32(s, w.counters.p + 7, 0, 1);   // the refit's work counter
        fill_i32(s, w.qCount.p + n, 0, 2);     // the hit and convex-pair counters
        bp_build(w);
        if (w.nDynCol == 0) fill_i32(s, w.worldRoot.p, -1, w.W);   // no tree: static lists only
        np_overlap_launch(s, A, w.nConvex > 0);
        PGX_HIP(hipEventRecord(w.qEv[1], s));
        PGX_HIP(hipMemcpyAsync(cnt.data(), w.qCount.p, sizeof(int32_t) * ((size_t)n + 2), hipMemcpyDeviceToHost, s));
        PGX_HIP(hipStreamSynchronize(s));
        PGX_HIP(hipEventElapsedTime(&ms0, w.qEv[0], w.qEv[1]));
        if (cnt[n + 1] <= listCap) break;
        if (attempt == 1) throw PgxError{PGX_ERR_HIP, "overlap: the convex-pair list overflowed twice"};
        listCap = cnt[n + 1];
    }
    long sum = 0;
    for (int i = 0; i < n; ++i) {
        offsets[i] = (int32_t)sum;
        sum += cnt[i];
    }
    offsets[n] = (int32_t)sum;
    *total = (int32_t)sum;
    if (sum != cnt[n]) throw PgxError{PGX_ERR_HIP, "overlap: the per-query hit counts disagree with the total"};
    if (deviceUs) *deviceUs = 1000.0 * ms0;
    if (sum > capacity) return PGX_ERR_CAPACITY;
    if (sum == 0) return PGX_OK;
    // the hits of each query ascending by collider: sort the keys (query, collider)
    PGX_HIP(hipEventRecord(w.qEv[2], s));
    radix_sort_u64(s, w.qKeys.p, w.qVals.p, (int)sum, qbits + cbits, w.sortT);
    PGX_HIP(hipEventRecord(w.qEv[3], s));
    std::vector<uint64_t> keys((size_t)sum);
    PGX_HIP(hipMemcpyAsync(keys.data(), w.qKeys.p, sizeof(uint64_t) * sum, hipMemcpyDeviceToHost, s));
    PGX_HIP(hipStreamSynchronize(s));
    const uint64_t cmask = (1ull << cbits) - 1;
    for (long k = 0; k < sum; ++k) colliders[k] = (int32_t)(keys[k] & cmask);
    if (deviceUs) {
        float ms1 = 0.f;
        PGX_HIP(hipEventElapsedTime(&ms1, w.qEv[2], w.qEv[3]));
        *deviceUs = 1000.0 * (ms0 + ms1);
    }
    return PGX_OK;
}

}  // namespace pgx
