// Serving: continuous batching of independent scenes in one device world.
//
// Every request is one world of a batched device world (worlds never interact; see the
// world id in the broadphase keys). At each step boundary the server retires requests
// that reached their horizon (their final state is copied out) and admits queued ones
// while the footprint budget allows. When the set changed, the next batched world's
// arrays are assembled on the device: each continuing request's segment of every array is
// gathered from the current world (body and collider indices renumbered), each admitted
// request's segments come from one packed upload of its converted descriptor, and the
// buffers swap. The persistent state of continuing requests -- body poses and velocities,
// the restitution contact cache, the bounds margin of their dynamic colliders, gear joint
// angles -- stays on the device. Worlds keep their admission order, so a continuing
// request's contact cache stays sorted after its world id is renumbered. The assembled
// world equals the one a descriptor concatenating the active requests would initialise.
//
// Entity ids of a request are shifted by a per-request base so that entity-keyed state
// (non-colliding pairs) never aliases across requests; the shift keeps their order, the
// only property the step depends on (pair orientation by raw entity value).
#include <cstring>
#include <deque>
#include <unordered_map>
#include <vector>
#include "pgx_internal.h"

using namespace pgx;

namespace {

struct HostScene {
    int nA = 0, nD = 0, nC = 0;
    std::vector<uint32_t> entity;
    std::vector<float> pos, quat;
    std::vector<int32_t> dynOfActor, actorOfDyn;
    std::vector<uint8_t> kinematic;
    std::vector<float> vel, angvel, invMass, com, invI;
    std::vector<int32_t> colActor, colIndex, colType, colData;
    std::vector<float> colPos, colQuat, colParams, colMat;
    std::vector<uint8_t> colTrigger, colEnable;
    std::vector<pgx_joint_desc> joints;
    uint32_t maxEntity = 0;
};

template <class T>
void take(std::vector<T>& v, const T* p, size_t n) { v.assign(p, p + n); }

struct Request {
    int id = 0;
    HostScene sc;
    uint32_t entityBase = 0;
    int horizon = 0, stepsDone = 0;
    double footprint = 0;
    // 0 queued, 1 active, 2 done (may still ride in the batch), 3 done and its result taken
    // while it still rides in the batch (the next boundary retires and erases it)
    int state = 0;
    bool inBatch = false;               // admitted and not yet retired at a boundary
    bool placed = false;                // has a placement in the current batched world
    // placement in the current batched world: world id, first dynamic body, first other
    // body (counted after all dynamic bodies), first collider, first entry of the dynamic
    // and of the other collider lists, first joint
    int world = 0, dOff = 0, sOff = 0, cOff = 0, dcOff = 0, scOff = 0, jointStart = 0;
    // device-layout facts kept from the conversion at admission: collider list lengths,
    // entity id and kinematic flag per internal body, joints with internal body indices
    int nDC = 0, nSC = 0;
    std::vector<uint32_t> bodyEntity;
    std::vector<uint8_t> bodyKinematic;
    std::vector<pgx_joint_desc> bodyJoints;
    std::vector<float4> rpos, rquat, rvel, rang;   // final dynamic state (done)
};

// estimated HBM bytes of one request: per collider the pair and manifold slots its
// share of the capacities (12 pair slots per collider) carries, plus per-body state
double footprint_of(int nA, int nD, int nC) {
    const double perPairSlot = 1136.0;   // narrowphase slot + manifold slot + contact cache
    return 12.0 * nC * perPairSlot + 400.0 * nA + 320.0 * nD + 200.0 * nC;
}

// One segment of an assembled array: `count` elements from the current world (src 0) or
// the staging upload (src 1) at srcStart to dstStart; flat = elements of the segments
// before it. Body indices are renumbered from the source layout (a = dynamic bodies in
// it, b / c = the segment owner's first dynamic / other body there) to the new one (d / e);
// collider indices move by `delta`; `world` is the owner's new world id.
struct Seg {
    int src, srcStart, dstStart, count, flat;
    int a, b, c, d, e, delta, world;
};
enum { FIX_NONE = 0, FIX_BODY = 1, FIX_COLLIDER = 2, FIX_WORLD = 3 };

template <class T, int FIX>
__global__ void k_gather(int nseg, const Seg* __restrict__ seg, int total, const T* __restrict__ cur,
                         const T* __restrict__ stage, T* __restrict__ dst, int dstDyn) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int lo = 0, hi = nseg - 1;   // the last segment starting at or before i (none is empty)
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (seg[mid].flat <= i) lo = mid; else hi = mid - 1;
    }
    const Seg g = seg[lo];
    const int k = i - g.flat;
    T v;
    if constexpr (FIX == FIX_WORLD) {
        v = (T)g.world;
    } else {
        v = (g.src ? stage : cur)[g.srcStart + k];
        if constexpr (FIX == FIX_BODY) {
            int x = (int)v;
            v = (T)(x < g.a ? x - g.b + g.d : x - g.a - g.c + dstDyn + g.e);
        } else if constexpr (FIX == FIX_COLLIDER) {
            v = (T)((int)v + g.delta);
        }
    }
    dst[g.dstStart + k] = v;
}

__global__ void k_iota(int n, int32_t* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = i;
}

__global__ void k_cache_flag(int n, const uint64_t* __restrict__ key, int oldLocalBits,
                             const int32_t* __restrict__ worldMap, int32_t* __restrict__ flag) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    flag[e] = worldMap[(int)(key[e] >> (2 * oldLocalBits))] >= 0 ? 1 : 0;
}

__global__ void k_cache_remap(int n, const uint64_t* __restrict__ key, const int32_t* __restrict__ tri,
                              const int32_t* __restrict__ num, const float4* __restrict__ lp, size_t lpStride,
                              int oldLocalBits,
                              int newLocalBits, const int32_t* __restrict__ worldMap,
                              const int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
                              uint64_t* __restrict__ nKey, int32_t* __restrict__ nTri, int32_t* __restrict__ nNum,
                              float4* __restrict__ nLp, int newCap) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n || !flag[e]) return;
    uint64_t k = key[e];
    uint64_t om = (1ull << oldLocalBits) - 1;
    int nw = worldMap[(int)(k >> (2 * oldLocalBits))];
    uint64_t l0 = (k >> oldLocalBits) & om, l1 = k & om;
    int q = pos[e];
    nKey[q] = ((uint64_t)nw << (2 * newLocalBits)) | (l0 << newLocalBits) | l1;
    nTri[q] = tri[e];
    nNum[q] = num[e];
    for (int j = 0; j < 4; ++j) nLp[(size_t)j * newCap + q] = lp[(size_t)j * lpStride + e];
}

}  // namespace

struct pgx_server {
    World w;
    pgx_config cfg{};
    double budget = 64.0 * (1ull << 30);
    int maxWorlds = 8192;
    int admitInterval = 1;
    int sinceBoundary = 0;
    int nextId = 1;
    uint32_t nextEntity = 1;
    std::deque<int> queue;
    std::vector<int> active;            // admission order = world order
    std::unordered_map<int, Request> reqs;
    double footprint = 0;
    bool initialised = false;
    int worlds = 0;                     // worlds of the current batched world
    // assembly: spare arrays the next world's arrays are gathered into (then swapped with the
    // world's), one packed upload (segment tables, admitted requests' arrays, world starts)
    // through a pinned host buffer, the previous contact cache and joint state
    DBuf<float4> xBpos, xBquat, xDvel, xDang, xDcom, xDinvI, xCparams, xClpos, xClquat, xCmat;
    DBuf<uint32_t> xBent, xCent;
    DBuf<int32_t> xCbody, xCworld, xClocal, xCflags, xCtype, xCindex, xDynCols, xStaticCols;
    DBuf<uint8_t> stage;
    uint8_t* hstage = nullptr;
    size_t hstageCap = 0;
    DBuf<float4> oLp, oGear;
    DBuf<uint64_t> oKey;
    DBuf<int32_t> oTri, oNum, oGearInit, worldMap, flag, fpos;
    DBuf<uint8_t> worldMask;
    ~pgx_server() { if (hstage) (void)hipHostFree(hstage); }
};

static int srv_fail(int code, const std::string& msg) { return report_error(code, msg); }

namespace {

// the descriptor a request was submitted with (its own copy)
pgx_scene_desc desc_of(const HostScene& h) {
    static const int32_t zero[2] = {0, 0};
    pgx_scene_desc d{};
    d.n_actors = h.nA; d.n_dyn = h.nD; d.n_colliders = h.nC;
    d.entity = h.entity.data(); d.pos = h.pos.data(); d.quat = h.quat.data(); d.dyn_of_actor = h.dynOfActor.data();
    d.actor_of_dyn = h.actorOfDyn.data(); d.kinematic = h.kinematic.data(); d.vel = h.vel.data();
    d.angvel = h.angvel.data(); d.inv_mass = h.invMass.data(); d.com = h.com.data(); d.inv_inertia = h.invI.data();
    d.col_actor = h.colActor.data(); d.col_index = h.colIndex.data(); d.col_pos = h.colPos.data();
    d.col_quat = h.colQuat.data(); d.col_type = h.colType.data(); d.col_params = h.colParams.data();
    d.col_material = h.colMat.data(); d.col_trigger = h.colTrigger.data(); d.col_enable = h.colEnable.data();
    d.col_data = h.colData.data();
    d.n_worlds = 1;
    d.convex_vert_start = zero; d.convex_face_start = zero; d.tri_vert_start = zero; d.tri_start = zero;
    return d;
}

// final dynamic state of a request leaving the batch
void read_result(pgx_server& S, Request& r) {
    World& w = S.w;
    int n = r.sc.nD;
    r.rpos.resize(n); r.rquat.resize(n); r.rvel.resize(n); r.rang.resize(n);
    if (n) {
        PGX_HIP(hipMemcpyAsync(r.rpos.data(), w.bpos.p + r.dOff, sizeof(float4) * n, hipMemcpyDeviceToHost, w.stream));
        PGX_HIP(hipMemcpyAsync(r.rquat.data(), w.bquat.p + r.dOff, sizeof(float4) * n, hipMemcpyDeviceToHost, w.stream));
        PGX_HIP(hipMemcpyAsync(r.rvel.data(), w.dvel.p + r.dOff, sizeof(float4) * n, hipMemcpyDeviceToHost, w.stream));
        PGX_HIP(hipMemcpyAsync(r.rang.data(), w.dang.p + r.dOff, sizeof(float4) * n, hipMemcpyDeviceToHost, w.stream));
    }
}

// Byte regions of the packed upload, 256-byte aligned.
struct Packing {
    size_t size = 0;
    template <class T>
    size_t region(size_t n) {
        size_t off = size;
        size = (off + n * sizeof(T) + 255) & ~(size_t)255;
        return off;
    }
};

template <class T>
void put(uint8_t* base, size_t region, size_t at, const std::vector<T>& v) {
    if (!v.empty()) std::memcpy(base + region + at * sizeof(T), v.data(), v.size() * sizeof(T));
}

// Gathers one array of the next world into `spare` (from `live`, the current world's, and
// the upload) and swaps the two.
template <class T, int FIX>
void gather(hipStream_t s, int nseg, const Seg* segs, int total, DBuf<T>& live, const uint8_t* stage, size_t region,
            DBuf<T>& spare, int dstDyn) {
    spare.alloc(total);
    if (total > 0 && nseg > 0) {
        hipLaunchKernelGGL((k_gather<T, FIX>), dim3(div_up(total, 256)), dim3(256), 0, s, nseg, segs, total,
                           (const T*)live.p, (const T*)(stage + region), spare.p, dstDyn);
        PGX_LAUNCH_CHECK();
    }
    live.swap(spare);
}

// Makes the batched world hold the active requests (S.active, in world order). Requests
// placed in the current world keep their state; the others are converted and uploaded.
void assemble(pgx_server& S) {
    World& w = S.w;
    hipStream_t s = w.stream;
    const int oldD = w.D, oldW = S.worlds, oldLocalBits = w.localBits;
    const int W = (int)S.active.size();
    // 1. conversion of the admitted requests (request-local layout, entity ids shifted)
    std::vector<SceneArrays> lay(W);
    for (int i = 0; i < W; ++i) {
        Request& r = S.reqs[S.active[i]];
        if (r.placed) continue;
        pgx_scene_desc d = desc_of(r.sc);
        convert_scene(&d, r.entityBase, lay[i]);
        const SceneArrays& L = lay[i];
        r.nDC = (int)L.dynCols.size();
        r.nSC = (int)L.staticCols.size();
        r.bodyEntity = L.bent;
        r.bodyKinematic = L.kinematic;
        r.bodyJoints = r.sc.joints;
        for (pgx_joint_desc& j : r.bodyJoints) {
            j.actor0 = L.actorToBody[j.actor0];
            j.actor1 = L.actorToBody[j.actor1];
        }
    }
    // 2. placements in the next world and the segment tables; the upload holds the admitted
    // requests one after the other (per request: dynamic bodies, then the others)
    int Dt = 0, St = 0, Ct = 0, DCt = 0, SCt = 0, Jt = 0, maxC = 1;
    for (int id : S.active) {
        const Request& r = S.reqs[id];
        Dt += r.sc.nD; St += r.sc.nA - r.sc.nD; Ct += r.sc.nC; DCt += r.nDC; SCt += r.nSC;
        Jt += (int)r.bodyJoints.size();
        maxC = std::max(maxC, r.sc.nC);
    }
    // checks the key widths before anything changes
    world_set_shape(w, &S.cfg, Dt + St, Dt, Ct, W, DCt, maxC);
    std::vector<Seg> gBody, gDyn, gDyn3, gCol, gDC, gSC, gJoint;
    std::vector<int32_t> colStart(W + 1, 0), staticStart(W + 1, 0), worldMapH(std::max(oldW, 1), -1);
    std::vector<uint8_t> mask(std::max(W, 1), 0);
    std::vector<int> upA(W), upD(W), upC(W), upDC(W), upSC(W);
    int uA = 0, uD = 0, uC = 0, uDC = 0, uSC = 0;
    int nd = 0, ns = 0, nc = 0, ndc = 0, nsc = 0, nj = 0;
    auto add = [](std::vector<Seg>& v, Seg g) {
        if (g.count <= 0) return;
        g.flat = v.empty() ? 0 : v.back().flat + v.back().count;
        v.push_back(g);
    };
    bool anyContinuing = false;
    for (int i = 0; i < W; ++i) {
        Request& r = S.reqs[S.active[i]];
        const int nD = r.sc.nD, nS = r.sc.nA - r.sc.nD, nC = r.sc.nC;
        Seg b{};
        b.world = i;
        b.d = nd; b.e = ns;
        if (r.placed) {
            anyContinuing = true;
            mask[i] = 1;
            worldMapH[r.world] = i;
            b.src = 0;
            b.a = oldD; b.b = r.dOff; b.c = r.sOff;
            b.delta = nc - r.cOff;
            Seg t = b;
            t.srcStart = r.dOff; t.dstStart = nd; t.count = nD; add(gBody, t);
            t.srcStart = oldD + r.sOff; t.dstStart = Dt + ns; t.count = nS; add(gBody, t);
            t.srcStart = r.dOff; t.dstStart = nd; t.count = nD; add(gDyn, t);
            t.srcStart = 3 * r.dOff; t.dstStart = 3 * nd; t.count = 3 * nD; add(gDyn3, t);
            t.srcStart = r.cOff; t.dstStart = nc; t.count = nC; add(gCol, t);
            t.srcStart = r.dcOff; t.dstStart = ndc; t.count = r.nDC; add(gDC, t);
            t.srcStart = r.scOff; t.dstStart = nsc; t.count = r.nSC; add(gSC, t);
            t.srcStart = r.jointStart; t.dstStart = nj; t.count = (int)r.bodyJoints.size(); add(gJoint, t);
        } else {
            upA[i] = uA; upD[i] = uD; upC[i] = uC; upDC[i] = uDC; upSC[i] = uSC;
            b.src = 1;
            b.a = nD; b.b = 0; b.c = 0;   // uploaded collider bodies are request-local
            b.delta = nc;
            Seg t = b;
            t.srcStart = uA; t.dstStart = nd; t.count = nD; add(gBody, t);
            t.srcStart = uA + nD; t.dstStart = Dt + ns; t.count = nS; add(gBody, t);
            t.srcStart = uD; t.dstStart = nd; t.count = nD; add(gDyn, t);
            t.srcStart = 3 * uD; t.dstStart = 3 * nd; t.count = 3 * nD; add(gDyn3, t);
            t.srcStart = uC; t.dstStart = nc; t.count = nC; add(gCol, t);
            t.srcStart = uDC; t.dstStart = ndc; t.count = r.nDC; add(gDC, t);
            t.srcStart = uSC; t.dstStart = nsc; t.count = r.nSC; add(gSC, t);
            uA += r.sc.nA; uD += nD; uC += nC; uDC += r.nDC; uSC += r.nSC;
        }
        r.world = i;
        r.dOff = nd; r.sOff = ns; r.cOff = nc; r.dcOff = ndc; r.scOff = nsc; r.jointStart = nj;
        nd += nD; ns += nS; nc += nC; ndc += r.nDC; nsc += r.nSC; nj += (int)r.bodyJoints.size();
        colStart[i + 1] = nc;
        staticStart[i + 1] = nsc;
    }
    // 3. the packed upload
    Packing P;
    const size_t rSeg[7] = {P.region<Seg>(gBody.size()), P.region<Seg>(gDyn.size()), P.region<Seg>(gDyn3.size()),
                            P.region<Seg>(gCol.size()), P.region<Seg>(gDC.size()), P.region<Seg>(gSC.size()),
                            P.region<Seg>(gJoint.size())};
    const size_t rColStart = P.region<int32_t>(W + 1), rStaticStart = P.region<int32_t>(W + 1);
    const size_t rWorldMap = P.region<int32_t>(worldMapH.size()), rMask = P.region<uint8_t>(mask.size());
    const size_t rBpos = P.region<float4>(uA), rBquat = P.region<float4>(uA), rBent = P.region<uint32_t>(uA);
    const size_t rDvel = P.region<float4>(uD), rDang = P.region<float4>(uD), rDcom = P.region<float4>(uD);
    const size_t rDinvI = P.region<float4>(3 * (size_t)uD);
    const size_t rCbody = P.region<int32_t>(uC), rClocal = P.region<int32_t>(uC), rCflags = P.region<int32_t>(uC);
    const size_t rCtype = P.region<int32_t>(uC), rCindex = P.region<int32_t>(uC), rCent = P.region<uint32_t>(uC);
    const size_t rCparams = P.region<float4>(uC), rClpos = P.region<float4>(uC), rClquat = P.region<float4>(uC);
    const size_t rCmat = P.region<float4>(uC);
    const size_t rDynCols = P.region<int32_t>(uDC), rStaticCols = P.region<int32_t>(uSC);
    if (P.size > S.hstageCap) {
        if (S.hstage) PGX_HIP(hipHostFree(S.hstage));
        S.hstage = nullptr;
        S.hstageCap = 0;
        PGX_HIP(hipHostMalloc((void**)&S.hstage, P.size, hipHostMallocDefault));
        S.hstageCap = P.size;
    }
    uint8_t* h = S.hstage;
    put(h, rSeg[0], 0, gBody); put(h, rSeg[1], 0, gDyn); put(h, rSeg[2], 0, gDyn3); put(h, rSeg[3], 0, gCol);
    put(h, rSeg[4], 0, gDC); put(h, rSeg[5], 0, gSC); put(h, rSeg[6], 0, gJoint);
    put(h, rColStart, 0, colStart); put(h, rStaticStart, 0, staticStart);
    put(h, rWorldMap, 0, worldMapH); put(h, rMask, 0, mask);
    for (int i = 0; i < W; ++i) {
        const Request& r = S.reqs[S.active[i]];
        if (r.placed) continue;
        const SceneArrays& L = lay[i];
        put(h, rBpos, upA[i], L.bpos); put(h, rBquat, upA[i], L.bquat); put(h, rBent, upA[i], L.bent);
        put(h, rDvel, upD[i], L.dvel); put(h, rDang, upD[i], L.dang); put(h, rDcom, upD[i], L.dcom);
        put(h, rDinvI, 3 * (size_t)upD[i], L.dinvI);
        put(h, rCbody, upC[i], L.cbody); put(h, rClocal, upC[i], L.clocal); put(h, rCflags, upC[i], L.cflags);
        put(h, rCtype, upC[i], L.ctype); put(h, rCindex, upC[i], L.cindex); put(h, rCent, upC[i], L.cent);
        put(h, rCparams, upC[i], L.cparams); put(h, rClpos, upC[i], L.clpos); put(h, rClquat, upC[i], L.clquat);
        put(h, rCmat, upC[i], L.cmat);
        put(h, rDynCols, upDC[i], L.dynCols); put(h, rStaticCols, upSC[i], L.staticCols);
    }
    S.stage.alloc(P.size);
    PGX_HIP(hipMemcpyAsync(S.stage.p, h, P.size, hipMemcpyHostToDevice, s));
    const uint8_t* st = S.stage.p;
    auto segs = [&](int k) { return (const Seg*)(st + rSeg[k]); };
    // 4. the state to carry that the world's re-initialisation replaces: contact cache and
    // joint state move to the server's buffers
    const int nCache = anyContinuing ? w.cacheCount[w.cacheCur] : 0;
    const size_t lpStride = w.manifoldCap;
    if (nCache) {
        const int cur = w.cacheCur;
        S.oKey.swap(w.cacheKey[cur]); S.oTri.swap(w.cacheTri[cur]); S.oNum.swap(w.cacheNum[cur]);
        S.oLp.swap(w.cacheLp[cur]);
    }
    if (!gJoint.empty()) { S.oGear.swap(w.jGear); S.oGearInit.swap(w.jGearInit); }
    // 5. the next world's arrays
    gather<float4, FIX_NONE>(s, (int)gBody.size(), segs(0), Dt + St, w.bpos, st, rBpos, S.xBpos, Dt);
    gather<float4, FIX_NONE>(s, (int)gBody.size(), segs(0), Dt + St, w.bquat, st, rBquat, S.xBquat, Dt);
    gather<uint32_t, FIX_NONE>(s, (int)gBody.size(), segs(0), Dt + St, w.bentity, st, rBent, S.xBent, Dt);
    gather<float4, FIX_NONE>(s, (int)gDyn.size(), segs(1), Dt, w.dvel, st, rDvel, S.xDvel, Dt);
    gather<float4, FIX_NONE>(s, (int)gDyn.size(), segs(1), Dt, w.dang, st, rDang, S.xDang, Dt);
    gather<float4, FIX_NONE>(s, (int)gDyn.size(), segs(1), Dt, w.dcom, st, rDcom, S.xDcom, Dt);
    gather<float4, FIX_NONE>(s, (int)gDyn3.size(), segs(2), 3 * Dt, w.dinvI, st, rDinvI, S.xDinvI, Dt);
    gather<int32_t, FIX_BODY>(s, (int)gCol.size(), segs(3), Ct, w.cbody, st, rCbody, S.xCbody, Dt);
    gather<int32_t, FIX_WORLD>(s, (int)gCol.size(), segs(3), Ct, w.cworld, st, 0, S.xCworld, Dt);
    gather<int32_t, FIX_NONE>(s, (int)gCol.size(), segs(3), Ct, w.clocal, st, rClocal, S.xClocal, Dt);
    gather<int32_t, FIX_NONE>(s, (int)gCol.size(), segs(3), Ct, w.cflags, st, rCflags, S.xCflags, Dt);
    gather<int32_t, FIX_NONE>(s, (int)gCol.size(), segs(3), Ct, w.ctype, st, rCtype, S.xCtype, Dt);
    gather<int32_t, FIX_NONE>(s, (int)gCol.size(), segs(3), Ct, w.cindex, st, rCindex, S.xCindex, Dt);
    gather<uint32_t, FIX_NONE>(s, (int)gCol.size(), segs(3), Ct, w.centity, st, rCent, S.xCent, Dt);
    gather<float4, FIX_NONE>(s, (int)gCol.size(), segs(3), Ct, w.cparams, st, rCparams, S.xCparams, Dt);
    gather<float4, FIX_NONE>(s, (int)gCol.size(), segs(3), Ct, w.clpos, st, rClpos, S.xClpos, Dt);
    gather<float4, FIX_NONE>(s, (int)gCol.size(), segs(3), Ct, w.clquat, st, rClquat, S.xClquat, Dt);
    gather<float4, FIX_NONE>(s, (int)gCol.size(), segs(3), Ct, w.cmat, st, rCmat, S.xCmat, Dt);
    gather<int32_t, FIX_COLLIDER>(s, (int)gDC.size(), segs(4), DCt, w.dynCols, st, rDynCols, S.xDynCols, Dt);
    gather<int32_t, FIX_COLLIDER>(s, (int)gSC.size(), segs(5), SCt, w.staticCols, st, rStaticCols, S.xStaticCols, Dt);
    w.worldColStart.alloc(W + 1); w.worldStaticStart.alloc(W + 1); w.colOfLocal.alloc(Ct);
    PGX_HIP(hipMemcpyAsync(w.worldColStart.p, st + rColStart, sizeof(int32_t) * (W + 1), hipMemcpyDeviceToDevice, s));
    PGX_HIP(hipMemcpyAsync(w.worldStaticStart.p, st + rStaticStart, sizeof(int32_t) * (W + 1),
                           hipMemcpyDeviceToDevice, s));
    if (Ct) {
        hipLaunchKernelGGL(k_iota, dim3(div_up(Ct, 256)), dim3(256), 0, s, Ct, w.colOfLocal.p);   // colliders by world
        PGX_LAUNCH_CHECK();
    }
    // 6. host state and the buffers that follow from the shape
    const int rounds = w.lastRounds;
    world_reset_host(w);
    w.lastRounds = rounds;   // sizes the next colouring's first batch of rounds only
    w.actorToBody.resize(Dt + St);
    for (int b = 0; b < Dt + St; ++b) w.actorToBody[b] = b;   // joints name internal bodies
    w.bodyToActor.clear();
    w.hostEntity.resize(Dt + St);
    w.hostKinematic.resize(Dt);
    for (int i = 0; i < W; ++i) {
        const Request& r = S.reqs[S.active[i]];
        const int nD = r.sc.nD;
        std::copy(r.bodyEntity.begin(), r.bodyEntity.begin() + nD, w.hostEntity.begin() + r.dOff);
        std::copy(r.bodyEntity.begin() + nD, r.bodyEntity.end(), w.hostEntity.begin() + Dt + r.sOff);
        std::copy(r.bodyKinematic.begin(), r.bodyKinematic.end(), w.hostKinematic.begin() + r.dOff);
    }
    {
        std::vector<int32_t> starts(W + 1, Dt);   // each request's dynamic bodies are one range
        for (int i = 0; i < W; ++i) starts[i] = S.reqs[S.active[i]].dOff;
        world_set_dyn_starts(w, starts, true);
    }
    w.hostColBody.clear();   // per-contact reports are not served
    w.hostColIndex.clear();
    world_alloc_derived(w, &S.cfg, 0);
    // creation-time bounds (no margin); a continuing request's dynamic colliders carry the
    // margin of its last step
    bp_compute_bounds(w, false);
    if (anyContinuing) bp_compute_bounds(w, true, st + rMask);
    // 7. joints (internal body indices) and the gear angles of continuing requests
    if (Jt) {
        std::vector<pgx_joint_desc> js;
        js.reserve(Jt);
        for (int i = 0; i < W; ++i) {
            const Request& r = S.reqs[S.active[i]];
            for (pgx_joint_desc j : r.bodyJoints) {
                for (int* a : {&j.actor0, &j.actor1})
                    *a = *a < r.sc.nD ? r.dOff + *a : Dt + r.sOff + (*a - r.sc.nD);
                js.push_back(j);
            }
        }
        world_add_joints(w, Jt, js.data());   // joint state starts at zero
        if (!gJoint.empty()) {
            const int n = gJoint.back().flat + gJoint.back().count;
            hipLaunchKernelGGL((k_gather<float4, FIX_NONE>), dim3(div_up(n, 256)), dim3(256), 0, s,
                               (int)gJoint.size(), segs(6), n, (const float4*)S.oGear.p, (const float4*)nullptr,
                               w.jGear.p, 0);
            hipLaunchKernelGGL((k_gather<int32_t, FIX_NONE>), dim3(div_up(n, 256)), dim3(256), 0, s,
                               (int)gJoint.size(), segs(6), n, (const int32_t*)S.oGearInit.p,
                               (const int32_t*)nullptr, w.jGearInit.p, 0);
            PGX_LAUNCH_CHECK();
        }
    }
    // 8. the contact cache of continuing requests, world ids renumbered
    if (nCache) {
        S.flag.alloc(nCache); S.fpos.alloc(nCache);
        const int32_t* wmap = (const int32_t*)(st + rWorldMap);
        hipLaunchKernelGGL(k_cache_flag, dim3(div_up(nCache, 256)), dim3(256), 0, s, nCache, S.oKey.p, oldLocalBits,
                           wmap, S.flag.p);
        exclusive_scan_i32(s, S.flag.p, S.fpos.p, nCache, w.counters.p + 15, w.scanT);
        hipLaunchKernelGGL(k_cache_remap, dim3(div_up(nCache, 256)), dim3(256), 0, s, nCache, S.oKey.p, S.oTri.p,
                           S.oNum.p, S.oLp.p, lpStride, oldLocalBits, w.localBits, wmap, S.flag.p, S.fpos.p,
                           w.cacheKey[0].p, w.cacheTri[0].p, w.cacheNum[0].p, w.cacheLp[0].p, w.manifoldCap);
        PGX_LAUNCH_CHECK();
        int kept = 0;
        PGX_HIP(hipMemcpyAsync(&kept, w.counters.p + 15, sizeof(int), hipMemcpyDeviceToHost, s));
        PGX_HIP(hipStreamSynchronize(s));
        w.cacheCur = 0;
        w.cacheCount[0] = kept;
    }
    PGX_HIP(hipStreamSynchronize(s));   // the pinned upload buffer is reused by the next assembly
    for (int id : S.active) S.reqs[id].placed = true;
    S.worlds = W;
}

template <class F>
int srv_guarded(F&& f) {
    try {
        return f();
    } catch (const HipError& e) {
        return srv_fail(PGX_ERR_HIP, std::string("HIP error ") + std::to_string((int)e.err) + " (" +
                                         hipGetErrorString(e.err) + ") at " + e.file + ":" + std::to_string(e.line) +
                                         ": " + e.expr);
    } catch (const PgxError& e) {
        return srv_fail(e.code, e.msg);
    } catch (const std::bad_alloc&) {
        return srv_fail(PGX_ERR_INVALID, "host out of memory");
    } catch (...) {
        return srv_fail(PGX_ERR_INVALID, "unknown error");
    }
}

}  // namespace

extern "C" {

int pgx_server_footprint(const pgx_scene_desc* sc, double* bytes) {
    if (!sc || !bytes) return srv_fail(PGX_ERR_INVALID, "NULL argument");
    *bytes = footprint_of(sc->n_actors, sc->n_dyn, sc->n_colliders);
    return PGX_OK;
}

int pgx_server_create(const pgx_config* cfg, const pgx_server_config* scfg, pgx_server** out) {
    if (!out) return srv_fail(PGX_ERR_INVALID, "out is NULL");
    *out = nullptr;
    return srv_guarded([&] {
        auto* S = new pgx_server();
        try {
            if (cfg) S->cfg = *cfg;
            else { S->cfg.num_substeps = 4; S->cfg.num_iterations = 2; S->cfg.gravity = 9.81f; }
            if (scfg && scfg->hbm_budget_bytes > 0) S->budget = scfg->hbm_budget_bytes;
            if (scfg && scfg->max_worlds > 0) S->maxWorlds = scfg->max_worlds;
            if (scfg && scfg->admit_interval > 0) S->admitInterval = scfg->admit_interval;
            S->w.device = S->cfg.device;
            PGX_HIP(hipSetDevice(S->w.device));
            PGX_HIP(hipStreamCreateWithFlags(&S->w.stream, hipStreamNonBlocking));
        } catch (...) {
            delete S;
            throw;
        }
        *out = S;
        return PGX_OK;
    });
}

int pgx_server_destroy(pgx_server* S) {
    if (!S) return PGX_OK;
    if (S->w.stream) {
        (void)hipSetDevice(S->w.device);
        (void)hipStreamSynchronize(S->w.stream);
    }
    hipStream_t st = S->w.stream;
    world_release_graph(S->w);
    delete S;   // device buffers are released by their destructors
    if (st) (void)hipStreamDestroy(st);
    return PGX_OK;
}

int pgx_server_submit(pgx_server* S, const pgx_scene_desc* sc, const pgx_joint_desc* joints, int n_joints,
                      int horizon_steps, int* request_id) {
    if (!S || !sc || !request_id || (n_joints > 0 && !joints)) return srv_fail(PGX_ERR_INVALID, "NULL argument");
    if (horizon_steps <= 0) return srv_fail(PGX_ERR_INVALID, "horizon must be positive");
    if (sc->n_convex > 0 || sc->n_trimesh > 0)
        return srv_fail(PGX_ERR_UNSUPPORTED, "scenes with convex or triangle meshes are not served");
    for (int c = 0; c < sc->n_colliders; ++c)
        if (sc->col_type[c] != PGX_SPHERE && sc->col_type[c] != PGX_CAPSULE && sc->col_type[c] != PGX_BOX)
            return srv_fail(PGX_ERR_UNSUPPORTED, "collider geometry not served");
    if (sc->actor_world) return srv_fail(PGX_ERR_INVALID, "a request is one world (actor_world must be NULL)");
    for (int d = 0; d < sc->n_dyn; ++d) {
        int a = sc->actor_of_dyn[d];
        if (a < 0 || a >= sc->n_actors || sc->dyn_of_actor[a] != d)
            return srv_fail(PGX_ERR_INVALID, "dynamic entry " + std::to_string(d) + ": actor mismatch");
    }
    for (int c = 0; c < sc->n_colliders; ++c)
        if (sc->col_actor[c] < 0 || sc->col_actor[c] >= sc->n_actors)
            return srv_fail(PGX_ERR_INVALID, "collider actor out of range");
    for (int j = 0; j < n_joints; ++j)
        if (joints[j].actor0 < 0 || joints[j].actor0 >= sc->n_actors || joints[j].actor1 < 0 ||
            joints[j].actor1 >= sc->n_actors || joints[j].type < PGX_JOINT_SPHERICAL || joints[j].type > PGX_JOINT_SERVO)
            return srv_fail(PGX_ERR_INVALID, "joint " + std::to_string(j) + " invalid");
    return srv_guarded([&] {
        Request r;
        r.id = S->nextId++;
        HostScene& h = r.sc;
        h.nA = sc->n_actors; h.nD = sc->n_dyn; h.nC = sc->n_colliders;
        take(h.entity, sc->entity, h.nA); take(h.pos, sc->pos, 3 * (size_t)h.nA); take(h.quat, sc->quat, 4 * (size_t)h.nA);
        take(h.dynOfActor, sc->dyn_of_actor, h.nA); take(h.actorOfDyn, sc->actor_of_dyn, h.nD);
        take(h.kinematic, sc->kinematic, h.nD); take(h.vel, sc->vel, 3 * (size_t)h.nD);
        take(h.angvel, sc->angvel, 3 * (size_t)h.nD); take(h.invMass, sc->inv_mass, h.nD);
        take(h.com, sc->com, 3 * (size_t)h.nD); take(h.invI, sc->inv_inertia, 9 * (size_t)h.nD);
        take(h.colActor, sc->col_actor, h.nC); take(h.colIndex, sc->col_index, h.nC);
        take(h.colPos, sc->col_pos, 3 * (size_t)h.nC); take(h.colQuat, sc->col_quat, 4 * (size_t)h.nC);
        take(h.colType, sc->col_type, h.nC); take(h.colParams, sc->col_params, 4 * (size_t)h.nC);
        take(h.colMat, sc->col_material, 3 * (size_t)h.nC); take(h.colTrigger, sc->col_trigger, h.nC);
        take(h.colEnable, sc->col_enable, h.nC); take(h.colData, sc->col_data, h.nC);
        if (n_joints) h.joints.assign(joints, joints + n_joints);
        for (uint32_t e : h.entity) h.maxEntity = std::max(h.maxEntity, e);
        if ((uint64_t)S->nextEntity + h.maxEntity + 1 > 0xffffffffull)
            throw PgxError{PGX_ERR_CAPACITY, "entity id space exhausted"};
        r.entityBase = S->nextEntity;
        S->nextEntity += h.maxEntity + 1;
        r.horizon = horizon_steps;
        r.footprint = footprint_of(h.nA, h.nD, h.nC);
        *request_id = r.id;
        S->queue.push_back(r.id);
        S->reqs.emplace(r.id, std::move(r));
        return PGX_OK;
    });
}

int pgx_server_step(pgx_server* S, float dt, pgx_server_step_info* info) {
    if (!S) return srv_fail(PGX_ERR_INVALID, "server is NULL");
    return srv_guarded([&] {
        World& w = S->w;
        PGX_HIP(hipSetDevice(w.device));
        int admitted = 0, retired = 0, finished = 0;
        // admission boundary: every admit_interval steps, or whenever nothing is running
        bool anyRunning = false;
        for (int id : S->active) anyRunning |= S->reqs[id].state == 1;
        if (S->sinceBoundary >= S->admitInterval || !anyRunning) {
            std::vector<int> keep;
            for (size_t i = 0; i < S->active.size(); ++i) {
                auto it = S->reqs.find(S->active[i]);
                Request& r = it->second;
                if (r.state == 2 || r.state == 3) {
                    S->footprint -= r.footprint;
                    r.inBatch = false;
                    ++retired;
                    if (r.state == 3) S->reqs.erase(it);   // its result was taken already
                } else {
                    keep.push_back(r.id);
                }
            }
            while (!S->queue.empty() && (int)keep.size() < S->maxWorlds) {
                Request& r = S->reqs[S->queue.front()];
                if (!keep.empty() && S->footprint + r.footprint > S->budget) break;
                S->queue.pop_front();
                r.state = 1;
                r.inBatch = true;
                S->footprint += r.footprint;
                keep.push_back(r.id);
                ++admitted;
            }
            S->active.swap(keep);
            if (retired || admitted) {
                if (S->active.empty()) {
                    S->initialised = false;
                    S->worlds = 0;
                } else {
                    assemble(*S);
                    S->initialised = true;
                }
            }
            S->sinceBoundary = 0;
        }
        long long bodies = 0;
        int advanced = 0;
        if (S->initialised) {
            world_step(w, dt);
            ++S->sinceBoundary;
            for (int id : S->active) {
                Request& r = S->reqs[id];
                if (r.state != 1) continue;
                ++r.stepsDone;
                ++advanced;
                bodies += r.sc.nD;
                if (r.stepsDone == r.horizon) {   // state right after the last step
                    read_result(*S, r);
                    r.state = 2;
                    ++finished;
                }
            }
            if (finished) PGX_HIP(hipStreamSynchronize(w.stream));
        }
        if (info) {
            info->active = advanced;
            info->admitted = admitted;
            info->retired = retired;
            info->finished = finished;
            info->queued = (int)S->queue.size();
            info->bodies = bodies;
            info->footprint_bytes = S->footprint;
        }
        return PGX_OK;
    });
}

int pgx_server_set_graph(pgx_server* S, int enabled) {
    if (!S) return srv_fail(PGX_ERR_INVALID, "server is NULL");
    S->w.graph = enabled != 0;
    return PGX_OK;
}

int pgx_server_set_tuning(pgx_server* S, const pgx_tuning* t) {
    if (!S || !t) return srv_fail(PGX_ERR_INVALID, "NULL argument");
    if ((t->solve_block != 64 && t->solve_block != 128 && t->solve_block != 256) || t->color_batch < 0 ||
        (t->pairs_per_lane != 1 && t->pairs_per_lane != 2 && t->pairs_per_lane != 4) || t->color_rounds < 1 ||
        t->color_rounds > 1024 || t->eager_rows_max < 0 || (t->pair_search != 0 && t->pair_search != 1))
        return srv_fail(PGX_ERR_INVALID, "tuning value out of range");
    S->w.tune.solveBlock = t->solve_block;
    S->w.tune.colorBatch = t->color_batch;
    S->w.tune.pairsPerLane = t->pairs_per_lane;
    S->w.tune.colorRounds = t->color_rounds;
    S->w.tune.eagerMax = t->eager_rows_max;
    S->w.tune.pairSearch = t->pair_search;
    return PGX_OK;
}

int pgx_server_done(pgx_server* S, int id) {
    if (!S) return srv_fail(PGX_ERR_INVALID, "server is NULL");
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return srv_fail(PGX_ERR_INVALID, "unknown request");
    if (it->second.state == 3) return srv_fail(PGX_ERR_INVALID, "request already taken");
    return it->second.state == 2 ? 1 : 0;
}

int pgx_server_take_result(pgx_server* S, int id, float* pos, float* quat, float* vel, float* angvel) {
    if (!S) return srv_fail(PGX_ERR_INVALID, "server is NULL");
    auto it = S->reqs.find(id);
    if (it == S->reqs.end() || it->second.state != 2) return srv_fail(PGX_ERR_INVALID, "request not finished");
    const Request& r = it->second;
    const HostScene& h = r.sc;
    for (int a = 0; a < h.nA; ++a) {
        int d = h.dynOfActor[a];
        if (pos) {
            if (d >= 0) { pos[3 * a] = r.rpos[d].x; pos[3 * a + 1] = r.rpos[d].y; pos[3 * a + 2] = r.rpos[d].z; }
            else std::memcpy(pos + 3 * a, h.pos.data() + 3 * a, 3 * sizeof(float));
        }
        if (quat) {
            if (d >= 0) { quat[4 * a] = r.rquat[d].x; quat[4 * a + 1] = r.rquat[d].y; quat[4 * a + 2] = r.rquat[d].z; quat[4 * a + 3] = r.rquat[d].w; }
            else std::memcpy(quat + 4 * a, h.quat.data() + 4 * a, 4 * sizeof(float));
        }
    }
    for (int d = 0; d < h.nD; ++d) {
        if (vel) { vel[3 * d] = r.rvel[d].x; vel[3 * d + 1] = r.rvel[d].y; vel[3 * d + 2] = r.rvel[d].z; }
        if (angvel) { angvel[3 * d] = r.rang[d].x; angvel[3 * d + 1] = r.rang[d].y; angvel[3 * d + 2] = r.rang[d].z; }
    }
    if (r.inBatch) {
        // still in the batched world until the next boundary, which retires it (its
        // footprint leaves the budget there) and erases it; the result is gone already
        Request& q = it->second;
        q.state = 3;
        std::vector<float4>().swap(q.rpos); std::vector<float4>().swap(q.rquat);
        std::vector<float4>().swap(q.rvel); std::vector<float4>().swap(q.rang);
    } else {
        S->reqs.erase(it);
    }
    return PGX_OK;
}

}  // extern "C"
