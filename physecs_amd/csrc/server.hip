// Serving: continuous batching of independent scenes in one device world.
//
// Every request is one world of a batched device world (worlds never interact; see the
// world id in the broadphase keys). At each step boundary the server retires requests
// that reached their horizon (their final state is copied out) and admits queued ones
// while the footprint budget allows; when the set changed, the batched world is
// re-initialised in place (device buffers are reused) from the concatenated descriptors,
// and the persistent state of continuing requests -- body poses and velocities, the
// restitution contact cache, the bounds margin of their dynamic colliders -- is carried
// over on the device. Worlds keep their admission order, so a continuing request's
// contact cache stays sorted after its world id is renumbered.
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
    int state = 0;                      // 0 queued, 1 active, 2 done (may still ride in the batch)
    int dynStart = 0;                   // placement in the current batched world
    int jointStart = 0;                 // first joint in the current batched world
    std::vector<float4> rpos, rquat, rvel, rang;   // final dynamic state (done)
};

// estimated HBM bytes of one request: per collider the pair and manifold slots its
// share of the capacities (12 pair slots per collider) carries, plus per-body state
double footprint_of(int nA, int nD, int nC) {
    const double perPairSlot = 1136.0;   // narrowphase slot + manifold slot + contact cache
    return 12.0 * nC * perPairSlot + 400.0 * nA + 320.0 * nD + 200.0 * nC;
}

__global__ void k_copy_segments(int nseg, const int4* __restrict__ seg, int total, const float4* __restrict__ s0,
                                const float4* __restrict__ s1, const float4* __restrict__ s2,
                                const float4* __restrict__ s3, float4* __restrict__ d0, float4* __restrict__ d1,
                                float4* __restrict__ d2, float4* __restrict__ d3) {
    // seg[i] = {source start, destination start, length, first flat index}
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (seg[mid].w <= i) lo = mid; else hi = mid - 1;
    }
    int4 g = seg[lo];
    int k = i - g.w;
    d0[g.y + k] = s0[g.x + k];
    d1[g.y + k] = s1[g.x + k];
    d2[g.y + k] = s2[g.x + k];
    d3[g.y + k] = s3[g.x + k];
}

__global__ void k_cache_flag(int n, const uint64_t* __restrict__ key, int oldLocalBits,
                             const int32_t* __restrict__ worldMap, int32_t* __restrict__ flag) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    flag[e] = worldMap[(int)(key[e] >> (2 * oldLocalBits))] >= 0 ? 1 : 0;
}

__global__ void k_cache_remap(int n, const uint64_t* __restrict__ key, const int32_t* __restrict__ tri,
                              const int32_t* __restrict__ num, const float4* __restrict__ lp, int oldLocalBits,
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
    for (int j = 0; j < 4; ++j) nLp[(size_t)j * newCap + q] = lp[(size_t)j * n + e];
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
    // staging of the carried-over state
    DBuf<float4> tPos, tQuat, tVel, tAng, tLp;
    DBuf<uint64_t> tKey;
    DBuf<int32_t> tTri, tNum, worldMap, flag, fpos;
    DBuf<uint8_t> worldMask;
    DBuf<int4> seg;
};

static int srv_fail(int code, const std::string& msg) { return report_error(code, msg); }

namespace {

// concatenation of the active requests' descriptors (entity ids shifted, world ids in
// admission order)
struct Combined {
    std::vector<uint32_t> entity;
    std::vector<float> pos, quat, vel, angvel, invMass, com, invI, colPos, colQuat, colParams, colMat;
    std::vector<int32_t> dynOfActor, actorOfDyn, colActor, colIndex, colType, colData, actorWorld;
    std::vector<uint8_t> kinematic, colTrigger, colEnable;
    std::vector<pgx_joint_desc> joints;
    int32_t zero[2] = {0, 0};
    float fzero[3] = {0, 0, 0};
    pgx_scene_desc desc{};
};

void build_combined(pgx_server& S, Combined& c) {
    int aOff = 0, dOff = 0, wd = 0;
    for (int id : S.active) {
        Request& r = S.reqs[id];
        const HostScene& h = r.sc;
        r.dynStart = dOff;
        for (int a = 0; a < h.nA; ++a) {
            c.entity.push_back(r.entityBase + h.entity[a]);
            c.dynOfActor.push_back(h.dynOfActor[a] >= 0 ? h.dynOfActor[a] + dOff : -1);
            c.actorWorld.push_back(wd);
        }
        c.pos.insert(c.pos.end(), h.pos.begin(), h.pos.end());
        c.quat.insert(c.quat.end(), h.quat.begin(), h.quat.end());
        for (int d = 0; d < h.nD; ++d) c.actorOfDyn.push_back(h.actorOfDyn[d] + aOff);
        c.kinematic.insert(c.kinematic.end(), h.kinematic.begin(), h.kinematic.end());
        c.vel.insert(c.vel.end(), h.vel.begin(), h.vel.end());
        c.angvel.insert(c.angvel.end(), h.angvel.begin(), h.angvel.end());
        c.invMass.insert(c.invMass.end(), h.invMass.begin(), h.invMass.end());
        c.com.insert(c.com.end(), h.com.begin(), h.com.end());
        c.invI.insert(c.invI.end(), h.invI.begin(), h.invI.end());
        for (int k = 0; k < h.nC; ++k) c.colActor.push_back(h.colActor[k] + aOff);
        c.colIndex.insert(c.colIndex.end(), h.colIndex.begin(), h.colIndex.end());
        c.colPos.insert(c.colPos.end(), h.colPos.begin(), h.colPos.end());
        c.colQuat.insert(c.colQuat.end(), h.colQuat.begin(), h.colQuat.end());
        c.colType.insert(c.colType.end(), h.colType.begin(), h.colType.end());
        c.colParams.insert(c.colParams.end(), h.colParams.begin(), h.colParams.end());
        c.colMat.insert(c.colMat.end(), h.colMat.begin(), h.colMat.end());
        c.colTrigger.insert(c.colTrigger.end(), h.colTrigger.begin(), h.colTrigger.end());
        c.colEnable.insert(c.colEnable.end(), h.colEnable.begin(), h.colEnable.end());
        c.colData.insert(c.colData.end(), h.colData.begin(), h.colData.end());
        r.jointStart = (int)c.joints.size();
        for (pgx_joint_desc j : h.joints) {
            j.actor0 += aOff;
            j.actor1 += aOff;
            c.joints.push_back(j);
        }
        aOff += h.nA;
        dOff += h.nD;
        ++wd;
    }
    pgx_scene_desc& d = c.desc;
    d.n_actors = aOff;
    d.entity = c.entity.data(); d.pos = c.pos.data(); d.quat = c.quat.data(); d.dyn_of_actor = c.dynOfActor.data();
    d.n_dyn = dOff;
    d.actor_of_dyn = c.actorOfDyn.data(); d.kinematic = c.kinematic.data(); d.vel = c.vel.data();
    d.angvel = c.angvel.data(); d.inv_mass = c.invMass.data(); d.com = c.com.data(); d.inv_inertia = c.invI.data();
    d.n_colliders = (int)c.colType.size();
    d.col_actor = c.colActor.data(); d.col_index = c.colIndex.data(); d.col_pos = c.colPos.data();
    d.col_quat = c.colQuat.data(); d.col_type = c.colType.data(); d.col_params = c.colParams.data();
    d.col_material = c.colMat.data(); d.col_trigger = c.colTrigger.data(); d.col_enable = c.colEnable.data();
    d.col_data = c.colData.data();
    d.n_worlds = wd;
    d.actor_world = c.actorWorld.data();
    d.n_convex = 0; d.n_trimesh = 0;
    d.convex_vert_start = c.zero; d.convex_face_start = c.zero; d.tri_vert_start = c.zero; d.tri_start = c.zero;
}

// final dynamic state of a request leaving the batch
void read_result(pgx_server& S, Request& r) {
    World& w = S.w;
    int n = r.sc.nD;
    r.rpos.resize(n); r.rquat.resize(n); r.rvel.resize(n); r.rang.resize(n);
    if (n) {
        PGX_HIP(hipMemcpyAsync(r.rpos.data(), w.bpos.p + r.dynStart, sizeof(float4) * n, hipMemcpyDeviceToHost, w.stream));
        PGX_HIP(hipMemcpyAsync(r.rquat.data(), w.bquat.p + r.dynStart, sizeof(float4) * n, hipMemcpyDeviceToHost, w.stream));
        PGX_HIP(hipMemcpyAsync(r.rvel.data(), w.dvel.p + r.dynStart, sizeof(float4) * n, hipMemcpyDeviceToHost, w.stream));
        PGX_HIP(hipMemcpyAsync(r.rang.data(), w.dang.p + r.dynStart, sizeof(float4) * n, hipMemcpyDeviceToHost, w.stream));
    }
}

// re-initialises the batched world for the current active set, carrying over the state
// of the requests in `continuing` (old world index -> request id)
void rebuild(pgx_server& S, const std::vector<std::pair<int, int>>& continuing, int oldW, int oldLocalBits) {
    World& w = S.w;
    hipStream_t s = w.stream;
    // 1. stage the continuing requests' dynamic state and the contact cache
    std::vector<int> oldDynStart(continuing.size()), nd(continuing.size());
    int total = 0;
    for (size_t i = 0; i < continuing.size(); ++i) {
        const Request& r = S.reqs[continuing[i].second];
        oldDynStart[i] = r.dynStart;
        nd[i] = r.sc.nD;
        total += r.sc.nD;
    }
    std::vector<int4> seg;
    if (total) {
        S.tPos.alloc(total); S.tQuat.alloc(total); S.tVel.alloc(total); S.tAng.alloc(total);
        int flat = 0;
        for (size_t i = 0; i < continuing.size(); ++i) {
            if (!nd[i]) continue;
            seg.push_back(make_int4(oldDynStart[i], flat, nd[i], flat));
            flat += nd[i];
        }
        S.seg.alloc(seg.size());
        PGX_HIP(hipMemcpyAsync(S.seg.p, seg.data(), sizeof(int4) * seg.size(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_copy_segments, dim3(div_up(total, 256)), dim3(256), 0, s, (int)seg.size(), S.seg.p, total,
                           w.bpos.p, w.bquat.p, w.dvel.p, w.dang.p, S.tPos.p, S.tQuat.p, S.tVel.p, S.tAng.p);
        PGX_LAUNCH_CHECK();
    }
    // gear joints keep angles across steps (GearJoint.h:7-14): stage every joint's state
    std::vector<float4> gearOld;
    std::vector<int32_t> gearInitOld;
    std::vector<int> oldJointStart(continuing.size());
    for (size_t i = 0; i < continuing.size(); ++i) oldJointStart[i] = S.reqs[continuing[i].second].jointStart;
    if (!continuing.empty() && w.J > 0) {
        gearOld.resize(w.J);
        gearInitOld.resize(w.J);
        PGX_HIP(hipMemcpyAsync(gearOld.data(), w.jGear.p, sizeof(float4) * w.J, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipMemcpyAsync(gearInitOld.data(), w.jGearInit.p, sizeof(int32_t) * w.J, hipMemcpyDeviceToHost, s));
    }
    int nCache = continuing.empty() ? 0 : w.cacheCount[w.cacheCur];
    if (nCache) {
        int cur = w.cacheCur;
        S.tKey.alloc(nCache); S.tTri.alloc(nCache); S.tNum.alloc(nCache); S.tLp.alloc(4 * (size_t)nCache);
        PGX_HIP(hipMemcpyAsync(S.tKey.p, w.cacheKey[cur].p, sizeof(uint64_t) * nCache, hipMemcpyDeviceToDevice, s));
        PGX_HIP(hipMemcpyAsync(S.tTri.p, w.cacheTri[cur].p, sizeof(int32_t) * nCache, hipMemcpyDeviceToDevice, s));
        PGX_HIP(hipMemcpyAsync(S.tNum.p, w.cacheNum[cur].p, sizeof(int32_t) * nCache, hipMemcpyDeviceToDevice, s));
        for (int j = 0; j < 4; ++j)
            PGX_HIP(hipMemcpyAsync(S.tLp.p + (size_t)j * nCache, w.cacheLp[cur].p + (size_t)j * w.manifoldCap,
                                   sizeof(float4) * nCache, hipMemcpyDeviceToDevice, s));
    }
    PGX_HIP(hipStreamSynchronize(s));
    // 2. re-initialise from the concatenated descriptors (new placements set here)
    Combined c;
    build_combined(S, c);
    world_init(w, &S.cfg, &c.desc);
    if (!c.joints.empty()) world_add_joints(w, (int)c.joints.size(), c.joints.data());
    if (!gearOld.empty() && w.J > 0) {
        std::vector<float4> gearNew(w.J, make_float4(0.f, 0.f, 0.f, 0.f));
        std::vector<int32_t> gearInitNew(w.J, 0);
        for (size_t i = 0; i < continuing.size(); ++i) {
            const Request& r = S.reqs[continuing[i].second];
            for (size_t k = 0; k < r.sc.joints.size(); ++k) {
                gearNew[r.jointStart + k] = gearOld[oldJointStart[i] + k];
                gearInitNew[r.jointStart + k] = gearInitOld[oldJointStart[i] + k];
            }
        }
        PGX_HIP(hipMemcpyAsync(w.jGear.p, gearNew.data(), sizeof(float4) * w.J, hipMemcpyHostToDevice, s));
        PGX_HIP(hipMemcpyAsync(w.jGearInit.p, gearInitNew.data(), sizeof(int32_t) * w.J, hipMemcpyHostToDevice, s));
        PGX_HIP(hipStreamSynchronize(s));   // the host vectors end with this scope
    }
    // 3. put the staged state back at the continuing requests' new placements
    std::vector<int> newWorldOf(oldW, -1);
    std::vector<uint8_t> mask(std::max(1, w.W), 0);
    {
        std::unordered_map<int, int> widx;
        for (size_t i = 0; i < S.active.size(); ++i) widx[S.active[i]] = (int)i;
        int flat = 0;
        seg.clear();
        for (size_t i = 0; i < continuing.size(); ++i) {
            int nw = widx[continuing[i].second];
            newWorldOf[continuing[i].first] = nw;
            mask[nw] = 1;
            if (!nd[i]) continue;
            seg.push_back(make_int4(flat, S.reqs[continuing[i].second].dynStart, nd[i], flat));
            flat += nd[i];
        }
    }
    if (total) {
        PGX_HIP(hipMemcpyAsync(S.seg.p, seg.data(), sizeof(int4) * seg.size(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_copy_segments, dim3(div_up(total, 256)), dim3(256), 0, s, (int)seg.size(), S.seg.p, total,
                           S.tPos.p, S.tQuat.p, S.tVel.p, S.tAng.p, w.bpos.p, w.bquat.p, w.dvel.p, w.dang.p);
        PGX_LAUNCH_CHECK();
    }
    if (!continuing.empty()) {
        // a continuing request's dynamic colliders carry the margin of its last step
        S.worldMask.alloc(mask.size());
        PGX_HIP(hipMemcpyAsync(S.worldMask.p, mask.data(), mask.size(), hipMemcpyHostToDevice, s));
        bp_compute_bounds(w, true, S.worldMask.p);
    }
    if (nCache) {
        S.worldMap.alloc(oldW);
        PGX_HIP(hipMemcpyAsync(S.worldMap.p, newWorldOf.data(), sizeof(int32_t) * oldW, hipMemcpyHostToDevice, s));
        S.flag.alloc(nCache); S.fpos.alloc(nCache);
        hipLaunchKernelGGL(k_cache_flag, dim3(div_up(nCache, 256)), dim3(256), 0, s, nCache, S.tKey.p, oldLocalBits,
                           S.worldMap.p, S.flag.p);
        exclusive_scan_i32(s, S.flag.p, S.fpos.p, nCache, w.counters.p + 15, w.scanT);
        hipLaunchKernelGGL(k_cache_remap, dim3(div_up(nCache, 256)), dim3(256), 0, s, nCache, S.tKey.p, S.tTri.p,
                           S.tNum.p, S.tLp.p, oldLocalBits, w.localBits, S.worldMap.p, S.flag.p, S.fpos.p,
                           w.cacheKey[0].p, w.cacheTri[0].p, w.cacheNum[0].p, w.cacheLp[0].p, w.manifoldCap);
        PGX_LAUNCH_CHECK();
        int kept = 0;
        PGX_HIP(hipMemcpyAsync(&kept, w.counters.p + 15, sizeof(int), hipMemcpyDeviceToHost, s));
        PGX_HIP(hipStreamSynchronize(s));
        w.cacheCur = 0;
        w.cacheCount[0] = kept;
    }
    PGX_HIP(hipStreamSynchronize(s));
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
            std::vector<std::pair<int, int>> continuing;   // (old world, request)
            for (size_t i = 0; i < S->active.size(); ++i) {
                Request& r = S->reqs[S->active[i]];
                if (r.state == 2) {
                    S->footprint -= r.footprint;
                    ++retired;
                } else {
                    keep.push_back(r.id);
                    continuing.push_back({(int)i, r.id});
                }
            }
            while (!S->queue.empty() && (int)keep.size() < S->maxWorlds) {
                Request& r = S->reqs[S->queue.front()];
                if (!keep.empty() && S->footprint + r.footprint > S->budget) break;
                S->queue.pop_front();
                r.state = 1;
                S->footprint += r.footprint;
                keep.push_back(r.id);
                ++admitted;
            }
            int oldW = (int)S->active.size();
            int oldLocalBits = w.localBits;
            S->active.swap(keep);
            if (retired || admitted) {
                if (S->active.empty()) S->initialised = false;
                else {
                    rebuild(*S, continuing, oldW, oldLocalBits);
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

int pgx_server_done(pgx_server* S, int id) {
    if (!S) return srv_fail(PGX_ERR_INVALID, "server is NULL");
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return srv_fail(PGX_ERR_INVALID, "unknown request");
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
    S->reqs.erase(it);
    return PGX_OK;
}

}  // extern "C"
