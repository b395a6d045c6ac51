// Oracle world step: CPU restatement of physecs::Scene::simulate (src/Physecs.cpp:112-561)
// for primitive shapes, with a C ABI for the tests and for the CPU-baseline timing.
//
// TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library, and only as the checker or the timed CPU baseline.
//
// What each part follows:
//   SAP sort + sweep             src/Physecs.cpp:121-172
//   narrowphase + contact setup  src/Physecs.cpp:181-317 (ThreadPool order: src/ThreadPool.cpp:3-10)
//   substep loop                 src/Physecs.cpp:363-531
//   contact rows                 src/ContactConstraints.cpp:4-124
//   cache swap + bounds          src/Physecs.cpp:553-559, 79-90
//   joint rows are not restated here (no joints in the plain configurations).
//   trigger pairs + events       src/Physecs.cpp:199-206 (overlap), 538-552 (enter/exit, swap)
// Differences by design: the order of contact constraints can be injected (set_order) so
// that a solver fed the GPU's graph color order can be compared with it; trigger events are
// kept as sorted lists (the reference calls its listeners in unordered_set order).
#include <algorithm>
#include <cmath>
#include <limits>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <map>
#include <set>
#include <tuple>
#include <unordered_set>
#include <vector>
#include "orc_collision.h"

using namespace orc;

namespace {

struct Actor { uint32_t entity; V3 pos; Q quat; int dyn; };
struct Dyn { int actor; bool kinematic; V3 v, w; float invMass; V3 com; M3 invI; };
struct Collider {
    int actor, index; V3 lpos; Q lquat; Geom geom;
    float friction, restitution, damping; bool trigger, enable; int data;
    bool alive = true;   // false after Scene::clearColliders
};
// ContactPair (include/Physecs/Physecs.h:90-99): entity0 < entity1 and each collider's
// index within its entity, the key of triggerCache
using TrigKey = std::tuple<uint32_t, int, uint32_t, int>;
struct BPEntry { int collider; uint32_t entity; int colliderIndex; Bounds bounds; bool enable; bool isDynamic; };
struct Pair { uint32_t e0; int c0; uint32_t e1; int c1; int g0, g1; };
struct PointC {
    V3 r0, r1, r0xn, r1xn, t, r0xt, r1xt;
    float tv, c, lN, lT;
    V3 r0xnt, r1xnt, r0xtt, r1xtt;
    float invEffN, invEffT;
};
struct CC {
    int a0, a1, d0, d1, b0, b1;
    V3 n; float friction; bool isSoft; float freq, damp; int numPoints;
    PointC pts[4];
    int32_t key[5];          // e0, c0, e1, c1, triangleIndex
    Manifold raw;            // narrowphase result (world points), for comparisons
};
struct Key {
    uint32_t e0; int c0; uint32_t e1; int c1; int tri;
    bool operator==(const Key& o) const { return e0 == o.e0 && c0 == o.c0 && e1 == o.e1 && c1 == o.c1 && tri == o.tri; }
};
struct KeyHash {
    size_t operator()(const Key& k) const {
        uint64_t h = ((uint64_t)k.e0 << 32) ^ (uint64_t)(uint32_t)k.e1;
        h ^= (uint64_t)(uint32_t)k.c0 * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)(uint32_t)k.c1 * 0xC2B2AE3D27D4EB4Full;
        h ^= (uint64_t)(uint32_t)k.tri * 0x165667B19E3779F9ull;
        return (size_t)h;
    }
};
struct CacheData { int numPoints; V3 lp[4]; float tv[4]; };

// ---- joints (include/Physecs/Joint.h, src/Joints/*.cpp, include/Physecs/Constraint1D*.h)
enum JointType { J_SPHERICAL = 0, J_UNIVERSAL = 1, J_REVOLUTE = 2, J_FIXED = 3, J_PRISMATIC = 4, J_GEAR = 5, J_SERVO = 6,
                 J_CUSTOM = 7 };
// A user Joint subclass (include/Physecs/Joint.h:42-70) through two host functions:
// getSolverDesc's layout (row flags in creation order, returns the row count) and
// MakeConstraintsFunc (writes the rows' fields). World-space data: p0, p1, r0, r1, u0, u1
// (30 floats, matrices column-major); a row: linear, angular0, angular1, targetVelocity,
// c, min, max, frequency, dampingRatio (15 floats).
typedef int (*CustomLayoutFn)(void* user, int joint, const float* ws, int32_t* flags);
typedef void (*CustomRowsFn)(void* user, int joint, const float* ws, float* rows, int n);
enum RowFlags { F_NONE = 0, F_SOFT = 1, F_ANGULAR = 2, F_LIMITED = 4 };
// flag-type order of Constraint1DContainer's collections (Constraint1DContainer.h:117)
static const int kFlagOrder[6] = {F_NONE, F_ANGULAR, F_SOFT, F_LIMITED, F_ANGULAR | F_SOFT, F_ANGULAR | F_LIMITED};

// Planted mistakes, for tests that show a pin catches them (0 = none, the restatement as is).
// Each one is a plausible misreading of the reference at the place it names.
enum Mutation {
    MUT_NONE = 0,
    MUT_SIMD_WARM_LOCAL = 1,        // Constraint1DW.cpp:137-161: warm impulse kept out of the written velocities
    MUT_SIMD_LIMIT_UNSCALED = 2,    // Constraint1DW.cpp:196-200: clamp to [min, max] instead of [min*h, max*h]
    MUT_SIMD_MASK_NONSTRICT = 3,    // Constraint1DW.cpp:140-142: <= instead of <
    MUT_SIMD_BIAS_FLOAT = 4,        // Constraint1DW.cpp:175: 0.2f * c / h instead of (float)(0.2 / h) * c
    MUT_SIMD_ZERO_WRITEBACK = 5,    // Constraint1DW.cpp:164,213-215: lanes with invEff 0 written back
    MUT_FRICTION_SOLVER_VEL = 6,    // ContactConstraints.cpp:92-101: friction from the solver velocities
    MUT_NO_RESTITUTION_CACHE = 7,   // Physecs.cpp:291-304: cached target velocities ignored
    MUT_RELAX_SOFT = 8,             // Physecs.cpp:515-520: soft contacts relaxed too
    MUT_JOINT_AXIS_ORDER = 9,       // Joints/*.cpp: angular row axis cross(a, b) instead of cross(b, a)
    MUT_JOINT_NO_COM = 10,          // Physecs.cpp:343-344: anchor not taken relative to the COM
    MUT_JOINT_UNIVERSAL_AXIS = 11,  // UniversalJoint.cpp:102-105: u1[1] instead of u1[2]
    MUT_FILTER_ARG_ORDER = 12,      // Physecs.cpp:200: the filter called with entity1's collider first
    MUT_CLEAR_DROPS_CACHE = 13,     // Physecs.cpp:725-737: the entity's contact cache entries erased too
};
static int g_mutation = MUT_NONE;
struct Joint {
    int type, a0, a1;                  // actor indices
    V3 p0; Q q0; V3 p1; Q q1;          // anchors in body space
    int driveEnabled; float driveVelocity, driveMaxTorque;
    // PrismaticJointData (include/Physecs/Joints/PrismaticJoint.h:8-17), ServoJointData
    // (ServoJoint.h:7-11), GearJointData (GearJoint.h:7-14)
    float lowerLimit, upperLimit, target, stiffness, damping, gearRatio;
    bool makeUpperLimit, makeLowerLimit;
    float persistentAngle0, persistentAngle1, virtualAngle0, virtualAngle1;
    bool gearInitialized;
    bool alive;
    int color;
};
struct Row {
    int flags, b0, b1;
    V3 linear, angular0, angular1;
    float tv, c, mn, mx, freq, damp;
    V3 linear0t, linear1t, angular0t, angular1t;
    float invEff, lambda;
};
struct JointRows {                     // one joint's rows, in solve order
    int joint;
    V3 r0, r1;                         // anchor relative to the COM (or the anchor) in body space
    int nRows;
    Row rows[7];
};

}  // namespace

// The reference's ThreadPool (src/ThreadPool.cpp:13-46): workers created once, started for
// each parallelFor, the caller taking tasks too and then spinning until every worker is done.
// (C++17 here: workers wait on a condition variable instead of std::atomic::wait.)
struct Pool {
    std::vector<std::thread> threads;
    std::mutex m;
    std::condition_variable cv;
    uint64_t gen = 0;
    bool quit = false;
    std::function<void()> job;
    std::atomic<int> running{0};
    explicit Pool(int n) {
        for (int t = 0; t < n; ++t) threads.emplace_back([this] { loop(); });
    }
    void loop() {
        uint64_t seen = 0;
        while (true) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return quit || gen != seen; });
                if (quit) return;
                seen = gen;
                f = job;
            }
            f();
            running.fetch_sub(1);
        }
    }
    void run(const std::function<void()>& f) {
        {
            std::lock_guard<std::mutex> lk(m);
            job = f;
            running.store((int)threads.size());
            ++gen;
        }
        cv.notify_all();
        f();
        while (running.load() != 0) std::this_thread::yield();
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(m);
            quit = true;
        }
        cv.notify_all();
        for (auto& t : threads) t.join();
    }
};

struct orc_world {
    std::vector<ConvexMesh> convex;
    std::vector<TriMesh> trimesh;
    std::vector<Actor> actors;
    std::vector<Dyn> dyn;
    std::vector<Collider> cols;
    std::vector<BPEntry> bp;
    std::vector<Pair> pairs;
    std::vector<CC> ccs;
    std::unordered_map<Key, CacheData, KeyHash> cache, cacheTemp;
    std::unordered_set<uint64_t> nonColliding;
    // Scene::contactFilter (include/Physecs/Physecs.h:174) tabulated over collider classes:
    // filterTable[k0 * filterK + k1] != 0 is TRIGGER for (collider of entity0, collider of
    // entity1); filterK = 0 is defaultContactFilter (src/Physecs.cpp:20-23)
    int filterK = 0;
    std::vector<uint8_t> filterTable;
    std::vector<int> filterClass;
    // triggerCache / triggerCacheTemp (include/Physecs/Physecs.h:170-171), and the last
    // step's events
    std::set<TrigKey> trig, trigTemp;
    std::vector<TrigKey> trigEnter, trigExit;
    std::vector<int32_t> order;   // injected manifold order (5 ints per manifold)
    std::vector<Joint> joints;
    std::unordered_map<uint32_t, unsigned char> jointBitsets;   // per entity, src/Physecs.cpp:696-705
    std::vector<JointRows> jrows[9];                           // per colour, joints in creation order
    int numSubSteps = 8, numIterations = 2, numThreads = 0;
    CustomLayoutFn custLayout = nullptr;
    CustomRowsFn custRows = nullptr;
    void* custUser = nullptr;
    int jointError = 0;                 // a custom layout that was refused (0: none)
    std::unique_ptr<Pool> pool;   // numThreads workers, made when the thread count is set
    float g = 9.81f;
    std::mutex mtx;
};

extern "C" {

struct orc_scene {
    int n_actors;
    const uint32_t* entity; const float* pos; const float* quat; const int32_t* dyn_of_actor;
    int n_dyn;
    const int32_t* actor_of_dyn; const uint8_t* kinematic; const float* vel; const float* angvel;
    const float* inv_mass; const float* com; const float* inv_inertia;
    int n_colliders;
    const int32_t* col_actor; const int32_t* col_index; const float* col_pos; const float* col_quat;
    const int32_t* col_type; const float* col_params; const float* col_material;
    const uint8_t* col_trigger; const uint8_t* col_enable; const int32_t* col_data;
    int n_worlds; const int32_t* actor_world;       // unused by the oracle (one world)
    // convex meshes; a CONVEX_MESH collider's params are {scale xyz, mesh index}
    int n_convex;
    const int32_t* convex_vert_start; const float* convex_verts;
    const int32_t* convex_face_start; const int32_t* face_idx_start; const int32_t* face_idx;
    const float* face_normal; const float* face_centroid;
    // triangle meshes; a TRIANGLE_MESH collider's params are {0, 0, 0, mesh index}
    int n_trimesh;
    const int32_t* tri_vert_start; const float* tri_verts; const int32_t* tri_start; const int32_t* tri_idx;
};

// ConvexMesh(vertices, faces) (include/Physecs/ConvexMesh.h:29-34, src/ConvexMesh.cpp:3-11):
// the vertex buffer is padded to a multiple of 4 with copies of the last vertex.
static ConvexMesh makeConvex(int nv, const float* verts, int nf, const int32_t* fstart, const int32_t* fidx,
                             const float* normals, const float* centroids) {
    ConvexMesh m;
    m.count = nv;
    for (int i = 0; i < nv; ++i) m.verts.push_back(v3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]));
    int padded = (nv + 3) / 4 * 4;
    V3 last = m.verts.back();
    m.verts.resize(padded, last);
    m.faces.resize(nf);
    for (int i = 0; i < nf; ++i) {
        m.faces[i].idx.assign(fidx + fstart[i], fidx + fstart[i + 1]);
        m.faces[i].normal = v3(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
        m.faces[i].centroid = v3(centroids[3 * i], centroids[3 * i + 1], centroids[3 * i + 2]);
    }
    return m;
}

orc_world* orc_world_create(const orc_scene* s) {
    auto* w = new orc_world();
    w->actors.resize(s->n_actors);
    for (int i = 0; i < s->n_actors; ++i) {
        auto& a = w->actors[i];
        a.entity = s->entity[i];
        a.pos = v3(s->pos[3 * i], s->pos[3 * i + 1], s->pos[3 * i + 2]);
        a.quat = {s->quat[4 * i], s->quat[4 * i + 1], s->quat[4 * i + 2], s->quat[4 * i + 3]};
        a.dyn = s->dyn_of_actor[i];
    }
    w->dyn.resize(s->n_dyn);
    for (int i = 0; i < s->n_dyn; ++i) {
        auto& d = w->dyn[i];
        d.actor = s->actor_of_dyn[i];
        d.kinematic = s->kinematic[i] != 0;
        d.v = v3(s->vel[3 * i], s->vel[3 * i + 1], s->vel[3 * i + 2]);
        d.w = v3(s->angvel[3 * i], s->angvel[3 * i + 1], s->angvel[3 * i + 2]);
        d.invMass = s->inv_mass[i];
        d.com = v3(s->com[3 * i], s->com[3 * i + 1], s->com[3 * i + 2]);
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) d.invI[c][r] = s->inv_inertia[9 * i + 3 * c + r];
    }
    for (int m = 0; m < s->n_convex; ++m) {
        int v0 = s->convex_vert_start[m], f0 = s->convex_face_start[m];
        w->convex.push_back(makeConvex(s->convex_vert_start[m + 1] - v0, s->convex_verts + 3 * v0,
                                       s->convex_face_start[m + 1] - f0, s->face_idx_start + f0, s->face_idx,
                                       s->face_normal + 3 * f0, s->face_centroid + 3 * f0));
    }
    for (int m = 0; m < s->n_trimesh; ++m) {
        int v0 = s->tri_vert_start[m], v1 = s->tri_vert_start[m + 1];
        std::vector<V3> verts;
        for (int i = v0; i < v1; ++i) verts.push_back(v3(s->tri_verts[3 * i], s->tri_verts[3 * i + 1], s->tri_verts[3 * i + 2]));
        std::vector<uint32_t> idx(s->tri_idx + 3 * s->tri_start[m], s->tri_idx + 3 * s->tri_start[m + 1]);
        w->trimesh.push_back(buildTriMesh(verts, idx));
    }
    w->cols.resize(s->n_colliders);
    w->bp.resize(s->n_colliders);
    for (int i = 0; i < s->n_colliders; ++i) {
        auto& c = w->cols[i];
        c.actor = s->col_actor[i];
        c.index = s->col_index[i];
        c.lpos = v3(s->col_pos[3 * i], s->col_pos[3 * i + 1], s->col_pos[3 * i + 2]);
        c.lquat = {s->col_quat[4 * i], s->col_quat[4 * i + 1], s->col_quat[4 * i + 2], s->col_quat[4 * i + 3]};
        c.geom.type = s->col_type[i];
        for (int k = 0; k < 4; ++k) c.geom.p[k] = s->col_params[4 * i + k];
        if (c.geom.type == CONVEX_MESH) c.geom.mesh = &w->convex.at((size_t)c.geom.p[3]);
        if (c.geom.type == TRIANGLE_MESH) c.geom.tmesh = &w->trimesh.at((size_t)c.geom.p[3]);
        c.friction = s->col_material[3 * i];
        c.restitution = s->col_material[3 * i + 1];
        c.damping = s->col_material[3 * i + 2];
        c.trigger = s->col_trigger[i] != 0;
        c.enable = s->col_enable[i] != 0;
        c.data = s->col_data[i];
        // src/Physecs.cpp:25-37: creation-time bounds carry no margin
        const Actor& a = w->actors[c.actor];
        int dy = a.dyn;
        bool isDynamic = dy >= 0 ? !w->dyn[dy].kinematic : false;
        Bounds b = getBounds(a.pos + qrot(a.quat, c.lpos), qmul(a.quat, c.lquat), c.geom);
        w->bp[i] = {i, a.entity, c.index, b, c.enable, isDynamic};
    }
    return w;
}

void orc_world_destroy(orc_world* w) { delete w; }

void orc_world_set_params(orc_world* w, int substeps, int iterations, float gravity, int threads) {
    w->numSubSteps = substeps;
    w->numIterations = iterations;
    w->g = gravity;
    if (threads != w->numThreads || (threads > 0 && !w->pool)) {
        w->pool.reset();
        if (threads > 0) w->pool.reset(new Pool(threads));
    }
    w->numThreads = threads;
}

// src/Physecs.cpp:788-793 (setCanCollide(e0, e1, false))
void orc_world_set_contact_filter(orc_world* w, int k, const uint8_t* table, const int32_t* classes) {
    if (k <= 0 || !table) { w->filterK = 0; return; }
    w->filterK = k;
    w->filterTable.assign(table, table + (size_t)k * k);
    w->filterClass.assign(classes, classes + w->cols.size());
}

void orc_world_set_non_colliding(orc_world* w, uint32_t e0, uint32_t e1) {
    uint32_t a = e0 < e1 ? e0 : e1, b = e0 < e1 ? e1 : e0;
    w->nonColliding.insert(((uint64_t)a << 32) | b);
}

void orc_world_set_order(orc_world* w, const int32_t* keys, int n) {
    w->order.assign(keys, keys + 5 * (size_t)n);
}



}  // extern "C"

// ---------------------------------------------------------------- step pieces
static void sap(orc_world* w) {
    auto& e = w->bp;
    // src/Physecs.cpp:121-133: stable insertion sort on min.x
    for (size_t i = 1; i < e.size(); ++i) {
        BPEntry x = e[i];
        size_t j = i;
        while (j > 0 && e[j - 1].bounds.min.x > x.bounds.min.x) { e[j] = e[j - 1]; --j; }
        if (j < i) e[j] = x;
    }
    // src/Physecs.cpp:134-172
    w->pairs.clear();
    for (size_t i = 0; i + 1 < e.size(); ++i) {
        auto& e0 = e[i];
        if (!e0.enable) continue;
        for (size_t j = i + 1; j < e.size(); ++j) {
            auto& e1 = e[j];
            if (!e1.enable) continue;
            if (e0.entity == e1.entity) continue;
            if (!(e0.isDynamic || e1.isDynamic)) continue;
            const Bounds& b0 = e0.bounds;
            const Bounds& b1 = e1.bounds;
            if (b0.max.x < b1.min.x) break;
            if (b0.max.y < b1.min.y || b0.min.y > b1.max.y) continue;
            if (b0.max.z < b1.min.z || b0.min.z > b1.max.z) continue;
            if (e0.entity < e1.entity) w->pairs.push_back({e0.entity, e0.colliderIndex, e1.entity, e1.colliderIndex, e0.collider, e1.collider});
            else w->pairs.push_back({e1.entity, e1.colliderIndex, e0.entity, e0.colliderIndex, e1.collider, e0.collider});
        }
    }
}

namespace {
bool overlapGeom(V3 p0, Q q0, const Geom& g0, V3 p1, Q q1, const Geom& g1);
}

// src/Physecs.cpp:183-316, one potential contact
static void narrowTask(orc_world* w, int index, std::vector<CC>& outCC, std::vector<std::pair<Key, CacheData>>& outCache) {
    const Pair& cp = w->pairs[index];
    const Collider& col0 = w->cols[cp.g0];
    const Collider& col1 = w->cols[cp.g1];
    const Actor& t0 = w->actors[col0.actor];
    const Actor& t1 = w->actors[col1.actor];
    V3 pos0 = t0.pos + qrot(t0.quat, col0.lpos);
    Q or0 = qmul(t0.quat, col0.lquat);
    V3 pos1 = t1.pos + qrot(t1.quat, col1.lpos);
    Q or1 = qmul(t1.quat, col1.lquat);
    // contactFilter(col0.isTrigger, col0.data, col1.isTrigger, col1.data), col0 the collider of
    // entity0 (src/Physecs.cpp:200); TRIGGER: overlap only (:199-206)
    const int k0 = w->filterK ? w->filterClass[g_mutation == MUT_FILTER_ARG_ORDER ? cp.g1 : cp.g0] : 0;
    const int k1 = w->filterK ? w->filterClass[g_mutation == MUT_FILTER_ARG_ORDER ? cp.g0 : cp.g1] : 0;
    const bool trigger = w->filterK ? w->filterTable[(size_t)k0 * w->filterK + k1] != 0
                                    : (col0.trigger || col1.trigger);
    if (trigger) {
        if (overlapGeom(pos0, or0, col0.geom, pos1, or1, col1.geom)) {
            std::lock_guard<std::mutex> lk(w->mtx);
            w->trigTemp.insert(TrigKey{cp.e0, cp.c0, cp.e1, cp.c1});
        }
        return;
    }
    uint64_t ek = ((uint64_t)cp.e0 << 32) | cp.e1;
    if (w->nonColliding.count(ek)) return;
    std::vector<Manifold> buf;
    if (!collision(pos0, or0, col0.geom, pos1, or1, col1.geom, buf)) return;
    const Dyn* dyn0 = t0.dyn >= 0 ? &w->dyn[t0.dyn] : nullptr;
    const Dyn* dyn1 = t1.dyn >= 0 ? &w->dyn[t1.dyn] : nullptr;
    V3 com0 = v3s(0), v0 = v3s(0), w0 = v3s(0);
    if (dyn0 && !dyn0->kinematic) { com0 = t0.pos + qrot(t0.quat, dyn0->com); v0 = dyn0->v; w0 = dyn0->w; }
    V3 com1 = v3s(0), v1 = v3s(0), w1 = v3s(0);
    if (dyn1 && !dyn1->kinematic) { com1 = t1.pos + qrot(t1.quat, dyn1->com); v1 = dyn1->v; w1 = dyn1->w; }
    for (auto& res : buf) {
        if (!res.numPoints) continue;
        V3 n = res.normal;
        Key key{cp.e0, cp.c0, cp.e1, cp.c1, res.triangleIndex};
        auto it = w->cache.find(key);
        const CacheData* prev = it != w->cache.end() ? &it->second : nullptr;
        CacheData cur{res.numPoints, {}, {}};
        float friction = (col0.friction + col1.friction) * 0.5f;
        bool isSoft; float frequency, damping, restitution;
        if (col0.damping != 0.f || col1.damping != 0.f) {
            isSoft = true;
            if (col0.damping != 0.f && col1.damping != 0.f) {
                frequency = gmin(col0.restitution, col1.restitution);
                damping = gmin(col0.damping, col1.damping);
            } else if (col0.damping != 0.f) {
                frequency = col0.restitution; damping = col0.damping;
            } else {
                frequency = col1.restitution; damping = col1.damping;
            }
            restitution = 0;
        } else {
            isSoft = false; frequency = 0; damping = 0;
            restitution = (col0.restitution + col1.restitution) * 0.5f;
        }
        CC cc;
        cc.a0 = col0.actor; cc.a1 = col1.actor;
        cc.d0 = t0.dyn; cc.d1 = t1.dyn;
        cc.b0 = dyn0 && !dyn0->kinematic ? t0.dyn : -1;
        cc.b1 = dyn1 && !dyn1->kinematic ? t1.dyn : -1;
        cc.n = n; cc.friction = friction; cc.isSoft = isSoft; cc.freq = frequency; cc.damp = damping;
        cc.numPoints = res.numPoints;
        cc.key[0] = (int32_t)cp.e0; cc.key[1] = cp.c0; cc.key[2] = (int32_t)cp.e1; cc.key[3] = cp.c1; cc.key[4] = res.triangleIndex;
        cc.raw = res;
        for (int k = 0; k < res.numPoints; ++k) {
            V3 r0 = res.points[k].p0 - com0;
            V3 r1 = res.points[k].p1 - com1;
            V3 rel = v1 + cross(w1, r1) - v0 - cross(w0, r0);
            float relN = dot(rel, n);
            r0 = qrot(qinverse(t0.quat), r0);
            r1 = qrot(qinverse(t1.quat), r1);
            float tv = 0;
            bool found = false;
            if (prev && g_mutation != MUT_NO_RESTITUTION_CACHE) {
                for (int i = 0; i < prev->numPoints; ++i) {
                    // glm::distance(r0, prev) < 0.1 compares in double (Physecs.cpp:294)
                    if ((double)distance(r0, prev->lp[i]) < 0.1) { found = true; tv = prev->tv[i]; break; }
                }
            }
            if (!found) tv = -restitution * relN;
            cur.lp[k] = r0; cur.tv[k] = tv;
            PointC& pc = cc.pts[k];
            std::memset(&pc, 0, sizeof(pc));
            pc.r0 = r0; pc.r1 = r1; pc.tv = tv;
        }
        outCC.push_back(cc);
        outCache.push_back({key, cur});
    }
}

static void narrowphase(orc_world* w) {
    w->ccs.clear();
    w->cacheTemp.clear();
    w->trigTemp.clear();
    int count = (int)w->pairs.size();
    if (w->numThreads <= 0) {
        // ThreadPool with no workers: the caller drains indices in descending order
        std::vector<CC> cc;
        std::vector<std::pair<Key, CacheData>> cd;
        for (int i = count - 1; i >= 0; --i) {
            cc.clear(); cd.clear();
            narrowTask(w, i, cc, cd);
            for (auto& x : cc) w->ccs.push_back(x);
            for (auto& x : cd) w->cacheTemp[x.first] = x.second;
        }
        return;
    }
    // src/ThreadPool.cpp:3-10,30-46: chunk = max(1, i / ((T+1)*2)), claimed by CAS, descending
    int T = w->numThreads;
    std::atomic<int> current(count);
    auto worker = [&]() {
        std::vector<CC> cc;
        std::vector<std::pair<Key, CacheData>> cd;
        int i = current.load();
        while (i) {
            int n = std::max(1, i / ((T + 1) * 2));
            if (!current.compare_exchange_weak(i, i - n)) continue;
            for (int j = 0; j < n; ++j) {
                cc.clear(); cd.clear();
                narrowTask(w, --i, cc, cd);
                if (!cc.empty()) {
                    std::lock_guard<std::mutex> lk(w->mtx);
                    for (auto& x : cc) w->ccs.push_back(x);
                    for (auto& x : cd) w->cacheTemp[x.first] = x.second;
                }
            }
        }
    };
    w->pool->run(worker);
}

static void applyOrder(orc_world* w) {
    if (w->order.empty()) return;
    size_t n = w->order.size() / 5;
    std::unordered_map<Key, size_t, KeyHash> pos;
    for (size_t i = 0; i < w->ccs.size(); ++i) {
        auto& k = w->ccs[i].key;
        pos[{(uint32_t)k[0], k[1], (uint32_t)k[2], k[3], k[4]}] = i;
    }
    std::vector<CC> out;
    out.reserve(w->ccs.size());
    std::vector<char> used(w->ccs.size(), 0);
    for (size_t i = 0; i < n; ++i) {
        const int32_t* k = &w->order[5 * i];
        auto it = pos.find({(uint32_t)k[0], k[1], (uint32_t)k[2], k[3], k[4]});
        if (it == pos.end()) continue;
        out.push_back(w->ccs[it->second]);
        used[it->second] = 1;
    }
    for (size_t i = 0; i < w->ccs.size(); ++i)
        if (!used[i]) out.push_back(w->ccs[i]);
    w->ccs.swap(out);
    w->order.clear();
}


// ---------------------------------------------------------------- joints
// Joint colouring at creation, src/Physecs.cpp:690-710: the lowest colour free on both
// entities (8 SIMD colours, colour 8 = scalar overflow); static entities take colours too.
static void addJoint(orc_world* w, const Joint& jin) {
    Joint j = jin;
    uint32_t e0 = w->actors[j.a0].entity, e1 = w->actors[j.a1].entity;
    uint32_t lo = e0 < e1 ? e0 : e1, hi = e0 < e1 ? e1 : e0;
    w->nonColliding.insert(((uint64_t)lo << 32) | hi);
    unsigned char& c0 = w->jointBitsets[e0];
    unsigned char& c1 = w->jointBitsets[e1];
    unsigned int un = (unsigned int)(c0 | c1);
    unsigned int inv = ~un;
    int color = 0;
    while (!((inv >> color) & 1u)) ++color;     // _BitScanForward(~colorsUnion)
    c0 = (unsigned char)(c0 | (1u << color));
    c1 = (unsigned char)(c1 | (1u << color));
    j.color = color;
    w->joints.push_back(j);
}

// The rows getSolverDesc declares, in creation order (src/Joints/*.cpp). A prismatic joint
// decides its limit row from the step-start pose (PrismaticJoint.cpp:123-150).
static int rowLayout(const orc_world* w, Joint& j, int* flags) {
    int n = 0;
    if (j.type == J_PRISMATIC) {
        const Actor& t0 = w->actors[j.a0];
        const Actor& t1 = w->actors[j.a1];
        V3 p0 = t0.pos + qrot(t0.quat, j.p0);
        V3 p1 = t1.pos + qrot(t1.quat, j.p1);
        V3 d = p1 - p0;
        V3 u00 = qrot(qmul(t0.quat, j.q0), v3(1, 0, 0));
        float dx = dot(d, u00);
        flags[n++] = F_NONE; flags[n++] = F_NONE;
        for (int k = 0; k < 3; ++k) flags[n++] = F_ANGULAR;
        if (dx > j.upperLimit) { j.makeUpperLimit = true; j.makeLowerLimit = false; flags[n++] = F_LIMITED; }
        else if (dx < j.lowerLimit) { j.makeUpperLimit = false; j.makeLowerLimit = true; flags[n++] = F_LIMITED; }
        else { j.makeUpperLimit = false; j.makeLowerLimit = false; }
        if (j.driveEnabled) flags[n++] = F_SOFT;
        return n;
    }
    if (j.type == J_GEAR) { flags[n++] = F_ANGULAR; return n; }                 // GearJoint.cpp:56-58
    for (int k = 0; k < 3; ++k) flags[n++] = F_NONE;                           // point-to-point
    if (j.type == J_UNIVERSAL) flags[n++] = F_ANGULAR;                         // UniversalJoint.cpp:102-105
    if (j.type == J_REVOLUTE) {                                                // RevoluteJoint.cpp:52-55
        flags[n++] = F_ANGULAR; flags[n++] = F_ANGULAR;
        if (j.driveEnabled) flags[n++] = F_ANGULAR | F_LIMITED;
    }
    if (j.type == J_FIXED) for (int k = 0; k < 3; ++k) flags[n++] = F_ANGULAR;  // FixedJoint.cpp:36-39
    if (j.type == J_SERVO) {                                                   // ServoJoint.cpp:50-53
        flags[n++] = F_ANGULAR; flags[n++] = F_ANGULAR; flags[n++] = F_ANGULAR | F_SOFT;
    }
    return n;
}

// JointSolverData::calculateWorldSpaceData (src/Joint.cpp:5-14) as 30 floats: p0, p1, r0,
// r1, then u0 and u1 by columns
static void jointWorldSpace(const orc_world* w, const Joint& j, const JointRows& jr, float* o) {
    const Actor& t0 = w->actors[j.a0];
    const Actor& t1 = w->actors[j.a1];
    M3 u0 = mat3_cast(qmul(t0.quat, j.q0));
    M3 u1 = mat3_cast(qmul(t1.quat, j.q1));
    V3 v[10] = {t0.pos + qrot(t0.quat, j.p0), t1.pos + qrot(t1.quat, j.p1), qrot(t0.quat, jr.r0), qrot(t1.quat, jr.r1),
                u0[0], u0[1], u0[2], u1[0], u1[1], u1[2]};
    for (int i = 0; i < 10; ++i) { o[3 * i] = v[i].x; o[3 * i + 1] = v[i].y; o[3 * i + 2] = v[i].z; }
}

// getSolverDesc of a custom joint: its rows' flags from the layout function
static int customLayout(orc_world* w, int index, const JointRows& jr, int* flags) {
    float ws[30];
    jointWorldSpace(w, w->joints[index], jr, ws);
    int32_t fl[7] = {0};
    const int n = w->custLayout ? w->custLayout(w->custUser, index, ws, fl) : -1;
    bool ok = n >= 0 && n <= 7;
    for (int r = 0; ok && r < n; ++r) ok = fl[r] == 0 || fl[r] == 1 || fl[r] == 2 || fl[r] == 4 || fl[r] == 3 || fl[r] == 6;
    if (!ok) { w->jointError = 1; return 0; }
    for (int r = 0; r < n; ++r) flags[r] = fl[r];
    return n;
}

// "create joint constraints", src/Physecs.cpp:321-355: rows are rebuilt every step (defaults
// from include/Physecs/Constraint1D.h:19-35, so the accumulated lambda restarts at 0). The
// joints are visited colour by colour, creation order inside a colour, as the reference
// walks jointGraph.colors (destroyJoint erases in place, Physecs.cpp:712-717): the order in
// which a custom joint's layout function is called.
static void buildJointRows(orc_world* w) {
    for (auto& c : w->jrows) c.clear();
    for (int col = 0; col < 9; ++col)
    for (size_t i = 0; i < w->joints.size(); ++i) {
        Joint& j = w->joints[i];
        if (!j.alive || j.color != col) continue;
        JointRows jr{};
        jr.joint = (int)i;
        const Actor& t0 = w->actors[j.a0];
        const Actor& t1 = w->actors[j.a1];
        const Dyn* d0 = t0.dyn >= 0 ? &w->dyn[t0.dyn] : nullptr;
        const Dyn* d1 = t1.dyn >= 0 ? &w->dyn[t1.dyn] : nullptr;
        int b0 = d0 && !d0->kinematic ? t0.dyn : -1;
        int b1 = d1 && !d1->kinematic ? t1.dyn : -1;
        jr.r0 = b0 >= 0 && g_mutation != MUT_JOINT_NO_COM ? j.p0 - d0->com : j.p0;
        jr.r1 = b1 >= 0 && g_mutation != MUT_JOINT_NO_COM ? j.p1 - d1->com : j.p1;
        int fl[7];
        jr.nRows = j.type == J_CUSTOM ? customLayout(w, (int)i, jr, fl) : rowLayout(w, j, fl);
        for (int r = 0; r < jr.nRows; ++r) {
            Row& row = jr.rows[r];
            std::memset(&row, 0, sizeof(row));
            row.flags = fl[r];
            row.b0 = b0; row.b1 = b1;
            row.mn = std::numeric_limits<float>::lowest();
            row.mx = std::numeric_limits<float>::max();
        }
        w->jrows[j.color].push_back(jr);
    }
}

// GearJoint.cpp:5-8: the signed difference of two angles, wrapped to [-pi, pi)
static float angleDiff(float angle0, float angle1) {
    const float pi = 3.14159265358979323846f, twoPi = 6.28318530717958647692f;
    const float diff = std::fmod(angle1 - angle0 + pi, twoPi) - pi;
    return diff < -pi ? diff + twoPi : diff;
}

static M3 mcols(V3 a, V3 b, V3 c) { M3 m; m[0] = a; m[1] = b; m[2] = c; return m; }

// calculateWorldSpaceData (src/Joint.cpp:5-14) + makeConstraints of each type
// (src/JointsUtil.h:8-28, src/Joints/*.cpp)
static void updateJointRows(orc_world* w, JointRows& jr) {
    Joint& j = w->joints[jr.joint];
    if (j.type == J_CUSTOM) {   // MakeConstraintsFunc through the rows function
        float ws[30], q[7 * 15];
        jointWorldSpace(w, j, jr, ws);
        for (int r = 0; r < jr.nRows; ++r) {
            const Row& R = jr.rows[r];
            float* o = q + 15 * r;
            const float f[15] = {R.linear.x, R.linear.y, R.linear.z, R.angular0.x, R.angular0.y, R.angular0.z,
                                 R.angular1.x, R.angular1.y, R.angular1.z, R.tv, R.c, R.mn, R.mx, R.freq, R.damp};
            std::memcpy(o, f, sizeof(f));
        }
        if (w->custRows) w->custRows(w->custUser, jr.joint, ws, q, jr.nRows);
        for (int r = 0; r < jr.nRows; ++r) {
            Row& R = jr.rows[r];
            const float* o = q + 15 * r;
            R.linear = v3(o[0], o[1], o[2]); R.angular0 = v3(o[3], o[4], o[5]); R.angular1 = v3(o[6], o[7], o[8]);
            R.tv = o[9]; R.c = o[10]; R.mn = o[11]; R.mx = o[12]; R.freq = o[13]; R.damp = o[14];
        }
        return;
    }
    const Actor& t0 = w->actors[j.a0];
    const Actor& t1 = w->actors[j.a1];
    M3 u0 = mat3_cast(qmul(t0.quat, j.q0));
    M3 u1 = mat3_cast(qmul(t1.quat, j.q1));
    V3 r0 = qrot(t0.quat, jr.r0);
    V3 r1 = qrot(t1.quat, jr.r1);
    V3 p0 = t0.pos + qrot(t0.quat, j.p0);
    V3 p1 = t1.pos + qrot(t1.quat, j.p1);
    V3 d = p1 - p0;
    Row* R = jr.rows;
    auto angularRow = [&](Row& row, V3 a, V3 b) {   // c = dot(a, b), axis = cross(b, a)
        V3 ax = g_mutation == MUT_JOINT_AXIS_ORDER ? cross(a, b) : cross(b, a);
        row.angular0 = ax; row.angular1 = ax; row.c = dot(a, b);
    };
    if (j.type == J_PRISMATIC) {                    // PrismaticJoint.cpp:6-92
        for (int k = 0; k < 2; ++k) {
            V3 u = u0[k + 1];
            R[k].linear = u; R[k].angular0 = cross(r0, u); R[k].angular1 = cross(r1, u); R[k].c = dot(d, u);
        }
        angularRow(R[2], u0[0], u1[1]);
        angularRow(R[3], u0[0], u1[2]);
        angularRow(R[4], u0[1], u1[2]);
        float dx = dot(d, u0[0]);
        V3 r0xx = cross(r0, u0[0]), r1xx = cross(r1, u0[0]);
        int r = 5;
        if (j.makeUpperLimit || j.makeLowerLimit) {
            Row& L = R[r++];
            L.linear = u0[0]; L.angular0 = r0xx; L.angular1 = r1xx;
            if (j.makeUpperLimit) { L.c = dx - j.upperLimit; L.mn = 0; }
            else { L.c = dx - j.lowerLimit; L.mx = 0; }
        }
        if (j.driveEnabled) {
            Row& S = R[r++];
            S.linear = u0[0]; S.angular0 = r0xx; S.angular1 = r1xx; S.c = dx - j.target;
            S.freq = j.stiffness; S.damp = j.damping;
        }
        return;
    }
    if (j.type == J_GEAR) {                         // GearJoint.cpp:10-47
        float angle0, angle1;
        {
            const V3 p1Proj = p1 + dot(p0 - p1, u0[0]) * u0[0];
            const V3 dir = normalize(p0 - p1Proj);
            const V3 n = cross(u0[0], dir);
            const M3 m = mcols(u0[0], -n, dir);
            const M3 u0t = mulmm(transpose(m), u0);
            angle0 = atan2_cr(u0t[1][2], u0t[2][2]);
        }
        {
            const V3 p0Proj = p0 + dot(p1 - p0, u1[0]) * u1[0];
            const V3 dir = normalize(p0Proj - p1);
            const V3 n = cross(u1[0], dir);
            const M3 m = mcols(u1[0], -n, dir);
            const M3 u1t = mulmm(transpose(u1), m);
            angle1 = atan2_cr(u1t[1][2], u1t[2][2]);
        }
        if (!j.gearInitialized) {
            j.persistentAngle0 = angle0;
            j.persistentAngle1 = angle1;
            j.gearInitialized = true;
        }
        j.virtualAngle0 += angleDiff(angle0, j.persistentAngle0);
        j.virtualAngle1 += angleDiff(angle1, j.persistentAngle1);
        j.persistentAngle0 = angle0;
        j.persistentAngle1 = angle1;
        R[0].angular0 = u0[0] * j.gearRatio;
        R[0].angular1 = -u1[0];
        R[0].c = j.virtualAngle0 * j.gearRatio - j.virtualAngle1;
        return;
    }
    R[0].linear = v3(1, 0, 0); R[0].angular0 = v3(0, r0.z, -r0.y); R[0].angular1 = v3(0, r1.z, -r1.y); R[0].c = d.x;
    R[1].linear = v3(0, 1, 0); R[1].angular0 = v3(-r0.z, 0, r0.x); R[1].angular1 = v3(-r1.z, 0, r1.x); R[1].c = d.y;
    R[2].linear = v3(0, 0, 1); R[2].angular0 = v3(r0.y, -r0.x, 0); R[2].angular1 = v3(r1.y, -r1.x, 0); R[2].c = d.z;
    if (j.type == J_UNIVERSAL) angularRow(R[3], u0[2], g_mutation == MUT_JOINT_UNIVERSAL_AXIS ? u1[1] : u1[2]);
    if (j.type == J_REVOLUTE || j.type == J_FIXED || j.type == J_SERVO) {
        angularRow(R[3], u0[0], u1[1]);
        angularRow(R[4], u0[0], u1[2]);
        if (j.type == J_FIXED) angularRow(R[5], u0[1], u1[2]);
        if (j.type == J_REVOLUTE && j.driveEnabled) {
            R[5].angular0 = u0[0]; R[5].angular1 = u0[0];
            R[5].tv = j.driveVelocity; R[5].mx = j.driveMaxTorque; R[5].mn = -j.driveMaxTorque;
        }
        if (j.type == J_SERVO) {                    // ServoJoint.cpp:9-37, glm::orientedAngle
            V3 x = u0[2], y = u1[2], ref = u0[0];
            const float angle = acos_cr(gclamp(dot(x, y), -1.f, 1.f));
            const float oriented = dot(ref, cross(x, y)) < 0.f ? -angle : angle;
            R[5].angular0 = u0[0]; R[5].angular1 = u0[0];
            R[5].c = oriented - j.target;
            R[5].freq = j.stiffness; R[5].damp = j.damping;
        }
    }
}

struct VelocityData;
struct PseudoVelocityData;
struct MassData;

struct VelocityData { V3 v, w; };
struct PseudoVelocityData { V3 pv, pw; int count; };
struct MassData { float invMass; M3 invI; };

// Constraint1D<flags>::preSolve (src/Constraint1D.cpp:5-52) and the per-lane meaning of
// Constraint1DW<flags>::preSolve (src/Constraint1DW.cpp:5-115): the same arithmetic; the
// SIMD form clamps LIMITED rows with _mm_min_ps/_mm_max_ps.
static void rowPreSolve(Row& r, const MassData* masses, PseudoVelocityData* pv, bool simd) {
    float im0 = 0, im1 = 0;
    M3 I0 = mzero(), I1 = mzero();
    if (r.b0 >= 0) { im0 = masses[r.b0].invMass; I0 = masses[r.b0].invI; }
    if (r.b1 >= 0) { im1 = masses[r.b1].invMass; I1 = masses[r.b1].invI; }
    bool ang = r.flags & F_ANGULAR;
    r.angular0t = mulmv(I0, r.angular0);
    r.angular1t = mulmv(I1, r.angular1);
    r.invEff = dot(r.angular0, r.angular0t) + dot(r.angular1, r.angular1t);
    if (!ang) {
        r.invEff = r.invEff + dot(r.linear, r.linear) * (im0 + im1);
        r.linear0t = im0 * r.linear;
        r.linear1t = im1 * r.linear;
    }
    if (r.flags & F_SOFT) return;
    if (r.c == 0.f || r.invEff == 0.f) return;
    float lambda = r.c / r.invEff;
    if (r.flags & F_LIMITED) {
        if (simd) { lambda = lambda > r.mn ? lambda : r.mn; lambda = lambda < r.mx ? lambda : r.mx; }
        else lambda = gclamp(lambda, r.mn, r.mx);
    }
    if (r.b0 >= 0) {
        if (!ang) pv[r.b0].pv = pv[r.b0].pv + lambda * r.linear0t;
        pv[r.b0].pw = pv[r.b0].pw + lambda * r.angular0t;
        ++pv[r.b0].count;
    }
    if (r.b1 >= 0) {
        if (!ang) pv[r.b1].pv = pv[r.b1].pv - lambda * r.linear1t;
        pv[r.b1].pw = pv[r.b1].pw - lambda * r.angular1t;
        ++pv[r.b1].count;
    }
}

// Constraint1DW<flags>::solve for one lane (src/Constraint1DW.cpp:117-233): strict warm-start
// test against (float)1e-4 and 10000, the warm impulse goes into the written-back
// velocities, bias (float)(0.2 / h), LIMITED clamp to [min*h, max*h].
static void rowSolveSimd(Row& r, VelocityData* vel, float h, bool warm) {
    bool ang = r.flags & F_ANGULAR;
    V3 v0 = v3s(0), w0 = v3s(0), v1 = v3s(0), w1 = v3s(0);
    if (r.b0 >= 0) { if (!ang) v0 = vel[r.b0].v; w0 = vel[r.b0].w; }
    if (r.b1 >= 0) { if (!ang) v1 = vel[r.b1].v; w1 = vel[r.b1].w; }
    const V3 v0in = v0, w0in = w0, v1in = v1, w1in = w1;
    if (!(r.flags & F_SOFT) && warm) {
        float ac = std::fabs(r.c), al = std::fabs(r.lambda);
        const bool inMask = g_mutation == MUT_SIMD_MASK_NONSTRICT ? (ac <= 1e-4f && al <= 10000.f)
                                                                  : (ac < 1e-4f && al < 10000.f);
        if (inMask) {
            float lam = r.lambda * 0.5f;
            if (!ang) { v0 = v0 + lam * r.linear0t; v1 = v1 - lam * r.linear1t; }
            w0 = w0 + lam * r.angular0t;
            w1 = w1 - lam * r.angular1t;
            r.lambda = lam;
        }
    }
    if (r.invEff == 0.f) {
        if (g_mutation == MUT_SIMD_ZERO_WRITEBACK) {
            if (r.b0 >= 0) { if (!ang) vel[r.b0].v = v0; vel[r.b0].w = w0; }
            if (r.b1 >= 0) { if (!ang) vel[r.b1].v = v1; vel[r.b1].w = w1; }
        }
        return;
    }
    if (g_mutation == MUT_SIMD_WARM_LOCAL && !(r.flags & (F_SOFT | F_LIMITED))) {
        // the relative velocity still sees the warm impulse, the write-back does not
        const float rvw = dot(r.angular1, w1) - dot(r.angular0, w0) + (ang ? 0.f : dot(r.linear, v1) - dot(r.linear, v0));
        const float lam = (rvw - r.tv + (float)(0.2 / (double)h) * r.c) * (1.f / r.invEff);
        r.lambda = r.lambda + lam;
        if (r.b0 >= 0) { if (!ang) vel[r.b0].v = v0in + lam * r.linear0t; vel[r.b0].w = w0in + lam * r.angular0t; }
        if (r.b1 >= 0) { if (!ang) vel[r.b1].v = v1in - lam * r.linear1t; vel[r.b1].w = w1in - lam * r.angular1t; }
        return;
    }
    float rv = dot(r.angular1, w1) - dot(r.angular0, w0);
    if (!ang) rv = rv + (dot(r.linear, v1) - dot(r.linear, v0));
    float effMass = 1.f / r.invEff;
    float lambda;
    if (r.flags & F_SOFT) {
        float twoPi = 2.f * 3.14159265358979323846f;
        float af = twoPi * r.freq;
        float stiffness = af * af * effMass;
        float damping = 2.f * af * r.damp * effMass;
        float gamma = 1.f / (damping + h * stiffness);
        float beta = h * stiffness * gamma;
        lambda = (rv + beta * r.c / h) / (r.invEff + gamma / h);
    } else {
        float bias = (float)(0.2 / (double)h);
        if (g_mutation == MUT_SIMD_BIAS_FLOAT) lambda = (rv - r.tv + 0.2f * r.c / h) * effMass;
        else lambda = (rv - r.tv + bias * r.c) * effMass;
    }
    float prev = r.lambda;
    if (r.flags & F_LIMITED) {
        float t = r.lambda + lambda;
        float lo = r.mn * h, hi = r.mx * h;
        if (g_mutation == MUT_SIMD_LIMIT_UNSCALED) { lo = r.mn; hi = r.mx; }
        t = t > lo ? t : lo;
        t = t < hi ? t : hi;
        r.lambda = t;
        lambda = r.lambda - prev;
    } else {
        r.lambda = r.lambda + lambda;
    }
    if (!ang) { v0 = v0 + lambda * r.linear0t; v1 = v1 - lambda * r.linear1t; }
    w0 = w0 + lambda * r.angular0t;
    w1 = w1 - lambda * r.angular1t;
    if (r.b0 >= 0) { if (!ang) vel[r.b0].v = v0; vel[r.b0].w = w0; }
    if (r.b1 >= 0) { if (!ang) vel[r.b1].v = v1; vel[r.b1].w = w1; }
}

// Constraint1D<flags>::solve (src/Constraint1D.cpp:54-127): non-strict warm-start test
// (|c| compared with the double 1e-4), the warm impulse only biases local copies, bias
// 0.2f * c / h, LIMITED clamp to [min, max].
static void rowSolveScalar(Row& r, VelocityData* vel, float h, bool warm) {
    if (r.invEff == 0.f) return;
    bool ang = r.flags & F_ANGULAR;
    V3 v0 = v3s(0), w0 = v3s(0), v1 = v3s(0), w1 = v3s(0);
    if (r.b0 >= 0) { if (!ang) v0 = vel[r.b0].v; w0 = vel[r.b0].w; }
    if (r.b1 >= 0) { if (!ang) v1 = vel[r.b1].v; w1 = vel[r.b1].w; }
    if (!(r.flags & F_SOFT)) {
        if (warm && !((double)gabs(r.c) > 1e-4 || gabs(r.lambda) > 10000.f)) {
            r.lambda = r.lambda * 0.5f;
            if (r.b0 >= 0) { if (!ang) v0 = v0 + r.lambda * r.linear0t; w0 = w0 + r.lambda * r.angular0t; }
            if (r.b1 >= 0) { if (!ang) v1 = v1 - r.lambda * r.linear1t; w1 = w1 - r.lambda * r.angular1t; }
        }
    }
    float rv = dot(r.angular1, w1) - dot(r.angular0, w0);
    if (!ang) rv = rv + (dot(r.linear, v1) - dot(r.linear, v0));
    float lambda;
    if (r.flags & F_SOFT) {
        float af = 2.f * 3.14159265358979323846f * r.freq;
        float stiffness = af * af / r.invEff;
        float damping = 2.f * af * r.damp / r.invEff;
        float gamma = 1.f / (damping + h * stiffness);
        float beta = h * stiffness / (damping + h * stiffness);
        lambda = (rv + beta * r.c / h) / (r.invEff + gamma / h);
    } else {
        lambda = (rv - r.tv + 0.2f * r.c / h) / r.invEff;
    }
    if (r.flags & F_LIMITED) {
        float prev = r.lambda;
        r.lambda = r.lambda + lambda;
        r.lambda = gclamp(r.lambda, r.mn, r.mx);
        lambda = r.lambda - prev;
    } else {
        r.lambda = r.lambda + lambda;
    }
    if (r.b0 >= 0) { if (!ang) vel[r.b0].v = vel[r.b0].v + lambda * r.linear0t; vel[r.b0].w = vel[r.b0].w + lambda * r.angular0t; }
    if (r.b1 >= 0) { if (!ang) vel[r.b1].v = vel[r.b1].v - lambda * r.linear1t; vel[r.b1].w = vel[r.b1].w - lambda * r.angular1t; }
}

// Joint rows of one colour in the reference's container order: colours 0-7 are SIMD rows
// of body-disjoint joints (per joint, its rows in flag-type order); colour 8 is the scalar
// overflow container, walked flag type by flag type over all its joints.
template <class F>
static void forColorRows(std::vector<JointRows>& jr, int color, F&& f) {
    if (color < 8) {
        for (auto& j : jr)
            for (int k = 0; k < 6; ++k)
                for (int r = 0; r < j.nRows; ++r)
                    if (j.rows[r].flags == kFlagOrder[k]) f(j.rows[r]);
    } else {
        for (int k = 0; k < 6; ++k)
            for (auto& j : jr)
                for (int r = 0; r < j.nRows; ++r)
                    if (j.rows[r].flags == kFlagOrder[k]) f(j.rows[r]);
    }
}

// src/ContactConstraints.cpp:4-31
static void preSolve(CC& c, const MassData* masses) {
    float im0 = 0; M3 I0 = mzero();
    if (c.b0 >= 0) { im0 = masses[c.b0].invMass; I0 = masses[c.b0].invI; }
    float im1 = 0; M3 I1 = mzero();
    if (c.b1 >= 0) { im1 = masses[c.b1].invMass; I1 = masses[c.b1].invI; }
    for (int i = 0; i < c.numPoints; ++i) {
        PointC& p = c.pts[i];
        p.r0xnt = mulmv(I0, p.r0xn);
        p.r1xnt = mulmv(I1, p.r1xn);
        p.r0xtt = mulmv(I0, p.r0xt);
        p.r1xtt = mulmv(I1, p.r1xt);
        p.invEffN = dot(c.n, c.n) * (im0 + im1) + dot(p.r0xn, p.r0xnt) + dot(p.r1xn, p.r1xnt);
        p.invEffT = dot(p.t, p.t) * (im0 + im1) + dot(p.r0xt, p.r0xtt) + dot(p.r1xt, p.r1xtt);
    }
}

// src/ContactConstraints.cpp:33-124
static void solveContact(orc_world* w, CC& c, VelocityData* vel, bool useBias, float h) {
    const Dyn* dyn0 = c.d0 >= 0 ? &w->dyn[c.d0] : nullptr;
    const Dyn* dyn1 = c.d1 >= 0 ? &w->dyn[c.d1] : nullptr;
    for (int i = 0; i < c.numPoints; ++i) {
        PointC& p = c.pts[i];
        if (p.invEffN == 0.f) continue;
        V3 v0 = v3s(0), w0 = v3s(0), v1 = v3s(0), w1 = v3s(0);
        if (c.b0 >= 0) { v0 = vel[c.b0].v; w0 = vel[c.b0].w; }
        if (c.b1 >= 0) { v1 = vel[c.b1].v; w1 = vel[c.b1].w; }
        float rv = dot(-c.n, v0) + dot(-p.r0xn, w0) + dot(c.n, v1) + dot(p.r1xn, w1);
        float effMass = 1.f / p.invEffN;
        float lambda;
        if (c.isSoft) {
            float af = 2.f * 3.14159265358979323846264338327950288f * c.freq;
            float stiffness = af * af * effMass;
            float damping = 2.f * af * c.damp * effMass;
            float gamma = 1.f / (damping + h * stiffness);
            float beta = h * stiffness / (damping + h * stiffness);
            lambda = (rv + beta * p.c / h) / (p.invEffN + gamma / h);
        } else {
            lambda = (rv - p.tv + (useBias ? 0.1f * p.c / h : 0.f)) * effMass;
        }
        float prev = p.lN;
        p.lN += lambda;
        p.lN = gmin(p.lN, 0.f);
        lambda = p.lN - prev;
        if (c.b0 >= 0) { vel[c.b0].v = vel[c.b0].v + lambda * dyn0->invMass * c.n; vel[c.b0].w = vel[c.b0].w + lambda * p.r0xnt; }
        if (c.b1 >= 0) { vel[c.b1].v = vel[c.b1].v - lambda * dyn1->invMass * c.n; vel[c.b1].w = vel[c.b1].w - lambda * p.r1xnt; }
    }
    for (int i = 0; i < c.numPoints; ++i) {
        PointC& p = c.pts[i];
        if (p.invEffT == 0.f) continue;
        V3 v0 = v3s(0), w0 = v3s(0), v1 = v3s(0), w1 = v3s(0);
        if (c.b0 >= 0) { v0 = dyn0->v; w0 = dyn0->w; }     // component velocities (quirk)
        if (c.b1 >= 0) { v1 = dyn1->v; w1 = dyn1->w; }
        if (g_mutation == MUT_FRICTION_SOLVER_VEL) {
            if (c.b0 >= 0) { v0 = vel[c.b0].v; w0 = vel[c.b0].w; }
            if (c.b1 >= 0) { v1 = vel[c.b1].v; w1 = vel[c.b1].w; }
        }
        float rv = dot(-p.t, v0) + dot(-p.r0xt, w0) + dot(p.t, v1) + dot(p.r1xt, w1);
        float lambda = rv / p.invEffT;
        float limit = c.friction * p.lN;
        float prev = p.lT;
        p.lT += lambda;
        p.lT = gclamp(p.lT, limit, -limit);
        lambda = p.lT - prev;
        if (c.b0 >= 0) { vel[c.b0].v = vel[c.b0].v + lambda * dyn0->invMass * p.t; vel[c.b0].w = vel[c.b0].w + lambda * p.r0xtt; }
        if (c.b1 >= 0) { vel[c.b1].v = vel[c.b1].v - lambda * dyn1->invMass * p.t; vel[c.b1].w = vel[c.b1].w - lambda * p.r1xtt; }
    }
}

// src/MathUtil.h:10-21
static V3 solve33(const M3& A, V3 b) {
    V3 c12 = cross(A[1], A[2]);
    float det = dot(A[0], c12);
    if (det == 0.f) return v3s(0);
    float id = 1.f / det;
    return v3(id * dot(b, c12), id * dot(A[0], cross(b, A[2])), id * dot(A[0], cross(A[1], b)));
}


// src/Physecs.cpp:450-467 for one body
static void integrateVelocity(Dyn& d, Q quat, float h, float g, MassData& md) {
    M3 rot = mat3_cast(quat);
    M3 invRot = transpose(rot);
    d.v = d.v + h * v3(0, -g, 0);
    V3 wl = mulmv(invRot, d.w);
    M3 I = inverse(d.invI);
    V3 f = h * cross(wl, mulmv(I, wl));
    M3 J = madd(I, mscale(h, msub(mulmm(matrixCross3(wl), I), matrixCross3(mulmv(I, wl)))));
    wl = wl - solve33(J, f);
    d.w = mulmv(rot, wl);
    md = {d.invMass, mulmm(mulmm(rot, d.invI), invRot)};
}

// src/Physecs.cpp:501-510 for one body
static void integratePosition(V3& pos, Q& quat, const VelocityData& vel, const PseudoVelocityData& pvel, V3 com, float h) {
    float s = pvel.count ? 1.f / pvel.count : 1.f;
    pos = pos + (h * vel.v + s * pvel.pv);
    V3 prevCom = qrot(quat, com);
    V3 hv = 0.5f * (h * vel.w + s * pvel.pw);
    quat = qadd(quat, qmul(qmake(0.f, hv.x, hv.y, hv.z), quat));
    quat = qnormalize(quat);
    pos = pos + (prevCom - qrot(quat, com));
}

static void substeps(orc_world* w, float timeStep) {
    size_t nd = w->dyn.size();
    std::vector<VelocityData> vel(nd);
    std::vector<PseudoVelocityData> pvel(nd);
    std::vector<MassData> mass(nd);
    float h = timeStep / w->numSubSteps;
    for (int m = 0; m < w->numSubSteps; ++m) {
        // update contact constraints, src/Physecs.cpp:368-426
        for (auto& c : w->ccs) {
            const Actor& t0 = w->actors[c.a0];
            const Actor& t1 = w->actors[c.a1];
            const Dyn* dyn0 = c.d0 >= 0 ? &w->dyn[c.d0] : nullptr;
            const Dyn* dyn1 = c.d1 >= 0 ? &w->dyn[c.d1] : nullptr;
            V3 com0 = v3s(0), v0 = v3s(0), w0 = v3s(0);
            if (dyn0 && !dyn0->kinematic) { com0 = t0.pos + qrot(t0.quat, dyn0->com); v0 = dyn0->v; w0 = dyn0->w; }
            V3 com1 = v3s(0), v1 = v3s(0), w1 = v3s(0);
            if (dyn1 && !dyn1->kinematic) { com1 = t1.pos + qrot(t1.quat, dyn1->com); v1 = dyn1->v; w1 = dyn1->w; }
            for (int k = 0; k < c.numPoints; ++k) {
                PointC& p = c.pts[k];
                V3 r0 = qrot(t0.quat, p.r0);
                V3 r1 = qrot(t1.quat, p.r1);
                V3 cp0 = com0 + r0;
                V3 cp1 = com1 + r1;
                float cn = dot(cp1 - cp0, c.n);
                V3 r0xn = cross(r0, c.n);
                V3 r1xn = cross(r1, c.n);
                V3 rel = v1 + cross(w1, r1) - v0 - cross(w0, r0);
                float relN = dot(rel, c.n);
                V3 t = rel - relN * c.n;
                float tl = length(t);
                if (tl != 0.f) t = t / tl;
                p.r0xn = r0xn; p.r1xn = r1xn; p.t = t;
                p.r0xt = cross(r0, t); p.r1xt = cross(r1, t);
                p.c = cn; p.lN = 0; p.lT = 0;
            }
        }
        // update joint constraints, src/Physecs.cpp:430-439
        for (auto& col : w->jrows)
            for (auto& jr : col) updateJointRows(w, jr);
        // integrate velocities, src/Physecs.cpp:443-468
        for (size_t i = 0; i < nd; ++i) {
            Dyn& d = w->dyn[i];
            if (d.kinematic) continue;
            integrateVelocity(d, w->actors[d.actor].quat, h, w->g, mass[i]);
            vel[i] = {d.v, d.w};
            pvel[i] = {v3s(0), v3s(0), 0};
        }
        for (auto& c : w->ccs) preSolve(c, mass.data());
        for (int color = 0; color < 9; ++color)
            forColorRows(w->jrows[color], color, [&](Row& r) { rowPreSolve(r, mass.data(), pvel.data(), color < 8); });
        for (int it = 0; it < w->numIterations; ++it) {
            for (auto& c : w->ccs) solveContact(w, c, vel.data(), true, h);
            for (int color = 0; color < 9; ++color)
                forColorRows(w->jrows[color], color, [&](Row& r) {
                    if (color < 8) rowSolveSimd(r, vel.data(), h, it == 0);
                    else rowSolveScalar(r, vel.data(), h, it == 0);
                });
        }
        // integrate positions, src/Physecs.cpp:494-511
        for (size_t i = 0; i < nd; ++i) {
            Dyn& d = w->dyn[i];
            if (d.kinematic) continue;
            Actor& tr = w->actors[d.actor];
            integratePosition(tr.pos, tr.quat, vel[i], pvel[i], d.com, h);
        }
        // relaxation, src/Physecs.cpp:516-519
        for (auto& c : w->ccs) {
            if (c.isSoft && g_mutation != MUT_RELAX_SOFT) continue;
            solveContact(w, c, vel.data(), false, c.isSoft ? h : 0.f);
        }
        // write back, src/Physecs.cpp:523-530
        for (size_t i = 0; i < nd; ++i) {
            Dyn& d = w->dyn[i];
            if (d.kinematic) continue;
            d.v = vel[i].v; d.w = vel[i].w;
        }
    }
}

extern "C" {

// physecs::Scene::simulate restated: SAP, narrowphase, (injected order), substeps,
// cache swap, bounds update with the 0.01 margin for non-kinematic dynamics.
void orc_world_step(orc_world* w, float timeStep) {
    sap(w);
    narrowphase(w);
    applyOrder(w);
    buildJointRows(w);
    substeps(w, timeStep);
    // src/Physecs.cpp:538-552: enter = new pairs, exit = pairs no longer overlapping
    w->trigEnter.clear();
    w->trigExit.clear();
    for (const auto& p : w->trigTemp)
        if (!w->trig.count(p)) w->trigEnter.push_back(p);
    for (const auto& p : w->trig)
        if (!w->trigTemp.count(p)) w->trigExit.push_back(p);
    w->trig.swap(w->trigTemp);
    w->cache.swap(w->cacheTemp);
    for (auto& e : w->bp) {
        const Collider& c = w->cols[e.collider];
        const Actor& a = w->actors[c.actor];
        if (a.dyn < 0 || w->dyn[a.dyn].kinematic) continue;
        Bounds b = getBounds(a.pos + qrot(a.quat, c.lpos), qmul(a.quat, c.lquat), c.geom);
        b.max = b.max + v3s(0.01f);
        b.min = b.min - v3s(0.01f);
        e.bounds = b;
    }
}

void orc_world_get_state(const orc_world* w, float* pos, float* quat, float* vel, float* angvel) {
    for (size_t i = 0; i < w->actors.size(); ++i) {
        const Actor& a = w->actors[i];
        pos[3 * i] = a.pos.x; pos[3 * i + 1] = a.pos.y; pos[3 * i + 2] = a.pos.z;
        quat[4 * i] = a.quat.x; quat[4 * i + 1] = a.quat.y; quat[4 * i + 2] = a.quat.z; quat[4 * i + 3] = a.quat.w;
    }
    for (size_t i = 0; i < w->dyn.size(); ++i) {
        const Dyn& d = w->dyn[i];
        vel[3 * i] = d.v.x; vel[3 * i + 1] = d.v.y; vel[3 * i + 2] = d.v.z;
        angvel[3 * i] = d.w.x; angvel[3 * i + 1] = d.w.y; angvel[3 * i + 2] = d.w.z;
    }
}

// bounds of every collider, by collider index: min xyz, max xyz
void orc_world_get_bounds(const orc_world* w, float* out) {
    for (auto& e : w->bp) {
        float* o = out + 6 * e.collider;
        o[0] = e.bounds.min.x; o[1] = e.bounds.min.y; o[2] = e.bounds.min.z;
        o[3] = e.bounds.max.x; o[4] = e.bounds.max.y; o[5] = e.bounds.max.z;
    }
}

int orc_world_num_pairs(const orc_world* w) { return (int)w->pairs.size(); }
// pairs of the last step as collider indices (g0 = collider of entity0)
void orc_world_get_pairs(const orc_world* w, int32_t* out) {
    for (size_t i = 0; i < w->pairs.size(); ++i) { out[2 * i] = w->pairs[i].g0; out[2 * i + 1] = w->pairs[i].g1; }
}

int orc_world_num_manifolds(const orc_world* w) { return (int)w->ccs.size(); }
// per manifold in solve order: keys[5]; ints[4] = b0, b1, numPoints, isSoft;
// floats[4 + 4*6 + 4*7] = normal(3), friction, raw points 4 x (p0, p1), rows 4 x (r0l, r1l, tv)
void orc_world_get_manifolds(const orc_world* w, int32_t* keys, int32_t* ints, float* f) {
    for (size_t i = 0; i < w->ccs.size(); ++i) {
        const CC& c = w->ccs[i];
        for (int k = 0; k < 5; ++k) keys[5 * i + k] = c.key[k];
        ints[4 * i] = c.b0; ints[4 * i + 1] = c.b1; ints[4 * i + 2] = c.numPoints; ints[4 * i + 3] = c.isSoft;
        float* o = f + 56 * i;
        o[0] = c.n.x; o[1] = c.n.y; o[2] = c.n.z; o[3] = c.friction;
        for (int k = 0; k < 4; ++k) {
            const ContactPoint& p = c.raw.points[k];
            float* q = o + 4 + 6 * k;
            bool on = k < c.numPoints;
            q[0] = on ? p.p0.x : 0; q[1] = on ? p.p0.y : 0; q[2] = on ? p.p0.z : 0;
            q[3] = on ? p.p1.x : 0; q[4] = on ? p.p1.y : 0; q[5] = on ? p.p1.z : 0;
            const PointC& r = c.pts[k];
            float* s = o + 28 + 7 * k;
            s[0] = on ? r.r0.x : 0; s[1] = on ? r.r0.y : 0; s[2] = on ? r.r0.z : 0;
            s[3] = on ? r.r1.x : 0; s[4] = on ? r.r1.y : 0; s[5] = on ? r.r1.z : 0;
            s[6] = on ? r.tv : 0;
        }
    }
}


// Single-manifold and single-body entry points with the same layouts as the reference
// harness (oracle/ref_harness.cpp), used to pin this restatement against the reference.
int orc_contact_solve(int b0, int b1, int numPoints, const float* in_f, const float* masses,
                      const float* vel, const float* comp, float h, int iters, int relax, float* out) {
    orc_world w;
    w.dyn.resize(2);
    for (int k = 0; k < 2; ++k) {
        w.dyn[k].invMass = masses[10 * k];
        w.dyn[k].v = v3(comp[6 * k], comp[6 * k + 1], comp[6 * k + 2]);
        w.dyn[k].w = v3(comp[6 * k + 3], comp[6 * k + 4], comp[6 * k + 5]);
        w.dyn[k].kinematic = false;
    }
    CC c{};
    c.d0 = 0; c.d1 = 1;
    c.b0 = b0 >= 0 ? 0 : -1; c.b1 = b1 >= 0 ? 1 : -1;
    c.n = v3(in_f[0], in_f[1], in_f[2]); c.friction = in_f[3]; c.isSoft = in_f[4] != 0.f;
    c.freq = in_f[5]; c.damp = in_f[6]; c.numPoints = numPoints;
    for (int k = 0; k < 4; ++k) {
        const float* p = in_f + 7 + 17 * k;
        PointC& pc = c.pts[k];
        pc.r0xn = v3(p[0], p[1], p[2]); pc.r1xn = v3(p[3], p[4], p[5]); pc.t = v3(p[6], p[7], p[8]);
        pc.r0xt = v3(p[9], p[10], p[11]); pc.r1xt = v3(p[12], p[13], p[14]);
        pc.c = p[15]; pc.tv = p[16];
    }
    MassData md[2];
    VelocityData vd[2];
    for (int k = 0; k < 2; ++k) {
        md[k].invMass = masses[10 * k];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) md[k].invI[i][j] = masses[10 * k + 1 + 3 * i + j];
        vd[k].v = v3(vel[6 * k], vel[6 * k + 1], vel[6 * k + 2]);
        vd[k].w = v3(vel[6 * k + 3], vel[6 * k + 4], vel[6 * k + 5]);
    }
    preSolve(c, md);
    for (int i = 0; i < iters; ++i) solveContact(&w, c, vd, true, h);
    for (int i = 0; i < relax; ++i) solveContact(&w, c, vd, false, 0.f);
    for (int k = 0; k < 2; ++k) {
        out[6 * k] = vd[k].v.x; out[6 * k + 1] = vd[k].v.y; out[6 * k + 2] = vd[k].v.z;
        out[6 * k + 3] = vd[k].w.x; out[6 * k + 4] = vd[k].w.y; out[6 * k + 5] = vd[k].w.z;
    }
    for (int k = 0; k < 4; ++k) { out[12 + 2 * k] = c.pts[k].lN; out[13 + 2 * k] = c.pts[k].lT; }
    return 0;
}

int orc_integrate_velocity(const float* quat, const float* v, const float* w, const float* invI, float h, float g, float* out) {
    Dyn d;
    d.v = v3(v[0], v[1], v[2]); d.w = v3(w[0], w[1], w[2]); d.invMass = 1.f;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) d.invI[i][j] = invI[3 * i + j];
    MassData md;
    integrateVelocity(d, {quat[0], quat[1], quat[2], quat[3]}, h, g, md);
    out[0] = d.v.x; out[1] = d.v.y; out[2] = d.v.z; out[3] = d.w.x; out[4] = d.w.y; out[5] = d.w.z;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) out[6 + 3 * i + j] = md.invI[i][j];
    return 0;
}

int orc_integrate_position(const float* pos, const float* quat, const float* v, const float* w,
                           const float* pv, const float* pw, int count, const float* com, float h, float* out) {
    V3 p = v3(pos[0], pos[1], pos[2]);
    Q q = {quat[0], quat[1], quat[2], quat[3]};
    VelocityData vd{v3(v[0], v[1], v[2]), v3(w[0], w[1], w[2])};
    PseudoVelocityData pd{v3(pv[0], pv[1], pv[2]), v3(pw[0], pw[1], pw[2]), count};
    integratePosition(p, q, vd, pd, v3(com[0], com[1], com[2]), h);
    out[0] = p.x; out[1] = p.y; out[2] = p.z; out[3] = q.x; out[4] = q.y; out[5] = q.z; out[6] = q.w;
    return 0;
}

// convex meshes for the single-call pin entry points below (params[3] = index)
static std::vector<ConvexMesh*> g_convex;

int orc_set_convex_mesh(int n_verts, const float* verts, int n_faces, const int32_t* face_start,
                        const int32_t* face_idx, const float* normals, const float* centroids) {
    g_convex.push_back(new ConvexMesh(makeConvex(n_verts, verts, n_faces, face_start, face_idx, normals, centroids)));
    return (int)g_convex.size() - 1;
}

static std::vector<TriMesh*> g_trimesh;

// builds a TriangleMesh for the pin entry points; returns its index (params[3])
int orc_set_tri_mesh(int n_verts, const float* verts, int n_tris, const uint32_t* idx) {
    std::vector<V3> v;
    for (int i = 0; i < n_verts; ++i) v.push_back(v3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]));
    g_trimesh.push_back(new TriMesh(buildTriMesh(v, std::vector<uint32_t>(idx, idx + 3 * n_tris))));
    return (int)g_trimesh.size() - 1;
}

// the built mesh: triangle vertex indices after the BVH permutation (3 per triangle) and
// the BVH nodes {min xyz, max xyz, triCount, index}
int orc_tri_mesh_layout(int mesh, uint32_t* triIdx, float* nodes) {
    const TriMesh& m = *g_trimesh.at((size_t)mesh);
    for (size_t t = 0; t < m.tris.size(); ++t)
        for (int k = 0; k < 3; ++k) triIdx[3 * t + k] = m.tris[t].idx[k];
    for (size_t n = 0; n < m.bvh.size(); ++n) {
        const TriNode& b = m.bvh[n];
        float* o = nodes + 8 * n;
        o[0] = b.bounds.min.x; o[1] = b.bounds.min.y; o[2] = b.bounds.min.z;
        o[3] = b.bounds.max.x; o[4] = b.bounds.max.y; o[5] = b.bounds.max.z;
        o[6] = (float)b.triCount; o[7] = (float)b.index;
    }
    return (int)m.bvh.size();
}

static Geom pinGeom(int type, const float* prm) {
    Geom g{type, {prm[0], prm[1], prm[2], prm[3]}};
    if (type == CONVEX_MESH) g.mesh = g_convex.at((size_t)prm[3]);
    if (type == TRIANGLE_MESH) g.tmesh = g_trimesh.at((size_t)prm[3]);
    return g;
}

// every manifold of one collision() call: returns the count; per manifold 30 floats
// {numPoints, normal(3), 4 x (p0, p1), triangleIndex, hit}
int orc_collision_multi(int type0, const float* prm0, const float* pos0, const float* q0, int type1,
                        const float* prm1, const float* pos1, const float* q1, int cap, float* out) {
    Geom g0 = pinGeom(type0, prm0), g1 = pinGeom(type1, prm1);
    std::vector<Manifold> buf;
    bool hit = collision(v3(pos0[0], pos0[1], pos0[2]), {q0[0], q0[1], q0[2], q0[3]}, g0,
                         v3(pos1[0], pos1[1], pos1[2]), {q1[0], q1[1], q1[2], q1[3]}, g1, buf);
    int n = 0;
    for (const Manifold& m : buf) {
        if (n >= cap) break;
        float* o = out + 30 * n++;
        std::memset(o, 0, sizeof(float) * 30);
        o[0] = (float)m.numPoints;
        o[1] = m.normal.x; o[2] = m.normal.y; o[3] = m.normal.z;
        for (int k = 0; k < m.numPoints && k < 4; ++k) {
            o[4 + 6 * k] = m.points[k].p0.x; o[5 + 6 * k] = m.points[k].p0.y; o[6 + 6 * k] = m.points[k].p0.z;
            o[7 + 6 * k] = m.points[k].p1.x; o[8 + 6 * k] = m.points[k].p1.y; o[9 + 6 * k] = m.points[k].p1.z;
        }
        o[28] = (float)m.triangleIndex;
        o[29] = hit ? 1.f : 0.f;
    }
    return hit ? n : -1;
}

long orc_epa_exhausted() { return epaExhausted; }

// single narrowphase call (for pinning against the reference).
// out: numPoints, normal(3), 4 x (p0, p1); returns the collision() bool.
int orc_collision(int type0, const float* prm0, const float* pos0, const float* q0,
                  int type1, const float* prm1, const float* pos1, const float* q1, float* out) {
    Geom g0 = pinGeom(type0, prm0);
    Geom g1 = pinGeom(type1, prm1);
    std::vector<Manifold> buf;
    bool hit = collision(v3(pos0[0], pos0[1], pos0[2]), {q0[0], q0[1], q0[2], q0[3]}, g0,
                         v3(pos1[0], pos1[1], pos1[2]), {q1[0], q1[1], q1[2], q1[3]}, g1, buf);
    std::memset(out, 0, sizeof(float) * 28);
    const Manifold& m = buf.back();
    out[0] = (float)m.numPoints;
    out[1] = m.normal.x; out[2] = m.normal.y; out[3] = m.normal.z;
    for (int k = 0; k < m.numPoints && k < 4; ++k) {
        float* o = out + 4 + 6 * k;
        o[0] = m.points[k].p0.x; o[1] = m.points[k].p0.y; o[2] = m.points[k].p0.z;
        o[3] = m.points[k].p1.x; o[4] = m.points[k].p1.y; o[5] = m.points[k].p1.z;
    }
    return hit ? 1 : 0;
}

int orc_bounds(int type, const float* prm, const float* pos, const float* q, float* out) {
    Geom g = pinGeom(type, prm);
    Bounds b = getBounds(v3(pos[0], pos[1], pos[2]), {q[0], q[1], q[2], q[3]}, g);
    out[0] = b.min.x; out[1] = b.min.y; out[2] = b.min.z;
    out[3] = b.max.x; out[4] = b.max.y; out[5] = b.max.z;
    return 0;
}

// out: sqrDist, t, q(3)
int orc_sqr_dist_segment_aabb(const float* p, const float* dir, float mn, float mx, const float* he, float* out) {
    V3 qq = v3s(0);
    float t = 0;
    float sq = sqrDistSegmentAABB(v3(p[0], p[1], p[2]), v3(dir[0], dir[1], dir[2]), mn, mx, v3(he[0], he[1], he[2]), t, qq);
    out[0] = sq; out[1] = t; out[2] = qq.x; out[3] = qq.y; out[4] = qq.z;
    return 0;
}

// Scene::createJoint<T> for the built-in types (include/Physecs/Physecs.h:209-214); the
// type-specific fields are the joints' data members, set through their setters
struct orc_joint {
    int32_t type, actor0, actor1;
    float anchor0[7];    // position xyz, orientation xyzw
    float anchor1[7];
    int32_t drive_enabled;
    float drive_velocity, drive_max_torque;
    float lower_limit, upper_limit, target, drive_stiffness, drive_damping, gear_ratio;
};

static void setParams(Joint& j, const orc_joint& s) {
    j.driveEnabled = s.drive_enabled; j.driveVelocity = s.drive_velocity; j.driveMaxTorque = s.drive_max_torque;
    j.lowerLimit = s.lower_limit; j.upperLimit = s.upper_limit; j.target = s.target;
    j.stiffness = s.drive_stiffness; j.damping = s.drive_damping; j.gearRatio = s.gear_ratio;
}

int orc_world_add_joints(orc_world* w, int n, const orc_joint* js) {
    for (int i = 0; i < n; ++i) {
        const orc_joint& s = js[i];
        Joint j{};
        j.type = s.type; j.a0 = s.actor0; j.a1 = s.actor1;
        j.p0 = v3(s.anchor0[0], s.anchor0[1], s.anchor0[2]);
        j.q0 = {s.anchor0[3], s.anchor0[4], s.anchor0[5], s.anchor0[6]};
        j.p1 = v3(s.anchor1[0], s.anchor1[1], s.anchor1[2]);
        j.q1 = {s.anchor1[3], s.anchor1[4], s.anchor1[5], s.anchor1[6]};
        setParams(j, s);
        j.alive = true;
        addJoint(w, j);
    }
    return 0;
}

// The joints' setters (PrismaticJoint::setUpperLimit, ServoJoint::setTargetAngle,
// GearJoint::setGearRatio, RevoluteJoint::setDriveEnabled, ...): take effect next step.
int orc_world_set_joint_params(orc_world* w, int i, const orc_joint* s) {
    if (i < 0 || i >= (int)w->joints.size() || !w->joints[i].alive) return -1;
    setParams(w->joints[i], *s);
    return 0;
}

// Scene::destroyJoint (src/Physecs.cpp:712-725): out of its colour, the entity pair may
// collide again, and the colour bit is freed on both entities.
int orc_world_destroy_joint(orc_world* w, int i) {
    if (i < 0 || i >= (int)w->joints.size() || !w->joints[i].alive) return -1;
    Joint& j = w->joints[i];
    j.alive = false;
    uint32_t e0 = w->actors[j.a0].entity, e1 = w->actors[j.a1].entity;
    uint32_t lo = e0 < e1 ? e0 : e1, hi = e0 < e1 ? e1 : e0;
    w->nonColliding.erase(((uint64_t)lo << 32) | hi);
    w->jointBitsets[e0] = (unsigned char)(w->jointBitsets[e0] & ~(1u << j.color));
    w->jointBitsets[e1] = (unsigned char)(w->jointBitsets[e1] & ~(1u << j.color));
    return 0;
}

int orc_world_joint_color(const orc_world* w, int i) { return w->joints[i].color; }

// the host functions of the world's custom joints (J_CUSTOM)
void orc_world_set_joint_callbacks(orc_world* w, CustomLayoutFn layout, CustomRowsFn rows, void* user) {
    w->custLayout = layout;
    w->custRows = rows;
    w->custUser = user;
}
// 1 when a custom joint's layout was refused since the last call (then it had no rows)
int orc_world_joint_error(orc_world* w) {
    const int e = w->jointError;
    w->jointError = 0;
    return e;
}

// Plants one of the Mutation mistakes in every later call (0 restores the restatement).
void orc_set_mutation(int m) { g_mutation = m; }
int orc_get_mutation(void) { return g_mutation; }

// The rows joint i would get at the world's current poses: buildJointRows's layout for that
// joint, then updateJointRows (the same two steps a step runs). Per row, 18 floats: flags, b0,
// b1, linear(3), angular0(3), angular1(3), c, targetVelocity, min, max, frequency, damping.
// The joint's own state (gear angles, prismatic limit choice) is left as it was. Returns the
// row count, or -1 for a destroyed or unknown joint.
int orc_world_joint_rows(orc_world* w, int i, float* out) {
    if (i < 0 || i >= (int)w->joints.size() || !w->joints[i].alive) return -1;
    const Joint saved = w->joints[i];
    Joint& j = w->joints[i];
    JointRows jr{};
    jr.joint = i;
    const Actor& t0 = w->actors[j.a0];
    const Actor& t1 = w->actors[j.a1];
    const Dyn* d0 = t0.dyn >= 0 ? &w->dyn[t0.dyn] : nullptr;
    const Dyn* d1 = t1.dyn >= 0 ? &w->dyn[t1.dyn] : nullptr;
    int b0 = d0 && !d0->kinematic ? t0.dyn : -1;
    int b1 = d1 && !d1->kinematic ? t1.dyn : -1;
    jr.r0 = b0 >= 0 && g_mutation != MUT_JOINT_NO_COM ? j.p0 - d0->com : j.p0;
    jr.r1 = b1 >= 0 && g_mutation != MUT_JOINT_NO_COM ? j.p1 - d1->com : j.p1;
    int fl[7];
    jr.nRows = rowLayout(w, j, fl);
    for (int r = 0; r < jr.nRows; ++r) {
        Row& row = jr.rows[r];
        std::memset(&row, 0, sizeof(row));
        row.flags = fl[r];
        row.b0 = b0; row.b1 = b1;
        row.mn = std::numeric_limits<float>::lowest();
        row.mx = std::numeric_limits<float>::max();
    }
    updateJointRows(w, jr);
    for (int r = 0; r < jr.nRows; ++r) {
        const Row& row = jr.rows[r];
        float* o = out + 18 * r;
        o[0] = (float)row.flags; o[1] = (float)row.b0; o[2] = (float)row.b1;
        o[3] = row.linear.x; o[4] = row.linear.y; o[5] = row.linear.z;
        o[6] = row.angular0.x; o[7] = row.angular0.y; o[8] = row.angular0.z;
        o[9] = row.angular1.x; o[10] = row.angular1.y; o[11] = row.angular1.z;
        o[12] = row.c; o[13] = row.tv; o[14] = row.mn; o[15] = row.mx; o[16] = row.freq; o[17] = row.damp;
    }
    w->joints[i] = saved;
    return jr.nRows;
}

// One joint row through this restatement, same layout as ref_row_scalar (oracle/ref_harness.cpp);
// simd != 0 runs the per-lane Constraint1DW semantics instead of Constraint1D's.
int orc_row_ex(int simd, int flags, int b0, int b1, const float* in, const float* masses, float* vel, float* pv,
               float h, int iters, float lambda0, int warmMask, float* out);

int orc_row(int simd, int flags, int b0, int b1, const float* in, const float* masses, float* vel, float* pv, float h,
            int iters, float* out) {
    return orc_row_ex(simd, flags, b0, b1, in, masses, vel, pv, h, iters, 0.f, 1, out);
}

// orc_row with a starting accumulated lambda (it carries over between the substeps of a step)
// and a per-solve warm-start mask (bit i: warm start on solve i)
int orc_row_ex(int simd, int flags, int b0, int b1, const float* in, const float* masses, float* vel, float* pv,
               float h, int iters, float lambda0, int warmMask, float* out) {
    Row r{};
    r.flags = flags; r.b0 = b0 >= 0 ? 0 : -1; r.b1 = b1 >= 0 ? 1 : -1;
    r.linear = v3(in[0], in[1], in[2]); r.angular0 = v3(in[3], in[4], in[5]); r.angular1 = v3(in[6], in[7], in[8]);
    r.tv = in[9]; r.c = in[10]; r.mn = in[11]; r.mx = in[12]; r.freq = in[13]; r.damp = in[14];
    MassData md[2];
    VelocityData vd[2];
    PseudoVelocityData pd[2];
    for (int k = 0; k < 2; ++k) {
        md[k].invMass = masses[10 * k];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) md[k].invI[i][j] = masses[10 * k + 1 + 3 * i + j];
        vd[k].v = v3(vel[6 * k], vel[6 * k + 1], vel[6 * k + 2]);
        vd[k].w = v3(vel[6 * k + 3], vel[6 * k + 4], vel[6 * k + 5]);
        pd[k].pv = v3(pv[7 * k], pv[7 * k + 1], pv[7 * k + 2]);
        pd[k].pw = v3(pv[7 * k + 3], pv[7 * k + 4], pv[7 * k + 5]);
        pd[k].count = (int)pv[7 * k + 6];
    }
    rowPreSolve(r, md, pd, simd != 0);
    r.lambda = lambda0;
    for (int i = 0; i < iters; ++i) {
        const bool warm = ((unsigned)warmMask >> i) & 1u;
        if (simd) rowSolveSimd(r, vd, h, warm);
        else rowSolveScalar(r, vd, h, warm);
    }
    for (int k = 0; k < 2; ++k) {
        vel[6 * k] = vd[k].v.x; vel[6 * k + 1] = vd[k].v.y; vel[6 * k + 2] = vd[k].v.z;
        vel[6 * k + 3] = vd[k].w.x; vel[6 * k + 4] = vd[k].w.y; vel[6 * k + 5] = vd[k].w.z;
        pv[7 * k] = pd[k].pv.x; pv[7 * k + 1] = pd[k].pv.y; pv[7 * k + 2] = pd[k].pv.z;
        pv[7 * k + 3] = pd[k].pw.x; pv[7 * k + 4] = pd[k].pw.y; pv[7 * k + 5] = pd[k].pw.z; pv[7 * k + 6] = (float)pd[k].count;
    }
    out[0] = r.invEff;
    out[1] = r.lambda;
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ scene raycast
namespace {

// intersectRayAABB (src/Raycast.cpp:5-30)
bool rayAABB(V3 o, V3 d, V3 mn, V3 mx, float& t) {
    float tMin = 0.f;
    float tMax = std::numeric_limits<float>::max();
    const float ov[3] = {o.x, o.y, o.z}, dv[3] = {d.x, d.y, d.z}, lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
    for (int i = 0; i < 3; ++i) {
        if (gabs(dv[i]) < 0.001f) {
            if (ov[i] < lo[i] || ov[i] > hi[i]) return false;   // parallel to the slab, outside it
        } else {
            float ood = 1.f / dv[i];
            float t1 = (lo[i] - ov[i]) * ood;
            float t2 = (hi[i] - ov[i]) * ood;
            if (t1 > t2) std::swap(t1, t2);
            if (t1 > tMin) tMin = t1;
            if (t2 < tMax) tMax = t2;
            if (tMin > tMax) return false;
        }
    }
    t = tMin;
    return true;
}

// intersectRaySphere (src/Raycast.cpp:32-46)
bool raySphere(V3 o, V3 d, V3 pos, float radius, float& t) {
    V3 m = o - pos;
    float b = dot(m, d);
    float c = dot(m, m) - radius * radius;
    if (c > 0.f && b > 0.f) return false;
    float discr = b * b - c;
    if (discr < 0.f) return false;
    t = -b - std::sqrt(discr);
    if (t < 0.f) t = 0.f;
    return true;
}

// intersectRayCapsule (src/Raycast.cpp:48-103): top hemisphere, cylinder, bottom hemisphere
bool rayCapsule(V3 o, V3 d0, V3 pos, Q ori, float hh, float radius, float& t) {
    Q inv = qinverse(ori);
    V3 p = qrot(inv, o - pos);
    V3 d = qrot(inv, d0);
    float pd = dot(p, d);
    float hdy = hh * d.y;
    float thpy = 2.f * hh * p.y;
    float phr = dot(p, p) + hh * hh - radius * radius;
    float tMin = std::numeric_limits<float>::max();
    bool hit = false;
    float b = pd - hdy, c = phr - thpy, discr = b * b - c;
    if (discr >= 0.f) {
        t = -b - std::sqrt(discr);
        if (t < 0.f) t = 0.f;
        if ((p + t * d).y >= hh) { hit = true; tMin = t; }
    }
    float a = d.x * d.x + d.z * d.z;
    b = p.x * d.x + p.z * d.z;
    c = p.x * p.x + p.z * p.z - radius * radius;
    discr = b * b - a * c;
    if (discr >= 0.f) {
        t = (-b - std::sqrt(discr)) / a;
        if (t < 0.f) t = 0.f;
        float y = (p + t * d).y;
        if (-hh <= y && y < hh && t < tMin) { hit = true; tMin = t; }
    }
    b = pd + hdy;
    c = phr + thpy;
    discr = b * b - c;
    if (discr >= 0.f) {
        t = -b - std::sqrt(discr);
        if (t < 0.f) t = 0.f;
        if ((p + t * d).y < -hh && t < tMin) { hit = true; tMin = t; }
    }
    t = tMin;
    return hit;
}

// intersectRayBox (src/Raycast.cpp:105-110)
bool rayBox(V3 o, V3 d, V3 pos, Q ori, V3 he, float& t) {
    Q inv = qinverse(ori);
    return rayAABB(qrot(inv, o - pos), qrot(inv, d), -he, he, t);
}

// intersectRayConvexMesh (src/Raycast.cpp:112-139): the faces' half-spaces, in face order
bool rayConvex(V3 o, V3 d, V3 pos, Q ori, const ConvexMesh* mesh, V3 scale, float& t) {
    float tMin = 0.f;
    float tMax = std::numeric_limits<float>::max();
    for (const ConvexFace& f : mesh->faces) {
        V3 planeOrig = pos + qrot(ori, scale * f.centroid);
        V3 planeNormal = qrot(ori, (v3s(1.f) / scale) * f.normal);
        float denom = dot(d, planeNormal);
        float dist = dot(planeOrig - o, planeNormal);
        if (denom == 0.f) {
            if (dist > 0) return false;
        } else {
            t = dist / denom;
            if (denom < 0) { if (t > tMin) tMin = t; }
            else { if (t < tMax) tMax = t; }
            if (tMin > tMax) return false;
        }
    }
    t = tMin;
    return true;
}

// intersectRayGeometry (src/Raycast.cpp:141-150); triangle meshes are never hit
bool rayGeometry(V3 o, V3 d, V3 pos, Q ori, const Geom& g, float& t) {
    switch (g.type) {
        case SPHERE: return raySphere(o, d, pos, g.p[0], t);
        case CAPSULE: return rayCapsule(o, d, pos, ori, g.p[0], g.p[1], t);
        case BOX: return rayBox(o, d, pos, ori, v3(g.p[0], g.p[1], g.p[2]), t);
        case CONVEX_MESH: return rayConvex(o, d, pos, ori, g.mesh, v3(g.p[0], g.p[1], g.p[2]), t);
        default: return false;
    }
}

// src/Overlap.cpp:19-24
bool ovSphereSphere(V3 pos0, float r0, V3 pos1, float r1) {
    V3 d = pos1 - pos0;
    float rs = r0 + r1;
    return !(dot(d, d) > rs * rs);
}

// src/Overlap.cpp:26-35
bool ovCapsuleCapsule(V3 pos0, Q or0, float hh0, float r0, V3 pos1, Q or1, float hh1, float r1) {
    V3 up = v3(0, 1, 0);
    V3 v0 = qrot(or0, up), v1 = qrot(or1, up);
    V3 a, b;
    closestPointsBetweenSegments(pos0, v0, -hh0, hh0, pos1, v1, -hh1, hh1, a, b);
    V3 d = b - a;
    float rs = r0 + r1;
    return !(dot(d, d) > rs * rs);
}

// src/Overlap.cpp:37-139: 15-axis separating-axis test
bool ovBoxBox(V3 pos0, Q or0, V3 he0, V3 pos1, Q or1, V3 he1) {
    M3 u0 = mat3_cast(or0), u1 = mat3_cast(or1);
    V3 t = pos1 - pos0;
    t = v3(dot(t, u0[0]), dot(t, u0[1]), dot(t, u0[2]));
    float r[3][3], ar[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { r[i][j] = dot(u0[i], u1[j]); ar[i][j] = gabs(r[i][j]); }
    const float a[3] = {he0.x, he0.y, he0.z}, b[3] = {he1.x, he1.y, he1.z}, tt[3] = {t.x, t.y, t.z};
    for (int i = 0; i < 3; ++i) {
        float ra = a[i], rb = b[0] * ar[i][0] + b[1] * ar[i][1] + b[2] * ar[i][2];
        if (gabs(tt[i]) - ra - rb > 0) return false;
    }
    for (int i = 0; i < 3; ++i) {
        float ra = a[0] * ar[0][i] + a[1] * ar[1][i] + a[2] * ar[2][i], rb = b[i];
        if (gabs(tt[0] * r[0][i] + tt[1] * r[1][i] + tt[2] * r[2][i]) - ra - rb > 0) return false;
    }
    // L = Ai x Bj: the other two axes of A and B, in the reference's operand order
    for (int i = 0; i < 3; ++i) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
        for (int j = 0; j < 3; ++j) {
            const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            // ra = he0[lo] * absR[hi][j] + he0[hi] * absR[lo][j], lo < hi the two indices != i
            const int lo = i1 < i2 ? i1 : i2, hi = i1 < i2 ? i2 : i1;
            const int blo = j1 < j2 ? j1 : j2, bhi = j1 < j2 ? j2 : j1;
            float ra = a[lo] * ar[hi][j] + a[hi] * ar[lo][j];
            float rb = b[blo] * ar[i][bhi] + b[bhi] * ar[i][blo];
            float l = gabs(tt[i2] * r[i1][j] - tt[i1] * r[i2][j]);
            if (l - ra - rb > 0) return false;
        }
    }
    return true;
}

// src/Overlap.cpp:145-153
bool ovSphereCapsule(V3 pos0, float r0, V3 pos1, Q or1, float hh1, float r1) {
    V3 dir = qrot(or1, v3(0, 1, 0));
    V3 q = closestPointOnSegment(pos0, pos1, dir, -hh1, hh1);
    V3 v = q - pos0;
    float rs = r0 + r1;
    return !(dot(v, v) > rs * rs);
}

// src/Overlap.cpp:155-169
bool ovSphereBox(V3 pos0, float r0, V3 pos1, Q or1, V3 he1) {
    M3 u1 = mat3_cast(or1);
    V3 d = pos0 - pos1;
    d = v3(dot(d, u1[0]), dot(d, u1[1]), dot(d, u1[2]));
    V3 q = pos1 + mulmv(u1, vclamp(d, -he1, he1));
    V3 v = q - pos0;
    return !(dot(v, v) > r0 * r0);
}

// src/Overlap.cpp:175-193
bool ovCapsuleBox(V3 pos0, Q or0, float hh0, float r0, V3 pos1, Q or1, V3 he1) {
    M3 u1 = mat3_cast(or1);
    V3 p = pos0 - pos1;
    p = v3(dot(p, u1[0]), dot(p, u1[1]), dot(p, u1[2]));
    V3 dir = qrot(or0, v3(0, 1, 0));
    V3 dirLc = v3(dot(dir, u1[0]), dot(dir, u1[1]), dot(dir, u1[2]));
    float t;
    V3 q;
    float sq = sqrDistSegmentAABB(p, dirLc, -hh0, hh0, he1, t, q);
    return !(sq >= r0 * r0);
}

// physecs::overlap (src/Overlap.cpp:203-235); a triangle mesh on either side never overlaps
bool overlapGeom(V3 p0, Q q0, const Geom& g0, V3 p1, Q q1, const Geom& g1) {
    const V3 s0 = v3(g0.p[0], g0.p[1], g0.p[2]), s1 = v3(g1.p[0], g1.p[1], g1.p[2]);
    switch (g0.type) {
        case SPHERE:
            switch (g1.type) {
                case SPHERE: return ovSphereSphere(p0, g0.p[0], p1, g1.p[0]);
                case CAPSULE: return ovSphereCapsule(p0, g0.p[0], p1, q1, g1.p[0], g1.p[1]);
                case BOX: return ovSphereBox(p0, g0.p[0], p1, q1, s1);
                case CONVEX_MESH: return overlapSphereConvex(p0, g0.p[0], p1, q1, g1.mesh, s1);
            }
            return false;
        case CAPSULE:
            switch (g1.type) {
                case SPHERE: return ovSphereCapsule(p1, g1.p[0], p0, q0, g0.p[0], g0.p[1]);
                case CAPSULE: return ovCapsuleCapsule(p0, q0, g0.p[0], g0.p[1], p1, q1, g1.p[0], g1.p[1]);
                case BOX: return ovCapsuleBox(p0, q0, g0.p[0], g0.p[1], p1, q1, s1);
                case CONVEX_MESH: return overlapCapsuleConvex(p0, q0, g0.p[0], g0.p[1], p1, q1, g1.mesh, s1);
            }
            return false;
        case BOX:
            switch (g1.type) {
                case SPHERE: return ovSphereBox(p1, g1.p[0], p0, q0, s0);
                case CAPSULE: return ovCapsuleBox(p1, q1, g1.p[0], g1.p[1], p0, q0, s0);
                case BOX: return ovBoxBox(p0, q0, s0, p1, q1, s1);
                case CONVEX_MESH: return overlapBoxConvex(p0, q0, s0, p1, q1, g1.mesh, s1);
            }
            return false;
        case CONVEX_MESH:
            switch (g1.type) {
                case SPHERE: return overlapSphereConvex(p1, g1.p[0], p0, q0, g0.mesh, s0);
                case CAPSULE: return overlapCapsuleConvex(p1, q1, g1.p[0], g1.p[1], p0, q0, g0.mesh, s0);
                case BOX: return overlapBoxConvex(p1, q1, s1, p0, q0, g0.mesh, s0);
                case CONVEX_MESH: return overlapConvexConvex(p0, q0, g0.mesh, s0, p1, q1, g1.mesh, s1);
            }
            return false;
    }
    return false;
}

}  // namespace

extern "C" {

// registry.patch<TransformComponent> of one entity: Scene::onRigidBodyMove ->
// updateBounds (src/Physecs.cpp:51-54,79-90): the pose, then every collider's bounds of the
// entity with the 0.01 margin, static or dynamic.
void orc_world_patch_transform(orc_world* w, int actor, const float* pos, const float* q) {
    Actor& a = w->actors.at((size_t)actor);
    a.pos = v3(pos[0], pos[1], pos[2]);
    a.quat = qmake(q[3], q[0], q[1], q[2]);
    for (auto& e : w->bp) {
        const Collider& c = w->cols[e.collider];
        if (c.actor != actor) continue;
        Bounds b = getBounds(a.pos + qrot(a.quat, c.lpos), qmul(a.quat, c.lquat), c.geom);
        b.max = b.max + v3s(0.01f);
        b.min = b.min - v3s(0.01f);
        e.bounds = b;
    }
}

// Scene::setIsKinematic (src/Physecs.cpp:753-770): the flag, zero velocities when it turns
// kinematic, and the broadphase entries' isDynamic.
void orc_world_set_kinematic(orc_world* w, int actor, int kinematic) {
    const Actor& a = w->actors.at((size_t)actor);
    if (a.dyn < 0) return;
    Dyn& d = w->dyn[a.dyn];
    d.kinematic = kinematic != 0;
    if (d.kinematic) { d.v = v3s(0); d.w = v3s(0); }
    for (auto& e : w->bp)
        if (w->cols[e.collider].actor == actor) e.isDynamic = !d.kinematic;
}

// The last step's trigger events: which = 0 enter, 1 exit; pairs (collider of entity0,
// collider of entity1) as the oracle's collider handles (the scene's collider order, then
// added colliders in the order added), ascending; -1 for a collider that is gone (an exit
// after Scene::clearColliders). Returns the count (at most cap pairs written).
int orc_world_trigger_events(const orc_world* w, int which, int cap, int32_t* out) {
    const auto& v = which == 0 ? w->trigEnter : w->trigExit;
    std::map<std::pair<uint32_t, int>, int> live;
    for (size_t c = 0; c < w->cols.size(); ++c)
        if (w->cols[c].alive) live[{w->actors[w->cols[c].actor].entity, w->cols[c].index}] = (int)c;
    auto handle = [&](uint32_t e, int i) {
        auto it = live.find({e, i});
        return it == live.end() ? -1 : it->second;
    };
    std::vector<std::pair<int, int>> pairs;
    for (const TrigKey& k : v)
        pairs.push_back({handle(std::get<0>(k), std::get<1>(k)), handle(std::get<2>(k), std::get<3>(k))});
    std::sort(pairs.begin(), pairs.end());
    for (size_t i = 0; i < pairs.size() && (int)i < cap; ++i) {
        out[2 * i] = pairs[i].first;
        out[2 * i + 1] = pairs[i].second;
    }
    return (int)pairs.size();
}

// The same events as the listeners get them (src/Physecs.cpp:538-551): rows (entity0,
// colliderIndex0, entity1, colliderIndex1), ascending.
int orc_world_trigger_events_ec(const orc_world* w, int which, int cap, uint32_t* out) {
    const auto& v = which == 0 ? w->trigEnter : w->trigExit;
    for (size_t i = 0; i < v.size() && (int)i < cap; ++i) {
        out[4 * i] = std::get<0>(v[i]);
        out[4 * i + 1] = (uint32_t)std::get<1>(v[i]);
        out[4 * i + 2] = std::get<2>(v[i]);
        out[4 * i + 3] = (uint32_t)std::get<3>(v[i]);
    }
    return (int)v.size();
}

// Scene::addCollider (src/Physecs.cpp:739-752): the collider takes its entity's next index
// (the count of the entity's colliders), and a broadphase entry with its bounds at the
// actor's current pose, without margin, isDynamic from the body's kinematic flag. The
// caches are untouched. Returns the new collider's handle.
int orc_world_add_collider(orc_world* w, int actor, int type, const float* prm, const float* lpos, const float* lquat,
                           const float* material, int trigger, int enable, int data) {
    const Actor& a = w->actors.at((size_t)actor);
    Collider c;
    c.actor = actor;
    c.index = 0;
    for (const Collider& o : w->cols)
        if (o.alive && o.actor == actor) ++c.index;
    c.lpos = v3(lpos[0], lpos[1], lpos[2]);
    c.lquat = {lquat[0], lquat[1], lquat[2], lquat[3]};
    c.geom.type = type;
    for (int k = 0; k < 4; ++k) c.geom.p[k] = prm[k];
    if (type == CONVEX_MESH) c.geom.mesh = &w->convex.at((size_t)prm[3]);
    if (type == TRIANGLE_MESH) c.geom.tmesh = &w->trimesh.at((size_t)prm[3]);
    c.friction = material[0];
    c.restitution = material[1];
    c.damping = material[2];
    c.trigger = trigger != 0;
    c.enable = enable != 0;
    c.data = data;
    const int handle = (int)w->cols.size();
    w->cols.push_back(c);
    if (w->filterK) w->filterClass.push_back(0);   // class 0 until the filter is set again
    const bool isDynamic = a.dyn >= 0 ? !w->dyn[a.dyn].kinematic : false;
    Bounds b = getBounds(a.pos + qrot(a.quat, c.lpos), qmul(a.quat, c.lquat), c.geom);
    w->bp.push_back({handle, a.entity, c.index, b, c.enable, isDynamic});
    return handle;
}

// Scene::clearColliders (src/Physecs.cpp:725-737): each of the entity's broadphase entries,
// in collider-index order, replaced by the last entry; the colliders are gone. The contact
// and trigger caches keep their keys, so a collider added later under the same index takes
// them over, as in the reference.
void orc_world_clear_colliders(orc_world* w, int actor) {
    std::vector<int> mine;
    for (size_t c = 0; c < w->cols.size(); ++c)
        if (w->cols[c].alive && w->cols[c].actor == actor) mine.push_back((int)c);
    std::sort(mine.begin(), mine.end(), [&](int x, int y) { return w->cols[x].index < w->cols[y].index; });
    for (int c : mine) {
        for (size_t i = 0; i < w->bp.size(); ++i) {
            if (w->bp[i].collider != c) continue;
            w->bp[i] = w->bp.back();
            w->bp.pop_back();
            break;
        }
        w->cols[c].alive = false;
    }
    if (g_mutation == MUT_CLEAR_DROPS_CACHE) {
        const uint32_t e = w->actors.at((size_t)actor).entity;
        for (auto it = w->cache.begin(); it != w->cache.end();)
            it = (it->first.e0 == e || it->first.e1 == e) ? w->cache.erase(it) : std::next(it);
    }
}

int orc_overlap(int type0, const float* prm0, const float* pos0, const float* q0, int type1, const float* prm1,
                const float* pos1, const float* q1) {
    Geom g0 = pinGeom(type0, prm0), g1 = pinGeom(type1, prm1);
    return overlapGeom(v3(pos0[0], pos0[1], pos0[2]), qmake(q0[3], q0[0], q0[1], q0[2]), g0,
                       v3(pos1[0], pos1[1], pos1[2]), qmake(q1[3], q1[0], q1[1], q1[2]), g1) ? 1 : 0;
}

// Scene::overlap (src/Physecs.cpp:628-651) over the world's colliders: a leaf is reached
// when the query's bounds intersect the leaf's bounds (every ancestor's box holds them), so
// the hits are those of this loop; they are listed by collider index (the reference lists
// them in its tree's order). filter != 0 keeps colliders whose data has a common bit with
// it. A convex query's params[3] is one of the world's convex meshes. Returns the hit count
// (at most cap written to out).
int orc_world_overlap(const orc_world* w, int type, const float* prm, const float* pos, const float* q, int filter,
                      int cap, int32_t* out) {
    if (type == TRIANGLE_MESH) return 0;   // overlaps nothing (src/Overlap.cpp:203-235)
    Geom g{type, {prm[0], prm[1], prm[2], prm[3]}};
    if (type == CONVEX_MESH) g.mesh = &w->convex.at((size_t)prm[3]);   // one of the scene's convex meshes
    V3 p = v3(pos[0], pos[1], pos[2]);
    Q qq = qmake(q[3], q[0], q[1], q[2]);
    Bounds b = getBounds(p, qq, g);
    std::vector<int> hits;
    for (const BPEntry& e : w->bp) {
        const Bounds& eb = e.bounds;
        if (b.max.x < eb.min.x || b.min.x > eb.max.x || b.max.y < eb.min.y || b.min.y > eb.max.y ||
            b.max.z < eb.min.z || b.min.z > eb.max.z)
            continue;   // intersects (src/BoundsUtil.cpp:87-92)
        const Collider& c = w->cols[e.collider];
        if (filter && !(c.data & filter)) continue;
        const Actor& a = w->actors[c.actor];
        if (overlapGeom(p, qq, g, a.pos + qrot(a.quat, c.lpos), qmul(a.quat, c.lquat), c.geom)) hits.push_back(e.collider);
    }
    std::sort(hits.begin(), hits.end());
    for (size_t i = 0; i < hits.size() && (int)i < cap; ++i) out[i] = hits[i];
    return (int)hits.size();
}

// Scene::overlapWithMinTranslationalDistance (src/Physecs.cpp:653-688): every collider
// whose bounds meet the query's bounds is tested with collision(collider, query); each
// manifold with points gives {collider, normal, mtd}, mtd = the largest
// dot(contact0 - contact1, normal) over its points, and at least 0. Listed by collider
// index, then in collision()'s manifold order (the reference lists them in its tree's
// order). A convex query's params[3] is one of the world's convex meshes. Writes at most
// cap hits: out_col[i], out_nm[4i..] = normal xyz, mtd. Returns the hit count.
struct MtdHit { int col; V3 n; float mtd; };

static std::vector<MtdHit> overlapMtd(const orc_world* w, V3 p, Q qq, const Geom& g) {
    std::vector<MtdHit> out;
    Bounds b = getBounds(p, qq, g);
    std::vector<int> cands;
    for (const BPEntry& e : w->bp) {
        const Bounds& eb = e.bounds;
        if (b.max.x < eb.min.x || b.min.x > eb.max.x || b.max.y < eb.min.y || b.min.y > eb.max.y ||
            b.max.z < eb.min.z || b.min.z > eb.max.z)
            continue;
        cands.push_back(e.collider);
    }
    std::sort(cands.begin(), cands.end());
    for (int ci : cands) {
        const Collider& c = w->cols[ci];
        const Actor& a = w->actors[c.actor];
        std::vector<Manifold> res;
        if (!collision(a.pos + qrot(a.quat, c.lpos), qmul(a.quat, c.lquat), c.geom, p, qq, g, res)) continue;
        for (const Manifold& m : res) {
            if (!m.numPoints) continue;
            float mtd = 0;
            for (int i = 0; i < m.numPoints; i++) {
                float dist = dot(m.points[i].p0 - m.points[i].p1, m.normal);
                if (dist > mtd) mtd = dist;
            }
            out.push_back({ci, m.normal, mtd});
        }
    }
    return out;
}

int orc_world_overlap_mtd(const orc_world* w, int type, const float* prm, const float* pos, const float* q, int cap,
                          int32_t* out_col, float* out_nm) {
    Geom g{type, {prm[0], prm[1], prm[2], prm[3]}};
    if (type == CONVEX_MESH) g.mesh = &w->convex.at((size_t)prm[3]);
    if (type == TRIANGLE_MESH) return -1;   // not taken by the device boundary either
    const std::vector<MtdHit> hits = overlapMtd(w, v3(pos[0], pos[1], pos[2]), qmake(q[3], q[0], q[1], q[2]), g);
    for (size_t k = 0; k < hits.size() && (int)k < cap; ++k) {
        out_col[k] = hits[k].col;
        out_nm[4 * k] = hits[k].n.x; out_nm[4 * k + 1] = hits[k].n.y; out_nm[4 * k + 2] = hits[k].n.z;
        out_nm[4 * k + 3] = hits[k].mtd;
    }
    return (int)hits.size();
}

// CharacterController::move (src/CharacterController.cpp:5-45) of one character whose
// entity stands at pos (updated) with orientation q. The step-up branch (:13-23) asks
// raycastClosest's unfiltered overload, which reports no hit (its leaf test calls an
// empty filter), so it never changes the translation. Returns CharacterCollisionFlags.
int orc_character_move(const orc_world* w, float* pos, const float* q, float height, float radius,
                       float slopeLimit, int numIterations, const float* translation) {
    Q ori = qmake(q[3], q[0], q[1], q[2]);
    V3 up = qrot(ori, v3(0, 1, 0));
    V3 position = v3(pos[0], pos[1], pos[2]) + v3(translation[0], translation[1], translation[2]);
    int flags = 0;
    Geom geom{CAPSULE, {height * 0.5f, radius, 0, 0}};
    for (int i = 0; i < numIterations; ++i) {
        for (const MtdHit& h : overlapMtd(w, position + up * (height * 0.5f + radius), ori, geom)) {
            float d = dot(h.n, up);
            if (d > 0.001f) flags |= 1;          // DOWN
            else if (d < -0.001f) flags |= 4;    // UP
            else flags |= 2;                     // SIDES
            if (d > slopeLimit) position = position + up * h.mtd / d;
            else position = position + h.n * h.mtd;
        }
    }
    pos[0] = position.x; pos[1] = position.y; pos[2] = position.z;
    return flags;
}

int orc_ray_aabb(const float* o, const float* d, const float* mn, const float* mx, float* t) {
    return rayAABB(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), v3(mn[0], mn[1], mn[2]), v3(mx[0], mx[1], mx[2]), *t) ? 1 : 0;
}

int orc_ray_geometry(const float* o, const float* d, int type, const float* prm, const float* pos, const float* q,
                     float* t) {
    Geom g = pinGeom(type, prm);
    return rayGeometry(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), v3(pos[0], pos[1], pos[2]),
                       qmake(q[3], q[0], q[1], q[2]), g, *t) ? 1 : 0;
}

// Scene::raycastClosest (src/Physecs.cpp:571-621) over the world's colliders with their
// current broadphase bounds. The reference walks its dynamic AABB tree; a leaf is reached
// when the ray enters the leaf's own bounds within maxDistance (every ancestor's box holds
// those bounds), so the accepted hits, and the closest of them, are the ones this loop
// over every collider finds. Equal distances go to the lower collider index here; the
// reference settles them by its tree's shape. The filter rejects one entity (`ignore`;
// 0xffffffff rejects none); the reference's own unfiltered overload passes an empty
// std::function, which its leaf test (`filter && filter(entity)`) never accepts.
// Returns the collider index hit, or -1; *tOut is the hit distance along d.
int orc_world_raycast(const orc_world* w, const float* o, const float* d, float maxDistance, uint32_t ignore,
                      float* tOut) {
    V3 ro = v3(o[0], o[1], o[2]), rd = v3(d[0], d[1], d[2]);
    int best = -1;
    float bestT = 0.f;
    for (const BPEntry& e : w->bp) {
        const Collider& c = w->cols[e.collider];
        const Actor& a = w->actors[c.actor];
        if (a.entity == ignore) continue;
        float t;
        if (!rayAABB(ro, rd, e.bounds.min, e.bounds.max, t) || t > maxDistance) continue;
        if (!rayGeometry(ro, rd, a.pos + qrot(a.quat, c.lpos), qmul(a.quat, c.lquat), c.geom, t)) continue;
        if (t > maxDistance) continue;
        if (best < 0 || t < bestT || (t == bestT && e.collider < best)) { best = e.collider; bestT = t; }
    }
    if (best >= 0) *tOut = bestT;
    return best;
}

}  // extern "C"
