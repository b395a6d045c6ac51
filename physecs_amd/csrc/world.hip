// C ABI of libphysecs_gpu.so (include/physecs_gpu.h): world lifetime, upload/download of
// the component state, and the per-step orchestration of Scene::simulate
// (src/Physecs.cpp:112-561) on one HIP stream.
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <numeric>
#include <unordered_map>
#include "pgx_internal.h"

using namespace pgx;

struct pgx_world {
    World w;
};

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

template <class F>
static int guarded(F&& f) {
    try {
        return f();
    } catch (const HipError& e) {
        return fail(PGX_ERR_HIP, "HIP error %d (%s) at %s:%d: %s", (int)e.err, hipGetErrorString(e.err), e.file,
                    e.line, e.expr);
    } catch (const PgxError& e) {
        return fail(e.code, "%s", e.msg.c_str());
    } catch (const std::bad_alloc&) {
        return fail(PGX_ERR_INVALID, "host out of memory");
    } catch (...) {
        return fail(PGX_ERR_INVALID, "unknown error");
    }
}

template <class T>
static void upload(DBuf<T>& b, const std::vector<T>& v, hipStream_t s) {
    b.alloc(v.size());
    if (!v.empty()) PGX_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
}

static int bits_for(long n) {
    int b = 1;
    while ((1L << b) < n) ++b;
    return b;
}

extern "C" {

const char* pgx_last_error(void) { return g_err.c_str(); }

int pgx_device_count(int* count) {
    if (!count) return fail(PGX_ERR_INVALID, "count is NULL");
    return guarded([&] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        *count = e == hipSuccess ? n : 0;
        return PGX_OK;
    });
}

int pgx_sort_u64(int device, const uint64_t* keys, const int32_t* vals, int n, int bits, uint64_t* keys_out,
                 int32_t* vals_out) {
    if (n < 0 || bits < 0 || bits > 64) return fail(PGX_ERR_INVALID, "n must be >= 0 and bits in [0, 64]");
    if (n > 0 && (!keys || !vals || !keys_out || !vals_out)) return fail(PGX_ERR_INVALID, "array is NULL");
    return guarded([&] {
        PGX_HIP(hipSetDevice(device));
        hipStream_t s = nullptr;
        PGX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard{s};
        DBuf<uint64_t> k;
        DBuf<int32_t> v;
        SortTemp t;
        k.alloc(n);
        v.alloc(n);
        if (n) {
            PGX_HIP(hipMemcpyAsync(k.p, keys, sizeof(uint64_t) * n, hipMemcpyHostToDevice, s));
            PGX_HIP(hipMemcpyAsync(v.p, vals, sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
        }
        radix_sort_u64(s, k.p, v.p, n, bits, t);
        if (n) {
            PGX_HIP(hipMemcpyAsync(keys_out, k.p, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, s));
            PGX_HIP(hipMemcpyAsync(vals_out, v.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
        }
        PGX_HIP(hipStreamSynchronize(s));
        return PGX_OK;
    });
}

}  // extern "C"

static int validate_scene(const pgx_scene_desc* sc) {
    if (!sc) return fail(PGX_ERR_INVALID, "scene is NULL");
    if (sc->n_actors < 0 || sc->n_dyn < 0 || sc->n_colliders < 0)
        return fail(PGX_ERR_INVALID, "negative counts");
    // validate the descriptor on the host before anything reaches the device
    for (int d = 0; d < sc->n_dyn; ++d) {
        int a = sc->actor_of_dyn[d];
        if (a < 0 || a >= sc->n_actors || sc->dyn_of_actor[a] != d)
            return fail(PGX_ERR_INVALID, "dynamic entry %d: actor_of_dyn / dyn_of_actor mismatch", d);
    }
    int W = sc->actor_world ? std::max(1, sc->n_worlds) : 1;
    for (int a = 0; a < sc->n_actors; ++a) {
        int wd = sc->actor_world ? sc->actor_world[a] : 0;
        if (wd < 0 || wd >= W) return fail(PGX_ERR_INVALID, "actor %d: world %d out of range", a, wd);
        int d = sc->dyn_of_actor[a];
        if (d < -1 || d >= sc->n_dyn) return fail(PGX_ERR_INVALID, "actor %d: dyn index %d out of range", a, d);
    }
    for (int c = 0; c < sc->n_colliders; ++c) {
        int a = sc->col_actor[c];
        if (a < 0 || a >= sc->n_actors) return fail(PGX_ERR_INVALID, "collider %d: actor out of range", c);
        int t = sc->col_type[c];
        if (t == PGX_CONVEX_MESH) {
            float mf = sc->col_params[4 * c + 3];
            if (!(mf >= 0.f) || mf != (float)(int)mf || (int)mf >= sc->n_convex)
                return fail(PGX_ERR_INVALID, "collider %d: convex mesh index %g out of range", c, (double)mf);
        } else if (t == PGX_TRIANGLE_MESH) {
            float mf = sc->col_params[4 * c + 3];
            if (!(mf >= 0.f) || mf != (float)(int)mf || (int)mf >= sc->n_trimesh)
                return fail(PGX_ERR_INVALID, "collider %d: triangle mesh index %g out of range", c, (double)mf);
        } else if (t != PGX_SPHERE && t != PGX_CAPSULE && t != PGX_BOX) {
            return fail(PGX_ERR_UNSUPPORTED, "collider %d: geometry type %d is not supported on the device path", c, t);
        }
    }
    for (int m = 0; m < sc->n_trimesh; ++m) {
        int nv = sc->tri_vert_start[m + 1] - sc->tri_vert_start[m];
        int t0 = sc->tri_start[m], t1 = sc->tri_start[m + 1];
        if (nv < 3 || t1 <= t0) return fail(PGX_ERR_INVALID, "triangle mesh %d: no vertices or triangles", m);
        for (long k = 3L * t0; k < 3L * t1; ++k)
            if (sc->tri_idx[k] < 0 || sc->tri_idx[k] >= nv)
                return fail(PGX_ERR_INVALID, "triangle mesh %d: vertex index out of range", m);
    }
    for (int m = 0; m < sc->n_convex; ++m) {
        int nv = sc->convex_vert_start[m + 1] - sc->convex_vert_start[m];
        int f0 = sc->convex_face_start[m], f1 = sc->convex_face_start[m + 1];
        if (nv < 1 || f1 <= f0) return fail(PGX_ERR_INVALID, "convex mesh %d: no vertices or faces", m);
        for (int f = f0; f < f1; ++f) {
            int k0 = sc->face_idx_start[f], k1 = sc->face_idx_start[f + 1];
            if (k1 - k0 < 3 || k1 - k0 > CVX_MAX_FACE_VERTS)
                return fail(PGX_ERR_INVALID, "convex mesh %d face %d: %d vertices (3..%d supported)", m, f - f0,
                            k1 - k0, CVX_MAX_FACE_VERTS);
            for (int k = k0; k < k1; ++k)
                if (sc->face_idx[k] < 0 || sc->face_idx[k] >= nv)
                    return fail(PGX_ERR_INVALID, "convex mesh %d face %d: vertex index out of range", m, f - f0);
        }
    }
    return PGX_OK;
}

namespace pgx {

// The device-layout arrays of a descriptor: bodies in internal order (the dynamic storage
// first, so body b == dynamic index b, then the other actors), colliders in descriptor
// order with world-local indices, the dynamic colliders (LBVH leaves) and the others
// grouped by world. Entity ids are shifted by entityBase.
void convert_scene(const pgx_scene_desc* sc, uint32_t entityBase, SceneArrays& o) {
    const int W = sc->actor_world ? std::max(1, sc->n_worlds) : 1;
    const int A = sc->n_actors, D = sc->n_dyn, C = sc->n_colliders;
    o.A = A; o.D = D; o.C = C; o.W = W;
    o.actorToBody.assign(A, -1);
    o.bodyToActor.clear();
    o.bodyToActor.reserve(A);
    for (int d = 0; d < D; ++d) { o.actorToBody[sc->actor_of_dyn[d]] = d; o.bodyToActor.push_back(sc->actor_of_dyn[d]); }
    for (int a = 0; a < A; ++a)
        if (sc->dyn_of_actor[a] < 0) { o.actorToBody[a] = (int)o.bodyToActor.size(); o.bodyToActor.push_back(a); }
    o.bpos.resize(A); o.bquat.resize(A); o.bent.resize(A);
    for (int b = 0; b < A; ++b) {
        int a = o.bodyToActor[b];
        o.bpos[b] = make_float4(sc->pos[3 * a], sc->pos[3 * a + 1], sc->pos[3 * a + 2], 0.f);
        o.bquat[b] = make_float4(sc->quat[4 * a], sc->quat[4 * a + 1], sc->quat[4 * a + 2], sc->quat[4 * a + 3]);
        o.bent[b] = sc->entity[a] + entityBase;
    }
    o.dvel.resize(D); o.dang.resize(D); o.dcom.resize(D); o.dinvI.resize(3 * (size_t)D);
    for (int d = 0; d < D; ++d) {
        o.dvel[d] = make_float4(sc->vel[3 * d], sc->vel[3 * d + 1], sc->vel[3 * d + 2], sc->inv_mass[d]);
        o.dang[d] = make_float4(sc->angvel[3 * d], sc->angvel[3 * d + 1], sc->angvel[3 * d + 2], 0.f);
        o.dcom[d] = make_float4(sc->com[3 * d], sc->com[3 * d + 1], sc->com[3 * d + 2], sc->kinematic[d] ? 1.f : 0.f);
        for (int k = 0; k < 3; ++k)
            o.dinvI[3 * d + k] = make_float4(sc->inv_inertia[9 * d + 3 * k], sc->inv_inertia[9 * d + 3 * k + 1],
                                             sc->inv_inertia[9 * d + 3 * k + 2], 0.f);
    }
    o.kinematic.assign(sc->kinematic, sc->kinematic + D);
    o.dynWorld.resize(D);
    for (int d = 0; d < D; ++d) o.dynWorld[d] = sc->actor_world ? sc->actor_world[sc->actor_of_dyn[d]] : 0;
    // colliders (descriptor order), world-local indices
    o.cbody.resize(C); o.cworld.resize(C); o.clocal.resize(C); o.cflags.resize(C); o.ctype.resize(C);
    o.cindex.resize(C); o.cent.resize(C); o.cparams.resize(C); o.clpos.resize(C); o.clquat.resize(C); o.cmat.resize(C);
    std::vector<int32_t> perWorld(W, 0);
    for (int c = 0; c < C; ++c) {
        int a = sc->col_actor[c];
        int wd = sc->actor_world ? sc->actor_world[a] : 0;
        o.cbody[c] = o.actorToBody[a];
        o.cworld[c] = wd;
        o.clocal[c] = perWorld[wd]++;
        int d = sc->dyn_of_actor[a];
        bool dyn = d >= 0 && !sc->kinematic[d];
        o.cflags[c] = (sc->col_enable[c] ? CF_ENABLE : 0) | (sc->col_trigger[c] ? CF_TRIGGER : 0) | (dyn ? CF_DYNAMIC : 0);
        o.ctype[c] = sc->col_type[c];
        o.cindex[c] = sc->col_index[c];
        o.cent[c] = sc->entity[a] + entityBase;
        o.cparams[c] = make_float4(sc->col_params[4 * c], sc->col_params[4 * c + 1], sc->col_params[4 * c + 2],
                                   sc->col_params[4 * c + 3]);
        o.clpos[c] = make_float4(sc->col_pos[3 * c], sc->col_pos[3 * c + 1], sc->col_pos[3 * c + 2], 0.f);
        o.clquat[c] = make_float4(sc->col_quat[4 * c], sc->col_quat[4 * c + 1], sc->col_quat[4 * c + 2],
                                  sc->col_quat[4 * c + 3]);
        o.cmat[c] = make_float4(sc->col_material[3 * c], sc->col_material[3 * c + 1], sc->col_material[3 * c + 2], 0.f);
    }
    o.wstart.assign(W + 1, 0);
    o.colOfLocal.resize(C);
    for (int i = 0; i < W; ++i) o.wstart[i + 1] = o.wstart[i] + perWorld[i];
    for (int c = 0; c < C; ++c) o.colOfLocal[o.wstart[o.cworld[c]] + o.clocal[c]] = c;
    // LBVH leaves: colliders of non-kinematic dynamic bodies; the rest per world
    o.dynCols.clear();
    o.wss.assign(W + 1, 0);
    for (int c = 0; c < C; ++c)
        if (o.cflags[c] & CF_DYNAMIC) o.dynCols.push_back(c); else ++o.wss[o.cworld[c] + 1];
    for (int i = 0; i < W; ++i) o.wss[i + 1] += o.wss[i];
    o.staticCols.resize(o.wss[W]);
    {
        std::vector<int32_t> fillp(o.wss.begin(), o.wss.end() - 1);
        for (int c = 0; c < C; ++c)
            if (!(o.cflags[c] & CF_DYNAMIC)) o.staticCols[fillp[o.cworld[c]]++] = c;
    }
    o.maxPerWorld = W ? *std::max_element(perWorld.begin(), perWorld.end()) : 1;
}

// Host state of a previous scene (joints, non-colliding pairs, contact cache, statistics)
// is cleared.
void world_reset_host(World& w) {
    w.J = 0; w.jCap = 0; w.jRowsTotal = 0;
    w.custList.clear(); w.custRowsHost.clear();
    w.meshColsHost.clear();   // the server's worlds bound meshes in k_bounds
    for (auto* v : {&w.hjType, &w.hjA0, &w.hjA1, &w.hjB0, &w.hjB1, &w.hjRowFlags, &w.hjNRows, &w.hjColor, &w.hjAlive})
        v->clear();
    for (auto* v : {&w.hjP0, &w.hjQ0, &w.hjP1, &w.hjQ1, &w.hjDrive, &w.hjPar0, &w.hjPar1}) v->clear();
    w.hjBits.clear();
    w.jColorStart.clear();
    w.hostNonColliding.clear();
    w.nonCollidingDirty = true;
    w.cacheCur = 0; w.cacheCount[0] = w.cacheCount[1] = 0;
    w.nConvex = 0; w.nTriMesh = 0; w.mmCap = 0; w.lastMeshSlots = 0;
    w.lastPairs = w.lastManifolds = w.lastColors = w.lastRounds = w.lastPoints = 0;
    w.lastQueues = 0; w.lastFlowPasses = 0;
}

// step parameters of the configuration; keys' bit widths of the world's shape
void world_set_shape(World& w, const pgx_config* cfg, int A, int D, int C, int W, int nDynCol, int maxPerWorld) {
    if (cfg) {
        if (cfg->num_substeps > 0) w.substeps = cfg->num_substeps;
        if (cfg->num_iterations >= 0) w.iterations = cfg->num_iterations;
        if (cfg->gravity != 0.f) w.g = cfg->gravity;
    }
    w.A = A; w.D = D; w.C = C; w.W = W;
    w.nDynCol = nDynCol;
    w.localBits = bits_for(std::max(2, maxPerWorld));
    w.worldBits = W > 1 ? bits_for(W) : 0;
    if (w.worldBits + 30 > 64 || w.worldBits + 2 * w.localBits > 64)
        throw PgxError{PGX_ERR_INVALID, "too many worlds or colliders per world for 64-bit keys"};
}

// First dynamic body of each world (W + 1 entries) for the per-world passes, which need
// each world's dynamic bodies to be one contiguous range (`contiguous`).
void world_set_dyn_starts(World& w, const std::vector<int32_t>& starts, bool contiguous) {
    w.wDynHost = starts;
    w.worldsContiguous = contiguous;
    w.maxWorldDyn = 0;
    for (int k = 0; k < 2; ++k) w.worldListHost[k].clear();
    for (size_t i = 0; i + 1 < starts.size(); ++i) {
        const int n = starts[i + 1] - starts[i];
        w.maxWorldDyn = std::max(w.maxWorldDyn, n);
        if (n > 0) w.worldListHost[n <= WORLD_SMALL ? 0 : 1].push_back((int32_t)i);
    }
    upload(w.wDyn, w.wDynHost, w.stream);
    for (int k = 0; k < 2; ++k) upload(w.worldList[k], w.worldListHost[k], w.stream);
}

// the same from the world of every dynamic body
void world_set_dyn_ranges(World& w, const std::vector<int32_t>& dynWorld) {
    std::vector<int32_t> starts(w.W + 1, 0);
    bool contiguous = true;
    for (size_t d = 0; d < dynWorld.size(); ++d) {
        ++starts[dynWorld[d] + 1];
        if (d > 0 && dynWorld[d] < dynWorld[d - 1]) contiguous = false;
    }
    for (int i = 0; i < w.W; ++i) starts[i + 1] += starts[i];
    world_set_dyn_starts(w, starts, contiguous);
}

// Every buffer whose size follows from the world's shape (bodies, colliders, worlds,
// dynamic colliders) and the pair capacity; reused where large enough.
void world_alloc_derived(World& w, const pgx_config* cfg, int nTrimesh) {
    const int D = w.D, C = w.C, W = w.W;
    w.refitReady.alloc(std::max(1, w.nDynCol));
    w.bvel.alloc(BV_WORDS * (size_t)D); w.pvel.alloc(D); w.pang.alloc(D); w.winvI.alloc(3 * (size_t)D);
    w.cbmin.alloc(C); w.cbmax.alloc(C);
    // broadphase buffers
    w.worldBoundsI.alloc(16);
    w.sceneMoments.alloc(6);
    {
        uint32_t t = 64;   // the grid's hash table: at least twice the dynamic colliders
        while (t < 2u * (uint32_t)std::max(w.nDynCol, 1)) t <<= 1;
        w.gridKey.alloc(t); w.gridStart.alloc(t); w.gridEnd.alloc(t);
        w.gridRec.alloc(2 * (size_t)std::max(w.nDynCol, 1));
    }
    w.mkeys.alloc(C); w.mvals.alloc(C);
    w.leafMin.alloc(C); w.leafMax.alloc(C); w.leafFlags.alloc(C); w.leafParent.alloc(C); w.leafEsc.alloc(C);
    size_t ni = C > 1 ? C - 1 : 1;
    w.nodeChild.alloc(2 * ni); w.nodeParent.alloc(ni); w.nodeEsc.alloc(ni); w.nodeFlags.alloc(ni);
    w.nodeCounter.alloc(ni); w.nodeRange.alloc(ni); w.nodeMin.alloc(ni); w.nodeMax.alloc(ni);
    w.nodeA.alloc(ni); w.nodeB.alloc(ni); w.nodeLast.alloc(ni);
    w.leafA.alloc(C); w.leafB.alloc(C); w.leafI.alloc(C);
    w.travA.alloc(ni + C); w.travB.alloc(ni + C); w.travC.alloc(ni + C);
    w.statA.alloc(std::max(1, C)); w.statB.alloc(std::max(1, C)); w.statL.alloc(std::max(1, C));
    w.worldRoot.alloc(W); w.worldStop.alloc(W); w.worldMStart.alloc(W);
    w.pairCap = (cfg && cfg->pair_capacity > 0) ? cfg->pair_capacity : std::max(1024, 12 * C);
    // manifolds of triangle-mesh pairs are stored after the pair slots
    w.mmCap = nTrimesh > 0 ? std::max(65536, 16 * w.nDynCol) : 0;
    w.manifoldCap = w.pairCap + w.mmCap;
    w.pkeys.alloc(w.pairCap); w.pvals.alloc(w.pairCap);
    w.counters.alloc(16);
    size_t ND = (size_t)w.pairCap + w.mmCap;
    w.npNum.alloc(ND); w.npNormal.alloc(ND); w.npPts.alloc(8 * ND);
    w.npIndex.alloc(w.pairCap);
    w.meshCount.alloc(w.pairCap); w.meshFirst.alloc(w.pairCap); w.npTri.alloc(std::max(1, w.mmCap));
    size_t M = w.manifoldCap;
    w.mfKey.alloc(M); w.mfPair.alloc(M); w.mfB0.alloc(M); w.mfB1.alloc(M); w.mfColor.alloc(M);
    w.mfSrc.alloc(M); w.mfTri.alloc(M); w.sWorld.alloc(M);
    w.claim.alloc(std::max(D, 1)); w.usedMask.alloc(2 * (size_t)std::max(D, 1));
    w.flowPub.alloc(6 * (size_t)std::max(D, 1));
    w.passBarrier.alloc(1);
    // versions start at zero; a launch's versions carry its epoch, which starts at one
    PGX_HIP(hipMemsetAsync(w.flowPub.p, 0, sizeof(uint64_t) * 6 * std::max(D, 1), w.stream));
    w.flowErr.alloc(1);
    PGX_HIP(hipMemsetAsync(w.flowErr.p, 0, sizeof(int32_t), w.stream));
    w.qStart.alloc(FLOW_QUEUES + 1); w.qList.alloc(M); w.qDone.alloc(FLOW_QUEUES);
    PGX_HIP(hipMemsetAsync(w.qDone.p, 0, sizeof(int32_t) * FLOW_QUEUES, w.stream));
    {
        int cus = 0;
        PGX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, w.device));
        w.nXcd = std::max(1, std::min(FLOW_QUEUES, cus / 32));   // 32 CUs per XCD on CDNA3 and CDNA4
    }

    w.colorKeys.alloc(M); w.colorPerm.alloc(M); w.colorRank.alloc(M); w.colorCount.alloc(160); w.colorPoints.alloc(160);
    w.sB0.alloc(M); w.sB1.alloc(M); w.sFlags.alloc(M); w.sA0.alloc(M); w.sA1.alloc(M); w.sCanon.alloc(M);
    w.sN.alloc(M); w.sSoft.alloc(M); w.sR0.alloc(4 * M); w.sR1.alloc(4 * M); w.sRow.alloc(4 * CONTACT_ROW_FIELDS * M); w.sLam.alloc(4 * M);
    for (int k = 0; k < 2; ++k) {
        w.cacheKey[k].alloc(M); w.cacheTri[k].alloc(M); w.cacheNum[k].alloc(M); w.cacheLp[k].alloc(4 * M);
    }
    w.nonColliding.alloc(1);
    w.heavy.alloc(w.pairCap);
}

// (Re)initialises a world from a validated descriptor on w's device and stream. Device
// buffers are reused where they are large enough.
void world_init(World& w, const pgx_config* cfg, const pgx_scene_desc* sc) {
    world_reset_host(w);
    SceneArrays o;
    convert_scene(sc, 0, o);
    world_set_shape(w, cfg, o.A, o.D, o.C, o.W, (int)o.dynCols.size(), o.maxPerWorld);
    w.actorToBody = o.actorToBody;
    w.bodyToActor = o.bodyToActor;
    w.hostColBody = o.cbody;
    w.hostKinematic = o.kinematic;
    w.hostColIndex = o.cindex;
    w.hostEntity = o.bent;
    world_set_dyn_ranges(w, o.dynWorld);

    hipStream_t s = w.stream;
    upload(w.bpos, o.bpos, s); upload(w.bquat, o.bquat, s); upload(w.bentity, o.bent, s);
    upload(w.dvel, o.dvel, s); upload(w.dang, o.dang, s); upload(w.dcom, o.dcom, s); upload(w.dinvI, o.dinvI, s);
    upload(w.cbody, o.cbody, s); upload(w.cworld, o.cworld, s); upload(w.clocal, o.clocal, s);
    upload(w.cflags, o.cflags, s); upload(w.ctype, o.ctype, s); upload(w.cindex, o.cindex, s);
    upload(w.centity, o.cent, s); upload(w.cparams, o.cparams, s); upload(w.clpos, o.clpos, s);
    upload(w.clquat, o.clquat, s); upload(w.cmat, o.cmat, s);
    std::vector<int32_t> cdata((size_t)o.C, 0);   // lives until the stream is synchronised below
    if (sc->col_data) std::copy(sc->col_data, sc->col_data + o.C, cdata.begin());
    upload(w.cdata, cdata, s);
    upload(w.worldColStart, o.wstart, s); upload(w.colOfLocal, o.colOfLocal, s);
    upload(w.dynCols, o.dynCols, s); upload(w.staticCols, o.staticCols, s); upload(w.worldStaticStart, o.wss, s);
    w.meshColsHost.clear();
    for (int c = 0; c < o.C; ++c)
        if (sc->col_type[c] == PGX_CONVEX_MESH || sc->col_type[c] == PGX_TRIANGLE_MESH) w.meshColsHost.push_back(c);
    if (!w.meshColsHost.empty()) upload(w.meshCols, w.meshColsHost, s);
    world_alloc_derived(w, cfg, sc->n_trimesh);
    // trigger lists: at most one entry per pair slot (pairCap is set by world_alloc_derived)
    w.hasTriggers = false;
    w.hostColOfLocal = o.colOfLocal;
    w.hostWorldColStart = o.wstart;
    bool anyTrigger = false;
    for (int c = 0; c < o.C && sc->col_trigger; ++c) anyTrigger = anyTrigger || sc->col_trigger[c] != 0;
    w.trigCur = 0;
    if (anyTrigger) world_enable_triggers(w);
    if (sc->n_convex > 0) {        // ConvexMesh vertex buffers padded to a multiple of 4 with the last vertex
        // (src/ConvexMesh.cpp:3-11); faces, normals and centroids as given
        std::vector<float4> cvv, cvn, cvc;
        std::vector<int4> cvm;
        std::vector<int2> cvf;
        std::vector<int32_t> cvfi;
        for (int m = 0; m < sc->n_convex; ++m) {
            int v0 = sc->convex_vert_start[m], nv = sc->convex_vert_start[m + 1] - v0;
            int first = (int)cvv.size();
            for (int i = 0; i < nv; ++i)
                cvv.push_back(make_float4(sc->convex_verts[3 * (v0 + i)], sc->convex_verts[3 * (v0 + i) + 1],
                                          sc->convex_verts[3 * (v0 + i) + 2], 0.f));
            while ((cvv.size() - first) % 4) cvv.push_back(cvv.back());
            int f0 = sc->convex_face_start[m], nf = sc->convex_face_start[m + 1] - f0;
            cvm.push_back(make_int4(first, nv, (int)cvf.size(), nf));
            for (int f = f0; f < f0 + nf; ++f) {
                int k0 = sc->face_idx_start[f], k1 = sc->face_idx_start[f + 1];
                cvf.push_back(make_int2((int)cvfi.size(), k1 - k0));
                for (int k = k0; k < k1; ++k) cvfi.push_back(sc->face_idx[k]);
                cvn.push_back(make_float4(sc->face_normal[3 * f], sc->face_normal[3 * f + 1],
                                          sc->face_normal[3 * f + 2], 0.f));
                cvc.push_back(make_float4(sc->face_centroid[3 * f], sc->face_centroid[3 * f + 1],
                                          sc->face_centroid[3 * f + 2], 0.f));
            }
        }
        w.nConvex = sc->n_convex;
        upload(w.cvVerts, cvv, s); upload(w.cvMesh, cvm, s); upload(w.cvFace, cvf, s);
        upload(w.cvFaceIdx, cvfi, s); upload(w.cvFaceN, cvn, s); upload(w.cvFaceC, cvc, s);
        w.epaLanes = EPA_LANES;
        size_t L = (size_t)w.epaLanes;
        w.epaV.alloc(9 * EPA_MAXV * L); w.epaFN.alloc(3 * EPA_MAXF * L);
        w.epaFI.alloc(3 * EPA_MAXF * L); w.epaE.alloc(2 * EPA_MAXE * L);
    }
    if (sc->n_trimesh > 0) {
        w.tmMaxDepth = upload_triangle_meshes(w, sc, s);
        if (w.tmMaxDepth + 2 > TM_STACK)
            throw PgxError{PGX_ERR_UNSUPPORTED, "triangle mesh BVH depth " + std::to_string(w.tmMaxDepth) +
                                                    " exceeds the device traversal stack"};
        w.meshPairs.alloc(w.pairCap);
        w.meshLanes = TM_LANES;
        w.tmWs.alloc((size_t)TM_WS_WORDS * w.meshLanes);
        // candidate triangles of all mesh pairs of a step, one contact record each
        w.tmCandCap = std::max(1 << 18, 4 * w.mmCap);
        w.tmCand.alloc(w.tmCandCap); w.tmPairOf.alloc(w.tmCandCap);
        w.tmRecs.alloc((size_t)(TM_CI_WORDS + 1) * w.tmCandCap);
        w.tmCandOff.alloc(w.pairCap); w.tmCandCount.alloc(w.pairCap);
        if (w.epaLanes < w.meshLanes) {   // convex bodies against triangles need EPA state per lane
            w.epaLanes = w.meshLanes;
            size_t L = (size_t)w.epaLanes;
            w.epaV.alloc(9 * EPA_MAXV * L); w.epaFN.alloc(3 * EPA_MAXF * L);
            w.epaFI.alloc(3 * EPA_MAXF * L); w.epaE.alloc(2 * EPA_MAXE * L);
        }
    }
    // creation-time bounds: no margin (src/Physecs.cpp:32)
    bp_compute_bounds(w, false);
    PGX_HIP(hipStreamSynchronize(s));
}

// trigger lists: at most one entry per pair slot (pairCap is set by world_alloc_derived)
void world_enable_triggers(World& w) {
    if (w.hasTriggers) return;
    w.trigList.alloc(w.pairCap); w.trigKeys[0].alloc(w.pairCap); w.trigKeys[1].alloc(w.pairCap);
    w.trigN.alloc(4);
    fill_i32(w.stream, w.trigN.p, 0, 4);
    w.trigCur = 0;
    w.hasTriggers = true;
}

}  // namespace pgx

extern "C" {

int pgx_world_create(const pgx_config* cfg, const pgx_scene_desc* sc, pgx_world** out) {
    if (!sc || !out) return fail(PGX_ERR_INVALID, "scene or out is NULL");
    *out = nullptr;
    int rc = validate_scene(sc);
    if (rc != PGX_OK) return rc;
    if (cfg && cfg->block_size != 0 && cfg->block_size != 64 && cfg->block_size != 128 && cfg->block_size != 256)
        return fail(PGX_ERR_INVALID, "block_size must be 0, 64, 128 or 256 (got %d)", cfg->block_size);
    return guarded([&] {
        auto* h = new pgx_world();
        World& w = h->w;
        try {
            if (cfg && cfg->block_size) w.tune.solveBlock = cfg->block_size;
            w.device = cfg ? cfg->device : 0;
            PGX_HIP(hipSetDevice(w.device));
            PGX_HIP(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
            world_init(w, cfg, sc);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        return PGX_OK;
    });
}

static void release_world(World& w) {
    DBuf<float4>* f4s[] = {&w.bpos, &w.bquat, &w.dvel, &w.dang, &w.dcom, &w.dinvI, &w.bvel, &w.pvel,
                           &w.pang, &w.winvI, &w.cparams, &w.clpos, &w.clquat, &w.cmat, &w.cbmin, &w.cbmax,
                           &w.leafMin, &w.leafMax, &w.nodeMin, &w.nodeMax, &w.npNormal, &w.npPts, &w.sN, &w.sR0,
                           &w.sR1, &w.sRow, &w.sLam, &w.cacheLp[0], &w.cacheLp[1]};
    for (auto* b : f4s) b->release();
    DBuf<int32_t>* i32s[] = {&w.cbody, &w.cworld, &w.clocal, &w.cflags, &w.ctype, &w.cindex, &w.worldColStart,
                             &w.colOfLocal, &w.mvals, &w.leafFlags, &w.leafParent, &w.leafEsc, &w.nodeChild,
                             &w.nodeParent, &w.nodeEsc, &w.nodeFlags, &w.nodeCounter, &w.worldRoot, &w.worldStop,
                             &w.pvals, &w.counters, &w.npNum, &w.npIndex, &w.mfPair, &w.mfB0, &w.mfB1, &w.mfColor,
                             &w.colorPerm, &w.colorRank, &w.colorCount, &w.worldMStart, &w.sB0, &w.sB1, &w.sFlags, &w.sA0,
                             &w.sA1, &w.sCanon, &w.cacheNum[0], &w.cacheNum[1], &w.sortT.valt, &w.scanT.sums,
                             &w.dynCols, &w.staticCols, &w.worldStaticStart, &w.refitReady};
    for (auto* b : i32s) b->release();
    DBuf<uint64_t>* u64s[] = {&w.mkeys, &w.pkeys, &w.mfKey, &w.claim, &w.usedMask, &w.colorKeys, &w.cacheKey[0],
                              &w.cacheKey[1], &w.nonColliding, &w.sortT.kalt};
    for (auto* b : u64s) b->release();
    w.bentity.release(); w.centity.release(); w.worldBoundsI.release(); w.sortT.hist.release(); w.sortT.scan.sums.release();
    w.nodeRange.release(); w.sSoft.release();
    DBuf<int32_t>* ji[] = {&w.jType, &w.jA0, &w.jA1, &w.jB0, &w.jB1, &w.jRowFlags, &w.jNRows, &w.jOrder};
    for (auto* b : ji) b->release();
    DBuf<float4>* jf[] = {&w.jP0, &w.jQ0, &w.jP1, &w.jQ1, &w.jDrive, &w.jR0, &w.jR1, &w.jRow};
    for (auto* b : jf) b->release();
    world_release_graph(w);
    for (auto& e : w.qEv)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : w.ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : w.solveEv) (void)hipEventDestroy(e);
    w.solveEv.clear();
    w.colorPoints.release();
    if (w.stream) (void)hipStreamDestroy(w.stream);
    w.stream = nullptr;
}

int pgx_world_destroy(pgx_world* h) {
    if (!h) return PGX_OK;
    int rc = guarded([&] {
        PGX_HIP(hipSetDevice(h->w.device));
        if (h->w.stream) PGX_HIP(hipStreamSynchronize(h->w.stream));
        return PGX_OK;
    });
    release_world(h->w);
    delete h;
    return rc;
}

int pgx_world_set_contact_filter(pgx_world* h, int n_classes, const uint8_t* table, const int32_t* collider_class) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    World& w = h->w;
    if (n_classes == 0 || !table) {
        w.filterK = 0;
        return PGX_OK;
    }
    if (n_classes < 0 || n_classes > PGX_MAX_FILTER_CLASSES)
        return fail(PGX_ERR_INVALID, "n_classes must be in [0, %d]", PGX_MAX_FILTER_CLASSES);
    if (!collider_class && w.C > 0) return fail(PGX_ERR_INVALID, "collider_class is NULL");
    for (int c = 0; c < w.C; ++c)
        if (collider_class[c] < 0 || collider_class[c] >= n_classes)
            return fail(PGX_ERR_INVALID, "collider %d has class %d outside [0, %d)", c, collider_class[c], n_classes);
    return guarded([&] {
        PGX_HIP(hipSetDevice(w.device));
        const size_t nt = (size_t)n_classes * n_classes;
        std::vector<uint8_t> t(table, table + nt);
        bool anyTrigger = false;
        for (uint8_t& v : t) { v = v ? 1 : 0; anyTrigger = anyTrigger || v; }
        // colliders keep descriptor order on the device
        std::vector<int32_t> cls(std::max(w.C, 1), 0);
        for (int c = 0; c < w.C; ++c) cls[c] = collider_class[c];
        w.filterClass.alloc(cls.size());
        w.filterTable.alloc(nt);
        PGX_HIP(hipMemcpyAsync(w.filterClass.p, cls.data(), sizeof(int32_t) * cls.size(), hipMemcpyHostToDevice,
                               w.stream));
        PGX_HIP(hipMemcpyAsync(w.filterTable.p, t.data(), nt, hipMemcpyHostToDevice, w.stream));
        if (anyTrigger) world_enable_triggers(w);
        PGX_HIP(hipStreamSynchronize(w.stream));
        w.filterK = n_classes;
        return PGX_OK;
    });
}

int pgx_world_set_params(pgx_world* h, int substeps, int iterations, float gravity) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    if (substeps <= 0 || iterations < 0) return fail(PGX_ERR_INVALID, "substeps must be > 0, iterations >= 0");
    h->w.substeps = substeps;
    h->w.iterations = iterations;
    h->w.g = gravity;
    return PGX_OK;
}

int pgx_world_set_can_collide(pgx_world* h, uint32_t e0, uint32_t e1, int canCollide) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    uint64_t k = ((uint64_t)std::min(e0, e1) << 32) | std::max(e0, e1);
    auto& v = h->w.hostNonColliding;
    auto it = std::lower_bound(v.begin(), v.end(), k);
    bool present = it != v.end() && *it == k;
    if (!canCollide && !present) v.insert(it, k);
    if (canCollide && present) v.erase(it);
    h->w.nonCollidingDirty = true;
    return PGX_OK;
}

}  // extern "C"

namespace pgx {

// The rows a joint declares, in creation order, as packed 4-bit flags (src/Joints/*.cpp
// getSolverDesc). A prismatic joint's limit row is decided on the device every step; the
// host layout holds its fixed rows and drive.
static int joint_layout(int type, bool drive, int* packed) {
    int fl[7], nr = 0;
    if (type == PGX_JOINT_CUSTOM) {   // declared by the layout callback at each step
        *packed = 0;
        return 0;
    }
    if (type == PGX_JOINT_PRISMATIC) {
        fl[nr++] = 0; fl[nr++] = 0; fl[nr++] = 2; fl[nr++] = 2; fl[nr++] = 2;
        if (drive) fl[nr++] = 1;                               // SOFT
    } else if (type == PGX_JOINT_GEAR) {
        fl[nr++] = 2;
    } else {
        for (int r = 0; r < 3; ++r) fl[nr++] = 0;              // point-to-point: NONE
        if (type == PGX_JOINT_UNIVERSAL) fl[nr++] = 2;         // ANGULAR
        if (type == PGX_JOINT_REVOLUTE) { fl[nr++] = 2; fl[nr++] = 2; if (drive) fl[nr++] = 2 | 4; }
        if (type == PGX_JOINT_FIXED) for (int r = 0; r < 3; ++r) fl[nr++] = 2;
        if (type == PGX_JOINT_SERVO) { fl[nr++] = 2; fl[nr++] = 2; fl[nr++] = 2 | 1; }
    }
    *packed = 0;
    for (int r = 0; r < nr; ++r) *packed |= fl[r] << (4 * r);
    return nr;
}

static void joint_params(World& w, int i, const pgx_joint_desc& d) {
    w.hjDrive[i] = make_float4(d.drive_enabled ? 1.f : 0.f, d.drive_velocity, d.drive_max_torque, 0.f);
    w.hjPar0[i] = make_float4(d.lower_limit, d.upper_limit, d.target, d.drive_stiffness);
    w.hjPar1[i] = make_float4(d.drive_damping, d.gear_ratio, 0.f, 0.f);
    int packed;
    w.hjNRows[i] = joint_layout(w.hjType[i], d.drive_enabled != 0, &packed);
    w.hjRowFlags[i] = packed;
}

// Grows a per-joint device buffer to n entries, keeping the first `keep` (state that lives
// on the device, such as the gear angles) and zeroing the rest.
template <class T>
static void grow_keep(DBuf<T>& b, int keep, int n, hipStream_t s) {
    T* np = nullptr;
    PGX_HIP(hipMalloc((void**)&np, sizeof(T) * std::max(n, 1)));
    PGX_HIP(hipMemsetAsync(np, 0, sizeof(T) * std::max(n, 1), s));
    if (keep > 0 && b.p) PGX_HIP(hipMemcpyAsync(np, b.p, sizeof(T) * keep, hipMemcpyDeviceToDevice, s));
    PGX_HIP(hipStreamSynchronize(s));
    b.release();
    b.p = np;
    b.n = (size_t)std::max(n, 1);
}

constexpr int ISLAND_MAX = 64;   // joints of one colour in one island: one wave's lanes
constexpr int ISLAND_MIN = 1;

// The live joints grouped by colour, creation order inside a colour, and the per-joint
// arrays, to the device.
static void upload_joints(World& w, int keepState) {
    std::vector<int32_t> order;
    order.reserve(w.J);
    w.jColorStart.assign(10, 0);
    w.jRowsTotal = 0;
    for (int j = 0; j < w.J; ++j)
        if (w.hjAlive[j]) { ++w.jColorStart[w.hjColor[j] + 1]; w.jRowsTotal += w.hjNRows[j]; }
    for (int c = 0; c < 9; ++c) w.jColorStart[c + 1] += w.jColorStart[c];
    order.assign(w.jColorStart[9], 0);
    std::vector<int> fillp(w.jColorStart.begin(), w.jColorStart.end() - 1);
    for (int j = 0; j < w.J; ++j)
        if (w.hjAlive[j]) order[fillp[w.hjColor[j]]++] = j;
    // islands of the SIMD colours (0-7): joints joined through dynamic bodies; each island's
    // joints by colour, in the colour order above inside a colour. One launch per joint pass
    // solves them all (k_joint_islands) when no island holds more than ISLAND_MAX joints of
    // one colour; per-colour launches otherwise.
    {
        std::vector<int> parent(w.D > 0 ? w.D : 1);
        std::iota(parent.begin(), parent.end(), 0);
        std::function<int(int)> root = [&](int x) {
            while (parent[x] != x) x = parent[x] = parent[parent[x]];
            return x;
        };
        for (int k = 0; k < w.jColorStart[8]; ++k) {
            const int j = order[k], b0 = w.hjB0[j], b1 = w.hjB1[j];
            if (b0 >= 0 && b1 >= 0) parent[root(b0)] = root(b1);
        }
        // island of each joint: its dynamic body's root, or a fresh island for a joint
        // between two non-dynamic bodies
        std::vector<int> islandOf(w.jColorStart[8]);
        std::unordered_map<int, int> idOfRoot;
        int nIsl = 0;
        for (int k = 0; k < w.jColorStart[8]; ++k) {
            const int j = order[k], b = w.hjB0[j] >= 0 ? w.hjB0[j] : w.hjB1[j];
            if (b < 0) { islandOf[k] = nIsl++; continue; }
            auto it = idOfRoot.emplace(root(b), nIsl);
            if (it.second) ++nIsl;
            islandOf[k] = it.first->second;
        }
        std::vector<int32_t> cnt((size_t)nIsl * 8, 0);   // joints per island and colour
        for (int k = 0; k < w.jColorStart[8]; ++k) ++cnt[(size_t)islandOf[k] * 8 + w.hjColor[order[k]]];
        const int most = cnt.empty() ? 0 : *std::max_element(cnt.begin(), cnt.end());
        w.jIslands = 0;
        if (nIsl >= ISLAND_MIN && most <= ISLAND_MAX) {
            std::vector<int32_t> cs((size_t)nIsl * 9), cursor((size_t)nIsl * 8), list(w.jColorStart[8]);
            int at = 0;
            for (int i = 0; i < nIsl; ++i) {
                for (int c = 0; c < 8; ++c) {
                    cs[(size_t)i * 9 + c] = cursor[(size_t)i * 8 + c] = at;
                    at += cnt[(size_t)i * 8 + c];
                }
                cs[(size_t)i * 9 + 8] = at;
            }
            for (int k = 0; k < w.jColorStart[8]; ++k) {   // colour order, creation order inside a colour
                const int j = order[k];
                list[cursor[(size_t)islandOf[k] * 8 + w.hjColor[j]]++] = j;
            }
            w.jIslands = nIsl;
            upload(w.jIslandCS, cs, w.stream);
            upload(w.jIslandJoint, list, w.stream);
        }
    }
    hipStream_t s = w.stream;
    w.jCap = w.J;
    upload(w.jType, w.hjType, s); upload(w.jA0, w.hjA0, s); upload(w.jA1, w.hjA1, s);
    upload(w.jB0, w.hjB0, s); upload(w.jB1, w.hjB1, s); upload(w.jRowFlags, w.hjRowFlags, s);
    upload(w.jNRows, w.hjNRows, s); upload(w.jOrder, order, s);
    upload(w.jP0, w.hjP0, s); upload(w.jQ0, w.hjQ0, s); upload(w.jP1, w.hjP1, s); upload(w.jQ1, w.hjQ1, s);
    upload(w.jDrive, w.hjDrive, s); upload(w.jPar0, w.hjPar0, s); upload(w.jPar1, w.hjPar1, s);
    if (keepState < w.J) {
        grow_keep(w.jGear, keepState, w.J, s);
        grow_keep(w.jGearInit, keepState, w.J, s);
    }
    w.jR0.alloc(w.J); w.jR1.alloc(w.J);
    w.jRow.alloc((size_t)JOINT_MAX_ROWS * 8 * w.J);
    // custom joints in the order the reference calls them: colour by colour, creation
    // order inside a colour (src/Physecs.cpp:322-325,431-433)
    w.custList.clear();
    for (int32_t j : order)
        if (w.hjType[j] == PGX_JOINT_CUSTOM) w.custList.push_back(j);
    w.custRowsHost.assign(w.custList.size() * JOINT_MAX_ROWS, pgx_joint_row{});
    if (!w.custList.empty()) upload(w.custListDev, w.custList, s);
    PGX_HIP(hipStreamSynchronize(s));
}

// Scene::createJoint for a batch of built-in joints (actor indices of the descriptor the
// world was initialised from); colours as Scene::addJoint assigns them.
void world_add_joints(World& w, int n, const pgx_joint_desc* js) {
    if (w.hjBits.empty()) w.hjBits.assign(w.A, 0);
    int before = w.J;
    for (int i = 0; i < n; ++i) {
        const pgx_joint_desc& d = js[i];
        int a0 = w.actorToBody[d.actor0], a1 = w.actorToBody[d.actor1];
        // Scene::addJoint (src/Physecs.cpp:690-710)
        uint32_t e0 = w.hostEntity[a0], e1 = w.hostEntity[a1];
        uint64_t k = ((uint64_t)std::min(e0, e1) << 32) | std::max(e0, e1);
        auto it = std::lower_bound(w.hostNonColliding.begin(), w.hostNonColliding.end(), k);
        if (it == w.hostNonColliding.end() || *it != k) w.hostNonColliding.insert(it, k);
        w.nonCollidingDirty = true;
        unsigned int inv = ~(unsigned int)(w.hjBits[a0] | w.hjBits[a1]);
        int color = 0;
        while (!((inv >> color) & 1u)) ++color;
        w.hjBits[a0] = (unsigned char)(w.hjBits[a0] | (1u << color));
        w.hjBits[a1] = (unsigned char)(w.hjBits[a1] | (1u << color));
        auto dyn = [&](int b) { return b < w.D && !w.hostKinematic[b] ? b : -1; };
        w.hjType.push_back(d.type); w.hjA0.push_back(a0); w.hjA1.push_back(a1);
        w.hjB0.push_back(dyn(a0)); w.hjB1.push_back(dyn(a1));
        w.hjColor.push_back(color);
        w.hjAlive.push_back(1);
        w.hjP0.push_back(make_float4(d.anchor0[0], d.anchor0[1], d.anchor0[2], 0.f));
        w.hjQ0.push_back(make_float4(d.anchor0[3], d.anchor0[4], d.anchor0[5], d.anchor0[6]));
        w.hjP1.push_back(make_float4(d.anchor1[0], d.anchor1[1], d.anchor1[2], 0.f));
        w.hjQ1.push_back(make_float4(d.anchor1[3], d.anchor1[4], d.anchor1[5], d.anchor1[6]));
        w.hjDrive.emplace_back(); w.hjPar0.emplace_back(); w.hjPar1.emplace_back();
        w.hjRowFlags.push_back(0); w.hjNRows.push_back(0);
        joint_params(w, (int)w.hjType.size() - 1, d);
    }
    w.J = (int)w.hjType.size();
    upload_joints(w, before);
}

}  // namespace pgx

extern "C" {

int pgx_world_add_joints(pgx_world* h, int n, const pgx_joint_desc* js) {
    if (!h || (n > 0 && !js)) return fail(PGX_ERR_INVALID, "NULL argument");
    World& w = h->w;
    for (int i = 0; i < n; ++i) {
        const pgx_joint_desc& d = js[i];
        if (d.type < PGX_JOINT_SPHERICAL || d.type > PGX_JOINT_CUSTOM)
            return fail(PGX_ERR_INVALID, "joint %d: unknown type %d", i, d.type);
        if (d.actor0 < 0 || d.actor0 >= w.A || d.actor1 < 0 || d.actor1 >= w.A)
            return fail(PGX_ERR_INVALID, "joint %d: actor out of range", i);
    }
    return guarded([&] {
        PGX_HIP(hipSetDevice(w.device));
        world_add_joints(w, n, js);
        return PGX_OK;
    });
}

int pgx_world_set_joint_params(pgx_world* h, int index, const pgx_joint_desc* d) {
    if (!h || !d) return fail(PGX_ERR_INVALID, "NULL argument");
    World& w = h->w;
    if (index < 0 || index >= w.J || !w.hjAlive[index]) return fail(PGX_ERR_INVALID, "joint %d does not exist", index);
    if (d->type != w.hjType[index]) return fail(PGX_ERR_INVALID, "joint %d: the type cannot change", index);
    return guarded([&] {
        PGX_HIP(hipSetDevice(w.device));
        PGX_HIP(hipStreamSynchronize(w.stream));   // the previous step reads the joint arrays
        joint_params(w, index, *d);
        upload_joints(w, w.J);
        return PGX_OK;
    });
}

int pgx_world_set_joint_callbacks(pgx_world* h, pgx_joint_layout_fn layout, pgx_joint_rows_fn rows, void* user) {
    if (!h) return fail(PGX_ERR_INVALID, "NULL argument");
    if ((layout == nullptr) != (rows == nullptr)) return fail(PGX_ERR_INVALID, "give both joint functions or neither");
    World& w = h->w;
    w.custLayout = layout;
    w.custRows = rows;
    w.custUser = layout ? user : nullptr;
    return PGX_OK;
}

int pgx_world_destroy_joint(pgx_world* h, int index) {
    if (!h) return fail(PGX_ERR_INVALID, "NULL argument");
    World& w = h->w;
    if (index < 0 || index >= w.J || !w.hjAlive[index]) return fail(PGX_ERR_INVALID, "joint %d does not exist", index);
    return guarded([&] {
        PGX_HIP(hipSetDevice(w.device));
        PGX_HIP(hipStreamSynchronize(w.stream));
        // Scene::destroyJoint (src/Physecs.cpp:712-725)
        w.hjAlive[index] = 0;
        int a0 = w.hjA0[index], a1 = w.hjA1[index];
        uint32_t e0 = w.hostEntity[a0], e1 = w.hostEntity[a1];
        uint64_t k = ((uint64_t)std::min(e0, e1) << 32) | std::max(e0, e1);
        auto it = std::lower_bound(w.hostNonColliding.begin(), w.hostNonColliding.end(), k);
        if (it != w.hostNonColliding.end() && *it == k) w.hostNonColliding.erase(it);
        w.nonCollidingDirty = true;
        unsigned char bit = (unsigned char)(1u << w.hjColor[index]);   // colour 8 sets no bit
        w.hjBits[a0] = (unsigned char)(w.hjBits[a0] & ~bit);
        w.hjBits[a1] = (unsigned char)(w.hjBits[a1] & ~bit);
        upload_joints(w, w.J);
        return PGX_OK;
    });
}

}  // extern "C"

namespace pgx {

// Everything a step carries into the next: body poses and velocities, collider bounds
// (the margin rule), the contact cache of the last step, gear joint angles, and the host
// counters and statistics. Autotuning steps the world from a snapshot and restores it.
struct Snapshot {
    DBuf<float4> pos, quat, vel, ang, bmin, bmax, lp, gear;
    DBuf<uint64_t> key, trigKeys[2];
    DBuf<int32_t> tri, num, gearInit, trigN;
    int trigCur = 0;
    int cacheCur = 0, cacheCount = 0, lastRounds = 0, lastPairs = 0, lastManifolds = 0, lastColors = 0;
    int lastMeshSlots = 0;
    long statSteps = 0, statSolveLaunches = 0, lastPointsHost = 0, statBatchLaunches = 0;
    double statSolveUs = 0, statSolveBytes = 0, statStepUs = 0, statBatchUs = 0, statBatchBytes = 0;
};

template <class T>
static void d2d(T* dst, const T* src, size_t n, hipStream_t s) {
    if (n) PGX_HIP(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
}

static void take_snapshot(World& w, Snapshot& z) {
    hipStream_t s = w.stream;
    const int cur = w.cacheCur, n = w.cacheCount[cur];
    z.pos.alloc(w.A); z.quat.alloc(w.A); z.vel.alloc(w.D); z.ang.alloc(w.D); z.bmin.alloc(w.C); z.bmax.alloc(w.C);
    d2d(z.pos.p, w.bpos.p, w.A, s); d2d(z.quat.p, w.bquat.p, w.A, s);
    d2d(z.vel.p, w.dvel.p, w.D, s); d2d(z.ang.p, w.dang.p, w.D, s);
    d2d(z.bmin.p, w.cbmin.p, w.C, s); d2d(z.bmax.p, w.cbmax.p, w.C, s);
    z.key.alloc(n); z.tri.alloc(n); z.num.alloc(n); z.lp.alloc(4 * (size_t)n);
    d2d(z.key.p, w.cacheKey[cur].p, n, s); d2d(z.tri.p, w.cacheTri[cur].p, n, s); d2d(z.num.p, w.cacheNum[cur].p, n, s);
    for (int j = 0; j < 4; ++j) d2d(z.lp.p + (size_t)j * n, w.cacheLp[cur].p + (size_t)j * w.manifoldCap, n, s);
    if (w.J > 0) {
        z.gear.alloc(w.J); z.gearInit.alloc(w.J);
        d2d(z.gear.p, w.jGear.p, w.J, s); d2d(z.gearInit.p, w.jGearInit.p, w.J, s);
    }
    if (w.hasTriggers) {   // the trigger lists, so that the events after a restore are unchanged
        z.trigN.alloc(4); d2d(z.trigN.p, w.trigN.p, 4, s);
        for (int k = 0; k < 2; ++k) { z.trigKeys[k].alloc(w.pairCap); d2d(z.trigKeys[k].p, w.trigKeys[k].p, w.pairCap, s); }
        z.trigCur = w.trigCur;
    }
    z.cacheCur = cur; z.cacheCount = n;
    z.lastRounds = w.lastRounds; z.lastPairs = w.lastPairs; z.lastManifolds = w.lastManifolds;
    z.lastColors = w.lastColors; z.lastMeshSlots = w.lastMeshSlots; z.lastPointsHost = w.lastPointsHost;
    z.statSteps = w.statSteps; z.statSolveLaunches = w.statSolveLaunches; z.statSolveUs = w.statSolveUs;
    z.statSolveBytes = w.statSolveBytes; z.statStepUs = w.statStepUs;
    z.statBatchLaunches = w.statBatchLaunches; z.statBatchUs = w.statBatchUs; z.statBatchBytes = w.statBatchBytes;
    PGX_HIP(hipStreamSynchronize(s));
}

static void restore_snapshot(World& w, const Snapshot& z) {
    hipStream_t s = w.stream;
    const int cur = z.cacheCur, n = z.cacheCount;
    d2d(w.bpos.p, z.pos.p, w.A, s); d2d(w.bquat.p, z.quat.p, w.A, s);
    d2d(w.dvel.p, z.vel.p, w.D, s); d2d(w.dang.p, z.ang.p, w.D, s);
    d2d(w.cbmin.p, z.bmin.p, w.C, s); d2d(w.cbmax.p, z.bmax.p, w.C, s);
    d2d(w.cacheKey[cur].p, z.key.p, n, s); d2d(w.cacheTri[cur].p, z.tri.p, n, s); d2d(w.cacheNum[cur].p, z.num.p, n, s);
    for (int j = 0; j < 4; ++j) d2d(w.cacheLp[cur].p + (size_t)j * w.manifoldCap, z.lp.p + (size_t)j * n, n, s);
    if (w.J > 0) { d2d(w.jGear.p, z.gear.p, w.J, s); d2d(w.jGearInit.p, z.gearInit.p, w.J, s); }
    if (w.hasTriggers) {
        d2d(w.trigN.p, z.trigN.p, 4, s);
        for (int k = 0; k < 2; ++k) d2d(w.trigKeys[k].p, z.trigKeys[k].p, w.pairCap, s);
        w.trigCur = z.trigCur;
    }
    w.cacheCur = cur; w.cacheCount[cur] = n;
    w.lastRounds = z.lastRounds; w.lastPairs = z.lastPairs; w.lastManifolds = z.lastManifolds;
    w.lastColors = z.lastColors; w.lastMeshSlots = z.lastMeshSlots; w.lastPointsHost = z.lastPointsHost;
    w.statSteps = z.statSteps; w.statSolveLaunches = z.statSolveLaunches; w.statSolveUs = z.statSolveUs;
    w.statSolveBytes = z.statSolveBytes; w.statStepUs = z.statStepUs;
    w.statBatchLaunches = z.statBatchLaunches; w.statBatchUs = z.statBatchUs; w.statBatchBytes = z.statBatchBytes;
    PGX_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ state save / load
// Everything a world carries from one step to the next, as one host byte string: poses,
// velocities, broadphase bounds (with their margins), the contact cache (the reference's
// contactCache, include/Physecs/Physecs.h:167), the gear joints' angle state
// (GearJointData, include/Physecs/Joints/GearJoint.h:7-14), the trigger lists
// (triggerCache) and the last step's sizes that the next step's launches start from, plus
// each collider's identity (world, entity, index within its entity). The structure
// (descriptor, joints, setCanCollide pairs, parameters) is not in it: world_load_state
// loads a state saved by a world created from the same descriptor and joints, and
// world_load_state_matched one saved by a world whose colliders differ.
namespace {
constexpr uint32_t STATE_MAGIC = 0x53584750u;   // "PGXS"
constexpr int32_t STATE_VERSION = 2;
struct StateHeader {
    uint32_t magic;
    int32_t version, A, D, C, J, W, pairCap, manifoldCap;
    int32_t cacheCount, hasTriggers, trigCur, trigN[4];
    int32_t lastPairs, lastManifolds, lastColors, lastRounds, lastPoints, lastMeshSlots;
    int32_t localBits, reserved;
};
// a collider's identity: the reference keys its contact and trigger caches by entity and
// collider index (CollisionPair / ContactPair, include/Physecs/Physecs.h:112-125)
struct ColId {
    int32_t world;
    uint32_t entity;
    int32_t index;
};

// the device sections in state order
enum Section { S_POS, S_QUAT, S_VEL, S_ANG, S_BMIN, S_BMAX, S_KEY, S_TRI, S_NUM, S_LP0, S_LP1, S_LP2, S_LP3, S_GEAR,
               S_GEAR_INIT, S_TRIG0, S_TRIG1, S_COUNT };

void section_bytes(const StateHeader& h, size_t out[S_COUNT]) {
    const size_t n = (size_t)h.cacheCount;
    out[S_POS] = out[S_QUAT] = 16 * (size_t)h.A;
    out[S_VEL] = out[S_ANG] = 16 * (size_t)h.D;
    out[S_BMIN] = out[S_BMAX] = 16 * (size_t)h.C;
    out[S_KEY] = 8 * n; out[S_TRI] = 4 * n; out[S_NUM] = 4 * n;
    out[S_LP0] = out[S_LP1] = out[S_LP2] = out[S_LP3] = 16 * n;
    out[S_GEAR] = 16 * (size_t)h.J; out[S_GEAR_INIT] = 4 * (size_t)h.J;
    out[S_TRIG0] = h.hasTriggers ? 8 * (size_t)h.trigN[1] : 0;
    out[S_TRIG1] = h.hasTriggers ? 8 * (size_t)h.trigN[3] : 0;
}

// byte offset of the first device section
size_t ids_end(const StateHeader& h) { return sizeof(StateHeader) + sizeof(ColId) * (size_t)h.C; }

size_t state_bytes(const StateHeader& h) {
    size_t sz[S_COUNT], n = ids_end(h);
    section_bytes(h, sz);
    for (size_t b : sz) n += b;
    return n;
}

StateHeader state_header(World& w) {
    StateHeader h{};
    h.magic = STATE_MAGIC; h.version = STATE_VERSION;
    h.A = w.A; h.D = w.D; h.C = w.C; h.J = w.J; h.W = w.W; h.pairCap = w.pairCap; h.manifoldCap = w.manifoldCap;
    h.cacheCount = w.cacheCount[w.cacheCur];
    h.hasTriggers = w.hasTriggers ? 1 : 0;
    h.trigCur = w.trigCur;
    if (w.hasTriggers) {
        PGX_HIP(hipMemcpyAsync(h.trigN, w.trigN.p, sizeof h.trigN, hipMemcpyDeviceToHost, w.stream));
        PGX_HIP(hipStreamSynchronize(w.stream));
    }
    h.lastPairs = w.lastPairs; h.lastManifolds = w.lastManifolds; h.lastColors = w.lastColors;
    h.lastRounds = w.lastRounds; h.lastPoints = w.lastPoints; h.lastMeshSlots = w.lastMeshSlots;
    h.localBits = w.localBits;
    return h;
}

// The device sections in order: (device pointer, bytes); `save` copies them out, `load` in.
template <class F>
void state_sections(World& w, const StateHeader& h, F&& f) {
    const int cur = w.cacheCur;
    void* dev[S_COUNT] = {w.bpos.p, w.bquat.p, w.dvel.p, w.dang.p, w.cbmin.p, w.cbmax.p, w.cacheKey[cur].p,
                          w.cacheTri[cur].p, w.cacheNum[cur].p, w.cacheLp[cur].p,
                          w.cacheLp[cur].p + (size_t)w.manifoldCap, w.cacheLp[cur].p + 2 * (size_t)w.manifoldCap,
                          w.cacheLp[cur].p + 3 * (size_t)w.manifoldCap, w.jGear.p, w.jGearInit.p,
                          h.hasTriggers ? (void*)w.trigKeys[0].p : nullptr,
                          h.hasTriggers ? (void*)w.trigKeys[1].p : nullptr};
    size_t sz[S_COUNT];
    section_bytes(h, sz);
    for (int k = 0; k < S_COUNT; ++k) f(dev[k], sz[k]);
}

// each descriptor collider's identity, from the host copies of the world's layout
std::vector<ColId> collider_ids(const World& w) {
    std::vector<ColId> ids((size_t)w.C);
    for (int wd = 0; wd < w.W; ++wd)
        for (int l = w.hostWorldColStart[wd]; l < w.hostWorldColStart[wd + 1]; ++l) ids[w.hostColOfLocal[l]].world = wd;
    for (int c = 0; c < w.C; ++c) {
        ids[c].entity = w.hostEntity[w.hostColBody[c]];
        ids[c].index = w.hostColIndex[c];
    }
    return ids;
}

StateHeader read_header(const void* buf, size_t bytes) {
    StateHeader h;
    if (bytes < sizeof h) throw PgxError{PGX_ERR_INVALID, "state buffer shorter than its header"};
    std::memcpy(&h, buf, sizeof h);
    if (h.magic != STATE_MAGIC || h.version != STATE_VERSION)
        throw PgxError{PGX_ERR_INVALID, "not a saved world state (magic or version)"};
    if (h.A < 0 || h.D < 0 || h.C < 0 || h.J < 0 || h.W < 1 || h.cacheCount < 0 || h.cacheCount > h.manifoldCap ||
        h.trigCur < 0 || h.trigCur > 1 || h.localBits < 1 || h.localBits > 31)
        throw PgxError{PGX_ERR_INVALID, "state header out of range"};
    for (int k = 0; k < 4; ++k)
        if (h.trigN[k] < 0 || h.trigN[k] > h.pairCap) throw PgxError{PGX_ERR_INVALID, "trigger counts out of range"};
    if (bytes < state_bytes(h)) throw PgxError{PGX_ERR_INVALID, "state buffer truncated"};
    return h;
}

// The last step's pair, manifold and mesh-slot lists live in buffers the state does not
// carry: after a load there is no last step to report (counts and manifolds read empty
// until the next step). The colouring-round count only sizes the next step's first batch
// of rounds.
void reset_last_step(World& w, const StateHeader& h) {
    w.lastPairs = 0; w.lastManifolds = 0; w.lastColors = 0; w.lastPoints = 0; w.lastMeshSlots = 0;
    w.lastPointsHost = 0;
    w.lastRounds = std::max(0, std::min(h.lastRounds, 1024));
}
}  // namespace

size_t world_state_size(World& w) { return state_bytes(state_header(w)); }

void world_save_state(World& w, void* buf, size_t bytes) {
    const StateHeader h = state_header(w);
    if (bytes < state_bytes(h)) throw PgxError{PGX_ERR_CAPACITY, "state buffer of " + std::to_string(bytes) +
                                                                     " bytes < " + std::to_string(state_bytes(h))};
    char* out = static_cast<char*>(buf);
    std::memcpy(out, &h, sizeof h);
    const std::vector<ColId> ids = collider_ids(w);
    if (!ids.empty()) std::memcpy(out + sizeof h, ids.data(), sizeof(ColId) * ids.size());
    size_t at = ids_end(h);
    state_sections(w, h, [&](void* dev, size_t nb) {
        if (nb) PGX_HIP(hipMemcpyAsync(out + at, dev, nb, hipMemcpyDeviceToHost, w.stream));
        at += nb;
    });
    PGX_HIP(hipStreamSynchronize(w.stream));
}

void world_load_state(World& w, const void* buf, size_t bytes) {
    const StateHeader h = read_header(buf, bytes);
    if (h.A != w.A || h.D != w.D || h.C != w.C || h.J != w.J || h.W != w.W ||
        h.hasTriggers != (w.hasTriggers ? 1 : 0) || h.localBits != w.localBits)
        throw PgxError{PGX_ERR_INVALID, "the state was saved by a world of another shape (actors, bodies, colliders, "
                                        "joints, worlds or triggers differ)"};
    if (h.cacheCount > w.manifoldCap) throw PgxError{PGX_ERR_INVALID, "state header out of range"};
    const char* in = static_cast<const char*>(buf);
    size_t at = ids_end(h);
    state_sections(w, h, [&](void* dev, size_t nb) {
        if (nb) PGX_HIP(hipMemcpyAsync(dev, in + at, nb, hipMemcpyHostToDevice, w.stream));
        at += nb;
    });
    if (w.hasTriggers) PGX_HIP(hipMemcpyAsync(w.trigN.p, h.trigN, sizeof h.trigN, hipMemcpyHostToDevice, w.stream));
    PGX_HIP(hipStreamSynchronize(w.stream));
    w.cacheCount[w.cacheCur] = h.cacheCount;
    w.trigCur = h.trigCur;
    reset_last_step(w, h);
}

// A state saved by a world whose colliders differ from this one's (Scene::addCollider and
// Scene::clearColliders, src/Physecs.cpp:725-752, and any other rebuild of a scene): the
// reference's caches are keyed by (entity, collider index), which neither edit changes for
// the other colliders, so every cache entry, trigger pair and bounds box goes to the
// collider of this world with the same world, entity and index, and what has no such
// collider is dropped. The contact cache stays sorted by its (new) keys, a pair's entries
// in their saved order. Poses and velocities stay as this world was created. The gear
// joints' state is loaded when both worlds have the same number of joints. Trigger pairs of
// the last step that lose a collider are written to `gone` (entity0, index0, entity1,
// index1; the lower entity first), as are all of them when this world keeps no trigger
// lists: the reference reports them as exits at the next step (src/Physecs.cpp:545-551).
// Nothing is changed when the state is rejected or `gone` is too small.
void world_load_state_matched(World& w, const void* buf, size_t bytes, int goneCap, uint32_t* gone, int32_t* nGone) {
    const StateHeader h = read_header(buf, bytes);
    if (h.W != w.W) throw PgxError{PGX_ERR_INVALID, "the state was saved by a world with another number of worlds"};
    const char* in = static_cast<const char*>(buf);
    std::vector<ColId> old((size_t)h.C);
    if (h.C) std::memcpy(old.data(), in + sizeof h, sizeof(ColId) * old.size());
    size_t sz[S_COUNT], off[S_COUNT];
    section_bytes(h, sz);
    off[0] = ids_end(h);
    for (int k = 1; k < S_COUNT; ++k) off[k] = off[k - 1] + sz[k - 1];

    // the saver's world-local collider order (descriptor order within each world)
    std::vector<int32_t> oldStart(h.W + 1, 0), oldOfLocal((size_t)h.C);
    for (const ColId& c : old) {
        if (c.world < 0 || c.world >= h.W) throw PgxError{PGX_ERR_INVALID, "state collider world out of range"};
        ++oldStart[c.world + 1];
    }
    for (int i = 0; i < h.W; ++i) oldStart[i + 1] += oldStart[i];
    {
        std::vector<int32_t> fillp(oldStart.begin(), oldStart.end() - 1);
        for (int c = 0; c < h.C; ++c) oldOfLocal[fillp[old[c].world]++] = c;
    }
    // this world's colliders by identity, and their world-local indices
    const std::vector<ColId> now = collider_ids(w);
    auto idKey = [](const ColId& c) {
        return ((uint64_t)(uint32_t)c.world << 32) ^ ((uint64_t)c.entity << 8) ^ (uint64_t)(uint32_t)c.index;
    };
    std::unordered_multimap<uint64_t, int32_t> byId;
    byId.reserve(now.size());
    for (int c = 0; c < w.C; ++c) byId.emplace(idKey(now[c]), c);
    std::vector<int32_t> local((size_t)w.C);
    for (int wd = 0; wd < w.W; ++wd)
        for (int l = w.hostWorldColStart[wd]; l < w.hostWorldColStart[wd + 1]; ++l)
            local[w.hostColOfLocal[l]] = l - w.hostWorldColStart[wd];
    std::vector<int32_t> map((size_t)h.C, -1);
    for (int c = 0; c < h.C; ++c) {
        auto r = byId.equal_range(idKey(old[c]));
        for (auto it = r.first; it != r.second; ++it) {
            const ColId& n = now[it->second];
            if (n.world == old[c].world && n.entity == old[c].entity && n.index == old[c].index) {
                map[c] = it->second;
                break;
            }
        }
    }
    // a saved key (world, local of the lower entity's collider, other local) in the saver's
    // layout to this world's: false when a collider is gone
    const uint64_t oldMask = (1ull << h.localBits) - 1;
    auto decode = [&](uint64_t key, int& c0, int& c1) {
        const uint64_t wd = key >> (2 * h.localBits);
        const uint64_t l0 = (key >> h.localBits) & oldMask, l1 = key & oldMask;
        if (wd >= (uint64_t)h.W || l0 >= (uint64_t)(oldStart[wd + 1] - oldStart[wd]) ||
            l1 >= (uint64_t)(oldStart[wd + 1] - oldStart[wd]))
            throw PgxError{PGX_ERR_INVALID, "state key out of range"};
        c0 = oldOfLocal[oldStart[wd] + l0];
        c1 = oldOfLocal[oldStart[wd] + l1];
    };
    auto translate = [&](uint64_t key, uint64_t& out) {
        int c0, c1;
        decode(key, c0, c1);
        const int n0 = map[c0], n1 = map[c1];
        if (n0 < 0 || n1 < 0) return false;
        out = ((uint64_t)now[n0].world << (2 * w.localBits)) | ((uint64_t)local[n0] << w.localBits) | (uint64_t)local[n1];
        return true;
    };

    // contact cache: the entries whose colliders both remain, stably sorted by their new keys
    const size_t n = (size_t)h.cacheCount;
    std::vector<uint64_t> keyIn(n);
    std::vector<int32_t> triIn(n), numIn(n);
    std::vector<float4> lpIn(4 * n);
    if (n) {
        std::memcpy(keyIn.data(), in + off[S_KEY], sz[S_KEY]);
        std::memcpy(triIn.data(), in + off[S_TRI], sz[S_TRI]);
        std::memcpy(numIn.data(), in + off[S_NUM], sz[S_NUM]);
        for (int j = 0; j < 4; ++j) std::memcpy(lpIn.data() + j * n, in + off[S_LP0 + j], sz[S_LP0 + j]);
    }
    std::vector<std::pair<uint64_t, int32_t>> kept;   // (new key, saved entry)
    kept.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t k;
        if (translate(keyIn[i], k)) kept.push_back({k, (int32_t)i});
    }
    std::stable_sort(kept.begin(), kept.end(),
                     [](const std::pair<uint64_t, int32_t>& a, const std::pair<uint64_t, int32_t>& b) {
                         return a.first < b.first;
                     });
    if (kept.size() > (size_t)w.manifoldCap)
        throw PgxError{PGX_ERR_CAPACITY, "the carried contact cache exceeds the world's manifold capacity"};
    const size_t m = kept.size(), mc = (size_t)w.manifoldCap;
    std::vector<uint64_t> keyOut(m);
    std::vector<int32_t> triOut(m), numOut(m);
    std::vector<float4> lpOut(4 * std::max<size_t>(m, 1));
    for (size_t i = 0; i < m; ++i) {
        const int32_t s = kept[i].second;
        keyOut[i] = kept[i].first; triOut[i] = triIn[s]; numOut[i] = numIn[s];
        for (int j = 0; j < 4; ++j) lpOut[j * m + i] = lpIn[j * n + s];
    }

    // trigger lists, and the last step's trigger pairs this world will not see again
    std::vector<uint64_t> trigOut[2];
    std::vector<uint32_t> goneIds;
    if (h.hasTriggers) {
        for (int k = 0; k < 2; ++k) {
            const size_t cnt = sz[S_TRIG0 + k] / 8;
            std::vector<uint64_t> keys(cnt);
            if (cnt) std::memcpy(keys.data(), in + off[S_TRIG0 + k], sz[S_TRIG0 + k]);
            for (uint64_t key : keys) {
                uint64_t t;
                const bool ok = translate(key, t);
                if (ok && w.hasTriggers) trigOut[k].push_back(t);
                if (k == h.trigCur && !(ok && w.hasTriggers)) {
                    int c0, c1;
                    decode(key, c0, c1);
                    goneIds.insert(goneIds.end(), {old[c0].entity, (uint32_t)old[c0].index, old[c1].entity,
                                                   (uint32_t)old[c1].index});
                }
            }
        }
    }
    const int32_t ng = (int32_t)(goneIds.size() / 4);
    if (nGone) *nGone = ng;
    if (ng > 0 && (ng > goneCap || !gone))
        throw PgxError{PGX_ERR_CAPACITY, std::to_string(ng) + " trigger pairs lose a collider; the gone list holds " +
                                             std::to_string(std::max(goneCap, 0))};

    // bounds: the carried boxes of the colliders that remain, the rest as created
    std::vector<float4> bmin((size_t)w.C), bmax((size_t)w.C);
    hipStream_t s = w.stream;
    if (w.C) {
        PGX_HIP(hipMemcpyAsync(bmin.data(), w.cbmin.p, 16 * (size_t)w.C, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipMemcpyAsync(bmax.data(), w.cbmax.p, 16 * (size_t)w.C, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipStreamSynchronize(s));
    }
    for (int c = 0; c < h.C; ++c) {
        if (map[c] < 0) continue;
        std::memcpy(&bmin[map[c]], in + off[S_BMIN] + 16 * (size_t)c, 16);
        std::memcpy(&bmax[map[c]], in + off[S_BMAX] + 16 * (size_t)c, 16);
    }

    // commit
    if (w.C) {
        PGX_HIP(hipMemcpyAsync(w.cbmin.p, bmin.data(), 16 * (size_t)w.C, hipMemcpyHostToDevice, s));
        PGX_HIP(hipMemcpyAsync(w.cbmax.p, bmax.data(), 16 * (size_t)w.C, hipMemcpyHostToDevice, s));
    }
    const int cur = w.cacheCur;
    if (m) {
        PGX_HIP(hipMemcpyAsync(w.cacheKey[cur].p, keyOut.data(), 8 * m, hipMemcpyHostToDevice, s));
        PGX_HIP(hipMemcpyAsync(w.cacheTri[cur].p, triOut.data(), 4 * m, hipMemcpyHostToDevice, s));
        PGX_HIP(hipMemcpyAsync(w.cacheNum[cur].p, numOut.data(), 4 * m, hipMemcpyHostToDevice, s));
        for (int j = 0; j < 4; ++j)
            PGX_HIP(hipMemcpyAsync(w.cacheLp[cur].p + j * mc, lpOut.data() + j * m, 16 * m, hipMemcpyHostToDevice, s));
    }
    if (h.J == w.J && w.J > 0) {
        PGX_HIP(hipMemcpyAsync(w.jGear.p, in + off[S_GEAR], sz[S_GEAR], hipMemcpyHostToDevice, s));
        PGX_HIP(hipMemcpyAsync(w.jGearInit.p, in + off[S_GEAR_INIT], sz[S_GEAR_INIT], hipMemcpyHostToDevice, s));
    }
    int32_t trigN[4] = {0, 0, 0, 0};
    if (w.hasTriggers) {
        for (int k = 0; k < 2; ++k) {
            trigN[2 * k + 1] = (int32_t)trigOut[k].size();
            if (!trigOut[k].empty())
                PGX_HIP(hipMemcpyAsync(w.trigKeys[k].p, trigOut[k].data(), 8 * trigOut[k].size(),
                                       hipMemcpyHostToDevice, s));
        }
        PGX_HIP(hipMemcpyAsync(w.trigN.p, trigN, sizeof trigN, hipMemcpyHostToDevice, s));
    }
    PGX_HIP(hipStreamSynchronize(s));
    if (ng > 0) std::memcpy(gone, goneIds.data(), sizeof(uint32_t) * goneIds.size());
    w.cacheCount[cur] = (int32_t)m;
    w.trigCur = h.hasTriggers ? h.trigCur : 0;
    reset_last_step(w, h);
}

// Picks the launch shapes (and between per-colour launches and per-world passes) by
// timing trial steps from the world's current state (one
// untimed step, then trial_steps timed, wall clock with the stream drained; the better of
// two such runs), one parameter at a time, in the order below. A value replaces the
// current one only when it is at least 3% faster, so timing noise does not move the
// choice. The state is restored afterwards and between trials, so every trial steps the
// same scene. The results never depend on these parameters.
void world_autotune(World& w, float dt, int trialSteps) {
    Snapshot z;
    take_snapshot(w, z);
    const int period = w.timingPeriod;
    w.timingPeriod = 0;
    auto trial = [&]() {
        double best = 1e30;
        for (int rep = 0; rep < 2; ++rep) {
            restore_snapshot(w, z);
            world_step(w, dt);
            PGX_HIP(hipStreamSynchronize(w.stream));
            auto t0 = std::chrono::steady_clock::now();
            for (int k = 0; k < trialSteps; ++k) world_step(w, dt);
            PGX_HIP(hipStreamSynchronize(w.stream));
            best = std::min(best, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        }
        return best;
    };
    try {
        struct Knob { int* field; std::vector<int> values; };
        std::vector<Knob> knobs = {{&w.tune.colorBatch, {0, 256, 512, 1024, 2048, 4096}},
                                   {&w.tune.solveBlock, {64, 128, 256}},
                                   {&w.tune.eagerMax, {0, 8192, 32768, 131072, 1 << 30}},
                                   {&w.tune.pairsPerLane, {1, 2, 4}},
                                   {&w.tune.colorRounds, {8, 16, 32}},
                                   {&w.tune.pairSearch, {0, 1}}};
        // per-colour launches or per-world passes (the same floats either way)
        if ((w.execMode == EXEC_COLORS || w.execMode == EXEC_WORLDS) && sv_world_pass_fits(w))
            knobs.insert(knobs.begin(), Knob{&w.execMode, {EXEC_COLORS, EXEC_WORLDS}});
        double best = trial();
        for (Knob& k : knobs) {
            const int keep = *k.field;
            int bestValue = keep;
            for (int v : k.values) {
                if (v == keep) continue;
                *k.field = v;
                double t = trial();
                if (t < 0.97 * best) { best = t; bestValue = v; }
            }
            *k.field = bestValue;
        }
    } catch (...) {
        w.timingPeriod = period;
        throw;
    }
    w.timingPeriod = period;
    restore_snapshot(w, z);
}

// The substep loop as one HIP graph launch. The loop's launch sequence follows the step's
// colours, so it is captured every step; the instantiated graph is updated in place when
// the sequence has the shape of the previous step's and instantiated anew otherwise.
static void substeps_graph(World& w, int M, float dt) {
    hipStream_t s = w.stream;
    sv_presize(w);   // no allocation under capture
    PGX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    try {
        sv_substeps(w, M, dt);
    } catch (...) {
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    PGX_HIP(hipStreamEndCapture(s, &g));
    bool updated = false;
    if (w.subExec) {
        hipGraphNode_t errNode = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        updated = hipGraphExecUpdate(w.subExec, g, &errNode, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
        if (!updated) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(w.subExec);
            w.subExec = nullptr;
        }
    }
    if (updated) {
        ++w.graphUpdates;
    } else {
        hipError_t e = hipGraphInstantiate(&w.subExec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            w.subExec = nullptr;
            (void)hipGraphDestroy(g);
            PGX_HIP(e);
        }
        ++w.graphInstantiations;
    }
    (void)hipGraphDestroy(g);   // the executable graph does not refer to it
    PGX_HIP(hipGraphLaunch(w.subExec, s));
}

void world_release_graph(World& w) {
    if (w.subExec) (void)hipGraphExecDestroy(w.subExec);
    w.subExec = nullptr;
}

// one Scene::simulate(dt) of every world in w (throws on error)
void world_step(World& w, float dt) {
    PGX_HIP(hipSetDevice(w.device));
    hipStream_t s = w.stream;
    if (w.nonCollidingDirty) {
        w.nonColliding.alloc(std::max<size_t>(1, w.hostNonColliding.size()));
        if (!w.hostNonColliding.empty())
            PGX_HIP(hipMemcpyAsync(w.nonColliding.p, w.hostNonColliding.data(),
                                   sizeof(uint64_t) * w.hostNonColliding.size(), hipMemcpyHostToDevice, s));
        w.nNonColliding = (int)w.hostNonColliding.size();
        w.nonCollidingDirty = false;
    }
    // per-launch events on every timingPeriod-th step only: the steps between them run bare
    w.timing = w.timingPeriod > 0 && w.timingCounter++ % w.timingPeriod == 0;
    if (w.timing && !w.ev[0])
        for (auto& e : w.ev) PGX_HIP(hipEventCreate(&e));
    if (w.timing) PGX_HIP(hipEventRecord(w.ev[0], s));
    int nPairs = bp_find_pairs(w);
    if (w.timing) PGX_HIP(hipEventRecord(w.ev[1], s));
    np_run(w, nPairs);
    int M = sv_setup_manifolds(w, nPairs);
    if (w.timing) PGX_HIP(hipEventRecord(w.ev[2], s));
    int nColors = sv_color(w, M);
    if (w.timing) PGX_HIP(hipEventRecord(w.ev[3], s));
    // custom joints hand their rows over at every substep, which a graph cannot wait for
    if (w.graph && !w.timing && w.execMode == EXEC_COLORS && w.custList.empty()) substeps_graph(w, M, dt);
    else sv_substeps(w, M, dt);
    if (w.timing) PGX_HIP(hipEventRecord(w.ev[4], s));
    sv_finish_step(w, M);
    if (w.timing) {
        PGX_HIP(hipEventRecord(w.ev[5], s));
        PGX_HIP(hipEventSynchronize(w.ev[5]));
        PGX_HIP(hipEventElapsedTime(&w.tBroad, w.ev[0], w.ev[1]));
        PGX_HIP(hipEventElapsedTime(&w.tNarrow, w.ev[1], w.ev[2]));
        PGX_HIP(hipEventElapsedTime(&w.tColor, w.ev[2], w.ev[3]));
        PGX_HIP(hipEventElapsedTime(&w.tSolve, w.ev[3], w.ev[4]));
        PGX_HIP(hipEventElapsedTime(&w.tTotal, w.ev[0], w.ev[5]));
        // the substep-loop kernel: duration (HIP events on the world's stream) and the
        // algorithmic bytes of this step's substep loop
        std::vector<int32_t> pts(160, 0);
        sv_color_points(w, nColors);
        PGX_HIP(hipMemcpyAsync(pts.data(), w.colorPoints.p, sizeof(int32_t) * 160, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipStreamSynchronize(s));
        long points = 0;
        for (int c = 0; c < nColors; ++c) points += pts[c];
        w.lastPointsHost = points;
        const bool coop = w.execMode == EXEC_COOPERATIVE && w.J == 0;
        if (coop && M > 0) {
            float ms = 0.f;
            PGX_HIP(hipEventElapsedTime(&ms, w.ev[6], w.ev[7]));
            w.statSolveUs += 1000.0 * ms;
            w.statSolveBytes += sv_step_bytes(w, M, points);
            ++w.statSolveLaunches;
        }
        for (int j = 0; !coop && j < w.solveEvUsed; ++j) {
            // k_solve: 284 B per contact point + 12 B per manifold of the launched colour;
            // k_solve_flow (colour -1): the same over every colour of the pass
            float ms = 0.f;
            PGX_HIP(hipEventElapsedTime(&ms, w.solveEv[2 * j], w.solveEv[2 * j + 1]));
            int c = w.solveEvColor[j], ce = w.solveEvColorEnd[j];
            if (ce > 0) {   // k_solve_batch: the same per-point model over its run of colours
                double bytes = 0;
                for (int k = c; k < ce; ++k) bytes += 284.0 * pts[k] + 12.0 * (w.colorStartHost[k + 1] - w.colorStartHost[k]);
                w.statBatchUs += 1000.0 * ms;
                w.statBatchBytes += bytes;
                ++w.statBatchLaunches;
                continue;
            }
            double bytes = c >= 0 ? 284.0 * pts[c] + 12.0 * (w.colorStartHost[c + 1] - w.colorStartHost[c])
                                  : 284.0 * points + 12.0 * M;
            w.statSolveUs += 1000.0 * ms;
            w.statSolveBytes += bytes;
            ++w.statSolveLaunches;
        }
        w.statStepUs += 1000.0 * w.tTotal;
        ++w.statSteps;
    }
    w.lastPairs = nPairs;
    w.lastManifolds = M;
    // colours that hold manifolds (inherited colours can leave gaps below the highest)
    int used = 0;
    for (int c = 0; c < nColors; ++c) used += w.colorStartHost[c + 1] > w.colorStartHost[c] ? 1 : 0;
    w.lastColors = used;
}

int report_error(int code, const std::string& msg) { return fail(code, "%s", msg.c_str()); }

}  // namespace pgx

extern "C" {

int pgx_world_step(pgx_world* h, float dt) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    return guarded([&] {
        world_step(h->w, dt);
        return PGX_OK;
    });
}

int pgx_world_sync(pgx_world* h) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    return guarded([&] {
        PGX_HIP(hipSetDevice(h->w.device));
        PGX_HIP(hipStreamSynchronize(h->w.stream));
        return PGX_OK;
    });
}

int pgx_world_read_state(pgx_world* h, float* pos, float* quat, float* vel, float* angvel) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    return guarded([&] {
        World& w = h->w;
        PGX_HIP(hipSetDevice(w.device));
        std::vector<float4> bp(w.A), bq(w.A), dv(w.D), da(w.D);
        hipStream_t s = w.stream;
        if (w.A) {
            PGX_HIP(hipMemcpyAsync(bp.data(), w.bpos.p, sizeof(float4) * w.A, hipMemcpyDeviceToHost, s));
            PGX_HIP(hipMemcpyAsync(bq.data(), w.bquat.p, sizeof(float4) * w.A, hipMemcpyDeviceToHost, s));
        }
        if (w.D) {
            PGX_HIP(hipMemcpyAsync(dv.data(), w.dvel.p, sizeof(float4) * w.D, hipMemcpyDeviceToHost, s));
            PGX_HIP(hipMemcpyAsync(da.data(), w.dang.p, sizeof(float4) * w.D, hipMemcpyDeviceToHost, s));
        }
        PGX_HIP(hipStreamSynchronize(s));
        for (int a = 0; a < w.A; ++a) {
            int b = w.actorToBody[a];
            if (pos) { pos[3 * a] = bp[b].x; pos[3 * a + 1] = bp[b].y; pos[3 * a + 2] = bp[b].z; }
            if (quat) { quat[4 * a] = bq[b].x; quat[4 * a + 1] = bq[b].y; quat[4 * a + 2] = bq[b].z; quat[4 * a + 3] = bq[b].w; }
        }
        for (int d = 0; d < w.D; ++d) {
            if (vel) { vel[3 * d] = dv[d].x; vel[3 * d + 1] = dv[d].y; vel[3 * d + 2] = dv[d].z; }
            if (angvel) { angvel[3 * d] = da[d].x; angvel[3 * d + 1] = da[d].y; angvel[3 * d + 2] = da[d].z; }
        }
        return PGX_OK;
    });
}

int pgx_world_write_state(pgx_world* h, const float* pos, const float* quat, const float* vel, const float* angvel) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    return guarded([&] {
        World& w = h->w;
        PGX_HIP(hipSetDevice(w.device));
        hipStream_t s = w.stream;
        std::vector<float4> bp(w.A), bq(w.A), dv(w.D), da(w.D);
        PGX_HIP(hipMemcpyAsync(bp.data(), w.bpos.p, sizeof(float4) * w.A, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipMemcpyAsync(bq.data(), w.bquat.p, sizeof(float4) * w.A, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipMemcpyAsync(dv.data(), w.dvel.p, sizeof(float4) * w.D, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipMemcpyAsync(da.data(), w.dang.p, sizeof(float4) * w.D, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipStreamSynchronize(s));
        for (int a = 0; a < w.A; ++a) {
            int b = w.actorToBody[a];
            if (pos) bp[b] = make_float4(pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], 0.f);
            if (quat) bq[b] = make_float4(quat[4 * a], quat[4 * a + 1], quat[4 * a + 2], quat[4 * a + 3]);
        }
        for (int d = 0; d < w.D; ++d) {
            if (vel) dv[d] = make_float4(vel[3 * d], vel[3 * d + 1], vel[3 * d + 2], dv[d].w);
            if (angvel) da[d] = make_float4(angvel[3 * d], angvel[3 * d + 1], angvel[3 * d + 2], 0.f);
        }
        upload(w.bpos, bp, s); upload(w.bquat, bq, s); upload(w.dvel, dv, s); upload(w.dang, da, s);
        PGX_HIP(hipStreamSynchronize(s));
        return PGX_OK;
    });
}

int pgx_world_patch_transforms(pgx_world* h, int n, const int32_t* actors, const float* pos, const float* quat) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    if (n < 0 || (n > 0 && (!actors || !pos || !quat))) return fail(PGX_ERR_INVALID, "NULL argument");
    World& w = h->w;
    for (int i = 0; i < n; ++i)
        if (actors[i] < 0 || actors[i] >= w.A) return fail(PGX_ERR_INVALID, "actor %d out of range", actors[i]);
    if (n == 0) return PGX_OK;
    return guarded([&] {
        PGX_HIP(hipSetDevice(w.device));
        hipStream_t s = w.stream;
        std::vector<float4> bp(w.A), bq(w.A);
        PGX_HIP(hipMemcpyAsync(bp.data(), w.bpos.p, sizeof(float4) * w.A, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipMemcpyAsync(bq.data(), w.bquat.p, sizeof(float4) * w.A, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipStreamSynchronize(s));
        std::vector<uint8_t> moved((size_t)w.A, 0), mask((size_t)std::max(w.C, 1), 0);
        for (int i = 0; i < n; ++i) {
            const int a = actors[i], b = w.actorToBody[a];
            bp[b] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.f);
            bq[b] = make_float4(quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]);
            moved[b] = 1;
        }
        for (int c = 0; c < w.C; ++c) mask[c] = moved[w.hostColBody[c]];
        upload(w.bpos, bp, s); upload(w.bquat, bq, s); upload(w.patchMask, mask, s);
        bp_compute_bounds(w, false, nullptr, w.patchMask.p);   // Scene::updateBounds (src/Physecs.cpp:79-90)
        PGX_HIP(hipStreamSynchronize(s));
        return PGX_OK;
    });
}

int pgx_world_get_counts(pgx_world* h, pgx_counts* out) {
    if (!h || !out) return fail(PGX_ERR_INVALID, "NULL argument");
    out->pairs = h->w.lastPairs;
    out->manifolds = h->w.lastManifolds;
    out->colors = h->w.lastColors;
    out->joint_rows = h->w.jRowsTotal;
    out->coloring_rounds = h->w.lastRounds;
    out->points = 0;
    return guarded([&] {
        World& w = h->w;
        PGX_HIP(hipSetDevice(w.device));
        PGX_HIP(hipStreamSynchronize(w.stream));   // the queries read the last step's results
        int M = w.lastManifolds;
        if (M) {
            std::vector<int32_t> f(M);
            PGX_HIP(hipMemcpy(f.data(), w.sFlags.p, sizeof(int32_t) * M, hipMemcpyDeviceToHost));
            long pts = 0;
            for (int v : f) pts += v & 255;
            out->points = (int)pts;
        }
        return PGX_OK;
    });
}

int pgx_world_get_bounds(pgx_world* h, float* out) {
    if (!h || !out) return fail(PGX_ERR_INVALID, "NULL argument");
    return guarded([&] {
        World& w = h->w;
        PGX_HIP(hipSetDevice(w.device));
        PGX_HIP(hipStreamSynchronize(w.stream));   // the queries read the last step's results
        std::vector<float4> mn(w.C), mx(w.C);
        if (w.C) {
            PGX_HIP(hipMemcpy(mn.data(), w.cbmin.p, sizeof(float4) * w.C, hipMemcpyDeviceToHost));
            PGX_HIP(hipMemcpy(mx.data(), w.cbmax.p, sizeof(float4) * w.C, hipMemcpyDeviceToHost));
        }
        for (int c = 0; c < w.C; ++c) {
            out[6 * c] = mn[c].x; out[6 * c + 1] = mn[c].y; out[6 * c + 2] = mn[c].z;
            out[6 * c + 3] = mx[c].x; out[6 * c + 4] = mx[c].y; out[6 * c + 5] = mx[c].z;
        }
        return PGX_OK;
    });
}

int pgx_world_get_pairs(pgx_world* h, int32_t* out, int cap, int* n) {
    if (!h || !n) return fail(PGX_ERR_INVALID, "NULL argument");
    return guarded([&] {
        World& w = h->w;
        PGX_HIP(hipSetDevice(w.device));
        PGX_HIP(hipStreamSynchronize(w.stream));   // the queries read the last step's results
        int P = w.lastPairs;
        *n = P;
        if (!out || P == 0) return PGX_OK;
        std::vector<uint64_t> keys(P);
        std::vector<int32_t> ws(w.W + 1), col(w.C);
        PGX_HIP(hipMemcpy(keys.data(), w.pkeys.p, sizeof(uint64_t) * P, hipMemcpyDeviceToHost));
        PGX_HIP(hipMemcpy(ws.data(), w.worldColStart.p, sizeof(int32_t) * (w.W + 1), hipMemcpyDeviceToHost));
        PGX_HIP(hipMemcpy(col.data(), w.colOfLocal.p, sizeof(int32_t) * w.C, hipMemcpyDeviceToHost));
        uint64_t lm = (1ull << w.localBits) - 1;
        for (int i = 0; i < P && i < cap; ++i) {
            uint64_t k = keys[i];
            int wd = (int)(k >> (2 * w.localBits));
            out[2 * i] = col[ws[wd] + (int)((k >> w.localBits) & lm)];
            out[2 * i + 1] = col[ws[wd] + (int)(k & lm)];
        }
        return PGX_OK;
    });
}

int pgx_world_get_manifolds(pgx_world* h, int32_t* keys, int32_t* ints, float* floats, int cap, int* n) {
    if (!h || !n) return fail(PGX_ERR_INVALID, "NULL argument");
    return guarded([&] {
        World& w = h->w;
        PGX_HIP(hipSetDevice(w.device));
        PGX_HIP(hipStreamSynchronize(w.stream));   // the queries read the last step's results
        int M = w.lastManifolds;
        *n = M;
        if (M == 0) return PGX_OK;
        size_t Mc = w.manifoldCap;
        std::vector<int32_t> canon(M), b0(M), b1(M), flags(M), mfPair(M), mfSrc(M), mfTri(M);
        std::vector<uint64_t> mkey(M);
        std::vector<float4> sn(M), r0(4 * Mc), r1(4 * Mc);
        PGX_HIP(hipMemcpy(canon.data(), w.sCanon.p, sizeof(int32_t) * M, hipMemcpyDeviceToHost));
        PGX_HIP(hipMemcpy(b0.data(), w.sB0.p, sizeof(int32_t) * M, hipMemcpyDeviceToHost));
        PGX_HIP(hipMemcpy(b1.data(), w.sB1.p, sizeof(int32_t) * M, hipMemcpyDeviceToHost));
        PGX_HIP(hipMemcpy(flags.data(), w.sFlags.p, sizeof(int32_t) * M, hipMemcpyDeviceToHost));
        PGX_HIP(hipMemcpy(mfPair.data(), w.mfPair.p, sizeof(int32_t) * M, hipMemcpyDeviceToHost));
        PGX_HIP(hipMemcpy(mfSrc.data(), w.mfSrc.p, sizeof(int32_t) * M, hipMemcpyDeviceToHost));
        PGX_HIP(hipMemcpy(mfTri.data(), w.mfTri.p, sizeof(int32_t) * M, hipMemcpyDeviceToHost));
        PGX_HIP(hipMemcpy(mkey.data(), w.mfKey.p, sizeof(uint64_t) * M, hipMemcpyDeviceToHost));
        PGX_HIP(hipMemcpy(sn.data(), w.sN.p, sizeof(float4) * M, hipMemcpyDeviceToHost));
        PGX_HIP(hipMemcpy(r0.data(), w.sR0.p, sizeof(float4) * 4 * Mc, hipMemcpyDeviceToHost));
        PGX_HIP(hipMemcpy(r1.data(), w.sR1.p, sizeof(float4) * 4 * Mc, hipMemcpyDeviceToHost));
        // narrowphase points: the pair slots, then the triangle-mesh manifold slots
        std::vector<float4> pts(8 * ((size_t)w.lastPairs + w.lastMeshSlots));
        if (w.lastPairs)
            PGX_HIP(hipMemcpy(pts.data(), w.npPts.p, sizeof(float4) * 8 * (size_t)w.lastPairs, hipMemcpyDeviceToHost));
        if (w.lastMeshSlots)
            PGX_HIP(hipMemcpy(pts.data() + 8 * (size_t)w.lastPairs, w.npPts.p + 8 * (size_t)w.pairCap,
                              sizeof(float4) * 8 * (size_t)w.lastMeshSlots, hipMemcpyDeviceToHost));
        std::vector<int32_t> ws(w.W + 1), col(w.C);
        PGX_HIP(hipMemcpy(ws.data(), w.worldColStart.p, sizeof(int32_t) * (w.W + 1), hipMemcpyDeviceToHost));
        PGX_HIP(hipMemcpy(col.data(), w.colOfLocal.p, sizeof(int32_t) * w.C, hipMemcpyDeviceToHost));
        uint64_t lm = (1ull << w.localBits) - 1;
        for (int s = 0; s < M && s < cap; ++s) {
            int m = canon[s];
            uint64_t k = mkey[m];
            int wd = (int)(k >> (2 * w.localBits));
            int c0 = col[ws[wd] + (int)((k >> w.localBits) & lm)];
            int c1 = col[ws[wd] + (int)(k & lm)];
            int num = flags[s] & 255;
            if (keys) {
                keys[5 * s] = (int32_t)w.hostEntity[w.hostColBody[c0]];
                keys[5 * s + 1] = w.hostColIndex[c0];
                keys[5 * s + 2] = (int32_t)w.hostEntity[w.hostColBody[c1]];
                keys[5 * s + 3] = w.hostColIndex[c1];
                keys[5 * s + 4] = mfTri[m];
            }
            if (ints) {
                ints[4 * s] = b0[s]; ints[4 * s + 1] = b1[s]; ints[4 * s + 2] = num; ints[4 * s + 3] = flags[s] >> 16;
            }
            if (floats) {
                float* o = floats + 56 * (size_t)s;
                std::memset(o, 0, 56 * sizeof(float));
                o[0] = sn[s].x; o[1] = sn[s].y; o[2] = sn[s].z; o[3] = sn[s].w;
                int p = mfSrc[m] < w.pairCap ? mfSrc[m] : w.lastPairs + (mfSrc[m] - w.pairCap);
                for (int k2 = 0; k2 < num; ++k2) {
                    float4 a = pts[8 * (size_t)p + 2 * k2], b = pts[8 * (size_t)p + 2 * k2 + 1];
                    float* q = o + 4 + 6 * k2;
                    q[0] = a.x; q[1] = a.y; q[2] = a.z; q[3] = b.x; q[4] = b.y; q[5] = b.z;
                    float4 ra = r0[(size_t)k2 * Mc + s], rb = r1[(size_t)k2 * Mc + s];
                    float* t = o + 28 + 7 * k2;
                    t[0] = ra.x; t[1] = ra.y; t[2] = ra.z; t[3] = rb.x; t[4] = rb.y; t[5] = rb.z; t[6] = ra.w;
                }
            }
        }
        return PGX_OK;
    });
}

int pgx_world_reset_stats(pgx_world* h) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    World& w = h->w;
    w.statSteps = w.statSolveLaunches = w.statBatchLaunches = 0;
    w.statSolveUs = w.statSolveBytes = w.statStepUs = w.statBatchUs = w.statBatchBytes = 0;
    return PGX_OK;
}

int pgx_world_get_stats(pgx_world* h, pgx_stats* out) {
    if (!h || !out) return fail(PGX_ERR_INVALID, "NULL argument");
    World& w = h->w;
    out->steps = w.statSteps;
    out->solve_launches = w.statSolveLaunches;
    out->solve_us = w.statSolveUs;
    out->solve_bytes = w.statSolveBytes;
    out->step_us = w.statStepUs;
    out->batch_launches = w.statBatchLaunches;
    out->batch_us = w.statBatchUs;
    out->batch_bytes = w.statBatchBytes;
    return PGX_OK;
}

int pgx_world_set_exec(pgx_world* h, int mode, int blocks) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    if (blocks < 0) return fail(PGX_ERR_INVALID, "blocks must be >= 0");
    if (mode < PGX_EXEC_COLORS || mode > PGX_EXEC_XPASS) return fail(PGX_ERR_INVALID, "unknown execution mode %d", mode);
    h->w.execMode = mode;
    h->w.persistentBlocksWanted = blocks;
    h->w.persistentBlocks = 0;
    return PGX_OK;
}

static int check_tuning(const pgx_tuning* t) {
    if (!t) return fail(PGX_ERR_INVALID, "tuning is NULL");
    if (t->solve_block != 64 && t->solve_block != 128 && t->solve_block != 256)
        return fail(PGX_ERR_INVALID, "solve_block must be 64, 128 or 256 (got %d)", t->solve_block);
    if (t->color_batch < 0) return fail(PGX_ERR_INVALID, "color_batch must be >= 0");
    if (t->pairs_per_lane != 1 && t->pairs_per_lane != 2 && t->pairs_per_lane != 4)
        return fail(PGX_ERR_INVALID, "pairs_per_lane must be 1, 2 or 4 (got %d)", t->pairs_per_lane);
    if (t->color_rounds < 1 || t->color_rounds > 1024) return fail(PGX_ERR_INVALID, "color_rounds must be in 1..1024");
    if (t->eager_rows_max < 0) return fail(PGX_ERR_INVALID, "eager_rows_max must be >= 0");
    if (t->pair_search != 0 && t->pair_search != 1)
        return fail(PGX_ERR_INVALID, "pair_search must be 0 or 1 (got %d)", t->pair_search);
    return PGX_OK;
}

static void tuning_in(World& w, const pgx_tuning* t) {
    w.tune.solveBlock = t->solve_block;
    w.tune.colorBatch = t->color_batch;
    w.tune.pairsPerLane = t->pairs_per_lane;
    w.tune.colorRounds = t->color_rounds;
    w.tune.eagerMax = t->eager_rows_max;
    w.tune.pairSearch = t->pair_search;
}

static void tuning_out(const World& w, pgx_tuning* t) {
    t->solve_block = w.tune.solveBlock;
    t->color_batch = w.tune.colorBatch;
    t->pairs_per_lane = w.tune.pairsPerLane;
    t->color_rounds = w.tune.colorRounds;
    t->eager_rows_max = w.tune.eagerMax;
    t->pair_search = w.tune.pairSearch;
}

int pgx_world_set_tuning(pgx_world* h, const pgx_tuning* t) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    int rc = check_tuning(t);
    if (rc != PGX_OK) return rc;
    tuning_in(h->w, t);
    return PGX_OK;
}

int pgx_world_get_tuning(pgx_world* h, pgx_tuning* t) {
    if (!h || !t) return fail(PGX_ERR_INVALID, "NULL argument");
    tuning_out(h->w, t);
    return PGX_OK;
}

int pgx_world_autotune(pgx_world* h, float dt, int trial_steps, pgx_tuning* chosen) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    if (trial_steps < 1) return fail(PGX_ERR_INVALID, "trial_steps must be >= 1");
    return guarded([&] {
        World& w = h->w;
        PGX_HIP(hipSetDevice(w.device));
        world_autotune(w, dt, trial_steps);
        if (chosen) tuning_out(w, chosen);
        return PGX_OK;
    });
}

int pgx_world_set_graph(pgx_world* h, int enabled) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    h->w.graph = enabled != 0;
    return PGX_OK;
}

int pgx_world_raycast(pgx_world* h, int n, const float* origins, const float* dirs, const float* max_distance,
                      const int32_t* ray_world, const uint32_t* ignore_entity, int32_t* hit_collider,
                      float* hit_distance, double* device_us) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    if (n < 0) return fail(PGX_ERR_INVALID, "ray count must be >= 0");
    if (n > 0 && (!origins || !dirs || !max_distance || !hit_collider || !hit_distance))
        return fail(PGX_ERR_INVALID, "NULL ray or result array");
    World& w = h->w;
    for (int i = 0; ray_world && i < n; ++i)
        if (ray_world[i] < 0 || ray_world[i] >= w.W) return fail(PGX_ERR_INVALID, "ray %d: world %d out of range", i, ray_world[i]);
    if (n == 0) return PGX_OK;
    return guarded([&] {
        PGX_HIP(hipSetDevice(w.device));
        qr_raycast(w, n, origins, dirs, max_distance, ray_world, ignore_entity, hit_collider, hit_distance, device_us);
        return PGX_OK;
    });
}

int pgx_world_overlap(pgx_world* h, int n, const int32_t* types, const float* params, const float* positions,
                      const float* orientations, const int32_t* query_world, const int32_t* filters, int capacity,
                      int32_t* offsets, int32_t* colliders, int32_t* total, double* device_us) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    if (n < 0 || capacity < 0) return fail(PGX_ERR_INVALID, "query count and capacity must be >= 0");
    if (!offsets || !total) return fail(PGX_ERR_INVALID, "NULL offsets or total");
    if (n > 0 && (!types || !params || !positions || !orientations))
        return fail(PGX_ERR_INVALID, "NULL query array");
    if (capacity > 0 && !colliders) return fail(PGX_ERR_INVALID, "NULL collider array with capacity > 0");
    World& w = h->w;
    for (int i = 0; i < n; ++i) {
        if (types[i] < PGX_SPHERE || types[i] > PGX_TRIANGLE_MESH)
            return fail(PGX_ERR_INVALID, "query %d: shape type %d out of range", i, types[i]);
        if (types[i] == PGX_CONVEX_MESH) {
            const float m = params[4 * i + 3];
            if (!(m >= 0.f && m < (float)w.nConvex) || m != (float)(int)m)
                return fail(PGX_ERR_INVALID, "query %d: convex mesh %g is not one of the scene's %d", i, (double)m,
                            w.nConvex);
        }
        if (query_world && (query_world[i] < 0 || query_world[i] >= w.W))
            return fail(PGX_ERR_INVALID, "query %d: world %d out of range", i, query_world[i]);
    }
    if (n == 0) {
        offsets[0] = 0;
        *total = 0;
        if (device_us) *device_us = 0.0;
        return PGX_OK;
    }
    return guarded([&] {
        PGX_HIP(hipSetDevice(w.device));
        int rc = qr_overlap(w, n, types, params, positions, orientations, query_world, filters, capacity, offsets,
                            colliders, total, device_us);
        if (rc == PGX_ERR_CAPACITY)
            return fail(PGX_ERR_CAPACITY, "overlap: %d hits exceed the capacity %d", *total, capacity);
        return rc;
    });
}

int pgx_world_get_trigger_events(pgx_world* h, int capacity, int32_t* enter, int32_t* n_enter, int32_t* exit,
                                 int32_t* n_exit) {
    if (!h || !n_enter || !n_exit) return fail(PGX_ERR_INVALID, "NULL argument");
    if (capacity < 0) return fail(PGX_ERR_INVALID, "capacity must be >= 0");
    if (capacity > 0 && (!enter || !exit)) return fail(PGX_ERR_INVALID, "NULL event array with capacity > 0");
    World& w = h->w;
    return guarded([&] {
        PGX_HIP(hipSetDevice(w.device));
        int rc = np_trigger_events(w, capacity, enter, n_enter, exit, n_exit);
        if (rc == PGX_ERR_CAPACITY)
            return fail(PGX_ERR_CAPACITY, "trigger events: %d enter / %d exit exceed the capacity %d", *n_enter,
                        *n_exit, capacity);
        return rc;
    });
}

int pgx_world_state_size(pgx_world* h, size_t* bytes) {
    if (!h || !bytes) return fail(PGX_ERR_INVALID, "NULL argument");
    return guarded([&] {
        PGX_HIP(hipSetDevice(h->w.device));
        *bytes = world_state_size(h->w);
        return PGX_OK;
    });
}

int pgx_world_save_state(pgx_world* h, void* buf, size_t bytes) {
    if (!h || !buf) return fail(PGX_ERR_INVALID, "NULL argument");
    return guarded([&] {
        PGX_HIP(hipSetDevice(h->w.device));
        world_save_state(h->w, buf, bytes);
        return PGX_OK;
    });
}

int pgx_world_load_state(pgx_world* h, const void* buf, size_t bytes) {
    if (!h || !buf) return fail(PGX_ERR_INVALID, "NULL argument");
    return guarded([&] {
        PGX_HIP(hipSetDevice(h->w.device));
        world_load_state(h->w, buf, bytes);
        return PGX_OK;
    });
}

int pgx_world_load_state_matched(pgx_world* h, const void* buf, size_t bytes, int gone_capacity, uint32_t* gone,
                                 int32_t* n_gone) {
    if (!h || !buf) return fail(PGX_ERR_INVALID, "NULL argument");
    if (n_gone) *n_gone = 0;
    return guarded([&] {
        PGX_HIP(hipSetDevice(h->w.device));
        world_load_state_matched(h->w, buf, bytes, gone_capacity, gone, n_gone);
        return PGX_OK;
    });
}

int pgx_world_overlap_mtd(pgx_world* h, int n, const int32_t* types, const float* params, const float* positions,
                          const float* orientations, const int32_t* query_world, int capacity, int32_t* offsets,
                          int32_t* colliders, float* normal_mtd, int32_t* total, double* device_us) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    if (n < 0 || capacity < 0) return fail(PGX_ERR_INVALID, "query count and capacity must be >= 0");
    if (!offsets || !total) return fail(PGX_ERR_INVALID, "NULL offsets or total");
    if (n > 0 && (!types || !params || !positions || !orientations))
        return fail(PGX_ERR_INVALID, "NULL query array");
    if (capacity > 0 && (!colliders || !normal_mtd)) return fail(PGX_ERR_INVALID, "NULL result arrays with capacity > 0");
    World& w = h->w;
    for (int i = 0; i < n; ++i) {
        if (types[i] == PGX_TRIANGLE_MESH)
            return fail(PGX_ERR_UNSUPPORTED, "query %d: triangle-mesh query shapes are not supported", i);
        if (types[i] < PGX_SPHERE || types[i] > PGX_CONVEX_MESH)
            return fail(PGX_ERR_INVALID, "query %d: shape type %d out of range", i, types[i]);
        if (types[i] == PGX_CONVEX_MESH) {
            const float m = params[4 * i + 3];
            if (!(m >= 0.f && m < (float)w.nConvex) || m != (float)(int)m)
                return fail(PGX_ERR_INVALID, "query %d: convex mesh %g is not one of the scene's %d", i, (double)m,
                            w.nConvex);
        }
        if (query_world && (query_world[i] < 0 || query_world[i] >= w.W))
            return fail(PGX_ERR_INVALID, "query %d: world %d out of range", i, query_world[i]);
    }
    if (n == 0) {
        offsets[0] = 0;
        *total = 0;
        if (device_us) *device_us = 0.0;
        return PGX_OK;
    }
    return guarded([&] {
        PGX_HIP(hipSetDevice(w.device));
        int rc = qr_overlap_mtd(w, n, types, params, positions, orientations, query_world, capacity, offsets,
                                colliders, normal_mtd, total, device_us);
        if (rc == PGX_ERR_CAPACITY)
            return fail(PGX_ERR_CAPACITY, "MTD overlap: %d hits exceed the capacity %d", *total, capacity);
        return rc;
    });
}

int pgx_world_character_move(pgx_world* h, int n, pgx_character* characters, const float* translations,
                             float delta_time, uint8_t* flags) {
    (void)delta_time;   // only the reference's step-up branch reads it, and that branch never runs
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    if (n < 0) return fail(PGX_ERR_INVALID, "character count must be >= 0");
    if (n > 0 && (!characters || !translations || !flags)) return fail(PGX_ERR_INVALID, "NULL argument");
    World& w = h->w;
    for (int i = 0; i < n; ++i) {
        const pgx_character& c = characters[i];
        if (c.world < 0 || c.world >= w.W) return fail(PGX_ERR_INVALID, "character %d: world %d out of range", i, c.world);
        if (c.num_iterations < 0) return fail(PGX_ERR_INVALID, "character %d: num_iterations < 0", i);
    }
    if (n == 0) return PGX_OK;
    return guarded([&] {
        PGX_HIP(hipSetDevice(w.device));
        qr_character_move(w, n, characters, translations, flags);
        return PGX_OK;
    });
}

int pgx_world_get_exec(pgx_world* h, int* mode) {
    if (!h || !mode) return fail(PGX_ERR_INVALID, "NULL argument");
    *mode = h->w.execMode;
    return PGX_OK;
}

int pgx_world_enable_timing(pgx_world* h, int enable) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    if (enable < 0) return fail(PGX_ERR_INVALID, "negative timing period");
    h->w.timingPeriod = enable;
    h->w.timingCounter = 0;
    return PGX_OK;
}

int pgx_world_get_timing(pgx_world* h, float* b, float* n, float* c, float* s, float* t) {
    if (!h) return fail(PGX_ERR_INVALID, "world is NULL");
    if (b) *b = h->w.tBroad * 1000.f;
    if (n) *n = h->w.tNarrow * 1000.f;
    if (c) *c = h->w.tColor * 1000.f;
    if (s) *s = h->w.tSolve * 1000.f;
    if (t) *t = h->w.tTotal * 1000.f;
    return PGX_OK;
}

}  // extern "C"
