// Internal declarations shared by the HIP translation units of libphysecs_gpu.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include <string>
#include <vector>
#include "../../include/physecs_gpu.h"
#include "pgx_math.h"

namespace pgx {

struct HipError {
    hipError_t err;
    const char* expr;
    const char* file;
    int line;
};
struct PgxError {
    int code;
    std::string msg;
};

#define PGX_HIP(x)                                                      \
    do {                                                                \
        hipError_t e_ = (x);                                            \
        if (e_ != hipSuccess) throw ::pgx::HipError{e_, #x, __FILE__, __LINE__}; \
    } while (0)
#define PGX_LAUNCH_CHECK() PGX_HIP(hipGetLastError())

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t cnt) {
        if (cnt <= n && p) return;
        release();
        PGX_HIP(hipMalloc((void**)&p, std::max<size_t>(cnt, 1) * sizeof(T)));
        n = cnt;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    operator T*() const { return p; }
    void swap(DBuf& o) { std::swap(p, o.p); std::swap(n, o.n); }
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
};

// Pinned host memory for read-backs (one allocation, freed with its owner).
// pinned host memory that grows by reallocation (contents not kept)
struct PinnedBytes {
    void* p = nullptr;
    size_t n = 0;
    void alloc(size_t bytes) {
        if (bytes <= n) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        PGX_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
        n = bytes;
    }
    PinnedBytes() = default;
    PinnedBytes(const PinnedBytes&) = delete;
    PinnedBytes& operator=(const PinnedBytes&) = delete;
    ~PinnedBytes() {
        if (p) (void)hipHostFree(p);
    }
};

struct PinnedInts {
    int32_t* p = nullptr;
    void alloc(size_t n) {
        if (!p) PGX_HIP(hipHostMalloc((void**)&p, sizeof(int32_t) * n, hipHostMallocDefault));
    }
    PinnedInts() = default;
    PinnedInts(const PinnedInts&) = delete;
    PinnedInts& operator=(const PinnedInts&) = delete;
    ~PinnedInts() {
        if (p) (void)hipHostFree(p);
    }
};

constexpr int WAVE = 64;
// What a contact or joint solve reads and writes for one dynamic body, in one 128-byte
// record (one cache line): solver v and w (velocityTemp), component v after the velocity
// integration (.w = inverse mass), component w, and the columns of the world inverse
// inertia of the substep (massTemp's R * invI * R^T). Bodies are gathered by index, so
// one line per body instead of one per array.
enum { BV_V = 0, BV_W = 1, BV_CV = 2, BV_CW = 3, BV_I0 = 4, BV_I1 = 5, BV_I2 = 6, BV_WORDS = 8 };
constexpr int CONTACT_ROW_FIELDS = 9;   // float4 rows per contact point (sRow)
__host__ __device__ __forceinline__ int bv(int body, int field) { return BV_WORDS * body + field; }
constexpr int JOINT_MAX_ROWS = 7;      // a prismatic joint with its limit and drive rows
// substep-loop execution: one launch per colour and pass; one cooperative launch per step
// (grid barriers between colours); one dependency-driven launch per pass
// (one device-wide queue), the same with a queue per XCD
// (one launch per pass); one workgroup per world per pass, bodies in LDS
enum { EXEC_COLORS = 0, EXEC_COOPERATIVE = 1, EXEC_FLOW = 2, EXEC_FLOW_XCD = 3, EXEC_WORLDS = 4, EXEC_PASS = 5, EXEC_XPASS = 6 };
constexpr int FLOW_QUEUES = 8;         // queues of a dependency-driven pass: at most one per XCD
constexpr int WORLD_SMALL = 256;       // per-world passes: worlds up to this many dynamic bodies share a launch
// convex narrowphase: EPA polytope capacity per lane and lanes of the convex kernel
constexpr int EPA_MAXV = 104;          // 4 simplex vertices + at most 100 insertions
constexpr int EPA_MAXF = 256;          // > 2 * EPA_MAXV - 4, the face count of a closed triangulation
constexpr int EPA_MAXE = 128;          // horizon (loose) edges of one insertion
constexpr int EPA_LANES = 65536;
constexpr int CVX_MAX_FACE_VERTS = 16;
// triangle-mesh narrowphase: per-lane capacities and lanes
constexpr int TM_MAXCAND = 256;        // triangles overlapping one body's local bounds
constexpr int TM_STACK = 64;           // BVH traversal stack (the build's depth is checked)
constexpr int TM_CI_WORDS = 14;        // one triangle contact record (+ 1 word: the test hit)
constexpr int TM_WS_WORDS = TM_MAXCAND * (2 + 3);   // per lane: candidates, order, covered vertices
constexpr int TM_LANES = 16384;
inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }

// ------------------------------------------------------------------ utilities (util.hip)
struct ScanTemp {
    DBuf<int32_t> sums;
};
struct SortTemp {
    DBuf<int32_t> hist;                  // per pass: digit x block counts
    DBuf<uint64_t> kalt;
    DBuf<int32_t> valt;
    ScanTemp scan;
};
// Stable LSD radix sort of (keys, vals) by the low `bits` bits of keys (8-bit digits, a
// histogram and a scatter launch per digit). The result is left in keys/vals.
void radix_sort_u64(hipStream_t s, uint64_t* keys, int32_t* vals, int n, int bits, SortTemp& t);
// The sort's scratch, in elements: alternate keys and values (n each), digit histograms,
// scan block sums (0 when the histograms are summed in the scatter itself).
struct RadixScratch {
    size_t keys = 0, vals = 0, hist = 0, sums = 0;
};
RadixScratch radix_sort_scratch(int n, int bits);
// The same sort with caller-provided scratch of radix_sort_scratch(n, bits) elements; it
// needs no zeroed memory (every table it reads it has written first) and allocates nothing.
void radix_sort_u64_into(hipStream_t s, uint64_t* keys, int32_t* vals, int n, int bits, uint64_t* kalt,
                         int32_t* valt, int32_t* hist, int32_t* sums);
// out[i] = sum(in[0..i-1]); *total (device) = sum(in[0..n-1]); in == out is allowed
void exclusive_scan_i32(hipStream_t s, const int32_t* in, int32_t* out, int n, int32_t* total, ScanTemp& t);
// the same with caller-provided block sums (scan_blocks(n) of them)
int scan_blocks(int n);
void exclusive_scan_i32_into(hipStream_t s, const int32_t* in, int32_t* out, int n, int32_t* total, int32_t* sums);

void fill_i32(hipStream_t s, int32_t* p, int32_t v, int n);
void fill_u64(hipStream_t s, uint64_t* p, uint64_t v, int n);

// ------------------------------------------------------------------ world state
// Collider flag bits
enum { CF_ENABLE = 1, CF_TRIGGER = 2, CF_DYNAMIC = 4 };
constexpr int LEAF_BIT = 0x40000000;

struct World {
    int device = 0;
    hipStream_t stream = nullptr;
    int A = 0, D = 0, C = 0, W = 1;       // actors, dynamic bodies, colliders, worlds
    int substeps = 8, iterations = 2;
    float g = 9.81f;
    int pairCap = 0, manifoldCap = 0;
    int localBits = 1, worldBits = 0;
    int blockSize = 256;

    // host-side maps
    std::vector<int> actorToBody, bodyToActor;   // descriptor actor <-> internal body
    std::vector<uint32_t> hostEntity;            // per internal body
    std::vector<int> hostColBody;
    std::vector<int> hostColIndex;
    std::vector<uint8_t> hostKinematic;
    std::vector<uint64_t> hostNonColliding;       // sorted (min entity << 32 | max entity)
    bool nonCollidingDirty = false;

    // bodies: internal body b; b < D are the dynamic-storage entries (same index)
    DBuf<float4> bpos, bquat;            // A
    DBuf<uint32_t> bentity;              // A
    DBuf<float4> dvel, dang;             // D component velocity (dvel.w = invMass)
    DBuf<float4> dcom;                   // D local com (w = kinematic ? 1 : 0)
    DBuf<float4> dinvI;                  // 3 x D columns of the local inverse inertia
    DBuf<float4> bvel;                   // BV_WORDS per dynamic body (record layout below)
    DBuf<float4> pvel, pang;             // D pseudo velocities (pang.w = constraint count)
    DBuf<float4> winvI;                  // 3 x D world inverse inertia (massTemp)

    // colliders, descriptor order
    DBuf<int32_t> cbody, cworld, clocal, cflags, ctype;
    DBuf<uint32_t> centity;
    DBuf<int32_t> cindex;
    DBuf<float4> cparams, clpos, clquat, cmat;
    DBuf<float4> cbmin, cbmax;
    DBuf<int32_t> worldColStart, colOfLocal;   // (world, local) -> collider
    int nDynCol = 0;                     // colliders of non-kinematic dynamic bodies (the LBVH leaves)
    DBuf<int32_t> dynCols;               // those colliders
    DBuf<int32_t> staticCols;            // the other colliders, grouped by world
    DBuf<int32_t> meshCols;              // convex and triangle-mesh colliders: their boxes one wave each
    std::vector<int32_t> meshColsHost;
    DBuf<int32_t> worldStaticStart;      // W + 1 offsets into staticCols
    DBuf<int32_t> refitReady;

    // broadphase
    DBuf<uint32_t> worldBoundsI;         // W x 6 orderable-int encoded min/max of centres
    DBuf<uint64_t> mkeys;
    DBuf<int32_t> mvals;
    DBuf<float4> leafMin, leafMax;       // sorted leaves
    DBuf<int32_t> leafFlags, leafParent, leafEsc;
    DBuf<int32_t> nodeChild;             // 2 x (C-1)
    DBuf<int32_t> nodeParent, nodeEsc, nodeFlags, nodeCounter;
    DBuf<int2> nodeRange;
    DBuf<float4> nodeMin, nodeMax;
    // traversal records: one load round per visited node or leaf
    DBuf<float4> nodeA, nodeB;           // {min, left child bits}, {max, escape bits}
    DBuf<int32_t> nodeLast;              // last leaf of the node's subtree
    DBuf<float4> leafA, leafB;           // {min, escape bits}, {max, flags bits}
    DBuf<int2> leafI;                    // {entity, world-local collider index}
    // The pair walk's records: internal nodes [0, n-1), leaves [n-1, 2n-1) of one array, so
    // every visit is the same three loads. A: {min, next if it overlaps}, B: {max, escape},
    // C: {last leaf of the subtree (a leaf's own index), entity, world-local index, flags}.
    DBuf<float4> travA, travB;
    DBuf<double> sceneMoments;           // sums and sums of squares of the dynamic centres (Morton box)
    long gridFallbacks = 0;              // steps whose grid had a cell run too long (tree walk used)
    DBuf<uint64_t> gridKey;              // the grid's hash table: cell key, run start and end
    DBuf<int32_t> gridStart, gridEnd;
    DBuf<float4> gridRec;                // the grid search's collider records: 2 float4 per dynamic collider
    DBuf<int4> travC;
    // the non-dynamic colliders in static-list order: {min, flags}, {max, entity}, local index
    DBuf<float4> statA, statB;
    DBuf<int32_t> statL;
    DBuf<int32_t> worldRoot, worldStop;
    DBuf<uint64_t> pkeys;
    DBuf<int32_t> pvals;
    DBuf<int32_t> counters;              // [0] pairs, [1] manifolds, [3] uncolored, [5] colour overflow, [6] bvh error, [7] refit ready
    SortTemp sortT;
    ScanTemp scanT;

    // narrowphase output, one slot per pair
    DBuf<int32_t> npNum;                 // numPoints (0 = no manifold)
    DBuf<float4> npNormal;               // normal
    DBuf<float4> npPts;                  // 8 per pair: p0,p1 for 4 points
    DBuf<int32_t> npIndex;               // exclusive scan of (numPoints > 0)
    DBuf<uint64_t> npClass;              // per pair: its shape-pair class (sort key)
    DBuf<int32_t> npOrder;               // pair slots grouped by shape-pair class

    // manifolds, canonical order (sorted by pair key)
    DBuf<uint64_t> mfKey;                // pair key
    DBuf<int32_t> mfPair;                // pair slot
    DBuf<int32_t> mfB0, mfB1, mfColor;
    DBuf<uint64_t> claim, claim2;        // D coloring claims, by round parity
    DBuf<int32_t> colorUncol;            // manifolds left uncoloured after each round
    DBuf<int32_t> mfPrev;                // per manifold: its previous step's cache entry, or -1
    DBuf<int32_t> colorFlag;             // set when inherited colours clash on a body
    DBuf<uint64_t> usedMask;             // 2 x D color masks
    DBuf<uint64_t> colorKeys;
    DBuf<int32_t> colorPerm;
    DBuf<int32_t> colorRank;            // per manifold: its slot in solve order
    DBuf<int32_t> colorCount;            // colour start offsets (129 + 1)
    DBuf<int32_t> worldMStart;           // W first manifold of each world

    // manifolds, solve (colour) order
    DBuf<int32_t> sB0, sB1, sFlags, sA0, sA1;   // sFlags: numPoints | soft<<8 | color<<16
    DBuf<int32_t> sCanon;                // canonical index
    DBuf<int32_t> batchLoc, batchList, batchN;   // runs of small colours: LDS slots of their bodies
    DBuf<int32_t> colorPack;             // the colouring's read-back, packed by k_color_start
    PinnedInts readHost;                 // pinned host memory for the step's read-backs (READ_HOST_INTS)
    bool batchLdsSet = false;            // k_solve_batch's LDS limit raised on this world's device
    DBuf<float4> sN;                     // normal, friction
    DBuf<float2> sSoft;                  // frequency, damping
    DBuf<float4> sR0, sR1;               // 4 x M: r0 local + targetVelocity, r1 local
    DBuf<float4> sRow;                   // 5 x 4 x M per-substep rows
    DBuf<float4> sLam;                   // 4 x M: lambda N, lambda T, target velocity

    // contact cache (canonical order, double-buffered)
    DBuf<uint64_t> cacheKey[2];
    DBuf<int32_t> cacheTri[2];           // triangle index of the cached manifold (-1: none)
    DBuf<int32_t> cacheNum[2];
    DBuf<float4> cacheLp[2];             // 4 x M: local r0 + targetVelocity
    int cacheCur = 0;
    int cacheCount[2] = {0, 0};

    DBuf<uint64_t> nonColliding;
    int nNonColliding = 0;

    // convex meshes: padded vertex buffers, {first vertex, count, first face, faces} per
    // mesh, {first index, count} per face, face normals / centroids
    int nConvex = 0;
    DBuf<float4> cvVerts, cvFaceN, cvFaceC;
    DBuf<int4> cvMesh;
    DBuf<int2> cvFace;
    DBuf<int32_t> cvFaceIdx;
    DBuf<int32_t> heavy;               // pairs with a convex mesh (filled by the primitive kernel)
    int epaLanes = 0;                  // lanes of the convex kernel = EPA workspaces
    DBuf<double> epaV, epaFN;
    DBuf<int32_t> epaFI, epaE;
    // static triangle meshes: vertices, triangles in the BVH build's order (mesh-local
    // vertex indices), triangle normals / centroids, BVH nodes (2 float4 each: {min,
    // triCount bits}, {max, index bits}), {first vertex, vertices, first triangle, first node}
    int nTriMesh = 0, tmMaxDepth = 0;
    DBuf<float4> tmVerts, tmTriN, tmTriC, tmNodes;
    DBuf<int4> tmMesh;
    DBuf<int4> tmTris;
    DBuf<int32_t> meshPairs;           // pairs with a triangle mesh (filled by the primitive kernel)
    int meshLanes = 0;                 // lanes of the mesh kernel
    DBuf<int32_t> tmWs;                // per-lane candidates, order, voided vertices
    int tmCandCap = 0;                 // candidate triangles per step (all mesh pairs)
    DBuf<int32_t> tmCand, tmPairOf;    // per candidate: triangle, mesh pair
    DBuf<int32_t> tmRecs;              // per candidate: contact record (TM_CI_WORDS + hit)
    DBuf<int32_t> tmCandOff, tmCandCount;   // per mesh pair: its candidates
    int mmCap = 0;                     // manifolds of mesh pairs, stored after the pair slots
    int lastMeshSlots = 0;
    DBuf<int32_t> npTri, meshFirst, meshCount;
    DBuf<int32_t> mfSrc, mfTri;        // per manifold: narrowphase data slot, triangle index

    // joints (descriptor order = creation order), colours as Scene::addJoint assigns them
    int J = 0, jCap = 0, jRowsTotal = 0;
    std::vector<int32_t> hjType, hjA0, hjA1, hjB0, hjB1, hjRowFlags, hjNRows, hjColor, hjAlive;
    std::vector<float4> hjP0, hjQ0, hjP1, hjQ1, hjDrive, hjPar0, hjPar1;
    std::vector<unsigned char> hjBits;           // per internal body: colours taken (src/Physecs.cpp:696-705)
    std::vector<int> jColorStart;                // 10 offsets into jOrder
    DBuf<int32_t> jType, jA0, jA1, jB0, jB1, jRowFlags, jNRows, jOrder;
    int jIslands = 0;                    // islands of the one-launch joint passes (0: per-colour launches)
    DBuf<int32_t> jIslandCS, jIslandJoint;   // per island 9 colour offsets; joints by island and colour
    DBuf<float4> jP0, jQ0, jP1, jQ1, jDrive, jR0, jR1, jRow;
    DBuf<float4> jPar0, jPar1;           // (lower, upper, target, stiffness), (damping, gear ratio)
    DBuf<float4> jGear;                  // gear state: persistent angle 0, 1, virtual angle 0, 1
    DBuf<int32_t> jGearInit;
    // user joint types (PGX_JOINT_CUSTOM): the host functions that declare and write their rows
    pgx_joint_layout_fn custLayout = nullptr;
    pgx_joint_rows_fn custRows = nullptr;
    void* custUser = nullptr;
    std::vector<int32_t> custList;               // live custom joints, colour order, creation order inside
    std::vector<pgx_joint_row> custRowsHost;     // JOINT_MAX_ROWS per custom joint, kept across a step's substeps
    DBuf<int32_t> custListDev;
    DBuf<float> custDev;                         // world-space data out, layouts and rows in
    PinnedBytes custHost;                        // the same, on the host

    // last step stats
    int lastPairs = 0, lastManifolds = 0, lastColors = 0, lastRounds = 0, lastPoints = 0;
    std::vector<int> colorStartHost;
    bool timing = false;                 // this step records events (set at each step from the period)
    int timingPeriod = 0;                // events on every timingPeriod-th step (0: none)
    long timingCounter = 0;
    int execMode = EXEC_COLORS;          // how the substep loop's contact passes are launched
    // launch shapes, picked by pgx_world_autotune or set by the caller; results do not
    // depend on them
    struct Tuning {
        int solveBlock = 64;             // threads per k_solve workgroup (64, 128, 256)
        int colorBatch = 512;            // runs of colours of at most this many manifolds: one launch
        int pairsPerLane = 1;            // narrowphase pairs per lane (1, 2, 4)
        int colorRounds = 4;             // least colouring rounds queued per read-back
        int eagerMax = 32768;            // colours of at most this many manifolds load rows eagerly
        int pairSearch = 0;              // broadphase pair search: 0 uniform grid, 1 tree walk
    } tune;
    bool graph = false;                  // substep loop as one HIP graph launch (colour mode)
    // per-world passes: first dynamic body of each world (W + 1), whether every world's
    // dynamic bodies are contiguous, the largest world, the LDS limit; per solve slot its
    // world; per (colour, world) its slots
    std::vector<int32_t> wDynHost;
    DBuf<int32_t> wDyn, sWorld, wcS, wcE;
    // scene queries: rays (origin + maxDistance, direction + world), rejected entities,
    // results, the events around a query's device work
    DBuf<float4> qRays;
    DBuf<uint32_t> qIgnore;
    DBuf<int32_t> qHit;
    DBuf<float> qT;
    hipEvent_t qEv[4] = {};
    // trigger pairs (defaultContactFilter's TRIGGER: overlap only, src/Physecs.cpp:199-206):
    // per step the trigger pairs' slots, and the keys of those that overlap, double-buffered
    // (list trigCur is this step's, the other the last step's); trigN[2 * k] = trigger pairs
    // found while list k was current, trigN[2 * k + 1] = list k's overlapping pairs
    bool hasTriggers = false;
    int trigCur = 0;
    DBuf<int32_t> trigList, trigN;
    DBuf<uint64_t> trigKeys[2];
    std::vector<int32_t> hostColOfLocal, hostWorldColStart;
    // contact filter (Scene::setContactFilter), tabulated: filterK classes of colliders
    // (their distinct (isTrigger, data)), filterTable[k0 * filterK + k1] != 0 makes a pair of
    // classes (collider of the lower entity first) a trigger pair; filterK = 0: the default
    // filter (a trigger on either side)
    int filterK = 0;
    DBuf<int32_t> filterClass;
    DBuf<uint8_t> filterTable;
    // overlap queries: Collider::data (descriptor order), query shapes, per-query hit
    // counts and offsets, the hit keys
    DBuf<int32_t> cdata;
    DBuf<uint8_t> patchMask;             // colliders of the entities a transform patch moved
    DBuf<float4> qShape;
    DBuf<int32_t> qType, qFilter, qCount;
    DBuf<uint64_t> qKeys;
    DBuf<int32_t> qVals;
    // MTD queries: counters, results, the convex and triangle-mesh pair lists
    DBuf<int32_t> qMtdN;
    DBuf<float4> qRes;
    DBuf<int2> qCvxPairs, qMeshPairs;
    DBuf<int32_t> qRecs;                 // MTD triangle-mesh pairs: TM_MAXCAND contact records per lane
    std::vector<int32_t> worldListHost[2];   // worlds of at most WORLD_SMALL dynamic bodies, and the rest
    DBuf<int32_t> worldList[2];
    bool worldsContiguous = false;
    int maxWorldDyn = 0, ldsMax = 0;
    hipGraphExec_t subExec = nullptr;    // the last instantiated substep-loop graph
    long graphInstantiations = 0, graphUpdates = 0;
    int persistentBlocks = 0, persistentBlocksWanted = 0;
    int passBlocksMax = 0;               // co-resident workgroups of the one-launch-per-pass kernel
    DBuf<int32_t> passBarrier;           // its grid barriers' arrival count
    int xpassBlocksMax = 0;              // co-resident workgroups of the one-XCD pass kernel
    DBuf<int32_t> xpassCtl;              // its control blocks, one per pass of a substep
    // dependency-driven passes
    int nXcd = 1;                        // XCDs of the device (CUs / 32)
    uint32_t flowEpoch = 0;              // per launch: versions of the published velocities
    int lastQueues = 0;                  // queues of the last step's passes (0: no pass ran)
    int lastFlowPasses = 0;              // dependency-driven launches of the last step
    DBuf<uint64_t> flowPub;              // 6 x D published velocity words
    DBuf<int32_t> tickets;               // work tickets, FLOW_QUEUES per pass of a substep
    DBuf<int32_t> flowErr;               // set when a dependency wait runs out
    DBuf<int32_t> qStart, qList, qDone;  // queues: offsets, solve positions (M), manifolds finished
    hipEvent_t ev[8] = {};
    float tBroad = 0, tNarrow = 0, tColor = 0, tSolve = 0, tTotal = 0;
    // per-launch timing of the contact solve (first substep of each timed step)
    DBuf<int32_t> colorPoints;           // contact points per colour (this step)
    std::vector<hipEvent_t> solveEv;
    std::vector<int> solveEvColor;       // the launch's colour (-1: every colour of the pass)
    std::vector<int> solveEvColorEnd;    // a k_solve_batch run: one past its last colour (else 0)
    int solveEvUsed = 0;
    long statSteps = 0, statSolveLaunches = 0, statBatchLaunches = 0;
    double statSolveUs = 0, statSolveBytes = 0, statStepUs = 0, statBatchUs = 0, statBatchBytes = 0;
    long lastPointsHost = 0;
};

// broadphase.hip
// bounds of every collider (creation: no margin) or, with the margin, of the dynamic
// colliders; worldMask (device, one byte per world) limits the update to the marked worlds;
// colMask (device, one byte per collider) updates the marked colliders, with the margin
void bp_compute_bounds(World& w, bool dynamicOnlyWithMargin, const uint8_t* worldMask = nullptr,
                       const uint8_t* colMask = nullptr);
int bp_find_pairs(World& w);          // returns pair count (host)
void bp_build(World& w);              // the LBVH from the current bounds (no pairs)
// queries.hip: Scene::raycastClosest for n rays (host arrays)
void qr_raycast(World& w, int n, const float* origins, const float* dirs, const float* maxDist,
                const int32_t* rayWorld, const uint32_t* ignoreEntity, int32_t* hitCollider, float* hitDistance,
                double* deviceUs);
// queries.hip: Scene::overlap for n query shapes (host arrays); returns PGX_OK or
// PGX_ERR_CAPACITY when the hits exceed `capacity` (offsets and *total are set either way)
int qr_overlap(World& w, int n, const int32_t* types, const float* params, const float* positions,
               const float* orientations, const int32_t* queryWorld, const int32_t* filters, int capacity,
               int32_t* offsets, int32_t* colliders, int32_t* total, double* deviceUs);
// queries.hip: CharacterController::move for n characters
void qr_character_move(World& w, int n, pgx_character* chars, const float* translations, uint8_t* flags);
// queries.hip: Scene::overlapWithMinTranslationalDistance for n query shapes; returns PGX_OK
// or PGX_ERR_CAPACITY (only *total set) when the hits exceed `capacity`
int qr_overlap_mtd(World& w, int n, const int32_t* types, const float* params, const float* positions,
                   const float* orientations, const int32_t* queryWorld, int capacity, int32_t* offsets,
                   int32_t* colliders, float* normalMtd, int32_t* total, double* deviceUs);

// device view of the convex mesh tables
struct ConvexView {
    const float4* verts;    // per mesh, padded to a multiple of 4 with copies of the last vertex
    const int4* mesh;       // {first padded vertex, vertex count, first face, face count}
    const int2* face;       // {first index, index count}
    const int32_t* faceIdx; // mesh-local vertex indices
    const float4* faceN;    // outward unit normal
    const float4* faceC;    // centroid
};

// what the overlap query kernel reads and writes
struct OverlapArgs {
    int n, nLeaves, cap, cbits;             // keys: (query << cbits) | collider
    const float4 *qPrm, *qPos, *qQuat;      // per query: params (.w mesh), position (.w world bits), orientation
    const int32_t *qType, *qFilter;
    const int32_t *worldRoot, *worldStop, *worldStaticStart, *staticCols, *mvals;
    const float4 *leafA, *nodeA, *nodeB;
    const float4 *cbmin, *cbmax, *cparams, *clpos, *clquat, *bpos, *bquat;
    const int32_t *cbody, *ctype, *cdata;
    ConvexView cv;
    int32_t* count;                          // per query: its hits
    int32_t* total;                          // all hits (device counter, zero before the launch)
    uint64_t* keys;                          // the first `cap` hits' keys
    int32_t* nCvx;                           // pairs with a convex mesh (device counter, zero before)
    int2* cvxPairs;                          // those pairs (query, collider), listCap of them
    int listCap;
};

// what the MTD overlap query kernels read and write
struct MtdArgs {
    int n, nLeaves, cap, cbits, listCap, recLanes;
    int32_t* recs;        // the triangle-mesh pass's contact records, TM_MAXCAND per lane
    const float4 *qPrm, *qPos, *qQuat;      // per query: params (.w mesh), position (.w world bits), orientation
    const int32_t* qType;
    const int32_t *worldRoot, *worldStop, *worldStaticStart, *staticCols, *mvals;
    const float4 *leafA, *nodeA, *nodeB;
    const float4 *cbmin, *cbmax, *cparams, *clpos, *clquat, *bpos, *bquat;
    const int32_t *cbody, *ctype;
    int32_t* counts;      // [0] results, [1] convex pairs, [2] mesh pairs, [3] errors (device, zero before)
    uint64_t* keys;       // (query, collider, manifold) keys of the first `cap` results
    int32_t* vals;        // their result index (the sort's payload)
    float4* res;          // normal xyz, mtd
    int2* convexPairs;    // (query, collider) with a convex mesh on either side, listCap
    int2* meshPairs;      // (query, triangle-mesh collider), listCap
};

// narrowphase.hip
void np_run(World& w, int nPairs);
// the MTD query's passes: the walk (primitive pairs solved, the others listed), the convex
// pairs, the triangle-mesh pairs
void np_mtd_launch(World& w, const MtdArgs& A);
// the last step's trigger events as (collider of entity0, collider of entity1) pairs,
// ascending; returns PGX_ERR_CAPACITY (counts set) when either list exceeds cap
int np_trigger_events(World& w, int cap, int32_t* enter, int32_t* nEnter, int32_t* exit, int32_t* nExit);
// the overlap query kernel, then (convex: the scene or a query has convex meshes) its
// convex pairs
void np_overlap_launch(hipStream_t s, const OverlapArgs& A, bool convex);
int upload_triangle_meshes(World& w, const pgx_scene_desc* sc, hipStream_t s);   // returns the BVH depth
// device-layout arrays of a scene descriptor (world.hip)
struct SceneArrays {
    int A = 0, D = 0, C = 0, W = 1, maxPerWorld = 1;
    std::vector<int> actorToBody, bodyToActor;
    std::vector<float4> bpos, bquat, dvel, dang, dcom, dinvI, cparams, clpos, clquat, cmat;
    std::vector<uint32_t> bent, cent;
    std::vector<int32_t> cbody, cworld, clocal, cflags, ctype, cindex, wstart, colOfLocal, dynCols, staticCols, wss;
    std::vector<uint8_t> kinematic;
    std::vector<int32_t> dynWorld;        // world of each dynamic body
};
void convert_scene(const pgx_scene_desc* sc, uint32_t entityBase, SceneArrays& o);
void world_reset_host(World& w);
void world_set_shape(World& w, const pgx_config* cfg, int A, int D, int C, int W, int nDynCol, int maxPerWorld);
void world_alloc_derived(World& w, const pgx_config* cfg, int nTrimesh);
void world_set_dyn_ranges(World& w, const std::vector<int32_t>& dynWorld);
void world_set_dyn_starts(World& w, const std::vector<int32_t>& starts, bool contiguous);
void world_init(World& w, const pgx_config* cfg, const pgx_scene_desc* sc);      // world.hip
void world_add_joints(World& w, int n, const pgx_joint_desc* js);
void world_step(World& w, float dt);
void world_release_graph(World& w);
void world_autotune(World& w, float dt, int trialSteps);
int report_error(int code, const std::string& msg);   // sets pgx_last_error(), returns code
void world_enable_triggers(World& w);                  // allocates the trigger lists (once)

// solver.hip
int sv_setup_manifolds(World& w, int nPairs);
void sv_color_points(World& w, int nColors);        // contact points per colour -> colorPoints          // returns manifold count
int sv_color(World& w, int nManifolds);                 // returns colour count
constexpr int READ_HOST_INTS = 2048;
void sv_presize(World& w);
void sv_substeps(World& w, int nManifolds, float dt);
void sv_finish_step(World& w, int nManifolds);
double sv_step_bytes(const World& w, int M, long points);
bool sv_world_pass_fits(World& w);                  // per-world passes possible for this world

// joints.hip
void jt_step_setup(World& w);
void jt_update(World& w, float h);
void jt_colors(World& w, float h, int phase, bool warm);

}  // namespace pgx
