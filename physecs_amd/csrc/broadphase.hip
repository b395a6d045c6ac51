// Broadphase: per-collider AABBs, 30-bit Morton codes (world id in the high bits), an LBVH
// (Karras 2012) rebuilt every step over the radix-sorted codes, and a stackless traversal
// over escape links that emits candidate pairs with wave-aggregated appends.
//
// It replaces the reference's sort-and-sweep (src/Physecs.cpp:119-172) and must produce
// the same pair SET: the same filters (enableSimulation, same entity, both non-dynamic)
// and the same inclusive AABB overlap test (Physecs.cpp:152-154). The bounds follow
// src/BoundsUtil.cpp:35-59 with the reference's margin rule: none at creation, +0.01 for
// the colliders of non-kinematic dynamic bodies after every step (Physecs.cpp:32,79-90).
#include "pgx_internal.h"
#include "shape_bounds.h"

namespace pgx {

// ------------------------------------------------------------------ bounds
__device__ __forceinline__ void collider_world_pose(const float4* __restrict__ bpos, const float4* __restrict__ bquat,
                                                    const float4* __restrict__ clpos,
                                                    const float4* __restrict__ clquat, int body, int c, f3& pos,
                                                    qf& q) {
    f3 p = xyz(bpos[body]);
    qf bq = q4(bquat[body]);
    pos = add(p, qrot(bq, xyz(clpos[c])));
    q = qmul(bq, q4(clquat[c]));
}

__global__ void k_bounds(int C, const int32_t* __restrict__ cbody, const int32_t* __restrict__ ctype,
                         const int32_t* __restrict__ cflags, const float4* __restrict__ cparams,
                         const float4* __restrict__ clpos, const float4* __restrict__ clquat,
                         const float4* __restrict__ bpos, const float4* __restrict__ bquat, int dynamicOnly,
                         const int32_t* __restrict__ cworld, const uint8_t* __restrict__ worldMask,
                         const uint8_t* __restrict__ colMask, const float4* __restrict__ cvVerts,
                         const int4* __restrict__ cvMesh, const float4* __restrict__ tmVerts,
                         const int4* __restrict__ tmMesh, int meshesApart, float4* __restrict__ bmin,
                         float4* __restrict__ bmax) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    if (meshesApart && (ctype[c] == PGX_CONVEX_MESH || ctype[c] == PGX_TRIANGLE_MESH)) return;   // k_bounds_mesh
    if (dynamicOnly && !(cflags[c] & CF_DYNAMIC)) return;
    if (worldMask && !worldMask[cworld[c]]) return;
    if (colMask && !colMask[c]) return;   // the colliders of moved entities only, with the margin
    f3 pos;
    qf q;
    collider_world_pose(bpos, bquat, clpos, clquat, cbody[c], c, pos, q);
    f3 mn, mx;
    shape_bounds(ctype[c], cparams[c], pos, q, cvVerts, cvMesh, tmVerts, tmMesh, mn, mx);
    if (dynamicOnly || colMask) {   // Bounds::addMargin(vec3(0.01f)), src/Bounds.cpp:23-26
        mx = add(mx, s3(0.01f));
        mn = sub(mn, s3(0.01f));
    }
    bmin[c] = f4(mn, 0.f);
    bmax[c] = f4(mx, 0.f);
}

// The box of one convex or triangle-mesh collider, one wave per collider: the lanes take
// every 64th vertex and the wave reduces. BoundsUtil folds min/max over the vertices in
// order and keeps the earlier of equal values (glm::min / max), which decides the sign of a
// zero; the reduction carries each candidate's vertex index and keeps the lower one on ties,
// so the bits are those of the sequential fold.
struct MinMax {
    float v;
    int i;
};
__device__ __forceinline__ MinMax pick_min(MinMax a, MinMax b) { return (b.v < a.v || (b.v == a.v && b.i < a.i)) ? b : a; }
__device__ __forceinline__ MinMax pick_max(MinMax a, MinMax b) { return (a.v < b.v || (b.v == a.v && b.i < a.i)) ? b : a; }

__global__ __launch_bounds__(WAVE) void k_bounds_mesh(int nMesh, const int32_t* __restrict__ meshCols,
                                                      const int32_t* __restrict__ cbody, const int32_t* __restrict__ ctype,
                                                      const int32_t* __restrict__ cflags, const float4* __restrict__ cparams,
                                                      const float4* __restrict__ clpos, const float4* __restrict__ clquat,
                                                      const float4* __restrict__ bpos, const float4* __restrict__ bquat,
                                                      int dynamicOnly, const int32_t* __restrict__ cworld,
                                                      const uint8_t* __restrict__ worldMask,
                                                      const uint8_t* __restrict__ colMask,
                                                      const float4* __restrict__ cvVerts, const int4* __restrict__ cvMesh,
                                                      const float4* __restrict__ tmVerts, const int4* __restrict__ tmMesh,
                                                      float4* __restrict__ bmin, float4* __restrict__ bmax) {
    const int c = meshCols[blockIdx.x];
    if (dynamicOnly && !(cflags[c] & CF_DYNAMIC)) return;
    if (worldMask && !worldMask[cworld[c]]) return;
    if (colMask && !colMask[c]) return;
    f3 pos;
    qf q;
    collider_world_pose(bpos, bquat, clpos, clquat, cbody[c], c, pos, q);
    // src/BoundsUtil.cpp:61-69 (convex: the padded vertex buffer) and :77-85 (triangle mesh)
    const float4 prm = cparams[c];
    const bool cvx = ctype[c] == PGX_CONVEX_MESH;
    const int4 mi = cvx ? cvMesh[(int)prm.w] : tmMesh[(int)prm.w];
    const float4* v = (cvx ? cvVerts : tmVerts) + mi.x;
    const int n = cvx ? ((mi.y + 3) & ~3) : mi.y;
    const f3 sc = cvx ? mk3(prm.x, prm.y, prm.z) : s3(1.f);
    const int lane = threadIdx.x;
    MinMax lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = {3.402823466e+38f, 0x7fffffff};
        hi[k] = {-3.402823466e+38f, 0x7fffffff};
    }
    for (int i = lane; i < n; i += WAVE) {   // each lane's vertices in order: its first extreme kept
        const f3 t = add(pos, qrot(q, cvx ? mulv(sc, xyz(v[i])) : xyz(v[i])));
        const float tv[3] = {t.x, t.y, t.z};
        for (int k = 0; k < 3; ++k) {
            if (tv[k] < lo[k].v) lo[k] = {tv[k], i};
            if (hi[k].v < tv[k]) hi[k] = {tv[k], i};
        }
    }
    for (int o = WAVE / 2; o >= 1; o >>= 1)
        for (int k = 0; k < 3; ++k) {
            lo[k] = pick_min(lo[k], MinMax{__shfl_xor(lo[k].v, o), __shfl_xor(lo[k].i, o)});
            hi[k] = pick_max(hi[k], MinMax{__shfl_xor(hi[k].v, o), __shfl_xor(hi[k].i, o)});
        }
    if (lane != 0) return;
    f3 mn = mk3(lo[0].v, lo[1].v, lo[2].v), mx = mk3(hi[0].v, hi[1].v, hi[2].v);
    if (dynamicOnly || colMask) {   // Bounds::addMargin(vec3(0.01f)), src/Bounds.cpp:23-26
        mx = add(mx, s3(0.01f));
        mn = sub(mn, s3(0.01f));
    }
    bmin[c] = f4(mn, 0.f);
    bmax[c] = f4(mx, 0.f);
}

void bp_compute_bounds(World& w, bool dynamicOnlyWithMargin, const uint8_t* worldMask, const uint8_t* colMask) {
    if (w.C == 0) return;
    const int nMesh = (int)w.meshColsHost.size();
    hipLaunchKernelGGL(k_bounds, dim3(div_up(w.C, 256)), dim3(256), 0, w.stream, w.C, w.cbody.p, w.ctype.p,
                       w.cflags.p, w.cparams.p, w.clpos.p, w.clquat.p, w.bpos.p, w.bquat.p,
                       dynamicOnlyWithMargin ? 1 : 0, w.cworld.p, worldMask, colMask, w.cvVerts.p, w.cvMesh.p,
                       w.tmVerts.p, w.tmMesh.p, nMesh > 0 ? 1 : 0, w.cbmin.p, w.cbmax.p);
    if (nMesh > 0)
        hipLaunchKernelGGL(k_bounds_mesh, dim3(nMesh), dim3(WAVE), 0, w.stream, nMesh, w.meshCols.p, w.cbody.p,
                           w.ctype.p, w.cflags.p, w.cparams.p, w.clpos.p, w.clquat.p, w.bpos.p, w.bquat.p,
                           dynamicOnlyWithMargin ? 1 : 0, w.cworld.p, worldMask, colMask, w.cvVerts.p, w.cvMesh.p,
                           w.tmVerts.p, w.tmMesh.p, w.cbmin.p, w.cbmax.p);
    PGX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ Morton codes
// Quantisation box: the AABB of all dynamic collider centres (the pair set does not depend
// on it; worlds are kept apart by the world id in the key's high bits).
__device__ __forceinline__ uint32_t ord_enc(float f) {   // order-preserving float -> uint
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_dec(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// (and, when given, the step's 16 broadphase counters cleared)
__global__ void k_scene_bounds_init(uint32_t* wb, double* mom, int32_t* counters) {
    if (threadIdx.x < 3) { wb[threadIdx.x] = 0xffffffffu; wb[3 + threadIdx.x] = 0u; wb[6 + threadIdx.x] = 0u; }
    if (threadIdx.x < 6) mom[threadIdx.x] = 0.0;
    if (counters && threadIdx.x < 16) counters[threadIdx.x] = 0;
}

// The dynamic colliders' centres: their box (min, max per axis) and their sums and sums of
// squares per axis (the Morton box and the grid below); and the colliders' largest extent
// per axis (wb[6..8], as float bits: extents are >= 0, so the bits order like the values). One block of 1024 threads reduces its share with
// wave shuffles and per-wave partials in LDS, then 6 integer and 6 double atomics per block.
__global__ __launch_bounds__(1024) void k_scene_bounds(int n, const int32_t* __restrict__ dynCols,
                                                       const float4* __restrict__ bmin,
                                                       const float4* __restrict__ bmax, uint32_t* __restrict__ wb,
                                                       double* __restrict__ mom) {
    __shared__ uint32_t red[9];
    __shared__ double part[1024 / WAVE][6];
    if (threadIdx.x < 3) { red[threadIdx.x] = 0xffffffffu; red[3 + threadIdx.x] = 0u; red[6 + threadIdx.x] = 0u; }
    __syncthreads();
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u}, ex[3] = {0u, 0u, 0u};
    double sm[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int c = dynCols[i];
        float4 a = bmin[c], b = bmax[c];
        const float cc[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
        const float ee[3] = {b.x - a.x, b.y - a.y, b.z - a.z};
        for (int k = 0; k < 3; ++k) {
            const uint32_t e = ord_enc(cc[k]);
            mn[k] = min(mn[k], e);
            mx[k] = max(mx[k], e);
            ex[k] = max(ex[k], __float_as_uint(fmaxf(ee[k], 0.f)));
            sm[k] += (double)cc[k];
            sm[3 + k] += (double)cc[k] * (double)cc[k];
        }
    }
    for (int off = WAVE / 2; off >= 1; off >>= 1) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = min(mn[k], (uint32_t)__shfl_xor((int)mn[k], off));
            mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], off));
            ex[k] = max(ex[k], (uint32_t)__shfl_xor((int)ex[k], off));
        }
        for (int k = 0; k < 6; ++k) sm[k] += __shfl_xor(sm[k], off);
    }
    const int wv = threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        for (int k = 0; k < 3; ++k) { atomicMin(&red[k], mn[k]); atomicMax(&red[3 + k], mx[k]); atomicMax(&red[6 + k], ex[k]); }
        for (int k = 0; k < 6; ++k) part[wv][k] = sm[k];
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        atomicMin(&wb[threadIdx.x], red[threadIdx.x]);
        atomicMax(&wb[3 + threadIdx.x], red[3 + threadIdx.x]);
        atomicMax(&wb[6 + threadIdx.x], red[6 + threadIdx.x]);
    }
    if (threadIdx.x < 6) {
        double t = 0.0;
        for (int w = 0; w < (int)blockDim.x / WAVE; ++w) t += part[w][threadIdx.x];
        atomicAdd(&mom[threadIdx.x], t);
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Morton code of each dynamic collider's centre, 10 bits per axis, over the centres' box cut
// to mean +- 4 standard deviations per axis: a few bodies far from the rest (thrown out of a
// pile) would otherwise stretch the box until thousands of the pile's bodies share a code
// and the tree over them loses its shape. Centres outside the cut box get the edge's code.
// The codes order the tree's leaves only; the pairs the walk finds do not depend on them.
__global__ void k_morton(int n, const int32_t* __restrict__ dynCols, const int32_t* __restrict__ cworld,
                         const float4* __restrict__ bmin, const float4* __restrict__ bmax,
                         const uint32_t* __restrict__ wb, const double* __restrict__ mom, uint64_t* __restrict__ keys,
                         int32_t* __restrict__ vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = dynCols[i];
    float4 a = bmin[c], b = bmax[c];
    float cc[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float lo = ord_dec(wb[k]), hi = ord_dec(wb[3 + k]);
        const double mean = mom[k] / n, var = fmax(mom[3 + k] / n - mean * mean, 0.0), sd = sqrt(var);
        const float clo = (float)(mean - 4.0 * sd), chi = (float)(mean + 4.0 * sd);
        if (clo > lo && clo < hi) lo = clo;
        if (chi < hi && chi > lo) hi = chi;
        float ext = hi - lo;
        float t = ext > 0.f ? (cc[k] - lo) / ext : 0.5f;
        t = fminf(fmaxf(t, 0.f), 1.f);
        q[k] = (uint32_t)fminf(t * 1024.f, 1023.f);
    }
    uint32_t m = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    keys[i] = ((uint64_t)cworld[c] << 30) | m;
    vals[i] = c;
}

__device__ __forceinline__ bool overlaps(float4 amin, float4 amax, float4 bmin, float4 bmax) {
    // inclusive, as the sweep's `max < min` rejections (src/Physecs.cpp:152-154)
    return !(amax.x < bmin.x || amin.x > bmax.x || amax.y < bmin.y || amin.y > bmax.y || amax.z < bmin.z ||
             amin.z > bmax.z);
}

// Workgroups are dispatched round-robin to the 8 XCDs (workgroup b on XCD b % 8), each
// with its own L2. The searches of Morton-adjacent leaves touch the same part of the tree (or the
// same cells), so workgroup b takes logical block xcd_block(b): XCD x searches for one
// contiguous eighth of the leaves and its L2 holds that region's data. (Any other dispatch order only costs L2 hits.)
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int x = b & 7, i = b >> 3, q = nb >> 3, r = nb & 7;
    return x * q + min(x, r) + i;
}

// ------------------------------------------------------------------ uniform grid of the dynamic colliders
// The step's pair search (pgx_tuning.pair_search 0). Cells are boxes of side h per axis, h
// at least the largest dynamic collider extent along that axis (with 1% and a few float
// ulps of slack), laid over the centres' box cut to mean +- 4 standard deviations (the
// Morton box); a collider sits in the cell of its centre, clamped into the grid. Two
// overlapping AABBs have centres at most h apart per axis, so their cells differ by at most
// one per axis; clamping only merges cells, which keeps that. A cell is a key
// world << 30 | Morton(cell x, y, z) (at most 1024 cells per axis), the colliders are sorted
// by it, and a hash table maps each occupied cell to its run [start, end) of the sorted
// colliders. A collider then tests the colliders of its own cell that come after it and all
// colliders of the 13 neighbour cells that lie forward of its cell in (dx, dy, dz) order:
// of two distinct neighbour cells exactly one is forward of the other, so each pair is found
// once. The tests are the sweep's inclusive AABB overlap, so the pair set is the SAP's.
constexpr int GRID_LANES = 16;      // per collider: its own cell, 13 forward cells, 2 halves of the statics
constexpr int GRID_SLOTS = 8;       // pairs a lane holds before its wave appends them
constexpr int GRID_MAX_RUN = 512;   // a fuller cell hands the step to the tree walk
constexpr uint64_t GRID_EMPTY = ~0ull;

__device__ __forceinline__ uint32_t compact10(uint32_t v) {   // every third bit, 10 of them
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((v >> (3 * b)) & 1u) << b;
    return r;
}

// murmur3's 64-bit finaliser (fmix64)
__device__ __forceinline__ uint32_t cell_hash(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return (uint32_t)k;
}

// the grid's origin and cell size per axis, the same in every thread of a step
__device__ __forceinline__ void grid_frame(const uint32_t* __restrict__ wb, const double* __restrict__ mom, int n,
                                           float lo[3], float h[3]) {
    for (int k = 0; k < 3; ++k) {
        float l = ord_dec(wb[k]), hi = ord_dec(wb[3 + k]);
        const float mag = fmaxf(fabsf(l), fabsf(hi));
        const double mean = mom[k] / n, var = fmax(mom[3 + k] / n - mean * mean, 0.0), sd = sqrt(var);
        const float clo = (float)(mean - 4.0 * sd), chi = (float)(mean + 4.0 * sd);
        if (clo > l && clo < hi) l = clo;
        if (chi < hi && chi > l) hi = chi;
        float cell = __uint_as_float(wb[6 + k]) * 1.01f + 1e-5f * mag + 1e-6f;
        const float span = hi - l;
        if (span > 1000.f * cell) cell = span / 1000.f;   // at most ~1000 cells per axis, each still >= the extent
        lo[k] = l;
        h[k] = cell;
    }
}

__global__ void k_cell_codes(int n, const int32_t* __restrict__ dynCols, const int32_t* __restrict__ cworld,
                             const float4* __restrict__ bmin, const float4* __restrict__ bmax,
                             const uint32_t* __restrict__ wb, const double* __restrict__ mom, uint64_t* __restrict__ keys,
                             int32_t* __restrict__ vals, int tableSize, uint64_t* __restrict__ tabKey) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < tableSize) tabKey[i] = GRID_EMPTY;   // the hash table starts empty (k_cell_insert fills it)
    if (i >= n) return;
    float lo[3], h[3];
    grid_frame(wb, mom, n, lo, h);
    int c = dynCols[i];
    float4 a = bmin[c], b = bmax[c];
    const float cc[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) q[k] = (uint32_t)fminf(fmaxf(floorf((cc[k] - lo[k]) / h[k]), 0.f), 1023.f);
    keys[i] = ((uint64_t)cworld[c] << 30) | (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    vals[i] = c;
}

// the first collider of each cell's run claims a table slot (linear probing)
__global__ void k_cell_insert(int n, const uint64_t* __restrict__ keys, uint32_t mask, uint64_t* __restrict__ tabKey,
                              int32_t* __restrict__ tabStart) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    if (i > 0 && keys[i - 1] == k) return;
    uint32_t slot = cell_hash(k) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {   // the table is at least twice the cells
        const unsigned long long prev = atomicCAS((unsigned long long*)&tabKey[slot], GRID_EMPTY, k);
        if (prev == GRID_EMPTY) {
            tabStart[slot] = i;
            return;
        }
        slot = (slot + 1) & mask;
    }
}

__device__ __forceinline__ int cell_find(uint64_t k, uint32_t mask, const uint64_t* __restrict__ tabKey) {
    uint32_t slot = cell_hash(k) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const uint64_t t = tabKey[slot];
        if (t == k) return (int)slot;
        if (t == GRID_EMPTY) return -1;
        slot = (slot + 1) & mask;
    }
    return -1;
}

// the last collider of each run records its end, and the longest run (counters[15]: one
// atomic per workgroup, after a wave and workgroup maximum)
__global__ __launch_bounds__(256) void k_cell_ends(int n, const uint64_t* __restrict__ keys, uint32_t mask,
                                                   const uint64_t* __restrict__ tabKey,
                                                   const int32_t* __restrict__ tabStart, int32_t* __restrict__ tabEnd,
                                                   int32_t* __restrict__ counters) {
    __shared__ int32_t waveMax[256 / WAVE];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int run = 0;
    if (i < n) {
        const uint64_t k = keys[i];
        if (i + 1 == n || keys[i + 1] != k) {
            const int slot = cell_find(k, mask, tabKey);
            if (slot < 0) {
                atomicOr(&counters[14], 2);   // never expected: the run start inserted it
            } else {
                tabEnd[slot] = i + 1;
                run = i + 1 - tabStart[slot];
            }
        }
    }
    for (int o = WAVE / 2; o >= 1; o >>= 1) run = max(run, __shfl_xor(run, o));
    if ((threadIdx.x & (WAVE - 1)) == 0) waveMax[threadIdx.x / WAVE] = run;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = 0;
        for (int w = 0; w < 256 / WAVE; ++w) m = max(m, waveMax[w]);
        if (m > 0) atomicMax(&counters[15], m);
    }
}

// The grid search's record of each dynamic collider in sorted order, one 32-byte line:
// {min xyz, world-local index}, {max xyz, entity}; a disabled collider gets an empty box, so
// it overlaps nothing. And the world's non-dynamic colliders, as k_pack writes them for the
// tree walk.
__global__ void k_grid_pack(int n, const int32_t* __restrict__ order, const float4* __restrict__ cbmin,
                            const float4* __restrict__ cbmax, const int32_t* __restrict__ cflags,
                            const uint32_t* __restrict__ centity, const int32_t* __restrict__ clocal,
                            float4* __restrict__ gridRec, int nStatic, const int32_t* __restrict__ staticCols,
                            float4* __restrict__ statA, float4* __restrict__ statB, int32_t* __restrict__ statL) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        const int c = order[t];
        float4 a = cbmin[c], b = cbmax[c];
        if (!(cflags[c] & CF_ENABLE)) {
            a = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
            b = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
        }
        gridRec[2 * t] = make_float4(a.x, a.y, a.z, __int_as_float(clocal[c]));
        gridRec[2 * t + 1] = make_float4(b.x, b.y, b.z, __int_as_float((int)centity[c]));
    }
    if (t < nStatic) {
        const int c = staticCols[t];
        const float4 a = cbmin[c], b = cbmax[c];
        statA[t] = make_float4(a.x, a.y, a.z, __int_as_float(cflags[c]));
        statB[t] = make_float4(b.x, b.y, b.z, __int_as_float((int)centity[c]));
        statL[t] = clocal[c];
    }
}

// GRID_LANES lanes per enabled dynamic collider s (sorted order): lane 0 its own cell after
// s, lanes 1-13 the forward neighbour cells, lanes 14-15 the two halves of its world's
// non-dynamic colliders. Two candidates per round of loads (more registers would cost the
// occupancy that keeps every wave of the 32k-collider grid resident at once). Found pairs
// wait in the lane's
// LDS slots; at the end the workgroup scans its lanes' counts and appends all its pairs with
// one atomic on the pair count (a lane whose slots fill before that appends with its wave).
// When the fullest cell holds more than GRID_MAX_RUN colliders (cells made large by a few
// large colliders) every lane leaves at once and counters[14] hands the step to the tree walk.
constexpr int GRID_BLOCK = 1024;
__global__ __launch_bounds__(GRID_BLOCK) void k_grid_pairs(
    int n, const uint64_t* __restrict__ keys, const float4* __restrict__ rec, uint32_t mask,
    const uint64_t* __restrict__ tabKey, const int32_t* __restrict__ tabStart, const int32_t* __restrict__ tabEnd,
    const int32_t* __restrict__ worldStaticStart, const float4* __restrict__ statA, const float4* __restrict__ statB,
    const int32_t* __restrict__ statL, int localBits, int32_t* __restrict__ counters, int cap,
    uint64_t* __restrict__ pkeys) {
    if (counters[15] > GRID_MAX_RUN) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&counters[14], 1);
        return;
    }
    __shared__ uint64_t found[GRID_SLOTS][GRID_BLOCK];
    __shared__ int32_t waveTotal[GRID_BLOCK / WAVE];
    __shared__ int32_t blockBase;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
    const long long g = (long long)xcd_block(blockIdx.x, gridDim.x) * GRID_BLOCK + tid;
    const int s = (int)(g / GRID_LANES), k = (int)(g % GRID_LANES);
    bool active = false, statics = false;
    int j = 0, end = 0, wd = 0, myLocal = 0;
    uint32_t myEnt = 0;
    float4 qmin = make_float4(0.f, 0.f, 0.f, 0.f), qmax = qmin;
    if (s < n) {
        qmin = rec[2 * s];
        qmax = rec[2 * s + 1];
    }
    if (s < n && qmin.x <= qmax.x) {   // an enabled collider (a disabled one has an empty box)
        const uint64_t key = keys[s];
        wd = (int)(key >> 30);
        myLocal = __float_as_int(qmin.w);
        myEnt = (uint32_t)__float_as_int(qmax.w);
        if (k < 14) {
            int dx = 0, dy = 0, dz = 0;
            if (k >= 1 && k <= 9) { dx = 1; dy = (k - 1) / 3 - 1; dz = (k - 1) % 3 - 1; }
            else if (k >= 10 && k <= 12) { dy = 1; dz = k - 11; }
            else if (k == 13) { dz = 1; }
            const uint32_t m = (uint32_t)(key & 0x3fffffffu);
            const int x = (int)compact10(m >> 2) + dx, y = (int)compact10(m >> 1) + dy, z = (int)compact10(m) + dz;
            if (x >= 0 && x < 1024 && y >= 0 && y < 1024 && z >= 0 && z < 1024) {
                const uint64_t nk = ((uint64_t)wd << 30) | (spread10((uint32_t)x) << 2) | (spread10((uint32_t)y) << 1) |
                                    spread10((uint32_t)z);
                const int slot = cell_find(nk, mask, tabKey);
                if (slot >= 0) {
                    j = k == 0 ? s + 1 : tabStart[slot];
                    end = tabEnd[slot];
                    active = j < end;
                }
            }
        } else {
            const int a = worldStaticStart[wd], b = worldStaticStart[wd + 1], mid = a + (b - a + 1) / 2;
            j = k == 14 ? a : mid;
            end = k == 14 ? mid : b;
            active = j < end;
            statics = true;
        }
    }
    int held = 0;
    auto pairKey = [&](uint32_t e2, int l2) {
        const int l0 = myEnt < e2 ? myLocal : l2;
        const int l1 = myEnt < e2 ? l2 : myLocal;
        return ((uint64_t)wd << (2 * localBits)) | ((uint64_t)l0 << localBits) | (uint64_t)l1;
    };
    auto waveFlush = [&]() {   // a lane's slots are full: the wave appends its pairs now
        int incl = held;
        for (int o = 1; o < WAVE; o <<= 1) {
            int t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        const int total = __shfl(incl, WAVE - 1);
        int base = 0;
        if (lane == WAVE - 1 && total) base = atomicAdd(counters, total);
        base = __shfl(base, WAVE - 1) + incl - held;
        for (int q = 0; q < held; ++q)
            if (base + q < cap) pkeys[base + q] = found[q][tid];
        held = 0;
    };
    // two candidates per round of loads: the cell's dynamic colliders (lanes 0-13) or the
    // world's non-dynamic ones (lanes 14-15)
    while (__ballot(active)) {
        bool hit0 = false, hit1 = false;
        uint64_t c0 = 0, c1 = 0;
        if (active) {
            const bool two = j + 1 < end;
            float4 a0, b0, a1, b1;
            bool ok0 = true, ok1 = true;
            if (statics) {
                a0 = statA[j]; b0 = statB[j];
                a1 = two ? statA[j + 1] : a0; b1 = two ? statB[j + 1] : b0;
                const int l0 = statL[j], l1 = two ? statL[j + 1] : l0;
                ok0 = (__float_as_int(a0.w) & CF_ENABLE) != 0;
                ok1 = (__float_as_int(a1.w) & CF_ENABLE) != 0;
                a0.w = __int_as_float(l0);
                a1.w = __int_as_float(l1);
            } else {
                a0 = rec[2 * j]; b0 = rec[2 * j + 1];
                a1 = two ? rec[2 * j + 2] : a0; b1 = two ? rec[2 * j + 3] : b0;
            }
            const uint32_t e0 = (uint32_t)__float_as_int(b0.w), e1 = (uint32_t)__float_as_int(b1.w);
            hit0 = ok0 && e0 != myEnt && overlaps(qmin, qmax, a0, b0);
            hit1 = two && ok1 && e1 != myEnt && overlaps(qmin, qmax, a1, b1);
            if (hit0) c0 = pairKey(e0, __float_as_int(a0.w));
            if (hit1) c1 = pairKey(e1, __float_as_int(a1.w));
            j += two ? 2 : 1;
            active = j < end;
        }
        if (__ballot(held + 2 > GRID_SLOTS)) waveFlush();
        if (hit0) found[held++][tid] = c0;
        if (hit1) found[held++][tid] = c1;
    }
    // the workgroup's remaining pairs: one atomic for all of them
    int incl = held;
    for (int o = 1; o < WAVE; o <<= 1) {
        int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == WAVE - 1) waveTotal[wv] = incl;
    __syncthreads();
    if (tid == 0) {
        int total = 0;
        for (int w = 0; w < GRID_BLOCK / WAVE; ++w) {
            const int t = waveTotal[w];
            waveTotal[w] = total;
            total += t;
        }
        blockBase = total ? atomicAdd(counters, total) : 0;
    }
    __syncthreads();
    const int base = blockBase + waveTotal[wv] + incl - held;
    for (int q = 0; q < held; ++q)
        if (base + q < cap) pkeys[base + q] = found[q][tid];
}

// ------------------------------------------------------------------ LBVH build
__device__ __forceinline__ int lcp(const uint64_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a != b) return __clzll((long long)(a ^ b));
    return 64 + __clz((int)((uint32_t)i ^ (uint32_t)j));
}

// Internal node i of the n-1 nodes: range, split, children (Karras 2012, Alg. 2).
__global__ void k_build(int n, const uint64_t* __restrict__ keys, int32_t* __restrict__ child,
                        int2* __restrict__ range, int32_t* __restrict__ nodeParent,
                        int32_t* __restrict__ leafParent, int32_t* __restrict__ worldRoot, int worldShift) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (lcp(keys, n, i, i + 1) - lcp(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = lcp(keys, n, i, i - d);
    int lmax = 2;
    while (lcp(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (lcp(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int first = min(i, j), last = max(i, j);
    int dnode = lcp(keys, n, first, last);
    int split = first, step = last - first;
    do {
        step = (step + 1) >> 1;
        int ns = split + step;
        if (ns < last && lcp(keys, n, first, ns) > dnode) split = ns;
    } while (step > 1);
    int left = (split == first) ? (split | LEAF_BIT) : split;
    int right = (split + 1 == last) ? ((split + 1) | LEAF_BIT) : (split + 1);
    child[2 * i] = left;
    child[2 * i + 1] = right;
    range[i] = make_int2(first, last);
    if (left & LEAF_BIT) leafParent[split] = i; else nodeParent[split] = i;
    if (right & LEAF_BIT) leafParent[split + 1] = i; else nodeParent[split + 1] = i;
    // A child whose leaves all belong to one world while this node spans two worlds is
    // that world's root (the world id forms the high bits of every key).
    int wFirst = (int)(keys[first] >> worldShift), wLast = (int)(keys[last] >> worldShift);
    if (wFirst != wLast) {
        int wl = (int)(keys[first] >> worldShift), wsplit = (int)(keys[split] >> worldShift);
        if (wl == wsplit) worldRoot[wl] = left;
        int wr = (int)(keys[split + 1] >> worldShift), wlast = (int)(keys[last] >> worldShift);
        if (wr == wlast) worldRoot[wr] = right;
    } else if (i == 0) {
        worldRoot[wFirst] = 0;
    }
}

// Leaf data in sorted order.
__global__ void k_leaves(int n, const int32_t* __restrict__ order, const float4* __restrict__ bmin,
                         const float4* __restrict__ bmax, const int32_t* __restrict__ cflags,
                         float4* __restrict__ lmin, float4* __restrict__ lmax, int32_t* __restrict__ lflags) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int c = order[s];
    lmin[s] = bmin[c];
    lmax[s] = bmax[c];
    lflags[s] = cflags[c];
}

// Bottom-up refit, stage 1. Block b owns leaves [b*1024, b*1024+1023]; an internal node
// whose leaf range lies inside that span is "local" (its index lies in the span too), and
// is completed in LDS by the second of its two children to arrive, with workgroup-scope
// ordering only. A climb that reaches a non-local node counts one finished child for it.
constexpr int REFIT_BLOCK = 1024;

__device__ __forceinline__ float4 fmin4(float4 a, float4 b) {
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.f);
}
__device__ __forceinline__ float4 fmax4(float4 a, float4 b) {
    return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), 0.f);
}

__global__ __launch_bounds__(REFIT_BLOCK) void k_refit_local(
    int n, const int32_t* __restrict__ child, const int2* __restrict__ range, const int32_t* __restrict__ nodeParent,
    const int32_t* __restrict__ leafParent, const float4* __restrict__ lmin, const float4* __restrict__ lmax,
    float4* __restrict__ nmin, float4* __restrict__ nmax, int32_t* __restrict__ gcount) {
    __shared__ int cnt[REFIT_BLOCK];
    __shared__ float4 smin[REFIT_BLOCK], smax[REFIT_BLOCK];
    const int B0 = blockIdx.x * REFIT_BLOCK;
    const int B1 = min(n, B0 + REFIT_BLOCK) - 1;
    cnt[threadIdx.x] = 0;
    __syncthreads();
    int s = B0 + threadIdx.x;
    if (s >= n) return;
    int node = leafParent[s];
    for (int guard = 0; node >= 0 && guard < REFIT_BLOCK; ++guard) {
        int2 r = range[node];
        if (r.x < B0 || r.y > B1) {
            atomicAdd(&gcount[node], 1);   // consumed by stage 2 after this kernel ends
            return;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (atomicAdd(&cnt[node - B0], 1) == 0) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        int a = child[2 * node], b = child[2 * node + 1];
        float4 amn = (a & LEAF_BIT) ? lmin[a & ~LEAF_BIT] : smin[a - B0];
        float4 amx = (a & LEAF_BIT) ? lmax[a & ~LEAF_BIT] : smax[a - B0];
        float4 bmn = (b & LEAF_BIT) ? lmin[b & ~LEAF_BIT] : smin[b - B0];
        float4 bmx = (b & LEAF_BIT) ? lmax[b & ~LEAF_BIT] : smax[b - B0];
        float4 mn = fmin4(amn, bmn), mx = fmax4(amx, bmx);
        smin[node - B0] = mn;
        smax[node - B0] = mx;
        nmin[node] = mn;
        nmax[node] = mx;
        node = nodeParent[node];
    }
}

// Stage 2a: list the non-local nodes whose two children were both finished in stage 1.
__global__ void k_refit_ready(int n, const int32_t* __restrict__ gcount, int32_t* __restrict__ ready,
                              int32_t* __restrict__ nReady) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n - 1 && gcount[i] == 2) ready[atomicAdd(nReady, 1)] = i;
}

// Stage 2b, one workgroup: climbs from the ready nodes; a parent (non-local as well) is
// completed by whichever climb brings its second child.
__global__ __launch_bounds__(1024) void k_refit_global(int n, const int32_t* __restrict__ child,
                                                       const int32_t* __restrict__ nodeParent,
                                                       const float4* __restrict__ lmin,
                                                       const float4* __restrict__ lmax, float4* __restrict__ nmin,
                                                       float4* __restrict__ nmax, int32_t* __restrict__ gcount,
                                                       const int32_t* __restrict__ ready,
                                                       const int32_t* __restrict__ nReady) {
    int nr = *nReady;
    for (int k = threadIdx.x; k < nr; k += blockDim.x) {
        int node = ready[k];
        for (int guard = 0; node >= 0 && guard < n; ++guard) {
            int a = child[2 * node], b = child[2 * node + 1];
            float4 amn = (a & LEAF_BIT) ? lmin[a & ~LEAF_BIT] : nmin[a];
            float4 amx = (a & LEAF_BIT) ? lmax[a & ~LEAF_BIT] : nmax[a];
            float4 bmn = (b & LEAF_BIT) ? lmin[b & ~LEAF_BIT] : nmin[b];
            float4 bmx = (b & LEAF_BIT) ? lmax[b & ~LEAF_BIT] : nmax[b];
            nmin[node] = fmin4(amn, bmn);
            nmax[node] = fmax4(amx, bmx);
            int p = nodeParent[node];
            if (p < 0) break;
            int old = __hip_atomic_fetch_add(&gcount[p], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old != 1) break;
            node = p;
        }
    }
}

// Escape link of every node: the next node in depth-first pre-order after its subtree.
__global__ void k_escape(int n, const int32_t* __restrict__ child, const int32_t* __restrict__ nodeParent,
                         const int32_t* __restrict__ leafParent, int32_t* __restrict__ nodeEsc,
                         int32_t* __restrict__ leafEsc) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int total = 2 * n - 1;
    if (t >= total) return;
    bool isLeaf = t >= n - 1;
    int self = isLeaf ? ((t - (n - 1)) | LEAF_BIT) : t;
    int cur = self;
    int esc = -1;
    for (int guard = 0; guard < n; ++guard) {
        int p = (cur & LEAF_BIT) ? leafParent[cur & ~LEAF_BIT] : nodeParent[cur];
        if (p < 0) break;
        if (child[2 * p] == cur) { esc = child[2 * p + 1]; break; }
        cur = p;
    }
    if (isLeaf) leafEsc[t - (n - 1)] = esc; else nodeEsc[t] = esc;
}

__global__ void k_world_stop(int W, const int32_t* __restrict__ worldRoot, const int32_t* __restrict__ nodeEsc,
                             const int32_t* __restrict__ leafEsc, int32_t* __restrict__ worldStop) {
    int wd = blockIdx.x * blockDim.x + threadIdx.x;
    if (wd >= W) return;
    int r = worldRoot[wd];
    if (r == -1) { worldStop[wd] = -1; return; }
    worldStop[wd] = (r & LEAF_BIT) ? leafEsc[r & ~LEAF_BIT] : nodeEsc[r];
}

// a child or escape link (a leaf carries LEAF_BIT, -1 ends the walk) as an index into the
// pair walk's one array of n - 1 internal nodes followed by n leaves
__device__ __forceinline__ int trav_index(int link, int n) {
    return link < 0 ? -1 : (link & LEAF_BIT) ? n - 1 + (link & ~LEAF_BIT) : link;
}

// Traversal records: a visited node (or leaf) is one round of independent loads. The
// queries walk nodeA/nodeB/leafA/leafB/leafI; the pair walk walks travA/travB/travC, and
// tests the non-dynamic colliders from statA/statB/statL.
__global__ void k_pack(int n, const float4* __restrict__ nmin, const float4* __restrict__ nmax,
                       const int32_t* __restrict__ child, const int32_t* __restrict__ nodeEsc,
                       const int2* __restrict__ range, const float4* __restrict__ lmin, const float4* __restrict__ lmax,
                       const int32_t* __restrict__ leafEsc, const int32_t* __restrict__ lflags,
                       const int32_t* __restrict__ order, const uint32_t* __restrict__ centity,
                       const int32_t* __restrict__ clocal, float4* __restrict__ nodeA, float4* __restrict__ nodeB,
                       int32_t* __restrict__ nodeLast, float4* __restrict__ leafA, float4* __restrict__ leafB,
                       int2* __restrict__ leafI, float4* __restrict__ travA, float4* __restrict__ travB,
                       int4* __restrict__ travC, int nStatic, const int32_t* __restrict__ staticCols,
                       const float4* __restrict__ cbmin, const float4* __restrict__ cbmax,
                       const int32_t* __restrict__ cflags, float4* __restrict__ statA, float4* __restrict__ statB,
                       int32_t* __restrict__ statL) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n - 1) {
        float4 a = nmin[t], b = nmax[t];
        nodeA[t] = make_float4(a.x, a.y, a.z, __int_as_float(child[2 * t]));
        nodeB[t] = make_float4(b.x, b.y, b.z, __int_as_float(nodeEsc[t]));
        nodeLast[t] = range[t].y;
        travA[t] = make_float4(a.x, a.y, a.z, __int_as_float(trav_index(child[2 * t], n)));
        travB[t] = make_float4(b.x, b.y, b.z, __int_as_float(trav_index(nodeEsc[t], n)));
        travC[t] = make_int4(range[t].y, 0, 0, CF_ENABLE);
    }
    if (t < n) {
        float4 a = lmin[t], b = lmax[t];
        leafA[t] = make_float4(a.x, a.y, a.z, __int_as_float(leafEsc[t]));
        leafB[t] = make_float4(b.x, b.y, b.z, __int_as_float(lflags[t]));
        int c = order[t];
        leafI[t] = make_int2((int)centity[c], clocal[c]);
        const float esc = __int_as_float(trav_index(leafEsc[t], n));
        if (lflags[t] & CF_ENABLE) {
            travA[n - 1 + t] = make_float4(a.x, a.y, a.z, esc);
            travB[n - 1 + t] = make_float4(b.x, b.y, b.z, esc);
        } else {   // a disabled leaf overlaps nothing
            travA[n - 1 + t] = make_float4(INFINITY, INFINITY, INFINITY, esc);
            travB[n - 1 + t] = make_float4(-INFINITY, -INFINITY, -INFINITY, esc);
        }
        travC[n - 1 + t] = make_int4(t, (int)centity[c], clocal[c], lflags[t]);
    }
    if (t < nStatic) {
        const int c = staticCols[t];
        const float4 a = cbmin[c], b = cbmax[c];
        statA[t] = make_float4(a.x, a.y, a.z, __int_as_float(cflags[c]));
        statB[t] = make_float4(b.x, b.y, b.z, __int_as_float((int)centity[c]));
        statL[t] = clocal[c];
    }
}

// ------------------------------------------------------------------ traversal

constexpr int TRAV_SLOTS = 8;   // pairs a lane holds before its wave appends them

// Each enabled dynamic leaf s walks the records after it in depth-first pre-order, starting
// at its own escape link (the pre-order successor of a leaf), up to its world's stop: those
// are exactly the subtrees whose leaves come after s, so each dynamic pair is found once,
// by its first leaf, and no walk descends from the root to its own leaf. Every walk step is
// two loads of one record ({min, next if taken}, {max, escape}), internal node or leaf
// alike, and one select: a record is taken when it overlaps; a taken node leads to its left
// child, a taken leaf is a candidate and leads over its escape link. A disabled leaf's
// record holds an empty box. Alongside the walk, one of the world's non-dynamic colliders
// (static and kinematic; their bounds never move during a step) per step is tested from its
// own records. Candidates from the walk are kept as leaf indices and resolved to pair keys
// (and same-entity colliders dropped) when the lane's slots are flushed.
__global__ __launch_bounds__(256) void k_traverse(
    int n, const uint64_t* __restrict__ mkeys, int worldShift, const float4* __restrict__ travA,
    const float4* __restrict__ travB, const int4* __restrict__ travC, const int32_t* __restrict__ worldStop,
    const int32_t* __restrict__ worldStaticStart, const float4* __restrict__ statA,
    const float4* __restrict__ statB, const int32_t* __restrict__ statL, int localBits,
    int32_t* __restrict__ pairCount, int cap, uint64_t* __restrict__ pkeys) {
    const int s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    bool active = false;
    float4 qmin = make_float4(0.f, 0.f, 0.f, 0.f), qmax = qmin;
    int node = -1, stop = -1, st = 0, stEnd = 0;
    int myLocal = 0;
    uint32_t myEnt = 0;
    int wd = 0;
    if (s < n) {
        const int self = n - 1 + s;
        const int4 c = travC[self];
        if (c.w & CF_ENABLE) {
            active = true;
            qmin = travA[self];
            qmax = travB[self];
            node = __float_as_int(qmax.w);   // the leaf's escape link: its pre-order successor
            myEnt = (uint32_t)c.y;
            myLocal = c.z;
            wd = (int)(mkeys[s] >> worldShift);
            stop = trav_index(worldStop[wd], n);
            st = worldStaticStart[wd];
            stEnd = worldStaticStart[wd + 1];
        }
    }
    const int lane = threadIdx.x & (WAVE - 1);
    // candidates wait in the lane's own LDS slots (a static pair as its key, a dynamic leaf
    // as its index, marked in treeMask); when a lane's slots are full (or the walk ends) the
    // wave resolves and appends all of its lanes' pairs with one atomic on the counter
    __shared__ uint64_t found[TRAV_SLOTS][256];   // slot-major: a wave's lanes hit consecutive words
    int held = 0;
    uint32_t treeMask = 0;
    auto pairKey = [&](uint32_t e2, int l2) {
        const int l0 = myEnt < e2 ? myLocal : l2;
        const int l1 = myEnt < e2 ? l2 : myLocal;
        return ((uint64_t)wd << (2 * localBits)) | ((uint64_t)l0 << localBits) | (uint64_t)l1;
    };
    auto flush = [&]() {
        // resolve the walk's leaves: entity and local index from their records (independent
        // loads), colliders of the lane's own entity dropped
        int4 rec[TRAV_SLOTS];
#pragma unroll
        for (int j = 0; j < TRAV_SLOTS; ++j)
            if ((treeMask >> j) & 1u) rec[j] = travC[n - 1 + (int)found[j][threadIdx.x]];
        int kept = 0;
#pragma unroll
        for (int j = 0; j < TRAV_SLOTS; ++j) {
            if (j >= held) break;
            uint64_t key = found[j][threadIdx.x];
            bool keep = true;
            if ((treeMask >> j) & 1u) {
                keep = (uint32_t)rec[j].y != myEnt;
                key = pairKey((uint32_t)rec[j].y, rec[j].z);
            }
            if (keep) found[kept++][threadIdx.x] = key;
        }
        int incl = kept;   // inclusive prefix sum of kept over the wave
        for (int o = 1; o < WAVE; o <<= 1) {
            int t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        int total = __shfl(incl, WAVE - 1);
        int base = 0;
        if (lane == WAVE - 1 && total) base = atomicAdd(pairCount, total);
        base = __shfl(base, WAVE - 1) + incl - kept;
        for (int j = 0; j < kept; ++j)
            if (base + j < cap) pkeys[base + j] = found[j][threadIdx.x];
        held = 0;
        treeMask = 0;
    };
    int visits = 0;
    const int maxVisits = 2 * n;   // a depth-first walk visits at most 2n - 1 records
    while (__ballot(active && (node != stop || st < stEnd))) {
        bool emitLeaf = false, emitStatic = false;
        int leaf = 0;
        uint64_t skey = 0;
        if (active && node != stop && ++visits > maxVisits) {
            atomicOr(&pairCount[6], 1);   // malformed tree: stop and report
            node = stop;
        }
        if (active && node != stop) {
            const float4 a = travA[node], b = travB[node];
            const bool take = overlaps(qmin, qmax, a, b);
            if (take && node >= n - 1) {
                emitLeaf = true;
                leaf = node - (n - 1);
            }
            node = take ? __float_as_int(a.w) : __float_as_int(b.w);
        }
        if (active && st < stEnd) {
            const float4 a = statA[st], b = statB[st];
            const int l = statL[st];
            const uint32_t en = (uint32_t)__float_as_int(b.w);
            if ((__float_as_int(a.w) & CF_ENABLE) && en != myEnt && overlaps(qmin, qmax, a, b)) {
                emitStatic = true;
                skey = pairKey(en, l);
            }
            ++st;
        }
        if (__ballot(held + 2 > TRAV_SLOTS)) flush();
        if (emitLeaf) {
            treeMask |= 1u << held;
            found[held++][threadIdx.x] = (uint64_t)leaf;
        }
        if (emitStatic) found[held++][threadIdx.x] = skey;
    }
    if (__ballot(held > 0)) flush();
}

__global__ void k_single_leaf_root(int32_t* worldRoot, int W, const uint64_t* mkeys, int worldShift) {
    // n == 1: the only leaf is its world's root
    if (threadIdx.x == 0 && blockIdx.x == 0) worldRoot[(int)(mkeys[0] >> worldShift)] = 0 | LEAF_BIT;
}

// The LBVH of every world's dynamic colliders from the current bounds: Morton keys,
// sorted leaves, Karras nodes, refit, escape links, packed traversal records.
void bp_build(World& w) {
    hipStream_t s = w.stream;
    int n = w.nDynCol;
    int W = w.W;
    int worldShift = 30;
    if (n > 0) {
        // quantisation box -> Morton codes of the dynamic colliders -> sort
        hipLaunchKernelGGL(k_scene_bounds_init, dim3(1), dim3(64), 0, s, w.worldBoundsI.p, w.sceneMoments.p,
                           nullptr);
        hipLaunchKernelGGL(k_scene_bounds, dim3(std::min(64, div_up(n, 1024))), dim3(1024), 0, s, n, w.dynCols.p,
                           w.cbmin.p, w.cbmax.p, w.worldBoundsI.p, w.sceneMoments.p);
        hipLaunchKernelGGL(k_morton, dim3(div_up(n, 256)), dim3(256), 0, s, n, w.dynCols.p, w.cworld.p, w.cbmin.p,
                           w.cbmax.p, w.worldBoundsI.p, w.sceneMoments.p, w.mkeys.p, w.mvals.p);
        PGX_LAUNCH_CHECK();
        radix_sort_u64(s, w.mkeys.p, w.mvals.p, n, 30 + w.worldBits, w.sortT);
        fill_i32(s, w.worldRoot.p, -1, W);
        hipLaunchKernelGGL(k_leaves, dim3(div_up(n, 256)), dim3(256), 0, s, n, w.mvals.p, w.cbmin.p, w.cbmax.p,
                           w.cflags.p, w.leafMin.p, w.leafMax.p, w.leafFlags.p);
        if (n == 1) {
            hipLaunchKernelGGL(k_single_leaf_root, dim3(1), dim3(64), 0, s, w.worldRoot.p, W, w.mkeys.p, worldShift);
            fill_i32(s, w.leafEsc.p, -1, 1);
            fill_i32(s, w.leafParent.p, -1, 1);
        } else {
            fill_i32(s, w.nodeParent.p, -1, n - 1);
            fill_i32(s, w.nodeCounter.p, 0, n - 1);
            hipLaunchKernelGGL(k_build, dim3(div_up(n - 1, 256)), dim3(256), 0, s, n, w.mkeys.p, w.nodeChild.p,
                               w.nodeRange.p, w.nodeParent.p, w.leafParent.p, w.worldRoot.p, worldShift);
            hipLaunchKernelGGL(k_refit_local, dim3(div_up(n, REFIT_BLOCK)), dim3(REFIT_BLOCK), 0, s, n, w.nodeChild.p,
                               w.nodeRange.p, w.nodeParent.p, w.leafParent.p, w.leafMin.p, w.leafMax.p, w.nodeMin.p,
                               w.nodeMax.p, w.nodeCounter.p);
            hipLaunchKernelGGL(k_refit_ready, dim3(div_up(n - 1, 256)), dim3(256), 0, s, n, w.nodeCounter.p,
                               w.refitReady.p, w.counters.p + 7);
            hipLaunchKernelGGL(k_refit_global, dim3(1), dim3(1024), 0, s, n, w.nodeChild.p, w.nodeParent.p,
                               w.leafMin.p, w.leafMax.p, w.nodeMin.p, w.nodeMax.p, w.nodeCounter.p, w.refitReady.p,
                               w.counters.p + 7);
            hipLaunchKernelGGL(k_escape, dim3(div_up(2 * n - 1, 256)), dim3(256), 0, s, n, w.nodeChild.p,
                               w.nodeParent.p, w.leafParent.p, w.nodeEsc.p, w.leafEsc.p);
        }
        hipLaunchKernelGGL(k_world_stop, dim3(div_up(W, 256)), dim3(256), 0, s, W, w.worldRoot.p, w.nodeEsc.p,
                           w.leafEsc.p, w.worldStop.p);
        const int nStatic = w.C - w.nDynCol;
        hipLaunchKernelGGL(k_pack, dim3(div_up(std::max(n, nStatic), 256)), dim3(256), 0, s, n, w.nodeMin.p,
                           w.nodeMax.p, w.nodeChild.p, w.nodeEsc.p, w.nodeRange.p, w.leafMin.p, w.leafMax.p,
                           w.leafEsc.p, w.leafFlags.p, w.mvals.p, w.centity.p, w.clocal.p, w.nodeA.p, w.nodeB.p,
                           w.nodeLast.p, w.leafA.p, w.leafB.p, w.leafI.p, w.travA.p, w.travB.p, w.travC.p, nStatic,
                           w.staticCols.p, w.cbmin.p, w.cbmax.p, w.cflags.p, w.statA.p, w.statB.p, w.statL.p);
        PGX_LAUNCH_CHECK();
    }
}

// The grid's hash table: at least twice the dynamic colliders (so at least twice the cells)
static uint32_t grid_table_size(int n) {
    uint32_t t = 64;
    while (t < 2u * (uint32_t)std::max(n, 1)) t <<= 1;
    return t;
}

// The step's pairs on the uniform grid: cell codes of the dynamic colliders, sorted; the
// cells' runs in the hash table; the grid search. (The tree walk's inputs are not built.)
static void bp_grid(World& w) {
    hipStream_t s = w.stream;
    const int n = w.nDynCol;
    const uint32_t T = grid_table_size(n);
    hipLaunchKernelGGL(k_scene_bounds_init, dim3(1), dim3(64), 0, s, w.worldBoundsI.p, w.sceneMoments.p,
                       w.counters.p);
    hipLaunchKernelGGL(k_scene_bounds, dim3(std::min(64, div_up(n, 1024))), dim3(1024), 0, s, n, w.dynCols.p,
                       w.cbmin.p, w.cbmax.p, w.worldBoundsI.p, w.sceneMoments.p);
    hipLaunchKernelGGL(k_cell_codes, dim3(div_up(std::max<int>(n, (int)T), 256)), dim3(256), 0, s, n, w.dynCols.p,
                       w.cworld.p, w.cbmin.p, w.cbmax.p, w.worldBoundsI.p, w.sceneMoments.p, w.mkeys.p, w.mvals.p,
                       (int)T, w.gridKey.p);
    PGX_LAUNCH_CHECK();
    radix_sort_u64(s, w.mkeys.p, w.mvals.p, n, 30 + w.worldBits, w.sortT);
    const int nStatic = w.C - w.nDynCol;
    hipLaunchKernelGGL(k_grid_pack, dim3(div_up(std::max(n, nStatic), 256)), dim3(256), 0, s, n, w.mvals.p,
                       w.cbmin.p, w.cbmax.p, w.cflags.p, w.centity.p, w.clocal.p, w.gridRec.p, nStatic, w.staticCols.p,
                       w.statA.p, w.statB.p, w.statL.p);
    hipLaunchKernelGGL(k_cell_insert, dim3(div_up(n, 256)), dim3(256), 0, s, n, w.mkeys.p, T - 1, w.gridKey.p,
                       w.gridStart.p);
    hipLaunchKernelGGL(k_cell_ends, dim3(div_up(n, 256)), dim3(256), 0, s, n, w.mkeys.p, T - 1, w.gridKey.p,
                       w.gridStart.p, w.gridEnd.p, w.counters.p);
    // 16 lanes per collider at every size: fewer lanes, each opening several cells one after
    // another, measured slower on the 857k-collider batch (506 against 390 us)
    const long long lanes = (long long)n * GRID_LANES;
    hipLaunchKernelGGL(k_grid_pairs, dim3((unsigned)div_up(lanes, GRID_BLOCK)), dim3(GRID_BLOCK), 0, s, n, w.mkeys.p,
                       w.gridRec.p, T - 1, w.gridKey.p, w.gridStart.p, w.gridEnd.p, w.worldStaticStart.p, w.statA.p,
                       w.statB.p, w.statL.p, w.localBits, w.counters.p, w.pairCap, w.pkeys.p);
    PGX_LAUNCH_CHECK();
}

int bp_find_pairs(World& w) {
    hipStream_t s = w.stream;
    int n = w.nDynCol;
    int worldShift = 30;
    // [6] traversal visit bound, [7] refit, [8] convex pairs, [9] EPA capacity errors,
    // [10] mesh pairs, [11] mesh manifold slots, [12] mesh capacity errors, [13] mesh
    // candidate triangles, [14] the grid handed the step to the tree walk (1) or lost a cell
    // (2, never expected), [15] the grid's fullest cell
    const bool grid = n > 0 && w.tune.pairSearch == 0;
    if (!grid) fill_i32(s, w.counters.p, 0, 16);   // the grid's first kernel clears them
    auto tree_walk = [&]() {
        bp_build(w);
        if (n > 0) {
            hipLaunchKernelGGL(k_traverse, dim3(div_up(n, 256)), dim3(256), 0, s, n, w.mkeys.p, worldShift, w.travA.p,
                               w.travB.p, w.travC.p, w.worldStop.p, w.worldStaticStart.p, w.statA.p, w.statB.p,
                               w.statL.p, w.localBits, w.counters.p, w.pairCap, w.pkeys.p);
            PGX_LAUNCH_CHECK();
        }
    };
    if (grid) bp_grid(w);
    else tree_walk();
    int cnt[16] = {0};
    w.readHost.alloc(READ_HOST_INTS);   // pinned: the copy is a plain DMA
    PGX_HIP(hipMemcpyAsync(w.readHost.p, w.counters.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
    PGX_HIP(hipStreamSynchronize(s));
    std::copy(w.readHost.p, w.readHost.p + 16, cnt);
    if (grid && (cnt[14] & 2)) throw PgxError{PGX_ERR_HIP, "broadphase grid lost a cell of its hash table"};
    if (grid && cnt[14]) {
        // a cell too full for the grid: the tree walk finds the step's pairs instead
        ++w.gridFallbacks;
        fill_i32(s, w.counters.p, 0, 1);
        tree_walk();
        PGX_HIP(hipMemcpyAsync(w.readHost.p, w.counters.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
        PGX_HIP(hipStreamSynchronize(s));
        std::copy(w.readHost.p, w.readHost.p + 16, cnt);
    }
    int nPairs = cnt[0];
    if (cnt[6]) throw PgxError{PGX_ERR_INVALID, "broadphase traversal exceeded its visit bound (malformed tree)"};
    if (nPairs > w.pairCap) throw PgxError{PGX_ERR_CAPACITY, "pair capacity exceeded: " + std::to_string(nPairs) +
                                                                 " > " + std::to_string(w.pairCap)};
    // canonical order: (world, local collider of the smaller entity, other local collider)
    radix_sort_u64(s, w.pkeys.p, w.pvals.p, nPairs, w.worldBits + 2 * w.localBits, w.sortT);
    return nPairs;
}

}  // namespace pgx
