// Contact setup, graph colouring and the TGS soft-step substep loop.
//
// Reference semantics (src/Physecs.cpp, src/ContactConstraints.cpp):
//   contact setup (material mix, body indices, body-local anchors, restitution target
//   with the (pair, triangle) cache matched within 0.1 m)      Physecs.cpp:215-315
//   update contact rows each substep                           Physecs.cpp:368-426
//   integrate velocities (gravity, implicit gyroscopic term)   Physecs.cpp:443-468
//   preSolve / solve (normal + one friction row per point)     ContactConstraints.cpp:4-124
//   integrate positions (COM pivot, pseudo velocities)         Physecs.cpp:494-511
//   relaxation (hard contacts, no bias) and write-back         Physecs.cpp:515-530
//
// The reference solves contact constraints sequentially in push order. Here the
// manifolds are graph-coloured on the device so that no two manifolds of one colour share
// a dynamic body; each colour is one kernel launch and is solved fully in parallel with
// the same result as a sequential pass over it in any order. The solve order is
// "colour 0, colour 1, ..." and is exported (pgx_world_get_manifolds) so that a
// sequential solver fed this order reproduces the device result bit for bit.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_ext.h>
#include "pgx_internal.h"

namespace cg = cooperative_groups;

namespace pgx {

// ------------------------------------------------------------------ compaction
// manifolds of pair p: one for a pair with points, meshCount[p] for a triangle-mesh pair
__global__ void k_flag_manifolds(int nPairs, const int32_t* __restrict__ npNum, const int32_t* __restrict__ meshCount,
                                 int32_t* __restrict__ flag) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nPairs) flag[p] = (npNum[p] > 0 ? 1 : 0) + meshCount[p];
}

__device__ __forceinline__ void pair_colliders(uint64_t key, int localBits, const int32_t* __restrict__ worldColStart,
                                               const int32_t* __restrict__ colOfLocal, int& c0, int& c1, int& wd) {
    uint64_t lmask = (1ull << localBits) - 1;
    wd = (int)(key >> (2 * localBits));
    int base = worldColStart[wd];
    c0 = colOfLocal[base + (int)((key >> localBits) & lmask)];
    c1 = colOfLocal[base + (int)(key & lmask)];
}

// Manifold list in pair order; a triangle-mesh pair contributes its manifolds in the order
// collisionTriangleMesh emitted them. mfSrc is the narrowphase data slot (the pair's own
// slot, or pairCap + the mesh manifold's slot), mfTri the triangle index (-1 otherwise).
__global__ void k_compact(int nPairs, const int32_t* __restrict__ npNum, const int32_t* __restrict__ npIndex,
                          const int32_t* __restrict__ meshFirst, const int32_t* __restrict__ meshCount,
                          const int32_t* __restrict__ npTri, int pairCap,
                          const uint64_t* __restrict__ pkeys, int localBits, const int32_t* __restrict__ worldColStart,
                          const int32_t* __restrict__ colOfLocal, const int32_t* __restrict__ cbody,
                          const int32_t* __restrict__ cflags, int32_t* __restrict__ mfPair,
                          uint64_t* __restrict__ mfKey, int32_t* __restrict__ mfB0, int32_t* __restrict__ mfB1,
                          int32_t* __restrict__ mfSrc, int32_t* __restrict__ mfTri) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nPairs) return;
    int nm = meshCount[p];
    if (npNum[p] == 0 && nm == 0) return;
    int m = npIndex[p];
    uint64_t key = pkeys[p];
    int c0, c1, wd;
    pair_colliders(key, localBits, worldColStart, colOfLocal, c0, c1, wd);
    int b0 = (cflags[c0] & CF_DYNAMIC) ? cbody[c0] : -1;
    int b1 = (cflags[c1] & CF_DYNAMIC) ? cbody[c1] : -1;
    if (nm == 0) {
        mfPair[m] = p; mfKey[m] = key; mfB0[m] = b0; mfB1[m] = b1; mfSrc[m] = p; mfTri[m] = -1;
        return;
    }
    int first = meshFirst[p];
    for (int j = 0; j < nm; ++j) {
        mfPair[m + j] = p; mfKey[m + j] = key; mfB0[m + j] = b0; mfB1[m + j] = b1;
        mfSrc[m + j] = pairCap + first + j;
        mfTri[m + j] = npTri[first + j];
    }
}

// ------------------------------------------------------------------ colouring
// Jones-Plassmann with largest-priority-first selection and smallest-available colour.
// Priorities are a bijective hash of the manifold's index within its world, so a world's
// colouring does not depend on the other worlds sharing the buffers.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// The colouring's start in one launch: each world's first manifold, and the colours, body
// claims, colour masks, per-round counts and flags cleared.
__global__ void k_color_begin(int M, int D, int rounds, const uint64_t* __restrict__ mfKey, int worldShift,
                              int32_t* __restrict__ wstart, int32_t* __restrict__ color,
                              uint64_t* __restrict__ claim, uint64_t* __restrict__ claim2, uint64_t* __restrict__ used,
                              int32_t* __restrict__ uncol, int32_t* __restrict__ overflow, int32_t* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) {
        const int wd = (int)(mfKey[i] >> worldShift);
        if (i == 0 || (int)(mfKey[i - 1] >> worldShift) != wd) wstart[wd] = i;
        color[i] = -1;
    }
    if (i < D) { claim[i] = 0ull; claim2[i] = 0ull; }
    if (i < 2 * D) used[i] = 0ull;
    if (i < rounds) uncol[i] = 0;
    if (i == 0) { *overflow = 0; *flag = 0; }
}

__global__ void k_color_claim(int M, int round, const uint64_t* __restrict__ mfKey, int worldShift,
                              const int32_t* __restrict__ wstart, const int32_t* __restrict__ color,
                              const int32_t* __restrict__ b0, const int32_t* __restrict__ b1,
                              unsigned long long* __restrict__ claim) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M || color[m] >= 0) return;
    int wd = (int)(mfKey[m] >> worldShift);
    unsigned long long v = ((unsigned long long)round << 32) | mix32((uint32_t)(m - wstart[wd]));
    int a = b0[m], b = b1[m];
    if (a >= 0) atomicMax(&claim[a], v);
    if (b >= 0) atomicMax(&claim[b], v);
}

// One Jones-Plassmann round r: a manifold whose round-r claim (priority tagged with r)
// holds on both its dynamic bodies takes the smallest colour free on both; every other
// uncoloured manifold claims its bodies for round r+1 in the other claim buffer and is
// marked in uncol[r]. Rounds after the last one that left manifolds uncoloured exit at
// once, so a batch of rounds can be queued without asking the host in between.
__global__ void k_color_round(int M, int round, const uint64_t* __restrict__ mfKey, int worldShift,
                              const int32_t* __restrict__ wstart, int32_t* __restrict__ color,
                              const int32_t* __restrict__ b0, const int32_t* __restrict__ b1,
                              const unsigned long long* __restrict__ claimIn, unsigned long long* __restrict__ claimOut,
                              uint64_t* __restrict__ used, int32_t* __restrict__ uncol, int32_t* __restrict__ counters) {
    if (round > 1 && uncol[round - 1] == 0) return;
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M || color[m] >= 0) return;
    int wd = (int)(mfKey[m] >> worldShift);
    uint32_t prio = mix32((uint32_t)(m - wstart[wd]));
    unsigned long long v = ((unsigned long long)round << 32) | prio;
    int a = b0[m], b = b1[m];
    bool won = (a < 0 || claimIn[a] == v) && (b < 0 || claimIn[b] == v);
    if (!won) {
        unsigned long long vn = ((unsigned long long)(round + 1) << 32) | prio;
        if (a >= 0) atomicMax(&claimOut[a], vn);
        if (b >= 0) atomicMax(&claimOut[b], vn);
        uncol[round] = 1;   // any writer writes 1: no atomic needed
        return;
    }
    uint64_t lo = 0, hi = 0;
    if (a >= 0) { lo |= used[2 * a]; hi |= used[2 * a + 1]; }
    if (b >= 0) { lo |= used[2 * b]; hi |= used[2 * b + 1]; }
    int c;
    if (~lo) c = __ffsll((long long)~lo) - 1;
    else if (~hi) c = 64 + __ffsll((long long)~hi) - 1;
    else { c = 127; atomicAdd(&counters[5], 1); }   // overflow: reported as an error by the host
    color[m] = c;
    uint64_t bl = c < 64 ? (1ull << c) : 0ull, bh = c >= 64 ? (1ull << (c - 64)) : 0ull;
    if (a >= 0) { used[2 * a] |= bl; used[2 * a + 1] |= bh; }
    if (b >= 0) { used[2 * b] |= bl; used[2 * b + 1] |= bh; }
}

// The previous step's colour of each manifold that persists (same pair key and triangle,
// found in the previous step's contact cache as k_setup finds it), and its bodies' used
// colours. Persisting manifolds were coloured properly among themselves last step, so
// their colours stay valid unless a body changed from kinematic to dynamic in between: a
// colour set twice on one body raises the clash flag, and k_color_clash then drops every
// inherited colour (the step colours from scratch). Only new manifolds go through the
// claim rounds.
constexpr int CACHE_NUM_MASK = 255, CACHE_COLOR_SHIFT = 8, CACHE_HAS_COLOR = 0x8000;
__global__ void k_color_inherit(int M, const uint64_t* __restrict__ mfKey, const int32_t* __restrict__ mfTri,
                                const uint64_t* __restrict__ prevKey, const int32_t* __restrict__ prevTri,
                                const int32_t* __restrict__ prevNum, int prevCount, const int32_t* __restrict__ b0,
                                const int32_t* __restrict__ b1, int32_t* __restrict__ color,
                                int32_t* __restrict__ mfPrev, unsigned long long* __restrict__ used,
                                int32_t* __restrict__ clash) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const uint64_t key = mfKey[m];
    const int tri = mfTri[m];
    int prev = -1;
    {   // the cache is in manifold order, so a pair's manifolds are one run of equal keys
        int lo = 0, hi = prevCount;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (prevKey[mid] < key) lo = mid + 1; else hi = mid;
        }
        for (; lo < prevCount && prevKey[lo] == key; ++lo)
            if (prevTri[lo] == tri) { prev = lo; break; }
    }
    mfPrev[m] = prev;
    int c = -1;
    if (prev >= 0) {
        const int pn = prevNum[prev];
        if (pn & CACHE_HAS_COLOR) c = (pn >> CACHE_COLOR_SHIFT) & 127;
    }
    color[m] = c;
    if (c < 0) return;
    const unsigned long long bit = 1ull << (c & 63);
    const int word = c >> 6;
    const int a = b0[m], b = b1[m];
    bool twice = false;
    if (a >= 0) twice |= (atomicOr(&used[2 * a + word], bit) & bit) != 0;
    if (b >= 0) twice |= (atomicOr(&used[2 * b + word], bit) & bit) != 0;
    if (twice) *clash = 1;
}

__global__ void k_color_clash(int M, int D, const int32_t* __restrict__ clash, int32_t* __restrict__ color,
                              uint64_t* __restrict__ used) {
    if (!*clash) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) color[i] = -1;
    if (i < 2 * D) used[i] = 0;
}

// Solve order: by colour, and within a colour by point count (byPoints), so the lanes of a
// wave solve manifolds of one point count and run the same number of rows. A colour's
// manifolds share no dynamic body, so their order within the colour changes no result. The
// per-world passes keep each world's manifolds of a colour together instead (byPoints 0).
constexpr int COLOR_KEY_SHIFT = 2;
__global__ void k_color_keys(int M, const int32_t* __restrict__ color, const int32_t* __restrict__ npNum,
                             const int32_t* __restrict__ mfSrc, int byPoints, uint64_t* __restrict__ keys,
                             int32_t* __restrict__ vals) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const int num = byPoints ? min(max(npNum[mfSrc[m]], 1), 4) : 1;
    keys[m] = ((uint64_t)color[m] << COLOR_KEY_SHIFT) | (uint64_t)(num - 1);
    vals[m] = m;
}

// start[c] = the first solve position with colour >= c, for every c in [0, n): colours with
// no manifold (inherited colours leave gaps) get empty ranges
// One workgroup. Also packs what the host reads back after the colouring rounds into one
// block, so a single copy brings it: [0, n) the starts, [n] the colour-overflow flag,
// [n+1, n+2] the last manifold's key (low, high word), [n+3, n+3+rounds] the count of
// uncoloured manifolds after each round.
__global__ void k_color_start(int M, int n, const uint64_t* __restrict__ keys, int32_t* __restrict__ start,
                              int rounds, const int32_t* __restrict__ uncol, const int32_t* __restrict__ overflow,
                              int32_t* __restrict__ pack) {
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        int lo = 0, hi = M;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (keys[mid] < ((uint64_t)c << COLOR_KEY_SHIFT)) lo = mid + 1; else hi = mid;
        }
        start[c] = lo;
        pack[c] = lo;
    }
    for (int r = threadIdx.x; r <= rounds; r += blockDim.x) pack[n + 3 + r] = uncol[r];
    if (threadIdx.x == 0) {
        pack[n] = *overflow;
        const uint64_t last = M > 0 ? keys[M - 1] : 0ull;
        pack[n + 1] = (int32_t)(uint32_t)last;
        pack[n + 2] = (int32_t)(uint32_t)(last >> 32);
    }
}

// ------------------------------------------------------------------ setup (solve order)

// each manifold's slot in solve order (the inverse of the colour sort's permutation)
__global__ void k_color_rank(int M, const int32_t* __restrict__ perm, int32_t* __restrict__ rank) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < M) rank[perm[s]] = s;
}

// One thread per manifold in manifold order, so its reads of the manifold arrays are
// coalesced; the solve-order arrays it writes are scattered instead (stores do not stall).
__global__ __launch_bounds__(256) void k_setup(
    int M, int Mcap, const int32_t* __restrict__ rank, const int32_t* __restrict__ mfSrc,
    const int32_t* __restrict__ mfTri, const uint64_t* __restrict__ mfKey, int localBits, const int32_t* __restrict__ worldColStart,
    const int32_t* __restrict__ colOfLocal, const int32_t* __restrict__ cbody, const int32_t* __restrict__ cflags,
    const float4* __restrict__ cmat, const int32_t* __restrict__ npNum, const float4* __restrict__ npNormal,
    const float4* __restrict__ npPts, const int32_t* __restrict__ mfColor, const int32_t* __restrict__ mfB0,
    const int32_t* __restrict__ mfB1, const float4* __restrict__ bpos, const float4* __restrict__ bquat,
    const float4* __restrict__ dvel, const float4* __restrict__ dang, const float4* __restrict__ dcom, int D,
    const int32_t* __restrict__ mfPrev, const int32_t* __restrict__ prevNum,
    const float4* __restrict__ prevLp, int prevCap, uint64_t* __restrict__ curKey,
    int32_t* __restrict__ curTri, int32_t* __restrict__ curNum,
    float4* __restrict__ curLp, int32_t* __restrict__ sB0, int32_t* __restrict__ sB1, int32_t* __restrict__ sA0,
    int32_t* __restrict__ sA1, int32_t* __restrict__ sFlags, int32_t* __restrict__ sCanon, float4* __restrict__ sN,
    float2* __restrict__ sSoft, float4* __restrict__ sR0, float4* __restrict__ sR1, int32_t* __restrict__ sWorld) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const int s = rank[m];
    int p = mfSrc[m];
    int tri = mfTri[m];
    uint64_t key = mfKey[m];
    int c0, c1, wd;
    pair_colliders(key, localBits, worldColStart, colOfLocal, c0, c1, wd);
    sWorld[s] = wd;
    int a0 = cbody[c0], a1 = cbody[c1];
    int b0 = mfB0[m], b1 = mfB1[m];
    float4 m0 = cmat[c0], m1 = cmat[c1];   // friction, restitution, damping
    float friction = (m0.x + m1.x) * 0.5f;
    bool soft;
    float freq, damp, rest;
    if (m0.z != 0.f || m1.z != 0.f) {
        soft = true;
        if (m0.z != 0.f && m1.z != 0.f) { freq = gmin(m0.y, m1.y); damp = gmin(m0.z, m1.z); }
        else if (m0.z != 0.f) { freq = m0.y; damp = m0.z; }
        else { freq = m1.y; damp = m1.z; }
        rest = 0.f;
    } else {
        soft = false; freq = 0.f; damp = 0.f;
        rest = (m0.y + m1.y) * 0.5f;
    }
    f3 n = xyz(npNormal[p]);
    int num = npNum[p];
    qf q0 = q4(bquat[a0]), q1 = q4(bquat[a1]);
    f3 com0 = s3(0.f), v0 = s3(0.f), w0 = s3(0.f), com1 = s3(0.f), v1 = s3(0.f), w1 = s3(0.f);
    if (b0 >= 0) { com0 = add(xyz(bpos[a0]), qrot(q0, xyz(dcom[b0]))); v0 = xyz(dvel[b0]); w0 = xyz(dang[b0]); }
    if (b1 >= 0) { com1 = add(xyz(bpos[a1]), qrot(q1, xyz(dcom[b1]))); v1 = xyz(dvel[b1]); w1 = xyz(dang[b1]); }
    // previous step's manifold of this (pair, triangle), src/Physecs.cpp:237, as
    // k_color_inherit found it
    const int prev = mfPrev[m];
    qf iq0 = qinverse(q0), iq1 = qinverse(q1);
    for (int k = 0; k < 4; ++k) {
        float4 r0o = make_float4(0.f, 0.f, 0.f, 0.f), r1o = r0o;
        if (k < num) {
            f3 p0 = xyz(npPts[8 * (size_t)p + 2 * k]), p1 = xyz(npPts[8 * (size_t)p + 2 * k + 1]);
            f3 r0 = sub(p0, com0), r1 = sub(p1, com1);
            f3 rel = sub(sub(add(v1, cross(w1, r1)), v0), cross(w0, r0));
            float relN = dot(rel, n);
            r0 = qrot(iq0, r0);
            r1 = qrot(iq1, r1);
            float tv = 0.f;
            bool found = false;
            if (prev >= 0) {
                int pn = prevNum[prev] & CACHE_NUM_MASK;
                for (int i = 0; i < pn; ++i) {
                    float4 lp = prevLp[(size_t)i * prevCap + prev];
                    f3 d = sub(xyz(lp), r0);
                    // glm::distance(...) < 0.1 is a double comparison (Physecs.cpp:294)
                    if ((double)sqrtf(dot(d, d)) < 0.1) { found = true; tv = lp.w; break; }
                }
            }
            if (!found) tv = -rest * relN;
            r0o = f4(r0, tv);
            r1o = f4(r1, 0.f);
            curLp[(size_t)k * Mcap + m] = r0o;
        }
        sR0[(size_t)k * Mcap + s] = r0o;
        sR1[(size_t)k * Mcap + s] = r1o;
    }
    curKey[m] = key;
    curTri[m] = tri;
    curNum[m] = num | (mfColor[m] << CACHE_COLOR_SHIFT) | CACHE_HAS_COLOR;
    sB0[s] = b0;
    sB1[s] = b1;
    sA0[s] = a0;
    sA1[s] = a1;
    sFlags[s] = num | (soft ? 256 : 0) | (mfColor[m] << 16);
    sCanon[s] = m;
    sN[s] = f4(n, friction);
    sSoft[s] = make_float2(freq, damp);
}

// contact points per colour (statistics). The manifolds are in colour order, so each
// thread sums a run of POINTS_RUN consecutive manifolds and adds to its block's per-colour
// counters in LDS when the colour changes; each block then adds its non-zero counters to
// the output, which starts at zero.
constexpr int POINTS_RUN = 16;
__global__ __launch_bounds__(256) void k_color_points(int M, const int32_t* __restrict__ sFlags,
                                                      int32_t* __restrict__ out) {
    __shared__ int32_t cnt[128];
    if (threadIdx.x < 128) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int first = (blockIdx.x * blockDim.x + threadIdx.x) * POINTS_RUN;
    int colour = -1, sum = 0;
    for (int i = first; i < min(M, first + POINTS_RUN); ++i) {
        const int f = sFlags[i], c = (f >> 16) & 127;
        if (c != colour) {
            if (sum) atomicAdd(&cnt[colour], sum);
            colour = c;
            sum = 0;
        }
        sum += f & 255;
    }
    if (sum) atomicAdd(&cnt[colour], sum);
    __syncthreads();
    if (threadIdx.x < 128 && cnt[threadIdx.x]) atomicAdd(&out[threadIdx.x], cnt[threadIdx.x]);
}

void sv_color_points(World& w, int nColors) {
    if (nColors <= 0) return;
    const int M = w.colorStartHost.back();
    fill_i32(w.stream, w.colorPoints.p, 0, 160);
    hipLaunchKernelGGL(k_color_points, dim3(std::max(1, div_up(M, 256 * POINTS_RUN))), dim3(256), 0, w.stream, M,
                       w.sFlags.p, w.colorPoints.p);
    PGX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ substep pieces
// Contact rows of point k of manifold slot s, field f at ((k * R_NF + f) * Mcap + s):
// r0 x n (.w = c), r1 x n (.w = inverse effective mass N), tangent (.w = inverse
// effective mass T), r0 x t (.w = effective mass N, 1 / invN), r1 x t (.w = the bias
// 0.1 c / h), and the rows times the bodies' world inverse inertia, as preSolve keeps them
// (src/ContactConstraints.cpp:21-24): r0xnt (.w = beta c / h of a soft contact), r1xnt
// (.w = invN + gamma / h of a soft contact), r0xtt, r1xtt. Everything here is fixed for the
// substep, so the solve's arithmetic on it is done once, with the operations and in the
// order ContactConstraints::solve does it (src/ContactConstraints.cpp:54-68). Per point
// also (lambda N, lambda T, target velocity) in sLam.
enum { R_R0XN = 0, R_R1XN, R_T, R_R0XT, R_R1XT, R_R0XNT, R_R1XNT, R_R0XTT, R_R1XTT, R_NF };
static_assert(R_NF == CONTACT_ROW_FIELDS, "sRow holds CONTACT_ROW_FIELDS float4 per point");
#define ROW(f, k) a.sRow[((size_t)((k) * R_NF + (f))) * a.Mcap + s]
#define LAM(k) a.sLam[(size_t)(k) * a.Mcap + s]

struct SubArgs {
    int M, Mcap, D, nColors, substeps, iterations;
    float h, g;
    const int32_t *sB0, *sB1, *sA0, *sA1, *sFlags, *colorStart;
    const float4 *sN, *sR0, *sR1;
    const float2* sSoft;
    float4 *sRow, *sLam;
    float4 *bpos, *bquat;                 // body poses (dynamic bodies come first)
    float4 *dvel, *dang;                  // component velocity at the start of the substep (dvel.w = invMass)
    float4* bvel;                         // per dynamic body: solver v, w (velocityTemp), component
                                          // v, w after gravity + gyroscopic term (friction reads them)
    const float4 *dcom, *dinvI;
    float4 *pvel, *pang;                  // pseudo velocities (pang.w = constraint count)
    // dependency-driven passes: per body, the colours of its manifolds (2 x 64 bits); a work
    // ticket per pass; a flag set if a wait ran out (never expected)
    const uint64_t* used;
    int32_t* tickets;
    int32_t* flowErr;
    int32_t* qDone;                       // manifolds finished per queue in this step
    int32_t* barrier;                     // arrivals at the pass kernel's grid barriers (this substep)
    int32_t* xctl;                        // control blocks of the one-XCD pass kernel, one per pass
    int xctlWords;                        // words of xctl cleared by k_prepare
};

__device__ __forceinline__ m3 world_inv_inertia(const SubArgs& a, int b) {
    m3 rot = mat3_cast(q4(a.bquat[b]));
    m3 invI = mk_m3(xyz(a.dinvI[3 * b]), xyz(a.dinvI[3 * b + 1]), xyz(a.dinvI[3 * b + 2]));
    return mulmm(mulmm(rot, invI), transpose(rot));   // rot * invI * invRot, src/Physecs.cpp:467
}

// update contact rows (src/Physecs.cpp:368-426) then preSolve (src/ContactConstraints.cpp:4-31)
// of one manifold; the world inverse inertia is rebuilt with the same operations as massTemp.
__device__ void prepare_manifold(const SubArgs& a, int s) {
    int b0 = a.sB0[s], b1 = a.sB1[s], a0 = a.sA0[s], a1 = a.sA1[s];
    int num = a.sFlags[s] & 255;
    f3 n = xyz(a.sN[s]);
    qf q0 = q4(a.bquat[a0]), q1 = q4(a.bquat[a1]);
    f3 com0 = s3(0.f), v0 = s3(0.f), w0 = s3(0.f), com1 = s3(0.f), v1 = s3(0.f), w1 = s3(0.f);
    float im0 = 0.f, im1 = 0.f;
    m3 I0 = mzero(), I1 = mzero();
    if (b0 >= 0) {
        com0 = add(xyz(a.bpos[a0]), qrot(q0, xyz(a.dcom[b0])));
        float4 vv = a.dvel[b0];
        v0 = xyz(vv); im0 = vv.w; w0 = xyz(a.dang[b0]);
        I0 = world_inv_inertia(a, b0);
    }
    if (b1 >= 0) {
        com1 = add(xyz(a.bpos[a1]), qrot(q1, xyz(a.dcom[b1])));
        float4 vv = a.dvel[b1];
        v1 = xyz(vv); im1 = vv.w; w1 = xyz(a.dang[b1]);
        I1 = world_inv_inertia(a, b1);
    }
    float nn = dot(n, n);
    const int sflags = a.sFlags[s];
    const bool soft = (sflags & 256) != 0;
    const float2 sp = a.sSoft[s];
    const float h = a.h;
    for (int k = 0; k < num; ++k) {
        f3 r0 = qrot(q0, xyz(a.sR0[(size_t)k * a.Mcap + s]));
        f3 r1 = qrot(q1, xyz(a.sR1[(size_t)k * a.Mcap + s]));
        f3 cp0 = add(com0, r0), cp1 = add(com1, r1);
        float cn = dot(sub(cp1, cp0), n);
        f3 r0xn = cross(r0, n), r1xn = cross(r1, n);
        f3 rel = sub(sub(add(v1, cross(w1, r1)), v0), cross(w0, r0));
        float relN = dot(rel, n);
        f3 t = sub(rel, mul(relN, n));
        float tl = length(t);
        if (tl != 0.f) t = divs(t, tl);
        f3 r0xt = cross(r0, t), r1xt = cross(r1, t);
        f3 r0xnt = mulmv(I0, r0xn), r1xnt = mulmv(I1, r1xn), r0xtt = mulmv(I0, r0xt), r1xtt = mulmv(I1, r1xt);
        float invN = nn * (im0 + im1) + dot(r0xn, r0xnt) + dot(r1xn, r1xnt);
        float invT = dot(t, t) * (im0 + im1) + dot(r0xt, r0xtt) + dot(r1xt, r1xtt);
        // src/ContactConstraints.cpp:54-68 on the substep's fixed values
        const float effMass = 1.f / invN;
        const float bias = 0.1f * cn / h;
        float softA = 0.f, softB = 0.f;
        if (soft) {
            float af = 2.f * 3.14159265358979323846f * sp.x;
            float stiffness = af * af * effMass;
            float damping = 2.f * af * sp.y * effMass;
            float gamma = 1.f / (damping + h * stiffness);
            float beta = h * stiffness / (damping + h * stiffness);
            softA = beta * cn / h;
            softB = invN + gamma / h;
        }
        ROW(R_R0XN, k) = f4(r0xn, cn);
        ROW(R_R1XN, k) = f4(r1xn, invN);
        ROW(R_T, k) = f4(t, invT);
        ROW(R_R0XT, k) = f4(r0xt, effMass);
        ROW(R_R1XT, k) = f4(r1xt, bias);
        ROW(R_R0XNT, k) = f4(r0xnt, softA);
        ROW(R_R1XNT, k) = f4(r1xnt, softB);
        ROW(R_R0XTT, k) = f4(r0xtt, 0.f);
        ROW(R_R1XTT, k) = f4(r1xtt, 0.f);
        LAM(k) = make_float4(0.f, 0.f, a.sR0[(size_t)k * a.Mcap + s].w, 0.f);   // lambdas start at 0
    }
}

// src/MathUtil.h:10-21
__device__ __forceinline__ f3 solve33(const m3& A, f3 b) {
    f3 c12 = cross(A.c1, A.c2);
    float det = dot(A.c0, c12);
    if (det == 0.f) return s3(0.f);
    float id = 1.f / det;
    return mk3(id * dot(b, c12), id * dot(A.c0, cross(b, A.c2)), id * dot(A.c0, cross(A.c1, b)));
}

// integrate velocities of one body, src/Physecs.cpp:443-468
__device__ void integrate_velocity(const SubArgs& a, int d) {
    if (a.dcom[d].w != 0.f) return;   // kinematic
    m3 rot = mat3_cast(q4(a.bquat[d]));
    m3 invRot = transpose(rot);
    float4 vv = a.dvel[d];
    f3 v = add(xyz(vv), mul(a.h, mk3(0.f, -a.g, 0.f)));
    m3 invI = mk_m3(xyz(a.dinvI[3 * d]), xyz(a.dinvI[3 * d + 1]), xyz(a.dinvI[3 * d + 2]));
    f3 wl = mulmv(invRot, xyz(a.dang[d]));
    m3 I = inverse(invI);
    f3 Iw = mulmv(I, wl);
    f3 f = mul(a.h, cross(wl, Iw));
    m3 J = madd(I, mscale(a.h, msub(mulmm(matrixCross3(wl), I), matrixCross3(Iw))));
    wl = sub(wl, solve33(J, f));
    f3 w = mulmv(rot, wl);
    a.bvel[bv(d, BV_CV)] = f4(v, vv.w);
    a.bvel[bv(d, BV_CW)] = f4(w, 0.f);
    a.bvel[bv(d, BV_V)] = f4(v, 0.f);
    a.bvel[bv(d, BV_W)] = f4(w, 0.f);
    const m3 wI = world_inv_inertia(a, d);   // the inertia prepare_manifold uses this substep
    a.bvel[bv(d, BV_I0)] = f4(wI.c0, 0.f);
    a.bvel[bv(d, BV_I1)] = f4(wI.c1, 0.f);
    a.bvel[bv(d, BV_I2)] = f4(wI.c2, 0.f);
    a.pvel[d] = make_float4(0.f, 0.f, 0.f, 0.f);
    a.pang[d] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// One manifold's normal then friction rows, src/ContactConstraints.cpp:33-124. Manifolds of
// one colour touch disjoint dynamic bodies, so a colour is solved in parallel exactly as a
// sequential pass over it would be.
// Everything one manifold's solve reads except the bodies' solver velocities: the
// manifold's rows, normal, friction, cached target velocities and lambdas, and both
// bodies' inverse masses and component velocities. All of it is fixed for the whole pass.
struct MInputs {
    int num, b0, b1;
    bool soft;
    f3 n;
    float friction, im0, im1;
    f3 cv0, cw0, cv1, cw1;
    float4 row[4][R_NF], lam[4];
};

// the manifold's own inputs: flags, bodies, normal; point slots' rows and lambdas (all four
// slots when ALL, else slot 0 with the flags and slots 1-3 once the point count is known)
template <bool ALL>
__device__ __forceinline__ void load_own(const SubArgs& a, int s, MInputs& in) {
    const int flags = a.sFlags[s];
    in.soft = (flags & 256) != 0;
    in.num = flags & 255;
    in.b0 = a.sB0[s];
    in.b1 = a.sB1[s];
    const float4 nf = a.sN[s];
    in.n = xyz(nf);
    in.friction = nf.w;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (ALL || k == 0 || k < in.num) {
#pragma unroll
            for (int f = 0; f < R_NF; ++f) in.row[k][f] = ROW(f, k);
            in.lam[k] = LAM(k);
        }
    }
}

// EAGER loads point slot 0's rows (every manifold has a first point) with the manifold's
// flags and body indices, without waiting for the point count; slots 1-3 wait for the
// count, in the same round of loads as the body records (which wait for the body indices
// anyway), so unused slots cost no traffic and the solve still takes two dependent rounds.
template <bool EAGER = false>
__device__ __forceinline__ void load_inputs(const SubArgs& a, int s, MInputs& in) {
    int flags = a.sFlags[s];
    in.soft = (flags & 256) != 0;
    in.num = flags & 255;
    in.b0 = a.sB0[s];
    in.b1 = a.sB1[s];
    float4 nf = a.sN[s];
    in.n = xyz(nf);
    in.friction = nf.w;
    in.im0 = 0.f; in.im1 = 0.f;
    in.cv0 = s3(0.f); in.cw0 = s3(0.f); in.cv1 = s3(0.f); in.cw1 = s3(0.f);
    if (in.b0 >= 0) {
        float4 c = a.bvel[bv(in.b0, BV_CV)];
        in.im0 = c.w; in.cv0 = xyz(c); in.cw0 = xyz(a.bvel[bv(in.b0, BV_CW)]);
    }
    if (in.b1 >= 0) {
        float4 c = a.bvel[bv(in.b1, BV_CV)];
        in.im1 = c.w; in.cv1 = xyz(c); in.cw1 = xyz(a.bvel[bv(in.b1, BV_CW)]);
    }
    // all rows of the manifold are loaded before any dependent arithmetic (memory-level
    // parallelism); unused point slots are never read back
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if ((EAGER && k == 0) || k < in.num) {
#pragma unroll
            for (int f = 0; f < R_NF; ++f) in.row[k][f] = ROW(f, k);
            in.lam[k] = LAM(k);
        }
    }
}

// ContactConstraints::solve (src/ContactConstraints.cpp:33-124) of one manifold on the
// bodies' solver velocities v0, w0, v1, w1 (updated in place); stores the accumulated
// lambdas in the manifold's rows. The effective mass, the bias and a soft contact's terms
// come from the rows (computed once per substep with the same operations).
__device__ __forceinline__ void solve_core(const SubArgs& a, int s, const MInputs& in, bool useBias, float h, f3& v0,
                                           f3& w0, f3& v1, f3& w1) {
    (void)h;
    const int num = in.num, b0 = in.b0, b1 = in.b1;
    const f3 n = in.n;
    const float im0 = in.im0, im1 = in.im1;
    float lN[4], lT[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= num) break;
        const float4* r = in.row[k];
        lN[k] = in.lam[k].x;
        float invN = r[R_R1XN].w;
        if (invN == 0.f) continue;
        f3 r0xn = xyz(r[R_R0XN]), r1xn = xyz(r[R_R1XN]);
        float rv = dot(neg(n), v0) + dot(neg(r0xn), w0) + dot(n, v1) + dot(r1xn, w1);
        float lambda;
        if (in.soft) lambda = (rv + r[R_R0XNT].w) / r[R_R1XNT].w;
        else lambda = (rv - in.lam[k].z + (useBias ? r[R_R1XT].w : 0.f)) * r[R_R0XT].w;
        float prev = lN[k];
        lN[k] = lN[k] + lambda;
        lN[k] = gmin(lN[k], 0.f);
        lambda = lN[k] - prev;
        if (b0 >= 0) { v0 = add(v0, mul(lambda * im0, n)); w0 = add(w0, mul(lambda, xyz(r[R_R0XNT]))); }
        if (b1 >= 0) { v1 = sub(v1, mul(lambda * im1, n)); w1 = sub(w1, mul(lambda, xyz(r[R_R1XNT]))); }
    }
    // friction rows read the component velocities (src/ContactConstraints.cpp:142-152)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= num) break;
        const float4* r = in.row[k];
        lT[k] = in.lam[k].y;
        float invT = r[R_T].w;
        if (invT == 0.f) continue;
        f3 t = xyz(r[R_T]), r0xt = xyz(r[R_R0XT]), r1xt = xyz(r[R_R1XT]);
        float rv = dot(neg(t), in.cv0) + dot(neg(r0xt), in.cw0) + dot(t, in.cv1) + dot(r1xt, in.cw1);
        float lambda = rv / invT;
        float limit = in.friction * lN[k];
        float prev = lT[k];
        lT[k] = lT[k] + lambda;
        lT[k] = gclamp(lT[k], limit, -limit);
        lambda = lT[k] - prev;
        if (b0 >= 0) { v0 = add(v0, mul(lambda * im0, t)); w0 = add(w0, mul(lambda, xyz(r[R_R0XTT]))); }
        if (b1 >= 0) { v1 = sub(v1, mul(lambda * im1, t)); w1 = sub(w1, mul(lambda, xyz(r[R_R1XTT]))); }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= num) break;
        *reinterpret_cast<float2*>(&LAM(k)) = make_float2(lN[k], lT[k]);
    }
}

// one manifold of a colour launch: every other manifold touching its bodies is in another
// launch, so plain accesses see the previous colours' results
template <bool EAGER = false>
__device__ void solve_manifold(const SubArgs& a, int s, bool useBias, float h, bool skipSoft) {
    MInputs in;
    load_inputs<EAGER>(a, s, in);
    if (skipSoft && in.soft) return;
    f3 v0 = s3(0.f), w0 = s3(0.f), v1 = s3(0.f), w1 = s3(0.f);
    if (in.b0 >= 0) { v0 = xyz(a.bvel[bv(in.b0, BV_V)]); w0 = xyz(a.bvel[bv(in.b0, BV_W)]); }
    if (in.b1 >= 0) { v1 = xyz(a.bvel[bv(in.b1, BV_V)]); w1 = xyz(a.bvel[bv(in.b1, BV_W)]); }
    solve_core(a, s, in, useBias, h, v0, w0, v1, w1);
    if (in.b0 >= 0) { a.bvel[bv(in.b0, BV_V)] = f4(v0, 0.f); a.bvel[bv(in.b0, BV_W)] = f4(w0, 0.f); }
    if (in.b1 >= 0) { a.bvel[bv(in.b1, BV_V)] = f4(v1, 0.f); a.bvel[bv(in.b1, BV_W)] = f4(w1, 0.f); }
}

// integrate positions of one body, src/Physecs.cpp:494-511
__device__ void integrate_position(const SubArgs& a, int d) {
    float4 cm = a.dcom[d];
    if (cm.w != 0.f) return;
    float4 pa = a.pang[d];
    float cnt = pa.w;
    float sc = cnt != 0.f ? 1.f / cnt : 1.f;
    f3 com = xyz(cm);
    f3 pos = xyz(a.bpos[d]);
    qf q = q4(a.bquat[d]);
    pos = add(pos, add(mul(a.h, xyz(a.bvel[bv(d, BV_V)])), mul(sc, xyz(a.pvel[d]))));
    f3 prevCom = qrot(q, com);
    f3 hv = mul(0.5f, add(mul(a.h, xyz(a.bvel[bv(d, BV_W)])), mul(sc, xyz(pa))));
    qf dq = qmul(mkq(0.f, hv.x, hv.y, hv.z), q);
    q = mkq(q.w + dq.w, q.x + dq.x, q.y + dq.y, q.z + dq.z);
    q = qnormalize(q);
    pos = add(pos, sub(prevCom, qrot(q, com)));
    a.bpos[d] = f4(pos, 0.f);
    a.bquat[d] = f4q(q);
}

// write back, src/Physecs.cpp:523-530
__device__ void write_back(const SubArgs& a, int d) {
    if (a.dcom[d].w != 0.f) return;
    float4 v = a.bvel[bv(d, BV_V)];
    a.dvel[d] = make_float4(v.x, v.y, v.z, a.dvel[d].w);
    a.dang[d] = a.bvel[bv(d, BV_W)];
}

// ------------------------------------------------------------------ discrete kernels
__global__ __launch_bounds__(256) void k_prepare(SubArgs a, int sub) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (a.iterations + 1) * FLOW_QUEUES) a.tickets[i] = 0;
    if (i == 0) a.barrier[0] = 0;
    for (int j = i; j < a.xctlWords; j += gridDim.x * blockDim.x) a.xctl[j] = 0;
    if (sub == 0 && i < FLOW_QUEUES) a.qDone[i] = 0;
    if (i < a.M) prepare_manifold(a, i);
    else if (i - a.M < a.D) integrate_velocity(a, i - a.M);
}
// XCD-homed slots. Workgroups are dispatched round-robin to the 8 XCDs (workgroup b on XCD
// b % 8), each with its own L2. k_prepare writes slot s's rows from workgroup s / 256, so
// from XCD (s / 256) % 8, the slot's home; the colour launches give each workgroup a chunk
// of BLOCK slots homed on its own XCD, so the rows (and the lambdas, rewritten in place every
// pass) are read from that XCD's L2 rather than from memory. Chunk k holds slots
// [k * BLOCK, (k + 1) * BLOCK); its home is (k / q) % 8 with q = 256 / BLOCK chunks per
// prepare workgroup, so home-x chunks come in runs of q, one run every 8q chunks. Workgroup
// b (XCD x = b % 8) takes the (b / 8)-th home-x chunk at or after chunk k0.
__device__ __forceinline__ int homed_chunk(int b, int k0, int q) {
    const int x = b & 7, i = b >> 3, period = 8 * q;
    const int g = k0 / period, r = k0 - g * period;
    const int m = q * g + min(max(r - q * x, 0), q) + i;   // home-x chunks below k0, plus i
    return period * (m / q) + q * x + (m % q);
}

template <int BLOCK, bool EAGER>
__global__ __launch_bounds__(BLOCK) void k_solve(SubArgs a, int begin, int end, int useBias, float h, int skipSoft) {
    const int k = homed_chunk(blockIdx.x, begin / BLOCK, 256 / BLOCK);
    int s = k * BLOCK + threadIdx.x;
    if (s >= begin && s < end) solve_manifold<EAGER>(a, s, useBias != 0, h, skipSoft != 0);
}

// ------------------------------------------------------------------ one launch per pass
// k_solve_pass runs every colour of one contact pass in one launch, a grid barrier between
// colours. Within it, the only data that passes from one colour to the next across CUs (and
// XCDs, whose L2 caches are not coherent with each other) is the bodies' solver velocities:
// they are read and written with relaxed agent-scope atomic accesses, which go to the
// device's coherence point, so the barrier needs no cache writeback. Each thread's
// accesses are complete (s_waitcnt 0) before its workgroup arrives; one lane per workgroup
// adds the arrival and polls the count with agent-scope loads. Every other input of a
// manifold is fixed for the pass and read with plain loads: a thread loads the next colour's
// manifold before it waits at the barrier. Lambdas are read and written by the same thread
// within the launch.
constexpr int PASS_THREADS = 256;
constexpr int PASS_SPIN_LIMIT = 1 << 22;   // ~seconds; never expected with co-resident blocks

__device__ __forceinline__ f3 ld_agent3(const float4* p) {
    uint32_t* q = reinterpret_cast<uint32_t*>(const_cast<float4*>(p));
    const uint32_t x = __hip_atomic_load(q + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t y = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t z = __hip_atomic_load(q + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return mk3(__uint_as_float(x), __uint_as_float(y), __uint_as_float(z));
}
__device__ __forceinline__ void st_agent3(float4* p, f3 v) {
    uint32_t* q = reinterpret_cast<uint32_t*>(p);
    __hip_atomic_store(q + 0, __float_as_uint(v.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, __float_as_uint(v.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 2, __float_as_uint(v.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the solve of one manifold whose fixed inputs are loaded, on velocities at the coherence point
__device__ __forceinline__ void solve_coherent(const SubArgs& a, int s, const MInputs& in, bool useBias, float h,
                                              bool skipSoft) {
    if (skipSoft && in.soft) return;
    f3 v0 = s3(0.f), w0 = s3(0.f), v1 = s3(0.f), w1 = s3(0.f);
    if (in.b0 >= 0) { v0 = ld_agent3(&a.bvel[bv(in.b0, BV_V)]); w0 = ld_agent3(&a.bvel[bv(in.b0, BV_W)]); }
    if (in.b1 >= 0) { v1 = ld_agent3(&a.bvel[bv(in.b1, BV_V)]); w1 = ld_agent3(&a.bvel[bv(in.b1, BV_W)]); }
    solve_core(a, s, in, useBias, h, v0, w0, v1, w1);
    if (in.b0 >= 0) { st_agent3(&a.bvel[bv(in.b0, BV_V)], v0); st_agent3(&a.bvel[bv(in.b0, BV_W)], w0); }
    if (in.b1 >= 0) { st_agent3(&a.bvel[bv(in.b1, BV_V)], v1); st_agent3(&a.bvel[bv(in.b1, BV_W)], w1); }
}

// grid barrier number `target` / gridDim.x of the substep: every workgroup has arrived
// `target` times in all. A wait that runs out (or that finds another one ran out) sets the
// error flag and lets the grid go on, so every wave still reaches the end of the kernel.
__device__ __forceinline__ void pass_barrier(int32_t* count, int target, int32_t* err) {
    __builtin_amdgcn_s_waitcnt(0);   // this thread's velocity stores are at the coherence point
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(count, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int spins = 0;
        while (__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > PASS_SPIN_LIMIT || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(PASS_THREADS) void k_solve_pass(SubArgs a, int nColors, int useBias, float h,
                                                            int skipSoft, int barrierBase) {
    const int tid = blockIdx.x * PASS_THREADS + threadIdx.x;
    const int nth = gridDim.x * PASS_THREADS;
    const bool bias = useBias != 0, skip = skipSoft != 0;
    int s = a.colorStart[0] + tid, e = a.colorStart[1];
    MInputs in;
    bool have = s < e;
    if (have) load_inputs<true>(a, s, in);
    for (int c = 0; c < nColors; ++c) {
        if (have) solve_coherent(a, s, in, bias, h, skip);
        for (int s2 = s + nth; s2 < e; s2 += nth) {
            MInputs in2;
            load_inputs<true>(a, s2, in2);
            solve_coherent(a, s2, in2, bias, h, skip);
        }
        if (c + 1 == nColors) break;
        s = e + tid;                 // the next colour starts where this one ends
        e = a.colorStart[c + 2];
        have = s < e;
        if (have) load_inputs<true>(a, s, in);
        pass_barrier(a.barrier, barrierBase + (c + 1) * (int)gridDim.x, a.flowErr);
    }
}

// ------------------------------------------------------------------ one launch per pass, on one XCD
// k_solve_xpass runs every colour of one contact pass in one launch, like k_solve_pass, but
// only on the CUs of one XCD, so the solver velocities handed from one colour to the next
// stay in that XCD's L2 instead of crossing to the device's coherence point:
// - every workgroup registers the XCD it runs on (hardware register XCC_ID) with one
//   agent-scope atomic add on a 64-bit word holding one byte per XCD; the byte's old value
//   is the workgroup's rank on its XCD. When all gridDim.x have registered, every workgroup
//   reads the word and the XCD with the most workgroups (the lowest on a tie) is chosen;
//   the workgroups of the other XCDs leave;
// - the participants solve the colours in order, a barrier between colours: velocities are
//   loaded with agent-scope loads, which skip the CU's L1 and are served by the XCD's L2,
//   where the other participants' plain stores land; stores are plain and complete
//   (s_waitcnt 0) before a workgroup arrives at the barrier (pass_barrier's count);
// - every wait is bounded and reports through flowErr.
// ctl (per launch, cleared by k_prepare): [0..1] the registration word, [2] the barrier count.
constexpr int XPASS_THREADS = 256;
constexpr int XPASS_CTL = 4;        // int32 words of one launch's control block
constexpr int XPASS_MAX_GRID = 248; // at most 31 workgroups per XCD: a byte never overflows

__device__ __forceinline__ f3 ld_l2_3(const float4* p) { return ld_agent3(p); }

__device__ __forceinline__ void solve_xcd(const SubArgs& a, int s, const MInputs& in, bool useBias, float h,
                                          bool skipSoft) {
    if (skipSoft && in.soft) return;
    f3 v0 = s3(0.f), w0 = s3(0.f), v1 = s3(0.f), w1 = s3(0.f);
    if (in.b0 >= 0) { v0 = ld_l2_3(&a.bvel[bv(in.b0, BV_V)]); w0 = ld_l2_3(&a.bvel[bv(in.b0, BV_W)]); }
    if (in.b1 >= 0) { v1 = ld_l2_3(&a.bvel[bv(in.b1, BV_V)]); w1 = ld_l2_3(&a.bvel[bv(in.b1, BV_W)]); }
    solve_core(a, s, in, useBias, h, v0, w0, v1, w1);
    if (in.b0 >= 0) { a.bvel[bv(in.b0, BV_V)] = f4(v0, 0.f); a.bvel[bv(in.b0, BV_W)] = f4(w0, 0.f); }
    if (in.b1 >= 0) { a.bvel[bv(in.b1, BV_V)] = f4(v1, 0.f); a.bvel[bv(in.b1, BV_W)] = f4(w1, 0.f); }
}

__global__ __launch_bounds__(XPASS_THREADS) void k_solve_xpass(SubArgs a, int nColors, int useBias, float h,
                                                              int skipSoft, int32_t* ctl) {
    __shared__ int sRank, sP;
    unsigned long long* reg = reinterpret_cast<unsigned long long*>(ctl);
    if (threadIdx.x == 0) {
        const int x = (int)(__builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 7);   // hwreg(HW_REG_XCC_ID)
        const unsigned long long old =
            __hip_atomic_fetch_add(reg, 1ull << (8 * x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int spins = 0;
        unsigned long long v = 0;
        for (;;) {
            v = __hip_atomic_load(reg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int total = 0;
            for (int k = 0; k < 8; ++k) total += (int)((v >> (8 * k)) & 255);
            if (total >= (int)gridDim.x) break;
            __builtin_amdgcn_s_sleep(1);
            if (++spins > PASS_SPIN_LIMIT || __hip_atomic_load(a.flowErr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_store(a.flowErr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                v = 0;
                break;
            }
        }
        int best = 0, bestN = 0;
        for (int k = 0; k < 8; ++k) {
            const int c = (int)((v >> (8 * k)) & 255);
            if (c > bestN) { best = k; bestN = c; }
        }
        sRank = (int)((old >> (8 * x)) & 255);
        sP = (x == best) ? bestN : 0;
    }
    __syncthreads();
    const int P = sP;
    if (P == 0) return;            // another XCD's workgroup (or registration failed)
    const int tid = sRank * XPASS_THREADS + threadIdx.x;
    const int nth = P * XPASS_THREADS;
    const bool bias = useBias != 0, skip = skipSoft != 0;
    int s = a.colorStart[0] + tid, e = a.colorStart[1];
    MInputs in;
    bool have = s < e;
    if (have) load_inputs<true>(a, s, in);
    for (int c = 0; c < nColors; ++c) {
        if (have) solve_xcd(a, s, in, bias, h, skip);
        for (int s2 = s + nth; s2 < e; s2 += nth) {
            MInputs in2;
            load_inputs<true>(a, s2, in2);
            solve_xcd(a, s2, in2, bias, h, skip);
        }
        if (c + 1 == nColors) break;
        s = e + tid;
        e = a.colorStart[c + 2];
        have = s < e;
        if (have) load_inputs<true>(a, s, in);
        pass_barrier(ctl + 2, (c + 1) * P, a.flowErr);
    }
}

// Consecutive small colours c0..c1-1 (solve slots [S, E)) in one launch of one workgroup:
// the colours are solved in order with a workgroup barrier between them, so the solve order
// is the colour order, exactly as with one launch per colour. The solver velocities of the
// run's bodies stay in LDS for the whole run: one gather before the first colour, one
// scatter after the last, and between colours only the manifolds' fixed inputs come from
// global memory. k_batch_map gives each run its bodies (batchList, from 2S on) and each
// manifold side its LDS slot (batchLoc, 2 per solve slot; -1 for a static side), once per
// step: a run's slots never change within the step.
constexpr int BATCH_THREADS = 512;      // k_solve's register footprint allows 2 waves per SIMD
constexpr int BATCH_MAP_THREADS = 1024;
constexpr int BATCH_HASH = 8192;        // LDS hash of body -> slot: at least twice a run's bodies
constexpr int BATCH_MAX_GROUPS = 16;
constexpr int BATCH_MAX_MANIFOLDS = 1264;   // 2 x 1264 bodies x 64 B of records in LDS
constexpr int BATCH_MAX_COLOURS = 128;

struct BatchGroups {
    int n;
    int S[BATCH_MAX_GROUPS], E[BATCH_MAX_GROUPS];
};

__device__ __forceinline__ uint32_t batch_hash(int b) { return ((uint32_t)b * 2654435761u) >> (32 - 13); }

// One workgroup, every run in turn: the run's bodies get slots 0..n-1 in first-claim order
// (the order does not reach the arithmetic) through an LDS hash table with linear probing.
__global__ __launch_bounds__(BATCH_MAP_THREADS) void k_batch_map(BatchGroups g, const int32_t* __restrict__ sB0,
                                                                 const int32_t* __restrict__ sB1,
                                                                 int32_t* __restrict__ batchLoc,
                                                                 int32_t* __restrict__ batchList,
                                                                 int32_t* __restrict__ batchN) {
    static_assert(BATCH_HASH == 1 << 13, "batch_hash gives 13 bits");
    __shared__ int32_t hk[BATCH_HASH], hv[BATCH_HASH];
    __shared__ int32_t count;
    for (int q = 0; q < g.n; ++q) {
        const int S = g.S[q], E = g.E[q], sides = 2 * (E - S);
        for (int i = threadIdx.x; i < BATCH_HASH; i += BATCH_MAP_THREADS) hk[i] = -1;
        if (threadIdx.x == 0) count = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < sides; i += BATCH_MAP_THREADS) {
            const int sl = S + (i >> 1);
            const int b = (i & 1) ? sB1[sl] : sB0[sl];
            if (b < 0) continue;
            uint32_t h = batch_hash(b);
            for (int probe = 0; probe < BATCH_HASH; ++probe) {   // the table is never full
                const int prev = atomicCAS(&hk[h], -1, b);
                if (prev == -1) {
                    const int slot = atomicAdd(&count, 1);
                    hv[h] = slot;
                    batchList[2 * S + slot] = b;
                    break;
                }
                if (prev == b) break;
                h = (h + 1) & (BATCH_HASH - 1);
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < sides; i += BATCH_MAP_THREADS) {
            const int sl = S + (i >> 1);
            const int b = (i & 1) ? sB1[sl] : sB0[sl];
            int slot = -1;
            if (b >= 0) {
                uint32_t h = batch_hash(b);
                for (int probe = 0; probe < BATCH_HASH && hk[h] != b; ++probe) h = (h + 1) & (BATCH_HASH - 1);
                slot = hv[h];
            }
            batchLoc[2 * sl + (i & 1)] = slot;
        }
        if (threadIdx.x == 0) batchN[q] = count;
        __syncthreads();
    }
}

// A workgroup barrier that orders LDS only: the global loads a thread has in flight stay in
// flight across it (s_waitcnt lgkmcnt(0); s_barrier).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The run's bodies' records (solver v and w, component v and w with the inverse mass:
// BV_V..BV_CW) are gathered into LDS once per launch, so a manifold's solve needs only its
// own rows from global memory: slot 0's with the flags, slots 1-3 once the point count is
// known (one CU's load bandwidth bounds these launches, so unused slots are not read). Each
// colour's rows are loaded before the barrier that waits for the previous colour's
// velocities, so their latency overlaps the wait.
constexpr int BATCH_BODY_F4 = BV_CW + 1;
__global__ __launch_bounds__(BATCH_THREADS) void k_solve_batch(SubArgs a, int c0, int c1, int S,
                                                               const int32_t* __restrict__ batchLoc,
                                                               const int32_t* __restrict__ batchList,
                                                               const int32_t* __restrict__ batchN, int useBias, float h,
                                                               int skipSoft) {
    extern __shared__ float4 rec[];       // [BATCH_BODY_F4][n]: record field f of body slot i at f * n + i
    __shared__ int32_t cs[BATCH_MAX_COLOURS + 1];   // the run's colour starts
    const int n = *batchN;
    const int32_t* list = batchList + 2 * S;
    for (int i = threadIdx.x; i <= c1 - c0; i += BATCH_THREADS) cs[i] = a.colorStart[c0 + i];
    for (int i = threadIdx.x; i < n; i += BATCH_THREADS) {
        const int b = list[i];
#pragma unroll
        for (int f = 0; f < BATCH_BODY_F4; ++f) rec[f * n + i] = a.bvel[bv(b, f)];
    }
    __syncthreads();
    auto solve = [&](int s, MInputs& in, int2 l) {
        if (skipSoft && in.soft) return;
        f3 v0 = s3(0.f), w0 = s3(0.f), v1 = s3(0.f), w1 = s3(0.f);
        in.im0 = 0.f; in.im1 = 0.f;
        in.cv0 = s3(0.f); in.cw0 = s3(0.f); in.cv1 = s3(0.f); in.cw1 = s3(0.f);
        if (in.b0 >= 0) {
            v0 = xyz(rec[BV_V * n + l.x]); w0 = xyz(rec[BV_W * n + l.x]);
            const float4 c = rec[BV_CV * n + l.x];
            in.im0 = c.w; in.cv0 = xyz(c); in.cw0 = xyz(rec[BV_CW * n + l.x]);
        }
        if (in.b1 >= 0) {
            v1 = xyz(rec[BV_V * n + l.y]); w1 = xyz(rec[BV_W * n + l.y]);
            const float4 c = rec[BV_CV * n + l.y];
            in.im1 = c.w; in.cv1 = xyz(c); in.cw1 = xyz(rec[BV_CW * n + l.y]);
        }
        solve_core(a, s, in, useBias != 0, h, v0, w0, v1, w1);
        if (in.b0 >= 0) { rec[BV_V * n + l.x] = f4(v0, 0.f); rec[BV_W * n + l.x] = f4(w0, 0.f); }
        if (in.b1 >= 0) { rec[BV_V * n + l.y] = f4(v1, 0.f); rec[BV_W * n + l.y] = f4(w1, 0.f); }
    };
    for (int c = 0; c < c1 - c0; ++c) {
        const int e = cs[c + 1];
        int s = cs[c] + (int)threadIdx.x;
        MInputs in;
        int2 l = make_int2(-1, -1);
        const bool have = s < e;
        if (have) {
            load_own<false>(a, s, in);
            l = *reinterpret_cast<const int2*>(batchLoc + 2 * s);
        }
        lds_barrier();   // the previous colour's velocities are in LDS
        if (have) solve(s, in, l);
        // colours larger than the workgroup: a thread's manifolds of one colour share no body
        for (s += BATCH_THREADS; s < e; s += BATCH_THREADS) {
            load_own<false>(a, s, in);
            solve(s, in, *reinterpret_cast<const int2*>(batchLoc + 2 * s));
        }
    }
    lds_barrier();
    for (int i = threadIdx.x; i < n; i += BATCH_THREADS) {
        const int b = list[i];
        a.bvel[bv(b, BV_V)] = rec[BV_V * n + i];
        a.bvel[bv(b, BV_W)] = rec[BV_W * n + i];
    }
}

// ------------------------------------------------------------------ per-world passes
// Worlds never share a body, and a world's dynamic bodies are a contiguous index range, so
// one workgroup can run a whole contact pass of one world: the solver velocities of its
// bodies in LDS, its manifolds colour by colour with a workgroup barrier between colours.
// Every body sees its manifolds in colour order, as with one launch per colour, so the
// floats are the same. No launch per colour and no global body gathers inside the pass.
struct WorldPassArgs {
    int W, nColors;
    const int32_t* wDyn;                  // W + 1: first dynamic body of each world
    const int32_t *wcS, *wcE;             // per colour and world: its solve slots [S, E)
    const int32_t* list;                  // the worlds of this launch (one workgroup each)
};

// the solve slots of every (colour, world) run; slots are in colour order and, within a
// colour, in world order. Absent runs stay [0, 0).
__global__ void k_world_ranges(int M, int W, const int32_t* __restrict__ sFlags, const int32_t* __restrict__ sWorld,
                               int32_t* __restrict__ wcS, int32_t* __restrict__ wcE) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    const int c = (sFlags[s] >> 16) & 127, wd = sWorld[s];
    if (s == 0 || ((sFlags[s - 1] >> 16) & 127) != c || sWorld[s - 1] != wd) wcS[c * W + wd] = s;
    if (s == M - 1 || ((sFlags[s + 1] >> 16) & 127) != c || sWorld[s + 1] != wd) wcE[c * W + wd] = s + 1;
}

// One wave per world: a world's colour holds a few dozen manifolds, and the solve's
// register footprint allows 8 waves per CU, so one-wave workgroups keep 8 worlds resident
// per CU (when their LDS fits). Rows are loaded eagerly: these colours are latency-bound.
constexpr int WORLD_THREADS = 64;
__global__ __launch_bounds__(WORLD_THREADS) void k_world_pass(SubArgs a, WorldPassArgs p, int useBias, float h,
                                                               int skipSoft) {
    extern __shared__ float4 wv[];        // [n] solver v, then [n] solver w of the world's bodies
    const int wd = p.list[blockIdx.x];
    const int d0 = p.wDyn[wd], n = p.wDyn[wd + 1] - d0;
    if (n == 0) return;
    for (int i = threadIdx.x; i < n; i += WORLD_THREADS) {
        wv[i] = a.bvel[bv(d0 + i, BV_V)];
        wv[n + i] = a.bvel[bv(d0 + i, BV_W)];
    }
    __syncthreads();
    for (int c = 0; c < p.nColors; ++c) {
        const int S = p.wcS[c * p.W + wd], E = p.wcE[c * p.W + wd];
        for (int s = S + (int)threadIdx.x; s < E; s += WORLD_THREADS) {
            MInputs in;
            load_inputs<true>(a, s, in);
            if (skipSoft && in.soft) continue;
            const int l0 = in.b0 - d0, l1 = in.b1 - d0;
            f3 v0 = s3(0.f), w0 = s3(0.f), v1 = s3(0.f), w1 = s3(0.f);
            if (in.b0 >= 0) { v0 = xyz(wv[l0]); w0 = xyz(wv[n + l0]); }
            if (in.b1 >= 0) { v1 = xyz(wv[l1]); w1 = xyz(wv[n + l1]); }
            solve_core(a, s, in, useBias != 0, h, v0, w0, v1, w1);
            if (in.b0 >= 0) { wv[l0] = f4(v0, 0.f); wv[n + l0] = f4(w0, 0.f); }
            if (in.b1 >= 0) { wv[l1] = f4(v1, 0.f); wv[n + l1] = f4(w1, 0.f); }
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += WORLD_THREADS) {
        a.bvel[bv(d0 + i, BV_V)] = wv[i];
        a.bvel[bv(d0 + i, BV_W)] = wv[n + i];
    }
}

// colours of body b, and its colours below colour c
__device__ __forceinline__ int colours_of(const uint64_t* used, int b) {
    return __popcll(used[2 * b]) + __popcll(used[2 * b + 1]);
}
__device__ __forceinline__ int colours_below(const uint64_t* used, int b, int c) {
    uint64_t lo = used[2 * b], hi = used[2 * b + 1];
    if (c < 64) return __popcll(lo & ((1ull << c) - 1ull));
    return __popcll(lo) + __popcll(hi & ((1ull << (c - 64)) - 1ull));
}

// ------------------------------------------------------------------ dependency-driven pass
// A pass (one of the substep's biased iterations, or the relaxation) over every colour in
// ONE launch. A manifold of colour c needs the velocities its bodies hold after their
// manifolds of lower colours in this pass; with r = (colours of body b below c) such
// manifolds before it, the r-th of them publishes body b's velocities under version
// (epoch, r). The first manifold of a body (r = 0) reads the plain solver velocities, which
// are complete when the launch starts (previous pass, joints); its last manifold of the
// pass writes them, their only writer in the launch.
//
// Published velocities: 6 words of {value, version} per body, each word readable on its
// own, so a reader accepts a record only when all six carry the version it waits for and
// needs no ordering between the words' stores. Records are agent-scope atomic accesses (a
// load that hits the CU's own cache may never see another CU's store within a launch).
// Queues: one device-wide queue, or one per XCD (world w on XCD w % XCDs; waves take
// manifolds only from their own XCD's queue, read from the hardware register XCC_ID), so
// that a world's hand-offs stay inside one XCD: a few hundred nanoseconds each, against
// microseconds between XCDs.
//
// The solve order is the colour order's, so the arithmetic is exactly that of one launch
// per colour. Waves take 64 manifolds at a time, in solve order, from the pass's ticket: a
// manifold waits only on manifolds of lower solve positions held by waves already running,
// so the waits resolve in order. Every wait is bounded; one that runs out sets flowErr and
// gives up (the host reports the error) rather than hang.
constexpr int FLOW_SPIN_LIMIT = 1 << 18;   // ~0.5 s of polling: a wait this long is a bug

struct FlowArgs {
    int pass;
    int useBias, skipSoft;
    float h;
    uint32_t epoch;                        // unique per launch (24 bits used)
    uint64_t* pub;                         // 6 words per body
    int xcdScope;                          // 1: a queue per XCD, 0: one queue
    int nQueues;
    const int32_t* qStart;                 // nQueues + 1 offsets into qList
    const int32_t* qList;                  // solve positions by queue (nullptr: identity)
};

__device__ __forceinline__ uint64_t pub_word(float v, uint32_t ver) {
    return (uint64_t)__float_as_uint(v) | ((uint64_t)ver << 32);
}

// body b's published velocities under version ver, if all six words carry it
__device__ __forceinline__ bool read_pub(const FlowArgs& f, int b, uint32_t ver, f3& v, f3& w) {
    uint64_t q[6];
    uint64_t* p = f.pub + 6 * (size_t)b;
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = __hip_atomic_load(p + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) ok = ok && (uint32_t)(q[k] >> 32) == ver;
    v = mk3(__uint_as_float((uint32_t)q[0]), __uint_as_float((uint32_t)q[1]), __uint_as_float((uint32_t)q[2]));
    w = mk3(__uint_as_float((uint32_t)q[3]), __uint_as_float((uint32_t)q[4]), __uint_as_float((uint32_t)q[5]));
    return ok;
}

// Hands body b's velocities on after one of its manifolds: to the next manifold of the
// pass (published under version ver), or, after its last manifold of the pass, to the
// kernels after the pass (the plain solver velocities, written back at the launch's end).
__device__ __forceinline__ void hand_on(const SubArgs& a, const FlowArgs& f, int b, bool last, uint32_t ver, f3 v,
                                        f3 w) {
    if (last) {
        a.bvel[bv(b, BV_V)] = f4(v, 0.f);
        a.bvel[bv(b, BV_W)] = f4(w, 0.f);
        return;
    }
    uint64_t* p = f.pub + 6 * (size_t)b;
    __hip_atomic_store(p + 0, pub_word(v.x, ver), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 1, pub_word(v.y, ver), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 2, pub_word(v.z, ver), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 3, pub_word(w.x, ver), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 4, pub_word(w.y, ver), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 5, pub_word(w.z, ver), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(256) void k_solve_flow(SubArgs a, FlowArgs f) {
    const int lane = threadIdx.x & (WAVE - 1);
    int queue = 0;
    if (f.xcdScope) {
        queue = (int)(__builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 15);   // hwreg(HW_REG_XCC_ID, 0, 4)
        if (queue >= f.nQueues) return;
    }
    const int qb = f.qStart[queue], qn = f.qStart[queue + 1] - qb;
    const uint32_t vbase = (f.epoch & 0xffffffu) << 8;
    const bool useBias = f.useBias != 0, skipSoft = f.skipSoft != 0;
    int done = 0;
    for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(a.tickets + f.pass * FLOW_QUEUES + queue, WAVE);
        base = __shfl(base, 0);
        if (base >= qn) break;
        const int idx = base + lane;
        bool pending = idx < qn;
        const int s = pending ? (f.qList ? f.qList[qb + idx] : qb + idx) : 0;
        MInputs in;
        int r0 = 0, r1 = 0;
        bool last0 = false, last1 = false;
        f3 v0 = s3(0.f), w0 = s3(0.f), v1 = s3(0.f), w1 = s3(0.f);
        bool have0 = true, have1 = true;
        if (pending) {
            // everything but the bodies' solver velocities is fixed for the pass: loaded
            // before the wait, off the critical path
            load_inputs(a, s, in);
            const int c = a.sFlags[s] >> 16;
            if (in.b0 >= 0) {
                r0 = colours_below(a.used, in.b0, c);
                last0 = r0 == colours_of(a.used, in.b0) - 1;
                have0 = r0 == 0;
                if (have0) { v0 = xyz(a.bvel[bv(in.b0, BV_V)]); w0 = xyz(a.bvel[bv(in.b0, BV_W)]); }
            }
            if (in.b1 >= 0) {
                r1 = colours_below(a.used, in.b1, c);
                last1 = r1 == colours_of(a.used, in.b1) - 1;
                have1 = r1 == 0;
                if (have1) { v1 = xyz(a.bvel[bv(in.b1, BV_V)]); w1 = xyz(a.bvel[bv(in.b1, BV_W)]); }
            }
        }
        int spins = 0;
        while (__ballot(pending)) {
            if (pending) {
                if (!have0) have0 = read_pub(f, in.b0, vbase | (uint32_t)r0, v0, w0);
                if (!have1) have1 = read_pub(f, in.b1, vbase | (uint32_t)r1, v1, w1);
                if (have0 && have1) {
                    // a soft manifold in the relaxation passes its bodies' velocities on
                    if (!(skipSoft && in.soft)) solve_core(a, s, in, useBias, f.h, v0, w0, v1, w1);
                    if (in.b0 >= 0) hand_on(a, f, in.b0, last0, vbase | (uint32_t)(r0 + 1), v0, w0);
                    if (in.b1 >= 0) hand_on(a, f, in.b1, last1, vbase | (uint32_t)(r1 + 1), v1, w1);
                    pending = false;
                    ++done;
                } else if (++spins > FLOW_SPIN_LIMIT) {
                    atomicOr(a.flowErr, 1);
                    pending = false;
                }
            }
            if (__ballot(pending)) __builtin_amdgcn_s_sleep(1);
        }
    }
    // manifolds finished per queue: the host checks that every queue was served
    for (int o = WAVE / 2; o > 0; o >>= 1) done += __shfl_xor(done, o);
    if (lane == 0 && done) atomicAdd(a.qDone + queue, done);
}

// The queues of the dependency-driven pass. One world, or one device-wide queue: queue 0
// is the whole solve order. Several worlds on XCD queues: world w goes to queue w % nq,
// keeping solve order inside each queue (a stable sort by queue).
__global__ void k_flow_identity(int M, int nq, int32_t* qStart) {
    int i = threadIdx.x;
    if (i <= nq) qStart[i] = i == 0 ? 0 : M;
}
__global__ void k_flow_keys(int M, const int32_t* __restrict__ sCanon, const uint64_t* __restrict__ mfKey,
                            int worldShift, int nq, uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    keys[s] = (mfKey[sCanon[s]] >> worldShift) % (uint64_t)nq;
    vals[s] = s;
}
__global__ void k_flow_starts(int M, int nq, const uint64_t* __restrict__ keys, int32_t* __restrict__ qStart) {
    int q = threadIdx.x;
    if (q > nq) return;
    int lo = 0, hi = M;   // first position with key >= q
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (keys[mid] < (uint64_t)q) lo = mid + 1;
        else hi = mid;
    }
    qStart[q] = lo;
}

__global__ __launch_bounds__(256) void k_integrate_positions(SubArgs a) {
    int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < a.D) integrate_position(a, d);
}
__global__ __launch_bounds__(256) void k_writeback(SubArgs a) {
    int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < a.D) write_back(a, d);
}

// ------------------------------------------------------------------ persistent kernel
// Every substep of one step in a single cooperative launch; grid-wide barriers separate the
// phases and the colours. Blocks are co-resident (cooperative launch), every block reaches
// every barrier the same number of times, and the kernel ends after the last substep.
__global__ __launch_bounds__(256) void k_step_persistent(SubArgs a) {
    cg::grid_group grid = cg::this_grid();
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nth = gridDim.x * blockDim.x;
    for (int sub = 0; sub < a.substeps; ++sub) {
        for (int i = tid; i < a.M + a.D; i += nth) {
            if (i < a.M) prepare_manifold(a, i);
            else integrate_velocity(a, i - a.M);
        }
        grid.sync();
        for (int it = 0; it < a.iterations; ++it) {
            for (int c = 0; c < a.nColors; ++c) {
                int b = a.colorStart[c], e = a.colorStart[c + 1];
                for (int s = b + tid; s < e; s += nth) solve_manifold(a, s, true, a.h, false);
                grid.sync();
            }
        }
        for (int d = tid; d < a.D; d += nth) integrate_position(a, d);
        grid.sync();
        for (int c = 0; c < a.nColors; ++c) {
            int b = a.colorStart[c], e = a.colorStart[c + 1];
            for (int s = b + tid; s < e; s += nth) solve_manifold(a, s, false, 0.f, true);
            grid.sync();
        }
        for (int d = tid; d < a.D; d += nth) write_back(a, d);
        grid.sync();
    }
}

// ------------------------------------------------------------------ host orchestration
int sv_setup_manifolds(World& w, int nPairs) {
    hipStream_t s = w.stream;
    if (nPairs == 0) return 0;
    // manifold index of every pair with points: exclusive scan of the hit flags (the pair
    // sort's value buffer is free at this point and holds the flags)
    hipLaunchKernelGGL(k_flag_manifolds, dim3(div_up(nPairs, 256)), dim3(256), 0, s, nPairs, w.npNum.p,
                       w.meshCount.p, w.pvals.p);
    exclusive_scan_i32(s, w.pvals.p, w.npIndex.p, nPairs, w.counters.p + 1, w.scanT);
    int M = 0, epaErrors = 0, flowErr = 0;
    int meshInfo[3] = {0, 0, 0};   // mesh pairs, mesh manifold slots, mesh capacity errors
    // the counters (manifold total, EPA and mesh counts) and the previous step's
    // dependency-driven passes: a wait that ran out, and every queue served by waves of its
    // XCD; to pinned host memory
    int qDone[FLOW_QUEUES] = {}, qStart[FLOW_QUEUES + 1] = {};
    const int lastQueues = w.lastQueues, lastPasses = w.lastFlowPasses;
    w.readHost.alloc(READ_HOST_INTS);
    int32_t* rb = w.readHost.p;
    PGX_HIP(hipMemcpyAsync(rb, w.counters.p, 16 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PGX_HIP(hipMemcpyAsync(rb + 16, w.flowErr.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (lastQueues) {
        PGX_HIP(hipMemcpyAsync(qDone, w.qDone.p, sizeof(qDone), hipMemcpyDeviceToHost, s));
        PGX_HIP(hipMemcpyAsync(qStart, w.qStart.p, sizeof(int) * (lastQueues + 1), hipMemcpyDeviceToHost, s));
    }
    hipLaunchKernelGGL(k_compact, dim3(div_up(nPairs, 256)), dim3(256), 0, s, nPairs, w.npNum.p, w.npIndex.p,
                       w.meshFirst.p, w.meshCount.p, w.npTri.p, w.pairCap, w.pkeys.p, w.localBits, w.worldColStart.p, w.colOfLocal.p, w.cbody.p, w.cflags.p,
                       w.mfPair.p, w.mfKey.p, w.mfB0.p, w.mfB1.p, w.mfSrc.p, w.mfTri.p);
    PGX_LAUNCH_CHECK();
    PGX_HIP(hipStreamSynchronize(s));
    M = rb[1];
    flowErr = rb[16];
    if (w.nConvex) epaErrors = rb[9];
    if (w.nTriMesh) std::copy(rb + 10, rb + 13, meshInfo);
    if (flowErr)
        throw PgxError{PGX_ERR_HIP, "contact solve: a wait of the previous step ran out (dependency-driven pass or "
                                    "grid barrier of the one-launch pass)"};
    for (int q = 0; q < lastQueues; ++q)
        if (qDone[q] != lastPasses * (qStart[q + 1] - qStart[q]))
            throw PgxError{PGX_ERR_HIP, "contact solve: queue " + std::to_string(q) + " of the previous step finished " +
                                            std::to_string(qDone[q]) + " of " +
                                            std::to_string(lastPasses * (qStart[q + 1] - qStart[q])) +
                                            " manifold solves (no wave ran on its XCD)"};
    if (epaErrors)
        throw PgxError{PGX_ERR_CAPACITY, "EPA polytope capacity exceeded in " + std::to_string(epaErrors) + " pairs"};
    if (meshInfo[2])
        throw PgxError{PGX_ERR_CAPACITY, "triangle-mesh narrowphase capacity exceeded in " +
                                             std::to_string(meshInfo[2]) + " pairs (candidates, stack or slots)"};
    w.lastMeshSlots = meshInfo[1];
    return M;
}

int sv_color(World& w, int M) {
    hipStream_t s = w.stream;
    const int prevRounds = w.lastRounds;
    w.lastRounds = 0;
    if (M == 0) {
        w.colorStartHost.assign(1, 0);
        return 0;
    }
    int worldShift = 2 * w.localBits;
    int nb = div_up(M, 256);
    constexpr int MAX_ROUNDS = 1024;
    static_assert(133 + MAX_ROUNDS + 1 <= READ_HOST_INTS, "the colouring's read-back fits the pinned buffer");
    w.colorUncol.alloc(MAX_ROUNDS + 1);
    w.colorPack.alloc(133 + MAX_ROUNDS + 1);
    w.readHost.alloc(READ_HOST_INTS);
    w.claim2.alloc(std::max(w.D, 1));
    w.mfPrev.alloc(w.manifoldCap);
    w.colorFlag.alloc(1);
    hipLaunchKernelGGL(k_color_begin, dim3(div_up(std::max({M, 2 * w.D, MAX_ROUNDS + 1}), 256)), dim3(256), 0, s, M,
                       w.D, MAX_ROUNDS + 1, w.mfKey.p, worldShift, w.worldMStart.p, w.mfColor.p, w.claim.p,
                       w.claim2.p, w.usedMask.p, w.colorUncol.p, w.counters.p + 5, w.colorFlag.p);
    unsigned long long* claims[2] = {(unsigned long long*)w.claim.p, (unsigned long long*)w.claim2.p};
    // colours of the manifolds that persist from the previous step; only new ones are
    // coloured in the rounds below
    const int prevCache = w.cacheCur;
    hipLaunchKernelGGL(k_color_inherit, dim3(nb), dim3(256), 0, s, M, w.mfKey.p, w.mfTri.p, w.cacheKey[prevCache].p,
                       w.cacheTri[prevCache].p, w.cacheNum[prevCache].p, w.cacheCount[prevCache], w.mfB0.p, w.mfB1.p,
                       w.mfColor.p,
                       w.mfPrev.p, (unsigned long long*)w.usedMask.p, w.colorFlag.p);
    hipLaunchKernelGGL(k_color_clash, dim3(div_up(std::max(M, 2 * w.D), 256)), dim3(256), 0, s, M, w.D, w.colorFlag.p,
                       w.mfColor.p, w.usedMask.p);
    // round-1 claims
    hipLaunchKernelGGL(k_color_claim, dim3(nb), dim3(256), 0, s, M, 1, w.mfKey.p, worldShift, w.worldMStart.p,
                       w.mfColor.p, w.mfB0.p, w.mfB1.p, claims[1]);
    // Rounds are queued in batches sized from the last step's count (a round after the last
    // one needed exits at once); the sort into colour order and the colour starts follow,
    // and one read-back tells whether the batch finished, with the counts and starts. Only
    // a batch that ran short is followed by another (and the sort is redone).
    int round = 0;
    int batch = std::min(MAX_ROUNDS, std::max(w.tune.colorRounds, prevRounds + 4));
    std::vector<int32_t> starts(130, 0), uc;
    uint64_t lastKey = 0;
    while (true) {
        for (int r = 0; r < batch && round < MAX_ROUNDS; ++r) {
            ++round;
            hipLaunchKernelGGL(k_color_round, dim3(nb), dim3(256), 0, s, M, round, w.mfKey.p, worldShift,
                               w.worldMStart.p, w.mfColor.p, w.mfB0.p, w.mfB1.p,
                               (const unsigned long long*)claims[round & 1], claims[(round + 1) & 1], w.usedMask.p,
                               w.colorUncol.p, w.counters.p);
        }
        // stable sort by colour and point count -> solve order
        hipLaunchKernelGGL(k_color_keys, dim3(nb), dim3(256), 0, s, M, w.mfColor.p, w.npNum.p, w.mfSrc.p,
                           w.execMode == EXEC_WORLDS ? 0 : 1, w.colorKeys.p, w.colorPerm.p);
        radix_sort_u64(s, w.colorKeys.p, w.colorPerm.p, M, 7 + COLOR_KEY_SHIFT, w.sortT);
        hipLaunchKernelGGL(k_color_start, dim3(1), dim3(256), 0, s, M, 130, w.colorKeys.p, w.colorCount.p, round,
                           w.colorUncol.p, w.counters.p + 5, w.colorPack.p);
        PGX_LAUNCH_CHECK();
        // one copy to pinned host memory brings the starts, the overflow flag, the last key
        // and the per-round counts
        const size_t packed = 133 + (size_t)round + 1;
        PGX_HIP(hipMemcpyAsync(w.readHost.p, w.colorPack.p, sizeof(int32_t) * packed, hipMemcpyDeviceToHost, s));
        PGX_HIP(hipStreamSynchronize(s));
        const int32_t* rb = w.readHost.p;
        const int ovf = rb[130];
        lastKey = (uint64_t)(uint32_t)rb[131] | ((uint64_t)(uint32_t)rb[132] << 32);
        std::copy(rb, rb + 130, starts.begin());
        uc.assign(rb + 133, rb + 133 + round + 1);
        if (ovf) throw PgxError{PGX_ERR_CAPACITY, "a body touches more than 127 contact colours"};
        if (uc[round] == 0) break;
        if (round >= MAX_ROUNDS) throw PgxError{PGX_ERR_INVALID, "graph colouring did not converge"};
        batch = 16;
    }
    // rounds actually used: the first whose count of uncoloured manifolds is zero
    int used = round;
    for (int r = 1; r <= round; ++r)
        if (uc[r] == 0) { used = r; break; }
    int nColors = (int)(lastKey >> COLOR_KEY_SHIFT) + 1;
    w.colorStartHost.assign(starts.begin(), starts.begin() + nColors + 1);
    w.lastRounds = used;
    // contact setup, written in solve order
    int cur = w.cacheCur ^ 1, prv = w.cacheCur;
    hipLaunchKernelGGL(k_color_rank, dim3(nb), dim3(256), 0, s, M, w.colorPerm.p, w.colorRank.p);
    hipLaunchKernelGGL(k_setup, dim3(nb), dim3(256), 0, s, M, w.manifoldCap, w.colorRank.p, w.mfSrc.p, w.mfTri.p,
                       w.mfKey.p,
                       w.localBits, w.worldColStart.p, w.colOfLocal.p, w.cbody.p, w.cflags.p, w.cmat.p, w.npNum.p,
                       w.npNormal.p, w.npPts.p, w.mfColor.p, w.mfB0.p, w.mfB1.p, w.bpos.p, w.bquat.p, w.dvel.p,
                       w.dang.p, w.dcom.p, w.D, w.mfPrev.p, w.cacheNum[prv].p, w.cacheLp[prv].p, w.manifoldCap,
                       w.cacheKey[cur].p, w.cacheTri[cur].p,
                       w.cacheNum[cur].p, w.cacheLp[cur].p,
                       w.sB0.p, w.sB1.p, w.sA0.p, w.sA1.p, w.sFlags.p, w.sCanon.p, w.sN.p, w.sSoft.p, w.sR0.p,
                       w.sR1.p, w.sWorld.p);
    PGX_LAUNCH_CHECK();
    return nColors;
}

static SubArgs make_args(World& w, int M, float dt) {
    SubArgs a;
    a.M = M; a.Mcap = w.manifoldCap; a.D = w.D;
    a.nColors = (int)w.colorStartHost.size() - 1;
    a.substeps = w.substeps; a.iterations = w.iterations;
    a.h = dt / w.substeps; a.g = w.g;
    a.sB0 = w.sB0.p; a.sB1 = w.sB1.p; a.sA0 = w.sA0.p; a.sA1 = w.sA1.p; a.sFlags = w.sFlags.p;
    a.colorStart = w.colorCount.p;
    a.sN = w.sN.p; a.sR0 = w.sR0.p; a.sR1 = w.sR1.p; a.sSoft = w.sSoft.p; a.sRow = w.sRow.p; a.sLam = w.sLam.p;
    a.bpos = w.bpos.p; a.bquat = w.bquat.p; a.dvel = w.dvel.p; a.dang = w.dang.p;
    a.bvel = w.bvel.p; a.dcom = w.dcom.p; a.dinvI = w.dinvI.p;
    a.pvel = w.pvel.p; a.pang = w.pang.p;
    a.used = w.usedMask.p; a.tickets = w.tickets.p; a.flowErr = w.flowErr.p; a.qDone = w.qDone.p;
    a.barrier = w.passBarrier.p;
    if (w.execMode == EXEC_XPASS) w.xpassCtl.alloc(XPASS_CTL * (size_t)(w.iterations + 1));
    a.xctl = w.xpassCtl.p;
    a.xctlWords = w.execMode == EXEC_XPASS ? XPASS_CTL * (w.iterations + 1) : 0;
    return a;
}

// Largest co-resident grid for the persistent kernel (cached per world).
static int persistent_blocks(World& w) {
    if (w.persistentBlocks > 0) return w.persistentBlocks;
    int perCU = 0, cus = 0;
    PGX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, (const void*)k_step_persistent, 256, 0));
    PGX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, w.device));
    int target = w.persistentBlocksWanted > 0 ? w.persistentBlocksWanted : cus;   // one block per CU by default
    w.persistentBlocks = std::max(1, std::min(target, perCU * cus));
    return w.persistentBlocks;
}

// Grid of the one-launch-per-pass kernel: enough threads for the largest colour in one
// round, at most one workgroup per CU, and never more than can be co-resident.
static int pass_blocks(World& w, int maxColour) {
    if (w.passBlocksMax == 0) {
        int perCU = 0, cus = 0;
        PGX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, (const void*)k_solve_pass, PASS_THREADS, 0));
        PGX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, w.device));
        w.passBlocksMax = std::max(1, std::min(cus, perCU * cus));
    }
    const int want = w.persistentBlocksWanted > 0 ? w.persistentBlocksWanted : div_up(std::max(maxColour, 1), PASS_THREADS);
    return std::max(1, std::min(want, w.passBlocksMax));
}

// Grid of the one-XCD pass kernel: about an eighth of it lands on each XCD, so 8 x enough
// workgroups for the largest colour in one round on one XCD, never more than can be
// co-resident, and at most XPASS_MAX_GRID so no XCD's registration byte can overflow.
static int xpass_blocks(World& w, int maxColour) {
    if (w.xpassBlocksMax == 0) {
        int perCU = 0, cus = 0;
        PGX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, (const void*)k_solve_xpass, XPASS_THREADS, 0));
        PGX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, w.device));
        w.xpassBlocksMax = std::max(8, std::min(XPASS_MAX_GRID, perCU * cus));
    }
    const int perXcd = w.persistentBlocksWanted > 0 ? div_up(w.persistentBlocksWanted, 8)
                                                    : div_up(std::max(maxColour, 1), XPASS_THREADS);
    return std::max(8, std::min(8 * perXcd, w.xpassBlocksMax));
}

// The start/stop events of one timed k_solve launch (timing mode, first substep of the
// step). They go to hipExtLaunchKernelGGL, which stamps them from the kernel's own dispatch
// packet: the kernel's execution time, without the launch latency an event recorded on the
// stream before and after the launch would add.
static void solve_events(World& w, int color, hipEvent_t* start, hipEvent_t* stop, int colorEnd = 0) {
    int idx = 2 * w.solveEvUsed;
    while ((int)w.solveEv.size() <= idx + 1) {
        hipEvent_t e;
        PGX_HIP(hipEventCreate(&e));
        w.solveEv.push_back(e);
    }
    *start = w.solveEv[idx];
    *stop = w.solveEv[idx + 1];
    if ((int)w.solveEvColor.size() <= w.solveEvUsed) {
        w.solveEvColor.resize(w.solveEvUsed + 1);
        w.solveEvColorEnd.resize(w.solveEvUsed + 1);
    }
    w.solveEvColor[w.solveEvUsed] = color;
    w.solveEvColorEnd[w.solveEvUsed] = colorEnd;
    ++w.solveEvUsed;
}

// per-world passes need every world's dynamic bodies contiguous and their solver
// velocities within one workgroup's LDS
bool sv_world_pass_fits(World& w) {
    if (!w.worldsContiguous) return false;
    if (w.ldsMax == 0) PGX_HIP(hipDeviceGetAttribute(&w.ldsMax, hipDeviceAttributeMaxSharedMemoryPerBlock, w.device));
    return 2 * sizeof(float4) * (size_t)std::max(1, w.maxWorldDyn) <= (size_t)w.ldsMax;
}

// one colour [b, e) in workgroups of `block` threads; with events, stamped from the
// launch's own dispatch packet
// Workgroups of a colour launch over slots [b, e) in chunks of bs homed on their XCD
// (homed_chunk): 8 per round, as many rounds as the XCD with the most chunks needs.
static int homed_blocks(int b, int e, int bs) {
    const int q = 256 / bs, period = 8 * q, k0 = b / bs, k1 = (e - 1) / bs + 1;   // chunks [k0, k1)
    auto below = [&](int k, int x) {   // home-x chunks below chunk k
        const int g = k / period, r = k - g * period;
        return q * g + std::min(std::max(r - q * x, 0), q);
    };
    int most = 0;
    for (int x = 0; x < 8; ++x) most = std::max(most, below(k1, x) - below(k0, x));
    return 8 * most;
}

static void launch_solve(int block, bool eager, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const SubArgs& a, int b,
                         int e, int useBias, float h, int skipSoft) {
    auto go = [&](auto kernel, int bs) {
        const int nb = homed_blocks(b, e, bs);
        if (e0) hipExtLaunchKernelGGL(kernel, dim3(nb), dim3(bs), 0, s, e0, e1, 0, a, b, e, useBias, h, skipSoft);
        else hipLaunchKernelGGL(kernel, dim3(nb), dim3(bs), 0, s, a, b, e, useBias, h, skipSoft);
    };
    if (eager) {
        if (block == 256) go(k_solve<256, true>, 256);
        else if (block == 128) go(k_solve<128, true>, 128);
        else go(k_solve<64, true>, 64);
    } else {
        if (block == 256) go(k_solve<256, false>, 256);
        else if (block == 128) go(k_solve<128, false>, 128);
        else go(k_solve<64, false>, 64);
    }
}

// Buffers and launch attributes of the substep loop, set before a graph capture of it (no
// allocation under capture); idempotent.
void sv_presize(World& w) {
    w.tickets.alloc((size_t)(w.iterations + 1) * FLOW_QUEUES);   // k_prepare clears them
    if (w.tune.colorBatch > 0) {
        w.batchLoc.alloc(2 * (size_t)w.manifoldCap);
        w.batchList.alloc(2 * (size_t)w.manifoldCap);
        w.batchN.alloc(BATCH_MAX_GROUPS);
        if (!w.batchLdsSet) {
            PGX_HIP(hipFuncSetAttribute((const void*)k_solve_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)(2 * BATCH_BODY_F4 * sizeof(float4) * BATCH_MAX_MANIFOLDS)));
            w.batchLdsSet = true;
        }
    }
}

void sv_substeps(World& w, int M, float dt) {
    hipStream_t s = w.stream;
    w.solveEvUsed = 0;
    sv_presize(w);
    SubArgs a = make_args(w, M, dt);
    if (w.execMode == EXEC_COOPERATIVE && w.J == 0) {
        int blocks = persistent_blocks(w);
        void* args[] = {&a};
        if (w.timing) PGX_HIP(hipEventRecord(w.ev[6], s));
        PGX_HIP(hipLaunchCooperativeKernel((const void*)k_step_persistent, dim3(blocks), dim3(256), args, 0, s));
        if (w.timing) PGX_HIP(hipEventRecord(w.ev[7], s));
        return;
    }
    int nbP = div_up(std::max(M + w.D, 1), 256), nbD = div_up(std::max(w.D, 1), 256);
    int nColors = a.nColors;
    const bool flow = w.execMode == EXEC_FLOW || w.execMode == EXEC_FLOW_XCD;
    FlowArgs f{};
    f.pub = w.flowPub.p;
    if (flow && M > 0) {
        f.xcdScope = w.execMode == EXEC_FLOW_XCD;
        f.nQueues = f.xcdScope ? std::min(w.W, w.nXcd) : 1;
        f.qStart = w.qStart.p;
        f.qList = nullptr;
        if (f.nQueues > 1) {
            // several worlds: a stable partition of the solve order by world % queues
            hipLaunchKernelGGL(k_flow_keys, dim3(div_up(M, 256)), dim3(256), 0, s, M, w.sCanon.p, w.mfKey.p,
                               2 * w.localBits, f.nQueues, w.colorKeys.p, w.qList.p);
            radix_sort_u64(s, w.colorKeys.p, w.qList.p, M, 4, w.sortT);
            hipLaunchKernelGGL(k_flow_starts, dim3(1), dim3(64), 0, s, M, f.nQueues, w.colorKeys.p, w.qStart.p);
            f.qList = w.qList.p;
        } else {
            hipLaunchKernelGGL(k_flow_identity, dim3(1), dim3(64), 0, s, M, f.nQueues, w.qStart.p);
        }
        PGX_LAUNCH_CHECK();
    }
    // per-world passes: the (colour, world) runs once per step
    const bool worlds = w.execMode == EXEC_WORLDS && sv_world_pass_fits(w) && M > 0;
    WorldPassArgs wp{};
    if (worlds) {
        wp.W = w.W; wp.nColors = nColors; wp.wDyn = w.wDyn.p;
        const size_t cells = (size_t)std::max(nColors, 1) * w.W;
        w.wcS.alloc(cells); w.wcE.alloc(cells);
        fill_i32(s, w.wcS.p, 0, (int)cells);
        fill_i32(s, w.wcE.p, 0, (int)cells);
        hipLaunchKernelGGL(k_world_ranges, dim3(div_up(M, 256)), dim3(256), 0, s, M, w.W, w.sFlags.p, w.sWorld.p,
                           w.wcS.p, w.wcE.p);
        PGX_LAUNCH_CHECK();
        wp.wcS = w.wcS.p; wp.wcE = w.wcE.p;
    }
    // worlds by size: each size class in its own launch, with LDS for its largest world
    const size_t worldLds[2] = {2 * sizeof(float4) * (size_t)WORLD_SMALL,
                                2 * sizeof(float4) * (size_t)std::max(1, w.maxWorldDyn)};
    if (worlds && worldLds[1] > 65536)
        PGX_HIP(hipFuncSetAttribute((const void*)k_world_pass, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)worldLds[1]));
    w.lastQueues = flow && M > 0 ? f.nQueues : 0;
    w.lastFlowPasses = flow && M > 0 ? w.substeps * (w.iterations + 1) : 0;
    // waves of the dependency-driven pass: enough to hold every manifold at once, at most
    // what the chip keeps resident comfortably; on XCD queues, 64 workgroups per XCD
    const int flowBlocks = f.xcdScope ? 64 * w.nXcd : std::max(1, std::min(div_up(std::max(M, 1), 256), 2048));
    // launch groups of the colour passes: (first colour, end colour, batched run index + 1);
    // a run holds consecutive colours of at most colorBatch manifolds each, at most
    // BATCH_MAX_MANIFOLDS in all (its bodies' velocities fit in LDS), two colours or more
    std::vector<int3> groups;
    BatchGroups bg{};
    for (int c = 0; c < nColors;) {
        const int n = w.colorStartHost[c + 1] - w.colorStartHost[c];
        if (w.tune.colorBatch > 0 && n <= w.tune.colorBatch && bg.n < BATCH_MAX_GROUPS) {
            int c1 = c + 1;
            while (c1 < nColors && w.colorStartHost[c1 + 1] - w.colorStartHost[c1] <= w.tune.colorBatch &&
                   w.colorStartHost[c1 + 1] - w.colorStartHost[c] <= BATCH_MAX_MANIFOLDS)
                ++c1;
            if (c1 - c >= 2) {
                bg.S[bg.n] = w.colorStartHost[c];
                bg.E[bg.n] = w.colorStartHost[c1];
                groups.push_back(make_int3(c, c1, ++bg.n));
                c = c1;
                continue;
            }
        }
        if (n > 0) groups.push_back(make_int3(c, c + 1, 0));
        ++c;
    }
    if (bg.n > 0) {   // the runs' body slots, fixed for the step
        sv_presize(w);
        hipLaunchKernelGGL(k_batch_map, dim3(1), dim3(BATCH_MAP_THREADS), 0, s, bg, w.sB0.p, w.sB1.p, w.batchLoc.p,
                           w.batchList.p, w.batchN.p);
        PGX_LAUNCH_CHECK();
    }
    int passIndex = 0;
    const bool onePass = w.execMode == EXEC_PASS && M > 0 && nColors > 0;
    const bool xcdPass = w.execMode == EXEC_XPASS && M > 0 && nColors > 0;
    int passGrid = 0;
    if (onePass || xcdPass) {
        int maxColour = 0;
        for (int c = 0; c < nColors; ++c) maxColour = std::max(maxColour, w.colorStartHost[c + 1] - w.colorStartHost[c]);
        passGrid = onePass ? pass_blocks(w, maxColour) : xpass_blocks(w, maxColour);
    }
    auto pass = [&](bool rec, int useBias, float h, int skipSoft) {
        const int p = passIndex++;
        if (M == 0) return;
        if (xcdPass) {
            int nc = nColors;
            int32_t* ctl = w.xpassCtl.p + XPASS_CTL * p;
            void* args[] = {&a, &nc, &useBias, &h, &skipSoft, &ctl};
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (rec) {
                solve_events(w, -1, &e0, &e1);
                PGX_HIP(hipEventRecord(e0, s));
            }
            PGX_HIP(hipLaunchCooperativeKernel((const void*)k_solve_xpass, dim3(passGrid), dim3(XPASS_THREADS), args,
                                               0, s));
            if (rec) PGX_HIP(hipEventRecord(e1, s));
            return;
        }
        if (onePass) {
            // barriers of this substep so far: (nColors - 1) per earlier pass, each counted
            // once per workgroup; the count is cleared by k_prepare
            int base = p * (nColors - 1) * passGrid;
            int nc = nColors;
            void* args[] = {&a, &nc, &useBias, &h, &skipSoft, &base};
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (rec) {
                solve_events(w, -1, &e0, &e1);
                PGX_HIP(hipEventRecord(e0, s));
            }
            PGX_HIP(hipLaunchCooperativeKernel((const void*)k_solve_pass, dim3(passGrid), dim3(PASS_THREADS), args, 0,
                                               s));
            if (rec) PGX_HIP(hipEventRecord(e1, s));
            return;
        }
        if (flow) {
            f.pass = p; f.useBias = useBias; f.skipSoft = skipSoft; f.h = h;
            f.epoch = ++w.flowEpoch;
            if (rec) {
                hipEvent_t e0, e1;
                solve_events(w, -1, &e0, &e1);
                hipExtLaunchKernelGGL(k_solve_flow, dim3(flowBlocks), dim3(256), 0, s, e0, e1, 0, a, f);
            } else {
                hipLaunchKernelGGL(k_solve_flow, dim3(flowBlocks), dim3(256), 0, s, a, f);
            }
            return;
        }
        if (worlds) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (rec) solve_events(w, -1, &e0, &e1);
            for (int k = 0; k < 2; ++k) {
                const int nw = (int)w.worldListHost[k].size();
                if (!nw) continue;
                WorldPassArgs q = wp;
                q.list = w.worldList[k].p;
                // the events bracket both launches: the first stamps the start, the last the end
                const bool first = k == 0 || w.worldListHost[0].empty(), last = k == 1 || w.worldListHost[1].empty();
                if (rec && (first || last))
                    hipExtLaunchKernelGGL(k_world_pass, dim3(nw), dim3(WORLD_THREADS), worldLds[k], s,
                                          first ? e0 : nullptr, last ? e1 : nullptr, 0, a, q, useBias, h, skipSoft);
                else
                    hipLaunchKernelGGL(k_world_pass, dim3(nw), dim3(WORLD_THREADS), worldLds[k], s, a, q, useBias, h,
                                       skipSoft);
            }
            return;
        }
        // one launch per colour (a colour's manifolds share no dynamic body), runs of small
        // colours in one single-workgroup launch
        for (const int3& g : groups) {
            if (g.z) {
                const int q = g.z - 1;
                // 2 bodies per manifold at most
                const size_t lds = 2 * BATCH_BODY_F4 * sizeof(float4) * (size_t)(bg.E[q] - bg.S[q]);
                if (rec) {   // a run of colours [g.x, g.y), stamped from its dispatch packet
                    hipEvent_t e0, e1;
                    solve_events(w, g.x, &e0, &e1, g.y);
                    hipExtLaunchKernelGGL(k_solve_batch, dim3(1), dim3(BATCH_THREADS), lds, s, e0, e1, 0, a, g.x, g.y,
                                          bg.S[q], w.batchLoc.p, w.batchList.p, w.batchN.p + q, useBias, h, skipSoft);
                } else {
                    hipLaunchKernelGGL(k_solve_batch, dim3(1), dim3(BATCH_THREADS), lds, s, a, g.x, g.y, bg.S[q],
                                       w.batchLoc.p, w.batchList.p, w.batchN.p + q, useBias, h, skipSoft);
                }
                continue;
            }
            const int c = g.x, b = w.colorStartHost[c], e = w.colorStartHost[c + 1];
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (rec) solve_events(w, c, &e0, &e1);
            launch_solve(w.tune.solveBlock, e - b <= w.tune.eagerMax, s, e0, e1, a, b, e, useBias, h, skipSoft);
        }
    };
    jt_step_setup(w);
    for (int sub = 0; sub < w.substeps; ++sub) {
        const bool rec = w.timing && sub == 0;
        passIndex = 0;
        // rows + preSolve of the contacts, velocity integration; then the joint rows, whose
        // preSolve follows the contacts' (src/Physecs.cpp:430-479)
        hipLaunchKernelGGL(k_prepare, dim3(nbP), dim3(256), 0, s, a, sub);
        jt_update(w, a.h);
        jt_colors(w, a.h, 0, false);
        for (int it = 0; it < w.iterations; ++it) {
            pass(rec, 1, a.h, 0);
            jt_colors(w, a.h, 1, it == 0);   // joint colours after the contacts (Physecs.cpp:482-490)
        }
        hipLaunchKernelGGL(k_integrate_positions, dim3(nbD), dim3(256), 0, s, a);
        pass(rec, 0, 0.f, 1);
        hipLaunchKernelGGL(k_writeback, dim3(nbD), dim3(256), 0, s, a);
        PGX_LAUNCH_CHECK();
    }
}

// algorithmic bytes of one step's substep loop (per-phase model in the measurement notes)
double sv_step_bytes(const World& w, int M, long points) {
    double N = w.D, C = (double)points, Mm = M;
    double perSub = 164 * Mm + 96 * C + 168 * N + 80 * Mm + 116 * C + (w.iterations + 1) * (284 * C + 12 * Mm) +
                    120 * N + 48 * N;
    return w.substeps * perSub;
}

void sv_finish_step(World& w, int M) {
    // contact cache of this step becomes the previous one (Physecs.cpp:553)
    w.cacheCur ^= 1;
    w.cacheCount[w.cacheCur] = M;
    bp_compute_bounds(w, true);
}

}  // namespace pgx
