// Device-wide primitives: a stable LSD radix sort of 64-bit keys with 32-bit values, and an
// exclusive scan, both written for 64-lane CDNA wavefronts.
#include "pgx_internal.h"

namespace pgx {

// ------------------------------------------------------------------ radix sort
// LSD radix sort, one pass per digit of up to 8 bits, each pass stable.
//
// Ranking (both paths): a wave walks its run of elements in item order; for item k, eight
// ballots over the digit bits give the lanes holding the same digit; the lane's rank is the
// wave's running count of that digit (in LDS) plus the same-digit lanes below it, and the
// lowest such lane adds the group's size to the count. The wave counts are then scanned
// across the block's waves per digit, so ranks follow input order.
//
// Per pass, k_rs_hist counts each tile's digits in LDS into H, and k_rs_scatter gives tile
// b's digit d the output offset sum(H[d' < d][*]) + sum(H[d][b' < b]). With up to
// RS_SUM_TILES tiles, H is tile-major (H[b * 256 + d]) and each block sums its own offsets,
// every thread loading its digit's count in every tile at once (coalesced across the
// block's digits, one memory round trip); with more tiles, H is digit-major (H[d * nb + b]),
// whose exclusive scan is the offset table, scanned in place before the scatter reads it.
constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / WAVE;
constexpr int RS_BINS = 256;
// up to this many tiles, every block sums its offsets from the histogram table itself
constexpr int RS_SUM_TILES = 64;

__device__ __forceinline__ uint32_t rs_digit(uint64_t k, int shift, int dbits) {
    return (uint32_t)(k >> shift) & ((1u << dbits) - 1u);
}

// rank of the lane's element among the wave's elements of the same digit so far; bumps the
// wave's count of the digit (counts: this wave's RS_BINS counters in LDS)
__device__ __forceinline__ int32_t rs_rank(uint32_t d, bool ok, int dbits, uint64_t lt, int32_t* counts) {
    uint64_t same = __ballot(ok);
    for (int b = 0; b < dbits; ++b) {
        const uint64_t bal = __ballot((d >> b) & 1u);
        same &= ((d >> b) & 1u) ? bal : ~bal;
    }
    const int32_t prior = counts[d];
    if (ok && (same & lt) == 0ull) counts[d] = prior + (int32_t)__popcll(same);
    return prior + (int32_t)__popcll(same & lt);
}

// The tile's digit counts with LDS atomics, every key load issued before the counting. (A
// ballot-ranked count per wave, free of the atomics' bank conflicts on nearly sorted keys,
// measured slower: 11.9 against 6.4 us per pass on the 32k scene's 16-key tiles.)
template <int ITEMS>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t* __restrict__ keys, int n, int shift, int dbits,
                                                        int nb, int32_t* __restrict__ H, int digitMajor) {
    __shared__ int32_t cnt[RS_BINS];
    constexpr int TILE = RS_THREADS * ITEMS;
    cnt[threadIdx.x] = 0;
    const int base = blockIdx.x * TILE;
    uint32_t d[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const int i = base + k * RS_THREADS + threadIdx.x;
        d[k] = i < n ? rs_digit(keys[i], shift, dbits) : RS_BINS;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k)
        if (d[k] < RS_BINS) atomicAdd(&cnt[d[k]], 1);
    __syncthreads();
    if (digitMajor) H[(size_t)threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
    else H[(size_t)blockIdx.x * RS_BINS + threadIdx.x] = cnt[threadIdx.x];
}

template <int ITEMS>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t* __restrict__ kin, const int32_t* __restrict__ vin,
                                                           uint64_t* __restrict__ kout, int32_t* __restrict__ vout, int n,
                                                           int shift, int dbits, int nb, const int32_t* __restrict__ H,
                                                           int scanned) {
    constexpr int TILE = RS_THREADS * ITEMS;
    __shared__ int32_t waveCnt[RS_WAVES][RS_BINS];
    __shared__ int32_t off[RS_BINS];
    __shared__ int32_t wsum[RS_WAVES];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
    for (int w = 0; w < RS_WAVES; ++w) waveCnt[w][tid] = 0;
    // this tile's offset for digit tid: all tiles' smaller digits + earlier tiles' digit tid
    if (scanned) {
        off[tid] = H[(size_t)tid * nb + blockIdx.x];
        __syncthreads();
    } else {
        // the digit's count in each of the nb <= RS_SUM_TILES tiles (tile-major), all loads
        // issued before the sums
        int32_t v[RS_SUM_TILES];
#pragma unroll
        for (int b = 0; b < RS_SUM_TILES; ++b) v[b] = b < nb ? H[(size_t)b * RS_BINS + tid] : 0;
        int32_t total = 0, pre = 0;
#pragma unroll
        for (int b = 0; b < RS_SUM_TILES; ++b) {
            total += v[b];
            pre += b < (int)blockIdx.x ? v[b] : 0;
        }
        int32_t x = total;
        for (int o = 1; o < WAVE; o <<= 1) {
            const int32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == WAVE - 1) wsum[wv] = x;
        __syncthreads();
        int32_t before = 0;
        for (int w = 0; w < RS_WAVES; ++w) before += w < wv ? wsum[w] : 0;
        off[tid] = before + x - total + pre;
    }
    // the wave's elements in input order: item k covers [k * 64, k * 64 + 64) of its run
    const int runBase = blockIdx.x * TILE + wv * WAVE * ITEMS;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint64_t key[ITEMS];
    int32_t val[ITEMS];
    int32_t rank[ITEMS];
    for (int k = 0; k < ITEMS; ++k) {
        const int i = runBase + k * WAVE + lane;
        const bool ok = i < n;
        key[k] = ok ? kin[i] : 0ull;
        val[k] = ok ? vin[i] : 0;
    }
    for (int k = 0; k < ITEMS; ++k)
        rank[k] = rs_rank(rs_digit(key[k], shift, dbits), runBase + k * WAVE + lane < n, dbits, lt, waveCnt[wv]);
    __syncthreads();
    {   // exclusive scan of the wave counts across waves, per digit
        int32_t run = 0;
        for (int w = 0; w < RS_WAVES; ++w) {
            const int32_t c = waveCnt[w][tid];
            waveCnt[w][tid] = run;
            run += c;
        }
    }
    __syncthreads();
    for (int k = 0; k < ITEMS; ++k) {
        if (runBase + k * WAVE + lane >= n) continue;
        const uint32_t d = rs_digit(key[k], shift, dbits);
        const int pos = off[d] + waveCnt[wv][d] + rank[k];
        kout[pos] = key[k];
        vout[pos] = val[k];
    }
}

// An input of at most one tile (RS_SINGLE keys): every digit pass in one workgroup, the keys
// and values held in LDS between passes, each pass ranked and placed as k_rs_scatter places
// a single tile (no other tile's counts), so the output is the same stable order in one
// launch instead of two per digit.
constexpr int RS_SINGLE_ITEMS = 16;
constexpr int RS_SINGLE = RS_THREADS * RS_SINGLE_ITEMS;
__global__ __launch_bounds__(RS_THREADS) void k_rs_single(uint64_t* __restrict__ keys, int32_t* __restrict__ vals,
                                                          int n, int bits) {
    constexpr int ITEMS = RS_SINGLE_ITEMS;
    __shared__ uint64_t sk[RS_SINGLE];
    __shared__ int32_t sv[RS_SINGLE];
    __shared__ int32_t waveCnt[RS_WAVES][RS_BINS];
    __shared__ int32_t off[RS_BINS];
    __shared__ int32_t wsum[RS_WAVES];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
    for (int i = tid; i < n; i += RS_THREADS) {
        sk[i] = keys[i];
        sv[i] = vals[i];
    }
    const int runBase = wv * WAVE * ITEMS;   // the wave's elements: item k covers [k * 64, k * 64 + 64) of its run
    const uint64_t lt = (1ull << lane) - 1ull;
    const int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p, dbits = min(8, bits - 8 * p);
        __syncthreads();   // the last pass's placement (or the load) is done, and its reads of the counts
        for (int w = 0; w < RS_WAVES; ++w) waveCnt[w][tid] = 0;
        uint64_t key[ITEMS];
        int32_t val[ITEMS];
        int32_t rank[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const int i = runBase + k * WAVE + lane;
            const bool ok = i < n;
            key[k] = ok ? sk[i] : 0ull;
            val[k] = ok ? sv[i] : 0;
        }
        __syncthreads();   // counters cleared
        for (int k = 0; k < ITEMS; ++k)
            rank[k] = rs_rank(rs_digit(key[k], shift, dbits), runBase + k * WAVE + lane < n, dbits, lt, waveCnt[wv]);
        __syncthreads();
        // digit tid: the waves' counts to exclusive prefixes across waves, and the digit's total
        int32_t total = 0;
        for (int w = 0; w < RS_WAVES; ++w) {
            const int32_t c = waveCnt[w][tid];
            waveCnt[w][tid] = total;
            total += c;
        }
        // the digit's offset: the exclusive scan of the totals over the digits
        int32_t x = total;
        for (int o = 1; o < WAVE; o <<= 1) {
            const int32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == WAVE - 1) wsum[wv] = x;
        __syncthreads();
        int32_t before = 0;
        for (int w = 0; w < RS_WAVES; ++w) before += w < wv ? wsum[w] : 0;
        off[tid] = before + x - total;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            if (runBase + k * WAVE + lane >= n) continue;
            const uint32_t d = rs_digit(key[k], shift, dbits);
            const int pos = off[d] + waveCnt[wv][d] + rank[k];
            sk[pos] = key[k];
            sv[pos] = val[k];
        }
    }
    __syncthreads();
    for (int i = tid; i < n; i += RS_THREADS) {
        keys[i] = sk[i];
        vals[i] = sv[i];
    }
}

__global__ void k_rs_copy(const uint64_t* __restrict__ ks, const int32_t* __restrict__ vs, uint64_t* __restrict__ kd,
                          int32_t* __restrict__ vd, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        kd[i] = ks[i];
        vd[i] = vs[i];
    }
}

// tiles: 512 keys up to 32 Ki keys, 1024 up to 64 Ki, 4096 above (small inputs in more
// blocks: a pass is a short dependent chain per block)
static int rs_items(int n) { return n <= 32 * 1024 ? 2 : n <= 64 * 1024 ? 4 : 16; }

RadixScratch radix_sort_scratch(int n, int bits) {
    RadixScratch r;
    if (n <= 1 || bits <= 0) return r;
    r.keys = (size_t)n;
    r.vals = (size_t)n;
    const int nb = div_up(n, RS_THREADS * rs_items(n));
    r.hist = (size_t)RS_BINS * nb;
    r.sums = nb > RS_SUM_TILES ? (size_t)scan_blocks(RS_BINS * nb) : 0;
    return r;
}

template <int ITEMS>
static void radix_sort_tiles(hipStream_t s, uint64_t* keys, int32_t* vals, int n, int bits, uint64_t* kalt,
                             int32_t* valt, int32_t* hist, int32_t* sums) {
    constexpr int TILE = RS_THREADS * ITEMS;
    const int nb = div_up(n, TILE);
    const int passes = div_up(bits, 8);
    const bool scanned = nb > RS_SUM_TILES;
    uint64_t* kb[2] = {keys, kalt};
    int32_t* vb[2] = {vals, valt};
    for (int p = 0; p < passes; ++p) {
        const int dbits = std::min(8, bits - 8 * p);
        hipLaunchKernelGGL(k_rs_hist<ITEMS>, dim3(nb), dim3(RS_THREADS), 0, s, kb[p & 1], n, 8 * p, dbits, nb, hist,
                           scanned ? 1 : 0);
        if (scanned) exclusive_scan_i32_into(s, hist, hist, RS_BINS * nb, nullptr, sums);
        hipLaunchKernelGGL(k_rs_scatter<ITEMS>, dim3(nb), dim3(RS_THREADS), 0, s, kb[p & 1], vb[p & 1],
                           kb[(p + 1) & 1], vb[(p + 1) & 1], n, 8 * p, dbits, nb, hist, scanned ? 1 : 0);
    }
    if (passes & 1)
        hipLaunchKernelGGL(k_rs_copy, dim3(div_up(n, 256)), dim3(256), 0, s, kalt, valt, keys, vals, n);
    PGX_LAUNCH_CHECK();
}

void radix_sort_u64_into(hipStream_t s, uint64_t* keys, int32_t* vals, int n, int bits, uint64_t* kalt,
                         int32_t* valt, int32_t* hist, int32_t* sums) {
    if (n <= 1 || bits <= 0) return;
    bits = std::min(bits, 64);
    if (n <= RS_SINGLE) {   // one tile: every pass in one launch
        hipLaunchKernelGGL(k_rs_single, dim3(1), dim3(RS_THREADS), 0, s, keys, vals, n, bits);
        PGX_LAUNCH_CHECK();
        return;
    }
    if (rs_items(n) == 2) {
        radix_sort_tiles<2>(s, keys, vals, n, bits, kalt, valt, hist, sums);
    } else if (rs_items(n) == 4) {
        radix_sort_tiles<4>(s, keys, vals, n, bits, kalt, valt, hist, sums);
    } else {
        radix_sort_tiles<16>(s, keys, vals, n, bits, kalt, valt, hist, sums);
    }
}

// Scratch sizes rounded up to a power of two: the sorts' sizes change from step to step, and
// growing a buffer to the exact size would free and allocate again (and hipFree waits for the
// device) every time a step sets a new maximum.
static size_t scratch_size(size_t n) {
    size_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

void radix_sort_u64(hipStream_t s, uint64_t* keys, int32_t* vals, int n, int bits, SortTemp& t) {
    const RadixScratch r = radix_sort_scratch(n, bits);
    if (!r.keys) return;
    t.kalt.alloc(scratch_size(r.keys));
    t.valt.alloc(scratch_size(r.vals));
    t.hist.alloc(scratch_size(std::max<size_t>(r.hist, 1)));
    t.scan.sums.alloc(scratch_size(std::max<size_t>(r.sums, 1)));
    radix_sort_u64_into(s, keys, vals, n, bits, t.kalt.p, t.valt.p, t.hist.p, t.scan.sums.p);
}

// ------------------------------------------------------------------ exclusive scan
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t* total) {
    __shared__ int32_t waveSums[SCAN_THREADS / WAVE];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wv = threadIdx.x / WAVE;
    int32_t x = v;
    for (int off = 1; off < WAVE; off <<= 1) {
        int32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == WAVE - 1) waveSums[wv] = x;
    __syncthreads();
    int32_t waveOff = 0, tot = 0;
    for (int k = 0; k < SCAN_THREADS / WAVE; ++k) {
        if (k < wv) waveOff += waveSums[k];
        tot += waveSums[k];
    }
    __syncthreads();
    *total = tot;
    return waveOff + x - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const int32_t* __restrict__ in, int n,
                                                              int32_t* __restrict__ sums) {
    int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int32_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i)
        if (base + i < n) s += in[base + i];
    int32_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_sums(int32_t* __restrict__ sums, int m, int32_t* total) {
    // one block: scans m block sums in chunks of SCAN_THREADS
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < m; base += SCAN_THREADS) {
        int i = base + threadIdx.x;
        int32_t v = i < m ? sums[i] : 0;
        int32_t tot;
        int32_t ex = block_exclusive_scan(v, &tot);
        int32_t c = carry;
        if (i < m) sums[i] = c + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const int32_t* __restrict__ in, int n,
                                                             const int32_t* __restrict__ sums,
                                                             int32_t* __restrict__ out) {
    int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int32_t v[SCAN_ITEMS];
    int32_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        s += v[i];
    }
    int32_t tot;
    int32_t ex = block_exclusive_scan(s, &tot) + sums[blockIdx.x];
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < n) out[base + i] = ex;
        ex += v[i];
    }
}

int scan_blocks(int n) { return div_up(std::max(n, 1), SCAN_TILE); }

void exclusive_scan_i32_into(hipStream_t s, const int32_t* in, int32_t* out, int n, int32_t* total, int32_t* sums) {
    if (n <= 0) {
        if (total) fill_i32(s, total, 0, 1);
        return;
    }
    int nb = scan_blocks(n);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, s, in, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_THREADS), 0, s, sums, nb, total);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_THREADS), 0, s, in, n, sums, out);
    PGX_LAUNCH_CHECK();
}

void exclusive_scan_i32(hipStream_t s, const int32_t* in, int32_t* out, int n, int32_t* total, ScanTemp& t) {
    if (n > 0) t.sums.alloc(scratch_size(scan_blocks(n)));
    exclusive_scan_i32_into(s, in, out, n, total, t.sums.p);
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void fill_i32(hipStream_t s, int32_t* p, int32_t v, int n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(div_up(n, 256)), dim3(256), 0, s, p, v, n);
    PGX_LAUNCH_CHECK();
}
void fill_u64(hipStream_t s, uint64_t* p, uint64_t v, int n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u64, dim3(div_up(n, 256)), dim3(256), 0, s, p, v, n);
    PGX_LAUNCH_CHECK();
}

}  // namespace pgx
