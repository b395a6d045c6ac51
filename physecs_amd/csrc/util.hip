// Device-wide primitives: stable LSD radix sort of 64-bit keys with 32-bit values, and an
// exclusive scan. Both are written for 64-lane CDNA wavefronts.
//
// Radix sort, per 8-bit digit:
//   1. hist:    each block counts its tile's digits in LDS  -> hist[digit * nBlocks + block]
//   2. scan:    exclusive scan of hist (digit-major), one block
//   3. scatter: each block ranks its tile stably. A tile is walked in rounds of 256 keys
//               (4 waves x 64 lanes); inside a wave, lanes holding the same digit are found
//               with 8 ballots (one per digit bit), so a key's rank is a popcount over the
//               lanes below it; per-wave digit counts in LDS give the offsets across waves
//               and rounds. Keys keep their input order within a digit (stable).
#include "pgx_internal.h"

namespace pgx {

constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;   // 2048 keys per block
constexpr int SORT_WAVES = SORT_THREADS / WAVE;

__global__ __launch_bounds__(SORT_THREADS) void k_sort_hist(const uint64_t* __restrict__ keys, int n, int shift,
                                                            uint32_t* __restrict__ hist, int nBlocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * SORT_TILE;
    for (int i = threadIdx.x; i < SORT_TILE; i += SORT_THREADS) {
        int idx = base + i;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nBlocks + blockIdx.x] = h[threadIdx.x];
}

// single block exclusive scan of m uint32 values (in place)
__global__ __launch_bounds__(1024) void k_scan_single_u32(uint32_t* __restrict__ a, int m) {
    __shared__ uint32_t partial[1024];
    int per = (m + 1023) / 1024;
    int beg = threadIdx.x * per;
    int end = min(beg + per, m);
    uint32_t s = 0;
    for (int i = beg; i < end; ++i) s += a[i];
    partial[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? partial[threadIdx.x - off] : 0;
        __syncthreads();
        partial[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = partial[threadIdx.x] - s;
    for (int i = beg; i < end; ++i) {
        uint32_t v = a[i];
        a[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(SORT_THREADS) void k_sort_scatter(const uint64_t* __restrict__ keys,
                                                               const int32_t* __restrict__ vals, int n, int shift,
                                                               const uint32_t* __restrict__ hist, int nBlocks,
                                                               uint64_t* __restrict__ okeys,
                                                               int32_t* __restrict__ ovals) {
    __shared__ uint32_t blockOff[256];      // global offset of this block's digits + running count
    __shared__ uint32_t waveCnt[SORT_WAVES][256];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wv = threadIdx.x / WAVE;
    blockOff[threadIdx.x] = hist[threadIdx.x * nBlocks + blockIdx.x];
    for (int k = 0; k < SORT_WAVES; ++k) waveCnt[k][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t lanesBelow = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int base = blockIdx.x * SORT_TILE;
    for (int r = 0; r < SORT_ITEMS; ++r) {
        int idx = base + r * SORT_THREADS + threadIdx.x;
        bool valid = idx < n;
        uint64_t key = valid ? keys[idx] : 0;
        int32_t val = valid ? vals[idx] : 0;
        uint32_t digit = valid ? (uint32_t)((key >> shift) & 255u) : 0u;
        // lanes of this wave with the same digit
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            uint64_t m = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? m : ~m;
        }
        int rankInWave = __popcll(peers & lanesBelow);
        if (valid && rankInWave == 0) waveCnt[wv][digit] = (uint32_t)__popcll(peers);
        __syncthreads();
        // thread t owns digit t: exclusive prefix across waves, then advance the block offset
        {
            uint32_t run = blockOff[threadIdx.x];
            for (int k = 0; k < SORT_WAVES; ++k) {
                uint32_t c = waveCnt[k][threadIdx.x];
                waveCnt[k][threadIdx.x] = run;
                run += c;
            }
            blockOff[threadIdx.x] = run;
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = waveCnt[wv][digit] + (uint32_t)rankInWave;
            okeys[pos] = key;
            ovals[pos] = val;
        }
        __syncthreads();
        for (int k = 0; k < SORT_WAVES; ++k) waveCnt[k][threadIdx.x] = 0;
        __syncthreads();
    }
}

void radix_sort_u64(hipStream_t s, uint64_t* keys, int32_t* vals, int n, int bits, SortTemp& t) {
    if (n <= 1 || bits <= 0) return;
    int nBlocks = div_up(n, SORT_TILE);
    t.hist.alloc((size_t)256 * nBlocks);
    t.kalt.alloc(n);
    t.valt.alloc(n);
    uint64_t* ka = keys;
    int32_t* va = vals;
    uint64_t* kb = t.kalt.p;
    int32_t* vb = t.valt.p;
    int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_sort_hist, dim3(nBlocks), dim3(SORT_THREADS), 0, s, ka, n, shift, t.hist.p, nBlocks);
        hipLaunchKernelGGL(k_scan_single_u32, dim3(1), dim3(1024), 0, s, t.hist.p, 256 * nBlocks);
        hipLaunchKernelGGL(k_sort_scatter, dim3(nBlocks), dim3(SORT_THREADS), 0, s, ka, va, n, shift, t.hist.p,
                           nBlocks, kb, vb);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    PGX_LAUNCH_CHECK();
    if (ka != keys) {
        PGX_HIP(hipMemcpyAsync(keys, ka, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, s));
        PGX_HIP(hipMemcpyAsync(vals, va, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
    }
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

void exclusive_scan_i32(hipStream_t s, const int32_t* in, int32_t* out, int n, int32_t* total, ScanTemp& t) {
    if (n <= 0) {
        if (total) fill_i32(s, total, 0, 1);
        return;
    }
    int nb = div_up(n, SCAN_TILE);
    t.sums.alloc(nb);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, s, in, n, t.sums.p);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_THREADS), 0, s, t.sums.p, nb, total);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_THREADS), 0, s, in, n, t.sums.p, out);
    PGX_LAUNCH_CHECK();
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
