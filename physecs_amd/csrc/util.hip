// Device-wide primitives: stable radix sort of 64-bit keys with 32-bit values (rocPRIM's
// onesweep radix sort through hipCUB, which is stable: equal keys keep their input order),
// and an exclusive scan written for 64-lane CDNA wavefronts.
#include <hipcub/hipcub.hpp>
#include "pgx_internal.h"

namespace pgx {

void radix_sort_u64(hipStream_t s, uint64_t* keys, int32_t* vals, int n, int bits, SortTemp& t) {
    if (n <= 1 || bits <= 0) return;
    t.kalt.alloc(n);
    t.valt.alloc(n);
    hipcub::DoubleBuffer<uint64_t> dk(keys, t.kalt.p);
    hipcub::DoubleBuffer<int32_t> dv(vals, t.valt.p);
    size_t bytes = 0;
    PGX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, dk, dv, n, 0, bits, s));
    t.temp.alloc(std::max<size_t>(bytes, 1));
    PGX_HIP(hipcub::DeviceRadixSort::SortPairs(t.temp.p, bytes, dk, dv, n, 0, bits, s));
    if (dk.Current() != keys) {
        PGX_HIP(hipMemcpyAsync(keys, dk.Current(), sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, s));
        PGX_HIP(hipMemcpyAsync(vals, dv.Current(), sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
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
