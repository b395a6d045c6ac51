// pgx_sort_u64_noalloc: the library's stable radix sort on device arrays, with its scratch
// (alternate keys and values, digit histograms, scan block sums) in one statically sized
// device array instead of allocated memory, for callers that cannot allocate (code under
// HIP graph capture, serving loops that must not call hipMalloc).
//
// The array is a module global: each GPU the library runs on holds its own copy, at its own
// address, until the process ends. Host code never passes the array's name to a kernel (in
// host code it is only a stand-in symbol): the first call on a device, with that device
// current, asks the runtime for the device address once and keeps it per device. Every call
// on a GPU shares that GPU's copy, so calls that overlap on one GPU (two streams, two
// threads) race; the sort needs no zeroed scratch (every table it reads it writes first).
#include <mutex>

#include "pgx_internal.h"

namespace pgx {

constexpr int NA_MAX_KEYS = 1 << 20;
constexpr size_t NA_ALIGN = 256;
constexpr size_t na_up(size_t b) { return (b + NA_ALIGN - 1) / NA_ALIGN * NA_ALIGN; }
// the largest tables any n <= NA_MAX_KEYS needs: 256 digit bins over the tiles of 4096 keys
// (smaller inputs use 1024-key tiles, at most 64 of them), and the block sums of the
// scanned table
constexpr size_t NA_HIST = 256ull * (NA_MAX_KEYS / 4096);
constexpr size_t NA_SUMS = (256ull * (NA_MAX_KEYS / 4096) + 2047) / 2048;
constexpr size_t NA_OFF_VALS = na_up(8ull * NA_MAX_KEYS);
constexpr size_t NA_OFF_HIST = NA_OFF_VALS + na_up(4ull * NA_MAX_KEYS);
constexpr size_t NA_OFF_SUMS = NA_OFF_HIST + na_up(4 * NA_HIST);
constexpr size_t NA_BYTES = NA_OFF_SUMS + na_up(4 * NA_SUMS);   // 12 MiB + 256 KiB + 256 B

}  // namespace pgx

static __device__ __align__(256) unsigned char scratch[pgx::NA_BYTES];

namespace pgx {

constexpr int NA_MAX_DEVICES = 64;
static std::mutex g_naMutex;
static unsigned char* g_naScratch[NA_MAX_DEVICES] = {};

// the device address of `scratch` on `device`, which must be current (looked up once per
// device)
static unsigned char* scratch_on(int device) {
    std::lock_guard<std::mutex> lock(g_naMutex);
    if (!g_naScratch[device]) {
        void* p = nullptr;
        PGX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw PgxError{PGX_ERR_HIP, "hipGetSymbolAddress returned no address for the sort scratch"};
        g_naScratch[device] = static_cast<unsigned char*>(p);
    }
    return g_naScratch[device];
}

}  // namespace pgx

extern "C" int pgx_sort_u64_noalloc(int device, uint64_t* keys, int32_t* vals, int n, int bits, void* stream) {
    using namespace pgx;
    if (n < 0 || n > NA_MAX_KEYS)
        return report_error(PGX_ERR_INVALID, "pgx_sort_u64_noalloc: n = " + std::to_string(n) +
                                                 " is outside [0, " + std::to_string(NA_MAX_KEYS) + "]");
    if (bits < 0 || bits > 64) return report_error(PGX_ERR_INVALID, "pgx_sort_u64_noalloc: bits must be in [0, 64]");
    if (device < 0 || device >= NA_MAX_DEVICES)
        return report_error(PGX_ERR_INVALID, "pgx_sort_u64_noalloc: device out of range");
    if (n > 1 && (!keys || !vals)) return report_error(PGX_ERR_INVALID, "pgx_sort_u64_noalloc: array is NULL");
    try {
        PGX_HIP(hipSetDevice(device));
        unsigned char* base = scratch_on(device);
        const RadixScratch r = radix_sort_scratch(n, bits);
        if (!r.keys) return PGX_OK;
        if (r.hist > NA_HIST || r.sums > NA_SUMS)
            throw PgxError{PGX_ERR_INVALID, "pgx_sort_u64_noalloc: scratch layout too small for this input"};
        radix_sort_u64_into(static_cast<hipStream_t>(stream), keys, vals, n, bits,
                            reinterpret_cast<uint64_t*>(base), reinterpret_cast<int32_t*>(base + NA_OFF_VALS),
                            reinterpret_cast<int32_t*>(base + NA_OFF_HIST),
                            reinterpret_cast<int32_t*>(base + NA_OFF_SUMS));
        return PGX_OK;
    } catch (const HipError& e) {
        return report_error(PGX_ERR_HIP, std::string("HIP error ") + hipGetErrorString(e.err) + " at " + e.expr);
    } catch (const PgxError& e) {
        return report_error(e.code, e.msg);
    } catch (...) {
        return report_error(PGX_ERR_INVALID, "pgx_sort_u64_noalloc: unknown error");
    }
}
