"""pgx_sort_u64_noalloc, the sort for callers that cannot allocate, against pgx_sort_u64 (the
same sort with allocated scratch) through the C ABI: two calls in a row on device arrays,
the second smaller than the first (the shared static scratch keeps the first call's
contents), both bit-identical to the allocating entry point; sizes beyond its limit are
rejected before anything runs. The test allocates its device arrays itself, through the HIP
runtime the library is linked against (found among the process's mapped libraries)."""
import ctypes

import numpy as np
import pytest

from gpu_util import log

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2
INVALID = -1   # PGX_ERR_INVALID


def _hip(api):
    api.lib()
    with open("/proc/self/maps") as f:
        paths = [line.split()[-1] for line in f if "libamdhip64" in line and line.split()[-1].startswith("/")]
    assert paths, "libamdhip64 not mapped"
    H = ctypes.CDLL(paths[0])
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    H.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    H.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return H


def _noalloc(api):
    fn = api.lib().pgx_sort_u64_noalloc
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    return fn


class _Dev:
    """A device copy of a host array, freed on exit."""

    def __init__(self, H, host):
        self.H, self.nbytes = H, host.nbytes
        self.p = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.p), max(self.nbytes, 1)) == 0
        assert H.hipMemcpy(self.p, host.ctypes.data, self.nbytes, H2D) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert self.H.hipMemcpy(out.ctypes.data, self.p, self.nbytes, D2H) == 0
        return out

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.H.hipFree(self.p)
        return False


@pytest.mark.parametrize("first,second", [((300001, 40), (5000, 12)), ((1 << 20, 64), (70000, 8))])
def test_noalloc_sort_equals_allocating_sort(first, second):
    from physecs_amd import api
    assert api.device_count() > 0, "no GPU visible"
    H = _hip(api)
    fn = _noalloc(api)
    stream = ctypes.c_void_p()
    assert H.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        for call, (n, bits) in enumerate((first, second)):
            log(f"[noalloc] call {call}: n={n} bits={bits}")
            rng = np.random.default_rng(n + bits)
            keys = rng.integers(0, 1 << min(bits, 63), n, dtype=np.uint64)
            if bits == 64:
                keys |= rng.integers(0, 2, n, dtype=np.uint64) << np.uint64(63)
            vals = np.arange(n, dtype=np.int32)
            want_k, want_v = api.sort_u64(keys, vals, bits)
            with _Dev(H, keys) as kd, _Dev(H, vals) as vd:
                rc = fn(0, kd.p, vd.p, n, bits, stream)
                assert rc == 0, api.lib().pgx_last_error()
                assert H.hipStreamSynchronize(stream) == 0
                assert np.array_equal(kd.get(keys), want_k)
                assert np.array_equal(vd.get(vals), want_v)
    finally:
        H.hipStreamDestroy(stream)


def test_noalloc_sort_rejects_oversized_input():
    from physecs_amd import api
    H = _hip(api)
    fn = _noalloc(api)
    small = np.zeros(16, np.uint64)
    with _Dev(H, small) as kd, _Dev(H, np.zeros(16, np.int32)) as vd:
        assert fn(0, kd.p, vd.p, (1 << 20) + 1, 32, None) == INVALID
        assert b"outside" in api.lib().pgx_last_error()
        assert fn(0, kd.p, vd.p, 16, 65, None) == INVALID
        assert np.array_equal(kd.get(small), small)      # nothing ran
