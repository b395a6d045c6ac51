"""The library's stable radix sort (pgx_sort_u64), the sort behind the broadphase's Morton
codes and canonical pair order, the narrowphase's shape-pair classes and the solve order
by colour, against numpy's stable argsort: sizes around the tile edges and both tile
shapes, key widths from 4 to 42 bits (odd and even digit counts), many equal keys (the
order among them must be the input order), bits above `bits` carried along unsorted, and
the many-tile path whose offsets come from a scanned histogram."""
import numpy as np
import pytest

from gpu_util import log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def _check(api, keys, bits):
    vals = np.arange(len(keys), dtype=np.int32)
    ko, vo = api.sort_u64(keys, vals, bits)
    mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    order = np.argsort(keys & mask, kind="stable")
    assert np.array_equal(vo, vals[order]), bits
    assert np.array_equal(ko, keys[order]), bits


@pytest.mark.parametrize("n", [2, 63, 64, 1023, 1024, 1025, 4095, 4096, 4097, 65535, 65536, 65537, 300001])
@pytest.mark.parametrize("bits", [4, 8, 30, 32])
def test_sort_matches_stable_argsort(api, n, bits):
    log(f"[sort] n={n} bits={bits}")
    rng = np.random.default_rng(n * 64 + bits)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64)
    _check(api, keys, bits)


@pytest.mark.parametrize("bits", [12, 42])
def test_sort_duplicates_and_high_bits(api, bits):
    """Few distinct low keys (long runs of equals) with random bits above `bits`: those bits
    must not take part in the order."""
    log(f"[sort] duplicates bits={bits}")
    rng = np.random.default_rng(bits)
    n = 200000
    low = rng.integers(0, 7, n, dtype=np.uint64) << np.uint64(bits - 3)
    high = rng.integers(0, 1 << 10, n, dtype=np.uint64) << np.uint64(bits)
    _check(api, low | high, bits)


def test_sort_many_tiles(api):
    """2,000,000 keys: over 64 tiles of 4096, so the offsets come from the scanned table."""
    log("[sort] n=2000000 bits=40")
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 1 << 40, 2000000, dtype=np.uint64)
    _check(api, keys, 40)


def test_sort_empty_and_one(api):
    for n in (0, 1):
        ko, vo = api.sort_u64(np.arange(n, dtype=np.uint64), np.arange(n, dtype=np.int32), 8)
        assert len(ko) == n
