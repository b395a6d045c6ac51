"""Scene overlap queries on the device (pgx_world_overlap) against the oracle's
Scene::overlap on the same state: the GPU world and the oracle are stepped in lockstep
(bit-identical states), then the same query shapes are tested on both. Every query's list
of overlapping colliders must agree exactly, with and without the data filter, for sphere,
capsule, box and convex-mesh queries (triangle-mesh queries overlap nothing); in batched
worlds each query is tested in its own world. A capacity below the hit count returns the
counts and PGX_ERR_CAPACITY."""
import numpy as np
import pytest

from gpu_util import lockstep, log
from physecs_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def queries(rng, n, lo, hi, n_convex):
    kinds = 4 if n_convex else 3
    t = rng.integers(0, kinds, n).astype(np.int32)
    t[rng.random(n) < 0.03] = scenes.TRIANGLE_MESH
    prm = np.zeros((n, 4), np.float32)
    for i in range(n):
        if t[i] == scenes.SPHERE:
            prm[i, 0] = rng.uniform(0.2, 2.0)
        elif t[i] == scenes.CAPSULE:
            prm[i, :2] = [rng.uniform(0.2, 2.0), rng.uniform(0.1, 0.8)]
        elif t[i] == scenes.BOX:
            prm[i, :3] = rng.uniform(0.2, 2.0, 3)
        elif t[i] == scenes.CONVEX_MESH:
            prm[i] = [*rng.uniform(0.5, 2.5, 3), int(rng.integers(0, n_convex))]
    pos = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    q = rng.normal(size=(n, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    return t, prm, pos, q


def compare(g, o, t, prm, pos, q, filters=None, world=None, col_offset=0):
    got, us = g.overlap(t, prm, pos, q, world=world, filters=filters)
    total = 0
    for i in range(len(t)):
        want = o.overlap(t[i], prm[i], pos[i], q[i], filt=0 if filters is None else int(filters[i])) + col_offset
        assert np.array_equal(got[i], want), (i, int(t[i]), got[i][:10], want[:10])
        total += len(want)
    return got, total, us


@pytest.mark.parametrize("name", ["mixed", "convex"])
def test_overlap_matches_oracle(api, name):
    sc = scenes.mixed_lattice(6, side=8) if name == "mixed" else scenes.convex_pile(3, n=60)
    rng = np.random.default_rng(21)
    arr = sc.arrays()
    arr["col_data"][:] = rng.integers(0, 8, len(arr["col_data"]))
    log(f"[overlap] {name}: {sc.n_dyn} bodies, 30 lockstep steps, then 2000 queries")
    g, o = lockstep(sc, 30, check_every=10)
    lo, hi = (-8, 14) if name == "mixed" else (-4, 5)
    t, prm, pos, q = queries(rng, 2000, lo, hi, len(sc.convex))
    got, total, us = compare(g, o, t, prm, pos, q)
    nonempty = sum(len(x) > 0 for x in got)
    log(f"[overlap] {name}: {total} hits, {nonempty} queries with hits, device {us:.0f} us")
    assert nonempty > 300 and total > 1000
    assert all(len(got[i]) == 0 for i in np.nonzero(t == scenes.TRIANGLE_MESH)[0])
    filters = rng.integers(0, 8, 2000).astype(np.int32)
    compare(g, o, t, prm, pos, q, filters=filters)


def test_overlap_capacity(api):
    sc = scenes.mixed_lattice(6, side=6)
    g = api.World(sc, substeps=4, iterations=2)
    g.step(1 / 60)
    t = np.array([scenes.BOX, scenes.SPHERE], np.int32)
    prm = np.array([[50, 50, 50, 0], [0.5, 0, 0, 0]], np.float32)
    pos = np.zeros((2, 3), np.float32)
    q = np.tile(np.array([0, 0, 0, 1], np.float32), (2, 1))
    full, _ = g.overlap(t, prm, pos, q)
    assert len(full[0]) == len(sc.arrays()["col_type"])          # the huge box meets every collider
    with pytest.raises(api.PgxError, match="CAPACITY"):
        g.overlap(t, prm, pos, q, capacity=len(full[0]) - 1)
    again, _ = g.overlap(t, prm, pos, q, capacity=len(full[0]) + len(full[1]))
    assert all(np.array_equal(a, b) for a, b in zip(full, again))


def test_overlap_in_batched_worlds(api):
    """Queries in one world of a batch equal the same queries in that world alone, and the
    alone world matches the oracle."""
    sc = scenes.batched_worlds(8, n_worlds=12, lo=8, hi=120, ragdoll_frac=0.5)
    log(f"[overlap] 12 worlds, {sc.n_dyn} bodies: 20 steps, queries per world vs each world alone")
    g = api.World(sc, substeps=4, iterations=2)
    for _ in range(20):
        g.step(1 / 60)
    arr = sc.arrays()
    col_world = arr["world"][arr["col_actor"]]
    rng = np.random.default_rng(12)
    for wd in (0, 3, 7, 11):
        sub = sc.world_subscene(wd)
        ga, o = lockstep(sub, 20, check_every=20)
        b = o.bounds()
        dyn = sub.arrays()["dyn_of_actor"][sub.arrays()["col_actor"]] >= 0
        lo, hi = b[dyn, :3].min(axis=0) - 1, b[dyn, 3:].max(axis=0) + 1   # around the world's bodies
        t, prm, pos, q = queries(rng, 400, lo, hi, 0)
        first = int(np.nonzero(col_world == wd)[0][0])
        gb, _ = g.overlap(t, prm, pos, q, world=np.full(400, wd, np.int32))
        got, total, _ = compare(ga, o, t, prm, pos, q)
        for i in range(400):
            assert np.array_equal(gb[i], got[i] + first), (wd, i)
        assert total > 50
