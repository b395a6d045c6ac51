"""Scene::overlapWithMinTranslationalDistance on the device (pgx_world_overlap_mtd)
against the oracle on lockstep states: per query the same colliders in the same order
(collider index, then manifold order), and normals and mtds bit for bit. Scenes: mixed
primitives, a convex pile (convex queries too: GJK/EPA), and convex bodies and chains on a
triangle-mesh terrain (the mesh pairs' BVH query, per-triangle tests and emission)."""
import numpy as np
import pytest

from gpu_util import bits, lockstep, log
from physecs_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def queries(rng, n, lo, hi, n_convex):
    t = rng.integers(0, 4 if n_convex else 3, n).astype(np.int32)
    prm = np.zeros((n, 4), np.float32)
    for i in range(n):
        if t[i] == scenes.SPHERE:
            prm[i, 0] = rng.uniform(0.2, 1.2)
        elif t[i] == scenes.CAPSULE:
            prm[i, :2] = [rng.uniform(0.2, 1.2), rng.uniform(0.1, 0.6)]
        elif t[i] == scenes.BOX:
            prm[i, :3] = rng.uniform(0.2, 1.2, 3)
        else:
            prm[i] = [*rng.uniform(0.5, 1.5, 3), int(rng.integers(0, n_convex))]
    pos = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    q = rng.normal(size=(n, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    return t, prm, pos, q


def compare(g, o, t, prm, pos, q):
    got, us = g.overlap_mtd(t, prm, pos, q)
    total = 0
    for i in range(len(t)):
        gc, gnm = got[i]
        oc, onm = o.overlap_mtd(t[i], prm[i], pos[i], q[i])
        assert np.array_equal(gc, oc), (i, int(t[i]), gc[:8], oc[:8])
        assert np.array_equal(bits(gnm), bits(onm)), (i, int(t[i]))
        total += len(gc)
    return got, total, us


@pytest.mark.parametrize("name", ["mixed", "convex", "terrain"])
def test_overlap_mtd_matches_oracle(api, name):
    if name == "mixed":
        sc, lo, hi = scenes.mixed_lattice(7, side=8), (-7, -0.5, -7), (7, 8, 7)
    elif name == "convex":
        sc, lo, hi = scenes.convex_pile(4, n=60), (-3, -0.5, -3), (3, 3, 3)
    else:
        sc, lo, hi = scenes.terrain_drop(3, quads=12, n=40), (-5, -2, -5), (5, 4, 5)
    log(f"[mtd] {name}: {sc.n_dyn} bodies, 30 lockstep steps, then 1500 queries")
    g, o = lockstep(sc, 30, check_every=10)
    rng = np.random.default_rng(4)
    t, prm, pos, q = queries(rng, 1500, lo, hi, len(sc.convex))
    got, total, us = compare(g, o, t, prm, pos, q)
    with_hits = sum(len(c) > 0 for c, _ in got)
    log(f"[mtd] {name}: {total} hits, {with_hits} queries with hits, device {us:.0f} us")
    assert with_hits > 200
    if name == "terrain":
        tri = np.nonzero(sc.arrays()["col_type"] == scenes.TRIANGLE_MESH)[0]
        assert sum(int(np.isin(c, tri).sum()) for c, _ in got) > 100   # the mesh path ran


def test_overlap_mtd_capacity_and_unsupported(api):
    sc = scenes.mixed_lattice(7, side=6)
    g = api.World(sc, substeps=4, iterations=2)
    g.step(1 / 60)
    t = np.array([scenes.BOX], np.int32)
    prm = np.array([[50, 50, 50, 0]], np.float32)
    full, _ = g.overlap_mtd(t, prm, np.zeros((1, 3), np.float32), np.array([[0, 0, 0, 1]], np.float32))
    assert len(full[0][0]) > 100
    with pytest.raises(api.PgxError, match="CAPACITY"):
        g.overlap_mtd(t, prm, np.zeros((1, 3), np.float32), np.array([[0, 0, 0, 1]], np.float32), capacity=10)
    with pytest.raises(api.PgxError, match="UNSUPPORTED"):
        g.overlap_mtd(np.array([scenes.TRIANGLE_MESH], np.int32), np.zeros((1, 4), np.float32),
                      np.zeros((1, 3), np.float32), np.array([[0, 0, 0, 1]], np.float32))
