"""Edge cases of the scene queries on the device against the oracle: a world of static
colliders only (no LBVH: the static lists alone), a world with no colliders at all among
batched worlds, and empty query batches."""
import numpy as np
import pytest

from gpu_util import bits, lockstep, log
from physecs_amd import scenes

pytestmark = pytest.mark.gpu
Q = np.array([[0, 0, 0, 1]], np.float32)


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def statics_only():
    sc = scenes.Scene()
    sc.add_static(scenes.BOX, [10, 0.5, 10], [0, -0.5, 0])
    sc.add_static(scenes.SPHERE, [1.0], [3, 1, 0])
    sc.add_static(scenes.CAPSULE, [1.0, 0.3], [-3, 1.3, 0])
    return sc


def test_static_only_world(api):
    sc = statics_only()
    log("[edge] a world of three static colliders: raycast, overlap, MTD, character")
    g, o = lockstep(sc, 2)
    origins = np.array([[-8, 1, 0], [0, 5, 0], [3, 5, 0], [-3, 5, 0.1]], np.float32)
    dirs = np.array([[1, 0, 0], [0, -1, 0], [0, -1, 0], [0, -1, 0]], np.float32)
    gc, gt, _ = g.raycast(origins, dirs, 100.0)
    oc, ot = o.raycast(origins, dirs, 100.0)
    assert np.array_equal(gc, oc) and np.array_equal(bits(gt), bits(ot)) and (gc >= 0).all()
    pos = np.array([[3, 2.2, 0], [-3, 0.2, 0], [0, 0.3, 0]], np.float32)
    t = np.full(3, scenes.SPHERE, np.int32)
    prm = np.array([[0.5], [0.5], [0.5]], np.float32)
    q = np.repeat(Q, 3, 0)
    got, _ = g.overlap(t, prm, pos, q)
    for i in range(3):
        assert np.array_equal(got[i], o.overlap(t[i], prm[i], pos[i], q[i]))
    mtd, _ = g.overlap_mtd(t, prm, pos, q)
    for i in range(3):
        oc, onm = o.overlap_mtd(t[i], prm[i], pos[i], q[i])
        assert np.array_equal(mtd[i][0], oc) and np.array_equal(bits(mtd[i][1]), bits(onm))
    assert sum(len(c) for c, _ in mtd) >= 3
    p2, f2 = g.character_move([[0, -0.05, 0]], Q, [[0.1, 0, 0]], 1.0, 0.3)
    op, of = o.character_move([0, -0.05, 0], Q[0], [0.1, 0, 0], 1.0, 0.3, float(np.float32(np.cos(np.pi / 4))), 2)
    assert np.array_equal(bits(p2[0]), bits(op)) and f2[0] == of == 1


def test_empty_world_and_empty_batches(api):
    sc = scenes.Scene()
    sc.n_worlds = 2
    sc.add_static(scenes.BOX, [10, 0.5, 10], [0, -0.5, 0], world=0)
    sc.add_body(scenes.SPHERE, [0.5], [0, 2, 0], world=0)
    g = api.World(sc, substeps=4, iterations=2)
    g.step(1 / 60)
    # world 1 holds no collider: every query there finds nothing
    col, _, _ = g.raycast([[0, 5, 0]], [[0, -1, 0]], 100.0, world=[1])
    assert col[0] == -1
    hits, _ = g.overlap([scenes.SPHERE], [[50.0]], [[0, 0, 0]], Q, world=[1])
    assert len(hits[0]) == 0
    mtd, _ = g.overlap_mtd([scenes.SPHERE], [[50.0]], [[0, 0, 0]], Q, world=[1])
    assert len(mtd[0][0]) == 0
    # and the same queries in world 0 do find the ground and the sphere
    hits, _ = g.overlap([scenes.SPHERE], [[50.0]], [[0, 0, 0]], Q, world=[0])
    assert list(hits[0]) == [0, 1]
    # empty batches
    c, t, _ = g.raycast(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), 1.0)
    assert len(c) == 0
    assert g.overlap(np.zeros(0, np.int32), np.zeros((0, 4), np.float32), np.zeros((0, 3), np.float32),
                     np.zeros((0, 4), np.float32))[0] == []
    assert g.overlap_mtd(np.zeros(0, np.int32), np.zeros((0, 4), np.float32), np.zeros((0, 3), np.float32),
                         np.zeros((0, 4), np.float32))[0] == []
    p, f = g.character_move(np.zeros((0, 3), np.float32), Q, np.zeros((0, 3), np.float32), 1.0, 0.3)
    assert len(p) == 0 and len(f) == 0
