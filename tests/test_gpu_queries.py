"""Scene raycasts on the device (pgx_world_raycast) against the oracle's
Scene::raycastClosest on the same state: the GPU world and the oracle are stepped in
lockstep (bit-identical states), then the same rays are cast on both. The closest hit
collider and its distance must agree bit for bit, with and without the entity filter and
with short maxDistance; in batched worlds each ray is cast in its own world."""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import bits, lockstep, log
from physecs_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def rays_into(rng, lo, hi, n):
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    target = rng.uniform(lo, hi, (n, 3))
    d = target - o
    norm = rng.random(n) < 0.5
    d[norm] /= np.linalg.norm(d[norm], axis=1, keepdims=True)
    axis = rng.random(n) < 0.1                      # some components below the parallel threshold
    d[axis, rng.integers(0, 3, int(axis.sum()))] = 0.0005
    return o, d.astype(np.float32)


def compare(g, o, origins, dirs, maxd, ignore=None, world=None, col_offset=0):
    gc, gt, us = g.raycast(origins, dirs, maxd, world=world, ignore=ignore)
    oc, ot = o.raycast(origins, dirs, maxd, ignore=ignore)
    oc = np.where(oc >= 0, oc + col_offset, -1)
    assert np.array_equal(gc, oc), np.nonzero(gc != oc)[0][:10]
    hit = gc >= 0
    assert np.array_equal(bits(gt[hit]), bits(ot[hit]))
    return gc, gt, us


@pytest.mark.parametrize("name", ["mixed", "convex"])
def test_raycast_matches_oracle(api, name):
    sc = scenes.mixed_lattice(6, side=8) if name == "mixed" else scenes.convex_pile(3, n=60)
    log(f"[raycast] {name}: {sc.n_dyn} bodies, 30 lockstep steps, then 4000 rays")
    g, o = lockstep(sc, 30, check_every=10)
    rng = np.random.default_rng(5)
    lo, hi = (-8, 14) if name == "mixed" else (-4, 5)
    origins, dirs = rays_into(rng, lo, hi, 4000)
    maxd = rng.uniform(1.0, 40.0, 4000).astype(np.float32)
    gc, gt, us = compare(g, o, origins, dirs, maxd)
    log(f"[raycast] {name}: {int((gc >= 0).sum())} hits, device {us:.0f} us")
    assert (gc >= 0).sum() > 500 and (gc < 0).sum() > 100
    # reject the entity each ray hit: the next closest hit (or none)
    ent = sc.arrays()["entity"][sc.arrays()["col_actor"]]
    ignore = np.where(gc >= 0, ent[np.maximum(gc, 0)], 0xFFFFFFFF).astype(np.uint32)
    compare(g, o, origins, dirs, maxd, ignore=ignore)
    # a maxDistance just below each hit distance loses that hit (rays starting inside a
    # shape hit at 0 and keep it)
    short = np.where(gc >= 0, np.nextafter(gt, np.float32(0)), maxd).astype(np.float32)
    c2, _, _ = compare(g, o, origins, dirs, short)
    lost = (gc >= 0) & (gt > 0)
    assert lost.sum() > 100 and not np.any(c2[lost] == gc[lost])


def test_raycast_in_batched_worlds(api):
    """Rays cast in one world of a batch equal the same rays cast in that world alone."""
    sc = scenes.batched_worlds(8, n_worlds=12, lo=8, hi=120, ragdoll_frac=0.5)
    log(f"[raycast] 12 worlds, {sc.n_dyn} bodies: 20 steps, rays per world vs each world alone")
    g = api.World(sc, substeps=4, iterations=2)
    for _ in range(20):
        g.step(1 / 60)
    arr = sc.arrays()
    col_world = arr["world"][arr["col_actor"]]
    rng = np.random.default_rng(9)
    for wd in (0, 3, 7, 11):
        # the batched world against the same world stepped alone on the device (their
        # states are bit-identical); the alone world against the oracle is the next test
        ga = api.World(sc.world_subscene(wd), substeps=4, iterations=2)
        for _ in range(20):
            ga.step(1 / 60)
        origins, dirs = rays_into(rng, -6, 12, 500)
        maxd = np.full(500, 50.0, np.float32)
        first = int(np.nonzero(col_world == wd)[0][0])
        gb, tb, _ = g.raycast(origins, dirs, maxd, world=np.full(500, wd, np.int32))
        gs, ts, _ = ga.raycast(origins, dirs, maxd)
        gs = np.where(gs >= 0, gs + first, -1)
        assert np.array_equal(gb, gs), wd
        assert np.array_equal(bits(tb[gb >= 0]), bits(ts[gs >= 0]))
        assert (gb >= 0).sum() > 20


def test_raycast_alone_world_matches_oracle(api):
    sc = scenes.batched_worlds(8, n_worlds=12, lo=8, hi=120, ragdoll_frac=0.5).world_subscene(3)
    g, o = lockstep(sc, 20, check_every=10)
    rng = np.random.default_rng(10)
    origins, dirs = rays_into(rng, -6, 12, 1500)
    compare(g, o, origins, dirs, np.full(1500, 50.0, np.float32))
