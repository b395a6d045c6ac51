"""The batched form: many independent worlds of uneven size in one set of device buffers.
A world's result must not depend on the other worlds it is stepped with: stepped in the
mix, it is bit-identical to the same world stepped alone, and that alone-run is stepped in
lockstep with the oracle, bit for bit (the oracle is fed the device's solve order)."""
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


def world_rows(scene, w):
    arr = scene.arrays()
    actors = np.nonzero(arr["world"] == w)[0]
    dyn = np.nonzero(np.isin(arr["actor_of_dyn"], actors))[0]
    return actors, dyn


@pytest.mark.parametrize("mode", ["colors", "xcd", "worlds", "pass", "xpass"])
def test_world_same_alone_and_in_mix(api, mode):
    """A world steps bit-identically alone and among 23 others, with the contact passes
    launched per colour, on per-XCD queues (worlds spread over the XCDs) or one workgroup
    per world."""
    sc = scenes.batched_worlds(3, n_worlds=24, lo=8, hi=160, ragdoll_frac=0.5)
    log(f"[batched] 24 worlds, {sc.n_dyn} bodies, {len(sc.joints)} joints: mixed run, exec {mode}")
    steps = 40
    g = api.World(sc, substeps=4, iterations=2)
    g.set_exec(mode)
    for _ in range(steps):
        g.step(1 / 60)
    pos, quat, vel, ang = g.state()
    arr = sc.arrays()
    has_joint = {int(arr["world"][a0]) for (_, a0, *_r) in sc.joints}
    picks = [0, 5, 11, 23] + sorted(has_joint)[:2]
    for w in picks:
        # alone, in lockstep with the oracle: batched == alone == oracle for the same world
        log(f"[batched] world {w} alone, lockstep with the oracle")
        sub = sc.world_subscene(w)
        ga, _ = lockstep(sub, steps, check_every=10)
        p1, q1, v1, w1 = ga.state()
        actors, dyn = world_rows(sc, w)
        assert np.array_equal(bits(pos[actors]), bits(p1)), w
        assert np.array_equal(bits(quat[actors]), bits(q1)), w
        assert np.array_equal(bits(vel[dyn]), bits(v1)), w
        assert np.array_equal(bits(ang[dyn]), bits(w1)), w
    assert ga.counts()["manifolds"] > 0


def test_ragdoll_world_lockstep_with_oracle(api):
    sc = scenes.batched_worlds(4, n_worlds=8, lo=8, hi=96, ragdoll_frac=1.0)
    sub = sc.world_subscene(2)
    log(f"[batched] ragdoll world alone vs oracle: {sub.n_dyn} bodies, {len(sub.joints)} joints")
    assert len(sub.joints) == 14
    lockstep(sub, 60)
