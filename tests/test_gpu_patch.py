"""pgx_world_patch_transforms (registry.patch<TransformComponent>: the pose, and the
moved entities' collider bounds with the 0.01 margin) against the oracle in lockstep:
dynamic bodies teleported and a static wall moved mid-run; the bounds right after the
patch and every later state bit for bit."""
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


def test_patch_matches_oracle(api):
    sc = scenes.mixed_lattice(9, side=7)
    arr = sc.arrays()
    dyn = np.nonzero(arr["dyn_of_actor"] >= 0)[0]
    rng = np.random.default_rng(2)
    movers = rng.choice(dyn, 12, replace=False)
    statics = np.nonzero(arr["dyn_of_actor"] < 0)[0]
    log(f"[patch] {len(dyn)} bodies: 30 lockstep steps, 12 bodies and a container wall patched at step 10")

    def on_step(k, g, o):
        if k != 9:
            return
        actors = np.concatenate([movers, statics[1:2]]).astype(np.int32)
        pos = g.state()[0][actors].copy()
        pos[:-1, 1] += 3.0                         # bodies lifted by 3 m
        pos[-1, 0] += 0.5                          # the wall moves in by 0.5 m
        quat = g.state()[1][actors].copy()
        quat[0] = [0.0, np.sin(0.3), 0.0, np.cos(0.3)]
        g.patch_transforms(actors, pos, quat)
        for a, p, q in zip(actors, pos, quat):
            o.patch_transform(a, p, q)
        assert np.array_equal(bits(g.bounds()), bits(o.bounds()))

    lockstep(sc, 30, check_every=5, on_step=on_step)


def test_set_kinematic_by_rebuild_carries_state(api):
    """Scene::setIsKinematic as a shim does it over this boundary: the descriptor's flag,
    a new world of the same shape, the saved state loaded into it and the body's velocities
    zeroed; then in lockstep with the oracle's setIsKinematic."""
    sc = scenes.mixed_lattice(11, side=6)
    arr = sc.arrays()
    dyn_actors = np.nonzero(arr["dyn_of_actor"] >= 0)[0]
    targets = dyn_actors[[3, 40, 100, 150]]
    log("[kinematic] 20 lockstep steps, 4 bodies made kinematic by a rebuild with the state carried, 20 more")
    g, o = lockstep(sc, 20, check_every=10)
    blob = g.save_state()
    d = arr["dyn_of_actor"][targets]
    arr["kinematic"][d] = 1
    g2 = api.World(sc, substeps=4, iterations=2)
    g2.load_state(blob)
    pos, quat, vel, ang = g2.state()
    vel[d] = 0
    ang[d] = 0
    g2.write_state(pos, quat, vel, ang)
    for a in targets:
        o.set_kinematic(a, True)
    for k in range(20):
        g2.step(1 / 60)
        o.set_order(g2.manifolds()[0])
        o.step(1 / 60)
    for x, y in zip(g2.state(), o.state()):
        assert np.array_equal(bits(x), bits(y))
    assert np.array_equal(bits(g2.state()[0][targets]), bits(pos[targets]))   # they stayed put
