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
