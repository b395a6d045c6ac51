"""Trigger pairs on the device (k_triggers, pgx_world_get_trigger_events) against the
oracle stepped in lockstep: after every step the enter and exit events must be the same
collider pairs, and the states stay bit-identical (a trigger pair makes no contact). Scenes:
a body falling through a trigger volume, a trigger body falling through a solid box, a
mixed lattice and a convex pile with some colliders flagged as triggers (sphere, capsule,
box and GJK overlap tests), and events preserved across the autotuner's trial steps."""
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


def run(sc, steps):
    seen = {"enter": 0, "exit": 0}

    def check(k, g, o):
        ge, gx = g.trigger_events()
        oe, ox = o.trigger_events()
        assert np.array_equal(ge, oe), (k, ge[:5], oe[:5])
        assert np.array_equal(gx, ox), (k, gx[:5], ox[:5])
        seen["enter"] += len(ge)
        seen["exit"] += len(gx)

    lockstep(sc, steps, check_every=10, on_step=check)
    return seen


def test_falling_through_volumes(api):
    sc = scenes.Scene()
    sc.add_static(scenes.BOX, [1, 1, 1], [0, 5, 0], trigger=True)
    sc.add_body(scenes.SPHERE, [0.25], [0.1, 10.0, 0.0])
    sc.add_body(scenes.CAPSULE, [0.3, 0.2], [4.0, 10.0, 0.3], trigger=True)
    sc.add_static(scenes.BOX, [1, 1, 1], [4, 5, 0])
    sc.add_static(scenes.BOX, [10, 0.5, 10], [0, -40, 0])
    log("[triggers] falling through volumes: 90 lockstep steps")
    seen = run(sc, 90)
    assert seen == {"enter": 2, "exit": 2}


@pytest.mark.parametrize("name", ["mixed", "convex"])
def test_trigger_events_match_oracle(api, name):
    sc = scenes.mixed_lattice(4, side=8) if name == "mixed" else scenes.convex_pile(5, n=60)
    arr = sc.arrays()
    rng = np.random.default_rng(8)
    arr["col_trigger"][:] = (rng.random(len(arr["col_trigger"])) < 0.15).astype(arr["col_trigger"].dtype)
    log(f"[triggers] {name}: {sc.n_dyn} bodies, {int(arr['col_trigger'].sum())} trigger colliders, 60 steps")
    seen = run(sc, 60)
    log(f"[triggers] {name}: {seen}")
    assert seen["enter"] > 20 and seen["exit"] > 5


def test_trigger_events_survive_autotune(api):
    sc = scenes.mixed_lattice(4, side=8)
    arr = sc.arrays()
    arr["col_trigger"][::7] = 1
    g = api.World(sc, substeps=4, iterations=2)
    for _ in range(30):
        g.step(1 / 60)
    before = g.trigger_events()
    state = g.state()
    g.autotune(1 / 60, 2)
    after = g.trigger_events()
    assert all(np.array_equal(a, b) for a, b in zip(before, after))
    assert all(np.array_equal(a, b) for a, b in zip(state, g.state()))
