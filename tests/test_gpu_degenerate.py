"""Degenerate scenes the reference leaves undefined or never meets, stepped on the device: a
scene with no actors at all, one with static colliders only, a kinematic body inside a
static wall (a pair with no dynamic side), and a custom joint that declares no rows. Each
steps without error, reports no contacts where there can be none, and the bodies that do
move match the oracle bit for bit."""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import bits, log
from physecs_amd import scenes

pytestmark = pytest.mark.gpu

DT = 1 / 60


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def test_empty_scene(api):
    log("[degenerate] a scene with no actors, 5 steps")
    g = api.World(scenes.Scene(), substeps=4, iterations=2)
    for _ in range(5):
        g.step(DT)
    g.sync()
    c = g.counts()
    assert c["pairs"] == 0 and c["manifolds"] == 0 and c["points"] == 0, c
    g.close()


def test_statics_only(api):
    log("[degenerate] static colliders only, overlapping, 5 steps")
    s = scenes.Scene()
    s.add_static(scenes.BOX, [10, 0.5, 10], [0, -0.5, 0])
    s.add_static(scenes.SPHERE, [1.0], [0, 0.2, 0])          # overlaps the ground: statics never pair
    s.add_static(scenes.CAPSULE, [0.5, 0.2], [1, 0, 0])
    g = api.World(s, substeps=4, iterations=2)
    for _ in range(5):
        g.step(DT)
    g.sync()
    c = g.counts()
    assert c["pairs"] == 0 and c["manifolds"] == 0, c
    pos, quat, _, _ = g.state()
    assert np.array_equal(bits(pos), bits(s.arrays()["pos"]))
    g.close()


def test_kinematic_against_statics(api):
    """A kinematic box (never integrated, src/Physecs.cpp:446) overlapping a static wall, and a
    dynamic sphere that falls onto it: the kinematic-static pair is skipped as both
    non-dynamic (src/Physecs.cpp:147), the sphere rests on the kinematic box. Lockstep with
    the oracle; the kinematic box does not move."""
    log("[degenerate] kinematic box in a static wall, a sphere on it, 60 steps")
    s = scenes.Scene()
    s.add_static(scenes.BOX, [0.1, 1.0, 2.0], [0.3, 1.0, 0])
    kin = s.add_actor([0.0, 0.5, 0])
    im, com, invI = scenes.mass_props(scenes.BOX, [0.3, 0.3, 0.3], 100.0)
    s.add_dynamic(kin, im, com, invI, kinematic=True)
    s.add_collider(kin, scenes.BOX, [0.3, 0.3, 0.3], (0.6, 0.0, 0.0))
    s.add_body(scenes.SPHERE, [0.2], [-0.05, 1.5, 0.05])
    g = api.World(s, substeps=4, iterations=2)
    o = ol.OracleWorld(s, substeps=4, iterations=2)
    for k in range(60):
        g.step(DT)
        keys, _, _ = g.manifolds()
        o.set_order(keys)
        o.step(DT)
        for n, x, y in zip(("pos", "quat", "vel", "angvel"), g.state(), o.state()):
            assert np.array_equal(bits(x), bits(y)), (k, n)
    pos = g.state()[0]
    assert np.array_equal(bits(pos[kin]), bits(np.float32([0.0, 0.5, 0])))
    assert 0.98 < pos[kin + 1][1] < 1.05, pos[kin + 1]     # resting on the box top (0.8) + radius
    g.close()


def test_custom_joint_without_rows(api):
    """A custom joint whose layout declares no rows joins nothing (its bodies fall apart as
    free bodies) but still keeps its pair from colliding, as any joint does."""
    log("[degenerate] custom joint with no rows, 30 steps")
    s = scenes.Scene()
    a = s.add_body(scenes.SPHERE, [0.2], [0, 2.0, 0])
    b = s.add_body(scenes.SPHERE, [0.2], [0, 2.3, 0])        # overlapping spheres
    s.add_joint(scenes.CUSTOM, a, b, (0, 0, 0), (0, 0, 0))
    g = api.World(s, substeps=4, iterations=2)
    calls = []
    g.set_joint_callbacks(lambda j, ws: calls.append(j) or [], lambda j, ws, r: None)
    for _ in range(30):
        g.step(DT)
    g.sync()
    assert len(calls) == 30
    assert g.counts()["manifolds"] == 0 and g.counts()["joint_rows"] == 0
    pos, _, vel, _ = g.state()
    assert np.allclose(pos[1, 1] - pos[0, 1], 0.3, atol=1e-5)   # both fall together, untouched
    g.close()
