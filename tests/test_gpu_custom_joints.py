"""User joint types (Joint subclasses, include/Physecs/Joint.h:42-70) on the device: the step
hands each custom joint's world-space data to the host, calls getSolverDesc's layout function
once per step and MakeConstraintsFunc's rows function once per substep (src/Physecs.cpp:
321-355,430-439), and solves the rows it gets back with the built-in joints' kernels.

- built-in joints restated as custom joints (tests/custom_joint_rows.py) step bit for bit as
  the built-in ones on the device, through the island launches, the per-colour launches (a
  chain with more than 64 joints of one colour) and the scalar overflow colour (a hub with
  ten joints);
- ropes (a layout that changes from step to step) and springs step bit for bit as the oracle
  with the same functions, and the functions are called in the same order;
- a refused layout fails the step; a world with custom joints and the graph option on steps
  as without it."""
import numpy as np
import pytest

import custom_joint_rows as cj
import oracle_lib as ol
from gpu_util import bits, log
from physecs_amd import scenes

pytestmark = pytest.mark.gpu

DT = 1 / 60
BUILTIN = (scenes.SPHERICAL, scenes.UNIVERSAL, scenes.REVOLUTE, scenes.FIXED, scenes.PRISMATIC)


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def _scene(name):
    if name == "chains":
        s = scenes.joint_scene(seed=1, chains=4, ragdolls=2)
        scenes.capsule_chain(s, [2.0, 3.0, 7.0], n=5, drive=(2.0, 50.0))
        cj.fixed_pairs(s, [-4.0, 1.0, 8.0])
        return s
    if name == "hub":          # ten joints on one body: two in the scalar overflow colour
        return cj.hub_scene()
    if name == "long_chain":   # 139 revolute joints in one island: colours 0 and 1 hold ~70 each
        s = scenes.Scene()
        s.add_static(scenes.BOX, [40, 0.5, 40], [0, -0.5, 0])
        scenes.capsule_chain(s, [-20.0, 2.0, 0.0], n=140, half_height=0.1, radius=0.05)
        return s
    return scenes.mechanism_scene(seed=2)


def _same(a, b, what):
    for n, x, y in zip(("pos", "quat", "vel", "angvel"), a, b):
        assert np.array_equal(bits(x), bits(y)), (what, n)


@pytest.mark.parametrize("name", ["chains", "mechanism", "long_chain", "hub"])
def test_restated_builtins_match_the_builtins(api, name):
    ref = _scene(name)
    cus, fns = cj.customised(_scene(name), BUILTIN)
    log(f"[custom joints] {name}: {len(fns.kinds)} joints restated, 60 steps")
    g0 = api.World(ref, substeps=4, iterations=2)
    g1 = api.World(cus, substeps=4, iterations=2)
    g1.set_joint_callbacks(fns.layout, fns.rows)
    for k in range(60):
        g0.step(DT)
        g1.step(DT)
        _same(g0.state(), g1.state(), (name, k))
    if name != "mechanism":   # a built-in prismatic joint's limit row is decided on the device, uncounted
        assert g1.counts()["joint_rows"] == g0.counts()["joint_rows"]
    for w in (g0, g1):
        w.close()


@pytest.mark.parametrize("name", ["ropes", "chains"])
def test_custom_joints_lockstep_with_the_oracle(api, name):
    if name == "ropes":
        s, fg = cj.rope_scene()
        _, fo = cj.rope_scene()
    else:
        s, fg = cj.customised(_scene("chains"), BUILTIN)
        _, fo = cj.customised(_scene("chains"), BUILTIN)
    log(f"[custom joints] {name} vs the oracle, 90 steps")
    g = api.World(s, substeps=4, iterations=2)
    o = ol.OracleWorld(s, substeps=4, iterations=2)
    g.set_joint_callbacks(fg.layout, fg.rows)
    o.set_joint_callbacks(fo.layout, fo.rows)
    for k in range(90):
        g.step(DT)
        keys, _, _ = g.manifolds()
        o.set_order(keys)
        o.step(DT)
        _same(g.state(), o.state(), (name, k))
    assert o.joint_error() == 0
    assert fg.calls == fo.calls
    if name == "ropes":
        assert fg.rope_layouts == {True, False}
    g.close()


@pytest.mark.parametrize("bad", [[0] * 8, [0, 5]])
def test_refused_layout_fails_the_step(api, bad):
    s = scenes.Scene()
    s.add_static(scenes.BOX, [5, 0.5, 5], [0, -0.5, 0])
    a = s.add_body(scenes.BOX, [0.2, 0.2, 0.2], [0, 2, 0])
    b = s.add_body(scenes.BOX, [0.2, 0.2, 0.2], [0, 2.5, 0])
    s.add_joint(scenes.CUSTOM, a, b, (0, 0.25, 0), (0, -0.25, 0))
    g = api.World(s, substeps=2, iterations=2)
    with pytest.raises(api.PgxError, match="PGX_ERR_INVALID"):
        g.step(DT)   # no functions yet
    g.close()
    g = api.World(s, substeps=2, iterations=2)
    g.set_joint_callbacks(lambda j, ws: bad, lambda j, ws, r: None)
    with pytest.raises(api.PgxError, match="PGX_ERR_INVALID"):
        g.step(DT)
    g.close()


def test_graph_option_with_custom_joints(api):
    s, f0 = cj.rope_scene(5)
    _, f1 = cj.rope_scene(5)
    log("[custom joints] ropes with the graph option on, 40 steps")
    g0 = api.World(s, substeps=4, iterations=2)
    g1 = api.World(s, substeps=4, iterations=2)
    g0.set_joint_callbacks(f0.layout, f0.rows)
    g1.set_joint_callbacks(f1.layout, f1.rows)
    g1.set_graph(True)
    for k in range(40):
        g0.step(DT)
        g1.step(DT)
        _same(g0.state(), g1.state(), k)
    for w in (g0, g1):
        w.close()
