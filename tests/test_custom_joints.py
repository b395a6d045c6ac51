"""User joint types (Joint subclasses, include/Physecs/Joint.h:42-70) on the oracle: the
layout function is getSolverDesc (called for every joint once per step, colour by colour and
in creation order inside a colour, src/Physecs.cpp:321-355) and the rows function is
MakeConstraintsFunc (every substep, src/Physecs.cpp:430-439). Built-in joints restated as
such user joints (tests/custom_joint_rows.py) step bit for bit as the built-in ones, which
are pinned against the reference's compiled row builders; a restatement with one sign
flipped does not, so the comparison sees the rows. Layouts the reference's containers cannot
hold are refused."""
import numpy as np
import pytest

import custom_joint_rows as cj
import oracle_lib as ol
from physecs_amd import scenes

DT = 1 / 60
BUILTIN = (scenes.SPHERICAL, scenes.UNIVERSAL, scenes.REVOLUTE, scenes.FIXED, scenes.PRISMATIC)


def _bits(x):
    return np.ascontiguousarray(x, np.float32).view(np.uint32)


def _joint_test_scene(name):
    if name == "chains":
        s = scenes.joint_scene(seed=1, chains=4, ragdolls=2)
        scenes.capsule_chain(s, [2.0, 3.0, 7.0], n=5, drive=(2.0, 50.0))
        cj.fixed_pairs(s, [-4.0, 1.0, 8.0])
        return s
    return scenes.mechanism_scene(seed=2)


def _run(scene, steps, fns=None):
    o = ol.OracleWorld(scene, substeps=4, iterations=2)
    if fns is not None:
        o.set_joint_callbacks(fns.layout, fns.rows)
    out = []
    for _ in range(steps):
        o.step(DT)
        out.append(o.state())
    return o, out


@pytest.mark.parametrize("name", ["chains", "mechanism"])
def test_restated_builtin_joints_step_bit_for_bit(name):
    ref = _joint_test_scene(name)
    cus, fns = cj.customised(_joint_test_scene(name), BUILTIN)
    assert len(fns.kinds) >= 10
    _, a = _run(ref, 60)
    o, b = _run(cus, 60, fns)
    assert o.joint_error() == 0
    for k, (x, y) in enumerate(zip(a, b)):
        for n, u, v in zip(("pos", "quat", "vel", "angvel"), x, y):
            assert np.array_equal(_bits(u), _bits(v)), (name, k, n)
    moved = np.abs(a[-1][0] - a[0][0]).max()
    assert moved > 0.1, moved


def test_a_flipped_sign_is_seen():
    ref = _joint_test_scene("chains")
    cus, fns = cj.customised(_joint_test_scene("chains"), BUILTIN)
    rows = fns.rows

    def flipped(joint, ws, r):
        rows(joint, ws, r)
        if fns.kinds[joint][0] == scenes.SPHERICAL:
            r[1, 10] = -r[1, 10]    # c of the y row
    fns.rows = flipped
    _, a = _run(ref, 30)
    _, b = _run(cus, 30, fns)
    assert any(not np.array_equal(_bits(x[0]), _bits(y[0])) for x, y in zip(a, b))


def test_call_order_and_counts():
    """Layouts once per step and rows once per substep, colour by colour, creation order
    inside a colour."""
    cus, fns = cj.customised(_joint_test_scene("chains"), BUILTIN)
    o, _ = _run(cus, 3, fns)
    colors = o.joint_colors()
    order = sorted(fns.kinds, key=lambda j: (colors[j], j))
    assert len(set(colors[j] for j in order)) >= 2
    want = []
    for _ in range(3):
        want += [(0, j) for j in order]
        for _ in range(4):
            want += [(1, j) for j in order]
    assert fns.calls == want


def test_ropes_and_springs_move_and_hold():
    """Joints the reference does not ship: a rope (its limit row only while taut) and a spring.
    Bodies stay within their rope lengths (plus a small solver error) once settled."""
    s, fns = cj.rope_scene()
    o, out = _run(s, 120, fns)
    assert o.joint_error() == 0
    assert fns.rope_layouts == {True, False}   # ropes went slack and taut: layouts changed between steps
    assert any(k == cj.ROPE for k, _ in fns.kinds.values()) and any(k == cj.SPRING for k, _ in fns.kinds.values())
    # chained bodies hang below their anchors
    pos = out[-1][0]
    arr = s.arrays()
    for j, (kind, p) in fns.kinds.items():
        _, a0, a1, _, _, _ = s.joints[j]
        if kind == cj.ROPE and arr["dyn_of_actor"][a0] < 0:
            d = np.linalg.norm(pos[a1] - pos[a0])
            assert d < p["length"] + 0.2, (j, d, p["length"])


@pytest.mark.parametrize("bad", [[0] * 8, [0, 5], [7]])
def test_refused_layouts(bad):
    s = scenes.Scene()
    s.add_static(scenes.BOX, [5, 0.5, 5], [0, -0.5, 0])
    a = s.add_body(scenes.BOX, [0.2, 0.2, 0.2], [0, 2, 0])
    b = s.add_body(scenes.BOX, [0.2, 0.2, 0.2], [0, 2.5, 0])
    s.add_joint(scenes.CUSTOM, a, b, (0, 0.25, 0), (0, -0.25, 0))
    o = ol.OracleWorld(s, substeps=2, iterations=2)
    o.set_joint_callbacks(lambda j, ws: bad, lambda j, ws, r: None)
    o.step(DT)
    assert o.joint_error() == 1


def test_overflow_colour_joints():
    """Ten joints on one hub: joints 9 and 10 take the scalar overflow colour (8), as custom
    joints as well as built-in ones."""
    ref = cj.hub_scene()
    cus, fns = cj.customised(cj.hub_scene(), BUILTIN)
    o0, a = _run(ref, 40)
    o1, b = _run(cus, 40, fns)
    assert 8 in o1.joint_colors()
    for k, (x, y) in enumerate(zip(a, b)):
        for u, v in zip(x, y):
            assert np.array_equal(_bits(u), _bits(v)), k
