"""Trigger pairs and their enter/exit events (src/Physecs.cpp:199-206, 538-552) in the
oracle, against closed forms: a pair with a trigger collider on either side makes no
contact, so a body falls freely through a trigger volume (and a trigger body through a
solid one), and with symplectic Euler substeps its height after n substeps is
y0 - g h^2 n (n + 1) / 2. The enter event comes at the first step whose start position
overlaps, the exit at the first step whose start position no longer does. The overlap
tests themselves are pinned against the reference's Overlap.cpp elsewhere."""
import numpy as np

import oracle_lib as ol
from physecs_amd import scenes

G, DT, SUB = 9.81, 1 / 60, 4


def height(y0, step):
    """Height at the start of `step` (0-based) after free fall from rest, float64."""
    h = DT / SUB
    n = SUB * step
    return y0 - G * h * h * n * (n + 1) / 2


def first_step(pred, y0, limit=400):
    for s in range(limit):
        if pred(height(y0, s)):
            return s
    raise AssertionError("never")


def events_by_step(o, steps):
    enter, exit_ = {}, {}
    for s in range(steps):
        o.step(DT)
        en, ex = o.trigger_events()
        for p in map(tuple, en):
            enter.setdefault(p, []).append(s)
        for p in map(tuple, ex):
            exit_.setdefault(p, []).append(s)
    return enter, exit_


def test_body_falls_through_trigger_volume():
    sc = scenes.Scene()
    vol = sc.add_static(scenes.BOX, [1, 1, 1], [0, 5, 0], trigger=True)        # y in [4, 6]
    ball = sc.add_body(scenes.SPHERE, [0.25], [0.1, 10.0, 0.0])
    sc.add_static(scenes.BOX, [10, 0.5, 10], [0, -40, 0])                       # far below
    o = ol.OracleWorld(sc, substeps=SUB, iterations=2)
    enter, exit_ = events_by_step(o, 90)
    s_in = first_step(lambda y: y - 0.25 <= 6.0, 10.0)
    s_out = first_step(lambda y: y + 0.25 < 4.0, 10.0)
    for s, edge in ((s_in, 6.25), (s_out, 3.75)):   # no boundary within rounding of a step
        assert abs(height(10.0, s) - edge) > 1e-3 and abs(height(10.0, s - 1) - edge) > 1e-3
    pair = (0, 1)                                    # (collider of the lower entity, other)
    assert enter == {pair: [s_in]} and exit_ == {pair: [s_out]}
    pos = o.state()[0]
    assert abs(pos[ball][1] - height(10.0, 90)) < 1e-3   # it fell freely through the volume


def test_trigger_body_falls_through_solid_box():
    sc = scenes.Scene()
    ball = sc.add_body(scenes.SPHERE, [0.25], [0.0, 10.0, 0.3], trigger=True)   # entity 0
    box = sc.add_static(scenes.BOX, [1, 1, 1], [0, 5, 0])                       # entity 1: solid
    o = ol.OracleWorld(sc, substeps=SUB, iterations=2)
    enter, exit_ = events_by_step(o, 90)
    s_in = first_step(lambda y: y - 0.25 <= 6.0, 10.0)
    s_out = first_step(lambda y: y + 0.25 < 4.0, 10.0)
    assert enter == {(0, 1): [s_in]} and exit_ == {(0, 1): [s_out]}


def test_no_events_without_overlap_and_none_repeated():
    sc = scenes.Scene()
    sc.add_static(scenes.BOX, [1, 1, 1], [0, 5, 0], trigger=True)
    sc.add_body(scenes.SPHERE, [0.25], [3.0, 10.0, 0.0])     # falls beside the volume: bounds
    o = ol.OracleWorld(sc, substeps=SUB, iterations=2)        # never meet, no pair at all
    enter, exit_ = events_by_step(o, 90)
    assert enter == {} and exit_ == {}
    # a body resting inside a volume: one enter, then nothing while it stays
    sc = scenes.Scene()
    sc.add_static(scenes.BOX, [10, 0.5, 10], [0, -0.5, 0])
    sc.add_static(scenes.BOX, [2, 2, 2], [0, 0, 0], trigger=True)
    sc.add_body(scenes.BOX, [0.3, 0.3, 0.3], [0, 0.3, 0])
    o = ol.OracleWorld(sc, substeps=SUB, iterations=2)
    enter, exit_ = events_by_step(o, 60)
    assert enter == {(1, 2): [0]} and exit_ == {}
