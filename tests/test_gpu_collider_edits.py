"""Scene::addCollider and Scene::clearColliders on a world that has been stepping
(src/Physecs.cpp:725-752): the world is created anew from the edited scene and
pgx_world_load_state_matched carries the old world's contact cache, trigger lists and
bounds to the colliders with the same entity and collider index, the keys the reference's
caches use. Against the oracle edited in place (its caches untouched, as the reference's),
every later step is bit-identical; a world created from the same edited scene without the
carried state is not. Trigger pairs that lose a collider come back from the load and are the
oracle's exits of the next step, and a collider cleared and added back under its index keeps
its trigger pair without an event. Through the compiled drop-in Scene, the same edits give
the same bodies bit for bit and the same listener callbacks."""
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


def _bouncy(name):
    if name == "cluster":
        sc = scenes.random_cluster(2, n=60)
    else:
        sc = scenes.terrain_drop(3, quads=12, n=40)
    sc.col_material[:, 1] = 0.4    # restitution: the cached target velocities matter
    return sc


def _step(g, o):
    g.step(DT)
    keys, _, _ = g.manifolds()
    o.set_order(keys)
    o.step(DT)


def _assert_same(g, o, k):
    for name, a, b in zip(("pos", "quat", "vel", "angvel"), g.state(), o.state()):
        assert np.array_equal(bits(a), bits(b)), (k, name)


def _collider(arr, c):
    """(type, params, keyword arguments) of descriptor collider c, for adding it again."""
    kw = dict(material=tuple(arr["col_material"][c]), lpos=tuple(arr["col_pos"][c]), lquat=tuple(arr["col_quat"][c]),
              trigger=bool(arr["col_trigger"][c]), enable=bool(arr["col_enable"][c]), data=int(arr["col_data"][c]))
    return int(arr["col_type"][c]), arr["col_params"][c].copy(), kw


def _touching(g, arr, n):
    """n dynamic actors in the last step's manifolds, first seen first."""
    keys, _, _ = g.manifolds()
    dyn = arr["dyn_of_actor"]
    seen = []
    for e0, _, e1, _, _ in keys.tolist():
        for e in (e0, e1):   # entity == actor index in these scenes
            if dyn[e] >= 0 and e not in seen:
                seen.append(e)
    assert len(seen) >= n
    return seen[:n]


def _edit(api, g, o, sc, clear, readd, add):
    """The same edits on the device (a new world from Scene.rebuilt plus the carried state)
    and on the oracle (in place). Returns (new world, new scene, trigger pairs gone)."""
    arr = sc.arrays()
    again = [(a,) + _collider(arr, int(np.nonzero(arr["col_actor"] == a)[0][0])) for a in readd]
    blob = g.save_state()
    s2, _ = sc.rebuilt(g.state(), clear=list(clear) + list(readd), add=again + list(add))
    g2 = api.World(s2, substeps=4, iterations=2)
    gone = g2.load_state_matched(blob)
    for a in list(clear) + list(readd):
        o.clear_colliders(a)
    for a, t, p, kw in again + list(add):
        o.add_collider(a, t, p, **kw)
    return g2, s2, gone


@pytest.mark.parametrize("name", ["cluster", "terrain"])
def test_collider_edits_keep_the_cache(api, name):
    sc = _bouncy(name)
    pre = 30 if name == "cluster" else 90   # the terrain's bodies land later
    log(f"[edits] {name}: {sc.n_dyn} bodies, {pre} steps, clear / re-add / add, 40 steps")
    g = api.World(sc, substeps=4, iterations=2)
    o = ol.OracleWorld(sc, substeps=4, iterations=2)
    for k in range(pre):
        _step(g, o)
    _assert_same(g, o, pre)
    arr = sc.arrays()
    a, b, c = _touching(g, arr, 3)
    add = [(a, scenes.SPHERE, [0.15], dict(material=(0.6, 0.4, 0.0), lpos=(0.0, 0.25, 0.0)))]
    g2, s2, gone = _edit(api, g, o, sc, clear=[b], readd=[c], add=add)
    assert len(gone) == 0
    g3 = api.World(s2, substeps=4, iterations=2)   # the same scene without the carried state
    differs = False
    for k in range(40):
        _step(g2, o)
        g3.step(DT)
        _assert_same(g2, o, pre + k)
        differs = differs or any(not np.array_equal(bits(x), bits(y)) for x, y in zip(g2.state(), g3.state()))
    assert differs, "the carried state made no difference"
    for w in (g, g2, g3):
        w.close()


def _trigger_scene(seed=4):
    """A static trigger box over the ground that the pile settles into, so trigger pairs last."""
    rng = np.random.default_rng(seed)
    s = scenes.Scene()
    s.add_static(scenes.BOX, [10, 0.5, 10], [0, -0.5, 0])
    s.add_static(scenes.BOX, [2.5, 1.0, 2.5], [0, 1.0, 0], trigger=True)
    for i in range(48):
        k = int(rng.integers(0, 3))
        prm = [0.3] if k == 0 else ([0.3, 0.2] if k == 1 else [0.25, 0.3, 0.2])
        s.add_body(k, prm, [rng.uniform(-2, 2), rng.uniform(0.5, 5.0), rng.uniform(-2, 2)],
                   scenes.random_quats(rng, 1)[0])
    return s


def _ec(pairs, arr):
    """Descriptor collider pairs to (entity0, index0, entity1, index1) rows."""
    ent, ca, ci = arr["entity"], arr["col_actor"], arr["col_index"]
    return sorted((int(ent[ca[x]]), int(ci[x]), int(ent[ca[y]]), int(ci[y])) for x, y in pairs.tolist())


def test_trigger_pairs_across_collider_edits(api):
    sc = _trigger_scene()
    log(f"[edits] trigger box: {sc.n_dyn} bodies, 40 steps, edits on bodies inside it, 30 steps")
    g = api.World(sc, substeps=4, iterations=2)
    o = ol.OracleWorld(sc, substeps=4, iterations=2)
    inside = set()
    for k in range(40):
        _step(g, o)
        en, ex = o.trigger_events_ec()
        inside |= {tuple(r) for r in en.tolist()}
        inside -= {tuple(r) for r in ex.tolist()}
        arr = sc.arrays()
        gen, gex = g.trigger_events()
        assert _ec(gen, arr) == sorted(tuple(r) for r in en.tolist()), k
        assert _ec(gex, arr) == sorted(tuple(r) for r in ex.tolist()), k
    _assert_same(g, o, 40)
    # dynamic entities in a trigger pair (entities 0 and 1: the ground and the trigger box)
    bodies = sorted(({r[0] for r in inside} | {r[2] for r in inside}) - {0, 1})
    assert len(bodies) >= 3, inside
    b, c, a = bodies[:3]
    add = [(a, scenes.SPHERE, [0.2], dict(trigger=True, lpos=(0.0, 0.3, 0.0)))]
    g2, s2, gone = _edit(api, g, o, sc, clear=[b], readd=[c], add=add)
    want_gone = sorted(r for r in inside if r[0] == b or r[2] == b)
    assert sorted(tuple(r) for r in gone.tolist()) == want_gone
    arr2 = s2.arrays()
    events = 0
    for k in range(30):
        _step(g2, o)
        en, ex = o.trigger_events_ec()
        gen, gex = g2.trigger_events()
        got_ex = _ec(gex, arr2) + ([tuple(r) for r in gone.tolist()] if k == 0 else [])
        assert _ec(gen, arr2) == sorted(tuple(r) for r in en.tolist()), k
        assert sorted(got_ex) == sorted(tuple(r) for r in ex.tolist()), k
        if k == 0:   # no pair that was inside enters again (the re-added collider's pairs held)
            assert not any(tuple(r) in inside for r in en.tolist())
        events += len(en) + len(ex)
        _assert_same(g2, o, 40 + k)
    log(f"  {len(want_gone)} pairs gone, {events} events after the edit")
    for w in (g, g2):
        w.close()


def test_scene_collider_edits_through_the_shim(api):
    """Scene::clearColliders / addCollider on the drop-in Scene, against the C ABI path above:
    the registry's bodies bit for bit and the listener callbacks, gone pairs included."""
    import shim_lib
    if not shim_lib.available():
        pytest.skip("oracle/_ref/libshimcheck.so not built")
    sc = _trigger_scene(6)
    arr = sc.arrays()
    log(f"[edits] drop-in Scene: {sc.n_dyn} bodies, edits at step 30, 60 steps")
    g = api.World(sc, substeps=4, iterations=2)
    reg, ents = shim_lib.replay(sc)
    with reg:
        reg.listen_triggers()
        inside = set()
        for k in range(30):
            g.step(DT)
            reg.simulate(DT, 4, 2)
            ge, gx = g.trigger_events()
            inside |= set(_ec(ge, arr))
            inside -= set(_ec(gx, arr))
        reg.take_triggers()
        bodies = sorted(({r[0] for r in inside} | {r[2] for r in inside}) - {0, 1})
        assert len(bodies) >= 2, inside
        b, c = bodies[:2]
        a = next(x for x in _touching(g, arr, 3) if x not in (b, c))
        cc = int(np.nonzero(arr["col_actor"] == c)[0][0])
        t, p, kw = _collider(arr, cc)
        blob = g.save_state()
        add = [(a, scenes.SPHERE, [0.2], dict(trigger=True, lpos=(0.0, 0.3, 0.0)))]
        s2, _ = sc.rebuilt(g.state(), clear=[b, c], add=[(c, t, p, kw)] + add)
        g.close()
        g = api.World(s2, substeps=4, iterations=2)
        gone = g.load_state_matched(blob)
        reg.clear_colliders(ents[b])
        reg.clear_colliders(ents[c])
        reg.add_collider(ents[c], t, p, **kw)
        reg.add_collider(ents[a], scenes.SPHERE, [0.2], trigger=True, lpos=(0.0, 0.3, 0.0))
        arr2 = s2.arrays()
        seen = 0
        for k in range(30):
            g.step(DT)
            reg.simulate(DT, 4, 2)
            got = reg.take_triggers()
            ge, gx = g.trigger_events()
            want = [(1,) + r for r in _ec(ge, arr2)] + [(0,) + r for r in _ec(gx, arr2)]
            if k == 0:
                want += [(0,) + tuple(int(v) for v in r) for r in gone.tolist()]
            assert sorted(tuple(int(v) for v in r) for r in got.tolist()) == sorted(want), k
            seen += len(want)
            pos, quat, vel, ang = g.state()
            for d in range(s2.n_dyn):
                act = int(arr2["actor_of_dyn"][d])
                pp, qq, vv, ww = reg.read(ents[act])
                assert np.array_equal(bits(pp), bits(pos[act])), (k, d)
                assert np.array_equal(bits(qq), bits(quat[act])), (k, d)
                assert np.array_equal(bits(vv), bits(vel[d])), (k, d)
                assert np.array_equal(bits(ww), bits(ang[d])), (k, d)
        assert len(gone) > 0
        log(f"  {seen} callbacks after the edit, {len(gone)} for gone pairs")
    g.close()
