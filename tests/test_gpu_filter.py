"""Scene::setContactFilter on the device (pgx_world_set_contact_filter) against the oracle
with the same filter: a mixed cluster whose colliders carry data 0, 1 or 2, some of them
triggers, under an asymmetric filter (argument 0 is the lower entity's collider), stepped
in lockstep bit for bit, with identical trigger events after every step; and the sphere
that falls through a floor its filter makes a trigger pair, as the oracle's closed form
says."""
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


def _filter(t0, d0, t1, d1):
    # asymmetric on purpose: (1, 2) is a trigger pair, (2, 1) is not
    return bool(t0 or t1) or (d0 == 1 and d1 == 2) or (d0 == 0 and d1 == 0)


def test_filtered_cluster_lockstep_with_events(api):
    sc = scenes.random_cluster(7, n=120, extent=2.5)
    arr = sc.arrays()
    arr["col_data"][:] = (np.arange(len(arr["col_data"])) * 7) % 3
    arr["col_trigger"][3::11] = 1
    log(f"[filter] cluster of {sc.n_dyn} bodies, asymmetric filter, 80 steps")

    def check_events(k, g, o):
        ge, oe = g.trigger_events(), o.trigger_events()
        for a, b in zip(ge, oe):
            assert np.array_equal(a, b), (k, a, b)

    g = api.World(sc, substeps=4, iterations=2)
    o = ol.OracleWorld(sc, substeps=4, iterations=2)
    g.set_contact_filter(_filter)
    o.set_contact_filter(_filter)
    entered = 0
    for k in range(80):
        g.step(1 / 60)
        keys, _, _ = g.manifolds()
        o.set_order(keys)
        o.step(1 / 60)
        check_events(k, g, o)
        entered += len(g.trigger_events()[0])
        for a, b in zip(g.state(), o.state()):
            assert np.array_equal(bits(a), bits(b)), k
    assert entered > 10


def test_filter_lets_a_body_fall_through(api):
    sc = scenes.Scene()
    sc.add_static(scenes.BOX, [10, 0.5, 10], [0, -0.5, 0], data=2)
    sc.add_body(scenes.SPHERE, [0.5], [0, 0.6, 0], data=1)
    g = api.World(sc, substeps=4, iterations=2)
    g.set_contact_filter(lambda t0, d0, t1, d1: d0 == 2 and d1 == 1)   # the floor is the lower entity
    for _ in range(90):
        g.step(1 / 60)
    h, n = 1 / 240, 360
    assert abs(g.state()[0][1][1] - (0.6 - 9.81 * h * h * n * (n + 1) / 2)) < 1e-3
    g.set_contact_filter(None)      # back to the default filter: the next contact is solid
    w2 = api.World(sc, substeps=4, iterations=2)
    w2.set_contact_filter(lambda t0, d0, t1, d1: d0 == 1 and d1 == 2)  # reverse order: no trigger
    for _ in range(90):
        w2.step(1 / 60)
    assert abs(w2.state()[0][1][1] - 0.5) < 0.01
