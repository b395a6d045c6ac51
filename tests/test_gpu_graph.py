"""The substep loop launched as one HIP graph per step. The loop's launches follow the
step's colours, so the graph is captured every step and its executable form is updated in
place (same launch sequence shape) or instantiated anew (the colours changed); both
happen in these runs. After every step the state must equal, bit for bit, that of the same
scene launched directly, with the program's own eager work (state read-backs, a joint
setter) and synchronisations between the graph launches. Served requests give the same
results with the graph on."""
import numpy as np
import pytest

from gpu_util import bits, log
from physecs_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def _scene(name):
    if name == "mixed":
        return scenes.mixed_lattice(2, side=10)
    if name == "batched":
        return scenes.batched_worlds(5, n_worlds=16, lo=8, hi=120, ragdoll_frac=0.5)
    return scenes.mechanism_scene(1)


@pytest.mark.parametrize("name", ["mixed", "batched", "mechanism"])
def test_graph_launch_matches_direct_launch(api, name):
    sc = _scene(name)
    log(f"[graph] {name}: {sc.n_dyn} bodies, {len(sc.joints)} joints, 60 steps graph vs direct")
    direct = api.World(sc, substeps=4, iterations=2)
    graph = api.World(sc, substeps=4, iterations=2)
    graph.set_graph(True)
    colours = set()
    for k in range(60):
        direct.step(1 / 60)
        graph.step(1 / 60)
        if name == "mechanism" and k == 30:
            servo = [i for i, j in enumerate(sc.joints) if j[0] == scenes.SERVO][0]
            for w in (direct, graph):
                w.set_joint_params(servo, target=0.3)
        a, b = direct.state(), graph.state()
        for x, y in zip(a, b):
            assert np.array_equal(bits(x), bits(y)), (name, k)
        colours.add(graph.counts()["colors"])
    log(f"[graph] {name}: colour counts seen {sorted(colours)}")
    assert len(colours) >= 1


def test_graph_served_requests_match(api):
    """A served request's result is the same with the batched world's substep loop as a
    graph (admission boundaries change the world between graph launches)."""
    pool = [scenes.served_request(2000 + k, n_bodies=int(n)) for k, n in enumerate([30, 200, 75, 400, 16, 120])]
    results = []
    for use_graph in (False, True):
        srv = api.Server(substeps=4, iterations=2, admit_interval=3)
        srv.set_graph(use_graph)
        ids = [srv.submit(sc, 20 + 7 * k) for k, sc in enumerate(pool)]
        out = {}
        for _ in range(200):
            srv.step(1 / 60)
            for i in ids:
                if i not in out and srv.done(i):
                    out[i] = srv.take_result(i)
            if len(out) == len(ids):
                break
        srv.close()
        assert len(out) == len(ids)
        results.append([out[i] for i in ids])
    for a, b in zip(*results):
        for x, y in zip(a, b):
            assert np.array_equal(bits(x), bits(y))
