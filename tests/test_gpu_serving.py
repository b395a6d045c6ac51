"""The served form: requests admitted while others run, stepped in the same launches,
retired at their horizon. Each request's result is bit-identical to the same scene
stepped alone for its horizon (continuing requests keep their device state and contact
cache across admissions and retirements), and that alone-run is stepped in lockstep with
the oracle."""
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


def alone(api, scene, steps):
    """The request's scene stepped alone for its horizon, in lockstep with the oracle, so a
    served result equal to it equals the oracle's too."""
    w, _ = lockstep(scene, steps, check_every=5)
    return w.state()


@pytest.mark.parametrize("interval", [1, 4])
def test_served_results_same_alone_and_in_mix(api, interval):
    plan = [  # (submit at server step, seed, bodies, chain, horizon)
        (0, 11, 40, False, 30), (0, 12, 120, True, 24), (5, 13, 16, False, 20),
        (5, 14, 300, False, 12), (9, 15, 64, True, 25), (20, 16, 200, False, 10),
    ]
    srv = api.Server(substeps=4, iterations=2, admit_interval=interval)
    ids = {}
    step = 0
    pending = list(plan)
    while pending or any(not srv.done(r) for r in ids.values()):
        for item in [p for p in pending if p[0] == step]:
            sc = scenes.served_request(item[1], n_bodies=item[2], chain=item[3])
            ids[item] = srv.submit(sc, item[4])
            pending.remove(item)
        info = srv.step(1 / 60)
        log(f"[serving] interval {interval} step {step}: {info}")
        step += 1
        assert step < 200
    for item, rid in ids.items():
        got = srv.take_result(rid)
        ref = alone(api, scenes.served_request(item[1], n_bodies=item[2], chain=item[3]), item[4])
        for name, a, b in zip(("pos", "quat", "vel", "angvel"), got, ref):
            assert np.array_equal(bits(a), bits(b)), (item, name)
    srv.close()


def test_served_mechanisms_keep_gear_state_across_boundaries(api):
    """Requests with prismatic, servo and gear joints: a gear joint keeps angles across steps
    (GearJoint.h:7-14), and the server carries them over when it rebuilds the batch for an
    admission, so each result still equals the scene stepped alone."""
    plan = [(0, 21, 40), (3, 22, 30), (6, 23, 18), (11, 24, 25)]   # (submit at step, seed, horizon)
    srv = api.Server(substeps=4, iterations=2, admit_interval=2)
    ids = {}
    step = 0
    pending = list(plan)
    while pending or any(not srv.done(r) for r in ids.values()):
        for item in [p for p in pending if p[0] == step]:
            ids[item] = srv.submit(scenes.mechanism_scene(item[1]), item[2])
            pending.remove(item)
        info = srv.step(1 / 60)
        log(f"[serving] mechanisms step {step}: {info}")
        step += 1
        assert step < 200
    for item, rid in ids.items():
        got = srv.take_result(rid)
        ref = alone(api, scenes.mechanism_scene(item[1]), item[2])
        for name, a, b in zip(("pos", "quat", "vel", "angvel"), got, ref):
            assert np.array_equal(bits(a), bits(b)), (item, name)
    srv.close()


def test_server_budget_limits_admission(api):
    sc = scenes.served_request(3, n_bodies=64)
    fp = api.Server.footprint(sc)
    srv = api.Server(budget_bytes=2.5 * fp)
    ids = [srv.submit(scenes.served_request(3 + k, n_bodies=64), 5) for k in range(5)]
    info = srv.step(1 / 60)
    assert info["active"] == 2 and info["queued"] == 3
    for _ in range(30):
        srv.step(1 / 60)
    assert all(srv.done(r) for r in ids)
    srv.close()


def test_results_taken_before_their_boundary_free_the_budget(api):
    """A result taken as soon as its request finishes, before the next admission boundary
    retires the request: the boundary still retires it and returns its footprint to the
    budget, so the queue behind it is admitted and drains, and the results equal the
    requests' alone-runs. (Taking it used to leave the request's footprint counted for good.)"""
    scs = [scenes.served_request(40 + k, n_bodies=48) for k in range(6)]
    fp = max(api.Server.footprint(sc) for sc in scs)
    srv = api.Server(substeps=4, iterations=2, budget_bytes=2.5 * fp, admit_interval=8)
    horizons = [5, 9, 6, 11, 7, 5]
    ids = [srv.submit(sc, h) for sc, h in zip(scs, horizons)]
    log("[served] 6 requests, results taken at once, budget for two")
    got = {}
    for k in range(200):
        info = srv.step(1 / 60)
        for r in ids:
            if r not in got and srv.done(r):
                got[r] = srv.take_result(r)
                with pytest.raises(api.PgxError):
                    srv.done(r)          # taken: gone for the caller
        if len(got) == len(ids) and info["active"] == 0:
            break
    assert len(got) == len(ids), (len(got), info)
    info = srv.step(1 / 60)
    assert info["footprint_bytes"] == 0 and info["queued"] == 0, info
    for r, sc, h in zip(ids, scs, horizons):
        want = alone(api, sc, h)
        for x, y in zip(got[r], want):
            assert np.array_equal(bits(x), bits(y)), r
    srv.close()
