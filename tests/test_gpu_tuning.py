"""Launch shapes (pgx_tuning) and the autotuner. Every tuning must give the same results
bit for bit: the contact solve order stays the colour order whether colours launch one by
one or a run of small colours is stepped through by one workgroup, and the narrowphase and
colouring results do not depend on pairs per lane or rounds per read-back. The autotuner
steps the world from a snapshot and must leave it exactly where it was."""
import numpy as np
import pytest

from gpu_util import bits, lockstep, log
from physecs_amd import scenes

pytestmark = pytest.mark.gpu

TUNINGS = [
    {"solve_block": 256, "color_batch": 1 << 30, "pairs_per_lane": 4, "color_rounds": 1, "eager_rows_max": 0},
    {"solve_block": 128, "color_batch": 512, "pairs_per_lane": 2, "color_rounds": 32, "eager_rows_max": 1 << 30},
    {"solve_block": 64, "color_batch": 64, "pairs_per_lane": 1, "color_rounds": 3, "eager_rows_max": 100},
]


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def _scene(name):
    if name == "mixed":
        return scenes.mixed_lattice(4, side=10)
    if name == "batched":
        return scenes.batched_worlds(6, n_worlds=16, lo=8, hi=120, ragdoll_frac=0.5)
    return scenes.mechanism_scene(2)


def _run(api, sc, steps, tuning=None):
    w = api.World(sc, substeps=4, iterations=2)
    if tuning:
        w.set_tuning(**tuning)
    out = []
    for _ in range(steps):
        w.step(1 / 60)
        out.append(w.state())
    return w, out


@pytest.mark.parametrize("name", ["mixed", "batched", "mechanism"])
def test_every_tuning_gives_the_same_results(api, name):
    sc = _scene(name)
    log(f"[tuning] {name}: {sc.n_dyn} bodies, 40 steps x {1 + len(TUNINGS)} tunings")
    _, ref = _run(api, sc, 40)
    for t in TUNINGS:
        w, got = _run(api, sc, 40, t)
        assert w.tuning() == t
        for k, (a, b) in enumerate(zip(ref, got)):
            for x, y in zip(a, b):
                assert np.array_equal(bits(x), bits(y)), (name, t, k)


def test_batched_colours_match_the_oracle():
    """Every run of two or more colours stepped through by one workgroup, against the
    oracle fed the device's solve order."""
    sc = scenes.box_pyramid(1, rows=10)
    log("[tuning] pyramid 10 rows, all colours batched, lockstep with the oracle")
    lockstep(sc, 40, tuning={"color_batch": 1 << 30, "eager_rows_max": 0})


def test_autotune_restores_the_world(api):
    sc = scenes.mixed_lattice(5, side=8)
    log(f"[tuning] autotune mid-run: {sc.n_dyn} bodies")
    a = api.World(sc, substeps=4, iterations=2)
    b = api.World(sc, substeps=4, iterations=2)
    for _ in range(15):
        a.step(1 / 60)
        b.step(1 / 60)
    before = b.state()
    chosen = b.autotune(1 / 60, trial_steps=2)
    log(f"[tuning] chosen {chosen}")
    assert chosen == b.tuning()
    assert chosen["solve_block"] in (64, 128, 256) and chosen["pairs_per_lane"] in (1, 2, 4)
    for x, y in zip(before, b.state()):
        assert np.array_equal(bits(x), bits(y))
    for k in range(15):
        a.step(1 / 60)
        b.step(1 / 60)
        for x, y in zip(a.state(), b.state()):
            assert np.array_equal(bits(x), bits(y)), k


def test_config_block_size_sets_solve_block(api):
    """pgx_config.block_size is the world's first solve_block; 0 keeps the default 64. Both
    worlds step to the same state."""
    import ctypes
    sc = scenes.box_pyramid(0, rows=4)
    states, blocks = [], []
    for bs in (0, 128):
        desc = sc.struct(api.PgxSceneDesc)
        desc.actor_world = ctypes.cast(None, api.I32)
        h = ctypes.c_void_p()
        cfg = api.PgxConfig(4, 2, 9.81, 0, 0, bs)
        api.check(api.lib().pgx_world_create(ctypes.byref(cfg), ctypes.byref(desc), ctypes.byref(h)))
        try:
            t = api.PgxTuning()
            api.check(api.lib().pgx_world_get_tuning(h, ctypes.byref(t)))
            blocks.append(t.solve_block)
            for _ in range(30):
                api.check(api.lib().pgx_world_step(h, 1 / 60))
            pos = np.zeros((sc.n_actors, 3), np.float32)
            api.check(api.lib().pgx_world_read_state(h, pos.ctypes.data_as(api.F), None, None, None))
            states.append(pos)
        finally:
            api.lib().pgx_world_destroy(h)
    assert blocks == [64, 128]
    assert np.array_equal(states[0].view(np.uint32), states[1].view(np.uint32))


def test_tuning_rejects_bad_values(api):
    w = api.World(scenes.box_pyramid(0, rows=3), substeps=4, iterations=2)
    start = w.tuning()
    for bad in ({"solve_block": 96}, {"pairs_per_lane": 3}, {"color_rounds": 0}, {"color_batch": -1},
                {"eager_rows_max": -1}):
        with pytest.raises(api.PgxError):
            w.set_tuning(**bad)
    assert w.tuning() == start
