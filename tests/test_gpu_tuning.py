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

# pair_search: 0 is the uniform-grid pair search (the default), 1 the walk of the LBVH
TUNINGS = [
    {"solve_block": 256, "color_batch": 1 << 30, "pairs_per_lane": 4, "color_rounds": 1, "eager_rows_max": 0,
     "pair_search": 1},
    {"solve_block": 128, "color_batch": 0, "pairs_per_lane": 2, "color_rounds": 32, "eager_rows_max": 1 << 30,
     "pair_search": 0},
    {"solve_block": 64, "color_batch": 64, "pairs_per_lane": 1, "color_rounds": 3, "eager_rows_max": 100,
     "pair_search": 1},
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
                {"eager_rows_max": -1}, {"pair_search": -1}, {"pair_search": 2}):
        with pytest.raises(api.PgxError):
            w.set_tuning(**bad)
    assert w.tuning() == start


def _giants(seed=9):
    """Small bodies in a pile and, far from it, a few dynamic boxes 40 m wide: the grid's cells
    would be as large as the boxes, its fullest cell far beyond its limit, so the step uses the
    tree walk."""
    rng = np.random.default_rng(seed)
    s = scenes.Scene()
    s.add_static(scenes.BOX, [400, 0.5, 400], [0, -0.5, 0])
    for i in range(3):
        s.add_body(scenes.BOX, [20.0, 1.0, 20.0], [300.0 * (i - 1) + (150.0 if i == 1 else 0.0), 1.2, 250.0],
                   density=10.0)
    for i in range(4000):   # about three cells of 40 m: over a thousand bodies in each
        k = int(rng.integers(0, 3))
        prm = [0.3] if k == 0 else ([0.3, 0.2] if k == 1 else [0.25, 0.3, 0.2])
        s.add_body(k, prm, [rng.uniform(-40, 40), rng.uniform(2.5, 12.0), rng.uniform(-15, 15)],
                   scenes.random_quats(rng, 1)[0])
    return s


@pytest.mark.parametrize("name", ["mixed32k", "batched", "batched_mid", "batched_large", "terrain", "giants"])
def test_pair_searches_find_the_same_pairs(api, name):
    """The broadphase's pair set from the uniform grid and from the LBVH walk is the same set,
    after a step and well into the pile, from 4k to 113k dynamic colliders; with a few giant
    bodies the grid hands its steps to the tree walk and the sets still agree."""
    from gpu_util import pair_set
    sc = {"mixed32k": lambda: scenes.mixed_lattice(0, side=32),
          "batched": lambda: scenes.batched_worlds(3, n_worlds=256, lo=8, hi=400, ragdoll_frac=0.25),
          "batched_mid": lambda: scenes.batched_worlds(5, n_worlds=300, lo=60, hi=400, ragdoll_frac=0.1),
          "batched_large": lambda: scenes.batched_worlds(4, n_worlds=600, lo=60, hi=400, ragdoll_frac=0.0),
          "terrain": lambda: scenes.convex_terrain(5, n_convex=48, n_chains=4, quads=16, amplitude=2.0,
                                                   drop=(0.3, 2.0)),
          "giants": _giants}[name]()
    log(f"[pairs] {name}: {sc.n_dyn} bodies, grid vs tree walk")
    worlds = [api.World(sc, substeps=4, iterations=2) for _ in range(2)]
    for w, mode in zip(worlds, (0, 1)):
        w.set_tuning(pair_search=mode)
    for k in range(60):
        for w in worlds:
            w.step(1 / 60)
        if k in (0, 59):
            sets = [pair_set(w.pairs()) for w in worlds]
            assert len(sets[0]) > 0
            assert sets[0] == sets[1], (name, k, [len(x) for x in sets])
            st = [x.state() for x in worlds]
            for a, b in zip(st[0], st[1]):
                assert np.array_equal(bits(a), bits(b)), (name, k)
    log(f"  {len(sets[0])} pairs after 60 steps")
