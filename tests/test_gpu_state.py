"""Saving and resuming a world's state (pgx_world_save_state / pgx_world_load_state): a
world resumed from a saved state, in a fresh world of the same scene or in the world that
saved it, steps on bit for bit as the world that saved it did. The scenes exercise every
piece of carried state: restitution (the contact cache's target velocities), gear joints
(their angle state), trigger colliders (the trigger lists and their events). States from a
world of another shape, truncated or foreign bytes are rejected."""
import numpy as np
import pytest

from gpu_util import bits, log
from physecs_amd import scenes

pytestmark = pytest.mark.gpu
DT = 1 / 60


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def make(name):
    if name == "restitution":
        sc = scenes.mixed_lattice(2, side=8)
        sc.arrays()["col_material"][:, 1] = 0.5
        return sc
    if name == "mechanism":
        return scenes.mechanism_scene(1)
    sc = scenes.mixed_lattice(3, side=8)
    sc.add_static(scenes.BOX, [8, 2, 8], [0, 2, 0], trigger=True)   # a volume the pile rests in
    sc.arrays()["col_trigger"][5::5] = 1                               # the container stays solid
    return sc


def same(a, b):
    return all(np.array_equal(bits(x), bits(y)) for x, y in zip(a, b))


@pytest.mark.parametrize("name", ["restitution", "mechanism", "triggers"])
def test_resume_matches_uninterrupted(api, name):
    sc = make(name)
    log(f"[state] {name}: 30 steps, save, 20 more; resume in a fresh world and in the same world")
    a = api.World(sc, substeps=4, iterations=2)
    for _ in range(30):
        a.step(DT)
    blob = a.save_state()
    events_at_save = a.trigger_events()
    state_at_save = a.state()
    for _ in range(20):
        a.step(DT)
    want_state, want_keys, want_events = a.state(), a.manifolds()[0], a.trigger_events()
    b = api.World(sc, substeps=4, iterations=2)
    b.step(DT)                                            # some other state first
    b.load_state(blob)
    assert all(np.array_equal(x, y) for x, y in zip(b.trigger_events(), events_at_save))
    for k in range(20):
        b.step(DT)
        if k == 0:
            resumed_first = b.trigger_events()
    assert same(b.state(), want_state)
    assert np.array_equal(b.manifolds()[0], want_keys)
    assert all(np.array_equal(x, y) for x, y in zip(b.trigger_events(), want_events))
    a.load_state(blob)                                    # the saving world, back in time
    for _ in range(20):
        a.step(DT)
    assert same(a.state(), want_state)
    # poses and velocities alone do not resume the world: the carried state matters
    c = api.World(sc, substeps=4, iterations=2)
    c.write_state(*state_at_save)
    for k in range(20):
        c.step(DT)
        if k == 0 and name == "triggers":   # every lasting overlap enters again
            assert len(c.trigger_events()[0]) > len(resumed_first[0]) + 20
    if name != "triggers":
        assert not same(c.state(), want_state)
    log(f"[state] {name}: {len(blob)} bytes")


def test_load_rejects_foreign_states(api):
    a = api.World(scenes.mixed_lattice(2, side=6), substeps=4, iterations=2)
    a.step(DT)
    blob = a.save_state()
    other = api.World(scenes.mixed_lattice(2, side=7), substeps=4, iterations=2)
    before = other.state()
    with pytest.raises(api.PgxError, match="INVALID"):
        other.load_state(blob)                            # another shape
    with pytest.raises(api.PgxError, match="INVALID"):
        a.load_state(blob[: len(blob) // 2])              # truncated
    with pytest.raises(api.PgxError, match="INVALID"):
        a.load_state(b"\0" * len(blob))                   # not a state
    assert same(other.state(), before)


def test_counts_and_manifolds_right_after_load(api):
    """A load leaves no last step to report: counts and manifolds read empty until the next
    step, whatever the loading world's own last step held, and the next step is the saver's."""
    sc = scenes.mixed_lattice(4, side=8)
    a = api.World(sc, substeps=4, iterations=2)
    for _ in range(10):
        a.step(DT)
    blob = a.save_state()
    a.step(DT)
    want = a.state()
    b = api.World(sc, substeps=4, iterations=2)
    for _ in range(40):                                   # a fuller contact list than the saver's
        b.step(DT)
    assert b.counts()["manifolds"] > 0
    b.load_state(blob)
    cnt = b.counts()
    assert cnt["pairs"] == 0 and cnt["manifolds"] == 0 and cnt["points"] == 0, cnt
    keys, ints, f = b.manifolds()
    assert len(keys) == 0 and len(ints) == 0 and len(f) == 0
    b.step(DT)
    assert same(b.state(), want)


def test_load_from_a_world_with_larger_capacity(api):
    """A state saved by a world with a larger pair (and so manifold) capacity loads where its
    contact cache fits, and is rejected, with nothing changed, where it does not."""
    sc = scenes.mixed_lattice(4, side=8)
    big = api.World(sc, substeps=4, iterations=2, pair_capacity=1 << 16)
    for _ in range(60):
        big.step(DT)
    blob = big.save_state()
    n_cache = big.counts()["manifolds"]
    big.step(DT)
    want = big.state()
    fits = api.World(sc, substeps=4, iterations=2, pair_capacity=max(4 * n_cache, 4096))
    fits.load_state(blob)
    fits.step(DT)
    assert same(fits.state(), want)
    small = api.World(sc, substeps=4, iterations=2, pair_capacity=max(1, n_cache // 4))
    before = small.state()
    with pytest.raises(api.PgxError, match="INVALID"):
        small.load_state(blob)
    assert same(small.state(), before)
