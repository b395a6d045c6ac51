"""CharacterController::move on the device (pgx_world_character_move: the MTD rounds on
the device, the pushes applied in hit order) against the oracle's move on lockstep states:
new positions bit for bit and the same collision flags, for characters walking over a
triangle-mesh terrain among bodies, and in batched worlds each in its own world."""
import numpy as np
import pytest

from gpu_util import bits, lockstep, log
from physecs_amd import scenes

pytestmark = pytest.mark.gpu
SLOPE = float(np.float32(np.cos(np.pi / 4)))


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def characters(rng, n, lo, hi):
    pos = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    yaw = rng.uniform(0, 2 * np.pi, n)
    quat = np.zeros((n, 4), np.float32)
    quat[:, 1] = np.sin(yaw / 2)
    quat[:, 3] = np.cos(yaw / 2)
    tr = (rng.normal(size=(n, 3)) * [0.2, 0.05, 0.2]).astype(np.float32)
    h = rng.uniform(0.8, 1.8, n).astype(np.float32)
    r = rng.uniform(0.2, 0.4, n).astype(np.float32)
    it = rng.integers(1, 4, n)
    return pos, quat, tr, h, r, it


@pytest.mark.parametrize("name", ["terrain", "mixed"])
def test_character_move_matches_oracle(api, name):
    if name == "terrain":
        sc, lo, hi = scenes.terrain_drop(6, quads=12, n=40), (-5, -2.5, -5), (5, 2.0, 5)
    else:
        sc, lo, hi = scenes.mixed_lattice(5, side=6), (-5, -0.5, -5), (5, 4, 5)
    log(f"[character] {name}: 20 lockstep steps, then 600 characters")
    g, o = lockstep(sc, 20, check_every=10)
    rng = np.random.default_rng(17)
    pos, quat, tr, h, r, it = characters(rng, 600, lo, hi)
    gp, gf = g.character_move(pos, quat, tr, h, r, SLOPE, it)
    moved = 0
    for i in range(600):
        op, of = o.character_move(pos[i], quat[i], tr[i], h[i], r[i], SLOPE, it[i])
        assert np.array_equal(bits(gp[i]), bits(op)), (i, gp[i], op)
        assert gf[i] == of, (i, gf[i], of)
        moved += int(of != 0)
    log(f"[character] {name}: {moved} of 600 touched something")
    assert moved > 100


def test_character_move_in_batched_worlds(api):
    sc = scenes.batched_worlds(8, n_worlds=6, lo=8, hi=80, ragdoll_frac=0.5)
    g = api.World(sc, substeps=4, iterations=2)
    for _ in range(20):
        g.step(1 / 60)
    rng = np.random.default_rng(3)
    for wd in (0, 5):
        ga, o = lockstep(sc.world_subscene(wd), 20, check_every=20)
        b = o.bounds()
        pos, quat, tr, h, r, it = characters(rng, 200, b[:, :3].min(axis=0), b[:, 3:].max(axis=0))
        gb, fb = g.character_move(pos, quat, tr, h, r, SLOPE, it, world=wd)
        gs, fs = ga.character_move(pos, quat, tr, h, r, SLOPE, it)
        assert np.array_equal(bits(gb), bits(gs)) and np.array_equal(fb, fs)
        for i in range(0, 200, 7):
            op, of = o.character_move(pos[i], quat[i], tr[i], h[i], r[i], SLOPE, it[i])
            assert np.array_equal(bits(gs[i]), bits(op)) and fs[i] == of
