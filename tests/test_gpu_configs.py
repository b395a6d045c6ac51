"""Whole workloads of each configuration class stepped in lockstep with the oracle, bit for
bit at every step (the oracle fed the device's solve order):

- convex hulls and capsule chains dropped on an fBm terrain mesh (the convex + triangle-mesh
  workload, at 48 hulls, 4 chains and 16 x 16 quads);
- a served request's scene (log-uniform bodies of the mixed shapes on a ground box, with an
  8-capsule revolute chain);
- the 32,768-body mixed lattice in its container, above the reference's 1024-body storage
  page."""
import numpy as np
import pytest

from gpu_util import bits, lockstep, log, manifold_dict, pair_set
import oracle_lib as ol
from physecs_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def _first_contact_step(api, sc, max_steps):
    """Steps the device and the oracle until the first step with manifolds, then compares
    the pair sets and the manifolds bit for bit."""
    g = api.World(sc, substeps=4, iterations=2)
    o = ol.OracleWorld(sc, substeps=4, iterations=2)
    for k in range(max_steps):
        g.step(1 / 60)
        keys, _, _ = g.manifolds()
        o.set_order(keys)
        o.step(1 / 60)
        if len(keys):
            break
    assert pair_set(g.pairs()) == pair_set(o.pairs())
    gm, om = manifold_dict(*g.manifolds()), manifold_dict(*o.manifolds())
    assert set(gm) == set(om) and len(gm) > 0
    for key in gm:
        assert np.array_equal(gm[key][0][:3], om[key][0][:3]), key
        assert np.array_equal(bits(gm[key][1]), bits(om[key][1])), key
    log(f"  first contacts at step {k}: {len(gm)} manifolds")


def test_convex_terrain_manifolds_at_first_contact(api):
    sc = scenes.convex_terrain(5, n_convex=48, n_chains=4, quads=16, amplitude=2.0, drop=(0.3, 1.5))
    log(f"[C3] convex terrain: {sc.n_dyn} bodies, {len(sc.joints)} joints, first contact")
    _first_contact_step(api, sc, 60)


@pytest.mark.parametrize("seed", [5, 6])
def test_convex_terrain_lockstep(api, seed):
    sc = scenes.convex_terrain(seed, n_convex=48, n_chains=4, quads=16, amplitude=2.0, drop=(0.3, 2.0))
    log(f"[C3] convex terrain seed {seed}: {sc.n_dyn} bodies, {len(sc.joints)} joints, 60 steps")
    g, _ = lockstep(sc, 60)
    cnt = g.counts()
    log(f"  counts {cnt}")
    assert cnt["manifolds"] > 40 and cnt["joint_rows"] > 0


@pytest.mark.parametrize("seed,n", [(3, 300), (8, 1200)])
def test_served_request_lockstep(api, seed, n):
    sc = scenes.served_request(seed, n_bodies=n, chain=True)
    log(f"[C5] served request seed {seed}: {sc.n_dyn} bodies, {len(sc.joints)} joints, 40 steps")
    g, _ = lockstep(sc, 40, check_every=4)
    assert g.counts()["manifolds"] > 0


@pytest.mark.parametrize("mode", ["colors", "pass", "xpass"])
def test_mixed_32k_lockstep(api, mode):
    sc = scenes.mixed_lattice(0, side=32)
    log(f"[C2] mixed lattice: {sc.n_dyn} bodies, 12 steps, exec {mode}")
    g, _ = lockstep(sc, 12, check_every=4, mode=mode)
    cnt = g.counts()
    log(f"  counts {cnt}")
    assert sc.n_dyn == 32768 and cnt["manifolds"] > 1000


@pytest.mark.parametrize("mode", ["pass", "xpass"])
def test_pass_mode_equals_colour_mode_settled(api, mode):
    """The 32,768-body scene well into its pile (about 20 colours of thousands of manifolds,
    so many barriers per pass): one launch per pass (on every XCD, or on one) equals one
    launch per colour bit for bit at every checked step (the colour mode is the one the
    locksteps above pin)."""
    sc = scenes.mixed_lattice(0, side=32)
    log(f"[C2] {mode} vs colours, 150 steps")
    a = api.World(sc, substeps=4, iterations=2)
    b = api.World(sc, substeps=4, iterations=2)
    b.set_exec(mode)
    for k in range(150):
        a.step(1 / 60)
        b.step(1 / 60)
        if (k + 1) % 25 == 0:
            for x, y in zip(a.state(), b.state()):
                assert np.array_equal(bits(x), bits(y)), k
    cnt = b.counts()
    log(f"  counts {cnt}")
    assert cnt["colors"] >= 8 and cnt["manifolds"] > 20000
