"""Parity of the HIP step (libphysecs_gpu.so, through its C ABI) with the CPU oracle on
seeded scenes. Pair sets, manifolds and whole steps are compared bit for bit; the oracle
solves the contacts in the device's graph-colour order (the reference solves them in push
order; any order is a valid Gauss-Seidel sweep, and fed the same order the two agree
exactly)."""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import bits, lockstep, log, manifold_dict, pair_set
from physecs_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pairs_bounds_manifolds_first_step(api, seed):
    log(f"[pairs/manifolds] random_cluster seed={seed}")
    sc = scenes.random_cluster(seed, n=150, extent=2.5)
    g = api.World(sc, substeps=4)
    o = ol.OracleWorld(sc, substeps=4)
    g.step(1 / 60)
    o.step(1 / 60)
    # broadphase: identical pair set (the reference's SAP vs the LBVH)
    gp, op = pair_set(g.pairs()), pair_set(o.pairs())
    assert len(gp) > 50
    assert gp == op, (len(gp - op), len(op - gp))
    # narrowphase + contact setup: identical manifolds by key
    gm = manifold_dict(*g.manifolds())
    om = manifold_dict(*o.manifolds())
    assert set(gm) == set(om)
    for k in gm:
        gi, gf = gm[k]
        oi, of = om[k]
        assert gi[0] == oi[0] and gi[1] == oi[1] and gi[2] == oi[2], k
        assert np.array_equal(bits(gf), bits(of)), (k, gf, of)


def test_coloring_is_proper(api):
    log("[coloring] mixed_lattice side=8")
    sc = scenes.mixed_lattice(3, side=8, with_container=True)
    g = api.World(sc, substeps=4)
    for _ in range(40):
        g.step(1 / 60)
    keys, ints, f = g.manifolds()
    colors = ints[:, 3]
    assert np.all(np.diff(colors) >= 0), "manifolds not in colour order"
    for c in np.unique(colors):
        sel = ints[colors == c]
        bodies = np.concatenate([sel[:, 0], sel[:, 1]])
        bodies = bodies[bodies >= 0]
        assert len(bodies) == len(np.unique(bodies)), f"colour {c} touches a body twice"
    cnt = g.counts()
    assert cnt["colors"] == len(np.unique(colors))
    log(f"  counts {cnt}")


def test_free_fall_closed_form(api):
    """No contacts: every substep adds -g*h to v.y exactly as the reference does."""
    sc = scenes.Scene()
    for i in range(10):
        sc.add_body(scenes.SPHERE, [0.3], [i * 2.0, 10.0, 0.0])
    g = api.World(sc, substeps=4)
    o = ol.OracleWorld(sc, substeps=4)
    for _ in range(30):
        g.step(1 / 60)
        o.step(1 / 60)
    gs, os_ = g.state(), o.state()
    for a, b in zip(gs, os_):
        assert np.array_equal(bits(a), bits(b))
    assert abs(gs[2][0, 1] - (-9.81 * 30 / 60)) < 1e-3


@pytest.mark.parametrize("mode", ["xcd", "flow", "colors", "cooperative", "worlds", "pass", "xpass"])
def test_pyramid_lockstep_bitexact(api, mode):
    """20-layer box pyramid, dt 1/60, 4 substeps, 2 iterations: 120 steps in lockstep, in
    every execution mode of the contact passes (dependency-driven, one launch per colour,
    one cooperative launch per step)."""
    log(f"[lockstep] box pyramid 120 steps exec={mode}")
    lockstep(scenes.box_pyramid(0), 120, mode=mode)


@pytest.mark.parametrize("seed", [0, 1])
def test_mixed_lockstep_bitexact(api, seed):
    """Spheres, capsules and boxes in a container: every pair type, 80 steps."""
    log(f"[lockstep] mixed lattice side=7 seed={seed}")
    lockstep(scenes.mixed_lattice(seed, side=7), 80, mode="flow" if seed == 0 else "colors")


def test_cluster_lockstep_with_restitution_and_soft(api):
    """Restitution (contact cache across steps) and soft contacts (damping > 0)."""
    log("[lockstep] cluster with restitution + soft materials")
    sc = scenes.random_cluster(5, n=80, extent=2.0)
    mat = sc.col_material
    rng = np.random.default_rng(5)
    for i in range(1, len(mat)):
        k = rng.integers(0, 3)
        if k == 0:
            mat[i] = np.array([0.4, 0.5, 0.0], np.float32)      # bouncy
        elif k == 1:
            mat[i] = np.array([0.6, 20.0, 0.7], np.float32)     # soft: restitution field is a frequency
    sc._frozen = None
    lockstep(sc, 60)


@pytest.mark.parametrize("seed,mode", [(0, "colors"), (1, "colors"), (0, "worlds"), (1, "pass"), (0, "xpass")])
def test_joints_lockstep_bitexact(api, seed, mode):
    """Revolute chains and ragdolls (spherical / universal / revolute rows, SIMD colours);
    contact passes per colour or per world with the joint colours between them."""
    log(f"[lockstep] joint scene seed={seed} exec {mode}")
    lockstep(scenes.joint_scene(seed, chains=6, ragdolls=3), 120, mode=mode)


def star_scene(spokes=11):
    """One hub body jointed to `spokes` bodies: colours 0-7 fill up and the rest go to the
    scalar overflow colour, whose joints share the hub."""
    sc = scenes.Scene()
    sc.add_static(scenes.BOX, [20, 0.5, 20], [0, -0.5, 0])
    hub = sc.add_body(scenes.SPHERE, [0.3], [0, 3.0, 0])
    for i in range(spokes):
        ang = 2 * np.pi * i / spokes
        d = np.array([np.cos(ang), 0.0, np.sin(ang)])
        b = sc.add_body(scenes.BOX, [0.15, 0.15, 0.15], [0, 3.0, 0] + 1.2 * d)
        jt = [scenes.SPHERICAL, scenes.UNIVERSAL, scenes.REVOLUTE, scenes.FIXED][i % 4]
        drive = (1.5, 50.0) if jt == scenes.REVOLUTE and i % 8 == 2 else None
        sc.add_joint(jt, hub, b, tuple(0.3 * d), tuple(-1.05 * d), drive=drive)
    return sc


def test_joint_overflow_colour_lockstep(api):
    """A hub with 11 joints: 8 SIMD colours + the sequential overflow colour, fixed joints and
    a revolute drive (ANGULAR | LIMITED rows)."""
    log("[lockstep] joint star (overflow colour)")
    sc = star_scene(11)
    o = ol.OracleWorld(sc, substeps=4)
    assert max(o.joint_colors()) == 8
    lockstep(sc, 90)


@pytest.mark.parametrize("seed", [0, 1])
def test_all_joint_types_lockstep_bitexact(api, seed):
    """Prismatic sliders (limit rows decided per step, drives) and pistons, servo arms, gear
    pairs on revolute wheels, boxes falling onto them. At step 40 the joints' setters change
    a drive, a servo target and a gear ratio; at step 80 a piston is destroyed. Every step
    bit-identical to the oracle."""
    log(f"[lockstep] mechanism scene seed={seed}")
    sc = scenes.mechanism_scene(seed)
    assert {j[0] for j in sc.joints} == {scenes.PRISMATIC, scenes.SERVO, scenes.REVOLUTE, scenes.GEAR}

    def hook(k, g, o):
        if k == 39:
            for w in (g, o):
                w.set_joint_params(0, drive_enabled=1, target=-0.5)
                w.set_joint_params(8, target=0.2)
                w.set_joint_params(12, gear_ratio=1.5)
        if k == 79:
            for w in (g, o):
                w.destroy_joint(5)
    lockstep(sc, 120, on_step=hook)


@pytest.mark.parametrize("seed", [0, 1])
def test_convex_manifolds_first_steps(api, seed):
    """Convex hulls among primitives: GJK/EPA in double precision on the device; pair sets
    and manifolds bit-identical to the oracle's over the first steps."""
    log(f"[convex] pile seed={seed}: manifolds")
    sc = scenes.convex_pile(seed, n=70, extent=1.6)
    g = api.World(sc, substeps=4)
    o = ol.OracleWorld(sc, substeps=4)
    for k in range(6):
        g.step(1 / 60)
        keys, ints, f = g.manifolds()
        o.set_order(keys)
        o.step(1 / 60)
        assert pair_set(g.pairs()) == pair_set(o.pairs()), k
        gm = manifold_dict(keys, ints, f)
        om = manifold_dict(*o.manifolds())
        assert set(gm) == set(om), k
        for key in gm:
            assert np.array_equal(bits(gm[key][1]), bits(om[key][1])), (k, key)
    assert len(gm) > 40


def test_convex_lockstep_bitexact(api):
    log("[lockstep] convex pile 90 steps")
    lockstep(scenes.convex_pile(3, n=50, extent=2.0), 90)


@pytest.mark.parametrize("seed", [0, 1])
def test_trimesh_manifolds_first_steps(api, seed):
    """Spheres, capsules, boxes and convex hulls on a triangle-mesh terrain: the device BVH
    query, per-triangle tests and face/edge/vertex voiding give the oracle's manifolds,
    triangle indices included, bit for bit."""
    log(f"[trimesh] terrain drop seed={seed}: manifolds")
    sc = scenes.terrain_drop(seed, quads=12, n=40, heights=(0.0, 0.4))
    g = api.World(sc, substeps=4)
    o = ol.OracleWorld(sc, substeps=4)
    ntri = 0
    for k in range(40):
        g.step(1 / 60)
        keys, ints, f = g.manifolds()
        o.set_order(keys)
        o.step(1 / 60)
        assert pair_set(g.pairs()) == pair_set(o.pairs()), k
        gm = manifold_dict(keys, ints, f)
        om = manifold_dict(*o.manifolds())
        assert set(gm) == set(om), (k, len(set(gm) - set(om)), len(set(om) - set(gm)))
        for key in gm:
            assert np.array_equal(bits(gm[key][1]), bits(om[key][1])), (k, key)
        ntri = max(ntri, int(np.sum(keys[:, 4] >= 0)))
    assert ntri > 20


def test_trimesh_lockstep_bitexact(api):
    log("[lockstep] terrain drop 120 steps")
    lockstep(scenes.terrain_drop(2, quads=12, n=40), 120)
