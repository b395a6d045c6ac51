"""The drop-in physecs::Scene (physecs_amd/shim) stepping on the device through
Scene::simulate: each step copies the registry's poses and velocities to the world, steps,
and writes them back into the TransformComponent and RigidBodyDynamicComponent of every
dynamic body; trigger listeners get their callbacks. Against the same scene stepped through
the C ABI directly (api.World, the path the oracle locksteps pin), every body's pose and
velocities read from the registry are bit-identical at every checked step, and the callbacks
are the trigger events. Scenes: every joint type (joints created through
Scene::createJoint, pairs through setCanCollide), a convex pile on a triangle-mesh terrain
(the terrain's BVH from the reference's own TriangleMesh constructor, carried by a cache),
and colliders whose data bits make trigger pairs under a contact filter set through
Scene::setContactFilter. User Joint subclasses made by createJoint<T> step as the built-in
joints they restate, and one that reads the registry in getSolverDesc holds its bodies.
Needs oracle/_ref/libshimcheck.so (built where the reference's headers are)."""
import numpy as np
import pytest

from gpu_util import bits, log
from physecs_amd import scenes
import shim_lib

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not shim_lib.available(), reason="oracle/_ref/libshimcheck.so not built")]


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def _joints_through_scene(api, reg, ents, sc):
    for (t, a0, a1, an0, an1, p) in sc.joints:
        d = api.PgxJointDesc()
        scenes.Scene.fill_joint(d, t, a0, a1, an0, an1, p)
        reg.create_joint(d, ents[a0], ents[a1])
    for e0, e1 in sc.non_colliding:
        reg.set_can_collide(ents[e0], ents[e1], False)


def _run(api, sc, steps, substeps=4, iterations=2, check_every=5, filter_fn=None, filter_id=0, triggers=False):
    g = api.World(sc, substeps=substeps, iterations=iterations)
    if filter_fn is not None:
        g.set_contact_filter(filter_fn)
    reg, ents = shim_lib.replay(sc)
    arr = sc.arrays()
    seen = 0
    with reg:
        _joints_through_scene(api, reg, ents, sc)
        if filter_id:
            reg.set_contact_filter(filter_id)
        if triggers:
            reg.listen_triggers()
        for k in range(steps):
            g.step(1 / 60)
            reg.simulate(1 / 60, substeps, iterations)
            if triggers:
                got = reg.take_triggers()
                ge, gx = g.trigger_events()
                want = [(flag, int(arr["entity"][arr["col_actor"][c0]]), int(arr["col_index"][c0]),
                         int(arr["entity"][arr["col_actor"][c1]]), int(arr["col_index"][c1]))
                        for flag, ev in ((1, ge), (0, gx)) for c0, c1 in ev]
                assert [tuple(r) for r in got.tolist()] == want, k
                seen += len(want)
            if (k + 1) % check_every and k != steps - 1:
                continue
            pos, quat, vel, ang = g.state()
            for d in range(sc.n_dyn):
                a = int(arr["actor_of_dyn"][d])
                p, q, v, w = reg.read(ents[a])
                assert np.array_equal(bits(p), bits(pos[a])), (k, d, "pos")
                assert np.array_equal(bits(q), bits(quat[a])), (k, d, "quat")
                assert np.array_equal(bits(v), bits(vel[d])), (k, d, "vel")
                assert np.array_equal(bits(w), bits(ang[d])), (k, d, "angvel")
    g.close()
    return seen


def test_scene_simulate_joints(api):
    sc = scenes.mechanism_scene(0)
    log(f"[shim] mechanism scene: {sc.n_dyn} bodies, {len(sc.joints)} joints, 40 steps")
    _run(api, sc, 40)


def test_scene_simulate_joint_chains(api):
    sc = scenes.joint_scene(1)
    log(f"[shim] joint chains: {sc.n_dyn} bodies, {len(sc.joints)} joints, 40 steps")
    _run(api, sc, 40)


def test_scene_simulate_convex_on_terrain(api):
    sc = scenes.terrain_drop(3, quads=12, n=40)
    log(f"[shim] terrain drop: {sc.n_dyn} bodies, {len(sc.trimesh)} triangle mesh, 60 steps")
    _run(api, sc, 60, check_every=10)


def test_scene_simulate_contact_filter_triggers(api):
    rng = np.random.default_rng(4)
    sc = scenes.Scene()
    sc.add_static(scenes.BOX, [10, 0.5, 10], [0, -0.5, 0], data=8)
    sc.add_static(scenes.BOX, [1.5, 1.0, 1.5], [0, 2.0, 0], data=1, trigger=True)
    for i in range(48):
        k = int(rng.integers(0, 3))
        prm = [0.3] if k == 0 else ([0.3, 0.2] if k == 1 else [0.25, 0.3, 0.2])
        sc.add_body(k, prm, [rng.uniform(-2, 2), rng.uniform(1.0, 6.0), rng.uniform(-2, 2)],
                    scenes.random_quats(rng, 1)[0], data=int(rng.choice([0, 1, 2, 4])))

    def share_bit(t0, d0, t1, d1):   # the filter oracle/shim_check.cpp sets as filter 1
        return t0 or t1 or (d0 & d1) != 0

    log("[shim] contact filter + trigger listeners: 50 bodies, 80 steps")
    seen = _run(api, sc, 80, check_every=10, filter_fn=share_bit, filter_id=1, triggers=True)
    log(f"  {seen} trigger callbacks")
    assert seen > 10


_USER_KIND = {scenes.SPHERICAL: "POINT", scenes.UNIVERSAL: "UNIVERSAL", scenes.REVOLUTE: "HINGE"}


def test_scene_user_joint_subclasses(api):
    """createJoint<T> with user Joint subclasses (the harness's PointJoint, UniversalPair and
    HingeJoint, written against the drop-in's Constraint1DContainer.h as against the
    reference's) in place of the built-in spherical, universal and revolute joints: their
    getSolverDesc and MakeConstraintsFunc run on the host every step and substep, and the
    bodies move bit for bit as with the built-in joints, through the same Scene, also after a
    collider edit has rebuilt the world and replayed the joint creations."""
    sc = scenes.joint_scene(1)
    assert {j[0] for j in sc.joints} <= set(_USER_KIND) and not any(j[5].get("drive_enabled") for j in sc.joints)
    log(f"[shim] user joint subclasses: {len(sc.joints)} joints, 40 steps")
    ref, ents0 = shim_lib.replay(sc)
    usr, ents1 = shim_lib.replay(sc)
    arr = sc.arrays()
    with ref, usr:
        _joints_through_scene(api, ref, ents0, sc)
        for (t, a0, a1, an0, an1, p) in sc.joints:
            usr.create_user_joint(getattr(usr, _USER_KIND[t]), ents1[a0], an0, ents1[a1], an1)
        for k in range(40):
            if k == 20:   # a collider edit rebuilds both worlds; the joints are created again
                for reg, ents in ((ref, ents0), (usr, ents1)):
                    reg.clear_colliders(ents[0])
                    reg.add_collider(ents[0], int(arr["col_type"][0]), arr["col_params"][0],
                                     material=tuple(arr["col_material"][0]), lpos=tuple(arr["col_pos"][0]),
                                     lquat=tuple(arr["col_quat"][0]))
            ref.simulate(1 / 60, 4, 2)
            usr.simulate(1 / 60, 4, 2)
            for d in range(sc.n_dyn):
                a = int(arr["actor_of_dyn"][d])
                for x, y in zip(ref.read(ents0[a]), usr.read(ents1[a])):
                    assert np.array_equal(bits(x), bits(y)), (k, d)


def test_scene_user_tether_and_destroy(api):
    """A user joint whose getSolverDesc reads the registry (as PrismaticJoint's does) and
    declares its limit row only while the rope is taut; destroyJoint(Joint*) lets go."""
    sc = scenes.Scene()
    sc.add_static(scenes.BOX, [10, 0.5, 10], [0, -0.5, 0])
    post = sc.add_static(scenes.SPHERE, [0.05], [0, 5, 0])
    box = sc.add_body(scenes.BOX, [0.2, 0.2, 0.2], [0.5, 4.5, 0], vel=(3.0, 0.0, 0.0))
    log("[shim] user tether: 120 steps, then destroyed, 60 steps")
    reg, ents = shim_lib.replay(sc)
    with reg:
        h = reg.create_user_joint(reg.TETHER, ents[post], (0, 0, 0, 0, 0, 0, 1), ents[box], (0, 0, 0, 0, 0, 0, 1),
                                  1.0)
        top = np.array([0, 5, 0], np.float32)
        far = 0.0
        for k in range(120):
            reg.simulate(1 / 60, 4, 2)
            far = max(far, float(np.linalg.norm(reg.read(ents[box])[0] - top)))
        assert reg.tether_taut_steps(h) > 10
        assert far < 1.05, far                     # held within its length
        y0 = float(reg.read(ents[box])[0][1])
        reg.destroy_user_joint(h)
        for k in range(60):
            reg.simulate(1 / 60, 4, 2)
        assert float(reg.read(ents[box])[0][1]) < y0 - 2.0   # falls once let go
