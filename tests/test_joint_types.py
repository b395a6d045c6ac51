"""Closed-form checks of the oracle's prismatic, servo and gear joints and of destroyJoint
(src/Joints/{Prismatic,Servo,Gear}Joint.cpp, src/Physecs.cpp:712-725). The reference's
joint files do not compile here (include/Physecs/SIMD.h is MSVC-only), so these rows are
pinned by what the mechanics fixes: free slides, limits, drive targets, gear velocity
ratios, free fall after a joint is destroyed."""
import numpy as np

import oracle_lib as ol
from physecs_amd import scenes


def _anchor_and_body(gtype=scenes.BOX, params=(0.2, 0.2, 0.2), pos=(0, 10, 0), vel=(0, 0, 0), angvel=(0, 0, 0)):
    sc = scenes.Scene()
    anchor = sc.add_static(scenes.SPHERE, [0.05], [0, 10, 0])
    b = sc.add_body(gtype, list(params), list(pos), vel=vel, angvel=angvel)
    return sc, anchor, b


def _no_collisions(sc):
    sc.arrays()["col_enable"][:] = 0
    sc._frozen = None


def _angle_about_x(q):
    return 2.0 * np.arctan2(q[0], q[3])


def test_prismatic_slides_freely_and_stops_at_the_limit():
    """No gravity; the body starts with velocity (1, 0.5, -0.3). The two linear rows take out
    the motion across the axis in the first step, the body then slides at 1 m/s along x
    (anchor frames are the identity, so every other row is exactly zero), and the limit row
    stops it at upper = 0.5. The row exists only in steps that start beyond the limit
    (PrismaticJoint.cpp:131-141), so the body settles within a centimetre of it."""
    sc, anchor, b = _anchor_and_body(vel=(1.0, 0.5, -0.3))
    _no_collisions(sc)
    sc.add_joint(scenes.PRISMATIC, anchor, b, (0, 0, 0), (0, 0, 0), lower_limit=-1.0, upper_limit=0.5)
    o = ol.OracleWorld(sc, substeps=4, iterations=2, gravity=0.0)
    o.step(1 / 60)
    v = o.state()[2][0]
    assert abs(v[1]) < 1e-6 and abs(v[2]) < 1e-6, v
    vx = v[0]
    assert 0.99 < vx <= 1.0
    for k in range(20):
        o.step(1 / 60)
        pos, _, vel, _ = o.state()
        assert vel[0][0] == vx and vel[0][1] == 0 and vel[0][2] == 0
    for k in range(60):
        o.step(1 / 60)
    pos, _, vel, _ = o.state()
    assert abs(pos[1][0] - 0.5) < 1e-2, pos[1]
    assert abs(vel[0][0]) < 1e-2, vel[0]


def test_prismatic_drive_reaches_its_target():
    sc, anchor, b = _anchor_and_body()
    _no_collisions(sc)
    sc.add_joint(scenes.PRISMATIC, anchor, b, (0, 0, 0), (0, 0, 0), lower_limit=-10.0, upper_limit=10.0,
                 drive_enabled=1, target=0.3, drive_stiffness=5.0, drive_damping=1.0)
    o = ol.OracleWorld(sc, substeps=4, iterations=2, gravity=0.0)
    for _ in range(240):
        o.step(1 / 60)
    pos, _, vel, _ = o.state()
    assert abs(pos[1][0] - 0.3) < 1e-3 and abs(vel[0][0]) < 1e-3, (pos[1], vel[0])
    assert abs(pos[1][1] - 10.0) < 1e-5 and abs(pos[1][2]) < 1e-5


def test_servo_turns_to_its_target_angle_and_follows_a_new_one():
    """The servo keeps the x axes aligned and drives the angle between the z axes about x
    (glm::orientedAngle) to the target; a later setTargetAngle moves it again."""
    sc, anchor, b = _anchor_and_body()
    _no_collisions(sc)
    j = sc.add_joint(scenes.SERVO, anchor, b, (0, 0, 0), (0, 0, 0), target=0.7)
    o = ol.OracleWorld(sc, substeps=4, iterations=2, gravity=0.0)
    for _ in range(180):
        o.step(1 / 60)
    q = o.state()[1][1]
    assert abs(_angle_about_x(q) - 0.7) < 2e-3, q
    assert abs(q[1]) < 1e-4 and abs(q[2]) < 1e-4
    o.set_joint_params(j, target=-0.4)
    for _ in range(180):
        o.step(1 / 60)
    q = o.state()[1][1]
    assert abs(_angle_about_x(q) + 0.4) < 2e-3, q


def test_gear_couples_wheel_speeds_through_angle_wraps():
    """Two wheels on revolute joints about x, side by side along z, joined by a gear of
    ratio 2: the gear row drives dot(u1x, w1) + ratio * dot(u0x, w0) to zero. The first wheel
    starts at 10 rad/s, so both wheels turn through several wraps of the gear's [-pi, pi)
    angle difference while the ratio holds."""
    sc = scenes.Scene()
    anchor = sc.add_static(scenes.SPHERE, [0.05], [0, 5, 0])
    a = sc.add_body(scenes.SPHERE, [0.4], [0, 5, -1], angvel=(10.0, 0, 0))
    b = sc.add_body(scenes.SPHERE, [0.4], [0, 5, 1])
    _no_collisions(sc)
    sc.add_joint(scenes.REVOLUTE, anchor, a, (0, 0, -1), (0, 0, 0))
    sc.add_joint(scenes.REVOLUTE, anchor, b, (0, 0, 1), (0, 0, 0))
    sc.add_joint(scenes.GEAR, a, b, (0, 0, 0), (0, 0, 0), gear_ratio=2.0)
    o = ol.OracleWorld(sc, substeps=4, iterations=2, gravity=0.0)
    turned = np.zeros(2)
    for k in range(600):
        o.step(1 / 60)
        w = o.state()[3]
        turned += w[:, 0] / 60
        if k > 60:
            assert abs(w[1][0] + 2.0 * w[0][0]) < 2e-2 * abs(w[0][0]), (k, w)
    assert abs(turned[0]) > 4 * np.pi and abs(turned[1]) > 8 * np.pi, turned
    w = o.state()[3]
    assert np.all(np.abs(w[:, 1:]) < 1e-4)


def test_destroyed_joint_frees_the_body():
    """After destroyJoint the body falls freely: each substep adds -g*h to its velocity."""
    sc, anchor, b = _anchor_and_body(gtype=scenes.SPHERE, params=(0.2,), pos=(0, 9, 0))
    _no_collisions(sc)
    j = sc.add_joint(scenes.SPHERICAL, anchor, b, (0, -1, 0), (0, 0, 0))
    o = ol.OracleWorld(sc, substeps=4, iterations=2)
    for _ in range(30):
        o.step(1 / 60)
    assert abs(o.state()[0][1][1] - 9.0) < 1e-2
    o.destroy_joint(j)
    v0 = o.state()[2][0][1]
    for _ in range(30):
        o.step(1 / 60)
    v1 = o.state()[2][0][1]
    assert abs((v1 - v0) + 9.81 * 30 / 60) < 1e-4, (v0, v1)
