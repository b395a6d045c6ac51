"""Custom joints (user Joint subclasses, include/Physecs/Joint.h:42-70) written in Python
the way a user writes one: getSolverDesc's row layout and MakeConstraintsFunc's rows, from
the joint's world-space data only (p0, p1, r0, r1, u0, u1; src/Joint.cpp:5-14).

Restated here as such user joints: the built-in spherical, universal, revolute, fixed and
prismatic joints (src/JointsUtil.h:8-28, src/Joints/SphericalJoint.cpp, UniversalJoint.cpp,
RevoluteJoint.cpp, FixedJoint.cpp, PrismaticJoint.cpp), so that a world whose joints are
replaced by these must step bit for bit as the world with the built-in ones; and two joints
the reference does not ship: a rope (a limit row only while the distance exceeds the length,
so its layout changes from step to step) and a soft spring.

All arithmetic is float32 (numpy scalars): products, sums and differences in the reference's
order (glm::cross, glm::dot as ((x + y) + z))."""
import copy

import numpy as np

from physecs_amd import scenes

f32 = np.float32
NONE, SOFT, ANGULAR, LIMITED = scenes.ROW_NONE, scenes.ROW_SOFT, scenes.ROW_ANGULAR, scenes.ROW_LIMITED
ROPE, SPRING = "rope", "spring"


def _v(a, i):
    return (a[i], a[i + 1], a[i + 2])


def world_space(ws):
    """(p0, p1, r0, r1, u0 columns, u1 columns) of a (30,) float32 world-space array."""
    return (_v(ws, 0), _v(ws, 3), _v(ws, 6), _v(ws, 9), (_v(ws, 12), _v(ws, 15), _v(ws, 18)),
            (_v(ws, 21), _v(ws, 24), _v(ws, 27)))


def sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def cross(x, y):
    return (x[1] * y[2] - y[1] * x[2], x[2] * y[0] - y[2] * x[0], x[0] * y[1] - y[0] * x[1])


def neg(a):
    return (-a[0], -a[1], -a[2])


def scale(s, a):
    return (s * a[0], s * a[1], s * a[2])


def put(row, lin=None, ang0=None, ang1=None, tv=None, c=None, mn=None, mx=None, freq=None, damp=None):
    """Constraint1DView's setters on one (15,) row: only the fields given are written."""
    if lin is not None:
        row[0:3] = lin
    if ang0 is not None:
        row[3:6] = ang0
    if ang1 is not None:
        row[6:9] = ang1
    for k, v in ((9, tv), (10, c), (11, mn), (12, mx), (13, freq), (14, damp)):
        if v is not None:
            row[k] = v


def point_to_point(rows, p0, p1, r0, r1):
    """createPointToPointConstraint, src/JointsUtil.h:8-28"""
    d = sub(p1, p0)
    z = f32(0)
    put(rows[0], lin=(1, 0, 0), ang0=(z, r0[2], -r0[1]), ang1=(z, r1[2], -r1[1]), c=d[0])
    put(rows[1], lin=(0, 1, 0), ang0=(-r0[2], z, r0[0]), ang1=(-r1[2], z, r1[0]), c=d[1])
    put(rows[2], lin=(0, 0, 1), ang0=(r0[1], -r0[0], z), ang1=(r1[1], -r1[0], z), c=d[2])


def angular(row, x, y):
    """the joints' angular rows: axis cross(y, x) on both bodies, c = dot(x, y)"""
    ax = cross(y, x)
    put(row, ang0=ax, ang1=ax, c=dot(x, y))


class Restated:
    """Layout and rows functions for a scene whose joints listed in `kinds` ({joint index:
    (kind, params)}, kind a built-in type, ROPE or SPRING) are custom joints."""

    def __init__(self, kinds):
        self.kinds = kinds
        self.calls = []          # (0 layout | 1 rows, joint), in call order
        self.rope_limit = {}     # per rope joint: the last step's layout has the limit row
        self.rope_layouts = set()   # every value rope_limit took

    def layout(self, joint, ws):
        self.calls.append((0, joint))
        kind, p = self.kinds[joint]
        p0, p1, r0, r1, u0, u1 = world_space(ws)
        if kind == scenes.SPHERICAL:
            return [NONE] * 3
        if kind == scenes.UNIVERSAL:
            return [NONE] * 3 + [ANGULAR]
        if kind == scenes.REVOLUTE:
            return [NONE] * 3 + [ANGULAR] * 2 + ([ANGULAR | LIMITED] if p.get("drive_enabled") else [])
        if kind == scenes.FIXED:
            return [NONE] * 3 + [ANGULAR] * 3
        if kind == scenes.PRISMATIC:   # PrismaticJoint.cpp:123-150: the limit row from the step-start pose
            dx = dot(sub(p1, p0), u0[0])
            fl = [NONE] * 2 + [ANGULAR] * 3
            p["_kind"] = 1 if dx > f32(p["upper_limit"]) else (2 if dx < f32(p["lower_limit"]) else 0)
            if p["_kind"]:
                fl.append(LIMITED)
            if p.get("drive_enabled"):
                fl.append(SOFT)
            return fl
        if kind == ROPE:
            d = sub(p1, p0)
            taut = np.sqrt(dot(d, d)) > f32(p["length"])
            self.rope_limit[joint] = taut
            self.rope_layouts.add(bool(taut))
            return [LIMITED] if taut else []
        if kind == SPRING:
            return [SOFT]
        raise ValueError(kind)

    def rows(self, joint, ws, rows):
        self.calls.append((1, joint))
        kind, p = self.kinds[joint]
        p0, p1, r0, r1, u0, u1 = world_space(ws)
        d = sub(p1, p0)
        if kind in (scenes.SPHERICAL, scenes.UNIVERSAL, scenes.REVOLUTE, scenes.FIXED):
            point_to_point(rows, p0, p1, r0, r1)
            if kind == scenes.UNIVERSAL:                       # UniversalJoint.cpp
                angular(rows[3], u0[2], u1[2])
            if kind in (scenes.REVOLUTE, scenes.FIXED):        # RevoluteJoint.cpp, FixedJoint.cpp
                angular(rows[3], u0[0], u1[1])
                angular(rows[4], u0[0], u1[2])
            if kind == scenes.FIXED:
                angular(rows[5], u0[1], u1[2])
            if kind == scenes.REVOLUTE and p.get("drive_enabled"):
                m = f32(p["drive_max_torque"])
                put(rows[5], ang0=u0[0], ang1=u0[0], tv=f32(p["drive_velocity"]), mx=m, mn=-m)
            return
        if kind == scenes.PRISMATIC:                           # PrismaticJoint.cpp:6-92
            for k in range(2):
                u = u0[k + 1]
                put(rows[k], lin=u, ang0=cross(r0, u), ang1=cross(r1, u), c=dot(d, u))
            angular(rows[2], u0[0], u1[1])
            angular(rows[3], u0[0], u1[2])
            angular(rows[4], u0[1], u1[2])
            dx = dot(d, u0[0])
            r0x, r1x = cross(r0, u0[0]), cross(r1, u0[0])
            r = 5
            if p["_kind"] == 1:
                put(rows[r], lin=u0[0], ang0=r0x, ang1=r1x, c=dx - f32(p["upper_limit"]), mn=f32(0))
                r += 1
            elif p["_kind"] == 2:
                put(rows[r], lin=u0[0], ang0=r0x, ang1=r1x, c=dx - f32(p["lower_limit"]), mx=f32(0))
                r += 1
            if p.get("drive_enabled"):
                put(rows[r], lin=u0[0], ang0=r0x, ang1=r1x, c=dx - f32(p["target"]),
                    freq=f32(p["drive_stiffness"]), damp=f32(p["drive_damping"]))
            return
        if kind in (ROPE, SPRING):
            length = np.sqrt(dot(d, d))
            n = scale(f32(1) / length, d) if length > f32(0) else (f32(1), f32(0), f32(0))
            if kind == ROPE:
                if len(rows):
                    put(rows[0], lin=n, ang0=cross(r0, n), ang1=cross(r1, n), c=length - f32(p["length"]), mn=f32(0))
            else:
                put(rows[0], lin=n, ang0=cross(r0, n), ang1=cross(r1, n), c=length - f32(p["rest"]),
                    freq=f32(p["frequency"]), damp=f32(p["damping"]))
            return
        raise ValueError(kind)


def customised(scene, types):
    """A copy of the scene whose joints of the given built-in types are custom joints, and
    the Restated functions that write their rows."""
    s2 = copy.copy(scene)   # shares the built arrays; only the joint list differs
    kinds = {}
    joints = []
    for i, (t, a0, a1, an0, an1, p) in enumerate(scene.joints):
        if t in types:
            kinds[i] = (t, dict(p))
            joints.append((scenes.CUSTOM, a0, a1, an0, an1, {}))
        else:
            joints.append((t, a0, a1, an0, an1, p))
    s2.joints = joints
    return s2, Restated(kinds)


def fixed_pairs(s, origin, n=3):
    """n pairs of boxes welded by fixed joints, tilted, above the ground."""
    rng = np.random.default_rng(7)
    for i in range(n):
        o = np.asarray(origin, float) + [1.2 * i, 0, 0]
        q = scenes.random_quats(rng, 1)[0]
        a = s.add_body(scenes.BOX, [0.2, 0.2, 0.2], o, q)
        b = s.add_body(scenes.BOX, [0.15, 0.3, 0.15], o + [0, 0.5, 0], q)
        s.add_joint(scenes.FIXED, a, b, (0, 0.2, 0), (0, -0.3, 0))


def rope_scene(seed=3):
    """Boxes hung from static anchors by ropes and springs, some chained, dropped onto a pile."""
    rng = np.random.default_rng(seed)
    s = scenes.Scene()
    s.add_static(scenes.BOX, [20, 0.5, 20], [0, -0.5, 0])
    kinds = {}
    for i in range(6):
        post = s.add_static(scenes.SPHERE, [0.05], [-5.0 + 2.0 * i, 4.0, 0.0])
        prev = post
        for k in range(3):
            b = s.add_body(scenes.BOX, list(rng.uniform(0.1, 0.25, 3)), [-5.0 + 2.0 * i + 0.4 * (k + 1), 3.6 - 0.5 * k, 0.2],
                           scenes.random_quats(rng, 1)[0], vel=rng.normal(size=3))
            j = s.add_joint(scenes.CUSTOM, prev, b, (0, 0, 0) if prev == post else (0.05, -0.1, 0), (0, 0.1, 0))
            if (i + k) % 2:
                kinds[j] = (ROPE, dict(length=float(rng.uniform(0.4, 0.8))))
            else:
                kinds[j] = (SPRING, dict(rest=float(rng.uniform(0.3, 0.6)), frequency=float(rng.uniform(2.0, 6.0)),
                                         damping=float(rng.uniform(0.2, 1.0))))
            prev = b
    for i in range(20):
        s.add_body(scenes.BOX, list(rng.uniform(0.15, 0.3, 3)), [rng.uniform(-5, 5), rng.uniform(0.3, 2.0), rng.uniform(-2, 2)],
                   scenes.random_quats(rng, 1)[0])
    return s, Restated(kinds)


def hub_scene(n=10):
    """A dynamic hub with n boxes joined to it by spherical joints: more than 8 joints on one
    body, so the last ones take the scalar overflow colour."""
    s = scenes.Scene()
    s.add_static(scenes.BOX, [10, 0.5, 10], [0, -0.5, 0])
    hub = s.add_body(scenes.SPHERE, [0.3], [0, 3.0, 0], vel=(0.5, 0.0, 0.2))
    for i in range(n):
        a = 2 * np.pi * i / n
        d = np.array([np.cos(a), 0.2 * (i % 3), np.sin(a)])
        b = s.add_body(scenes.BOX, [0.1, 0.1, 0.1], [0, 3.0, 0] + 0.8 * d)
        s.add_joint(scenes.SPHERICAL if i % 2 else scenes.REVOLUTE, hub, b, tuple(0.4 * d), (0, 0, 0))
    return s
