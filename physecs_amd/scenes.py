"""Seeded synthetic scenes in the scene-descriptor layout that the C ABI takes
(include/physecs_gpu.h, pgx_scene_desc).

A scene is a set of actors (entities with a TransformComponent), a dynamic storage
(RigidBodyDynamicComponent entries, in storage order) and colliders (in broadphase-entry
creation order), with the same fields as the reference's components
(include/Physecs/Components.h:6-17, include/Physecs/Colliders.h:9-58, src/Transform.h:6-10).

Mass properties follow src/MassUtil.cpp for one collider at the body origin: mass =
density x true volume, inertia from getInertiaSphere/Capsule/Box (the box formula keeps
the reference's use of half-extents where full extents belong, MassUtil.cpp:22-28).
Random draws use numpy's PCG64 with the given seed, in the order written below.
"""
import math

import numpy as np

SPHERE, CAPSULE, BOX, CONVEX_MESH, TRIANGLE_MESH = range(5)
SPHERICAL, UNIVERSAL, REVOLUTE, FIXED = range(4)


def _inertia_sphere(m, r):
    return np.diag([2.0 * m * r * r / 5.0] * 3)


def _inertia_capsule(m, hh, r):
    vc = 2 * r * r * math.pi * hh
    vs = 4 * r ** 3 * math.pi / 3
    v = vc + vs
    mc, ms = vc * m / v, vs * m / v
    ixx = mc * (hh * hh / 3 + r * r / 4) + ms * (hh * hh + 3 * hh * r / 4 + 2 * r * r / 5)
    iyy = mc * r * r / 2 + ms * 2 * r * r / 5
    return np.diag([ixx, iyy, ixx])


def _inertia_box(m, he):
    x, y, z = he[0] ** 2, he[1] ** 2, he[2] ** 2
    return np.diag([m / 12 * (y + z), m / 12 * (x + z), m / 12 * (x + y)])


def true_volume(t, p):
    if t == SPHERE:
        return 4.0 / 3.0 * math.pi * p[0] ** 3
    if t == CAPSULE:
        return math.pi * p[1] ** 2 * 2 * p[0] + 4.0 / 3.0 * math.pi * p[1] ** 3
    return 8.0 * p[0] * p[1] * p[2]


def mass_props(t, p, mass):
    """(invMass, com, invInertia column-major) for one collider at the body origin."""
    if t == SPHERE:
        inertia = _inertia_sphere(mass, p[0])
    elif t == CAPSULE:
        inertia = _inertia_capsule(mass, p[0], p[1])
    else:
        inertia = _inertia_box(mass, p[:3])
    inv = np.diag(1.0 / np.diag(inertia))          # diagonal for one collider at the origin
    return np.float32(1.0 / mass), np.zeros(3, np.float32), inv.T.astype(np.float32).ravel()


def random_quats(rng, n):
    """Uniform random unit quaternions (Shoemake), returned as x,y,z,w."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)
    return q.astype(np.float32)


class Scene:
    """Scene descriptor builder. Actors, dynamic storage and colliders are appended in
    creation order; entity ids are the creation index unless given."""

    def __init__(self):
        self.entity, self.pos, self.quat, self.dyn_of_actor, self.world = [], [], [], [], []
        self.actor_of_dyn, self.kinematic, self.vel, self.angvel = [], [], [], []
        self.inv_mass, self.com, self.inv_inertia = [], [], []
        self.col_actor, self.col_index, self.col_pos, self.col_quat = [], [], [], []
        self.col_type, self.col_params, self.col_material = [], [], []
        self.col_trigger, self.col_enable, self.col_data = [], [], []
        self.non_colliding = []
        self.joints = []        # (type, actor0, actor1, anchor0[7], anchor1[7], drive_enabled, drive_v, drive_max)
        self._ncol = {}
        self.n_worlds = 1
        self._frozen = None

    # -- building
    def add_actor(self, pos, quat=(0, 0, 0, 1), world=0, entity=None):
        a = len(self.entity)
        self.entity.append(a if entity is None else entity)
        self.pos.append(np.asarray(pos, np.float32))
        self.quat.append(np.asarray(quat, np.float32))
        self.dyn_of_actor.append(-1)
        self.world.append(world)
        self.n_worlds = max(self.n_worlds, world + 1)
        return a

    def add_dynamic(self, actor, inv_mass, com, inv_inertia, vel=(0, 0, 0), angvel=(0, 0, 0), kinematic=False):
        d = len(self.actor_of_dyn)
        self.dyn_of_actor[actor] = d
        self.actor_of_dyn.append(actor)
        self.kinematic.append(1 if kinematic else 0)
        self.vel.append(np.asarray(vel, np.float32))
        self.angvel.append(np.asarray(angvel, np.float32))
        self.inv_mass.append(np.float32(inv_mass))
        self.com.append(np.asarray(com, np.float32))
        self.inv_inertia.append(np.asarray(inv_inertia, np.float32).ravel())
        return d

    def add_collider(self, actor, gtype, params, material=(0.6, 0.0, 0.0), lpos=(0, 0, 0), lquat=(0, 0, 0, 1),
                     trigger=False, enable=True, data=0):
        idx = self._ncol.get(actor, 0)
        self._ncol[actor] = idx + 1
        self.col_actor.append(actor)
        self.col_index.append(idx)
        self.col_pos.append(np.asarray(lpos, np.float32))
        self.col_quat.append(np.asarray(lquat, np.float32))
        self.col_type.append(gtype)
        p = np.zeros(4, np.float32)
        p[: len(params)] = params
        self.col_params.append(p)
        self.col_material.append(np.asarray(material, np.float32))
        self.col_trigger.append(1 if trigger else 0)
        self.col_enable.append(1 if enable else 0)
        self.col_data.append(data)

    def add_body(self, gtype, params, pos, quat=(0, 0, 0, 1), density=1000.0, vel=(0, 0, 0), angvel=(0, 0, 0),
                 material=(0.6, 0.0, 0.0), world=0):
        a = self.add_actor(pos, quat, world)
        mass = density * true_volume(gtype, params)
        im, com, invI = mass_props(gtype, params, mass)
        self.add_dynamic(a, im, com, invI, vel, angvel)
        self.add_collider(a, gtype, params, material)
        return a

    def add_static(self, gtype, params, pos, quat=(0, 0, 0, 1), material=(0.6, 0.0, 0.0), world=0):
        a = self.add_actor(pos, quat, world)
        self.add_collider(a, gtype, params, material)
        return a

    def add_joint(self, jtype, a0, a1, p0, p1, q0=(0, 0, 0, 1), q1=(0, 0, 0, 1), drive=None):
        """Scene::createJoint<T>(entity0, anchor0Pos, anchor0Or, entity1, anchor1Pos, anchor1Or);
        anchors in body space. drive = (velocity, max torque) enables a revolute drive."""
        an0 = list(p0) + list(q0)
        an1 = list(p1) + list(q1)
        d = (1, drive[0], drive[1]) if drive else (0, 0.0, 0.0)
        self.joints.append((jtype, a0, a1, an0, an1) + d)

    def joint_array(self, cls):
        """ctypes array of joint descriptors (same field layout for the oracle and the C ABI)."""
        arr = (cls * max(1, len(self.joints)))()
        for i, (t, a0, a1, an0, an1, de, dv, dm) in enumerate(self.joints):
            arr[i].type, arr[i].actor0, arr[i].actor1 = t, a0, a1
            for k in range(7):
                arr[i].anchor0[k] = an0[k]
                arr[i].anchor1[k] = an1[k]
            arr[i].drive_enabled, arr[i].drive_velocity, arr[i].drive_max_torque = de, dv, dm
        return arr

    # -- views
    @property
    def n_actors(self):
        return len(self.entity)

    @property
    def n_dyn(self):
        return len(self.actor_of_dyn)

    @property
    def n_colliders(self):
        return len(self.col_actor)

    def arrays(self):
        if self._frozen is None:
            def f(x, shape, dt=np.float32):
                return np.ascontiguousarray(np.asarray(x, dt).reshape(shape)) if len(x) else np.zeros(shape, dt)
            A, D, C = self.n_actors, self.n_dyn, self.n_colliders
            self._frozen = dict(
                entity=f(self.entity, (A,), np.uint32), pos=f(self.pos, (A, 3)), quat=f(self.quat, (A, 4)),
                dyn_of_actor=f(self.dyn_of_actor, (A,), np.int32), world=f(self.world, (A,), np.int32),
                actor_of_dyn=f(self.actor_of_dyn, (D,), np.int32), kinematic=f(self.kinematic, (D,), np.uint8),
                vel=f(self.vel, (D, 3)), angvel=f(self.angvel, (D, 3)), inv_mass=f(self.inv_mass, (D,)),
                com=f(self.com, (D, 3)), inv_inertia=f(self.inv_inertia, (D, 9)),
                col_actor=f(self.col_actor, (C,), np.int32), col_index=f(self.col_index, (C,), np.int32),
                col_pos=f(self.col_pos, (C, 3)), col_quat=f(self.col_quat, (C, 4)),
                col_type=f(self.col_type, (C,), np.int32), col_params=f(self.col_params, (C, 4)),
                col_material=f(self.col_material, (C, 3)), col_trigger=f(self.col_trigger, (C,), np.uint8),
                col_enable=f(self.col_enable, (C,), np.uint8), col_data=f(self.col_data, (C,), np.int32))
        return self._frozen

    def struct(self, cls):
        """Fill a ctypes Structure whose field names match the arrays (see tests/oracle_lib.py
        and physecs_amd/api.py)."""
        import ctypes
        arr = self.arrays()
        kw = {}
        for name, ctype in cls._fields_:
            if name == "n_actors":
                kw[name] = self.n_actors
            elif name == "n_dyn":
                kw[name] = self.n_dyn
            elif name == "n_colliders":
                kw[name] = self.n_colliders
            elif name == "n_worlds":
                kw[name] = self.n_worlds
            elif name == "actor_world":
                kw[name] = arr["world"].ctypes.data_as(ctype)
            else:
                kw[name] = arr[name].ctypes.data_as(ctype)
        return cls(**kw)

    def world_subscene(self, w):
        """The actors, dynamics and colliders of world w as a standalone Scene (same order)."""
        arr = self.arrays()
        s = Scene()
        amap = {}
        for a in range(self.n_actors):
            if arr["world"][a] != w:
                continue
            amap[a] = s.add_actor(arr["pos"][a], arr["quat"][a], 0, entity=int(arr["entity"][a]))
        for d in range(self.n_dyn):
            a = int(arr["actor_of_dyn"][d])
            if a in amap:
                s.add_dynamic(amap[a], arr["inv_mass"][d], arr["com"][d], arr["inv_inertia"][d], arr["vel"][d],
                              arr["angvel"][d], bool(arr["kinematic"][d]))
        for c in range(self.n_colliders):
            a = int(arr["col_actor"][c])
            if a in amap:
                s.add_collider(amap[a], int(arr["col_type"][c]), arr["col_params"][c][:3], arr["col_material"][c],
                               arr["col_pos"][c], arr["col_quat"][c], bool(arr["col_trigger"][c]),
                               bool(arr["col_enable"][c]), int(arr["col_data"][c]))
        return s


# ---------------------------------------------------------------- configurations
def box_pyramid(seed=0, rows=20, jitter=0.01):
    """20-layer box pyramid: 210 dynamic 1x1x1 m boxes (density 1000, friction 0.6,
    restitution 0) on a static 100x1x100 m ground box; per-box xz jitter U(-j, j)."""
    rng = np.random.default_rng(seed)
    s = Scene()
    s.add_static(BOX, [50, 0.5, 50], [0, -0.5, 0])
    for k in range(rows):
        for i in range(rows - k):
            x = (i - (rows - 1 - k) / 2.0) * 1.0 + rng.uniform(-jitter, jitter)
            z = rng.uniform(-jitter, jitter)
            s.add_body(BOX, [0.5, 0.5, 0.5], [x, 0.5 + k, z])
    return s


def _mixed_shape(rng, kind):
    if kind == SPHERE:
        return [rng.uniform(0.2, 0.5)]
    if kind == CAPSULE:
        r = rng.uniform(0.15, 0.3)
        hh = rng.uniform(0.2, 0.6)
        return [hh, r]
    return list(rng.uniform(0.2, 0.6, 3))


def container(s, half=25.0, height=60.0, world=0):
    """Open container of 5 static boxes around [-half, half]^2: floor + four 1 m walls."""
    s.add_static(BOX, [half + 1, 0.5, half + 1], [0, -0.5, 0], world=world)
    h = height / 2
    s.add_static(BOX, [0.5, h, half + 1], [half + 0.5, h, 0], world=world)
    s.add_static(BOX, [0.5, h, half + 1], [-half - 0.5, h, 0], world=world)
    s.add_static(BOX, [half + 1, h, 0.5], [0, h, half + 0.5], world=world)
    s.add_static(BOX, [half + 1, h, 0.5], [0, h, -half - 0.5], world=world)


def mixed_lattice(seed=0, side=32, spacing=1.5, jitter=0.05, with_container=True):
    """side^3 dynamic bodies, 1/3 each of spheres, capsules and boxes (seeded shuffle),
    on a jittered lattice at `spacing`, uniform random orientations, v ~ N(0,1)^3,
    inside an open container of 5 static boxes (32^3 = 32,768 bodies by default)."""
    rng = np.random.default_rng(seed)
    n = side ** 3
    kinds = np.array([SPHERE, CAPSULE, BOX] * (n // 3 + 1))[:n]
    rng.shuffle(kinds)
    quats = random_quats(rng, n)
    vels = rng.normal(size=(n, 3)).astype(np.float32)
    jit = rng.uniform(-jitter, jitter, size=(n, 3))
    s = Scene()
    if with_container:
        half = max(25.0, (side - 1) * spacing / 2 + 1.75)
        container(s, half=half)
    off = (side - 1) * spacing / 2
    i = 0
    for iy in range(side):
        for iz in range(side):
            for ix in range(side):
                p = [ix * spacing - off + jit[i, 0], 1.0 + iy * spacing + jit[i, 1], iz * spacing - off + jit[i, 2]]
                s.add_body(int(kinds[i]), _mixed_shape(rng, int(kinds[i])), p, quats[i], vel=vels[i])
                i += 1
    return s


def batched_worlds(seed=0, n_worlds=4096, lo=8, hi=1024, spacing=1.5):
    """n_worlds independent worlds with body counts floor(exp(U(ln lo, ln hi))), each a
    seeded mix of the mixed-lattice shapes on a lattice above its own static ground box.
    All worlds share coordinates; the world id keeps them apart."""
    rng = np.random.default_rng(seed)
    s = Scene()
    counts = np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), n_worlds))).astype(int)
    for w in range(n_worlds):
        n = int(counts[w])
        side = int(math.ceil(n ** (1 / 3)))
        half = (side - 1) * spacing / 2 + 1.0
        s.add_static(BOX, [half + 1, 0.5, half + 1], [0, -0.5, 0], world=w)
        kinds = rng.integers(0, 3, n)
        quats = random_quats(rng, n)
        vels = rng.normal(size=(n, 3)).astype(np.float32)
        off = (side - 1) * spacing / 2
        for i in range(n):
            ix, iz, iy = i % side, (i // side) % side, i // (side * side)
            p = [ix * spacing - off, 1.0 + iy * spacing, iz * spacing - off]
            s.add_body(int(kinds[i]), _mixed_shape(rng, int(kinds[i])), p, quats[i], vel=vels[i], world=w)
    return s


def capsule_chain(s, origin, n=8, radius=0.1, half_height=0.3, axis="x", world=0, material=(0.6, 0.0, 0.0),
                  drive=None):
    """n capsules end to end, joined by n-1 revolute joints at their touching ends. The
    capsules lie along world x (axis="x") or hang along -y (axis="y")."""
    seg = 2 * (half_height + radius)
    if axis == "x":
        q = (0.0, 0.0, -math.sqrt(0.5), math.sqrt(0.5))   # local +y -> world +x
        step = np.array([seg, 0.0, 0.0])
    else:
        q = (0.0, 0.0, 0.0, 1.0)
        step = np.array([0.0, -seg, 0.0])
    actors = []
    for i in range(n):
        p = np.asarray(origin, float) + i * step
        actors.append(s.add_body(CAPSULE, [half_height, radius], p, q, world=world, material=material))
    end = half_height + radius
    for i in range(n - 1):
        if axis == "x":
            s.add_joint(REVOLUTE, actors[i], actors[i + 1], (0, end, 0), (0, -end, 0), drive=drive)
        else:
            s.add_joint(REVOLUTE, actors[i], actors[i + 1], (0, -end, 0), (0, end, 0), drive=drive)
    return actors


def ragdoll(s, origin, world=0, material=(0.6, 0.0, 0.0)):
    """A 15-capsule ragdoll with 14 joints (spherical, universal, revolute)."""
    o = np.asarray(origin, float)
    r, hh = 0.09, 0.12
    parts = {}

    def body(name, p, horizontal=False, rr=r, h=hh):
        q = (0.0, 0.0, -math.sqrt(0.5), math.sqrt(0.5)) if horizontal else (0.0, 0.0, 0.0, 1.0)
        parts[name] = s.add_body(CAPSULE, [h, rr], o + np.asarray(p, float), q, world=world, material=material)

    body("pelvis", (0, 1.0, 0)); body("abdomen", (0, 1.3, 0)); body("chest", (0, 1.6, 0))
    body("neck", (0, 1.86, 0), h=0.03, rr=0.05); body("head", (0, 2.05, 0), rr=0.11, h=0.04)
    for sgn, side in ((1, "l"), (-1, "r")):
        body("uarm" + side, (sgn * 0.42, 1.7, 0), horizontal=True)
        body("farm" + side, (sgn * 0.84, 1.7, 0), horizontal=True)
        body("hand" + side, (sgn * 1.13, 1.7, 0), horizontal=True, h=0.03, rr=0.06)
        body("thigh" + side, (sgn * 0.15, 0.62, 0))
        body("shin" + side, (sgn * 0.15, 0.2, 0))
    J = s.add_joint
    J(UNIVERSAL, parts["pelvis"], parts["abdomen"], (0, 0.15, 0), (0, -0.15, 0))
    J(UNIVERSAL, parts["abdomen"], parts["chest"], (0, 0.15, 0), (0, -0.15, 0))
    J(SPHERICAL, parts["chest"], parts["neck"], (0, 0.21, 0), (0, -0.05, 0))
    J(SPHERICAL, parts["neck"], parts["head"], (0, 0.05, 0), (0, -0.14, 0))
    for sgn, side in ((1, "l"), (-1, "r")):
        # horizontal capsules: local +y points along world x
        J(SPHERICAL, parts["chest"], parts["uarm" + side], (sgn * 0.21, 0.1, 0), (0, -sgn * 0.21, 0))
        J(REVOLUTE, parts["uarm" + side], parts["farm" + side], (0, sgn * 0.21, 0), (0, -sgn * 0.21, 0))
        J(UNIVERSAL, parts["farm" + side], parts["hand" + side], (0, sgn * 0.21, 0), (0, -sgn * 0.08, 0))
        J(SPHERICAL, parts["pelvis"], parts["thigh" + side], (sgn * 0.15, -0.1, 0), (0, 0.21, 0))
        J(REVOLUTE, parts["thigh" + side], parts["shin" + side], (0, -0.21, 0), (0, 0.21, 0))
    return parts


def joint_scene(seed=0, chains=6, ragdolls=2, ground=True):
    """Capsule chains and ragdolls dropped on a ground box, for joint parity tests."""
    rng = np.random.default_rng(seed)
    s = Scene()
    if ground:
        s.add_static(BOX, [30, 0.5, 30], [0, -0.5, 0])
    for c in range(chains):
        horizontal = c % 2 == 0
        top = rng.uniform(1.0, 4.0) if horizontal else rng.uniform(7.0, 9.0)   # a hanging chain is 6.4 m long
        capsule_chain(s, [rng.uniform(-6, 3), top, rng.uniform(-6, 6)], n=8, axis="x" if horizontal else "y")
    for r in range(ragdolls):
        ragdoll(s, [rng.uniform(-5, 5), rng.uniform(0.5, 2.0), rng.uniform(-5, 5)])
    return s


def random_cluster(seed=0, n=60, extent=3.0, kinds=(SPHERE, CAPSULE, BOX), ground=True):
    """Small dense random scene for parity tests: n bodies of mixed shapes in a box of
    side `extent` above a ground box, random orientations and velocities."""
    rng = np.random.default_rng(seed)
    s = Scene()
    if ground:
        s.add_static(BOX, [20, 0.5, 20], [0, -0.5, 0])
    for i in range(n):
        k = int(rng.choice(kinds))
        p = [rng.uniform(-extent, extent), rng.uniform(0.3, 0.3 + 2 * extent), rng.uniform(-extent, extent)]
        s.add_body(k, _mixed_shape(rng, k), p, random_quats(rng, 1)[0], vel=rng.normal(size=3))
    return s
