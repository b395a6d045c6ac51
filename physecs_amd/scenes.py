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
import ctypes
import math

import numpy as np

SPHERE, CAPSULE, BOX, CONVEX_MESH, TRIANGLE_MESH = range(5)
SPHERICAL, UNIVERSAL, REVOLUTE, FIXED, PRISMATIC, GEAR, SERVO, CUSTOM = range(8)
# Constraint1D flags (include/Physecs/Constraint1D.h:11-16)
ROW_NONE, ROW_SOFT, ROW_ANGULAR, ROW_LIMITED = 0, 1, 2, 4

# Defaults of the joints' data members (include/Physecs/Joints/*.h)
JOINT_DEFAULTS = {
    REVOLUTE: dict(drive_enabled=0, drive_velocity=0.0, drive_max_torque=0.0),
    PRISMATIC: dict(upper_limit=1.0, lower_limit=0.0, drive_enabled=0, target=0.0, drive_stiffness=5.0,
                    drive_damping=1.0),
    GEAR: dict(gear_ratio=1.0),
    SERVO: dict(target=0.0, drive_stiffness=30.0, drive_damping=1.0),
}
JOINT_FIELDS = ("drive_enabled", "drive_velocity", "drive_max_torque", "lower_limit", "upper_limit", "target",
                "drive_stiffness", "drive_damping", "gear_ratio")


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


def convex_hull(points):
    """ConvexMesh data for the convex hull of `points` (scipy's Qhull): vertices (the hull's
    points only), faces as polygons of coplanar hull facets ordered clockwise seen from
    outside, outward unit normals and face centroids (vertex means)."""
    from scipy.spatial import ConvexHull
    pts = np.asarray(points, np.float64)
    h = ConvexHull(pts)
    used = np.unique(h.simplices)
    remap = {int(v): i for i, v in enumerate(used)}
    verts = pts[used]
    inner = verts.mean(axis=0)
    groups = {}
    for simplex, eq in zip(h.simplices, h.equations):
        key = tuple(np.round(eq, 6))
        groups.setdefault(key, set()).update(int(v) for v in simplex)
    faces, normals, centroids = [], [], []
    for key, vs in groups.items():
        n = np.asarray(key[:3], np.float64)
        n /= np.linalg.norm(n)
        idx = [remap[v] for v in vs]
        c = verts[idx].mean(axis=0)
        # order around the face; clockwise seen from outside = counter-clockwise about -n
        u = verts[idx[0]] - c
        u /= np.linalg.norm(u)
        w = np.cross(n, u)
        ang = [math.atan2(np.dot(verts[i] - c, w), np.dot(verts[i] - c, u)) for i in idx]
        order = [idx[k] for k in np.argsort(ang)[::-1]]
        if np.dot(n, c - inner) < 0:
            n = -n
        faces.append(order)
        normals.append(n)
        centroids.append(c)
    return verts.astype(np.float32), faces, np.asarray(normals, np.float32), np.asarray(centroids, np.float32)


def ellipsoid_hull(rng, n_points=24, lo=0.3, hi=0.8):
    """Hull of n_points uniform on an ellipsoid with semi-axes U(lo, hi)^3."""
    axes = rng.uniform(lo, hi, 3)
    d = rng.normal(size=(n_points, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return convex_hull(d * axes)


def hull_volume(verts, faces):
    c = verts.mean(axis=0)
    vol = 0.0
    for f in faces:
        for i in range(1, len(f) - 1):
            vol += abs(np.dot(np.cross(verts[f[0]] - c, verts[f[i]] - c), verts[f[i + 1]] - c)) / 6.0
    return vol


def _fan(verts, faces, center):
    """(v0, v1, v2) arrays of the fan triangulation of every face, in face order."""
    a, b, c = [], [], []
    for f in faces:
        for i in range(1, len(f) - 1):
            a.append(f[0]); b.append(f[i]); c.append(f[i + 1])
    return verts[a], verts[b], verts[c]


def _inertia_tetrahedra(m, v):
    """src/MassUtil.cpp:31-72 for a batch: m (k,), v (k, 4, 3) -> (k, 3, 3) about the origin."""
    d = np.zeros((len(m), 3))
    for i in range(4):
        for j in range(i + 1):
            d += v[:, j] * v[:, i]
    a, b, c = (d[:, 1] + d[:, 2]) / 10, (d[:, 0] + d[:, 2]) / 10, (d[:, 0] + d[:, 1]) / 10
    ap = np.zeros(len(m)); bp = np.zeros(len(m)); cp = np.zeros(len(m))
    for i in range(4):
        for j in range(4):
            k = 2 if i == j else 1
            ap += k * v[:, i, 1] * v[:, j, 2]
            bp += k * v[:, i, 0] * v[:, j, 2]
            cp += k * v[:, i, 0] * v[:, j, 1]
    ap, bp, cp = ap / 20, bp / 20, cp / 20
    I = np.zeros((len(m), 3, 3))
    I[:, 0, 0], I[:, 1, 1], I[:, 2, 2] = a, b, c
    I[:, 0, 1] = I[:, 1, 0] = -bp
    I[:, 0, 2] = I[:, 2, 0] = -cp
    I[:, 1, 2] = I[:, 2, 1] = -ap
    return m[:, None, None] * I


def convex_mass_props(verts, faces, scale, mass):
    """(com, invInertia column-major) for one convex collider at the body origin, as
    src/MassUtil.cpp:96-123,156-177 computes them: the fan centre is the mean over the
    vertex buffer padded to a multiple of 4 with the last vertex, divided by the
    unpadded count."""
    sv = np.asarray(verts, np.float64) * np.asarray(scale, np.float64)
    n = len(sv)
    padded = np.concatenate([sv, np.repeat(sv[-1:], (-n) % 4, axis=0)])
    center = padded.sum(axis=0) / n
    v0, v1, v2 = _fan(sv, faces, center)
    vol = np.abs(np.einsum("ij,ij->i", np.cross(v0 - center, v1 - center), v2 - center)) / 6.0
    total = vol.sum()
    com = ((center + v0 + v1 + v2) / 4.0 * vol[:, None]).sum(axis=0) / total
    d = -com
    tets = np.stack([np.broadcast_to(center + d, v0.shape), v0 + d, v1 + d, v2 + d], axis=1)
    I = _inertia_tetrahedra(mass * vol / total, tets).sum(axis=0)
    inv = np.linalg.inv(I)
    return com.astype(np.float32), inv.T.astype(np.float32).ravel()


_FIELDS = {
    # name: (dtype, per-row shape)
    "entity": (np.uint32, ()), "pos": (np.float32, (3,)), "quat": (np.float32, (4,)),
    "dyn_of_actor": (np.int32, ()), "world": (np.int32, ()),
    "actor_of_dyn": (np.int32, ()), "kinematic": (np.uint8, ()), "vel": (np.float32, (3,)),
    "angvel": (np.float32, (3,)), "inv_mass": (np.float32, ()), "com": (np.float32, (3,)),
    "inv_inertia": (np.float32, (9,)),
    "col_actor": (np.int32, ()), "col_index": (np.int32, ()), "col_pos": (np.float32, (3,)),
    "col_quat": (np.float32, (4,)), "col_type": (np.int32, ()), "col_params": (np.float32, (4,)),
    "col_material": (np.float32, (3,)), "col_trigger": (np.uint8, ()), "col_enable": (np.uint8, ()),
    "col_data": (np.int32, ()),
}


class Scene:
    """Scene descriptor builder. Actors, dynamic storage and colliders are appended in
    creation order; entity ids are the creation index unless given. Fields are kept as
    lists of numpy chunks (one row per single append, many rows per bulk append)."""

    def __init__(self):
        self._chunks = {k: [] for k in _FIELDS}
        self._n = {"actor": 0, "dyn": 0, "col": 0}
        self._dyn_patch = {}          # actor -> dyn index, for actors added before their dynamic part
        self.non_colliding = []
        self.joints = []        # (type, actor0, actor1, anchor0[7], anchor1[7], params dict)
        self._ncol = {}
        self.n_worlds = 1
        self._frozen = None
        self.convex = []        # (verts (n,3), faces [list of vertex indices], normals (f,3), centroids (f,3))
        self.trimesh = []       # (verts (n,3), tris (t,3))

    def _put(self, name, rows):
        dt, shp = _FIELDS[name]
        a = np.asarray(rows, dt).reshape((-1,) + shp)
        self._chunks[name].append(a)
        self._frozen = None

    # -- building
    def add_actor(self, pos, quat=(0, 0, 0, 1), world=0, entity=None):
        a = self._n["actor"]
        self._n["actor"] += 1
        self._put("entity", [a if entity is None else entity])
        self._put("pos", pos)
        self._put("quat", quat)
        self._put("dyn_of_actor", [-1])
        self._put("world", [world])
        self.n_worlds = max(self.n_worlds, world + 1)
        return a

    def add_dynamic(self, actor, inv_mass, com, inv_inertia, vel=(0, 0, 0), angvel=(0, 0, 0), kinematic=False):
        d = self._n["dyn"]
        self._n["dyn"] += 1
        self._dyn_patch[actor] = d
        self._put("actor_of_dyn", [actor])
        self._put("kinematic", [1 if kinematic else 0])
        self._put("vel", vel)
        self._put("angvel", angvel)
        self._put("inv_mass", [inv_mass])
        self._put("com", com)
        self._put("inv_inertia", np.asarray(inv_inertia, np.float32).ravel())
        return d

    def add_collider(self, actor, gtype, params, material=(0.6, 0.0, 0.0), lpos=(0, 0, 0), lquat=(0, 0, 0, 1),
                     trigger=False, enable=True, data=0):
        idx = self._ncol.get(actor, 0)
        self._ncol[actor] = idx + 1
        self._n["col"] += 1
        p = np.zeros(4, np.float32)
        p[: len(params)] = params
        self._put("col_actor", [actor])
        self._put("col_index", [idx])
        self._put("col_pos", lpos)
        self._put("col_quat", lquat)
        self._put("col_type", [gtype])
        self._put("col_params", p)
        self._put("col_material", material)
        self._put("col_trigger", [1 if trigger else 0])
        self._put("col_enable", [1 if enable else 0])
        self._put("col_data", [data])

    def add_body(self, gtype, params, pos, quat=(0, 0, 0, 1), density=1000.0, vel=(0, 0, 0), angvel=(0, 0, 0),
                 material=(0.6, 0.0, 0.0), world=0, data=0, trigger=False):
        a = self.add_actor(pos, quat, world)
        mass = density * true_volume(gtype, params)
        im, com, invI = mass_props(gtype, params, mass)
        self.add_dynamic(a, im, com, invI, vel, angvel)
        self.add_collider(a, gtype, params, material, data=data, trigger=trigger)
        return a

    def add_static(self, gtype, params, pos, quat=(0, 0, 0, 1), material=(0.6, 0.0, 0.0), world=0, data=0,
                   trigger=False):
        a = self.add_actor(pos, quat, world)
        self.add_collider(a, gtype, params, material, data=data, trigger=trigger)
        return a

    def add_bodies_bulk(self, gtypes, params, pos, quats, vels, worlds, density=1000.0, material=(0.6, 0.0, 0.0)):
        """Appends many one-collider dynamic bodies at once (the same fields add_body writes)."""
        n = len(gtypes)
        gtypes = np.asarray(gtypes, np.int32)
        params = np.asarray(params, np.float64).reshape(n, 4)
        a0, d0 = self._n["actor"], self._n["dyn"]
        actors = np.arange(a0, a0 + n)
        self._put("entity", actors.astype(np.uint32))
        self._put("pos", pos)
        self._put("quat", quats)
        self._put("dyn_of_actor", np.arange(d0, d0 + n))
        self._put("world", worlds)
        self.n_worlds = max(self.n_worlds, int(np.max(worlds)) + 1 if n else 1)
        vol = np.where(gtypes == SPHERE, 4.0 / 3.0 * np.pi * params[:, 0] ** 3,
                       np.where(gtypes == CAPSULE, np.pi * params[:, 1] ** 2 * 2 * params[:, 0]
                                + 4.0 / 3.0 * np.pi * params[:, 1] ** 3, 8.0 * params[:, 0] * params[:, 1] * params[:, 2]))
        m = density * vol
        r = params[:, 0]
        isph = 2.0 * m * r * r / 5.0
        cr, ch = params[:, 1], params[:, 0]
        vc = 2 * cr * cr * np.pi * ch
        vs = 4 * cr ** 3 * np.pi / 3
        with np.errstate(invalid="ignore", divide="ignore"):   # rows that are not capsules
            mc, ms = vc * m / (vc + vs), vs * m / (vc + vs)
        cxx = mc * (ch * ch / 3 + cr * cr / 4) + ms * (ch * ch + 3 * ch * cr / 4 + 2 * cr * cr / 5)
        cyy = mc * cr * cr / 2 + ms * 2 * cr * cr / 5
        x2, y2, z2 = params[:, 0] ** 2, params[:, 1] ** 2, params[:, 2] ** 2
        bxx, byy, bzz = m / 12 * (y2 + z2), m / 12 * (x2 + z2), m / 12 * (x2 + y2)
        ixx = np.where(gtypes == SPHERE, isph, np.where(gtypes == CAPSULE, cxx, bxx))
        iyy = np.where(gtypes == SPHERE, isph, np.where(gtypes == CAPSULE, cyy, byy))
        izz = np.where(gtypes == SPHERE, isph, np.where(gtypes == CAPSULE, cxx, bzz))
        invI = np.zeros((n, 9))
        invI[:, 0], invI[:, 4], invI[:, 8] = 1 / ixx, 1 / iyy, 1 / izz
        self._put("actor_of_dyn", actors)
        self._put("kinematic", np.zeros(n))
        self._put("vel", vels)
        self._put("angvel", np.zeros((n, 3)))
        self._put("inv_mass", 1.0 / m)
        self._put("com", np.zeros((n, 3)))
        self._put("inv_inertia", invI)
        self._put("col_actor", actors)
        self._put("col_index", np.zeros(n))
        self._put("col_pos", np.zeros((n, 3)))
        self._put("col_quat", np.tile([0, 0, 0, 1], (n, 1)))
        self._put("col_type", gtypes)
        self._put("col_params", params)
        self._put("col_material", np.tile(material, (n, 1)))
        self._put("col_trigger", np.zeros(n))
        self._put("col_enable", np.ones(n))
        self._put("col_data", np.zeros(n))
        for a in range(a0, a0 + n):
            self._ncol[a] = 1
        self._n["actor"] += n
        self._n["dyn"] += n
        self._n["col"] += n
        return actors

    def add_convex_mesh(self, verts, faces, normals, centroids):
        """Registers a ConvexMesh (faces clockwise seen from outside); returns its index."""
        self.convex.append((np.asarray(verts, np.float32).reshape(-1, 3), [list(map(int, f)) for f in faces],
                            np.asarray(normals, np.float32).reshape(-1, 3),
                            np.asarray(centroids, np.float32).reshape(-1, 3)))
        self._frozen = None
        return len(self.convex) - 1

    def add_triangle_mesh(self, verts, tris):
        """Registers a static TriangleMesh; returns its index."""
        self.trimesh.append((np.asarray(verts, np.float32).reshape(-1, 3), np.asarray(tris, np.int32).reshape(-1, 3)))
        self._frozen = None
        return len(self.trimesh) - 1

    def add_convex_body(self, mesh, pos, quat=(0, 0, 0, 1), scale=(1, 1, 1), density=1000.0, vel=(0, 0, 0),
                        angvel=(0, 0, 0), material=(0.6, 0.0, 0.0), world=0):
        """A dynamic body with one convex-mesh collider; mass = density x hull volume,
        centre of mass and inertia from src/MassUtil.cpp's convex branch."""
        verts, faces, _, _ = self.convex[mesh]
        a = self.add_actor(pos, quat, world)
        mass = density * hull_volume(verts * np.asarray(scale, np.float64), faces)
        com, invI = convex_mass_props(verts, faces, scale, mass)
        self.add_dynamic(a, np.float32(1.0 / mass), com, invI, vel, angvel)
        self.add_collider(a, CONVEX_MESH, [scale[0], scale[1], scale[2], float(mesh)], material)
        return a

    def add_joint(self, jtype, a0, a1, p0, p1, q0=(0, 0, 0, 1), q1=(0, 0, 0, 1), drive=None, **params):
        """Scene::createJoint<T>(entity0, anchor0Pos, anchor0Or, entity1, anchor1Pos, anchor1Or);
        anchors in body space. drive = (velocity, max torque) enables a revolute drive; other
        keyword arguments set the joint's data members (JOINT_FIELDS), defaults as the
        reference's. Returns the joint's index."""
        an0 = list(p0) + list(q0)
        an1 = list(p1) + list(q1)
        p = dict(JOINT_DEFAULTS.get(jtype, {}))
        if drive:
            p.update(drive_enabled=1, drive_velocity=drive[0], drive_max_torque=drive[1])
        for k, v in params.items():
            if k not in JOINT_FIELDS:
                raise ValueError(f"unknown joint parameter {k}")
            p[k] = v
        self.joints.append((jtype, a0, a1, an0, an1, p))
        return len(self.joints) - 1

    @staticmethod
    def fill_joint(d, t, a0, a1, an0, an1, p):
        """Fills one ctypes joint descriptor (the same layout for the oracle and the C ABI)."""
        d.type, d.actor0, d.actor1 = t, a0, a1
        for k in range(7):
            d.anchor0[k] = an0[k]
            d.anchor1[k] = an1[k]
        for f in JOINT_FIELDS:
            setattr(d, f, p.get(f, 0))

    def joint_array(self, cls):
        """ctypes array of joint descriptors (same field layout for the oracle and the C ABI)."""
        arr = (cls * max(1, len(self.joints)))()
        for i, j in enumerate(self.joints):
            self.fill_joint(arr[i], *j)
        return arr

    def rebuilt(self, state=None, clear=(), add=()):
        """The scene as Scene::clearColliders and Scene::addCollider leave it
        (src/Physecs.cpp:725-752), for a world created anew: poses and velocities from `state`
        ((pos, quat, vel, angvel) as World.state() gives them), the colliders of each actor in
        `clear` removed, then each (actor, gtype, params, keyword arguments of add_collider) of
        `add` appended with its entity's next collider index. The colliders are then ordered
        as an entt::registry holds them when each actor's RigidBodyCollisionComponent was
        emplaced in the order the actors first appear here: by that order, then by index
        within the entity. Actors, bodies, joints, non-colliding pairs and meshes are kept.
        Returns (scene, collider map): the map gives each old collider's index in the new
        scene, or -1."""
        import copy
        arr = self.arrays()
        s = copy.copy(self)
        s._chunks = {k: [np.array(v, copy=True)] for k, v in arr.items() if k in _FIELDS}
        s._dyn_patch = {}
        s._frozen = None
        s._n = dict(self._n)
        s.non_colliding = list(self.non_colliding)
        s.joints = list(self.joints)
        if state is not None:
            for name, v in zip(("pos", "quat", "vel", "angvel"), state):
                s._chunks[name] = [np.ascontiguousarray(v, np.float32).reshape(s._chunks[name][0].shape)]
        keep = ~np.isin(arr["col_actor"], np.asarray(list(clear), np.int32))
        for k in _FIELDS:
            if k.startswith("col_"):
                s._chunks[k] = [s._chunks[k][0][keep]]
        s._n["col"] = int(keep.sum())
        colmap = np.full(self.n_colliders, -1, np.int64)
        colmap[keep] = np.arange(s._n["col"])
        actors = s._chunks["col_actor"][0]
        s._ncol = {int(a): int(c) for a, c in zip(*np.unique(actors, return_counts=True))}
        for a in clear:
            s._ncol[int(a)] = 0
        for actor, gtype, params, kw in add:
            s.add_collider(actor, gtype, params, **kw)
        new = s.arrays()
        rank = np.full(self.n_actors, self.n_actors, np.int64)
        first = np.unique(arr["col_actor"], return_index=True)
        rank[first[0]] = np.argsort(np.argsort(first[1]))
        order = np.lexsort((new["col_index"], np.where(rank[new["col_actor"]] < self.n_actors,
                                                        rank[new["col_actor"]], self.n_actors + new["col_actor"])))
        for k in _FIELDS:
            if k.startswith("col_"):
                s._chunks[k] = [new[k][order].copy()]
        s._frozen = None
        inv = np.empty_like(order)
        inv[order] = np.arange(len(order))
        colmap[keep] = inv[colmap[keep]]
        return s, colmap

    # -- views
    @property
    def n_actors(self):
        return self._n["actor"]

    @property
    def n_dyn(self):
        return self._n["dyn"]

    @property
    def n_colliders(self):
        return self._n["col"]

    @property
    def col_material(self):
        """Writable view of the colliders' materials (friction, restitution, damping)."""
        return self.arrays()["col_material"]

    @property
    def col_enable(self):
        return self.arrays()["col_enable"]

    def arrays(self):
        if self._frozen is None:
            out = {}
            for k, (dt, shp) in _FIELDS.items():
                ch = self._chunks[k]
                out[k] = np.ascontiguousarray(np.concatenate(ch) if ch else np.zeros((0,) + shp, dt))
            for a, d in self._dyn_patch.items():
                out["dyn_of_actor"][a] = d
            out.update(self._mesh_arrays())
            self._chunks = {k: [v] for k, v in out.items()}   # keep the frozen arrays as the single chunk
            self._dyn_patch = {}
            self._frozen = out
        return self._frozen

    def _mesh_arrays(self):
        cvs, cfs, fis, fi, fn, fc = [0], [0], [0], [], [], []
        verts = []
        for v, faces, normals, centroids in self.convex:
            verts.append(v)
            cvs.append(cvs[-1] + len(v))
            cfs.append(cfs[-1] + len(faces))
            for f in faces:
                fi.extend(f)
                fis.append(fis[-1] + len(f))
            fn.append(normals)
            fc.append(centroids)
        tvs, ts, tv, ti = [0], [0], [], []
        for v, t in self.trimesh:
            tv.append(v)
            ti.append(t)
            tvs.append(tvs[-1] + len(v))
            ts.append(ts[-1] + len(t))
        f32 = lambda xs: np.ascontiguousarray(np.concatenate(xs) if xs else np.zeros((1, 3)), np.float32)
        i32 = lambda xs: np.ascontiguousarray(xs if len(xs) else [0], np.int32)
        return {
            "convex_vert_start": i32(cvs), "convex_verts": f32(verts), "convex_face_start": i32(cfs),
            "face_idx_start": i32(fis), "face_idx": i32(fi), "face_normal": f32(fn), "face_centroid": f32(fc),
            "tri_vert_start": i32(tvs), "tri_verts": f32(tv), "tri_start": i32(ts),
            "tri_idx": np.ascontiguousarray(np.concatenate(ti) if ti else np.zeros((1, 3)), np.int32),
        }

    def struct(self, cls):
        """Fill a ctypes Structure whose field names match the arrays (see tests/oracle_lib.py
        and physecs_amd/api.py)."""
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
            elif name == "n_convex":
                kw[name] = len(self.convex)
            elif name == "n_trimesh":
                kw[name] = len(self.trimesh)
            elif name == "actor_world":
                kw[name] = arr["world"].ctypes.data_as(ctype)
            elif name == "tri_bvh_cache":   # per triangle mesh: cached build bytes, or None
                caches = getattr(self, "tri_bvh_cache", None)
                if caches:
                    self._cache_bufs = [None if c is None else ctypes.create_string_buffer(bytes(c), len(c))
                                        for c in caches]
                    ptrs = (ctypes.c_void_p * len(caches))(
                        *[None if b is None else ctypes.cast(b, ctypes.c_void_p).value for b in self._cache_bufs])
                    self._cache_ptrs = ptrs
                    kw[name] = ctypes.cast(ptrs, ctype)
                else:
                    kw[name] = None
            elif name == "tri_bvh_cache_bytes":
                caches = getattr(self, "tri_bvh_cache", None)
                if caches:
                    self._cache_lens = (ctypes.c_size_t * len(caches))(*[0 if c is None else len(c) for c in caches])
                    kw[name] = ctypes.cast(self._cache_lens, ctype)
                else:
                    kw[name] = None
            else:
                kw[name] = arr[name].ctypes.data_as(ctype)
        return cls(**kw)

    def world_subscene(self, w):
        """The actors, dynamics, colliders and joints of world w as a standalone Scene, in the
        same relative order and with the same entity ids."""
        arr = self.arrays()
        s = Scene()
        s.convex = list(self.convex)
        s.trimesh = list(self.trimesh)
        amap = {}
        for a in np.nonzero(arr["world"] == w)[0]:
            amap[int(a)] = s.add_actor(arr["pos"][a], arr["quat"][a], 0, entity=int(arr["entity"][a]))
        for d in np.nonzero(arr["world"][arr["actor_of_dyn"]] == w)[0]:
            a = int(arr["actor_of_dyn"][d])
            s.add_dynamic(amap[a], arr["inv_mass"][d], arr["com"][d], arr["inv_inertia"][d], arr["vel"][d],
                          arr["angvel"][d], bool(arr["kinematic"][d]))
        for c in np.nonzero(arr["world"][arr["col_actor"]] == w)[0]:
            a = int(arr["col_actor"][c])
            s.add_collider(amap[a], int(arr["col_type"][c]), arr["col_params"][c][:3], arr["col_material"][c],
                           arr["col_pos"][c], arr["col_quat"][c], bool(arr["col_trigger"][c]),
                           bool(arr["col_enable"][c]), int(arr["col_data"][c]))
        for (t, a0, a1, an0, an1, p) in self.joints:
            if a0 in amap and a1 in amap:
                s.joints.append((t, amap[a0], amap[a1], an0, an1, dict(p)))
        return s


# ---------------------------------------------------------------- configurations
def contact_filter_table(scene, fn):
    """A contact filter fn(is_trigger0, data0, is_trigger1, data1) -> True for TRIGGER (the
    reference's ContactType (*)(bool, int, bool, int), include/Physecs/Physecs.h:43,174),
    tabulated over the scene's distinct (isTrigger, data) values: (classes, table, per-
    collider class) for pgx_world_set_contact_filter. Argument 0 is the collider of the
    pair's lower entity, as src/Physecs.cpp:200 calls it."""
    arr = scene.arrays()
    keys = list(zip([bool(t) for t in arr["col_trigger"].tolist()], [int(d) for d in arr["col_data"].tolist()]))
    distinct = sorted(set(keys))
    index = {k: i for i, k in enumerate(distinct)}
    classes = np.array([index[k] for k in keys], np.int32)
    k = len(distinct)
    table = np.zeros((k, k), np.uint8)
    for i, (t0, d0) in enumerate(distinct):
        for j, (t1, d1) in enumerate(distinct):
            table[i, j] = 1 if fn(t0, d0, t1, d1) else 0
    return k, np.ascontiguousarray(table), classes


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


def batched_worlds(seed=0, n_worlds=4096, lo=8, hi=1024, spacing=1.5, ragdoll_frac=0.0):
    """n_worlds independent worlds with body counts floor(exp(U(ln lo, ln hi))), each a
    seeded mix of the mixed-lattice shapes on a lattice above its own static ground box;
    a fraction of the worlds also holds a 15-capsule ragdoll standing on a static pad of
    its own, 8 m beside the ground box. (Dropped onto the pile, the light ragdoll links
    end up between bodies up to a hundred times heavier, and the reference's solver with
    its default two iterations blows such a ragdoll up within a second.) All worlds share
    coordinates; the world id keeps them apart. Worlds are laid out one after another
    (ground, bodies, then pad and ragdoll)."""
    rng = np.random.default_rng(seed)
    s = Scene()
    counts = np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), n_worlds))).astype(int)
    has_rd = rng.random(n_worlds) < ragdoll_frac
    for w in range(n_worlds):
        n = int(counts[w])
        side = int(math.ceil(n ** (1 / 3)))
        half = (side - 1) * spacing / 2 + 1.0
        s.add_static(BOX, [half + 2, 0.5, half + 2], [0, -0.5, 0], world=w)
        kinds = rng.integers(0, 3, n)
        prm = np.zeros((n, 4))
        prm[kinds == SPHERE, 0] = rng.uniform(0.2, 0.5, int(np.sum(kinds == SPHERE)))
        nc = int(np.sum(kinds == CAPSULE))
        prm[kinds == CAPSULE, 1] = rng.uniform(0.15, 0.3, nc)
        prm[kinds == CAPSULE, 0] = rng.uniform(0.2, 0.6, nc)
        prm[kinds == BOX, :3] = rng.uniform(0.2, 0.6, (int(np.sum(kinds == BOX)), 3))
        quats = random_quats(rng, n)
        vels = rng.normal(size=(n, 3))
        i = np.arange(n)
        off = (side - 1) * spacing / 2
        pos = np.stack([(i % side) * spacing - off, 1.0 + (i // (side * side)) * spacing,
                        ((i // side) % side) * spacing - off], axis=1)
        s.add_bodies_bulk(kinds, prm, pos, quats, vels, np.full(n, w))
        if has_rd[w]:
            s.add_static(BOX, [2.0, 0.5, 2.0], [half + 10.0, -0.5, 0.0], world=w)
            ragdoll(s, [half + 10.0, 0.3, 0.0], world=w)
    return s


def served_request(seed, n_bodies=None, chain=None, spacing=1.5):
    """One serving request: a seeded scene of n_bodies (log-uniform in [16, 4096] when not
    given) of the mixed-lattice shapes above its own static ground box; 10% of requests
    (or chain=True) also hold an 8-capsule revolute chain."""
    rng = np.random.default_rng(seed)
    if n_bodies is None:
        n_bodies = int(np.floor(np.exp(rng.uniform(math.log(16), math.log(4096)))))
    if chain is None:
        chain = rng.random() < 0.1
    s = Scene()
    n = n_bodies
    side = int(math.ceil(n ** (1 / 3)))
    half = (side - 1) * spacing / 2 + 1.0
    s.add_static(BOX, [half + 6, 0.5, half + 6], [0, -0.5, 0])
    kinds = rng.integers(0, 3, n)
    prm = np.zeros((n, 4))
    prm[kinds == SPHERE, 0] = rng.uniform(0.2, 0.5, int(np.sum(kinds == SPHERE)))
    nc = int(np.sum(kinds == CAPSULE))
    prm[kinds == CAPSULE, 1] = rng.uniform(0.15, 0.3, nc)
    prm[kinds == CAPSULE, 0] = rng.uniform(0.2, 0.6, nc)
    prm[kinds == BOX, :3] = rng.uniform(0.2, 0.6, (int(np.sum(kinds == BOX)), 3))
    i = np.arange(n)
    off = (side - 1) * spacing / 2
    pos = np.stack([(i % side) * spacing - off, 1.0 + (i // (side * side)) * spacing,
                    ((i // side) % side) * spacing - off], axis=1)
    pos += rng.uniform(-0.05, 0.05, pos.shape)
    s.add_bodies_bulk(kinds, prm, pos, random_quats(rng, n), rng.normal(size=(n, 3)), np.zeros(n, int))
    if chain:
        top = 1.0 + (n // (side * side) + 1) * spacing + 1.0
        capsule_chain(s, [-3.2, top, 0.0], n=8, axis="x")
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


def _value_noise(rng, n, cells):
    """Value noise on an (n x n) vertex grid from a (cells+1)^2 lattice of U(-1,1) values,
    bilinear with smoothstep weights."""
    lat = rng.uniform(-1, 1, (cells + 1, cells + 1))
    t = np.linspace(0, cells, n, endpoint=True)
    i0 = np.minimum(np.floor(t).astype(int), cells - 1)
    f = t - i0
    f = f * f * (3 - 2 * f)
    a = lat[i0][:, i0]
    b = lat[i0 + 1][:, i0]
    c = lat[i0][:, i0 + 1]
    d = lat[i0 + 1][:, i0 + 1]
    fx, fz = f[:, None], f[None, :]
    return (a * (1 - fx) + b * fx) * (1 - fz) + (c * (1 - fx) + d * fx) * fz


def terrain_mesh(seed=0, quads=256, spacing=1.0, amplitude=8.0, octaves=5, lacunarity=2.0, gain=0.5):
    """A heightfield of quads x quads cells (two triangles each, normals +y) centred on the
    origin; heights = amplitude * fBm of value noise (octaves, lacunarity, gain)."""
    rng = np.random.default_rng(seed)
    n = quads + 1
    h = np.zeros((n, n))
    amp, cells, norm = 1.0, 4, 0.0
    for _ in range(octaves):
        h += amp * _value_noise(rng, n, min(cells, quads))
        norm += amp
        amp *= gain
        cells = int(cells * lacunarity)
    h = amplitude * h / norm
    half = quads * spacing / 2
    xs = np.arange(n) * spacing - half
    verts = np.zeros((n, n, 3), np.float32)
    verts[:, :, 0] = xs[:, None]
    verts[:, :, 2] = xs[None, :]
    verts[:, :, 1] = h
    i, j = np.meshgrid(np.arange(quads), np.arange(quads), indexing="ij")
    v00 = (i * n + j).ravel()
    v01 = (i * n + j + 1).ravel()
    v10 = ((i + 1) * n + j).ravel()
    v11 = ((i + 1) * n + j + 1).ravel()
    tris = np.stack([np.stack([v00, v01, v10], 1), np.stack([v10, v01, v11], 1)], 1).reshape(-1, 3)
    return verts.reshape(-1, 3), tris.astype(np.int32), h


def terrain_drop(seed=0, quads=12, n=40, convex_frac=0.4, amplitude=2.0, heights=(0.5, 3.0)):
    """Bodies of every shape dropped onto a small static triangle-mesh terrain, from
    heights U(heights) above the terrain under them."""
    rng = np.random.default_rng(seed)
    s = Scene()
    verts, tris, h = terrain_mesh(seed, quads=quads, amplitude=amplitude)
    tm = s.add_triangle_mesh(verts, tris)
    a = s.add_actor([0, 0, 0])
    s.add_collider(a, TRIANGLE_MESH, [0, 0, 0, float(tm)])
    meshes = [s.add_convex_mesh(*ellipsoid_hull(rng, lo=0.25, hi=0.5)) for _ in range(4)]
    half = quads / 2 - 1.5
    for _ in range(n):
        x, z = rng.uniform(-half, half, 2)
        i, j = int(round(x + quads / 2)), int(round(z + quads / 2))
        y = float(h[max(0, i - 1):i + 2, max(0, j - 1):j + 2].max()) + 0.6 + rng.uniform(*heights)
        q = random_quats(rng, 1)[0]
        v = rng.normal(size=3) * 0.3
        if rng.random() < convex_frac:
            s.add_convex_body(meshes[int(rng.integers(0, 4))], [x, y, z], q, vel=v)
        else:
            k = int(rng.integers(0, 3))
            prm = {SPHERE: [rng.uniform(0.2, 0.4)], CAPSULE: [rng.uniform(0.2, 0.4), rng.uniform(0.1, 0.2)],
                   BOX: list(rng.uniform(0.15, 0.4, 3))}[k]
            s.add_body(k, prm, [x, y, z], q, vel=v)
    return s


def convex_terrain(seed=0, n_convex=8192, n_chains=512, quads=256, amplitude=8.0, drop=(2.0, 40.0)):
    """8,192 convex bodies (each the hull of 24 points on an ellipsoid with semi-axes
    U(0.3, 0.8)) and 512 chains of 8 capsules (r 0.1, half-height 0.3, 7 revolute joints)
    dropped from U(2, 40) m above a static quads x quads fBm terrain (two triangles per
    quad, 1 m spacing, heights 8 * fBm of value noise, 5 octaves)."""
    rng = np.random.default_rng(seed)
    s = Scene()
    verts, tris, h = terrain_mesh(seed, quads=quads, amplitude=amplitude)
    tm = s.add_triangle_mesh(verts, tris)
    a = s.add_actor([0, 0, 0])
    s.add_collider(a, TRIANGLE_MESH, [0, 0, 0, float(tm)])
    half = quads / 2 - 4.0

    def ground(x, z):
        i = int(np.clip(round(x + quads / 2), 0, quads))
        j = int(np.clip(round(z + quads / 2), 0, quads))
        return float(h[max(0, i - 1):i + 2, max(0, j - 1):j + 2].max())

    for _ in range(n_convex):
        mesh = s.add_convex_mesh(*ellipsoid_hull(rng))
        x, z = rng.uniform(-half, half, 2)
        y = ground(x, z) + rng.uniform(*drop)
        s.add_convex_body(mesh, [x, y, z], random_quats(rng, 1)[0])
    for _ in range(n_chains):
        x, z = rng.uniform(-half, half - 7.0, 2)
        top = max(ground(x + k, z) for k in range(0, 8)) + rng.uniform(*drop)
        capsule_chain(s, [x, top, z], n=8, radius=0.1, half_height=0.3, axis="x")
    return s


def convex_pile(seed=0, n=60, n_meshes=6, extent=2.5, convex_frac=0.6, height=1.0):
    """Convex hulls (24 points on random ellipsoids) mixed with spheres, capsules and boxes,
    in a loose pile above a static ground box, for convex-mesh parity tests."""
    rng = np.random.default_rng(seed)
    s = Scene()
    s.add_static(BOX, [20, 0.5, 20], [0, -0.5, 0])
    meshes = [s.add_convex_mesh(*ellipsoid_hull(rng)) for _ in range(n_meshes)]
    for i in range(n):
        pos = [rng.uniform(-extent, extent), height + rng.uniform(0, 2 * extent), rng.uniform(-extent, extent)]
        q = random_quats(rng, 1)[0]
        v = rng.normal(size=3) * 0.5
        if rng.random() < convex_frac:
            sc = (1.0, 1.0, 1.0) if rng.random() < 0.7 else tuple(rng.uniform(0.7, 1.3, 3))
            s.add_convex_body(meshes[int(rng.integers(0, n_meshes))], pos, q, scale=sc, vel=v)
        else:
            k = int(rng.integers(0, 3))
            prm = {SPHERE: [rng.uniform(0.2, 0.5)], CAPSULE: [rng.uniform(0.2, 0.6), rng.uniform(0.15, 0.3)],
                   BOX: list(rng.uniform(0.2, 0.6, 3))}[k]
            s.add_body(k, prm, pos, q, vel=v)
    return s


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


def _axis_quat(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    s = np.sin(angle / 2)
    return [float(axis[0] * s), float(axis[1] * s), float(axis[2] * s), float(np.cos(angle / 2))]


def mechanism_scene(seed=0, ground=True):
    """Prismatic sliders and pistons, servo arms and gear pairs above a ground box, with
    boxes dropped onto them: every joint type of the reference in one scene, anchor frames
    tilted, limits and drives on. Returns the scene; its joints are in creation order."""
    rng = np.random.default_rng(seed)
    s = Scene()
    if ground:
        s.add_static(BOX, [30, 0.5, 30], [0, -0.5, 0])
    for i in range(4):
        # a slider on a tilted rail from a static post; gravity pulls it to its lower limit
        x = -8.0 + 2.5 * i
        post = s.add_static(BOX, [0.1, 0.1, 0.1], [x, 3.0, -4.0])
        q = _axis_quat([0, 0, 1], -np.pi / 2 + rng.uniform(0.3, 1.2))   # slide axis x tilted towards -y
        body = s.add_body(BOX, [0.25, 0.15, 0.15], [x, 3.0, -4.0], random_quats(rng, 1)[0] if i == 3 else (0, 0, 0, 1),
                          vel=rng.normal(size=3))
        drive = dict(drive_enabled=1, target=float(rng.uniform(-1.0, 0.5))) if i % 2 else {}
        s.add_joint(PRISMATIC, post, body, (0, 0, 0), (0, 0, 0), q, q, lower_limit=-1.5, upper_limit=0.8, **drive)
    for i in range(3):
        # a piston between two dynamic bodies
        x = -6.0 + 3.0 * i
        a = s.add_body(BOX, [0.3, 0.3, 0.3], [x, 1.0, 0.0])
        b = s.add_body(CAPSULE, [0.4, 0.15], [x, 2.2, 0.0])
        q = _axis_quat([0, 0, 1], np.pi / 2)                             # slide axis x -> y
        s.add_joint(PRISMATIC, a, b, (0, 0.3, 0), (0, -0.6, 0), q, q, lower_limit=0.1, upper_limit=1.0,
                    drive_enabled=int(i == 1), target=0.5, drive_stiffness=8.0, drive_damping=0.5)
    for i in range(3):
        # a servo arm on a static base
        base = s.add_static(SPHERE, [0.1], [3.0 + 2.0 * i, 2.0, 3.0])
        arm = s.add_body(BOX, [0.1, 0.6, 0.1], [3.0 + 2.0 * i, 1.4, 3.0])
        s.add_joint(SERVO, base, arm, (0, 0, 0), (0, 0.6, 0), target=float(rng.uniform(-1.0, 1.0)),
                    drive_stiffness=float(rng.uniform(5.0, 30.0)))
    for i in range(2):
        # two wheels on revolute joints to a frame, geared; one spinning
        z = 6.0 + 3.0 * i
        frame = s.add_static(BOX, [0.1, 0.1, 0.1], [0.0, 2.0, z])
        w0 = s.add_body(CAPSULE, [0.05, 0.5], [0.5, 2.0, z - 0.6], (0, 0, 0, 1), angvel=(6.0 * (i + 1), 0, 0))
        w1 = s.add_body(CAPSULE, [0.05, 0.5], [0.5, 2.0, z + 0.6])
        s.add_joint(REVOLUTE, frame, w0, (0.5, 0, -0.6), (0, 0, 0))
        s.add_joint(REVOLUTE, frame, w1, (0.5, 0, 0.6), (0, 0, 0))
        s.add_joint(GEAR, w0, w1, (0, 0, 0), (0, 0, 0), gear_ratio=float(rng.uniform(0.5, 3.0)))
    for i in range(12):
        p = [rng.uniform(-8, 8), rng.uniform(3.5, 6.0), rng.uniform(-5, 9)]
        s.add_body(BOX, list(rng.uniform(0.15, 0.35, 3)), p, random_quats(rng, 1)[0])
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
