"""ctypes access to oracle/_ref/libshimcheck.so: an entt::registry and the drop-in
physecs::Scene (physecs_amd/shim), built against the reference's headers (oracle/Makefile,
target shimcheck). replay() builds, in the registry, the scene a scenes.Scene describes, the
way an app builds one: entities in creation order, a TransformComponent each, the dynamic
components in the descriptor's dynamic-storage order, the collision components in collider
order, meshes built by the reference's ConvexMesh and TriangleMesh constructors."""
import ctypes
import os
import subprocess

import numpy as np

from physecs_amd import api

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SHIM_PATH = os.path.join(ROOT, "oracle", "_ref", "libshimcheck.so")
REFERENCE = "/root/reference"

F = ctypes.POINTER(ctypes.c_float)
I32 = ctypes.POINTER(ctypes.c_int32)
U32 = ctypes.POINTER(ctypes.c_uint32)
U8 = ctypes.POINTER(ctypes.c_uint8)
vp = ctypes.c_void_p
_lib = None


def build():
    """Builds the library where the reference's headers are (needs the built libphysecs_gpu.so)."""
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "ref", "shimcheck", f"REF={REFERENCE}"],
                   check=True, capture_output=True)


def available():
    return os.path.exists(SHIM_PATH)


def lib():
    global _lib
    if _lib is None:
        api.lib()            # the same libphysecs_gpu.so the shim links
        L = ctypes.CDLL(SHIM_PATH)
        L.sc_new.restype = vp
        L.sc_free.argtypes = [vp]
        L.sc_create.argtypes = [vp]
        L.sc_create.restype = ctypes.c_uint32
        L.sc_transform.argtypes = [vp, ctypes.c_uint32, F, F]
        L.sc_dynamic.argtypes = [vp, ctypes.c_uint32, ctypes.c_int, F, F, ctypes.c_float, F, F]
        L.sc_convex_mesh.argtypes = [vp, F, ctypes.c_int, I32, I32, F, F, ctypes.c_int]
        L.sc_triangle_mesh.argtypes = [vp, F, ctypes.c_int, U32, ctypes.c_int]
        L.sc_collision.argtypes = [vp, ctypes.c_uint32, ctypes.c_int, I32, F, F, F, F, U8, U8, I32]
        L.sc_set_mass_props.argtypes = [vp, ctypes.c_uint32, ctypes.c_float]
        L.sc_erase_dynamic.argtypes = [vp, ctypes.c_uint32]
        L.sc_destroy.argtypes = [vp, ctypes.c_uint32]
        L.sc_dynamic_order.argtypes = [vp, U32, ctypes.c_int]
        L.sc_gather.argtypes = [vp]
        L.sc_gather.restype = ctypes.POINTER(api.PgxSceneDesc)
        L.sc_simulate.argtypes = [vp, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.sc_error.restype = ctypes.c_char_p
        L.sc_read.argtypes = [vp, ctypes.c_uint32, F, F, F, F]
        L.sc_listen_triggers.argtypes = [vp]
        L.sc_take_triggers.argtypes = [vp, U32, ctypes.c_int]
        L.sc_set_contact_filter.argtypes = [vp, ctypes.c_int]
        L.sc_create_joint.argtypes = [vp, ctypes.POINTER(api.PgxJointDesc), ctypes.c_uint32, ctypes.c_uint32]
        L.sc_set_can_collide.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        L.sc_add_collider.argtypes = [vp, ctypes.c_uint32, ctypes.c_int, F, F, F, F, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int]
        L.sc_clear_colliders.argtypes = [vp, ctypes.c_uint32]
        L.sc_set_kinematic.argtypes = [vp, ctypes.c_uint32, ctypes.c_int]
        L.sc_create_user_joint.argtypes = [vp, ctypes.c_int, ctypes.c_uint32, F, ctypes.c_uint32, F, ctypes.c_float]
        L.sc_destroy_user_joint.argtypes = [vp, ctypes.c_int]
        L.sc_tether_taut_steps.argtypes = [vp, ctypes.c_int]
        _lib = L
    return _lib


def _f(a):
    return np.ascontiguousarray(a, np.float32).ctypes.data_as(F)


class Registry:
    """An entt::registry with a drop-in Scene on it."""

    def __init__(self):
        self.L = lib()
        self.h = self.L.sc_new()
        self._keep = []

    def close(self):
        if self.h:
            self.L.sc_free(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def create(self):
        return int(self.L.sc_create(self.h))

    def transform(self, e, pos, quat):
        self.L.sc_transform(self.h, e, _f(pos), _f(quat))

    def dynamic(self, e, inv_mass, com, inv_inertia_rows, vel=(0, 0, 0), angvel=(0, 0, 0), kinematic=False):
        self.L.sc_dynamic(self.h, e, int(kinematic), _f(vel), _f(angvel), float(inv_mass), _f(com),
                          _f(np.asarray(inv_inertia_rows, np.float32).reshape(9)))

    def convex_mesh(self, verts, faces, normals, centroids):
        fs = np.cumsum([0] + [len(f) for f in faces]).astype(np.int32)
        fi = np.ascontiguousarray([i for f in faces for i in f] or [0], np.int32)
        v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        return int(self.L.sc_convex_mesh(self.h, _f(v), len(v), fs.ctypes.data_as(I32), fi.ctypes.data_as(I32),
                                         _f(normals), _f(centroids), len(faces)))

    def triangle_mesh(self, verts, tris):
        v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
        return int(self.L.sc_triangle_mesh(self.h, _f(v), len(v), t.ctypes.data_as(U32), len(t)))

    def collision(self, e, types, params, lpos, lquat, material, trigger, enable, data):
        n = len(types)
        t = np.ascontiguousarray(types, np.int32)
        prm = np.ascontiguousarray(params, np.float32).reshape(n, 4)
        tr = np.ascontiguousarray(trigger, np.uint8)
        en = np.ascontiguousarray(enable, np.uint8)
        d = np.ascontiguousarray(data, np.int32)
        self.L.sc_collision(self.h, e, n, t.ctypes.data_as(I32), _f(prm), _f(np.reshape(lpos, (n, 3))),
                            _f(np.reshape(lquat, (n, 4))), _f(np.reshape(material, (n, 3))), tr.ctypes.data_as(U8),
                            en.ctypes.data_as(U8), d.ctypes.data_as(I32))

    def set_mass_props(self, e, mass):
        self.L.sc_set_mass_props(self.h, e, float(mass))

    def erase_dynamic(self, e):
        self.L.sc_erase_dynamic(self.h, e)

    def destroy(self, e):
        self.L.sc_destroy(self.h, e)

    def dynamic_order(self):
        out = np.zeros(1 << 16, np.uint32)
        n = self.L.sc_dynamic_order(self.h, out.ctypes.data_as(U32), len(out))
        return out[:n].copy()

    def gather(self):
        """The descriptor the Scene gathers now, as numpy arrays (copies)."""
        d = self.L.sc_gather(self.h).contents
        na, nd, nc = d.n_actors, d.n_dyn, d.n_colliders
        arr = lambda p, n, dt: np.ctypeslib.as_array(p, (n,)).astype(dt) if n else np.zeros(0, dt)
        out = dict(
            entity=arr(d.entity, na, np.uint32), pos=arr(d.pos, 3 * na, np.float32).reshape(-1, 3),
            quat=arr(d.quat, 4 * na, np.float32).reshape(-1, 4), dyn_of_actor=arr(d.dyn_of_actor, na, np.int32),
            actor_of_dyn=arr(d.actor_of_dyn, nd, np.int32), kinematic=arr(d.kinematic, nd, np.uint8),
            vel=arr(d.vel, 3 * nd, np.float32).reshape(-1, 3), angvel=arr(d.angvel, 3 * nd, np.float32).reshape(-1, 3),
            inv_mass=arr(d.inv_mass, nd, np.float32), com=arr(d.com, 3 * nd, np.float32).reshape(-1, 3),
            inv_inertia=arr(d.inv_inertia, 9 * nd, np.float32).reshape(-1, 9),
            col_actor=arr(d.col_actor, nc, np.int32), col_index=arr(d.col_index, nc, np.int32),
            col_pos=arr(d.col_pos, 3 * nc, np.float32).reshape(-1, 3),
            col_quat=arr(d.col_quat, 4 * nc, np.float32).reshape(-1, 4), col_type=arr(d.col_type, nc, np.int32),
            col_params=arr(d.col_params, 4 * nc, np.float32).reshape(-1, 4),
            col_material=arr(d.col_material, 3 * nc, np.float32).reshape(-1, 3),
            col_trigger=arr(d.col_trigger, nc, np.uint8), col_enable=arr(d.col_enable, nc, np.uint8),
            col_data=arr(d.col_data, nc, np.int32), n_worlds=d.n_worlds, n_convex=d.n_convex, n_trimesh=d.n_trimesh)
        cvs = arr(d.convex_vert_start, d.n_convex + 1, np.int32)
        cfs = arr(d.convex_face_start, d.n_convex + 1, np.int32)
        nf = int(cfs[-1])
        fis = arr(d.face_idx_start, nf + 1, np.int32)
        out.update(convex_vert_start=cvs, convex_verts=arr(d.convex_verts, 3 * int(cvs[-1]), np.float32).reshape(-1, 3),
                   convex_face_start=cfs, face_idx_start=fis, face_idx=arr(d.face_idx, int(fis[-1]), np.int32),
                   face_normal=arr(d.face_normal, 3 * nf, np.float32).reshape(-1, 3),
                   face_centroid=arr(d.face_centroid, 3 * nf, np.float32).reshape(-1, 3))
        tvs = arr(d.tri_vert_start, d.n_trimesh + 1, np.int32)
        ts = arr(d.tri_start, d.n_trimesh + 1, np.int32)
        out.update(tri_vert_start=tvs, tri_verts=arr(d.tri_verts, 3 * int(tvs[-1]), np.float32).reshape(-1, 3),
                   tri_start=ts, tri_idx=arr(d.tri_idx, 3 * int(ts[-1]), np.int32).reshape(-1, 3))
        caches = []
        for m in range(d.n_trimesh):
            n = d.tri_bvh_cache_bytes[m]
            caches.append(ctypes.string_at(d.tri_bvh_cache[m], n))
        out["tri_bvh_cache"] = caches
        return out

    # -- stepping through Scene::simulate
    def simulate(self, dt, substeps, iterations, steps=1):
        if self.L.sc_simulate(self.h, float(dt), substeps, iterations, steps) != 0:
            raise RuntimeError(self.L.sc_error().decode())

    def read(self, e):
        p, q, v, w = (np.zeros(3, np.float32), np.zeros(4, np.float32), np.zeros(3, np.float32),
                      np.zeros(3, np.float32))
        self.L.sc_read(self.h, e, _f(p), _f(q), _f(v), _f(w))
        return p, q, v, w

    def listen_triggers(self):
        self.L.sc_listen_triggers(self.h)

    def take_triggers(self):
        out = np.zeros((1 << 14, 5), np.uint32)
        n = self.L.sc_take_triggers(self.h, out.ctypes.data_as(U32), len(out))
        return out[: min(n, len(out))].copy()

    def set_contact_filter(self, which):
        self.L.sc_set_contact_filter(self.h, which)

    def create_joint(self, desc, e0, e1):
        return int(self.L.sc_create_joint(self.h, ctypes.byref(desc), e0, e1))

    # user Joint subclasses in the harness: createJoint<T> on the drop-in Scene
    POINT, UNIVERSAL, HINGE, TETHER = range(4)

    def create_user_joint(self, kind, e0, anchor0, e1, anchor1, param=0.0):
        h = int(self.L.sc_create_user_joint(self.h, kind, e0, _f(anchor0), e1, _f(anchor1), param))
        assert h >= 0
        return h

    def destroy_user_joint(self, handle):
        self.L.sc_destroy_user_joint(self.h, handle)

    def tether_taut_steps(self, handle):
        return int(self.L.sc_tether_taut_steps(self.h, handle))

    def set_can_collide(self, e0, e1, can):
        self.L.sc_set_can_collide(self.h, e0, e1, int(can))

    def add_collider(self, e, gtype, params, material=(0.6, 0.0, 0.0), lpos=(0, 0, 0), lquat=(0, 0, 0, 1),
                     trigger=False, enable=True, data=0):
        """Scene::addCollider (a mesh collider's fourth parameter names a registry mesh)."""
        p = np.zeros(4, np.float32)
        p[: len(params)] = params
        self.L.sc_add_collider(self.h, e, int(gtype), _f(p), _f(lpos), _f(lquat), _f(material), int(bool(trigger)),
                               int(bool(enable)), int(data))

    def clear_colliders(self, e):
        self.L.sc_clear_colliders(self.h, e)

    def set_kinematic(self, e, kinematic):
        self.L.sc_set_kinematic(self.h, e, int(bool(kinematic)))


def replay(scene):
    """A Registry holding `scene` (a scenes.Scene) as an app builds it. Returns (registry,
    entity of each actor)."""
    arr = scene.arrays()
    reg = Registry()
    ents = [reg.create() for _ in range(scene.n_actors)]
    for a, e in enumerate(ents):
        reg.transform(e, arr["pos"][a], arr["quat"][a])
    for d in range(scene.n_dyn):
        a = int(arr["actor_of_dyn"][d])
        # the descriptor holds the inverse inertia column-major (element (r, c) at 3c + r)
        rows = arr["inv_inertia"][d].reshape(3, 3).T
        reg.dynamic(ents[a], arr["inv_mass"][d], arr["com"][d], rows, arr["vel"][d], arr["angvel"][d],
                    bool(arr["kinematic"][d]))
    cvx = [reg.convex_mesh(v, f, n, c) for (v, f, n, c) in scene.convex]
    tri = [reg.triangle_mesh(v, t) for (v, t) in scene.trimesh]
    seen = []
    for a in arr["col_actor"].tolist():
        if a not in seen:
            seen.append(a)
    for a in seen:
        cs = np.nonzero(arr["col_actor"] == a)[0]
        prm = arr["col_params"][cs].copy()
        for k, c in enumerate(cs):   # mesh colliders name the registry's mesh objects
            t = int(arr["col_type"][c])
            if t == 3:
                prm[k, 3] = cvx[int(prm[k, 3])]
            elif t == 4:
                prm[k, 3] = tri[int(prm[k, 3])]
        reg.collision(ents[a], arr["col_type"][cs], prm, arr["col_pos"][cs], arr["col_quat"][cs],
                      arr["col_material"][cs], arr["col_trigger"][cs], arr["col_enable"][cs], arr["col_data"][cs])
    return reg, ents
