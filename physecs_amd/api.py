"""ctypes binding of libphysecs_gpu.so (include/physecs_gpu.h).

The library is built in-tree (physecs_amd/lib/libphysecs_gpu.so) by physecs_amd.build.
There is no CPU fallback: loading fails loudly when the library is missing, and creating
a world fails with the HIP error when no GPU is present.
"""
import atexit
import ctypes
import os
import sys
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libphysecs_gpu.so")

F = ctypes.POINTER(ctypes.c_float)
I32 = ctypes.POINTER(ctypes.c_int32)
U32 = ctypes.POINTER(ctypes.c_uint32)
U8 = ctypes.POINTER(ctypes.c_uint8)

PGX_OK = 0
ERR_CAPACITY = -3
ERRORS = {-1: "PGX_ERR_INVALID", -2: "PGX_ERR_HIP", -3: "PGX_ERR_CAPACITY", -4: "PGX_ERR_UNSUPPORTED"}

# every symbol declared in include/physecs_gpu.h
EXPORTS = [
    "pgx_last_error", "pgx_device_count", "pgx_sort_u64", "pgx_sort_u64_noalloc", "pgx_world_create", "pgx_world_destroy", "pgx_world_set_params", "pgx_world_set_contact_filter",
    "pgx_world_set_can_collide", "pgx_world_step", "pgx_world_sync", "pgx_world_read_state",
    "pgx_world_write_state", "pgx_world_get_counts", "pgx_world_get_bounds", "pgx_world_get_pairs",
    "pgx_world_get_manifolds", "pgx_world_enable_timing", "pgx_world_get_timing", "pgx_world_reset_stats",
    "pgx_world_get_stats", "pgx_world_set_exec", "pgx_world_add_joints", "pgx_world_set_joint_params",
    "pgx_world_destroy_joint", "pgx_world_set_graph", "pgx_world_set_tuning", "pgx_world_get_tuning",
    "pgx_world_autotune", "pgx_world_get_exec", "pgx_world_raycast", "pgx_world_overlap",
    "pgx_world_get_trigger_events", "pgx_world_overlap_mtd", "pgx_world_character_move", "pgx_mass_properties",
    "pgx_world_patch_transforms", "pgx_convex_hull", "pgx_trimesh_cache", "pgx_trimesh_cache_validate", "pgx_trimesh_cache_pack", "pgx_world_state_size", "pgx_world_save_state", "pgx_world_load_state",
    "pgx_world_load_state_matched", "pgx_world_set_joint_callbacks",
    "pgx_server_create", "pgx_server_destroy", "pgx_server_submit", "pgx_server_step", "pgx_server_done",
    "pgx_server_take_result", "pgx_server_footprint", "pgx_server_set_graph",
    "pgx_server_set_tuning",
]


class PgxConfig(ctypes.Structure):
    _fields_ = [("num_substeps", ctypes.c_int), ("num_iterations", ctypes.c_int), ("gravity", ctypes.c_float),
                ("device", ctypes.c_int), ("pair_capacity", ctypes.c_int), ("block_size", ctypes.c_int)]


class PgxSceneDesc(ctypes.Structure):
    _fields_ = [
        ("n_actors", ctypes.c_int), ("entity", U32), ("pos", F), ("quat", F), ("dyn_of_actor", I32),
        ("n_dyn", ctypes.c_int), ("actor_of_dyn", I32), ("kinematic", U8), ("vel", F), ("angvel", F),
        ("inv_mass", F), ("com", F), ("inv_inertia", F),
        ("n_colliders", ctypes.c_int), ("col_actor", I32), ("col_index", I32), ("col_pos", F), ("col_quat", F),
        ("col_type", I32), ("col_params", F), ("col_material", F), ("col_trigger", U8), ("col_enable", U8),
        ("col_data", I32), ("n_worlds", ctypes.c_int), ("actor_world", I32),
        ("n_convex", ctypes.c_int), ("convex_vert_start", I32), ("convex_verts", F), ("convex_face_start", I32),
        ("face_idx_start", I32), ("face_idx", I32), ("face_normal", F), ("face_centroid", F),
        ("n_trimesh", ctypes.c_int), ("tri_vert_start", I32), ("tri_verts", F), ("tri_start", I32), ("tri_idx", I32),
        ("tri_bvh_cache", ctypes.POINTER(ctypes.c_void_p)), ("tri_bvh_cache_bytes", ctypes.POINTER(ctypes.c_size_t)),
    ]


class PgxJointDesc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("actor0", ctypes.c_int32), ("actor1", ctypes.c_int32),
                ("anchor0", ctypes.c_float * 7), ("anchor1", ctypes.c_float * 7), ("drive_enabled", ctypes.c_int32),
                ("drive_velocity", ctypes.c_float), ("drive_max_torque", ctypes.c_float),
                ("lower_limit", ctypes.c_float), ("upper_limit", ctypes.c_float), ("target", ctypes.c_float),
                ("drive_stiffness", ctypes.c_float), ("drive_damping", ctypes.c_float), ("gear_ratio", ctypes.c_float)]


# User joint types (PGX_JOINT_CUSTOM): getSolverDesc's layout and MakeConstraintsFunc as host
# functions. The world-space data is 30 floats (p0, p1, r0, r1, then u0 and u1 by columns:
# pgx_joint_world_space); a row is 15 floats (linear, angular0, angular1, target velocity, c,
# min, max, frequency, damping ratio: pgx_joint_row). The oracle takes the same functions.
JOINT_LAYOUT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, F, I32)
JOINT_ROWS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, F, F, ctypes.c_int)
JOINT_MAX_ROWS = 7
JOINT_WS_FLOATS, JOINT_ROW_FLOATS = 30, 15


def joint_callbacks(layout, rows):
    """C function pointers for a custom joint's Python functions:
    layout(joint, ws) -> row flags in creation order (PGX_ROW_* values), called once per step;
    rows(joint, ws, rows) writes into rows, an (n, 15) float32 array holding the joint's rows as
    the previous call of the step left them; called every substep. ws is a (30,) float32 array.
    An exception in either function fails the step (it is printed)."""
    import traceback

    def lay(user, joint, ws, flags):
        try:
            fl = [int(f) for f in layout(joint, np.ctypeslib.as_array(ws, shape=(JOINT_WS_FLOATS,)))]
            if len(fl) > JOINT_MAX_ROWS:
                return len(fl)   # refused by the caller, nothing written
            for i, f in enumerate(fl):
                flags[i] = f
            return len(fl)
        except Exception:
            traceback.print_exc()
            return -1

    def wr(user, joint, ws, out, n):
        try:
            if n > 0:
                rows(joint, np.ctypeslib.as_array(ws, shape=(JOINT_WS_FLOATS,)),
                     np.ctypeslib.as_array(out, shape=(n, JOINT_ROW_FLOATS)))
        except Exception:
            traceback.print_exc()

    return JOINT_LAYOUT_FN(lay), JOINT_ROWS_FN(wr)


class PgxCounts(ctypes.Structure):
    _fields_ = [("pairs", ctypes.c_int), ("manifolds", ctypes.c_int), ("points", ctypes.c_int),
                ("colors", ctypes.c_int), ("joint_rows", ctypes.c_int), ("coloring_rounds", ctypes.c_int)]


class PgxStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_long), ("solve_launches", ctypes.c_long), ("solve_us", ctypes.c_double),
                ("solve_bytes", ctypes.c_double), ("step_us", ctypes.c_double), ("batch_launches", ctypes.c_long),
                ("batch_us", ctypes.c_double), ("batch_bytes", ctypes.c_double)]


class PgxServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_double), ("max_worlds", ctypes.c_int), ("admit_interval", ctypes.c_int)]


class PgxServerStepInfo(ctypes.Structure):
    _fields_ = [("active", ctypes.c_int), ("admitted", ctypes.c_int), ("retired", ctypes.c_int),
                ("finished", ctypes.c_int), ("queued", ctypes.c_int), ("bodies", ctypes.c_longlong),
                ("footprint_bytes", ctypes.c_double)]


class PgxTuning(ctypes.Structure):
    _fields_ = [("solve_block", ctypes.c_int), ("color_batch", ctypes.c_int), ("pairs_per_lane", ctypes.c_int),
                ("color_rounds", ctypes.c_int), ("eager_rows_max", ctypes.c_int), ("pair_search", ctypes.c_int)]


class PgxCharacter(ctypes.Structure):
    """pgx_character (include/physecs_gpu.h)."""
    _fields_ = [("position", ctypes.c_float * 3), ("orientation", ctypes.c_float * 4), ("height", ctypes.c_float),
                ("radius", ctypes.c_float), ("max_step_height", ctypes.c_float), ("slope_limit", ctypes.c_float),
                ("num_iterations", ctypes.c_int32), ("world", ctypes.c_int32)]


CHARACTER_DTYPE = np.dtype([("position", np.float32, 3), ("orientation", np.float32, 4), ("height", np.float32),
                            ("radius", np.float32), ("max_step_height", np.float32), ("slope_limit", np.float32),
                            ("num_iterations", np.int32), ("world", np.int32)])
assert CHARACTER_DTYPE.itemsize == ctypes.sizeof(PgxCharacter)


class PgxError(RuntimeError):
    pass


# Handles still open. They are closed by close(), by the garbage collector, or at the latest
# by _close_open_handles, which atexit runs at the start of interpreter shutdown while the
# HIP runtime and any profiler attached to it are fully alive. Destructors that run later,
# during finalization, make no HIP call: the process is ending and the driver reclaims the
# device memory.
_open_handles = weakref.WeakSet()


@atexit.register
def _close_open_handles():
    for obj in list(_open_handles):
        obj.close()


def _finalizing():
    return sys.is_finalizing()


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m physecs_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.pgx_last_error.restype = ctypes.c_char_p
        L.pgx_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.pgx_world_create.argtypes = [ctypes.POINTER(PgxConfig), ctypes.POINTER(PgxSceneDesc), ctypes.POINTER(vp)]
        L.pgx_world_destroy.argtypes = [vp]
        L.pgx_world_set_params.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float]
        L.pgx_world_set_can_collide.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        L.pgx_world_step.argtypes = [vp, ctypes.c_float]
        L.pgx_world_sync.argtypes = [vp]
        L.pgx_world_read_state.argtypes = [vp, F, F, F, F]
        L.pgx_world_write_state.argtypes = [vp, F, F, F, F]
        L.pgx_world_get_counts.argtypes = [vp, ctypes.POINTER(PgxCounts)]
        L.pgx_world_get_bounds.argtypes = [vp, F]
        L.pgx_world_get_pairs.argtypes = [vp, I32, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.pgx_world_get_manifolds.argtypes = [vp, I32, I32, F, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.pgx_world_enable_timing.argtypes = [vp, ctypes.c_int]
        L.pgx_world_get_timing.argtypes = [vp, F, F, F, F, F]
        L.pgx_world_reset_stats.argtypes = [vp]
        L.pgx_world_get_stats.argtypes = [vp, ctypes.POINTER(PgxStats)]
        L.pgx_world_set_exec.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.pgx_world_add_joints.argtypes = [vp, ctypes.c_int, ctypes.POINTER(PgxJointDesc)]
        L.pgx_world_set_joint_params.argtypes = [vp, ctypes.c_int, ctypes.POINTER(PgxJointDesc)]
        L.pgx_world_destroy_joint.argtypes = [vp, ctypes.c_int]
        L.pgx_world_set_joint_callbacks.argtypes = [vp, JOINT_LAYOUT_FN, JOINT_ROWS_FN, vp]
        L.pgx_world_set_graph.argtypes = [vp, ctypes.c_int]
        L.pgx_server_set_graph.argtypes = [vp, ctypes.c_int]
        L.pgx_world_set_tuning.argtypes = [vp, ctypes.POINTER(PgxTuning)]
        L.pgx_world_get_tuning.argtypes = [vp, ctypes.POINTER(PgxTuning)]
        L.pgx_world_autotune.argtypes = [vp, ctypes.c_float, ctypes.c_int, ctypes.POINTER(PgxTuning)]
        L.pgx_world_get_exec.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.pgx_world_raycast.argtypes = [vp, ctypes.c_int, F, F, F, ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int32), F,
                                        ctypes.POINTER(ctypes.c_double)]
        L.pgx_world_overlap.argtypes = [vp, ctypes.c_int, I32, F, F, F, I32, I32, ctypes.c_int, I32, I32, I32,
                                        ctypes.POINTER(ctypes.c_double)]
        L.pgx_world_get_trigger_events.argtypes = [vp, ctypes.c_int, I32, I32, I32, I32]
        L.pgx_world_overlap_mtd.argtypes = [vp, ctypes.c_int, I32, F, F, F, I32, ctypes.c_int, I32, I32, F, I32,
                                            ctypes.POINTER(ctypes.c_double)]
        L.pgx_world_character_move.argtypes = [vp, ctypes.c_int, ctypes.POINTER(PgxCharacter), F, ctypes.c_float,
                                               ctypes.POINTER(ctypes.c_uint8)]
        L.pgx_mass_properties.argtypes = [ctypes.c_int, I32, F, F, F, ctypes.POINTER(ctypes.c_uint8), ctypes.c_int,
                                          I32, F, I32, I32, I32, ctypes.c_float, F, F, F]
        L.pgx_world_patch_transforms.argtypes = [vp, ctypes.c_int, I32, F, F]
        L.pgx_convex_hull.argtypes = [F, ctypes.c_int, F, I32, I32, I32, F, F, I32]
        L.pgx_trimesh_cache.argtypes = [F, ctypes.c_int, I32, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                        ctypes.POINTER(ctypes.c_size_t)]
        L.pgx_world_state_size.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
        L.pgx_world_save_state.argtypes = [vp, ctypes.c_void_p, ctypes.c_size_t]
        L.pgx_world_load_state.argtypes = [vp, ctypes.c_void_p, ctypes.c_size_t]
        L.pgx_world_load_state_matched.argtypes = [vp, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                                   ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int32)]
        L.pgx_server_set_tuning.argtypes = [vp, ctypes.POINTER(PgxTuning)]
        L.pgx_server_create.argtypes = [ctypes.POINTER(PgxConfig), ctypes.POINTER(PgxServerConfig), ctypes.POINTER(vp)]
        L.pgx_server_destroy.argtypes = [vp]
        L.pgx_server_submit.argtypes = [vp, ctypes.POINTER(PgxSceneDesc), ctypes.POINTER(PgxJointDesc), ctypes.c_int,
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.pgx_server_step.argtypes = [vp, ctypes.c_float, ctypes.POINTER(PgxServerStepInfo)]
        L.pgx_server_done.argtypes = [vp, ctypes.c_int]
        L.pgx_server_take_result.argtypes = [vp, ctypes.c_int, F, F, F, F]
        L.pgx_server_footprint.argtypes = [ctypes.POINTER(PgxSceneDesc), ctypes.POINTER(ctypes.c_double)]
        for name in EXPORTS:
            if name != "pgx_last_error":
                getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def check(rc):
    if rc != PGX_OK:
        msg = lib().pgx_last_error().decode(errors="replace")
        raise PgxError(f"{ERRORS.get(rc, rc)}: {msg}")


def mass_properties(types, params, local_pos, local_quat, mass, is_trigger=None, meshes=()):
    """pgx_mass_properties: physecs::setMassProps for one body of colliders (host code, no
    device needed). meshes: convex meshes as (verts, faces, ...) tuples, indexed by a convex
    collider's params[3]. Returns (inv_mass, com (3,), inv_inertia (9,) column-major)."""
    t = np.ascontiguousarray(np.atleast_1d(types), np.int32)
    n = len(t)
    prm = np.zeros((n, 4), np.float32)
    pa = np.asarray(params, np.float32).reshape(n, -1)
    prm[:, : pa.shape[1]] = pa
    lp = np.ascontiguousarray(local_pos, np.float32).reshape(n, 3)
    lq = np.ascontiguousarray(local_quat, np.float32).reshape(n, 4)
    tr = None if is_trigger is None else np.ascontiguousarray(is_trigger, np.uint8)
    vs, fs, ist, idx, verts = [0], [0], [0], [], []
    for m in meshes:
        verts.append(np.asarray(m[0], np.float32).reshape(-1, 3))
        vs.append(vs[-1] + len(verts[-1]))
        fs.append(fs[-1] + len(m[1]))
        for f in m[1]:
            idx.extend(int(i) for i in f)
            ist.append(ist[-1] + len(f))
    i32 = lambda a: np.ascontiguousarray(a if len(a) else [0], np.int32)
    cv = np.ascontiguousarray(np.concatenate(verts) if verts else np.zeros((1, 3)), np.float32)
    vs, fs, ist, idx = i32(vs), i32(fs), i32(ist), i32(idx)
    ip = lambda a: a.ctypes.data_as(I32)
    im = np.zeros(1, np.float32)
    com = np.zeros(3, np.float32)
    inv = np.zeros(9, np.float32)
    check(lib().pgx_mass_properties(n, ip(t), _f(prm), _f(lp), _f(lq),
                                    None if tr is None else tr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                    len(meshes), ip(vs), _f(cv), ip(fs), ip(ist), ip(idx), float(mass), _f(im),
                                    _f(com), _f(inv)))
    return im[0], com, inv


def convex_hull(points):
    """pgx_convex_hull: (verts (k, 3), faces [lists of vertex indices], normals, centroids)
    of the convex hull of points (host code, no device needed)."""
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    n = len(p)
    verts = np.zeros((max(n, 1), 3), np.float32)
    fs = np.zeros(2 * n + 1, np.int32)
    fi = np.zeros(max(6 * n, 1), np.int32)
    nrm = np.zeros((max(2 * n, 1), 3), np.float32)
    cen = np.zeros((max(2 * n, 1), 3), np.float32)
    nv, nf = ctypes.c_int32(0), ctypes.c_int32(0)
    check(lib().pgx_convex_hull(_f(p), n, _f(verts), ctypes.byref(nv), fs.ctypes.data_as(I32), fi.ctypes.data_as(I32),
                                _f(nrm), _f(cen), ctypes.byref(nf)))
    faces = [list(fi[fs[f]:fs[f + 1]]) for f in range(nf.value)]
    return verts[: nv.value].copy(), faces, nrm[: nf.value].copy(), cen[: nf.value].copy()


def trimesh_cache(verts, tris):
    """pgx_trimesh_cache: the TriangleMesh build of one mesh as bytes (host code), for
    Scene.tri_bvh_cache."""
    v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
    n = ctypes.c_size_t(0)
    check(lib().pgx_trimesh_cache(_f(v), len(v), t.ctypes.data_as(I32), len(t), None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value)
    check(lib().pgx_trimesh_cache(_f(v), len(v), t.ctypes.data_as(I32), len(t), buf, n.value, ctypes.byref(n)))
    return buf.raw


def trimesh_cache_validate(verts, tris, cache):
    """pgx_trimesh_cache_validate: PGX_OK (0) or PGX_ERR_INVALID for `cache` against the mesh."""
    v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
    L = lib()
    L.pgx_trimesh_cache_validate.argtypes = [F, ctypes.c_int, I32, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
    return L.pgx_trimesh_cache_validate(_f(v), len(v), t.ctypes.data_as(I32), len(t), bytes(cache), len(cache))


def device_count():
    n = ctypes.c_int(0)
    check(lib().pgx_device_count(ctypes.byref(n)))
    return n.value


def sort_u64(keys, vals, bits, device=0):
    """pgx_sort_u64: the library's stable radix sort of (keys, vals) by the low `bits` bits."""
    keys = np.ascontiguousarray(keys, np.uint64)
    vals = np.ascontiguousarray(vals, np.int32)
    ko = np.zeros_like(keys)
    vo = np.zeros_like(vals)
    L = lib()
    L.pgx_sort_u64.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), I32, ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(ctypes.c_uint64), I32]
    P64 = ctypes.POINTER(ctypes.c_uint64)
    check(L.pgx_sort_u64(device, keys.ctypes.data_as(P64), vals.ctypes.data_as(I32), len(keys), bits,
                         ko.ctypes.data_as(P64), vo.ctypes.data_as(I32)))
    return ko, vo


def _f(a):
    return None if a is None else a.ctypes.data_as(F)


class World:
    """One pgx_world: a scene resident on the GPU, stepped by Scene::simulate semantics."""

    def __init__(self, scene, substeps=4, iterations=2, gravity=9.81, device=0, pair_capacity=0):
        self.L = lib()
        self.scene = scene
        self._arr = scene.arrays()
        desc = scene.struct(PgxSceneDesc)
        if scene.n_worlds <= 1:
            desc.actor_world = ctypes.cast(None, I32)
            desc.n_worlds = 1
        cfg = PgxConfig(substeps, iterations, gravity, device, pair_capacity, 0)
        h = ctypes.c_void_p()
        check(self.L.pgx_world_create(ctypes.byref(cfg), ctypes.byref(desc), ctypes.byref(h)))
        self.h = h
        _open_handles.add(self)
        check(self.L.pgx_world_set_params(self.h, substeps, iterations, gravity))
        for e0, e1 in getattr(scene, "non_colliding", []):
            check(self.L.pgx_world_set_can_collide(self.h, e0, e1, 0))
        if getattr(scene, "joints", None):
            arr = scene.joint_array(PgxJointDesc)
            check(self.L.pgx_world_add_joints(self.h, len(scene.joints), arr))

    def close(self):
        """pgx_world_destroy: waits for the world's stream and frees its device memory."""
        if getattr(self, "h", None):
            self.L.pgx_world_destroy(self.h)
            self.h = None
        _open_handles.discard(self)

    def __del__(self):
        if not _finalizing():
            self.close()

    def step(self, dt):
        check(self.L.pgx_world_step(self.h, dt))

    def sync(self):
        check(self.L.pgx_world_sync(self.h))

    def state(self):
        s = self.scene
        pos = np.zeros((s.n_actors, 3), np.float32)
        quat = np.zeros((s.n_actors, 4), np.float32)
        vel = np.zeros((s.n_dyn, 3), np.float32)
        ang = np.zeros((s.n_dyn, 3), np.float32)
        check(self.L.pgx_world_read_state(self.h, _f(pos), _f(quat), _f(vel), _f(ang)))
        return pos, quat, vel, ang

    def write_state(self, pos=None, quat=None, vel=None, ang=None):
        c = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        pos, quat, vel, ang = c(pos), c(quat), c(vel), c(ang)
        check(self.L.pgx_world_write_state(self.h, _f(pos), _f(quat), _f(vel), _f(ang)))

    def patch_transforms(self, actors, pos, quat):
        """pgx_world_patch_transforms: registry.patch<TransformComponent> of these actors
        (their poses, and their colliders' bounds with the margin)."""
        a = np.ascontiguousarray(np.atleast_1d(actors), np.int32)
        p = np.ascontiguousarray(pos, np.float32).reshape(len(a), 3)
        q = np.ascontiguousarray(quat, np.float32).reshape(len(a), 4)
        check(self.L.pgx_world_patch_transforms(self.h, len(a), a.ctypes.data_as(I32), _f(p), _f(q)))

    def set_joint_params(self, index, **params):
        """The joint's setters: updates the data members named in params (JOINT_FIELDS) of joint
        `index` (creation order); takes effect from the next step."""
        from . import scenes
        if not hasattr(self, "_joints"):
            self._joints = [tuple(j[:5]) + (dict(j[5]),) for j in self.scene.joints]
        t, a0, a1, an0, an1, p = self._joints[index]
        for k, v in params.items():
            if k not in scenes.JOINT_FIELDS:
                raise ValueError(f"unknown joint parameter {k}")
            p[k] = v
        d = PgxJointDesc()
        scenes.Scene.fill_joint(d, t, a0, a1, an0, an1, p)
        check(self.L.pgx_world_set_joint_params(self.h, index, ctypes.byref(d)))

    def destroy_joint(self, index):
        """Scene::destroyJoint for joint `index` (creation order)."""
        check(self.L.pgx_world_destroy_joint(self.h, index))

    def set_joint_callbacks(self, layout, rows):
        """The host functions of the world's custom joints (joint_callbacks); None, None
        removes them."""
        if layout is None:
            self._joint_fns = None
            check(self.L.pgx_world_set_joint_callbacks(self.h, JOINT_LAYOUT_FN(), JOINT_ROWS_FN(), None))
            return
        self._joint_fns = joint_callbacks(layout, rows)   # kept alive with the world
        check(self.L.pgx_world_set_joint_callbacks(self.h, self._joint_fns[0], self._joint_fns[1], None))

    def counts(self):
        c = PgxCounts()
        check(self.L.pgx_world_get_counts(self.h, ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in PgxCounts._fields_}

    def bounds(self):
        out = np.zeros((self.scene.n_colliders, 6), np.float32)
        check(self.L.pgx_world_get_bounds(self.h, _f(out)))
        return out

    def pairs(self):
        n = ctypes.c_int(0)
        check(self.L.pgx_world_get_pairs(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros((n.value, 2), np.int32)
        if n.value:
            check(self.L.pgx_world_get_pairs(self.h, out.ctypes.data_as(I32), n.value, ctypes.byref(n)))
        return out

    def manifolds(self):
        n = ctypes.c_int(0)
        check(self.L.pgx_world_get_manifolds(self.h, None, None, None, 0, ctypes.byref(n)))
        m = n.value
        keys = np.zeros((m, 5), np.int32)
        ints = np.zeros((m, 4), np.int32)
        f = np.zeros((m, 56), np.float32)
        if m:
            check(self.L.pgx_world_get_manifolds(self.h, keys.ctypes.data_as(I32), ints.ctypes.data_as(I32),
                                                 _f(f), m, ctypes.byref(n)))
        return keys, ints, f

    def enable_timing(self, on=True, period=1):
        """Per-launch HIP events on every `period`-th step (off with on=False)."""
        check(self.L.pgx_world_enable_timing(self.h, int(period) if on else 0))

    EXEC_MODES = {"colors": 0, "cooperative": 1, "flow": 2, "xcd": 3, "worlds": 4, "pass": 5, "xpass": 6}

    def set_exec(self, mode="colors", blocks=0):
        """How the contact passes launch: "colors" (one launch per colour, the default),
        "flow" (one dependency-driven launch per pass), "xcd" (the same with each world's
        manifolds on one XCD), "cooperative" (one launch per step), "worlds" (one
        workgroup per world per pass, bodies in LDS), "pass" (one launch per pass, grid
        barriers between colours) or "xpass" (one launch per pass on the CUs of one XCD,
        barriers between colours, velocities in that XCD's L2)."""
        check(self.L.pgx_world_set_exec(self.h, self.EXEC_MODES[mode], blocks))

    def set_contact_filter(self, fn):
        """Scene::setContactFilter: fn(is_trigger0, data0, is_trigger1, data1) -> True for
        TRIGGER, tabulated over the scene's colliders (scenes.contact_filter_table); None
        restores the default filter."""
        L = self.L
        L.pgx_world_set_contact_filter.argtypes = [ctypes.c_void_p, ctypes.c_int, U8, I32]
        if fn is None:
            check(L.pgx_world_set_contact_filter(self.h, 0, None, None))
            return
        from physecs_amd import scenes
        k, table, classes = scenes.contact_filter_table(self.scene, fn)
        self._filter = (table, classes)
        check(L.pgx_world_set_contact_filter(self.h, k, table.ctypes.data_as(U8), classes.ctypes.data_as(I32)))

    def set_graph(self, enabled=True):
        """Launch each step's substep loop as one HIP graph (colour mode); same results."""
        check(self.L.pgx_world_set_graph(self.h, 1 if enabled else 0))

    def tuning(self):
        t = PgxTuning()
        check(self.L.pgx_world_get_tuning(self.h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in PgxTuning._fields_}

    def set_tuning(self, **values):
        """Launch shapes (solve_block, color_batch, pairs_per_lane, color_rounds,
        eager_rows_max, pair_search); the results do not depend on them."""
        t = PgxTuning(**dict(self.tuning(), **values))
        check(self.L.pgx_world_set_tuning(self.h, ctypes.byref(t)))

    def autotune(self, dt, trial_steps=4):
        """pgx_world_autotune: picks the fastest launch shapes from trial steps of this
        world's current state, which it restores; returns the chosen values."""
        t = PgxTuning()
        check(self.L.pgx_world_autotune(self.h, dt, int(trial_steps), ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in PgxTuning._fields_}

    def raycast(self, origins, dirs, max_dist, world=None, ignore=None):
        """pgx_world_raycast: Scene::raycastClosest for every ray. Returns (collider index
        or -1, hit distance, device microseconds)."""
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = len(o)
        md = np.ascontiguousarray(np.broadcast_to(np.asarray(max_dist, np.float32), (n,)))
        wd = None if world is None else np.ascontiguousarray(np.broadcast_to(np.asarray(world, np.int32), (n,)))
        ig = None if ignore is None else np.ascontiguousarray(np.asarray(ignore, np.uint32))
        col = np.full(n, -1, np.int32)
        t = np.zeros(n, np.float32)
        us = ctypes.c_double(0)
        check(self.L.pgx_world_raycast(self.h, n, _f(o), _f(d), _f(md),
                                       None if wd is None else wd.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                       None if ig is None else ig.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                       col.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _f(t), ctypes.byref(us)))
        return col, t, us.value

    def overlap(self, types, params, positions, orientations, world=None, filters=None, capacity=None):
        """pgx_world_overlap: Scene::overlap for every query shape. Returns (list of
        ascending collider-index arrays, one per query; device microseconds). Without a
        capacity, the first call sizes the hit buffer from the returned total."""
        t = np.ascontiguousarray(np.atleast_1d(types), np.int32)
        n = len(t)
        prm = np.zeros((n, 4), np.float32)
        pa = np.asarray(params, np.float32).reshape(n, -1) if n else np.zeros((0, 4), np.float32)
        prm[:, : pa.shape[1]] = pa
        pos = np.ascontiguousarray(positions, np.float32).reshape(n, 3)
        q = np.ascontiguousarray(orientations, np.float32).reshape(n, 4)
        i32 = ctypes.POINTER(ctypes.c_int32)
        wd = None if world is None else np.ascontiguousarray(np.broadcast_to(np.asarray(world, np.int32), (n,)))
        fl = None if filters is None else np.ascontiguousarray(np.broadcast_to(np.asarray(filters, np.int32), (n,)))
        off = np.zeros(n + 1, np.int32)
        total = ctypes.c_int32(0)
        us = ctypes.c_double(0)
        cap = 4 * n + 1024 if capacity is None else int(capacity)
        while True:
            cols = np.zeros(max(cap, 1), np.int32)
            rc = self.L.pgx_world_overlap(self.h, n, t.ctypes.data_as(i32), _f(prm), _f(pos), _f(q),
                                          None if wd is None else wd.ctypes.data_as(i32),
                                          None if fl is None else fl.ctypes.data_as(i32), cap,
                                          off.ctypes.data_as(i32), cols.ctypes.data_as(i32), ctypes.byref(total),
                                          ctypes.byref(us))
            if rc == ERR_CAPACITY and capacity is None and total.value > cap:
                cap = total.value
                continue
            check(rc)
            break
        return [cols[off[i]:off[i + 1]].copy() for i in range(n)], us.value

    def overlap_mtd(self, types, params, positions, orientations, world=None, capacity=None):
        """pgx_world_overlap_mtd: Scene::overlapWithMinTranslationalDistance for every query.
        Returns (per query: (collider indices, (k, 4) normal xyz + mtd); device microseconds)."""
        t = np.ascontiguousarray(np.atleast_1d(types), np.int32)
        n = len(t)
        prm = np.zeros((n, 4), np.float32)
        pa = np.asarray(params, np.float32).reshape(n, -1) if n else np.zeros((0, 4), np.float32)
        prm[:, : pa.shape[1]] = pa
        pos = np.ascontiguousarray(positions, np.float32).reshape(n, 3)
        q = np.ascontiguousarray(orientations, np.float32).reshape(n, 4)
        i32 = ctypes.POINTER(ctypes.c_int32)
        wd = None if world is None else np.ascontiguousarray(np.broadcast_to(np.asarray(world, np.int32), (n,)))
        off = np.zeros(n + 1, np.int32)
        total = ctypes.c_int32(0)
        us = ctypes.c_double(0)
        cap = 4 * n + 1024 if capacity is None else int(capacity)
        while True:
            cols = np.zeros(max(cap, 1), np.int32)
            nm = np.zeros((max(cap, 1), 4), np.float32)
            rc = self.L.pgx_world_overlap_mtd(self.h, n, t.ctypes.data_as(i32), _f(prm), _f(pos), _f(q),
                                              None if wd is None else wd.ctypes.data_as(i32), cap,
                                              off.ctypes.data_as(i32), cols.ctypes.data_as(i32), _f(nm),
                                              ctypes.byref(total), ctypes.byref(us))
            if rc == ERR_CAPACITY and capacity is None and total.value > cap:
                cap = total.value
                continue
            check(rc)
            break
        return [(cols[off[i]:off[i + 1]].copy(), nm[off[i]:off[i + 1]].copy()) for i in range(n)], us.value

    def character_move(self, positions, orientations, translations, height, radius, slope_limit=np.cos(np.pi / 4),
                       num_iterations=2, world=0, max_step_height=0.3, dt=1 / 60):
        """pgx_world_character_move: CharacterController::move for n characters. Returns
        (new positions (n, 3), CharacterCollisionFlags (n,))."""
        p = np.asarray(positions, np.float32).reshape(-1, 3)
        n = len(p)
        arr = np.zeros(max(n, 1), CHARACTER_DTYPE)   # the pgx_character layout
        arr["position"][:n] = p
        arr["orientation"][:n] = np.broadcast_to(np.asarray(orientations, np.float32), (n, 4))
        arr["height"][:n], arr["radius"][:n], arr["slope_limit"][:n] = height, radius, slope_limit
        arr["max_step_height"][:n], arr["num_iterations"][:n], arr["world"][:n] = max_step_height, num_iterations, world
        tr = np.ascontiguousarray(np.broadcast_to(np.asarray(translations, np.float32), (n, 3)))
        flags = np.zeros(max(n, 1), np.uint8)
        check(self.L.pgx_world_character_move(self.h, n, arr.ctypes.data_as(ctypes.POINTER(PgxCharacter)), _f(tr),
                                              float(dt), flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return arr["position"][:n].copy(), flags[:n].copy()

    def trigger_events(self, capacity=1024):
        """pgx_world_get_trigger_events: the last step's (enter, exit) events, each an
        (n, 2) array of (collider of the lower entity, collider of the other), ascending."""
        i32 = ctypes.POINTER(ctypes.c_int32)
        while True:
            en = np.zeros((max(capacity, 1), 2), np.int32)
            ex = np.zeros((max(capacity, 1), 2), np.int32)
            ne, nx = ctypes.c_int32(0), ctypes.c_int32(0)
            rc = self.L.pgx_world_get_trigger_events(self.h, capacity, en.ctypes.data_as(i32), ctypes.byref(ne),
                                                     ex.ctypes.data_as(i32), ctypes.byref(nx))
            if rc == ERR_CAPACITY:
                capacity = max(ne.value, nx.value)
                continue
            check(rc)
            return en[: ne.value].copy(), ex[: nx.value].copy()

    def save_state(self):
        """pgx_world_save_state: the world's state as bytes."""
        n = ctypes.c_size_t(0)
        check(self.L.pgx_world_state_size(self.h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        check(self.L.pgx_world_save_state(self.h, buf, n.value))
        return buf.raw

    def load_state(self, state):
        """pgx_world_load_state: replaces the world's state with bytes from save_state."""
        buf = ctypes.create_string_buffer(bytes(state), len(state))
        check(self.L.pgx_world_load_state(self.h, buf, len(state)))

    def load_state_matched(self, state):
        """pgx_world_load_state_matched: the contact cache, trigger lists and bounds of a state
        saved by a world whose colliders differ, carried to the colliders with the same world,
        entity and index. Returns the last step's trigger pairs that lost a collider, as rows
        (entity0, index0, entity1, index1)."""
        buf = ctypes.create_string_buffer(bytes(state), len(state))
        cap = 64
        while True:
            gone = np.zeros((cap, 4), np.uint32)
            n = ctypes.c_int32(0)
            rc = self.L.pgx_world_load_state_matched(self.h, buf, len(state), cap,
                                                     gone.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                                     ctypes.byref(n))
            if rc == ERR_CAPACITY and n.value > cap:
                cap = n.value
                continue
            check(rc)
            return gone[: n.value].copy()

    def exec_mode(self):
        m = ctypes.c_int(0)
        check(self.L.pgx_world_get_exec(self.h, ctypes.byref(m)))
        return {v: k for k, v in self.EXEC_MODES.items()}[m.value]

    def reset_stats(self):
        check(self.L.pgx_world_reset_stats(self.h))

    def stats(self):
        st = PgxStats()
        check(self.L.pgx_world_get_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in PgxStats._fields_}

    def timing(self):
        v = [ctypes.c_float() for _ in range(5)]
        check(self.L.pgx_world_get_timing(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(["broadphase_us", "narrowphase_us", "coloring_us", "solver_us", "total_us"],
                        [x.value for x in v]))


def _one_world_desc(scene):
    desc = scene.struct(PgxSceneDesc)
    desc.actor_world = ctypes.cast(None, I32)
    desc.n_worlds = 1
    return desc


class Server:
    """pgx_server: continuous batching of independent scenes ("requests") in one device
    world. submit() queues a scene with a horizon; every step() retires finished requests,
    admits queued ones within the HBM footprint budget and steps all admitted requests once."""

    def __init__(self, substeps=4, iterations=2, gravity=9.81, device=0, budget_bytes=0.0, max_worlds=0,
                 admit_interval=1):
        self.L = lib()
        cfg = PgxConfig(substeps, iterations, gravity, device, 0, 0)
        scfg = PgxServerConfig(budget_bytes, max_worlds, admit_interval)
        h = ctypes.c_void_p()
        check(self.L.pgx_server_create(ctypes.byref(cfg), ctypes.byref(scfg), ctypes.byref(h)))
        self.h = h
        _open_handles.add(self)
        self.scenes = {}

    def close(self):
        """pgx_server_destroy: waits for the server's stream and frees its device memory."""
        if getattr(self, "h", None):
            self.L.pgx_server_destroy(self.h)
            self.h = None
        _open_handles.discard(self)

    def __del__(self):
        if not _finalizing():
            self.close()

    @staticmethod
    def footprint(scene):
        b = ctypes.c_double(0)
        desc = _one_world_desc(scene)
        check(lib().pgx_server_footprint(ctypes.byref(desc), ctypes.byref(b)))
        return b.value

    def submit(self, scene, horizon):
        scene.arrays()
        desc = _one_world_desc(scene)
        nj = len(scene.joints)
        js = scene.joint_array(PgxJointDesc) if nj else None
        rid = ctypes.c_int(0)
        check(self.L.pgx_server_submit(self.h, ctypes.byref(desc), js, nj, int(horizon), ctypes.byref(rid)))
        self.scenes[rid.value] = scene
        return rid.value

    def set_graph(self, enabled=True):
        """Launch the batched world's substep loop as one HIP graph per step."""
        check(self.L.pgx_server_set_graph(self.h, 1 if enabled else 0))

    def set_tuning(self, tuning):
        """Launch shapes of the batched world, e.g. World.autotune()'s choice."""
        t = PgxTuning(**tuning)
        check(self.L.pgx_server_set_tuning(self.h, ctypes.byref(t)))

    def step(self, dt):
        info = PgxServerStepInfo()
        check(self.L.pgx_server_step(self.h, dt, ctypes.byref(info)))
        return {k: getattr(info, k) for k, _ in PgxServerStepInfo._fields_}

    def done(self, rid):
        rc = self.L.pgx_server_done(self.h, rid)
        if rc < 0:
            check(rc)
        return rc == 1

    def take_result(self, rid):
        s = self.scenes.pop(rid)
        pos = np.zeros((s.n_actors, 3), np.float32)
        quat = np.zeros((s.n_actors, 4), np.float32)
        vel = np.zeros((s.n_dyn, 3), np.float32)
        ang = np.zeros((s.n_dyn, 3), np.float32)
        check(self.L.pgx_server_take_result(self.h, rid, _f(pos), _f(quat), _f(vel), _f(ang)))
        return pos, quat, vel, ang
