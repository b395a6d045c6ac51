"""ctypes loaders for the CPU oracle (oracle/liborc.so) and for the reference harness
(oracle/_ref/libphysecs_ref.so). Test infrastructure: imported only by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg."""
import ctypes
import os
import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_PATH = os.path.join(ROOT, "oracle", "liborc.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libphysecs_ref.so")

F = ctypes.POINTER(ctypes.c_float)
I32 = ctypes.POINTER(ctypes.c_int32)
U32 = ctypes.POINTER(ctypes.c_uint32)
U8 = ctypes.POINTER(ctypes.c_uint8)


def fptr(a):
    return a.ctypes.data_as(F)


def iptr(a):
    return a.ctypes.data_as(I32)


class SceneStruct(ctypes.Structure):
    """Mirror of the scene descriptor used by both the oracle and the GPU C ABI."""
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
    ]


def _ref_cpu_ok():
    """The reference harness is built with AVX-512VL (the reference's GJK uses it)."""
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        return "avx512vl" in flags
    except OSError:
        return False


_orc = None
_ref = None


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_PATH):
            raise RuntimeError(f"oracle library missing: {ORC_PATH} (run `make -C oracle`)")
        lib = ctypes.CDLL(ORC_PATH)
        lib.orc_world_create.restype = ctypes.c_void_p
        lib.orc_world_create.argtypes = [ctypes.POINTER(SceneStruct)]
        for name in ("orc_world_destroy",):
            getattr(lib, name).argtypes = [ctypes.c_void_p]
        lib.orc_world_set_params.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int]
        lib.orc_world_set_non_colliding.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
        lib.orc_world_set_order.argtypes = [ctypes.c_void_p, I32, ctypes.c_int]
        lib.orc_world_step.argtypes = [ctypes.c_void_p, ctypes.c_float]
        lib.orc_world_get_state.argtypes = [ctypes.c_void_p, F, F, F, F]
        lib.orc_world_get_bounds.argtypes = [ctypes.c_void_p, F]
        lib.orc_world_num_pairs.argtypes = [ctypes.c_void_p]
        lib.orc_world_get_pairs.argtypes = [ctypes.c_void_p, I32]
        lib.orc_world_num_manifolds.argtypes = [ctypes.c_void_p]
        lib.orc_world_get_manifolds.argtypes = [ctypes.c_void_p, I32, I32, F]
        lib.orc_collision.argtypes = [ctypes.c_int, F, F, F, ctypes.c_int, F, F, F, F]
        lib.orc_bounds.argtypes = [ctypes.c_int, F, F, F, F]
        lib.orc_sqr_dist_segment_aabb.argtypes = [F, F, ctypes.c_float, ctypes.c_float, F, F]
        lib.orc_contact_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, F, F, F, F, ctypes.c_float,
                                          ctypes.c_int, ctypes.c_int, F]
        lib.orc_integrate_velocity.argtypes = [F, F, F, F, ctypes.c_float, ctypes.c_float, F]
        lib.orc_integrate_position.argtypes = [F, F, F, F, F, F, ctypes.c_int, F, ctypes.c_float, F]
        _orc = lib
    return _orc


def ref_available():
    return os.path.exists(REF_PATH) and _ref_cpu_ok()


def ref():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_PATH)
        lib.ref_collision.argtypes = [ctypes.c_int, F, F, F, ctypes.c_int, F, F, F, F]
        lib.ref_bounds.argtypes = [ctypes.c_int, F, F, F, F]
        lib.ref_sqr_dist_segment_aabb.argtypes = [F, F, ctypes.c_float, ctypes.c_float, F, F]
        lib.ref_mass_props.argtypes = [ctypes.c_int, F, F, F, ctypes.c_float, F]
        lib.ref_mass_props_multi.argtypes = [ctypes.c_int, I32, F, F, F, ctypes.POINTER(ctypes.c_uint8),
                                             ctypes.c_float, F]
        lib.ref_contact_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, F, F, F, F, ctypes.c_float,
                                          ctypes.c_int, ctypes.c_int, F]
        lib.ref_integrate_velocity.argtypes = [F, F, F, F, ctypes.c_float, ctypes.c_float, F]
        lib.ref_integrate_position.argtypes = [F, F, F, F, F, F, ctypes.c_int, F, ctypes.c_float, F]
        _ref = lib
    return _ref


def _f32(x, n):
    a = np.zeros(n, np.float32)
    x = np.asarray(x, np.float32).ravel()
    a[: len(x)] = x
    return a


def ray_aabb(which, o, d, mn, mx):
    """intersectRayAABB of `which` ("orc" or "ref"): (hit, t)."""
    lib = orc() if which == "orc" else ref()
    fn = getattr(lib, which + "_ray_aabb")
    fn.argtypes = [F, F, F, F, F]
    t = np.full(1, -1.0, np.float32)
    hit = fn(fptr(_f32(o, 3)), fptr(_f32(d, 3)), fptr(_f32(mn, 3)), fptr(_f32(mx, 3)), fptr(t))
    return hit, t[0]


def ray_geometry(which, o, d, t, prm, pos, q):
    """intersectRayGeometry of `which`: (hit, t); convex params[3] = register_convex id."""
    lib = orc() if which == "orc" else ref()
    fn = getattr(lib, which + "_ray_geometry")
    fn.argtypes = [F, F, ctypes.c_int, F, F, F, F]
    out = np.full(1, -1.0, np.float32)
    hit = fn(fptr(_f32(o, 3)), fptr(_f32(d, 3)), int(t), fptr(_f32(prm, 4)), fptr(_f32(pos, 3)), fptr(_f32(q, 4)),
             fptr(out))
    return hit, out[0]


def overlap(which, t0, p0, pos0, q0, t1, p1, pos1, q1):
    """physecs::overlap of `which` ("orc" or "ref") for one pair: 0 or 1."""
    lib = orc() if which == "orc" else ref()
    fn = getattr(lib, which + "_overlap")
    fn.argtypes = [ctypes.c_int, F, F, F, ctypes.c_int, F, F, F]
    return fn(int(t0), fptr(_f32(p0, 4)), fptr(_f32(pos0, 3)), fptr(_f32(q0, 4)),
              int(t1), fptr(_f32(p1, 4)), fptr(_f32(pos1, 3)), fptr(_f32(q1, 4)))


def collision(lib, prefix, t0, p0, pos0, q0, t1, p1, pos1, q1):
    """Returns (hit, out[28]) for one narrowphase call of `lib` (orc or ref)."""
    out = np.zeros(28, np.float32)
    args = [_f32(p0, 4), _f32(pos0, 3), _f32(q0, 4), _f32(p1, 4), _f32(pos1, 3), _f32(q1, 4)]
    fn = getattr(lib, prefix + "_collision")
    hit = fn(t0, fptr(args[0]), fptr(args[1]), fptr(args[2]), t1, fptr(args[3]), fptr(args[4]), fptr(args[5]),
             fptr(out))
    return hit, out


class OrcJoint(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("actor0", ctypes.c_int32), ("actor1", ctypes.c_int32),
                ("anchor0", ctypes.c_float * 7), ("anchor1", ctypes.c_float * 7), ("drive_enabled", ctypes.c_int32),
                ("drive_velocity", ctypes.c_float), ("drive_max_torque", ctypes.c_float),
                ("lower_limit", ctypes.c_float), ("upper_limit", ctypes.c_float), ("target", ctypes.c_float),
                ("drive_stiffness", ctypes.c_float), ("drive_damping", ctypes.c_float), ("gear_ratio", ctypes.c_float)]


class OracleWorld:
    """The oracle's restated Scene::simulate over a scene descriptor (see physecs_amd.scenes)."""

    def __init__(self, scene, substeps=4, iterations=2, gravity=9.81, threads=0):
        self.lib = orc()
        self._keep = scene.arrays()
        self.scene = scene
        st = scene.struct(SceneStruct)
        self.h = self.lib.orc_world_create(ctypes.byref(st))
        self.lib.orc_world_set_params(self.h, substeps, iterations, gravity, threads)
        for e0, e1 in getattr(scene, "non_colliding", []):
            self.lib.orc_world_set_non_colliding(self.h, e0, e1)
        if getattr(scene, "joints", None):
            arr = scene.joint_array(OrcJoint)
            self.lib.orc_world_add_joints.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(OrcJoint)]
            self.lib.orc_world_add_joints(self.h, len(scene.joints), arr)

    def set_contact_filter(self, fn):
        """The oracle's contactFilter, tabulated as for the device (scenes.contact_filter_table)."""
        from physecs_amd import scenes
        self.lib.orc_world_set_contact_filter.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                                          ctypes.POINTER(ctypes.c_uint8), I32]
        if fn is None:
            self.lib.orc_world_set_contact_filter(self.h, 0, None, None)
            return
        k, table, classes = scenes.contact_filter_table(self.scene, fn)
        self._filter = (table, classes)
        self.lib.orc_world_set_contact_filter(self.h, k, table.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                              iptr(classes))

    def set_joint_params(self, index, **params):
        """The joint's setters on the oracle (see api.World.set_joint_params)."""
        from physecs_amd import scenes
        if not hasattr(self, "_joints"):
            self._joints = [tuple(j[:5]) + (dict(j[5]),) for j in self.scene.joints]
        t, a0, a1, an0, an1, p = self._joints[index]
        p.update(params)
        d = OrcJoint()
        scenes.Scene.fill_joint(d, t, a0, a1, an0, an1, p)
        self.lib.orc_world_set_joint_params.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(OrcJoint)]
        assert self.lib.orc_world_set_joint_params(self.h, index, ctypes.byref(d)) == 0

    def destroy_joint(self, index):
        self.lib.orc_world_destroy_joint.argtypes = [ctypes.c_void_p, ctypes.c_int]
        assert self.lib.orc_world_destroy_joint(self.h, index) == 0

    def set_joint_callbacks(self, layout, rows):
        """The custom joints' host functions, as api.World.set_joint_callbacks."""
        from physecs_amd.api import JOINT_LAYOUT_FN, JOINT_ROWS_FN, joint_callbacks
        self._joint_fns = joint_callbacks(layout, rows)
        self.lib.orc_world_set_joint_callbacks.argtypes = [ctypes.c_void_p, JOINT_LAYOUT_FN, JOINT_ROWS_FN,
                                                            ctypes.c_void_p]
        self.lib.orc_world_set_joint_callbacks(self.h, self._joint_fns[0], self._joint_fns[1], None)

    def joint_error(self):
        self.lib.orc_world_joint_error.argtypes = [ctypes.c_void_p]
        return self.lib.orc_world_joint_error(self.h)

    def joint_colors(self):
        self.lib.orc_world_joint_color.argtypes = [ctypes.c_void_p, ctypes.c_int]
        return [self.lib.orc_world_joint_color(self.h, i) for i in range(len(self.scene.joints))]

    def joint_rows(self, index):
        """The rows joint `index` gets at the current poses, as an (n, 18) array: flags, b0, b1,
        linear, angular0, angular1, c, target velocity, min, max, frequency, damping."""
        self.lib.orc_world_joint_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, F]
        buf = np.zeros((7, 18), np.float32)
        n = self.lib.orc_world_joint_rows(self.h, index, fptr(buf))
        assert n >= 0
        return buf[:n].copy()

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_world_destroy(self.h)
            self.h = None

    def set_order(self, keys):
        keys = np.ascontiguousarray(keys, np.int32)
        self.lib.orc_world_set_order(self.h, iptr(keys), len(keys))

    def step(self, dt):
        self.lib.orc_world_step(self.h, dt)

    def state(self):
        a, d = self.scene.n_actors, self.scene.n_dyn
        pos = np.zeros((a, 3), np.float32)
        quat = np.zeros((a, 4), np.float32)
        vel = np.zeros((d, 3), np.float32)
        ang = np.zeros((d, 3), np.float32)
        self.lib.orc_world_get_state(self.h, fptr(pos), fptr(quat), fptr(vel), fptr(ang))
        return pos, quat, vel, ang

    def bounds(self):
        out = np.zeros((self.scene.n_colliders, 6), np.float32)
        self.lib.orc_world_get_bounds(self.h, fptr(out))
        return out

    def pairs(self):
        n = self.lib.orc_world_num_pairs(self.h)
        out = np.zeros((n, 2), np.int32)
        if n:
            self.lib.orc_world_get_pairs(self.h, iptr(out))
        return out

    def raycast(self, origins, dirs, max_dist, ignore=None):
        """Scene::raycastClosest for each ray: (collider index or -1, hit distance)."""
        fn = self.lib.orc_world_raycast
        fn.argtypes = [ctypes.c_void_p, F, F, ctypes.c_float, ctypes.c_uint32, F]
        fn.restype = ctypes.c_int
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        md = np.broadcast_to(np.asarray(max_dist, np.float32), (len(o),))
        ig = np.full(len(o), 0xFFFFFFFF, np.uint32) if ignore is None else np.asarray(ignore, np.uint32)
        col = np.full(len(o), -1, np.int32)
        t = np.zeros(len(o), np.float32)
        for i in range(len(o)):
            ti = np.zeros(1, np.float32)
            col[i] = fn(self.h, fptr(np.ascontiguousarray(o[i])), fptr(np.ascontiguousarray(d[i])),
                        np.float32(md[i]), int(ig[i]), fptr(ti))
            t[i] = ti[0] if col[i] >= 0 else 0.0
        return col, t

    def overlap(self, t, prm, pos, q, filt=0, cap=4096):
        """Scene::overlap of one query shape: the overlapping collider indices, ascending."""
        fn = self.lib.orc_world_overlap
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int, F, F, F, ctypes.c_int, ctypes.c_int, I32]
        fn.restype = ctypes.c_int
        out = np.zeros(cap, np.int32)
        n = fn(self.h, int(t), fptr(_f32(prm, 4)), fptr(_f32(pos, 3)), fptr(_f32(q, 4)), int(filt), cap, iptr(out))
        assert n <= cap
        return out[:n].copy()

    def overlap_mtd(self, t, prm, pos, q, cap=4096):
        """Scene::overlapWithMinTranslationalDistance of one query shape: (collider indices,
        (n, 4) normal xyz + mtd), by collider, then manifold order."""
        fn = self.lib.orc_world_overlap_mtd
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int, F, F, F, ctypes.c_int, I32, F]
        fn.restype = ctypes.c_int
        cols = np.zeros(cap, np.int32)
        nm = np.zeros((cap, 4), np.float32)
        n = fn(self.h, int(t), fptr(_f32(prm, 4)), fptr(_f32(pos, 3)), fptr(_f32(q, 4)), cap, iptr(cols), fptr(nm))
        assert 0 <= n <= cap
        return cols[:n].copy(), nm[:n].copy()

    def character_move(self, pos, quat, translation, height, radius, slope_limit, num_iterations):
        """CharacterController::move of one character: (new position, flags)."""
        fn = self.lib.orc_character_move
        fn.argtypes = [ctypes.c_void_p, F, F, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int, F]
        fn.restype = ctypes.c_int
        p = _f32(pos, 3)
        flags = fn(self.h, fptr(p), fptr(_f32(quat, 4)), float(height), float(radius), float(slope_limit),
                   int(num_iterations), fptr(_f32(translation, 3)))
        return p, flags

    def patch_transform(self, actor, pos, quat):
        """registry.patch<TransformComponent>: the pose and the entity's bounds (with margin)."""
        fn = self.lib.orc_world_patch_transform
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int, F, F]
        fn(self.h, int(actor), fptr(_f32(pos, 3)), fptr(_f32(quat, 4)))

    def set_kinematic(self, actor, kinematic):
        """Scene::setIsKinematic."""
        fn = self.lib.orc_world_set_kinematic
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        fn(self.h, int(actor), int(bool(kinematic)))

    def trigger_events(self):
        """The last step's trigger (enter, exit) events: (n, 2) collider pairs, ascending."""
        fn = self.lib.orc_world_trigger_events
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, I32]
        fn.restype = ctypes.c_int
        out = []
        for which in (0, 1):
            n = fn(self.h, which, 0, iptr(np.zeros(2, np.int32)))
            buf = np.zeros((max(n, 1), 2), np.int32)
            fn(self.h, which, n, iptr(buf))
            out.append(buf[:n].copy())
        return tuple(out)

    def trigger_events_ec(self):
        """The last step's trigger (enter, exit) events as the listeners get them: (n, 4) rows
        (entity0, colliderIndex0, entity1, colliderIndex1)."""
        fn = self.lib.orc_world_trigger_events_ec
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)]
        fn.restype = ctypes.c_int
        out = []
        for which in (0, 1):
            n = fn(self.h, which, 0, np.zeros(4, np.uint32).ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
            buf = np.zeros((max(n, 1), 4), np.uint32)
            fn(self.h, which, n, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
            out.append(buf[:n].copy())
        return tuple(out)

    def add_collider(self, actor, gtype, params, material=(0.6, 0.0, 0.0), lpos=(0, 0, 0), lquat=(0, 0, 0, 1),
                     trigger=False, enable=True, data=0):
        """Scene::addCollider: returns the new collider's handle (its place after the scene's
        colliders)."""
        fn = self.lib.orc_world_add_collider
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, F, F, F, F, ctypes.c_int, ctypes.c_int,
                       ctypes.c_int]
        fn.restype = ctypes.c_int
        p = np.zeros(4, np.float32)
        p[: len(params)] = params
        return fn(self.h, int(actor), int(gtype), fptr(p), fptr(_f32(lpos, 3)), fptr(_f32(lquat, 4)),
                  fptr(_f32(material, 3)), int(bool(trigger)), int(bool(enable)), int(data))

    def clear_colliders(self, actor):
        """Scene::clearColliders."""
        fn = self.lib.orc_world_clear_colliders
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int]
        fn(self.h, int(actor))

    def manifolds(self):
        n = self.lib.orc_world_num_manifolds(self.h)
        keys = np.zeros((n, 5), np.int32)
        ints = np.zeros((n, 4), np.int32)
        f = np.zeros((n, 56), np.float32)
        if n:
            self.lib.orc_world_get_manifolds(self.h, iptr(keys), iptr(ints), fptr(f))
        return keys, ints, f


def _row_args(lib):
    lib.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, F, F, F, F, ctypes.c_float, ctypes.c_int, F]


def row_solve(which, flags, b0, b1, inp, masses, vel, pv, h, iters):
    """One joint row: which = 'ref' (reference Constraint1D), 'scalar' or 'simd' (oracle)."""
    vel = np.array(vel, np.float32)
    pv = np.array(pv, np.float32)
    out = np.zeros(2, np.float32)
    inp = np.ascontiguousarray(inp, np.float32)
    masses = np.ascontiguousarray(masses, np.float32)
    if which == "ref":
        fn = ref().ref_row_scalar
        fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, F, F, F, F, ctypes.c_float, ctypes.c_int, F]
        fn(flags, b0, b1, fptr(inp), fptr(masses), fptr(vel), fptr(pv), np.float32(h), iters, fptr(out))
    else:
        fn = orc().orc_row
        fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, F, F, F, F, ctypes.c_float,
                       ctypes.c_int, F]
        fn(1 if which == "simd" else 0, flags, b0, b1, fptr(inp), fptr(masses), fptr(vel), fptr(pv), np.float32(h),
           iters, fptr(out))
    return vel, pv, out


# Planted mistakes in the oracle (oracle/orc_world.cpp, enum Mutation), for tests that show
# a pin fails on a plausible misreading of the reference.
MUT_SIMD_WARM_LOCAL, MUT_SIMD_LIMIT_UNSCALED, MUT_SIMD_MASK_NONSTRICT, MUT_SIMD_BIAS_FLOAT = 1, 2, 3, 4
MUT_SIMD_ZERO_WRITEBACK, MUT_FRICTION_SOLVER_VEL, MUT_NO_RESTITUTION_CACHE, MUT_RELAX_SOFT = 5, 6, 7, 8
MUT_JOINT_AXIS_ORDER, MUT_JOINT_NO_COM, MUT_JOINT_UNIVERSAL_AXIS, MUT_FILTER_ARG_ORDER = 9, 10, 11, 12
MUT_CLEAR_DROPS_CACHE = 13


class mutation:
    """`with mutation(m):` runs the oracle with planted mistake m, then restores it."""

    def __init__(self, m):
        self.m = m

    def __enter__(self):
        orc().orc_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        orc().orc_set_mutation(0)
        return False


def row_solve_ex(which, flags, b0, b1, inp, masses, vel, pv, h, iters, lambda0, warm_mask):
    """row_solve with a starting accumulated lambda and a per-solve warm-start mask (bit i:
    warm start on solve i); which = 'ref' (reference Constraint1D), 'scalar' or 'simd'."""
    vel = np.array(vel, np.float32)
    pv = np.array(pv, np.float32)
    out = np.zeros(2, np.float32)
    inp = np.ascontiguousarray(inp, np.float32)
    masses = np.ascontiguousarray(masses, np.float32)
    if which == "ref":
        fn = ref().ref_row_scalar_ex
        fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, F, F, F, F, ctypes.c_float, ctypes.c_int,
                       ctypes.c_float, ctypes.c_int, F]
        fn(flags, b0, b1, fptr(inp), fptr(masses), fptr(vel), fptr(pv), np.float32(h), iters, np.float32(lambda0),
           warm_mask, fptr(out))
    else:
        fn = orc().orc_row_ex
        fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, F, F, F, F, ctypes.c_float,
                       ctypes.c_int, ctypes.c_float, ctypes.c_int, F]
        fn(1 if which == "simd" else 0, flags, b0, b1, fptr(inp), fptr(masses), fptr(vel), fptr(pv), np.float32(h),
           iters, np.float32(lambda0), warm_mask, fptr(out))
    return vel, pv, out


_convex_ids = {}


def register_convex(mesh):
    """Registers a convex mesh (verts, faces, normals, centroids) with both the oracle and,
    where built, the reference harness; returns the index both use (params[3])."""
    key = id(mesh)
    if key in _convex_ids:
        return _convex_ids[key][1]
    verts, faces, normals, centroids = mesh
    v = np.ascontiguousarray(verts, np.float32)
    fs = np.cumsum([0] + [len(f) for f in faces]).astype(np.int32)
    fi = np.ascontiguousarray(np.concatenate([np.asarray(f, np.int32) for f in faces]), np.int32)
    n = np.ascontiguousarray(normals, np.float32)
    c = np.ascontiguousarray(centroids, np.float32)
    args = (len(v), fptr(v), len(faces), iptr(fs), iptr(fi), fptr(n), fptr(c))
    lib = orc()
    lib.orc_set_convex_mesh.argtypes = [ctypes.c_int, F, ctypes.c_int, I32, I32, F, F]
    i0 = lib.orc_set_convex_mesh(*args)
    if ref_available():
        r = ref()
        r.ref_set_convex_mesh.argtypes = [ctypes.c_int, F, ctypes.c_int, I32, I32, F, F]
        i1 = r.ref_set_convex_mesh(*args)
        assert i0 == i1, (i0, i1)
    _convex_ids[key] = (mesh, i0)
    return i0


def register_trimesh(verts, tris):
    """Builds the triangle mesh in the oracle and, where built, in the reference harness
    (its own TriangleMesh constructor); returns the index both use (params[3])."""
    v = np.ascontiguousarray(verts, np.float32)
    t = np.ascontiguousarray(tris, np.uint32)
    U = ctypes.POINTER(ctypes.c_uint32)
    lib = orc()
    lib.orc_set_tri_mesh.argtypes = [ctypes.c_int, F, ctypes.c_int, U]
    i0 = lib.orc_set_tri_mesh(len(v), fptr(v), len(t), t.ctypes.data_as(U))
    if ref_available():
        r = ref()
        r.ref_set_tri_mesh.argtypes = [ctypes.c_int, F, ctypes.c_int, U]
        i1 = r.ref_set_tri_mesh(len(v), fptr(v), len(t), t.ctypes.data_as(U))
        assert i0 == i1, (i0, i1)
    return i0


def trimesh_layout(which, mesh, n_tris):
    U = ctypes.POINTER(ctypes.c_uint32)
    idx = np.zeros((n_tris, 3), np.uint32)
    nodes = np.zeros((2 * n_tris + 1, 8), np.float32)
    fn = orc().orc_tri_mesh_layout if which == "orc" else ref().ref_tri_mesh_layout
    fn.argtypes = [ctypes.c_int, U, F]
    n = fn(mesh, idx.ctypes.data_as(U), fptr(nodes))
    return idx, nodes[:n]


def collision_multi(which, t0, p0, pos0, q0, t1, p1, pos1, q1, cap=64):
    out = np.zeros((cap, 30), np.float32)
    args = [_f32(p0, 4), _f32(pos0, 3), _f32(q0, 4), _f32(p1, 4), _f32(pos1, 3), _f32(q1, 4)]
    fn = orc().orc_collision_multi if which == "orc" else ref().ref_collision_multi
    fn.argtypes = [ctypes.c_int, F, F, F, ctypes.c_int, F, F, F, ctypes.c_int, F]
    n = fn(t0, fptr(args[0]), fptr(args[1]), fptr(args[2]), t1, fptr(args[3]), fptr(args[4]), fptr(args[5]), cap,
           fptr(out))
    return n, out[:max(n, 0)]
