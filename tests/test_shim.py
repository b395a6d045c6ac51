"""The drop-in physecs::Scene (physecs_amd/shim/physecs_gpu_scene.{h,cpp}) compiled against the
reference's own headers, entt 3.12.2 and glm 0.9.9.8 (oracle/Makefile, target shimcheck),
gathering its descriptor from an entt::registry built the way an app builds one:

- for the scenes the parity tests use, the gathered descriptor equals scenes.py's field for
  field, bit for bit (actors by entity, the dynamic storage's packed order, colliders in
  collision-storage order then list order, convex tables);
- triangle meshes go through the reference's own TriangleMesh constructor and reach the
  descriptor as a cache of its post-build order and BVH; that cache equals, byte for byte,
  the one pgx_trimesh_cache builds from the input order (the library's own restatement);
- the dynamic storage order after erasures and destructions is entt's packed order
  (swap-and-pop), which is the reference's body index (src/Physecs.cpp:117,333-334);
- the inverse inertia leaves glm::mat3 column-major (a non-symmetric matrix pins it), and
  for bodies whose components the reference's setMassProps filled it equals
  pgx_mass_properties bit for bit.

Host code only; skipped where the reference's headers are absent."""
import os

import numpy as np
import pytest

from physecs_amd import api, scenes
import shim_lib

pytestmark = pytest.mark.skipif(not os.path.isdir(shim_lib.REFERENCE), reason="needs the reference's headers")

HEADER = 48   # bytes of a triangle-mesh cache's header


@pytest.fixture(scope="module", autouse=True)
def built():
    shim_lib.build()
    assert shim_lib.available()


def bits(a):
    a = np.asarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


FIELDS = ("entity", "pos", "quat", "dyn_of_actor", "actor_of_dyn", "kinematic", "vel", "angvel", "inv_mass", "com",
          "inv_inertia", "col_actor", "col_index", "col_pos", "col_quat", "col_type", "col_params", "col_material",
          "col_trigger", "col_enable", "col_data", "convex_vert_start", "convex_verts", "convex_face_start",
          "face_idx_start", "face_idx", "face_normal", "face_centroid", "tri_vert_start", "tri_verts", "tri_start")


def _first_use(scene):
    """scene with its meshes renumbered in the order colliders first use them (the Scene
    meets meshes as pointers, in collider order; a mesh no collider uses is not uploaded)."""
    arr = scene.arrays()
    out = scenes.Scene()
    out.__dict__.update({k: v for k, v in scene.__dict__.items() if k not in ("convex", "trimesh", "_frozen")})
    prm = arr["col_params"].copy()
    for gtype, attr in ((3, "convex"), (4, "trimesh")):
        order = []
        for c in np.nonzero(arr["col_type"] == gtype)[0]:
            m = int(prm[c, 3])
            if m not in order:
                order.append(m)
            prm[c, 3] = order.index(m)
        setattr(out, attr, [getattr(scene, attr)[m] for m in order])
    out._chunks = dict(scene._chunks)
    out._chunks["col_params"] = [prm]
    out._frozen = None
    return out


def _expected(scene):
    arr = scene.arrays()
    out = {k: np.asarray(arr[k]) for k in FIELDS}
    out["inv_inertia"] = out["inv_inertia"].reshape(-1, 9)
    if not scene.convex:   # scenes.py pads empty tables with one row
        for k in ("convex_verts", "face_idx", "face_normal", "face_centroid"):
            out[k] = out[k][:0]
    if not scene.trimesh:
        out["tri_verts"] = out["tri_verts"][:0]
    return out


def _compare(scene):
    scene = _first_use(scene)
    reg, ents = shim_lib.replay(scene)
    with reg:
        got = reg.gather()
    want = _expected(scene)
    assert np.array_equal(np.asarray(ents, np.uint32), want["entity"])
    for k in FIELDS:
        g, w = np.asarray(got[k]), np.asarray(want[k], got[k].dtype).reshape(got[k].shape)
        assert np.array_equal(bits(g), bits(w)), k
    assert got["n_worlds"] == 1 and got["n_convex"] == len(scene.convex) and got["n_trimesh"] == len(scene.trimesh)
    # triangle meshes: the reference's build, carried by the cache, equals the library's own
    arr = scene.arrays()
    for m, (v, t) in enumerate(scene.trimesh):
        own = api.trimesh_cache(v, t)
        cache = got["tri_bvh_cache"][m]
        assert len(cache) == len(own)
        assert cache[8:32] == own[8:32]               # nv, nt, nodes, depth, length
        assert cache[HEADER:] == own[HEADER:], m      # post-build triangles, normals, centroids, BVH
        t0, t1 = int(arr["tri_start"][m]), int(arr["tri_start"][m + 1])
        built = np.frombuffer(own[HEADER:HEADER + 16 * len(t)], np.int32).reshape(-1, 4)[:, :3]
        assert np.array_equal(got["tri_idx"][t0:t1], built)
        assert sorted(map(tuple, built.tolist())) == sorted(map(tuple, np.asarray(t).tolist()))
        assert api.trimesh_cache_validate(v, got["tri_idx"][t0:t1], cache) == 0
    return got


def _grouped(scene):
    """Colliders contiguous per actor, as one collision component per entity lists them."""
    ca = scene.arrays()["col_actor"]
    starts = np.r_[True, ca[1:] != ca[:-1]]
    return len(set(ca.tolist())) == int(starts.sum())


@pytest.mark.parametrize("make", [
    lambda: scenes.mechanism_scene(0),
    lambda: scenes.joint_scene(1),
    lambda: scenes.convex_pile(2, n=40),
    lambda: scenes.terrain_drop(3, quads=10, n=30),
    lambda: scenes.served_request(4, n_bodies=200, chain=True),
    lambda: scenes.random_cluster(5, n=50),
], ids=["mechanism", "joints", "convex_pile", "terrain_drop", "served_request", "cluster"])
def test_gathered_descriptor_equals_scenes(make):
    sc = make()
    assert _grouped(sc)
    _compare(sc)


def test_multi_collider_triggers_kinematic_statics():
    """Actors of several colliders with local poses, triggers, disabled colliders, data bits,
    a kinematic body and statics, bodies whose dynamic parts are created out of entity order."""
    rng = np.random.default_rng(7)
    sc = scenes.Scene()
    actors = [sc.add_actor(rng.uniform(-3, 3, 3), scenes.random_quats(rng, 1)[0]) for _ in range(12)]
    for a in actors[::-1][:8]:   # dynamic storage in reverse creation order
        invI = np.diag(rng.uniform(0.5, 2.0, 3)).astype(np.float32)
        sc.add_dynamic(a, rng.uniform(0.1, 1.0), rng.uniform(-0.1, 0.1, 3), invI, rng.normal(size=3),
                       rng.normal(size=3), kinematic=(a == actors[6]))
    for a in actors:
        for k in range(1 + a % 3):
            t = int(rng.integers(0, 3))
            prm = [0.3] if t == 0 else ([0.4, 0.2] if t == 1 else list(rng.uniform(0.2, 0.5, 3)))
            sc.add_collider(a, t, prm, material=tuple(rng.uniform(0, 1, 3)), lpos=rng.uniform(-0.5, 0.5, 3),
                            lquat=scenes.random_quats(rng, 1)[0], trigger=bool(k == 1 and a % 2),
                            enable=bool(a != 4 or k == 0), data=int(rng.integers(0, 8)))
    got = _compare(sc)
    assert got["kinematic"].sum() == 1 and got["col_trigger"].sum() > 0 and (got["col_enable"] == 0).any()


def _swap_and_pop(order, e):
    i = order.index(e)
    order[i] = order[-1]
    order.pop()


def test_dynamic_storage_order_after_erasures():
    """entt keeps a storage packed by moving its last element into an erased slot; the Scene
    takes that packed order as the body index, as the reference does, and recycled entities
    (a new version of a destroyed entity's index) keep their full value."""
    with shim_lib.Registry() as reg:
        ents = []
        dyn, col = [], []
        for i in range(40):
            e = reg.create()
            ents.append(e)
            reg.transform(e, (i, 0, 0), (0, 0, 0, 1))
            if i % 5 != 4:
                reg.dynamic(e, 1.0 + i, (0, 0, 0), np.eye(3) * (i + 1))
                dyn.append(e)
            reg.collision(e, [0], [[0.5, 0, 0, 0]], [0, 0, 0], [0, 0, 0, 1], [0.6, 0, 0], [0], [1], [i])
            col.append(e)
        for e in (ents[3], ents[17], ents[0]):
            reg.erase_dynamic(e)
            _swap_and_pop(dyn, e)
        for e in (ents[8], ents[21]):       # destroyed: out of both storages
            reg.destroy(e)
            _swap_and_pop(dyn, e)
            _swap_and_pop(col, e)
        for i in range(2):                   # recycled indices, new versions
            e = reg.create()
            assert e >> 20 >= 1 and (e & 0xFFFFF) in (8, 21)
            reg.transform(e, (100 + i, 0, 0), (0, 0, 0, 1))
            reg.dynamic(e, 0.5, (0, 0, 0), np.eye(3))
            reg.collision(e, [2], [[0.2, 0.3, 0.4, 0]], [0, 0, 0], [0, 0, 0, 1], [0.6, 0, 0], [0], [1], [99])
            dyn.append(e)
            col.append(e)
        assert reg.dynamic_order().tolist() == dyn
        got = reg.gather()
    ent = got["entity"]
    assert np.all(np.diff(ent.astype(np.int64)) > 0)             # actors by entity value
    assert ent[got["actor_of_dyn"]].tolist() == dyn              # dynamic storage order
    assert ent[got["col_actor"]].tolist() == col                 # collision storage order
    for d, a in enumerate(got["actor_of_dyn"]):
        assert got["dyn_of_actor"][a] == d
    # the body at storage slot d carries its own components
    inv_mass = {int(e): 1.0 + ents.index(e) if e in ents else 0.5 for e in dyn}
    assert np.array_equal(got["inv_mass"], np.asarray([inv_mass[e] for e in dyn], np.float32))


def test_inverse_inertia_is_column_major():
    """glm::mat3 element (row r, column c) is m[c][r]; the descriptor takes element (r, c)
    at 3c + r, as pgx_mass_properties writes it and the device reads it."""
    M = np.arange(1, 10, dtype=np.float32).reshape(3, 3) / 7   # not symmetric
    with shim_lib.Registry() as reg:
        e = reg.create()
        reg.transform(e, (0, 0, 0), (0, 0, 0, 1))
        reg.dynamic(e, 1.0, (0, 0, 0), M)
        reg.collision(e, [0], [[0.5, 0, 0, 0]], [0, 0, 0], [0, 0, 0, 1], [0.6, 0, 0], [0], [1], [0])
        got = reg.gather()
    assert np.array_equal(got["inv_inertia"][0], M.T.ravel())
    assert not np.array_equal(got["inv_inertia"][0], M.ravel())


def test_mass_props_from_reference_equal_library():
    """Bodies whose RigidBodyDynamicComponent the reference's setMassProps filled (one and
    several colliders with local poses, a convex mesh, a trigger that adds no mass): the
    gathered inv_mass, com and inverse inertia equal pgx_mass_properties' bit for bit."""
    rng = np.random.default_rng(11)
    hull = scenes.ellipsoid_hull(rng)
    bodies = [
        ([0], [[0.4, 0, 0, 0]], [[0, 0, 0]], [[0, 0, 0, 1]], [0]),
        ([1, 2], [[0.3, 0.2, 0, 0], [0.2, 0.3, 0.4, 0]], [[0.1, 0.2, 0], [-0.3, 0, 0.2]],
         [scenes.random_quats(rng, 1)[0], scenes.random_quats(rng, 1)[0]], [0, 0]),
        ([3], [[1.0, 0.8, 1.2, 0]], [[0, 0, 0]], [[0, 0, 0, 1]], [0]),
        ([2, 0], [[0.3, 0.3, 0.3, 0], [0.5, 0, 0, 0]], [[0, 0, 0], [1, 0, 0]], [[0, 0, 0, 1], [0, 0, 0, 1]], [0, 1]),
    ]
    with shim_lib.Registry() as reg:
        m = reg.convex_mesh(*hull)
        want = []
        for i, (types, prm, lp, lq, trig) in enumerate(bodies):
            e = reg.create()
            reg.transform(e, (i, 0, 0), (0, 0, 0, 1))
            reg.dynamic(e, 1.0, (0, 0, 0), np.eye(3))
            p = np.asarray(prm, np.float32)
            p[np.asarray(types) == 3, 3] = m
            n = len(types)
            reg.collision(e, types, p, lp, lq, [[0.6, 0, 0]] * n, trig, [1] * n, [0] * n)
            mass = 3.0 + i
            reg.set_mass_props(e, mass)
            want.append(api.mass_properties(types, p, lp, lq, mass, is_trigger=trig, meshes=[hull]))
        got = reg.gather()
    for d, (im, com, inv) in enumerate(want):
        assert bits(got["inv_mass"][d]) == bits(np.float32(im)), d
        assert np.array_equal(bits(got["com"][d]), bits(com)), d
        assert np.array_equal(bits(got["inv_inertia"][d]), bits(inv)), d
