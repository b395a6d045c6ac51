"""Checks the CPU oracle, bit for bit, against the committed golden vectors in tests/golden/:
outputs of the reference's own compiled functions on seeded inputs, written by
tests/golden/make_golden.py. Runs where the reference sources are absent."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402


def load(name):
    return np.load(os.path.join(HERE, "golden", name))


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def convex_ids(d):
    return [ol.register_convex(m) for m in mg.meshes_from_tables(d)]


def test_collision_primitive_golden():
    """src/Collision.cpp:895-1016, all 9 ordered sphere/capsule/box pairs."""
    d = load("collision.npz")
    orc = ol.orc()
    types, cases, hits, outs = d["prim_types"], d["prim_cases"], d["prim_hits"], d["prim_outs"]
    assert (hits != 0).sum() > len(hits) // 4
    for k in range(len(cases)):
        c = mg.unpack_case(cases[k])
        h, o = ol.collision(orc, "orc", int(types[k, 0]), *c[:3], int(types[k, 1]), *c[3:])
        assert h == hits[k], k
        if h:
            assert np.array_equal(bits(o), outs[k]), (k, types[k])


def test_collision_convex_golden():
    """GJK/EPA (src/GJK.h, src/EPA.h) and the convex contact generators."""
    d = load("collision.npz")
    ids = convex_ids(d)
    orc = ol.orc()
    types, cases, hits, outs = d["cvx_types"], d["cvx_cases"], d["cvx_hits"], d["cvx_outs"]
    multi = 0
    for k in range(len(cases)):
        tt = (int(types[k, 0]), int(types[k, 1]))
        c = mg.unpack_case(mg.live_case(tt, cases[k], ids, -1))
        h, o = ol.collision(orc, "orc", tt[0], *c[:3], tt[1], *c[3:])
        assert h == hits[k], k
        if h:
            multi += o[0] > 1
            assert np.array_equal(bits(o), outs[k]), (k, tt)
    assert multi > 100


def test_trimesh_bvh_and_collision_golden():
    """TriangleMesh's BVH build (src/TriangleMesh.cpp:7-161) and collisionTriangleMesh
    (src/CollisionTriangleMesh.cpp:882-956): node list, triangle permutation, and every
    case's manifold list in order."""
    d = load("trimesh.npz")
    ids = convex_ids(d)
    m = ol.register_trimesh(d["tm_verts"], d["tm_tris"])
    idx, nodes = ol.trimesh_layout("orc", m, len(d["tm_tris"]))
    assert np.array_equal(idx, d["bvh_idx"])
    assert np.array_equal(bits(nodes), d["bvh_nodes"])
    types, cases, counts, outs = d["col_types"], d["col_cases"], d["col_counts"], d["col_outs"]
    off = 0
    for k in range(len(cases)):
        tt = (int(types[k, 0]), int(types[k, 1]))
        c = mg.unpack_case(mg.live_case(tt, cases[k], ids, m))
        n, o = ol.collision_multi("orc", tt[0], *c[:3], tt[1], *c[3:])
        assert n == counts[k], k
        assert np.array_equal(bits(o).reshape(-1, 30), outs[off:off + max(n, 0)]), (k, tt)
        off += max(n, 0)
    assert off == len(outs) and (counts > 1).sum() > 50


def test_bounds_and_segment_aabb_golden():
    """getBounds (src/BoundsUtil.cpp:23-85) and sqrDistSegmentAABB
    (src/GeomUtil.cpp:129-467, with the sums kept from :273,326)."""
    d = load("kernels.npz")
    ids = convex_ids(d)
    orc = ol.orc()
    for t, p, pose, out in zip(d["bounds_types"], d["bounds_params"], d["bounds_poses"], d["bounds_outs"]):
        p = np.array(p, np.float32)
        if t == 3:
            p[3] = ids[int(p[3])]
        pos, q = np.ascontiguousarray(pose[:3]), np.ascontiguousarray(pose[3:])
        a = np.zeros(6, np.float32)
        orc.orc_bounds(int(t), ol.fptr(p), ol.fptr(pos), ol.fptr(q), ol.fptr(a))
        assert np.array_equal(bits(a), out)
    for x, out in zip(d["seg_ins"], d["seg_outs"]):
        p, dr, he = (np.ascontiguousarray(x[i:i + 3]) for i in (0, 3, 6))
        hh = np.float32(x[9])
        a = np.zeros(5, np.float32)
        orc.orc_sqr_dist_segment_aabb(ol.fptr(p), ol.fptr(dr), -hh, hh, ol.fptr(he), ol.fptr(a))
        assert np.array_equal(bits(a), out)


def test_contact_solver_golden():
    """ContactConstraints preSolve, two biased iterations and the relaxation pass
    (src/ContactConstraints.cpp:4-124), hard and soft."""
    d = load("kernels.npz")
    orc = ol.orc()
    for k in range(len(d["cs_meta"])):
        b0, npts, soft = (int(x) for x in d["cs_meta"][k])
        inf, masses = np.array(d["cs_infs"][k]), np.array(d["cs_masses"][k])
        vel, comp = np.array(d["cs_vels"][k]), np.array(d["cs_comps"][k])
        a = np.zeros(20, np.float32)
        orc.orc_contact_solve(b0, 1, npts, ol.fptr(inf), ol.fptr(masses), ol.fptr(vel), ol.fptr(comp),
                              np.float32(1 / 240), 2, 0 if soft else 1, ol.fptr(a))
        assert np.array_equal(bits(a), d["cs_outs"][k]), k


def test_integrators_golden():
    """Gyroscopic velocity integration and position integration (src/Physecs.cpp:450-467,
    501-510)."""
    d = load("kernels.npz")
    orc = ol.orc()
    h = np.float32(1 / 240)
    for x, out in zip(d["int_vin"], d["int_vout"]):
        q, v, w, invI = (np.ascontiguousarray(x[a:b]) for a, b in ((0, 4), (4, 7), (7, 10), (10, 19)))
        a = np.zeros(15, np.float32)
        orc.orc_integrate_velocity(ol.fptr(q), ol.fptr(v), ol.fptr(w), ol.fptr(invI), h, np.float32(9.81), ol.fptr(a))
        assert np.array_equal(bits(a), out)
    for x, cnt, out in zip(d["int_pin"], d["int_pcnt"], d["int_pout"]):
        pos, q, v, w, pv, pw, com = (np.ascontiguousarray(x[a:b]) for a, b in
                                     ((0, 3), (3, 7), (7, 10), (10, 13), (13, 16), (16, 19), (19, 22)))
        a = np.zeros(7, np.float32)
        orc.orc_integrate_position(ol.fptr(pos), ol.fptr(q), ol.fptr(v), ol.fptr(w), ol.fptr(pv), ol.fptr(pw),
                                   int(cnt), ol.fptr(com), h, ol.fptr(a))
        assert np.array_equal(bits(a), out)


def test_joint_rows_golden():
    """Constraint1D<flags>::preSolve/solve (src/Constraint1D.cpp) over 3 iterations: warm
    start, soft and limited rows, static first body."""
    d = load("kernels.npz")
    for k in range(len(d["row_meta"])):
        flags, b0 = (int(x) for x in d["row_meta"][k])
        v, pv, o = ol.row_solve("scalar", flags, b0, 1, d["row_inps"][k], d["row_masses"][k], d["row_vels"][k],
                                np.zeros(14, np.float32), 1 / 240, 3)
        assert np.array_equal(bits(np.concatenate([v, pv, o])), d["row_outs"][k]), (k, flags)


def test_golden_catches_a_planted_mistake():
    """Mistakes planted in the oracle's inputs (gravity 9.8 for 9.81; one biased iteration
    dropped from the contact solve) change the outputs the fixtures hold."""
    d = load("kernels.npz")
    orc = ol.orc()
    wrong = 0
    for x, out in zip(d["int_vin"][:50], d["int_vout"][:50]):
        q, v, w, invI = (np.ascontiguousarray(x[a:b]) for a, b in ((0, 4), (4, 7), (7, 10), (10, 19)))
        a = np.zeros(15, np.float32)
        orc.orc_integrate_velocity(ol.fptr(q), ol.fptr(v), ol.fptr(w), ol.fptr(invI), np.float32(1 / 240),
                                   np.float32(9.8), ol.fptr(a))
        wrong += not np.array_equal(bits(a), out)
    assert wrong == 50
    wrong = 0
    for k in range(50):
        b0, npts, soft = (int(x) for x in d["cs_meta"][k])
        inf, masses = np.array(d["cs_infs"][k]), np.array(d["cs_masses"][k])
        vel, comp = np.array(d["cs_vels"][k]), np.array(d["cs_comps"][k])
        a = np.zeros(20, np.float32)
        orc.orc_contact_solve(b0, 1, npts, ol.fptr(inf), ol.fptr(masses), ol.fptr(vel), ol.fptr(comp),
                              np.float32(1 / 240), 1, 0 if soft else 1, ol.fptr(a))
        wrong += not np.array_equal(bits(a), d["cs_outs"][k])
    assert wrong >= 35   # cases whose every impulse clamps to zero are unchanged
