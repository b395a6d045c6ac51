"""The library's host-side TriangleMesh build, as pgx_trimesh_cache writes it, against the
golden vectors written from the reference's own TriangleMesh constructor
(src/TriangleMesh.cpp:7-164): the post-build triangle order and every BVH node (bounds,
triangle count, index), bit for bit. Host code: no device needed."""
import os

import numpy as np

from physecs_amd import api

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = 32   # magic, version, vertices, triangles, nodes, depth (4 bytes each), hash (8)


def test_cache_matches_reference_build():
    d = np.load(os.path.join(HERE, "golden", "trimesh.npz"))
    verts, tris = d["tm_verts"], d["tm_tris"].astype(np.int32)
    blob = api.trimesh_cache(verts, tris)
    assert blob == api.trimesh_cache(verts, tris)                 # deterministic
    head = np.frombuffer(blob[:24], np.int32)
    nt, nodes = int(head[3]), int(head[4])
    assert nt == len(tris) and len(blob) == HEADER + 16 * (3 * nt + 2 * nodes)
    body = np.frombuffer(blob[HEADER:], np.uint32)
    t = body[: 4 * nt].reshape(nt, 4)
    assert np.array_equal(t[:, :3], d["bvh_idx"])                 # post-build triangle order
    nd = np.frombuffer(blob[HEADER + 16 * 3 * nt:], np.float32).reshape(nodes, 2, 4)
    ref = d["bvh_nodes"].view(np.float32)
    assert len(ref) == nodes
    assert np.array_equal(nd[:, 0, :3].view(np.uint32), ref[:, :3].view(np.uint32))      # min
    assert np.array_equal(nd[:, 1, :3].view(np.uint32), ref[:, 3:6].view(np.uint32))     # max
    assert np.array_equal(nd[:, 0, 3].view(np.int32), ref[:, 6].astype(np.int32))       # triangle count
    assert np.array_equal(nd[:, 1, 3].view(np.int32), ref[:, 7].astype(np.int32))       # index
