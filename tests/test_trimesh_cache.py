"""The library's host-side TriangleMesh build, as pgx_trimesh_cache writes it, against the
golden vectors written from the reference's own TriangleMesh constructor
(src/TriangleMesh.cpp:7-164): the post-build triangle order and every BVH node (bounds,
triangle count, index), bit for bit. Host code: no device needed."""
import os

import numpy as np

from physecs_amd import api

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = 48   # magic, version, vertices, triangles, nodes, depth (4 bytes each), bytes, hash, payload hash (8 each)


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


def _fnv(b):
    h = 1469598103934665603
    for x in b:
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _mesh():
    d = np.load(os.path.join(HERE, "golden", "trimesh.npz"))
    return d["tm_verts"], d["tm_tris"].astype(np.int32)


def test_validate_accepts_its_own_cache_and_rejects_others():
    verts, tris = _mesh()
    blob = api.trimesh_cache(verts, tris)
    assert api.trimesh_cache_validate(verts, tris, blob) == 0
    other = tris.copy()
    other[0] = other[0][::-1]
    assert api.trimesh_cache_validate(verts, other, blob) == -1          # built from another mesh
    assert api.trimesh_cache_validate(verts, tris, blob[:-16]) == -1     # truncated
    assert api.trimesh_cache_validate(verts, tris, blob[: HEADER - 1]) == -1
    assert api.trimesh_cache_validate(verts, tris, blob + b"\0" * 16) == -1   # longer than it says
    bad = bytearray(blob)
    bad[HEADER + 5] ^= 1                                                 # one payload bit
    assert api.trimesh_cache_validate(verts, tris, bytes(bad)) == -1


def _patched(blob, nt, node, field, value):
    """The cache with node `node`'s triangle count (field 0) or index (field 1) replaced and
    the payload hash redone, so only the structural checks stand between it and the device."""
    b = bytearray(blob)
    off = HEADER + 16 * 3 * nt + 32 * node + (12 if field == 0 else 28)
    b[off:off + 4] = np.int32(value).tobytes()
    b[40:48] = np.uint64(_fnv(bytes(b[HEADER:]))).tobytes()
    return bytes(b)


def test_validate_rejects_out_of_range_structure():
    verts, tris = _mesh()
    blob = api.trimesh_cache(verts, tris)
    nt, nodes = len(tris), int(np.frombuffer(blob[16:20], np.int32)[0])
    nd = np.frombuffer(blob[HEADER + 16 * 3 * nt:], np.int32).reshape(nodes, 8)
    internal = int(np.nonzero(nd[:, 3] == 0)[0][0])
    leaf = int(np.nonzero(nd[:, 3] > 0)[0][0])
    assert api.trimesh_cache_validate(verts, tris, _patched(blob, nt, 0, 1, nd[0, 7])) == 0   # unchanged: accepted
    assert api.trimesh_cache_validate(verts, tris, _patched(blob, nt, internal, 1, nodes - 1)) == -1   # child past the end
    assert api.trimesh_cache_validate(verts, tris, _patched(blob, nt, internal, 1, internal)) == -1    # child before itself
    assert api.trimesh_cache_validate(verts, tris, _patched(blob, nt, leaf, 1, nt)) == -1             # triangles past the end
    assert api.trimesh_cache_validate(verts, tris, _patched(blob, nt, leaf, 0, -3)) == -1             # negative count
