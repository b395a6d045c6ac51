"""The library's host-side pgx_convex_hull against scipy's Qhull (through
physecs_amd.scenes.convex_hull, the hull the test scenes use): the same hull vertices in
the same order, the same faces (polygons of coplanar facets) with the same clockwise
winding seen from outside, normals and centroids, for ellipsoid clouds, a cube (square
faces) and random clouds with interior points; and degenerate clouds rejected."""
import numpy as np
import pytest

from physecs_amd import api, scenes


def canon(face):
    """A polygon's vertex cycle from its smallest index, keeping the winding."""
    k = int(np.argmin(face))
    return tuple(int(v) for v in face[k:] + face[:k])


def same_hull(points):
    points = np.asarray(points, np.float32)          # both builders see the same float inputs
    v0, f0, n0, c0 = scenes.convex_hull(points)
    v1, f1, n1, c1 = api.convex_hull(points)
    assert np.array_equal(v0, v1)
    a = {canon(list(f)): (n, c) for f, n, c in zip(f0, n0, c0)}
    b = {canon(list(f)): (n, c) for f, n, c in zip(f1, n1, c1)}
    assert a.keys() == b.keys()
    for k in a:
        # normals: each builder takes a face's plane from one of its triangles, in double
        assert np.allclose(a[k][0], b[k][0], atol=1e-5) and np.allclose(a[k][1], b[k][1], atol=1e-6)
    return v1, f1, n1


def test_matches_qhull():
    rng = np.random.default_rng(12)
    for n in (8, 24, 60, 200):
        for _ in range(5):
            d = rng.normal(size=(n, 3))
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            same_hull(d * rng.uniform(0.3, 0.8, 3))
    for _ in range(5):   # interior points too
        same_hull(rng.uniform(-1, 1, (100, 3)))
    v, f, nrm = same_hull([[x, y, z] for x in (-.5, .5) for y in (-.5, .5) for z in (-.5, .5)])
    assert len(f) == 6 and all(len(x) == 4 for x in f)
    for face, n in zip(f, nrm):   # clockwise seen from outside
        e = np.cross(v[face[1]] - v[face[0]], v[face[2]] - v[face[0]])
        assert np.dot(e, n) < 0


def test_degenerate_clouds_rejected():
    with pytest.raises(api.PgxError):
        api.convex_hull([[0, 0, 0], [1, 0, 0], [0, 1, 0]])                 # too few
    with pytest.raises(api.PgxError):
        api.convex_hull([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]])      # flat
