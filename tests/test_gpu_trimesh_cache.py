"""Worlds created with pgx_scene_desc.tri_bvh_cache (the TriangleMesh build loaded from
pgx_trimesh_cache bytes instead of built at creation) step bit for bit as worlds that build
it; a cache written for another mesh is rejected at creation."""
import numpy as np
import pytest

from gpu_util import bits, log
from physecs_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from physecs_amd import api as a
    assert a.device_count() > 0, "no GPU visible"
    return a


def test_cached_build_steps_identically(api):
    built = scenes.terrain_drop(4, quads=16, n=50)
    cached = scenes.terrain_drop(4, quads=16, n=50)
    cached.tri_bvh_cache = [api.trimesh_cache(v, t) for v, t in cached.trimesh]
    log("[trimesh cache] terrain with 50 bodies: 40 steps built vs cached")
    a = api.World(built, substeps=4, iterations=2)
    b = api.World(cached, substeps=4, iterations=2)
    for _ in range(40):
        a.step(1 / 60)
        b.step(1 / 60)
    assert all(np.array_equal(bits(x), bits(y)) for x, y in zip(a.state(), b.state()))
    assert np.array_equal(a.manifolds()[0], b.manifolds()[0])
    other = scenes.terrain_drop(5, quads=16, n=50)
    other.tri_bvh_cache = cached.tri_bvh_cache                     # built for another mesh
    with pytest.raises(api.PgxError, match="INVALID"):
        api.World(other, substeps=4, iterations=2)
