"""The library's host-side pgx_mass_properties (physecs::setMassProps, src/MassUtil.cpp)
against the reference: bit for bit on the golden vectors written from the reference's own
compiled MassUtil.cpp for compound bodies of 1-4 colliders (spheres, capsules, boxes and
scaled convex meshes at random local poses, some of them triggers), live against it where
it is built, and against closed forms for single primitives at the origin. Host code: no
device needed."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol
from physecs_amd import api, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402


def bits(x):
    return np.ascontiguousarray(x, np.float32).view(np.uint32)


def library(types, prm, lpos, lq, trig, mass, meshes):
    im, com, inv = api.mass_properties(types, prm, lpos, lq, mass, is_trigger=trig, meshes=meshes)
    return np.concatenate([[im], com, inv]).astype(np.float32)


def test_golden_compound_bodies():
    d = np.load(os.path.join(HERE, "golden", "massprops.npz"))
    meshes = mg.meshes_from_tables(d)
    starts = np.concatenate([[0], np.cumsum(d["mp_count"])])
    kinds = set()
    for b in range(len(d["mp_count"])):
        s, e = starts[b], starts[b + 1]
        out = library(d["mp_types"][s:e], d["mp_params"][s:e], d["mp_lpos"][s:e], d["mp_lquat"][s:e],
                      d["mp_trig"][s:e], d["mp_mass"][b], meshes)
        assert np.array_equal(bits(out), d["mp_out"][b]), b
        kinds.update(int(t) for t in d["mp_types"][s:e])
    assert kinds == {0, 1, 2, 3} and len(d["mp_count"]) >= 1000


@pytest.mark.skipif(not ol.ref_available(), reason="reference harness not built")
def test_live_against_reference():
    meshes = mg.convex_meshes()
    ids = [ol.register_convex(m) for m in meshes]
    rng = np.random.default_rng(44)
    import ctypes
    for _ in range(500):
        types, prm, lpos, lq, trig, mass = mg.massprops_case(rng, len(ids))
        live = prm.copy()
        live[types == 3, 3] = [ids[int(m)] for m in prm[types == 3, 3]]
        want = np.zeros(13, np.float32)
        ol.ref().ref_mass_props_multi(len(types), types.ctypes.data_as(ol.I32), ol.fptr(live), ol.fptr(lpos),
                                      ol.fptr(lq), trig.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), mass,
                                      ol.fptr(want))
        got = library(types, prm, lpos, lq, trig, mass, meshes)
        assert np.array_equal(bits(got), bits(want))


def test_closed_forms():
    q = [[0, 0, 0, 1]]
    im, com, inv = api.mass_properties([scenes.SPHERE], [[0.5]], [[0, 0, 0]], q, 2.0)
    assert im == np.float32(0.5) and not com.any()
    assert np.allclose(inv.reshape(3, 3), np.eye(3) / (0.4 * 2.0 * 0.25), rtol=1e-6)
    im, com, inv = api.mass_properties([scenes.BOX], [[0.5, 1.0, 1.5]], [[0, 0, 0]], q, 12.0)
    assert np.allclose(np.diag(inv.reshape(3, 3)), 1 / np.array([1 + 2.25, 0.25 + 2.25, 0.25 + 1]), rtol=1e-6)
    # two equal spheres: the centre of mass halfway, the parallel-axis terms m d^2 on x and z
    im, com, inv = api.mass_properties([scenes.SPHERE] * 2, [[0.5], [0.5]], [[0, -1, 0], [0, 1, 0]], q * 2, 2.0)
    assert np.allclose(com, 0, atol=1e-7)
    i_xx = 2 * (0.4 * 1.0 * 0.25) + 2 * 1.0 * 1.0
    assert np.allclose(np.diag(inv.reshape(3, 3)), [1 / i_xx, 1 / (2 * 0.4 * 0.25), 1 / i_xx], rtol=1e-6)
