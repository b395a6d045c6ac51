"""Helpers shared by the GPU tests: run the device world and the oracle side by side."""
import sys

import numpy as np

import oracle_lib as ol


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pair_set(pairs):
    return set(map(tuple, np.asarray(pairs).tolist()))


def manifold_dict(keys, ints, f):
    out = {}
    for i in range(len(keys)):
        out[tuple(keys[i].tolist())] = (ints[i], f[i])
    return out


def lockstep(scene, steps, substeps=4, iterations=2, dt=1 / 60, check_every=1, on_step=None, mode="colors",
             tuning=None):
    """Steps the GPU world and the oracle from the same scene; before each oracle step the
    oracle is fed the device's solve order. Returns (gpu, oracle) after asserting the
    states are bit-identical at every checked step."""
    from physecs_amd import api
    g = api.World(scene, substeps=substeps, iterations=iterations)
    g.set_exec(mode)
    if tuning:
        g.set_tuning(**tuning)
    o = ol.OracleWorld(scene, substeps=substeps, iterations=iterations)
    for k in range(steps):
        g.step(dt)
        keys, ints, f = g.manifolds()
        o.set_order(keys)
        o.step(dt)
        if on_step:
            on_step(k, g, o)
        if (k + 1) % check_every == 0 or k == steps - 1:
            gs, os_ = g.state(), o.state()
            for name, a, b in zip(("pos", "quat", "vel", "angvel"), gs, os_):
                if not np.array_equal(bits(a), bits(b)):
                    diff = np.abs(a.astype(np.float64) - b.astype(np.float64))
                    idx = np.unravel_index(np.argmax(diff), diff.shape)
                    raise AssertionError(f"step {k}: {name} differs, max |d| = {diff.max():.3e} at {idx}: "
                                         f"gpu {a[idx[0]]} oracle {b[idx[0]]}")
    return g, o
