"""Benchmark of the rigid-body step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config plain32k|pyramid|batched|convex|served|triggers|raycast|overlap|character|snapshot|meshcache]
    python bench.py --copy-calibration      # a plain 1 GiB device copy, for scaling PMC byte counters

A step is one Scene::simulate(1/60) with 4 substeps and 2 iterations over the whole scene,
inputs resident in HBM. The default workload is the 32,768-body mixed scene (1/3 spheres,
capsules, boxes on a jittered 32^3 lattice at 1.5 m in an open 5-box container, 600 steps:
10 warm-up + 590 timed). Prints one JSON line (metric body-substeps/s, higher is better).

The timed window is always the end of the workload's run: with --warmup W and --steps K
adding up to less than the workload's length (600 steps for the default workload), the
first 600 - W - K steps run untimed (config.settle_steps), then W warm-up steps, then the K
timed ones, so a short run times the settled pile, not the first fall.

roofline: the dominant kernel is k_solve, the contact solve of one graph colour (one launch
per colour and pass: numIterations biased passes + 1 relaxation, every substep). Its
algorithmic bytes per launch are 284 B per contact point + 12 B per manifold of the colour
(rows 108 + scalars 24 + both bodies' solver and component v,w 96 read; lambdas 8 + both
bodies' v,w 48 written). Its duration comes from HIP start/stop events that
hipExtLaunchKernelGGL stamps from each launch's own dispatch packet, for every launch of the
first substep of every 16th timed step (--timing-period), on the world's stream: the events
and their reads slow the step that carries them (2.33 against 1.89 ms on the settled
default workload), so the steps between run without them, and p50/p99 are taken over
those (config.p99_step_ms_all over all timed steps). --exec flow launches one
k_solve_flow per pass instead (every colour in one launch: a manifold solves once its
bodies have taken the lower colours' solves), so a launch covers all of the pass's points
and manifolds; --exec cooperative times the single-launch substep kernel k_step_persistent
with the per-substep phase model.
traffic: HBM bytes per k_solve launch from the committed PMC summary (profiles/), which
records the window its counters were taken over; it is reported only when this run times
the same window of the same workload (config.window), and is null otherwise.

cpu_baseline: the CPU oracle (a restatement of the reference's Scene::simulate with the
reference's ThreadPool narrowphase chunking, nproc-1 workers + caller) on the same scene,
started from the device state after the warm-up, timed for about 15 s.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# PMC summaries of the roofline kernel, one per timed window they were summarised over
TRAFFIC_SUMMARIES = {("plain32k", "colors"): "k_solve_plain32k_traffic_*.md",
                     ("batched", "colors"): "k_solve_batched4096_traffic.md"}
# steps of each workload's run; the timed window is its end
RUN_STEPS = {"plain32k": 600, "pyramid": 300, "batched": 300, "convex": 600, "triggers": 200}


def measured_traffic(config, mode, window):
    """HBM bytes per launch of the roofline kernel, from the committed PMC summary of this
    window (FETCH_SIZE and WRITE_SIZE per dispatch over the timed window's launches, each
    scaled by the same counter's reading for a copy of known size). None when no summary
    was taken over this window."""
    import glob
    pattern = TRAFFIC_SUMMARIES.get((config, mode))
    if pattern is None:
        return None
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", pattern))):
        fields = {}
        with open(path) as f:
            for line in f:
                if ":" in line and not line.startswith(("#", "|", " ")):
                    k, v = line.split(":", 1)
                    fields[k.strip()] = v.strip()
        if fields.get("window") == window and "traffic_bytes_per_launch" in fields:
            return float(fields["traffic_bytes_per_launch"])
    return None


def copy_calibration():
    """Ten device-to-device copies of 1 GiB (reads 1 GiB, writes 1 GiB each): the reference
    reading for FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(10):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": True, "bytes_per_copy": 4 * n, "copies": 10}), flush=True)


def make_scene(name):
    from physecs_amd import scenes
    if name == "plain32k":
        return scenes.mixed_lattice(0, side=32), {"workload": "mixed 32768 bodies (sphere/capsule/box) in 5-box container",
                                                  "bodies": 32768, "substeps": 4, "iterations": 2, "dt": 1 / 60}
    if name == "pyramid":
        return scenes.box_pyramid(0), {"workload": "20-layer box pyramid (210 boxes)", "bodies": 210,
                                       "substeps": 4, "iterations": 2, "dt": 1 / 60}
    if name == "batched":
        s = scenes.batched_worlds(0, n_worlds=4096, ragdoll_frac=0.25)
        return s, {"workload": "4096 uneven worlds (8..1024 bodies, log-uniform; 25% with a 15-capsule ragdoll)",
                   "bodies": s.n_dyn, "worlds": 4096, "joints": len(s.joints), "substeps": 4, "iterations": 2,
                   "dt": 1 / 60}
    if name == "convex":
        s = scenes.convex_terrain(0)
        return s, {"workload": "8192 convex hulls + 512 revolute capsule chains on a 131072-triangle fBm terrain",
                   "bodies": s.n_dyn, "joints": len(s.joints), "triangles": 131072, "substeps": 4,
                   "iterations": 2, "dt": 1 / 60}
    if name == "triggers":
        s = scenes.mixed_lattice(0, side=32)
        s.arrays()["col_trigger"][5::10] = 1   # every tenth body collider a trigger (the container stays solid)
        return s, {"workload": "mixed 32768 bodies, every tenth a trigger (overlap tests and enter/exit lists)",
                   "bodies": 32768, "triggers": int(s.arrays()["col_trigger"].sum()), "substeps": 4, "iterations": 2,
                   "dt": 1 / 60}
    raise SystemExit(f"unknown config {name}")


def cpu_baseline(scene, state, substeps, iterations, dt, budget_s=15.0):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    pos, quat, vel, ang = state
    arr = scene.arrays()
    saved = {k: arr[k].copy() for k in ("pos", "quat", "vel", "angvel")}
    arr["pos"][:] = pos
    arr["quat"][:] = quat
    arr["vel"][:] = vel
    arr["angvel"][:] = ang
    threads = max(0, (os.cpu_count() or 1) - 1)
    o = ol.OracleWorld(scene, substeps=substeps, iterations=iterations, threads=threads)
    for k, v in saved.items():
        arr[k][:] = v
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step(dt)
        steps += 1
        if time.perf_counter() - t0 >= budget_s or steps >= 200:
            break
    el = time.perf_counter() - t0
    return {"value": scene.n_dyn * substeps * steps / el, "unit": "body-substeps/s", "cores": threads + 1,
            "kind": "port", "sample": f"{steps} steps of the same scene from the device state after warm-up, "
                                      f"{el:.1f} s; narrowphase on {threads} workers + caller, solver single-threaded"}


def cpu_baseline_worlds(scene, state, substeps, iterations, dt, budget_s=15.0):
    """The batched workload's CPU baseline: a sequential loop over per-world Scenes. Each
    sampled world is its own oracle world (the world's bodies, colliders and joints, from the
    device state after warm-up), stepped on its own; worlds are taken in a fixed stride over
    the batch so the sample spans the size range, each stepped 5 times, until the budget."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    pos, quat, vel, ang = state
    arr = scene.arrays()
    saved = {k: arr[k].copy() for k in ("pos", "quat", "vel", "angvel")}
    arr["pos"][:] = pos
    arr["quat"][:] = quat
    arr["vel"][:] = vel
    arr["angvel"][:] = ang
    threads = max(0, (os.cpu_count() or 1) - 1)
    order = [w for k in range(8) for w in range(k, scene.n_worlds, 8)]   # strided over the batch
    body_substeps, worlds, el = 0, 0, 0.0
    try:
        for w in order:
            sub = scene.world_subscene(w)
            o = ol.OracleWorld(sub, substeps=substeps, iterations=iterations, threads=threads)
            t0 = time.perf_counter()
            for _ in range(5):
                o.step(dt)
            el += time.perf_counter() - t0
            body_substeps += sub.n_dyn * substeps * 5
            worlds += 1
            del o
            if el >= budget_s:
                break
    finally:
        for k, v in saved.items():
            arr[k][:] = v
    return {"value": body_substeps / el, "unit": "body-substeps/s", "cores": threads + 1, "kind": "port",
            "sample": f"{worlds} of {scene.n_worlds} worlds (stride 8 over the batch), each its own oracle Scene "
                      f"from the device state after warm-up, 5 steps each, one world after another, {el:.1f} s of "
                      f"steps; narrowphase on {threads} workers + caller, solver single-threaded"}


def run_served(args):
    """Poisson request stream (2000 req/s) of seeded scenes with 16..4096 bodies (log-uniform,
    10% with an 8-capsule revolute chain) and horizons U(60, 600) steps, served by one
    pgx_server; the request loop runs in wall-clock time for --seconds. A request is the
    scene of one of 64 pre-built seeds (scene building is host work outside the loop)."""
    from physecs_amd import api, scenes
    rng = np.random.default_rng(0)
    pool = [scenes.served_request(1000 + k) for k in range(64)]
    for sc in pool:
        sc.arrays()
    rate, seconds = 2000.0, args.seconds
    n_req = int(rate * seconds * 1.2) + 16
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    horizons = rng.integers(60, 601, n_req)
    picks = rng.integers(0, len(pool), n_req)
    budget = 16e9
    srv = api.Server(substeps=4, iterations=2, budget_bytes=budget, max_worlds=4096, admit_interval=args.admit_interval)
    srv.set_graph(args.graph)
    # warm the device path with a first batch outside the timed window
    warm = [srv.submit(pool[k], 2) for k in range(8)]
    for _ in range(4):
        srv.step(1 / 60)
    for r in warm:
        srv.take_result(r)
    lat, advanced, outstanding = [], 0, []
    nxt = 0
    t0 = time.perf_counter()
    steps = 0
    completed = 0
    active_sum = 0
    while True:
        now = time.perf_counter() - t0
        if now >= seconds:
            break
        while nxt < n_req and arrivals[nxt] <= now and len(outstanding) < 20000:
            outstanding.append(srv.submit(pool[picks[nxt]], int(horizons[nxt])))
            nxt += 1
        ts = time.perf_counter()
        info = srv.step(1 / 60)
        if info["active"] == 0:   # nothing admitted yet: not a step of any request
            continue
        lat.append(time.perf_counter() - ts)
        advanced += info["bodies"]
        active_sum += info["active"]
        steps += 1
        if steps % 32 == 0:
            keep = []
            for r in outstanding:
                if srv.done(r):
                    srv.take_result(r)
                    completed += 1
                else:
                    keep.append(r)
            outstanding = keep
    el = time.perf_counter() - t0
    srv.close()
    bss = advanced * 4 / el
    line = {
        "metric": "body-substeps/s (served: Poisson 2000 req/s, 16..4096 bodies, 1 GPU)", "value": bss,
        "unit": "body-substeps/s", "n_gpus": 1, "steps": steps, "warmup": 4, "ms_per_step": 1000 * el / max(1, steps),
        "p50_step_ms": 1000 * float(np.percentile(lat, 50)), "p99_step_ms": 1000 * float(np.percentile(lat, 99)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded request scenes)",
        "config": {"workload": "served stream", "seconds": seconds, "requests_arrived": int(nxt),
                   "requests_completed": completed, "mean_active_requests": active_sum / max(1, steps),
                   "hbm_budget_bytes": budget, "admit_interval": args.admit_interval, "graph": args.graph,
                   "substeps": 4,
                   "iterations": 2, "dt": 1 / 60, "parallelism": "single GPU"},
    }
    print(json.dumps(line), flush=True)


def run_raycast(args):
    """Scene::raycastClosest for 1M rays (pgx_world_raycast) into the 32,768-body scene
    after 100 steps: rays from a box around the container toward random points inside it,
    half of them normalised, maxDistance 200. value = rays per second of device time (the
    LBVH rebuild from the current bounds plus the ray kernel, HIP events); the wall time of
    the whole call, host arrays and PCIe included, is reported beside it."""
    from physecs_amd import api, scenes
    sc = scenes.mixed_lattice(0, side=32)
    w = api.World(sc, substeps=4, iterations=2)
    for k in range(100):
        w.step(1 / 60)
    w.sync()
    rng = np.random.default_rng(0)
    n = 1 << 20
    o = rng.uniform((-30, 0, -30), (30, 60, 30), (n, 3)).astype(np.float32)
    tgt = rng.uniform((-25, 0, -25), (25, 50, 25), (n, 3))
    d = tgt - o
    half = rng.random(n) < 0.5
    d[half] /= np.linalg.norm(d[half], axis=1, keepdims=True)
    d = d.astype(np.float32)
    maxd = np.full(n, 200.0, np.float32)
    w.raycast(o[:1024], d[:1024], maxd[:1024])   # warm-up
    dev, wall, hits = [], [], 0
    for k in range(args.steps):
        t0 = time.perf_counter()
        col, t, us = w.raycast(o, d, maxd)
        wall.append(time.perf_counter() - t0)
        dev.append(us * 1e-6)
        hits = int((col >= 0).sum())
    value = n / float(np.median(dev))
    line = {"metric": "rays/s (Scene::raycastClosest, 32k bodies, 1 GPU)", "value": value, "unit": "rays/s",
            "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": 1000 * float(np.median(dev)),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded scene and rays)",
            "config": {"workload": "1,048,576 rays into 32,768 mixed bodies after 100 steps", "rays": n,
                       "hits": hits, "device_ms_median": 1000 * float(np.median(dev)),
                       "wall_ms_median_with_pcie": 1000 * float(np.median(wall)), "parallelism": "single GPU"},
            "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def run_overlap(args):
    """Scene::overlap for 1M query shapes (pgx_world_overlap) in the 32,768-body scene after
    100 steps: spheres, capsules and boxes of 0.2-1.5 m at random poses inside the
    container. value = queries per second of device time (the LBVH rebuild from the current
    bounds, the query kernel and the sort of the hits, HIP events); the wall time of the
    whole call, host arrays and PCIe included, is reported beside it."""
    from physecs_amd import api, scenes
    sc = scenes.mixed_lattice(0, side=32)
    w = api.World(sc, substeps=4, iterations=2)
    for k in range(100):
        w.step(1 / 60)
    w.sync()
    rng = np.random.default_rng(0)
    n = 1 << 20
    t = rng.integers(0, 3, n).astype(np.int32)
    prm = np.zeros((n, 4), np.float32)
    prm[:, :3] = rng.uniform(0.2, 1.5, (n, 3))
    pos = rng.uniform((-25, 0, -25), (25, 50, 25), (n, 3)).astype(np.float32)
    q = rng.normal(size=(n, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    hits, _ = w.overlap(t[:1024], prm[:1024], pos[:1024], q[:1024])   # warm-up
    cap = 16 * n
    dev, wall, total = [], [], 0
    for k in range(args.steps):
        t0 = time.perf_counter()
        hits, us = w.overlap(t, prm, pos, q, capacity=cap)
        wall.append(time.perf_counter() - t0)
        dev.append(us * 1e-6)
        total = int(sum(len(h) for h in hits))
    value = n / float(np.median(dev))
    line = {"metric": "queries/s (Scene::overlap, 32k bodies, 1 GPU)", "value": value, "unit": "queries/s",
            "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": 1000 * float(np.median(dev)),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded scene and query shapes)",
            "config": {"workload": "1,048,576 sphere/capsule/box overlap queries in 32,768 mixed bodies after 100 "
                                   "steps", "queries": n, "hits": total,
                       "device_ms_median": 1000 * float(np.median(dev)),
                       "wall_ms_median_with_pcie": 1000 * float(np.median(wall)), "parallelism": "single GPU"},
            "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def run_character(args):
    """CharacterController::move for 65,536 characters (pgx_world_character_move) on the
    convex-terrain scene after 60 steps: each stands 5 cm into the terrain under it, yaw
    random, and walks 0.1 m; 2 MTD rounds. value = characters
    per second of wall time (the device MTD rounds plus the host pushes); the device time
    of one round's MTD query of the same capsules is reported beside it."""
    from physecs_amd import api, scenes
    print("building the convex terrain scene", file=sys.stderr, flush=True)
    sc = scenes.convex_terrain(0)
    w = api.World(sc, substeps=4, iterations=2)
    for k in range(60):
        w.step(1 / 60)
    w.sync()
    rng = np.random.default_rng(0)
    n = 1 << 16
    # the terrain's surface under each character: the heightfield's two triangles per cell
    _, _, hf = scenes.terrain_mesh(0, quads=256, amplitude=8.0)
    xz = rng.uniform(-120, 120, (n, 2))
    i0, j0 = np.floor(xz[:, 0] + 128).astype(int), np.floor(xz[:, 1] + 128).astype(int)
    u, v = xz[:, 0] + 128 - i0, xz[:, 1] + 128 - j0
    h00, h01, h10, h11 = hf[i0, j0], hf[i0, j0 + 1], hf[i0 + 1, j0], hf[i0 + 1, j0 + 1]
    ground = np.where(u + v <= 1, h00 + (h10 - h00) * u + (h01 - h00) * v,
                      h11 + (h01 - h11) * (1 - u) + (h10 - h11) * (1 - v))
    pos = np.stack([xz[:, 0], ground - 0.05, xz[:, 1]], 1).astype(np.float32)   # 5 cm into the ground
    yaw = rng.uniform(0, 2 * np.pi, n)
    quat = np.zeros((n, 4), np.float32)
    quat[:, 1], quat[:, 3] = np.sin(yaw / 2), np.cos(yaw / 2)
    tr = np.stack([np.cos(yaw), np.zeros(n), np.sin(yaw)], 1).astype(np.float32) * 0.1
    h, r = 1.6, 0.35
    w.character_move(pos[:256], quat[:256], tr[:256], h, r)   # warm-up
    wall, flags = [], None
    for k in range(args.steps):
        t0 = time.perf_counter()
        _, flags = w.character_move(pos, quat, tr, h, r)
        wall.append(time.perf_counter() - t0)
    centre = (pos + tr + np.array([0, h * 0.5 + r, 0])).astype(np.float32)   # yaw-only: up is +y
    caps = np.tile(np.array([h * 0.5, r, 0, 0], np.float32), (n, 1))
    _, us = w.overlap_mtd(np.full(n, scenes.CAPSULE, np.int32), caps, centre, quat)
    value = n / float(np.median(wall))
    line = {"metric": "characters/s (CharacterController::move, 2 rounds, 1 GPU)", "value": value,
            "unit": "characters/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
            "ms_per_step": 1000 * float(np.median(wall)), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene and characters)",
            "config": {"workload": f"{n} capsule characters on the 131072-triangle terrain among 12288 bodies",
                       "characters": n, "touching_ground": int((flags & 1).astype(bool).sum()),
                       "mtd_round_device_ms": us / 1000.0, "parallelism": "single GPU"},
            "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def run_meshcache(args):
    """World creation for the convex-terrain scene (131,072-triangle terrain, 12,288 bodies,
    3,584 joints) with the TriangleMesh build done at creation and loaded from
    pgx_trimesh_cache bytes. value = creation time with the cache, ms (median)."""
    from physecs_amd import api, scenes
    sc = scenes.convex_terrain(0)
    api.World(sc, substeps=4, iterations=2).sync()   # warm-up: device, library, scene arrays
    t0 = time.perf_counter()
    caches = [api.trimesh_cache(v, t) for v, t in sc.trimesh]
    build_ms = 1000 * (time.perf_counter() - t0)
    times = {}
    for mode in ("build", "cached"):
        sc.tri_bvh_cache = caches if mode == "cached" else None
        ts = []
        for k in range(args.steps):
            t0 = time.perf_counter()
            w = api.World(sc, substeps=4, iterations=2)
            w.sync()
            ts.append(time.perf_counter() - t0)
            del w
        times[mode] = 1000 * float(np.median(ts))
    line = {"metric": "world creation ms (convex terrain, TriangleMesh build cached)", "value": times["cached"],
            "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": times["cached"],
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded scene)",
            "config": {"workload": "create the world of 12288 bodies on a 131072-triangle terrain",
                       "creation_ms_building": times["build"], "creation_ms_cached": times["cached"],
                       "cache_build_ms": build_ms, "cache_bytes": sum(len(c) for c in caches),
                       "parallelism": "single GPU"},
            "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def run_snapshot(args):
    """pgx_world_save_state + pgx_world_load_state of the 32,768-body scene after 100 steps
    (poses, velocities, bounds, contact cache). value = bytes per second through a save and a
    load (device to pageable host memory and back, PCIe included)."""
    from physecs_amd import api, scenes
    sc = scenes.mixed_lattice(0, side=32)
    w = api.World(sc, substeps=4, iterations=2)
    for k in range(100):
        w.step(1 / 60)
    w.sync()
    blob = w.save_state()
    times = []
    for k in range(args.steps):
        t0 = time.perf_counter()
        blob = w.save_state()
        w.load_state(blob)
        times.append(time.perf_counter() - t0)
    value = 2 * len(blob) / float(np.median(times))
    line = {"metric": "bytes/s (world state save + load, 32k bodies, 1 GPU)", "value": value, "unit": "B/s",
            "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": 1000 * float(np.median(times)),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bytes",
            "data": "synthetic (seeded scene)",
            "config": {"workload": "save and load the state of 32768 mixed bodies after 100 steps",
                       "state_bytes": len(blob), "parallelism": "single GPU"},
            "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's run after "
                    "the warm-up, 590 for the 32k scene)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="plain32k")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--blocks", type=int, default=0, help="cooperative grid size (0 = one block per CU)")
    ap.add_argument("--exec", default="colors", choices=["flow", "xcd", "colors", "cooperative", "worlds", "pass", "xpass"],
                    help="how the contact passes launch")
    ap.add_argument("--graph", action="store_true", help="launch each step's substep loop as one HIP graph")
    ap.add_argument("--timing-period", type=int, default=16,
                    help="per-launch HIP events on every n-th timed step (the others run without events)")
    ap.add_argument("--tune", default=None, help="launch shapes: 'auto' (pgx_world_autotune after the warm-up), "
                    "'default', or a JSON object of pgx_tuning fields; by default 'auto' for the batched worlds "
                    "and 'default' otherwise")
    ap.add_argument("--seconds", type=float, default=60.0, help="served: length of the request stream")
    ap.add_argument("--admit-interval", type=int, default=8, help="served: steps between admission boundaries")
    ap.add_argument("--copy-calibration", action="store_true", help="run only the 1 GiB copy used to scale PMC counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    try:
        dispatch(args)
    finally:
        if "physecs_amd.api" in sys.modules:
            sys.modules["physecs_amd.api"]._close_open_handles()   # explicit teardown before shutdown


def dispatch(args):
    if args.steps is None:   # the workload's whole run after the warm-up
        args.steps = RUN_STEPS.get(args.config, 600) - args.warmup
    if args.config == "served":
        run_served(args)
        return
    if args.config == "raycast":
        if args.steps == 590:
            args.steps = 10
        run_raycast(args)
        return
    if args.config == "overlap":
        if args.steps == 590:
            args.steps = 10
        run_overlap(args)
        return
    if args.config in ("character", "snapshot", "meshcache"):
        if args.steps == 590:
            args.steps = 10 if args.config != "meshcache" else 5
        {"character": run_character, "snapshot": run_snapshot, "meshcache": run_meshcache}[args.config](args)
        return
    run_world(args)


def run_world(args):
    from physecs_amd import api
    print(f"building the {args.config} scene", file=sys.stderr, flush=True)
    scene, cfg = make_scene(args.config)
    print(f"scene built: {scene.n_dyn} dynamic bodies", file=sys.stderr, flush=True)
    sub, it, dt = cfg["substeps"], cfg["iterations"], cfg["dt"]
    total = RUN_STEPS.get(args.config, args.warmup + args.steps)
    settle = max(0, total - args.warmup - args.steps)
    first = settle + args.warmup + 1
    window = f"steps {first}-{first + args.steps - 1} of {max(total, settle + args.warmup + args.steps)}"
    w = api.World(scene, substeps=sub, iterations=it)
    w.set_exec(args.exec, blocks=args.blocks)
    w.set_graph(args.graph)
    for k in range(settle):
        w.step(dt)
        if (k + 1) % 100 == 0:
            w.sync()
            print(f"settle step {k + 1}/{settle}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        w.step(dt)
    w.sync()
    if args.tune is None:
        # the library's autotuner where the default launch shapes are not the fastest: the
        # batch's worlds and the 210-box pyramid (one small world: per-world passes)
        args.tune = "auto" if args.config in ("batched", "pyramid") else "default"
    if args.tune == "auto":
        chosen = w.autotune(dt, trial_steps=4)   # restores the world's state
        print(f"autotune chose {chosen}, exec {w.exec_mode()}", file=sys.stderr, flush=True)
    elif args.tune != "default":
        w.set_tuning(**json.loads(args.tune))
    warm_state = w.state()
    # the kernels' durations: HIP events around every contact-solve launch of the first
    # substep, on every timing_period-th timed step; the steps between run without events
    period = max(1, args.timing_period)
    w.enable_timing(not args.graph, period=period)
    w.reset_stats()
    lat = []
    w.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        try:
            w.step(dt)
        except api.PgxError:
            print(f"timed step {k + 1} (step {args.warmup + k + 1} of the scene) failed", file=sys.stderr, flush=True)
            raise
        w.sync()
        lat.append(time.perf_counter() - ts)
        if (k + 1) % 100 == 0:
            print(f"timed step {k + 1}/{args.steps}", file=sys.stderr, flush=True)
    w.sync()
    el = time.perf_counter() - t0
    if args.graph:
        # graph launches carry no per-launch events: the kernel durations come from the
        # steps right after the timed ones, launched directly with the same colours
        w.enable_timing(True)
        for _ in range(min(args.steps, 50)):
            w.step(dt)
        w.sync()
    st = w.stats()
    counts = w.counts()
    tuning, exec_mode, phase = w.tuning(), w.exec_mode(), w.timing()
    w.close()
    bss = scene.n_dyn * sub * args.steps / el
    # the dominant contact-solve kernel: the per-colour k_solve, or for a scene whose colours
    # all hold few manifolds (the pyramid) the runs of colours in one workgroup
    kernel = {"flow": "k_solve_flow", "xcd": "k_solve_flow", "colors": "k_solve", "cooperative": "k_step_persistent",
              "worlds": "k_world_pass", "pass": "k_solve_pass", "xpass": "k_solve_xpass"}[exec_mode]
    s_us, s_bytes, s_n = st["solve_us"], st["solve_bytes"], st["solve_launches"]
    if st["batch_us"] > s_us:
        kernel, s_us, s_bytes, s_n = "k_solve_batch", st["batch_us"], st["batch_bytes"], st["batch_launches"]
    achieved = s_bytes / (s_us * 1e-6) / 1e9 if s_us > 0 else 0.0
    # step latency over the steps without events (those with events wait for their reads)
    timed = [k % period == 0 and not args.graph for k in range(args.steps)]
    bare = [t for t, ev in zip(lat, timed) if not ev] or lat
    line = {
        "metric": "body-substeps/s (32k bodies, 1 GPU)" if args.config == "plain32k" else "body-substeps/s",
        "value": bss, "unit": "body-substeps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000 * el / args.steps, "p50_step_ms": 1000 * float(np.percentile(bare, 50)),
        "p99_step_ms": 1000 * float(np.percentile(bare, 99)), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generator)",
        "config": dict(cfg, parallelism="single GPU", graph=args.graph, tuning=tuning, tuned=args.tune,
                       exec=exec_mode, settle_steps=settle, window=window,
                       steps_with_events=int(sum(timed)), timing_period=period,
                       p99_step_ms_all=1000 * float(np.percentile(lat, 99)),
                       points_per_body=counts["points"] / scene.n_dyn,
                       last_step_counts=counts,
                       phase_us_last_step=phase),
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.config, exec_mode, window) if kernel != "k_solve_batch" else None,
                     "avg_launch_us": s_us / max(1, s_n),
                     "bytes_per_launch": s_bytes / max(1, s_n),
                     "launches_timed": s_n,
                     "k_solve_batch": {"avg_launch_us": st["batch_us"] / max(1, st["batch_launches"]),
                                       "bytes_per_launch": st["batch_bytes"] / max(1, st["batch_launches"]),
                                       "launches_timed": st["batch_launches"]}},
    }
    if not args.no_cpu:
        base = cpu_baseline_worlds if scene.n_worlds > 1 else cpu_baseline
        line["cpu_baseline"] = base(scene, warm_state, sub, it, dt)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
