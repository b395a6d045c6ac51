"""CPU-side checks of the C ABI boundary and host logic (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "physecs_gpu.h")).read()
    return sorted(set(re.findall(r"\b(pgx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from physecs_amd import api
    L = api.lib()
    syms = header_symbols()
    assert len(syms) >= 10
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(api.EXPORTS) == syms


def test_no_device_reports_error_not_crash():
    from physecs_amd import api, scenes
    if api.device_count() > 0:
        pytest.skip("a GPU is present")
    sc = scenes.box_pyramid(0, rows=2)
    with pytest.raises(api.PgxError):
        api.World(sc)


def test_invalid_descriptor_rejected_on_host():
    from physecs_amd import api, scenes
    sc = scenes.box_pyramid(0, rows=2)
    arr = sc.arrays()
    arr["col_type"][0] = 7
    desc = sc.struct(api.PgxSceneDesc)
    desc.actor_world = ctypes.cast(None, api.I32)
    h = ctypes.c_void_p()
    rc = api.lib().pgx_world_create(None, ctypes.byref(desc), ctypes.byref(h))
    assert rc == -4
    assert b"geometry type 7" in api.lib().pgx_last_error()


def test_config_block_size_checked_on_host():
    """pgx_config.block_size: 0, 64, 128 or 256 threads per contact-solve workgroup; any
    other value is rejected before anything touches a device."""
    from physecs_amd import api, scenes
    sc = scenes.box_pyramid(0, rows=2)
    desc = sc.struct(api.PgxSceneDesc)
    desc.actor_world = ctypes.cast(None, api.I32)
    h = ctypes.c_void_p()
    cfg = api.PgxConfig(4, 2, 9.81, 0, 0, 96)
    assert api.lib().pgx_world_create(ctypes.byref(cfg), ctypes.byref(desc), ctypes.byref(h)) == -1
    assert b"block_size" in api.lib().pgx_last_error()


def test_scene_generators_shapes():
    from physecs_amd import scenes
    p = scenes.box_pyramid(0)
    assert p.n_dyn == 210 and p.n_colliders == 211
    m = scenes.mixed_lattice(0, side=4)
    assert m.n_dyn == 64 and m.n_colliders == 69
    kinds = np.bincount(m.arrays()["col_type"][5:], minlength=3)
    assert abs(kinds[0] - kinds[1]) <= 1 and abs(kinds[1] - kinds[2]) <= 1
    b = scenes.batched_worlds(0, n_worlds=16)
    assert b.n_worlds == 16
    assert b.n_colliders == b.n_actors


def test_mass_props_match_reference_formulas():
    import oracle_lib as ol
    from physecs_amd import scenes
    if not ol.ref_available():
        pytest.skip("oracle/_ref not built here")
    ref = ol.ref()
    for t, p in ((0, [0.4]), (1, [0.5, 0.2]), (2, [0.3, 0.5, 0.2])):
        pr = np.zeros(4, np.float32)
        pr[:len(p)] = p
        mass = 1000.0 * scenes.true_volume(t, pr)
        im, com, invI = scenes.mass_props(t, pr, mass)
        out = np.zeros(13, np.float32)
        z = np.zeros(3, np.float32)
        q = np.array([0, 0, 0, 1], np.float32)
        ref.ref_mass_props(t, ol.fptr(pr), ol.fptr(z), ol.fptr(q), np.float32(mass), ol.fptr(out))
        assert abs(out[0] - im) <= 1e-6 * abs(im)
        np.testing.assert_allclose(out[4:], invI, rtol=2e-5)


def test_joint_colours_follow_add_joint():
    """Scene::addJoint colouring (src/Physecs.cpp:690-710): lowest colour free on both
    entities, static bodies included, colour 8 once eight are taken."""
    import oracle_lib as ol
    from physecs_amd import scenes
    sc = scenes.Scene()
    ground = sc.add_static(scenes.BOX, [5, 0.5, 5], [0, -0.5, 0])
    bodies = [sc.add_body(scenes.SPHERE, [0.2], [i, 1, 0]) for i in range(12)]
    for b in bodies[:10]:
        sc.add_joint(scenes.SPHERICAL, ground, b, (0, 0.5, 0), (0, 0, 0))
    sc.add_joint(scenes.SPHERICAL, bodies[0], bodies[1], (0, 0, 0), (0, 0, 0))
    o = ol.OracleWorld(sc)
    assert o.joint_colors() == [0, 1, 2, 3, 4, 5, 6, 7, 8, 8, 2]


def test_oracle_joint_pendulum_closed_form():
    """A point mass on a spherical joint to a static anchor keeps its distance to the anchor
    (within the solver's position error) and swings with the closed-form small-angle period."""
    import oracle_lib as ol
    from physecs_amd import scenes
    sc = scenes.Scene()
    anchor = sc.add_static(scenes.SPHERE, [0.05], [0, 10, 0])
    L = 2.0
    theta0 = 0.1
    b = sc.add_body(scenes.SPHERE, [0.05], [L * np.sin(theta0), 10 - L * np.cos(theta0), 0])
    sc.col_enable[0] = 0
    sc.col_enable[1] = 0
    sc._frozen = None
    sc.add_joint(scenes.SPHERICAL, anchor, b, (0, 0, 0), (-L * np.sin(theta0), L * np.cos(theta0), 0))
    o = ol.OracleWorld(sc, substeps=8, iterations=2)
    xs = []
    for _ in range(600):
        o.step(1 / 120)
        p = o.state()[0][1]
        xs.append(p[0])
        assert abs(np.linalg.norm(p - np.array([0, 10, 0])) - L) < 5e-3
    xs = np.array(xs)
    crossings = np.where(np.diff(np.sign(xs)) != 0)[0]
    period = 2 * np.mean(np.diff(crossings)) / 120
    assert abs(period - 2 * np.pi * np.sqrt(L / 9.81)) < 0.05


def test_bulk_bodies_match_single_appends():
    """Scene.add_bodies_bulk writes the same descriptor fields as add_body, one by one."""
    from physecs_amd import scenes
    rng = np.random.default_rng(7)
    n = 60
    kinds = rng.integers(0, 3, n)
    prm = np.zeros((n, 4))
    for i, k in enumerate(kinds):
        prm[i, :3] = [rng.uniform(0.2, 0.5), rng.uniform(0.15, 0.3), rng.uniform(0.2, 0.6)] if k == 2 else \
            ([rng.uniform(0.2, 0.6), rng.uniform(0.15, 0.3), 0] if k == 1 else [rng.uniform(0.2, 0.5), 0, 0])
    pos = rng.normal(size=(n, 3))
    q = scenes.random_quats(rng, n)
    v = rng.normal(size=(n, 3))
    a, b = scenes.Scene(), scenes.Scene()
    a.add_static(scenes.BOX, [5, 0.5, 5], [0, -0.5, 0])
    b.add_static(scenes.BOX, [5, 0.5, 5], [0, -0.5, 0])
    for i in range(n):
        a.add_body(int(kinds[i]), list(prm[i, :3]), pos[i], q[i], vel=v[i])
    b.add_bodies_bulk(kinds, prm, pos, q, v, np.zeros(n, int))
    A, B = a.arrays(), b.arrays()
    for k in A:
        np.testing.assert_allclose(A[k], B[k], rtol=2e-7, err_msg=k)


def test_device_introsort_matches_std_sort(tmp_path):
    """physecs_amd/csrc/introsort.h repeats std::sort's exact sequence of compares and moves:
    on arrays full of equal keys (where an unstable sort's output order is implementation
    detail) it leaves the same permutation as this C++ library's std::sort, at every size
    from the insertion-sort cutoff through the heap-sort fallback."""
    import subprocess
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>
#include "introsort.h"
struct Arr { std::vector<int>& v; const std::vector<float>& key;
  int get(int i) { return v[i]; } void set(int i, int x) { v[i] = x; }
  bool less(int a, int b) { return key[a] < key[b]; } };
struct E { float k; int id; bool operator<(const E& o) const { return k < o.k; } };
int main() {
  std::mt19937 rng(7);
  for (int trial = 0; trial < 4000; ++trial) {
    int n = 1 + rng() % 300;
    int distinct = 1 + rng() % (trial % 3 == 0 ? 4 : 40);
    std::vector<float> key(n);
    for (auto& k : key) k = (float)(rng() % distinct);
    if (trial % 7 == 0) std::sort(key.begin(), key.end());          // presorted runs
    if (trial % 11 == 0) std::reverse(key.begin(), key.end());
    std::vector<E> e(n);
    for (int i = 0; i < n; ++i) e[i] = {key[i], i};
    std::sort(e.begin(), e.end());
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    Arr a{idx, key};
    pgx::introsort(a, n);
    for (int i = 0; i < n; ++i) if (idx[i] != e[i].id) { printf("mismatch trial %d n %d at %d\n", trial, n, i); return 1; }
  }
  printf("ok\n");
  return 0;
}
''')
    exe = tmp_path / "t"
    inc = os.path.join(ROOT, "physecs_amd", "csrc")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", inc, str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror of a header struct has the C compiler's size and field offsets."""
    import subprocess
    from physecs_amd import api
    mirrors = {"pgx_config": api.PgxConfig, "pgx_scene_desc": api.PgxSceneDesc, "pgx_joint_desc": api.PgxJointDesc,
               "pgx_counts": api.PgxCounts, "pgx_stats": api.PgxStats, "pgx_server_config": api.PgxServerConfig,
               "pgx_server_step_info": api.PgxServerStepInfo, "pgx_tuning": api.PgxTuning,
               "pgx_character": api.PgxCharacter}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "physecs_gpu.h"', "int main(void) {"]
    for cname, cls in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(line.split()[:2]): int(line.split()[2]) for line in out if line}
    for cname, cls in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


@pytest.mark.parametrize("lang,cc", [("c", "gcc"), ("c++", "g++")])
def test_header_is_self_contained(tmp_path, lang, cc):
    """include/physecs_gpu.h compiles as the first and only include, as C99 and C++17, with
    warnings as errors (a caller's binding includes nothing before it)."""
    import subprocess
    std = "-std=c99" if lang == "c" else "-std=c++17"
    src = tmp_path / ("only." + ("c" if lang == "c" else "cpp"))
    src.write_text('#include "physecs_gpu.h"\nint main(void) { return PGX_OK; }\n')
    subprocess.run([cc, std, "-Wall", "-Wextra", "-pedantic", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-c", str(src), "-o", str(tmp_path / "only.o")], check=True)
