"""Pins of the oracle's Constraint1DW rows, the SIMD joint rows of colours 0-7
(src/Constraint1DW.cpp:5-233). That file does not compile here (include/Physecs/SIMD.h uses
MSVC-only __m128 members), so each place where it departs from the scalar Constraint1D
(src/Constraint1D.cpp, which does compile, oracle/_ref) is pinned by a hand-derived
expectation on inputs where every float operation is exact:

- the warm-start impulse is written back with the lane's velocities (:137-161, :213-228);
- LIMITED rows clamp the accumulated lambda to [min*h, max*h] (:196-200), preSolve to
  [min, max] (:66-68);
- the warm-start masks are strict: |c| < 1e-4f and |lambda| < 10000 (:140-142);
- the bias factor is one rounding of 0.2 / h, taken in double (:175);
- a lane whose inverse effective mass is 0 keeps its velocities (:164, :213-215).

Where none of these applies, the SIMD row equals the compiled scalar row. Each pin is also
run with the matching mistake planted in the oracle (tests/oracle_lib.mutation) and must
fail there."""
import numpy as np
import pytest

import oracle_lib as ol

NONE, SOFT, ANGULAR, LIMITED = 0, 1, 2, 4
FMAX = float(np.finfo(np.float32).max)
needs_ref = pytest.mark.skipif(not ol.ref_available(), reason="oracle/_ref not built here")


def row(lin=(0, 0, 0), ang0=(0, 0, 0), ang1=(0, 0, 0), tv=0.0, c=0.0, mn=-FMAX, mx=FMAX, freq=0.0, damp=0.0):
    return np.array(list(lin) + list(ang0) + list(ang1) + [tv, c, mn, mx, freq, damp], np.float32)


def masses(im0=1.0, I0=None, im1=0.0, I1=None):
    m = np.zeros(20, np.float32)
    m[0] = im0
    m[1:10] = (np.eye(3) if I0 is None else np.asarray(I0)).ravel()
    m[10] = im1
    m[11:20] = (np.zeros((3, 3)) if I1 is None else np.asarray(I1)).ravel()
    return m


def run(which, flags, inp, m, vel, b1=-1, h=0.25, iters=1, lambda0=0.0, warm=1):
    """One row on bodies (0, b1): returns (v0, w0, v1, w1, pv (14), invEff, lambda)."""
    vel = np.zeros(12, np.float32) if vel is None else np.asarray(vel, np.float32)
    v, pv, out = ol.row_solve_ex(which, flags, 0, b1, inp, m, vel, np.zeros(14, np.float32), np.float32(h), iters,
                                 lambda0, warm)
    return v[0:3], v[3:6], v[6:9], v[9:12], pv, out[0], out[1]


def planted(m, check):
    """check() passes on the restatement and fails with mistake m planted."""
    check()
    with ol.mutation(m):
        with pytest.raises(AssertionError):
            check()


# ---------------------------------------------------------------- warm start write-back

def _warm_one_body():
    # v0 = 1 along the row; lambda0 = 2, c = 0: the warm half (1) is applied, then the solve
    # takes the relative velocity (-2) out. SIMD writes back 1 + 1 - 2 = 0; the scalar row
    # applies the warm half to local copies only and writes 1 - 2 = -1.
    inp = row(lin=(1, 0, 0))
    v0, _, _, _, _, ie, lam = run("simd", NONE, inp, masses(), [1, 0, 0, 0, 0, 0] + [0] * 6, lambda0=2.0)
    assert ie == 1.0
    assert lam == -1.0
    assert np.array_equal(v0, [0, 0, 0]), v0


def _warm_two_bodies():
    # invEff = |ang0|^2 + |ang1|^2 + |lin|^2 (1 + 1) = 4; lambda0 = 8 -> warm half 4 on every
    # component, relative velocity -16, lambda = -16 / 4 = -4 cancels it: SIMD ends at rest.
    inp = row(lin=(1, 0, 0), ang0=(0, 1, 0), ang1=(0, 0, 1))
    m = masses(1.0, np.eye(3), 1.0, np.eye(3))
    v0, w0, v1, w1, _, ie, lam = run("simd", NONE, inp, m, None, b1=1, lambda0=8.0)
    assert ie == 4.0 and lam == 0.0
    for x in (v0, w0, v1, w1):
        assert np.array_equal(x, [0, 0, 0]), (v0, w0, v1, w1)


def test_simd_warm_impulse_is_written_back():
    planted(ol.MUT_SIMD_WARM_LOCAL, _warm_one_body)
    planted(ol.MUT_SIMD_WARM_LOCAL, _warm_two_bodies)


def test_scalar_warm_impulse_stays_local():
    """The oracle's scalar row (colour 8) on the same inputs: the warm half biases the
    relative velocity but is not written (src/Constraint1D.cpp:72-87)."""
    v0, _, _, _, _, _, lam = run("scalar", NONE, row(lin=(1, 0, 0)), masses(), [1, 0, 0, 0, 0, 0] + [0] * 6,
                                 lambda0=2.0)
    assert lam == -1.0 and np.array_equal(v0, [-1, 0, 0])
    inp = row(lin=(1, 0, 0), ang0=(0, 1, 0), ang1=(0, 0, 1))
    v0, w0, v1, w1, _, _, lam = run("scalar", NONE, inp, masses(1.0, np.eye(3), 1.0, np.eye(3)), None, b1=1,
                                    lambda0=8.0)
    assert lam == 0.0
    assert np.array_equal(v0, [-4, 0, 0]) and np.array_equal(v1, [4, 0, 0])
    assert np.array_equal(w0, [0, -4, 0]) and np.array_equal(w1, [0, 0, 4])


@needs_ref
def test_scalar_warm_matches_compiled_reference():
    for inp, m, b1, lam0 in ((row(lin=(1, 0, 0)), masses(), -1, 2.0),
                             (row(lin=(1, 0, 0), ang0=(0, 1, 0), ang1=(0, 0, 1)),
                              masses(1.0, np.eye(3), 1.0, np.eye(3)), 1, 8.0)):
        a = run("ref", NONE, inp, m, [1, 0, 0, 0, 0, 0] + [0] * 6, b1=b1, lambda0=lam0)
        b = run("scalar", NONE, inp, m, [1, 0, 0, 0, 0, 0] + [0] * 6, b1=b1, lambda0=lam0)
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), np.asarray(y))


# ---------------------------------------------------------------- LIMITED clamp

def _limited_scaled():
    h = 0.25
    # rv = 0.5 -> lambda 0.5; SIMD clamps the total to max * h = 0.25
    v0, _, _, _, _, _, lam = run("simd", LIMITED, row(lin=(1, 0, 0), mn=-1, mx=1), masses(),
                                 [-0.5, 0, 0, 0, 0, 0] + [0] * 6, h=h)
    assert lam == 0.25 and np.array_equal(v0, [-0.25, 0, 0]), (lam, v0)
    # the other side: rv = -0.5, clamp to min * h = -0.25
    v0, _, _, _, _, _, lam = run("simd", LIMITED, row(lin=(1, 0, 0), mn=-1, mx=1), masses(),
                                 [0.5, 0, 0, 0, 0, 0] + [0] * 6, h=h)
    assert lam == -0.25 and np.array_equal(v0, [0.25, 0, 0]), (lam, v0)
    # angular limited row, the same numbers on w0
    _, w0, _, _, _, _, lam = run("simd", ANGULAR | LIMITED, row(ang0=(1, 0, 0), ang1=(1, 0, 0), mn=-1, mx=1),
                                 masses(), [0, 0, 0, -0.5, 0, 0] + [0] * 6, h=h)
    assert lam == 0.25 and np.array_equal(w0, [-0.25, 0, 0]), (lam, w0)


def test_simd_limited_clamp_is_scaled_by_h():
    planted(ol.MUT_SIMD_LIMIT_UNSCALED, _limited_scaled)


def test_scalar_limited_clamp_is_not_scaled():
    v0, _, _, _, _, _, lam = run("scalar", LIMITED, row(lin=(1, 0, 0), mn=-1, mx=1), masses(),
                                 [-0.5, 0, 0, 0, 0, 0] + [0] * 6)
    assert lam == 0.5 and np.array_equal(v0, [0, 0, 0])


@pytest.mark.parametrize("which", ["simd", "scalar"])
def test_limited_presolve_clamps_to_min_max(which):
    """preSolve's NGS lambda = c / invEff = 2 is clamped to [min, max] = [-1, 1] in both forms
    (Constraint1DW.cpp:66-68 with _mm_min_ps/_mm_max_ps; Constraint1D.cpp:33-35)."""
    _, _, _, _, pv, ie, lam = run(which, LIMITED, row(lin=(1, 0, 0), c=2.0, mn=-1, mx=1), masses(), None, iters=0)
    assert ie == 1.0 and lam == 0.0
    assert np.array_equal(pv[0:3], [1, 0, 0]) and pv[6] == 1, pv


# ---------------------------------------------------------------- strict warm-start masks

def _masks_strict():
    c = np.float32(1e-4)
    below = np.nextafter(c, np.float32(0))
    inp = lambda cc: row(lin=(1, 0, 0), c=cc)
    # c == 1e-4f: no warm start, lambda = 2 + bias * c
    lam = run("simd", NONE, inp(c), masses(), None, lambda0=2.0)[-1]
    assert abs(lam - 2.0) < 1e-3, lam
    # the float below: warm start halves lambda and the solve takes it out
    lam = run("simd", NONE, inp(below), masses(), None, lambda0=2.0)[-1]
    assert abs(lam) < 1e-3, lam
    # |lambda| == 10000: no warm start; the float below: warm start
    lam = run("simd", NONE, inp(0.0), masses(), None, lambda0=10000.0)[-1]
    assert lam == 10000.0, lam
    lam = run("simd", NONE, inp(0.0), masses(), None, lambda0=float(np.nextafter(np.float32(10000), 0)))[-1]
    assert lam == 0.0, lam


def test_simd_warm_masks_are_strict():
    planted(ol.MUT_SIMD_MASK_NONSTRICT, _masks_strict)


@pytest.mark.parametrize("which", ["scalar"] + (["ref"] if ol.ref_available() else []))
def test_scalar_warm_masks_are_inclusive(which):
    """Constraint1D warm-starts unless |c| > 1e-4 (the double) or |lambda| > 10000, so both
    boundary values warm-start there."""
    lam = run(which, NONE, row(lin=(1, 0, 0), c=np.float32(1e-4)), masses(), None, lambda0=2.0)[-1]
    assert abs(lam) < 1e-3, lam
    lam = run(which, NONE, row(lin=(1, 0, 0)), masses(), None, lambda0=10000.0)[-1]
    assert lam == 0.0, lam


# ---------------------------------------------------------------- bias factor

def _bias_one_rounding():
    h = np.float32(1 / 240)
    b = np.float32(0.2 / np.float64(h))
    rng = np.random.default_rng(7)
    cs = rng.uniform(1e-3, 1e-1, 400).astype(np.float32)
    differ = 0
    for c in cs:
        # rv = 0, invEff = 1: lambda = bias * c, one float product
        v0, _, _, _, _, _, lam = run("simd", NONE, row(lin=(1, 0, 0), c=c), masses(), None, h=h, warm=0)
        want = np.float32(b * c)
        assert lam.view(np.uint32) == want.view(np.uint32), (c, lam, want)
        assert v0[0].view(np.uint32) == want.view(np.uint32)
        differ += np.float32(np.float32(np.float32(0.2) * c) / h) != want
    assert differ > 20, differ     # the two roundings part on these inputs, so the check can tell


def test_simd_bias_is_one_rounding_of_0_2_over_h():
    planted(ol.MUT_SIMD_BIAS_FLOAT, _bias_one_rounding)


# ---------------------------------------------------------------- zero inverse effective mass

def _zero_lane():
    # invI0 = diag(-1, 1, 1) and angular0 = x cancel the linear part: invEff = -1 + 1 = 0, while
    # linear0t = x and angular0t = -x are not 0, so a written warm impulse would show.
    inp = row(lin=(1, 0, 0), ang0=(1, 0, 0))
    m = masses(1.0, np.diag([-1.0, 1.0, 1.0]))
    v0, w0, _, _, _, ie, lam = run("simd", NONE, inp, m, [0.5, 0, 0, 0.25, 0, 0] + [0] * 6, lambda0=2.0)
    assert ie == 0.0
    assert lam == 1.0                       # the warm half is kept in lambda
    assert np.array_equal(v0, [0.5, 0, 0]) and np.array_equal(w0, [0.25, 0, 0]), (v0, w0)


def test_simd_zero_inverse_mass_lane_keeps_velocities():
    planted(ol.MUT_SIMD_ZERO_WRITEBACK, _zero_lane)


@pytest.mark.parametrize("which", ["scalar"] + (["ref"] if ol.ref_available() else []))
def test_scalar_zero_inverse_mass_returns_first(which):
    """Constraint1D::solve returns before its warm start when invEff is 0 (:55)."""
    inp = row(lin=(1, 0, 0), ang0=(1, 0, 0))
    m = masses(1.0, np.diag([-1.0, 1.0, 1.0]))
    v0, w0, _, _, _, ie, lam = run(which, NONE, inp, m, [0.5, 0, 0, 0.25, 0, 0] + [0] * 6, lambda0=2.0)
    assert ie == 0.0 and lam == 2.0
    assert np.array_equal(v0, [0.5, 0, 0]) and np.array_equal(w0, [0.25, 0, 0])


# ---------------------------------------------------------------- soft rows

@pytest.mark.parametrize("flags", [SOFT, ANGULAR | SOFT])
def test_simd_soft_rows_never_warm_start(flags):
    inp = row(lin=(1, 0, 0), ang0=(0, 1, 0), c=np.float32(1e-5), freq=5.0, damp=0.7)
    a = run("simd", flags, inp, masses(), [0.3, 0, 0, 0, 0.1, 0] + [0] * 6, lambda0=2.0, warm=1)
    b = run("simd", flags, inp, masses(), [0.3, 0, 0, 0, 0.1, 0] + [0] * 6, lambda0=2.0, warm=0)
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x), np.asarray(y))


def _random_rows(rng, n, flags, c_scale):
    for _ in range(n):
        inp = np.zeros(15, np.float32)
        inp[0:9] = rng.normal(size=9)
        inp[9] = rng.normal() * 0.1
        inp[10] = rng.normal() * c_scale
        inp[11], inp[12] = -abs(rng.normal()) * 100, abs(rng.normal()) * 100
        inp[13], inp[14] = rng.uniform(1, 30), rng.uniform(0.1, 1)
        m = np.zeros(20, np.float32)
        for b in range(2):
            m[10 * b] = rng.uniform(0.01, 2)
            A = rng.normal(size=(3, 3))
            m[10 * b + 1: 10 * b + 10] = (A @ A.T * 0.3).astype(np.float32).ravel()
        vel = rng.normal(size=12).astype(np.float32)
        yield inp, m, vel


def _ulps(a, b):
    a = np.asarray(a, np.float32).astype(np.float64)
    b = np.asarray(b, np.float32).astype(np.float64)
    sp = np.spacing(np.maximum(np.abs(a), np.abs(b)).astype(np.float32)).astype(np.float64)
    return np.max(np.abs(a - b) / np.maximum(sp, 1e-45))


@needs_ref
@pytest.mark.parametrize("flags", [SOFT, ANGULAR | SOFT])
def test_simd_soft_rows_agree_with_compiled_scalar(flags):
    """The soft formula is the scalar's with effMass = 1 / invEff multiplied in instead of
    divided (:181-188): the results agree within a few float roundings."""
    rng = np.random.default_rng(31 + flags)
    h = np.float32(1 / 240)
    for inp, m, vel in _random_rows(rng, 2000, flags, 0.01):
        a = run("ref", flags, inp, m, vel, b1=1, h=h, iters=3)
        b = run("simd", flags, inp, m, vel, b1=1, h=h, iters=3)
        for x, y in zip(a[:4], b[:4]):
            assert np.allclose(x, y, rtol=2e-5, atol=2e-6), (x, y)
        assert np.isclose(a[6], b[6], rtol=2e-5, atol=2e-7), (a[6], b[6])


@needs_ref
@pytest.mark.parametrize("flags", [NONE, ANGULAR, SOFT, LIMITED, ANGULAR | SOFT, ANGULAR | LIMITED])
def test_simd_presolve_matches_compiled_scalar(flags):
    """preSolve is the same arithmetic in both forms (:5-115 vs Constraint1D.cpp:5-52): the
    inverse effective mass and the pseudo velocities agree bit for bit."""
    rng = np.random.default_rng(50 + flags)
    for inp, m, vel in _random_rows(rng, 2000, flags, 0.01):
        a = run("ref", flags, inp, m, vel, b1=1, iters=0)
        b = run("simd", flags, inp, m, vel, b1=1, iters=0)
        assert a[5].view(np.uint32) == b[5].view(np.uint32)
        assert np.array_equal(a[4].view(np.uint32), b[4].view(np.uint32))


@needs_ref
@pytest.mark.parametrize("flags", [NONE, ANGULAR])
def test_simd_hard_rows_match_compiled_scalar_where_nothing_differs(flags):
    """Hard rows without warm start, h a power of two (so (float)(0.2 / h) * c rounds as
    0.2f * c / h does) and invEff a power of two (so multiplying by 1 / invEff rounds as
    dividing by it does): bit-exact with the compiled Constraint1D. With other invEff the
    one product per solve may part by an ulp, and the velocities follow it within 1e-6."""
    rng = np.random.default_rng(70 + flags)
    exact = 0
    for inp, m, vel in _random_rows(rng, 1500, flags, 0.05):
        inp[10] = np.float32(0.01) * np.sign(inp[10] or 1) + inp[10]   # |c| > 1e-4: no warm start
        a = run("ref", flags, inp, m, vel, b1=1, h=0.25, iters=3)
        b = run("simd", flags, inp, m, vel, b1=1, h=0.25, iters=3)
        for x, y in zip(a[:4], b[:4]):
            assert np.allclose(x, y, rtol=1e-5, atol=1e-6)
        # the same row with invEff forced to a power of two: scale the angular part away
        inp2 = inp.copy()
        inp2[3:9] = 0
        m2 = m.copy()
        m2[0], m2[10] = 0.5, 0.5          # invEff = |lin|^2 (0.5 + 0.5) -> make |lin| = 1
        inp2[0:3] = (1, 0, 0) if rng.random() < 0.5 else (0, 0, 1)
        if flags == ANGULAR:
            inp2[0:3] = 0
            inp2[3:6] = (0, 1, 0)
            inp2[6:9] = (0, 1, 0)
            m2[1:10] = np.diag([0.25, 0.5, 0.25]).ravel()
            m2[11:20] = np.diag([0.25, 0.5, 0.25]).ravel()
        a = run("ref", flags, inp2, m2, vel, b1=1, h=0.25, iters=3)
        b = run("simd", flags, inp2, m2, vel, b1=1, h=0.25, iters=3)
        assert a[5] == 1.0 and b[5] == 1.0
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x, np.float32).view(np.uint32), np.asarray(y, np.float32).view(np.uint32))
        exact += 1
    assert exact == 1500
