"""Pins the oracle's patch evaluation and noise against mathematics:
FP64 Bernstein closed form, corner interpolation, partition of unity, the
Bezier -> Gregory mapping of tools/bez2mscene_greg.py:111, and lattice
properties of the value noise."""
import ctypes as C

import numpy as np

from micropolis_amd.scenes import BEZ_TO_GREGORY


def _eval(oracle, cps, ptype, pid, u, v):
    out = np.zeros(4, np.float32)
    oracle.load().or_eval_patch(oracle.fptr(cps), ptype, pid, u, v, oracle.fptr(out))
    return out


def _eval64(oracle, cps, ptype, pid, u, v):
    c64 = np.ascontiguousarray(cps.astype(np.float64))
    out = np.zeros(4, np.float64)
    oracle.load().or_eval_patch_f64(c64.ctypes.data, ptype, pid, u, v, out.ctypes.data)
    return out


def _bernstein_np(P, u, v):
    """Independent FP64 tensor-product Bernstein sum (k = 4*i + j, i along u)."""
    bu = np.array([(1 - u) ** 3, 3 * u * (1 - u) ** 2, 3 * u * u * (1 - u), u ** 3])
    bv = np.array([(1 - v) ** 3, 3 * v * (1 - v) ** 2, 3 * v * v * (1 - v), v ** 3])
    return np.einsum("i,j,ijc->c", bu, bv, P.reshape(4, 4, -1))


def test_bezier_matches_fp64_closed_form(oracle):
    rng = np.random.default_rng(0)
    cps = oracle.as_cps(rng.uniform(-1, 1, (5, 16, 3)))
    for pid in range(5):
        for u, v in rng.uniform(0, 1, (20, 2)):
            f32 = _eval(oracle, cps, 0, pid, u, v)
            ref = _bernstein_np(cps[pid].astype(np.float64), float(np.float32(u)), float(np.float32(v)))
            assert np.allclose(f32, ref, rtol=0, atol=2e-6)
            assert np.allclose(_eval64(oracle, cps, 0, pid, float(np.float32(u)), float(np.float32(v))), ref,
                               atol=1e-12)


def test_bezier_corners_are_exact_and_weights_sum_to_one(oracle):
    rng = np.random.default_rng(1)
    cps = oracle.as_cps(rng.uniform(-3, 3, (1, 16, 3)))
    for (u, v), k in (((0, 0), 0), ((1, 0), 12), ((0, 1), 3), ((1, 1), 15)):
        assert np.array_equal(_eval(oracle, cps, 0, 0, u, v), cps[0, k])
    for u, v in rng.uniform(0, 1, (50, 2)):
        w = _eval(oracle, cps, 0, 0, u, v)[3]
        assert abs(w - 1.0) < 4e-7


def test_gregory_of_a_bezier_equals_the_bezier(oracle):
    rng = np.random.default_rng(2)
    bez = rng.uniform(-1, 1, (3, 16, 3))
    greg = bez[:, BEZ_TO_GREGORY]
    cb, cg = oracle.as_cps(bez), oracle.as_cps(greg)
    for pid in range(3):
        for u, v in list(rng.uniform(0, 1, (20, 2))) + [(0, 0), (1, 0), (0, 1), (1, 1), (0.5, 0)]:
            gb = _eval64(oracle, cg, 1, pid, u, v)
            bb = _eval64(oracle, cb, 0, pid, u, v)
            assert np.allclose(gb, bb, atol=1e-12)
            g32 = _eval(oracle, cg, 1, pid, u, v)
            assert np.all(np.isfinite(g32))
            assert np.allclose(g32, bb, atol=5e-6)


def test_gregory_corner_singularities_give_corner_points(oracle):
    rng = np.random.default_rng(3)
    P = rng.uniform(-1, 1, (1, 20, 3))
    cps = oracle.as_cps(P)
    # corners P0..P3 sit at (0,0), (0,1), (1,1), (1,0) (utility.h:101-123)
    for (u, v), k in (((0, 0), 0), ((0, 1), 1), ((1, 1), 2), ((1, 0), 3)):
        assert np.array_equal(_eval(oracle, cps, 1, 0, u, v), cps[0, k])


def test_gregory_interior_uses_both_face_points(oracle):
    """With f+ != f-, the interior blends them rationally: moving f- moves
    the surface except on the edges where its weight vanishes."""
    rng = np.random.default_rng(4)
    P = rng.uniform(-1, 1, (1, 20, 3))
    Q = P.copy()
    Q[0, 16] += 0.5           # fm of corner 0 (paired with fp P12 in F0)
    a, b = oracle.as_cps(P), oracle.as_cps(Q)
    assert not np.allclose(_eval(oracle, a, 1, 0, 0.3, 0.3), _eval(oracle, b, 1, 0, 0.3, 0.3))
    assert np.array_equal(_eval(oracle, a, 1, 0, 0.0, 0.4), _eval(oracle, b, 1, 0, 0.0, 0.4))


def test_value_noise_lattice_and_bounds(oracle):
    lib = oracle.load()
    prm = oracle.make_params(displacement=1, displacement_frequency=1.0, displacement_octaves=1,
                             displacement_seed=9)
    vals = [lib.or_noise(C.byref(prm), float(x), float(y), float(z))
            for x in range(-3, 3) for y in range(-3, 3) for z in range(-3, 3)]
    vals = np.array(vals)
    assert np.all(vals >= -1) and np.all(vals < 1)
    assert abs(vals.mean()) < 0.2 and vals.std() > 0.3
    # continuity across a lattice cell boundary and inside it
    for p in ([0.0, 0.5, 0.5], [0.37, 1.2, -0.8]):
        a = lib.or_noise(C.byref(prm), p[0] - 1e-4, p[1], p[2])
        b = lib.or_noise(C.byref(prm), p[0] + 1e-4, p[1], p[2])
        assert abs(a - b) < 2e-3
    prm4 = oracle.make_params(displacement=1, displacement_frequency=4.0, displacement_octaves=4)
    rng = np.random.default_rng(5)
    v4 = np.array([lib.or_noise(C.byref(prm4), *map(float, p)) for p in rng.uniform(-2, 2, (200, 3))])
    assert np.all(np.abs(v4) <= 1.0)


def _displaced(oracle, cps, u, v, prm):
    out = np.zeros(4, np.float32)
    oracle.load().or_displaced_point(oracle.fptr(cps), 0, 0, u, v, C.byref(prm), oracle.fptr(out))
    return out


def _unit_normal64(P, u, v):
    """FP64 unit(cross(dP/dv, dP/du)) of the Bernstein patch (i along u)."""
    b = lambda t: np.array([(1 - t) ** 3, 3 * t * (1 - t) ** 2, 3 * t * t * (1 - t), t ** 3])
    db = lambda t: np.array([-3 * (1 - t) ** 2, 3 * (1 - t) * (1 - 3 * t), 3 * t * (2 - 3 * t), 3 * t * t])
    G = P.reshape(4, 4, 3).astype(np.float64)
    pu = np.einsum("i,j,ijc->c", db(u), b(v), G)
    pv = np.einsum("i,j,ijc->c", b(u), db(v), G)
    n = np.cross(pv, pu)
    return n / np.linalg.norm(n)


def test_displacement_follows_the_surface_normal(oracle):
    """P + amp * noise(P) * n with n the unit normal cross(dP/dv, dP/du)."""
    lib = oracle.load()
    prm = oracle.make_params(displacement=1, displacement_amplitude=0.05, displacement_frequency=3.0,
                             displacement_octaves=4, displacement_seed=11)
    # planar patch z = 0.3: displacement along z only, x and y unchanged
    P = np.zeros((1, 16, 3), np.float32)
    for i in range(4):
        for j in range(4):
            P[0, 4 * i + j] = (i / 3.0, j / 3.0, 0.3)
    cps = oracle.as_cps(P)
    for u, v in ((0.2, 0.7), (0.5, 0.5), (0.9, 0.1)):
        base = _eval(oracle, cps, 0, 0, u, v)
        d = _displaced(oracle, cps, u, v, prm)
        assert d[0] == base[0] and d[1] == base[1]
        nz = lib.or_noise(C.byref(prm), float(base[0]), float(base[1]), float(base[2]))
        # cross(dP/dv, dP/du) of this parametrisation points along -z
        np.testing.assert_allclose(d[2] - base[2], -0.05 * nz, rtol=1e-5, atol=1e-7)
    # curved patches: direction and length against the FP64 normal
    rng = np.random.default_rng(12)
    for trial in range(20):
        P = rng.uniform(-1, 1, (1, 16, 3)).astype(np.float32)
        cps = oracle.as_cps(P)
        u, v = rng.uniform(0.05, 0.95, 2).astype(np.float32)
        base = _eval(oracle, cps, 0, 0, u, v)
        d = (_displaced(oracle, cps, u, v, prm) - base)[:3].astype(np.float64)
        nz = lib.or_noise(C.byref(prm), float(base[0]), float(base[1]), float(base[2]))
        np.testing.assert_allclose(d, 0.05 * nz * _unit_normal64(P[0], float(u), float(v)), rtol=0, atol=2e-6)
