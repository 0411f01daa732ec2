"""Pins the split schedules: each oracle schedule is checked against an
independent Python restatement driven by the (separately pinned) bound()
predicate; the literal LOCAL kernel emulation, with and without output
overflow, yields the same DRAW multiset; the scan follows the reference's
own self-test (src/micropolis/main.cpp:343-400: srand(43), rand()%8+1)."""
import ctypes as C
import ctypes.util

import numpy as np
import pytest

from micropolis_amd import scenes


def small_scene():
    w = scenes.c1_sphere()
    o = w.views[0].objects[0]
    return o.patches, o.matrix


def setup(oracle, ppp=1024, limit=8.0):
    P, mv = small_scene()
    view = oracle.make_view(60.0, 0.01, 1024, 768)
    prm = oracle.make_params(reyes_patches_per_pass=ppp)
    return oracle.as_cps(P), mv, view, prm, limit


def py_bound(oracle, cps, mv, view, prm, limit, r):
    mvc = oracle.colmajor(mv)
    return oracle.load().or_bound(oracle.fptr(cps), 0, int(r[0]), int(r[1]), *map(float, r[2:]),
                                  oracle.fptr(mvc), C.byref(view), C.byref(prm), limit)


def children(r, f):
    pid, d, u0, v0, u1, v1 = r
    cx = np.float32((np.float32(u0) + np.float32(u1)) * np.float32(0.5))
    cy = np.float32((np.float32(v0) + np.float32(v1)) * np.float32(0.5))
    if f & 4:
        return (pid, d + 1, u0, v0, cx, v1), (pid, d + 1, cx, v0, u1, v1)
    return (pid, d + 1, u0, v0, u1, cy), (pid, d + 1, u0, cy, u1, v1)


def as_tuples(arr):
    return [(int(a["pid"]), int(a["depth"]), np.float32(a["umin"]), np.float32(a["vmin"]),
             np.float32(a["umax"]), np.float32(a["vmax"])) for a in arr]


def brute_force_set(oracle, cps, mv, view, prm, limit):
    out = []

    def rec(r):
        f = py_bound(oracle, cps, mv, view, prm, limit, r)
        if f & 1:
            out.append(r)
        elif f & 2:
            a, b = children(r, f)
            rec(a)
            rec(b)
    for pid in range(cps.shape[0]):
        rec((pid, 0, np.float32(0), np.float32(0), np.float32(1), np.float32(1)))
    return out


def test_local_and_balanced_equal_brute_force_set(oracle):
    cps, mv, view, prm, limit = setup(oracle)
    ref = sorted(brute_force_set(oracle, cps, mv, view, prm, limit))
    for m in ("LOCAL", "BALANCED"):
        r, pc, nb, nd = oracle.split(m, cps, 0, mv, view, prm, limit)
        assert sorted(as_tuples(r)) == ref
        assert nd == 0 and pc[0] == len(ref)


def test_breadth_level_order_and_truncation(oracle):
    cps, mv, view, prm, limit = setup(oracle, ppp=300)
    level = [(p, 0, np.float32(0), np.float32(0), np.float32(1), np.float32(1)) for p in range(cps.shape[0])]
    emitted, counts, dropped = [], [], 0
    while level:
        draws, nxt = [], []
        for r in level:
            f = py_bound(oracle, cps, mv, view, prm, limit, r)
            if f & 1:
                draws.append(r)
            elif f & 2:
                nxt.extend(children(r, f))
        counts.append(len(draws))
        emitted.extend(draws[:300])
        dropped += max(0, len(draws) - 300)
        level = nxt
    r, pc, nb, nd = oracle.split("BREADTH", cps, 0, mv, view, prm, limit)
    assert as_tuples(r) == emitted
    assert pc.tolist() == counts and nd == dropped and dropped > 0


def test_bounded_lifo_batches(oracle):
    cps, mv, view, prm, limit = setup(oracle, ppp=100)
    stack = [(p, 0, np.float32(0), np.float32(0), np.float32(1), np.float32(1)) for p in range(cps.shape[0])]
    emitted, counts = [], []
    while stack:
        b = min(len(stack), 100)
        batch = stack[len(stack) - b:]
        del stack[len(stack) - b:]
        draws, kids = [], []
        for r in batch:
            f = py_bound(oracle, cps, mv, view, prm, limit, r)
            if f & 1:
                draws.append(r)
            elif f & 2:
                kids.extend(children(r, f))
        stack.extend(kids)
        emitted.extend(draws)
        counts.append(len(draws))
    r, pc, nb, nd = oracle.split("BOUNDED", cps, 0, mv, view, prm, limit)
    assert as_tuples(r) == emitted and pc.tolist() == counts


@pytest.mark.parametrize("wgs,wgc,batch", [(64, 128, 100000), (4, 7, 100000), (64, 16, 500), (8, 3, 97)])
def test_local_kernel_emulation_same_multiset(oracle, wgs, wgc, batch):
    cps, mv, view, prm, limit = setup(oracle)
    ref = sorted(brute_force_set(oracle, cps, mv, view, prm, limit))
    cap = 1 << 16
    out = np.zeros(cap, oracle.RANGE_DTYPE)
    pc = np.zeros(4096, np.int32)
    npass, nb = C.c_int32(), C.c_long()
    n = oracle.load().or_split_local_emulated(
        oracle.fptr(cps), 0, cps.shape[0], oracle.fptr(oracle.colmajor(mv)), C.byref(view), C.byref(prm),
        limit, wgs, wgc, batch, out.ctypes.data, cap, pc.ctypes.data, 4096, C.byref(npass), C.byref(nb))
    assert n == len(ref)
    assert sorted(as_tuples(out[:n])) == ref
    if batch < len(ref):
        assert npass.value > 1          # output overflow forced extra passes


def test_canonical_sort_orders_by_patch_then_umin_vmin(oracle):
    cps, mv, view, prm, limit = setup(oracle)
    r, _, _, _ = oracle.split("BREADTH", cps, 0, mv, view, prm, limit)
    s = oracle.canonical_sort(r)
    keys = list(zip(s["pid"], s["umin"], s["vmin"]))
    assert keys == sorted(keys) and len(set(keys)) == len(keys)


def test_glibc_rand_sequence_matches_libc(oracle):
    libc = C.CDLL(ctypes.util.find_library("c"))
    libc.srand(43)
    expect = [libc.rand() for _ in range(2000)]
    got = np.zeros(2000, np.int32)
    oracle.load().or_glibc_rand(43, got.ctypes.data, 2000)
    assert got.tolist() == expect


@pytest.mark.parametrize("n", [1, 2, 20, 100, 128, 200, 512, 800, 1000, 1024, 2048, 4096, 5000,
                               128 * 128, 128 * 128 * 128, 50000])
def test_prefix_sum_reference_self_test(oracle, n):
    vals = np.zeros(n, np.int32)
    oracle.load().or_glibc_rand(43, vals.ctypes.data, n)
    vals = vals % 8 + 1
    out = np.zeros(n, np.int32)
    total = oracle.load().or_prefix_sum(vals.ctypes.data, out.ctypes.data, n)
    assert np.array_equal(out, np.cumsum(vals, dtype=np.int64).astype(np.int32))
    assert total == int(vals.sum())
