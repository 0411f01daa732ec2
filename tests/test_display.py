"""The display conversion table (mp_display_thresholds) against the oracle's
byte conversion (gamma 1/2.2 of shaders/tex_draw.frag:25-37, powf, 8 bits):
the byte of a value must be the number of thresholds at or below it, checked
at every threshold's neighbourhood, over every binade and at the special
values.  No GPU needed: the table is computed on the host."""
import ctypes as C

import numpy as np
import pytest

from micropolis_amd import _lib


@pytest.fixture(scope="module")
def table():
    lib = _lib.load()
    t = np.zeros(255, np.float32)
    assert lib.mp_display_thresholds(t.ctypes.data_as(C.POINTER(C.c_float))) == 0
    return t


@pytest.fixture(scope="module")
def byte(oracle):
    f = oracle.load().or_to_srgb8
    return lambda a: np.array([f(float(x)) for x in a], np.int64)


def table_byte(t, a):
    return np.searchsorted(t, np.asarray(a, np.float32), side="right")


def test_thresholds_increase(table):
    assert (table > 0).all() and np.isfinite(table).all()
    assert (np.diff(table) > 0).all()


def test_neighbourhoods_of_every_threshold(table, byte):
    bits = table.view(np.int32).astype(np.int64)
    offs = np.arange(-64, 65)
    cand = np.unique((bits[:, None] + offs[None, :]).ravel())
    cand = cand[(cand > 0) & (cand < 0x7f800000)].astype(np.int32).view(np.float32)
    assert np.array_equal(byte(cand), table_byte(table, cand))


def test_every_binade_and_special_values(table, byte):
    rng = np.random.default_rng(5)
    exps = np.arange(0, 255, dtype=np.int64)
    mant = rng.integers(0, 1 << 23, size=(exps.size, 16))
    cand = ((exps[:, None] << 23) | mant).ravel().astype(np.int32).view(np.float32)
    special = np.array([0.0, -0.0, -1.0, -np.inf, np.inf, np.nan, 1e-45, 1.0, 0.5, 2.0],
                       np.float32)
    cand = np.concatenate([cand, special, rng.random(4000).astype(np.float32)])
    got = table_byte(table, cand)
    got[np.isnan(cand)] = 0
    assert np.array_equal(byte(cand), got)
