"""Host builds of device routines and the generators (CPU)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import rtrex_cuda_amd as R
from rtrex_cuda_amd import build as B
from rtrex_cuda_amd import graphs as G


def _check_lib():
    L = C.CDLL(B.CHECK_LIB)
    for f in (L.rtrex_hostcheck_introsort, L.rtrex_hostcheck_stdsort):
        f.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_double), C.c_int64]
        f.restype = None
    return L


@pytest.mark.parametrize("seed", range(8))
def test_introsort_reproduces_std_sort_ties(seed):
    """The extraction's candidate order must be std::sort's, ties included
    (Decomposition.h:829 sorts with an unstable comparison)."""
    L = _check_lib()
    rng = np.random.default_rng(seed)
    for trial in range(300):
        n = int(rng.integers(0, 2500 if trial % 10 == 0 else 120))
        kinds = int(rng.integers(1, 9))
        w = rng.integers(0, kinds, n).astype(np.float64) * 0.125
        if trial % 7 == 0:
            w = np.arange(n, 0, -1, dtype=np.float64)
        a = np.arange(n, dtype=np.uint32)
        b = a.copy()
        L.rtrex_hostcheck_introsort(a.ctypes.data_as(C.POINTER(C.c_uint32)),
                                    w.ctypes.data_as(C.POINTER(C.c_double)), n)
        L.rtrex_hostcheck_stdsort(b.ctypes.data_as(C.POINTER(C.c_uint32)),
                                  w.ctypes.data_as(C.POINTER(C.c_double)), n)
        assert np.array_equal(a, b)


def _simple(n, u, v):
    assert np.all(u < v) and np.all(v < n)
    k = u.astype(np.uint64) << 32 | v
    assert np.all(np.diff(k.astype(np.int64)) > 0)  # sorted, no duplicates


def test_generators_deterministic_and_simple():
    a = G.chung_lu(2000, 8000, 2.5, 100, seed=5)
    b = G.chung_lu(2000, 8000, 2.5, 100, seed=5)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    _simple(*a)
    c = G.chung_lu(2000, 8000, 2.5, 100, seed=6)
    assert not np.array_equal(a[1], c[1])
    p = G.planted(3000, 50, 5, 30, 0.4, 2.5, 3000, 50, seed=3)
    _simple(*p)
    r = G.rmat(12, 8, 0.57, 0.19, 0.19, seed=1)
    _simple(*r)


def test_config1_shape():
    n, u, v = G.config1()
    assert n == 10_000 and len(u) == 100_000
    deg = np.bincount(np.concatenate([u, v]), minlength=n)
    assert deg.max() <= 500


def test_config4_graph_shape():
    for i in range(4):
        n, u, v = G.config4_graph(i)
        assert 50 <= n <= 20_000
        _simple(n, u, v)


def test_weights_sidecar():
    w = G.weights(1000, "uniform", 1)
    assert w.dtype == np.float32 and np.all((w > 0) & (w <= 1))
    w = G.weights(1000, "lognormal", 1)
    assert np.all((w > 0) & (w <= 10))


def test_server_job_table_under_concurrent_submit_and_wait():
    """The serving loop's job table (jobtable.hpp): submits that rehash the
    map while waiters hold claims, two waiters per id (one refused), every
    job waited on exactly once, nothing left behind."""
    L = C.CDLL(B.CHECK_LIB)
    L.rtrex_hostcheck_jobtable.restype = C.c_int64
    L.rtrex_hostcheck_jobtable.argtypes = [C.c_int, C.c_int, C.c_int]
    for _ in range(3):
        assert L.rtrex_hostcheck_jobtable(4, 400, 6) == 0
