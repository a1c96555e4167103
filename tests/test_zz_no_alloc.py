"""rtrex_cluster_stats_noalloc (no allocation; per-GPU static accumulators)
against rtrex_cluster_stats on the same handles, through the C ABI: two calls
in a row, the second on a smaller graph (fewer clusters), so that what the
first call left in the shared array must not leak into the second; the
cluster limit is refused with RTREX_EINVAL."""
from __future__ import annotations

import numpy as np
import pytest

import rtrex_cuda_amd as R
from conftest import golden_graph

pytestmark = pytest.mark.gpu


def _stats_both(name, eps):
    n, u, v = golden_graph(name)
    print(f"[gpu] no-alloc stats {name} eps {eps}", flush=True)
    with R.DeviceGraph.from_edges(n, u, v) as g:
        _, k = g.decompose(eps)
        a = g.populate_stats(k)
        b = g.populate_stats_noalloc(k)
        b2 = g.populate_stats_noalloc(k)  # same handle again: same results
    return k, a, b, b2


def test_noalloc_equals_allocating_entry_point_twice_second_smaller():
    k1, a1, b1, c1 = _stats_both("clean_astro", 0.3)   # 12,849 clusters
    k2, a2, b2, c2 = _stats_both("small_test", 0.3)    # 7 clusters
    assert k2 < k1
    for a, b, c in ((a1, b1, c1), (a2, b2, c2)):
        for x, y, z in zip(a, b, c):
            assert np.array_equal(x, y) and np.array_equal(x, z)


def test_noalloc_refuses_more_clusters_than_it_holds():
    n = (1 << 20) + 5  # isolated vertices: one singleton cluster each
    u = np.array([0], np.uint32)
    v = np.array([1], np.uint32)
    print("[gpu] no-alloc stats limit", flush=True)
    with R.DeviceGraph.from_edges(n, u, v) as g:
        _, k = g.decompose(0.3)
        assert k > (1 << 20)
        with pytest.raises(R.RtrexError) as e:
            g.populate_stats_noalloc(k)
        assert e.value.status == 1 and "at most 1048576" in str(e.value)
