"""GPU parity on a skewed graph: R-MAT scale 15 (edge factor 16, a,b,c =
0.57,0.19,0.19, seed 5; 32,768 vertices, about 442k edges, hubs up to degree
~5.9k) against the CPU oracle.  Its hubs span every count class: lists staged
whole in LDS (degree <= 4096) and the huge-hub kernel above (upper tree levels
in LDS, leaf blocks read from the hub's list), with the default plan and with
a plan that sends every hub above degree 64 to the huge-hub kernel, and with
both at 64 lanes per edge for every hub (their filtered paths), and through
the five-level tree the hubs above degree 32760 take.  Counts
and FP64 supports bit for bit, survivors, clusters and stats exactly.  At eps
0.3 the clean removes every edge (R-MAT has little clustering); eps 0.05 and
0.02 leave 132k and 347k survivors and non-trivial clusters."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import rtrex_cuda_amd as R
from rtrex_cuda_amd import graphs as G

pytestmark = pytest.mark.gpu

PLANS = {
    "default": {},
    "huge_above_64": {"l1_max": 64, "l2_max": 64},
    # 64 lanes per edge for every hub: the filtered large- and huge-hub paths
    "large_64_lanes": {"l1_max": 64, "g_thr": [0, 0, 0, 0]},
    "huge_64_lanes": {"l1_max": 64, "l2_max": 64, "g_thr": [0, 0, 0, 0]},
    # every hub above degree 64 through the five-level tree of the largest hubs
    "five_levels_above_64": {"l1_max": 64, "l2_max": 64, "l3_max": 64},
    "five_levels_64_lanes": {"l1_max": 64, "l2_max": 64, "l3_max": 64, "g_thr": [0, 0, 0, 0]},
}

_graph = {}
_oracle = {}


def graph():
    if not _graph:
        n, u, v = G.rmat(15, 16, 0.57, 0.19, 0.19, seed=5)
        off, nb = O.csr_from_undirected(n, u, v)
        _graph.update(n=n, u=u, v=v, off=off, nb=nb, count=O.count(off, nb))
    return _graph


def oracle_decomposition(eps):
    if eps not in _oracle:
        g = graph()
        _oracle[eps] = O.decompose(g["off"], g["nb"], eps, with_clusters=True)
    return _oracle[eps]


@pytest.mark.parametrize("plan", sorted(PLANS))
@pytest.mark.parametrize("eps", [0.3, 0.05, 0.02])
def test_rmat15_against_oracle(plan, eps):
    g = graph()
    d = oracle_decomposition(eps)
    c = g["count"]
    print(f"[gpu] rmat15 plan={plan} eps={eps}", flush=True)
    with R.DeviceGraph.from_edges(g["n"], g["u"], g["v"]) as dg:
        dg.set_count_params(PLANS[plan])
        total = dg.gpu_common_nbr()
        pe, pv, _ = dg.sync_weights()
        assert total == c.total
        assert np.array_equal(pe.view(np.uint64), c.perEdge.view(np.uint64))
        assert np.array_equal(pv.view(np.uint64), c.perVertex.view(np.uint64))
        dg.gpu_initial_clean(eps)
        st = dg.sync_edge_status()
        assert np.array_equal(st == ord("Y"), d.status_after_clean == ord("Y"))
        cm, k = dg.disjoint_extract(eps)
        nv, ne, cut = dg.populate_stats(k)
        assert dg.audit()["ambiguous_clean"] == 0
    assert k == d.n_clusters and np.array_equal(cm, d.cluster_map)
    stats = np.array([[x[0], x[1], x[2]] for x in d.clusters], np.int64)
    assert np.array_equal(np.stack([nv, ne, cut], 1), stats)


@pytest.mark.parametrize("cascade_min", [1, 64])
@pytest.mark.parametrize("eps", [0.05, 0.02])
def test_rmat15_cascades_in_workgroups(cascade_min, eps, monkeypatch):
    """The extraction's cascades handed to the workgroup kernel (8 waves per
    cascade, the pop's triangle list split over the waves) from a stack of
    cascade_min entries on, through the RTREX_CASCADE_WG_MIN setting read at
    handle creation: 1 sends every cascade there, 64 hands cascades over in
    the middle.  Clusters and stats equal the oracle's."""
    g = graph()
    d = oracle_decomposition(eps)
    monkeypatch.setenv("RTREX_CASCADE_WG_MIN", str(cascade_min))
    print(f"[gpu] rmat15 cascades in workgroups from {cascade_min}, eps={eps}", flush=True)
    with R.DeviceGraph.from_edges(g["n"], g["u"], g["v"]) as dg:
        cm, k = dg.decompose(eps)
        nv, ne, cut = dg.populate_stats(k)
        a = dg.audit()
    assert a["ambiguous_clean"] == 0 and a["ambiguous_ratio"] == 0
    assert a["edges_removed"] == len(g["u"])
    assert k == d.n_clusters and np.array_equal(cm, d.cluster_map)
    stats = np.array([[x[0], x[1], x[2]] for x in d.clusters], np.int64)
    assert np.array_equal(np.stack([nv, ne, cut], 1), stats)
