"""Count-plan parameters and the autotuner: any valid setting gives the
same supports bit for bit (only the time changes), invalid ones are refused,
and the tuned setting decomposes exactly as the default."""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

import rtrex_cuda_amd as R
from conftest import golden_graph, golden_meta, golden_result
from rtrex_cuda_amd import graphs as G

pytestmark = pytest.mark.gpu


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


SETTINGS = [
    {},
    {"hub_min": 0, "hub_max": 64, "l0_max": 64, "l1_max": 64},
    {"hub_min": 4, "l0_max": 16, "l1_max": 128, "g_thr": [1, 2, 4, 8]},
    {"hub_min": 64, "hub_max": 1024, "unit_min": 1, "unit_hash_mult": 1,
     "g_thr": [100, 200, 300, 400]},
    {"hub_min": 1, "hub_max": 4096, "l0_max": 0, "l1_max": 0},  # all hubs: 4096-entry trees
    # all hubs through the huge-hub kernel (upper tree levels staged, leaf in HBM)
    {"hub_min": 0, "hub_max": 32760, "l0_max": 0, "l1_max": 0, "l2_max": 0},
    {"hub_min": 2, "hub_max": 32760, "l0_max": 8, "l1_max": 32, "l2_max": 64,
     "g_thr": [2, 4, 8, 16]},
    # 64 lanes per edge: the filtered paths of the 4096-entry and huge-hub kernels
    {"hub_min": 1, "hub_max": 4096, "l0_max": 0, "l1_max": 0, "g_thr": [0, 0, 0, 0]},
    {"hub_min": 0, "hub_max": 32760, "l0_max": 0, "l1_max": 0, "l2_max": 0,
     "g_thr": [0, 0, 0, 0]},
    # all hubs through the five-level tree (levels from every 64th entry up in
    # LDS, the 8th-entry pivots and the leaf read from the list)
    {"hub_min": 0, "hub_max": 262136, "l0_max": 0, "l1_max": 0, "l2_max": 0, "l3_max": 0},
    {"hub_min": 0, "hub_max": 262136, "l0_max": 0, "l1_max": 0, "l2_max": 0, "l3_max": 0,
     "g_thr": [0, 0, 0, 0]},
]


@pytest.mark.parametrize("i", range(len(SETTINGS)))
def test_supports_do_not_depend_on_parameters(i):
    meta = golden_meta()["clean_astro_eps0.3"]
    n, u, v = golden_graph("clean_astro")
    print(f"[gpu] params {SETTINGS[i]}", flush=True)
    with R.DeviceGraph.from_edges(n, u, v) as g:
        g.set_count_params(SETTINGS[i])
        total = g.gpu_common_nbr()
        pe, pv, _ = g.sync_weights()
        assert total == meta["count_total"]
        assert sha(pe) == meta["count_perEdge_sha"] and sha(pv) == meta["count_perVertex_sha"]
        g.gpu_initial_clean(0.3)
        cm, k = g.disjoint_extract(0.3)
    assert np.array_equal(cm, golden_result("clean_astro_eps0.3")["cluster_map"])


def test_invalid_parameters_rejected():
    n, u, v = golden_graph("small_test")
    with R.DeviceGraph.from_edges(n, u, v) as g:
        for bad in ({"hub_max": 262137}, {"l3_max": 32761}, {"l2_max": 512, "l3_max": 256},
                    {"l0_max": 65}, {"l1_max": 600},
                    {"l2_max": 5000}, {"l1_max": 256, "l2_max": 128},
                    {"g_thr": [5, 4, 3, 2]}, {"hub_min": 100, "hub_max": 50}):
            with pytest.raises(R.RtrexError) as e:
                g.set_count_params(bad)
            assert e.value.status == 1


def test_autotune_keeps_results():
    n, u, v = G.config2()
    print("[gpu] autotune config2", flush=True)
    with R.DeviceGraph.from_edges(n, u, v) as g:
        cm0, k0 = g.decompose(0.3)
        params, best, start = g.autotune(repeats=1)
        assert best <= start
        assert g.count_params() == params
        cm1, k1 = g.decompose(0.3)
    assert k0 == k1 and np.array_equal(cm0, cm1)
