"""Serving (rtrex_server_*): every job's result in a concurrent, batched mix
equals the same graph decomposed alone through the plain path; admission by
HBM footprint; a bad job fails alone; an oversized job runs outside the pool."""
from __future__ import annotations

import numpy as np
import pytest

import rtrex_cuda_amd as R
from rtrex_cuda_amd import graphs as G

pytestmark = pytest.mark.gpu

EPS = 0.3


def _mix(count: int, seed: int):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        r = rng.random()
        if r < 0.7:
            m = int(np.exp(rng.uniform(np.log(1_000), np.log(20_000))))
            out.append(G.chung_lu(max(100, m // 5), m, 2.5, 60.0, seed=seed * 1000 + i))
        else:
            n = int(rng.integers(2_000, 12_000))
            out.append(G.planted(n, 0, 5, 40, 0.4, 2.5, 3 * n, max(8.0, n ** 0.5),
                                 seed=seed * 1000 + i))
    return out


def _alone(n, u, v, eps=EPS):
    with R.DeviceGraph.from_edges(n, u, v) as g:
        return g.decompose(eps)


def test_results_equal_alone_in_the_mix():
    graphs = _mix(48, 5)
    print(f"[gpu] serve mix of {len(graphs)} graphs", flush=True)
    # A budget of a few launches: admission has to wait, launches overlap.
    with R.Server(workers=4, hbm_budget_bytes=256 << 20, small_job_edges=16_384,
                  batch_max_edges=1 << 18) as s:
        ids = [s.submit(n, u, v, EPS) for n, u, v in graphs]
        res = [s.wait(j) for j in ids]
    for (n, u, v), (cm, k, info) in zip(graphs, res):
        cm0, k0 = _alone(n, u, v)
        assert k == k0 and np.array_equal(cm, cm0)
    infos = [r[2] for r in res]
    assert any(i["batch_graphs"] > 1 for i in infos), "no job was batched"
    assert all(i["pooled"] == 1 for i in infos)
    assert all(i["submit_s"] <= i["start_s"] <= i["end_s"] for i in infos)
    # launches of different workers overlapped in time
    spans = sorted((i["start_s"], i["end_s"], i["worker"]) for i in infos)
    assert any(a[1] > b[0] and a[2] != b[2] for a, b in zip(spans, spans[1:]))


def test_bad_job_fails_alone():
    good = _mix(6, 7)
    n, u, v = good[0]
    bad_u = np.concatenate([u, u[:1]])  # duplicate edge
    bad_v = np.concatenate([v, v[:1]])
    with R.Server(workers=2, hbm_budget_bytes=128 << 20, small_job_edges=1 << 20) as s:
        ids = [s.submit(*g, EPS) for g in good[:3]]
        jb = s.submit(n, bad_u, bad_v, EPS)
        ids += [s.submit(*g, EPS) for g in good[3:]]
        with pytest.raises(R.RtrexError) as e:
            s.wait(jb)
        assert e.value.status == 1
        res = [s.wait(j) for j in ids]
    for (gn, gu, gv), (cm, k, _) in zip(good, res):
        cm0, k0 = _alone(gn, gu, gv)
        assert k == k0 and np.array_equal(cm, cm0)


def test_oversized_job_runs_outside_the_pool():
    n, u, v = G.chung_lu(3_000, 20_000, 2.5, 60.0, seed=3)
    with R.Server(workers=1, hbm_budget_bytes=1 << 20) as s:
        cm, k, info = s.wait(s.submit(n, u, v, EPS))
    assert info["pooled"] == 0
    cm0, k0 = _alone(n, u, v)
    assert k == k0 and np.array_equal(cm, cm0)
