"""GPU parity of the per-seed entry points (the reference's gpuExtractInner
and gpuMarkClusterEdges boundary, gpu_extract.cuh:28-59) against the CPU
oracle on the same state:

  rtrex_extract_inner  one seed's inner loop (Decomposition.h:728-827): the
                       candidates (ascending id), their FP64 weights, the
                       internal and potential weight, bit for bit, for every
                       seed with alive edges (or a spread sample) of the
                       golden graphs after the initial clean, after deletions,
                       and on R-MAT scale 15 (seeds with over 64 alive
                       neighbours and over 1,024 candidates, the paths that
                       leave registers and LDS);
  rtrex_cluster_edges  a cluster's deletion pairs (Decomposition.h:897-914)
                       for member lists in a given order, and those pairs fed
                       to rtrex_delete_and_clean give the oracle's state.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import rtrex_cuda_amd as R
from conftest import golden_graph
from rtrex_cuda_amd import graphs as G

pytestmark = pytest.mark.gpu


def _seed_sample(off, nb, st, limit, rng, max_alive=256):
    """Seeds with 1 to max_alive alive neighbours (the pair loop is quadratic
    in them, on one wave and in the oracle): all of them, or the busiest
    quarter of limit and a random rest."""
    src = np.repeat(np.arange(len(off) - 1), np.diff(off))
    alive = np.bincount(src[st != ord("N")], minlength=len(off) - 1)
    seeds = np.flatnonzero((alive > 0) & (alive <= max_alive))
    if len(seeds) > limit:
        top = seeds[np.argsort(alive[seeds], kind="stable")[-(limit // 4):]]
        rest = rng.choice(np.setdiff1d(seeds, top), limit - len(top), replace=False)
        seeds = np.unique(np.concatenate([top, rest]))
    return seeds


def _check_seeds(dg, off, nb, eps, seeds):
    st = dg.sync_edge_status()
    big = 0
    for u in seeds:
        c, w, i, p = dg.extract_inner(int(u), eps)
        oc, ow, oi, op = O.extract_inner(off, nb, st, int(u), eps)
        assert np.array_equal(c, oc), u
        assert np.array_equal(w.view(np.uint64), ow.view(np.uint64)), u
        assert (np.float64(i).view(np.uint64), np.float64(p).view(np.uint64)) == \
            (np.float64(oi).view(np.uint64), np.float64(op).view(np.uint64)), u
        big = max(big, len(c))
    return big


@pytest.mark.parametrize("name,eps", [("small_test", 0.3), ("random_graph_rtr", 0.1),
                                      ("clean_email", 0.3), ("clean_email", 0.1),
                                      ("clean_astro", 0.1)])
def test_extract_inner_matches_oracle(name, eps):
    n, u, v = golden_graph(name)
    off, nb = O.csr_from_undirected(n, u, v)
    print(f"[gpu] extract_inner {name} eps={eps}", flush=True)
    with R.DeviceGraph.from_edges(n, u, v) as dg:
        dg.gpu_common_nbr()
        dg.gpu_initial_clean(eps)
        st = dg.sync_edge_status()
        seeds = _seed_sample(off, nb, st, 1500, np.random.default_rng(1))
        _check_seeds(dg, off, nb, eps, seeds)
        # the call leaves the state as it was: the extraction still matches
        cm, k = dg.disjoint_extract(eps)
    ref = O.decompose(off, nb, eps, with_clusters=False)
    assert k == ref.n_clusters and np.array_equal(cm, ref.cluster_map)


def test_extract_inner_large_neighbourhoods():
    """R-MAT scale 15 at eps 0.02: hubs keep hundreds of alive neighbours and
    reach thousands of candidates."""
    n, u, v = G.rmat(15, 16, 0.57, 0.19, 0.19, seed=5)
    off, nb = O.csr_from_undirected(n, u, v)
    eps = 0.02
    print("[gpu] extract_inner rmat15", flush=True)
    with R.DeviceGraph.from_edges(n, u, v) as dg:
        dg.gpu_common_nbr()
        dg.gpu_initial_clean(eps)
        st = dg.sync_edge_status()
        seeds = _seed_sample(off, nb, st, 160, np.random.default_rng(2))
        src = np.repeat(np.arange(n), np.diff(off))
        alive = np.bincount(src[st != ord("N")], minlength=n)
        assert alive[seeds].max() > 64  # neighbours read per pair, not from registers
        big = _check_seeds(dg, off, nb, eps, seeds)
    assert big > 1024  # candidates sorted outside LDS


def test_cluster_edges_then_delete_matches_oracle():
    """Clusters taken as given member lists: their pairs equal the oracle's,
    and deleting them (rtrex_delete_and_clean) leaves the oracle's statuses;
    extract_inner on that state matches again."""
    n, u, v = golden_graph("clean_email")
    off, nb = O.csr_from_undirected(n, u, v)
    eps = 0.1
    state = O.initial_clean(off, nb, eps)
    rng = np.random.default_rng(3)
    print("[gpu] cluster_edges clean_email", flush=True)
    with R.DeviceGraph.from_edges(n, u, v) as dg:
        dg.gpu_common_nbr()
        dg.gpu_initial_clean(eps)
        for step in range(6):
            st = dg.sync_edge_status()
            src = np.repeat(np.arange(n), np.diff(off))
            alive_v = np.unique(src[st == ord("Y")])
            z = int(rng.choice(alive_v))
            nbrs = nb[off[z]:off[z + 1]][st[off[z]:off[z + 1]] == ord("Y")]
            members = np.concatenate([[z], rng.permutation(nbrs)[:5]]).astype(np.int64)
            pairs = dg.cluster_edges(members)
            want = O.cluster_edges(off, nb, st, members)
            assert np.array_equal(pairs, want), step
            assert np.array_equal(dg.sync_edge_status(), st)  # statuses untouched
            dg.gpu_delete_and_clean(pairs, eps)
            state = O.delete_and_clean(off, nb, eps, pairs, state)
            got = dg.sync_edge_status()
            assert np.array_equal(got == ord("Y"), state.status == ord("Y")), step
            seeds = _seed_sample(off, nb, got, 200, rng)
            _check_seeds(dg, off, nb, eps, seeds)


def test_per_seed_arguments_rejected():
    n, u, v = golden_graph("small_test")
    with R.DeviceGraph.from_edges(n, u, v) as dg:
        with pytest.raises(R.RtrexError):  # before the clean
            dg.extract_inner(0, 0.3)
        dg.gpu_common_nbr()
        dg.gpu_initial_clean(0.3)
        for bad in (-1, n):
            with pytest.raises(R.RtrexError):
                dg.extract_inner(bad, 0.3)
        with pytest.raises(R.RtrexError):
            dg.extract_inner(0, 0.0)
        with pytest.raises(R.RtrexError):
            dg.cluster_edges([0, 0])
        with pytest.raises(R.RtrexError):
            dg.cluster_edges([n])
        assert dg.cluster_edges([]).shape == (0, 2)
