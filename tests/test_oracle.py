"""Pins the CPU oracle (oracle/rtrex_oracle.cpp).

The oracle is the checker of every GPU parity test, so it is itself pinned
here, on the CPU, by:
  * the reference's own outputs (tests/golden, written by make_golden.py from
    the reference sources compiled in the build container): CSR, triangle
    totals, FP64 supports bit for bit, survivors, cluster ids and stats;
  * brute force (trace(A^3)/6 triangle counts) and the K_n closed form;
  * the order-independent fixpoint that defines the clean phase, checked in
    exact rational arithmetic;
  * planted mistakes in a Python copy of the fold order, which the golden
    hashes must reject (so the pins can tell orders apart);
  * connected_components: the partition hash of the connected-components file
    the reference's own driver wrote, scipy's connected_components, and a
    closed form (cliques, a path, isolated vertices, whose ids would be -1
    if a start vertex were left unlabelled);
  * extract_inner (one seed's inner loop, the function the seed loop runs,
    so the golden clusters pin it too): values worked out by hand from
    Decomposition.h:728-827 on a four-vertex graph, with and without a
    deleted edge; cluster_edges (the seed loop's worklist builder) likewise
    from Decomposition.h:897-914.
"""
from __future__ import annotations

import hashlib
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from conftest import (GOLDEN_TAGS, golden_graph, golden_meta, golden_result,
                      survivors_of)


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _csr(name):
    n, u, v = golden_graph(name)
    return O.csr_from_undirected(n, u, v)


@pytest.mark.parametrize("tag", GOLDEN_TAGS)
def test_oracle_matches_reference_outputs(tag):
    meta = golden_meta()[tag]
    ref = golden_result(tag)
    name = tag.rsplit("_eps", 1)[0]
    off, nb = _csr(name)
    assert sha(np.concatenate([off, nb])) == meta["csr_sha"]  # makeCSR
    cnt = O.count(off, nb)
    assert cnt.total == meta["count_total"]
    assert sha(cnt.perEdge) == meta["count_perEdge_sha"]
    assert sha(cnt.perVertex) == meta["count_perVertex_sha"]
    cl = O.initial_clean(off, nb, meta["eps"], cnt)
    assert cl.total == meta["clean_total"]
    assert sha(cl.perEdge) == meta["clean_perEdge_sha"]
    assert sha(cl.perVertex) == meta["clean_perVertex_sha"]
    assert np.array_equal(survivors_of(off, nb, cl.status), ref["survivors"])
    d = O.decompose(off, nb, meta["eps"])
    assert d.n_clusters == meta["n_clusters"]
    assert np.array_equal(d.cluster_map, ref["cluster_map"])
    stats = np.array([[c[0], c[1], c[2]] for c in d.clusters], np.int64).reshape(-1, 3)
    assert np.array_equal(stats, ref["stats"])


def _random_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    a = np.triu(rng.random((n, n)) < p, 1)
    u, v = np.nonzero(a)
    return n, u.astype(np.uint32), v.astype(np.uint32)


@pytest.mark.parametrize("n,p,seed", [(30, 0.3, 1), (60, 0.15, 2), (45, 0.5, 3), (80, 0.08, 4)])
def test_triangle_total_brute_force(n, p, seed):
    n, u, v = _random_graph(n, p, seed)
    a = np.zeros((n, n), np.int64)
    a[u, v] = a[v, u] = 1
    off, nb = O.csr_from_undirected(n, u, v)
    assert O.count(off, nb).total == int(np.trace(a @ a @ a)) // 6


@pytest.mark.parametrize("k", [4, 5, 7, 10, 13])
def test_clique_closed_form(k):
    iu = np.array([(a, b) for a in range(k) for b in range(a + 1, k)], np.uint32)
    off, nb = O.csr_from_undirected(k, iu[:, 0], iu[:, 1])
    c = O.count(off, nb)
    t = 1.0 / float((k - 1) ** 3)
    s = 0.0
    for _ in range(k - 2):  # each edge: k-2 triangles of equal weight
        s += t
    assert np.all(c.perEdge == s)
    s = 0.0
    for _ in range((k - 1) * (k - 2) // 2):
        s += t
    assert np.all(c.perVertex == s)
    assert c.total == k * (k - 1) * (k - 2) // 6


def test_planted_fold_order_is_detected():
    """A descending-z fold (a plausible mistake) must not reproduce the
    reference's perEdge bits on clean_email."""
    off, nb = _csr("clean_email")
    meta = golden_meta()["clean_email_eps0.3"]
    deg = np.diff(off)
    pe = np.zeros(len(nb))
    nbl = [nb[off[i]:off[i + 1]] for i in range(len(off) - 1)]
    for x in range(len(off) - 1):
        for s in range(off[x], off[x + 1]):
            y = nb[s]
            if y < x:
                continue
            z = np.intersect1d(nbl[x], nbl[y])[::-1]  # wrong order
            acc = 0.0
            for w in z:
                acc += 1.0 / float(int(deg[x]) * int(deg[y]) * int(deg[w]))
            pe[s] = acc
            pe[off[y] + np.searchsorted(nbl[y], x)] = acc
    assert sha(pe) != meta["count_perEdge_sha"]
    assert np.allclose(pe, O.count(off, nb).perEdge, rtol=1e-12, atol=0)


def _exact_peel(off, nb, eps, status0):
    """Least fixpoint of 'delete edges below threshold' in exact rational
    arithmetic, starting from the reference's initial decisions."""
    n = len(off) - 1
    deg = np.diff(off)
    nbs = [set(nb[off[i]:off[i + 1]].tolist()) for i in range(n)]
    alive = {(x, y) for x in range(n) for y in nbs[x] if x < y}
    src = np.repeat(np.arange(n), np.diff(off))
    for i in range(len(nb)):
        if status0[i] and src[i] < nb[i]:
            alive.discard((int(src[i]), int(nb[i])))
    thr = {e: Fraction(eps) * Fraction(1.0 / float(int(deg[e[0]]) * int(deg[e[1]])))
           for e in alive}

    def sup(e):
        x, y = e
        s = Fraction(0)
        for z in nbs[x] & nbs[y]:
            if (min(x, z), max(x, z)) in alive and (min(y, z), max(y, z)) in alive:
                s += Fraction(1.0 / float(int(deg[x]) * int(deg[y]) * int(deg[z])))
        return s

    changed = True
    while changed:
        changed = False
        for e in sorted(alive):
            if sup(e) < thr[e]:
                alive.discard(e)
                changed = True
    return alive


@pytest.mark.parametrize("seed,eps", [(11, 0.2937), (12, 0.1713), (13, 0.4111)])
def test_clean_is_the_exact_fixpoint(seed, eps):
    n, u, v = _random_graph(40, 0.3, seed)
    off, nb = O.csr_from_undirected(n, u, v)
    cnt = O.count(off, nb)
    deg = np.diff(off)
    src = np.repeat(np.arange(n), np.diff(off))
    thr = eps * (1.0 / (deg[src].astype(np.float64) * deg[nb]))
    initial = cnt.perEdge < thr  # identical FP values to the reference's
    cl = O.initial_clean(off, nb, eps, cnt)
    exact = _exact_peel(off, nb, eps, initial)
    got = {tuple(map(int, e)) for e in survivors_of(off, nb, cl.status)}
    assert got == exact


def test_small_test_known_clusters():
    off, nb = _csr("small_test")
    d = O.decompose(off, nb, 0.3)
    sets = sorted(sorted(map(int, c[3])) for c in d.clusters)
    assert [1, 3, 4, 6] in sets
    assert sum(1 for s in sets if len(s) == 1) == 6


def test_disjoint_cliques_one_cluster_each():
    edges = []
    base = 0
    for k in (4, 5, 6, 7):
        edges += [(base + a, base + b) for a in range(k) for b in range(a + 1, k)]
        base += k
    e = np.array(edges, np.uint32)
    off, nb = O.csr_from_undirected(base, e[:, 0], e[:, 1])
    d = O.decompose(off, nb, 0.3)
    sizes = sorted(c[0] for c in d.clusters)
    assert sizes == [4, 5, 6, 7]


def test_delete_and_clean_removes_given_edge():
    off, nb = _csr("clique_k6")
    cnt = O.count(off, nb)
    cl = O.initial_clean(off, nb, 0.3, cnt)
    st = O.delete_and_clean(off, nb, 0.3, np.array([[0, 1]]), cl)
    s = off[0] + np.searchsorted(nb[off[0]:off[1]], 1)
    assert st.status[s] == ord("N")
    assert st.total == cl.total - 4  # K6: edge {0,1} lies in 4 triangles


# ---------------------------------------------------------------- connectedComp
def _golden_files():
    import json
    import os
    from conftest import GOLDEN
    with open(os.path.join(GOLDEN, "files.json")) as f:
        return json.load(f)


GRAPH_NAMES = sorted({t.rsplit("_eps", 1)[0] for t in GOLDEN_TAGS})


@pytest.mark.parametrize("name", GRAPH_NAMES)
def test_components_match_reference_driver(name):
    """The partition equals the one the reference's own driver wrote
    (connectedComp, Decomposition.h:1023-1068, via RTRex.cpp:203-204), and
    component ids follow the BFS start order: ascending smallest vertex."""
    from rtrex_cuda_amd import validate as V
    off, nb = _csr(name)
    comp, k = O.connected_components(off, nb)
    assert comp.min() >= 0 and comp.max() == k - 1
    assert V.canonical_cluster_sha(comp) == _golden_files()[f"{name}_eps0.3"][
        "connected-components-sets"]
    first = np.full(k, len(comp))
    np.minimum.at(first, comp, np.arange(len(comp)))
    assert np.all(np.diff(first) > 0)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_components_equal_textbook_routine(seed):
    """Same partition as scipy's connected_components on sparse random graphs
    with isolated vertices and several components."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components as sp_cc
    n, u, v = _random_graph(300, 0.004, seed)
    off, nb = O.csr_from_undirected(n, u, v)
    comp, k = O.connected_components(off, nb)
    a = coo_matrix((np.ones(len(u)), (u, v)), shape=(n, n))
    k2, lab = sp_cc(a, directed=False)
    assert k == k2
    # same partition: the label pairs form a bijection
    pairs = set(zip(comp.tolist(), lab.tolist()))
    assert len(pairs) == k


def test_components_closed_form():
    """Disjoint cliques, a path and isolated vertices: components known."""
    edges, base, want = [], 0, []
    for k in (3, 5):
        edges += [(base + a, base + b) for a in range(k) for b in range(a + 1, k)]
        want.append(list(range(base, base + k)))
        base += k
    want += [[base], [base + 1]]  # two isolated vertices
    base += 2
    edges += [(base + i, base + i + 1) for i in range(6)]  # a path of 7
    want.append(list(range(base, base + 7)))
    base += 7
    e = np.array(edges, np.uint32)
    off, nb = O.csr_from_undirected(base, e[:, 0], e[:, 1])
    comp, k = O.connected_components(off, nb)
    assert k == len(want)
    for i, w in enumerate(want):
        assert np.all(comp[w] == i)


@pytest.mark.parametrize("tag", ["clean_astro_eps0.1", "clean_email_eps0.3", "config1_eps0.1"])
def test_extract_from_state_equals_decompose(tag):
    """The extraction entered from the initial clean's state is the whole
    decomposition's (the same seed loop; pinned by the golden clusters)."""
    meta = golden_meta()[tag]
    off, nb = _csr(tag.rsplit("_eps", 1)[0])
    cl = O.initial_clean(off, nb, meta["eps"])
    cm, k = O.extract_from_state(off, nb, meta["eps"], cl)
    assert k == meta["n_clusters"]
    assert np.array_equal(cm, golden_result(tag)["cluster_map"])


def _fold(xs):
    acc = 0.0
    for x in xs:
        acc += x
    return acc


def test_extract_inner_worked_example():
    """Seed 0 of the graph 0-1, 0-2, 1-2, 1-3, 2-3 (degrees 2, 3, 3, 2) at eps
    0.3, worked out from the reference's loop (Decomposition.h:728-805): the
    pairs of alive neighbours are (1,1), (1,2), (2,2) (k >= j, so a neighbour
    pairs with itself); each adds 1/(2*3*3) to internal; the pair (1,1)
    enumerates N(1) and meets 2, a neighbour of 0 larger than both, adding
    1/(3*3*3); vertex 3 is a candidate reached from all three pairs, 1/18 each.
    A deleted edge {1,3} leaves only the pair (2,2) reaching it."""
    off, nb = O.csr_from_undirected(4, np.array([0, 0, 1, 1, 2]), np.array([1, 2, 2, 3, 3]))
    st = np.full(len(nb), ord("Y"), np.uint8)
    cand, wgt, internal, potential = O.extract_inner(off, nb, st, 0, 0.3)
    t = 1.0 / 18.0
    assert list(cand) == [3]
    assert wgt[0] == _fold([t, t, t]) and potential == wgt[0]
    assert internal == _fold([t, 1.0 / 27.0, t, t])
    src = np.repeat(np.arange(4), np.diff(off))
    dead = ((src == 1) & (nb == 3)) | ((src == 3) & (nb == 1))
    st[dead] = ord("N")
    cand, wgt, internal, potential = O.extract_inner(off, nb, st, 0, 0.3)
    assert list(cand) == [3] and wgt[0] == t and potential == t
    assert internal == _fold([t, 1.0 / 27.0, t, t])
    # seed 3 once {1,3} is deleted: its one alive neighbour 2 pairs with
    # itself; N(2) = {0, 1, 3} gives candidates 0 (1/(2*3*3)) and 1
    # (1/(3*3*3)), 1 no longer being a neighbour of 3
    cand, wgt, internal, potential = O.extract_inner(off, nb, st, 3, 0.3)
    assert list(cand) == [0, 1] and list(wgt) == [t, 1.0 / 27.0]
    assert internal == t and potential == t + 1.0 / 27.0


def test_cluster_edges_worked_example():
    """Members [1, 0] of the same graph (Decomposition.h:897-914): member 1
    lists its three 'Y' slots (1,0), (1,2), (1,3) and marks them and their
    partners 'S'; member 0 then skips (0,1) and lists (0,2).  A deleted slot
    pair is never listed; a member order [0, 1] lists {0,1} from 0."""
    off, nb = O.csr_from_undirected(4, np.array([0, 0, 1, 1, 2]), np.array([1, 2, 2, 3, 3]))
    st = np.full(len(nb), ord("Y"), np.uint8)
    assert O.cluster_edges(off, nb, st, [1, 0]).tolist() == [[1, 0], [1, 2], [1, 3], [0, 2]]
    assert O.cluster_edges(off, nb, st, [0, 1]).tolist() == [[0, 1], [0, 2], [1, 2], [1, 3]]
    src = np.repeat(np.arange(4), np.diff(off))
    st[((src == 1) & (nb == 2)) | ((src == 2) & (nb == 1))] = ord("N")
    assert O.cluster_edges(off, nb, st, [1, 0]).tolist() == [[1, 0], [1, 3], [0, 2]]
