"""The tolerance report and canonical hashes (rtrex_cuda_amd.validate), the
counterpart of the reference's RTRex_CUDA/validate.py, on the CPU."""
from __future__ import annotations

import math
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT
from rtrex_cuda_amd import validate as V


def _naive_match(ref, other, threshold=0.5):
    # the rule as RTRex_CUDA/validate.py:86-106 states it, over every candidate
    avail = list(range(len(other)))
    out = []
    for i, c in enumerate(ref):
        best, bj = 0, -1
        for j in avail:
            s = V.jaccard(c, other[j])
            if s > best:
                best, bj = s, j
        if best >= threshold:
            out.append((i, bj, best))
            avail.remove(bj)
    return out


@pytest.mark.parametrize("seed", range(6))
def test_match_equals_scan_of_every_candidate(seed):
    rng = np.random.default_rng(seed)
    n = 120
    a = rng.integers(0, 25, n)
    b = a.copy()
    flip = rng.random(n) < 0.25
    b[flip] = rng.integers(0, 25, flip.sum())
    ref = [s for s in V.clusters_from_map(np.unique(a, return_inverse=True)[1])]
    oth = [s for s in V.clusters_from_map(np.unique(b, return_inverse=True)[1])]
    got = V.match_clusters(ref, oth)
    want = _naive_match(ref, oth)
    assert [(i, j) for i, j, _ in got] == [(i, j) for i, j, _ in want]
    assert np.allclose([s for *_, s in got], [s for *_, s in want])


def test_tie_takes_lowest_candidate():
    ref = [{1, 2, 3, 4}]
    oth = [{9}, {3, 4}, {1, 2}]  # both candidates 0.5
    assert V.match_clusters(ref, oth) == [(0, 1, 0.5)]


def test_js_divergence_closed_forms():
    h = {1: 5, 2: 3, 7: 1}
    assert V.js_divergence(h, h) == 0.0
    assert math.isclose(V.js_divergence({1: 4}, {2: 9}), 1.0)  # disjoint supports
    # p = (1/2, 1/2), q = (1, 0): 0.5 * (0.5 log2(2/3)*... ) computed by hand
    p, q = [0.5, 0.5], [1.0, 0.0]
    m = [0.75, 0.25]
    want = 0.5 * (p[0] * math.log2(p[0] / m[0]) + p[1] * math.log2(p[1] / m[1])
                  + q[0] * math.log2(q[0] / m[0]))
    assert math.isclose(V.js_divergence({1: 1, 2: 1}, {1: 3}), want)


def test_report_bands_and_identity():
    ref = [{0, 1, 2}, {3, 4}, {5}, {6}]
    r = V.report(ref, ref)
    assert r["jsd"] == 0.0 and r["jsd_band"] == "excellent"
    assert r["clustered_ratio"] == 1.0 and r["clustered_band"] == "within 2%"
    assert r["ref_matched"] == 4 and r["mean_jaccard"] == 1.0  # singletons match too
    worse = [{0, 1, 2}, {3}, {4}, {5}, {6}]
    r = V.report(ref, worse)
    assert r["clustered_ratio"] == 3 / 5 and r["clustered_band"] == "outside 5%"


def test_canonical_hash_ignores_ids_and_order():
    cm = np.array([2, 0, 2, 1, 0, 1, 3])
    perm = np.array([3, 1, 0, 2])
    assert V.canonical_cluster_sha(cm) == V.canonical_cluster_sha(perm[cm])
    sets = V.clusters_from_map(cm)
    assert V.canonical_cluster_sha(cm) == V.canonical_sets_sha(sets[::-1])
    moved = cm.copy()
    moved[6] = 1  # a vertex moved to another cluster changes the hash
    assert V.canonical_cluster_sha(moved) != V.canonical_cluster_sha(cm)


def test_cli_on_two_files(tmp_path):
    a = tmp_path / "a.txt"
    b = tmp_path / "b.txt"
    a.write_text("1 3 4 6 \n0 \n2 \n5 \n")
    b.write_text("6 4 3 1 \n5 \n2 \n0 \n")
    out = subprocess.run([sys.executable, "-m", "rtrex_cuda_amd.validate", str(a), str(b)],
                         cwd=ROOT, capture_output=True, text=True, check=True).stdout
    assert "Identical partitions: True" in out
    assert "JS-divergence: 0.000000" in out
