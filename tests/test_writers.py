"""Output writers (rtrex_write_*; host code, no GPU) against the files the
reference's own driver writes (tests/golden/files.json, from oracle/_ref/RTRex
built from RTRex_Sequential/clustering/RTRex.cpp).  The clusters and the
order in which vertices joined them come from the oracle here; the GPU path
is checked the same way in test_gpu_writers.py."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O
import rtrex_cuda_amd as R
from conftest import GOLDEN, golden_graph, golden_meta

FILES = json.load(open(os.path.join(GOLDEN, "files.json")))
SUFFIXES = ("decomposition", "stats", "cut", "freq")


def sha_file(p) -> str:
    with open(p, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def check_files(tmp_path, tag, prefix, kind, n, suffixes=SUFFIXES):
    for suf in suffixes:
        exp = FILES[tag][f"{kind}-{suf}"]
        p = tmp_path / f"{prefix}-{suf}.txt"
        if len(exp) > 2:
            # The reference prints uninitialised memory on its "Size = n"
            # frequency line (RTRex.cpp:41-43); compare the other lines.
            with open(p, "rb") as f:
                lines = f.read().split(b"\n")
            keep = [ln for ln in lines if not ln.startswith(b"Size = %d," % n)]
            assert len(keep) == len(lines) - 1
            assert hashlib.sha256(b"\n".join(keep)).hexdigest() == exp[2], (tag, kind, suf)
            continue
        assert sha_file(p) == exp[0], (tag, kind, suf, os.path.getsize(p), exp[1])


@pytest.mark.parametrize("tag", sorted(FILES))
def test_writers_match_reference_driver(tag, tmp_path):
    name = tag.rsplit("_eps", 1)[0]
    eps = golden_meta()[tag]["eps"]
    n, u, v = golden_graph(name)
    off, nb = O.csr_from_undirected(n, u, v)
    d = O.decompose(off, nb, eps)
    R.write_decomposition(str(tmp_path / "RTRex-g"), off, nb, d.cluster_map, d.member_rank,
                          d.n_clusters)
    check_files(tmp_path, tag, "RTRex-g", "RTRex", n)
    R.write_components(str(tmp_path / "cc-g"), off, nb)
    check_files(tmp_path, tag, "cc-g", "connected-components", n)
    R.write_survivors(str(tmp_path / "surv.txt"), off, nb, d.status_after_clean)
    assert sha_file(tmp_path / "surv.txt") == FILES[tag]["survivors"][0]


@pytest.mark.parametrize("thr", [0, 2, 10])
@pytest.mark.parametrize("tag", sorted(FILES))
def test_expanded_match_reference_driver(tag, thr, tmp_path):
    """expandClusters (mode e) on the oracle's clusters: the expanded files
    and the scores file equal the reference driver's at each threshold."""
    name = tag.rsplit("_eps", 1)[0]
    eps = golden_meta()[tag]["eps"]
    n, u, v = golden_graph(name)
    off, nb = O.csr_from_undirected(n, u, v)
    d = O.decompose(off, nb, eps)
    eo, sc = R.write_expanded(str(tmp_path / "RTRex-expanded-g"), off, nb, d.cluster_map,
                              d.member_rank, d.n_clusters, thr)
    check_files(tmp_path, tag, "RTRex-expanded-g", f"RTRex-expanded-t{thr}", n,
                SUFFIXES + ("scores",))
    moved = sc >= 0
    assert np.all(eo[~moved] == d.cluster_map[~moved])
    assert np.all(eo[moved] != d.cluster_map[moved])


def test_writer_rejects_bad_ranks(tmp_path):
    n, u, v = golden_graph("small_test")
    off, nb = O.csr_from_undirected(n, u, v)
    d = O.decompose(off, nb, 0.3)
    bad = d.member_rank.copy()
    bad[:] = 0  # two members of a cluster at the same position
    with pytest.raises(R.RtrexError):
        R.write_decomposition(str(tmp_path / "x"), off, nb, d.cluster_map, bad, d.n_clusters)
