"""The output files of the GPU path -- clusters and member order from the
device extraction, written by rtrex_write_* and by the rtrex command-line
driver -- against the files the reference's own driver writes
(tests/golden/files.json)."""
from __future__ import annotations

import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import rtrex_cuda_amd as R
from conftest import GOLDEN, golden_graph, golden_meta
from test_writers import FILES, check_files, sha_file

pytestmark = pytest.mark.gpu

# Every golden case is decided exactly (see test_gpu_parity).
ORDER_DEPENDENT: set = set()
CLI = os.path.join(os.path.dirname(R.HIP_LIB), "rtrex")


@pytest.mark.parametrize("tag", [t for t in sorted(FILES) if t not in ORDER_DEPENDENT])
def test_gpu_files_match_reference_driver(tag, tmp_path):
    name = tag.rsplit("_eps", 1)[0]
    eps = golden_meta()[tag]["eps"]
    n, u, v = golden_graph(name)
    print(f"[gpu] files {tag}", flush=True)
    with R.DeviceGraph.from_edges(n, u, v) as g:
        off, nb = g.csr()
        g.gpu_common_nbr()
        g.gpu_initial_clean(eps)
        st = g.sync_edge_status()
        cm, k = g.disjoint_extract(eps)
        rank = g.member_order()
    R.write_decomposition(str(tmp_path / "RTRex-g"), off, nb, cm, rank, k)
    check_files(tmp_path, tag, "RTRex-g", "RTRex", n)
    R.write_survivors(str(tmp_path / "surv.txt"), off, nb, st)
    assert sha_file(tmp_path / "surv.txt") == FILES[tag]["survivors"][0]


@pytest.mark.parametrize("tag", ["small_test_eps0.3", "clean_astro_eps0.3", "config1_eps0.1"])
def test_cli_matches_reference_driver(tag, tmp_path):
    name = tag.rsplit("_eps", 1)[0]
    eps = golden_meta()[tag]["eps"]
    n, u, v = golden_graph(name)
    ef = tmp_path / "g.txt"
    O.write_escape(str(ef), n, u, v)
    print(f"[gpu] cli {tag}", flush=True)
    r = subprocess.run([CLI, str(ef), "g", str(eps), "s"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    check_files(tmp_path, tag, "RTRex-g", "RTRex", n)
    check_files(tmp_path, tag, "connected-components-g", "connected-components", n)
    assert sha_file(tmp_path / "amazon_cleaned_edges.txt") == FILES[tag]["survivors"][0]


@pytest.mark.parametrize("tag", ["clean_email_eps0.3", "clean_astro_eps0.1"])
def test_cli_mode_e_matches_reference_driver(tag, tmp_path):
    """The driver's mode e (expandClusters at threshold 2): the RTRex,
    RTRex-expanded (with scores) and connected-components files."""
    name = tag.rsplit("_eps", 1)[0]
    eps = golden_meta()[tag]["eps"]
    n, u, v = golden_graph(name)
    ef = tmp_path / "g.txt"
    O.write_escape(str(ef), n, u, v)
    print(f"[gpu] cli mode e {tag}", flush=True)
    r = subprocess.run([CLI, str(ef), "g", str(eps), "e", "2"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    check_files(tmp_path, tag, "RTRex-g", "RTRex", n)
    check_files(tmp_path, tag, "RTRex-expanded-g", "RTRex-expanded-t2", n,
                ("decomposition", "stats", "cut", "freq", "scores"))
    check_files(tmp_path, tag, "connected-components-g", "connected-components", n)
