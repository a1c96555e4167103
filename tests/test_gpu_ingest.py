"""Edge-list text parsed on the device (rtrex_graph_load, rtrex_snap_convert).

Escape (loadGraph_Escape, GraphIO.cpp:19-88): the graph built from a file
equals the reference's CSR (golden hash) whatever comment lines, blank lines,
CRLF line ends and extra columns (weights) the file holds; the loader's
errors are reported as the host reader reports them.

SNAP (RTRex_CUDA/snap_to_rtr.py:47-142): files made from a known graph by
relabelling its vertices (ids with gaps, 1-based, negative, permuted) and
adding what the conversion must drop -- comments ('#', '%', indented), blank
lines, lines with one token or a non-integer token, self-loops, duplicate and
reversed edges -- convert to the graph relabelled by sorted id order, as
computed here independently with numpy (np.unique ranks), and the Escape file
written is byte for byte that graph's."""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

import oracle as O
import rtrex_cuda_amd as R
from conftest import golden_graph, golden_meta

pytestmark = pytest.mark.gpu


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _escape_text(n, u, v, noise: bool, rng) -> str:
    lines = []
    if noise:
        lines += ["# a comment line", "", "   \t"]
    lines.append(f"{n} {len(u)}")
    for i, (a, b) in enumerate(zip(u.tolist(), v.tolist())):
        if noise and i % 97 == 0:
            lines.append("# comment between edges")
        if noise and i % 89 == 0:
            lines.append("")
        extra = f" {rng.random():.4f}" if noise and i % 3 == 0 else ""
        lines.append(f"{a} {b}{extra}")
    sep = "\r\n" if noise else "\n"
    return sep.join(lines) + sep


@pytest.mark.parametrize("name", ["small_test", "clean_email", "config1"])
@pytest.mark.parametrize("noise", [False, True])
def test_escape_load_equals_reference_csr(name, noise, tmp_path):
    n, u, v = golden_graph(name)
    p = tmp_path / "g.txt"
    p.write_text(_escape_text(n, u, v, noise, np.random.default_rng(1)))
    print(f"[gpu] ingest escape {name} noise={noise}", flush=True)
    with R.DeviceGraph.load(str(p), "escape") as g:
        off, nb = g.csr()
        total = g.gpu_common_nbr()
    meta = golden_meta()[f"{name}_eps0.3"]
    assert sha(np.concatenate([off, nb])) == meta["csr_sha"]
    assert total == meta["count_total"]


@pytest.mark.parametrize("text,status", [
    ("3 3\n0 1\n1 2\n", 4),                 # fewer edge lines than the header says
    ("3 2\n0 1\n1 x\n", 4),                 # an edge line without two integers
    ("# only a comment\n\n", 4),           # no header
    ("3 2\n0 1\n1 3\n", 1),                 # vertex id >= n
    ("3 2\n0 1\n-1 2\n", 1),                # negative id
    ("3 2\n0 1\n1 0\n", 1),                 # the same edge twice
    ("3 2\n0 0\n1 2\n", 1),                 # self-loop
])
def test_escape_load_errors(text, status, tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(R.RtrexError) as e:
        R.DeviceGraph.load(str(p), "escape")
    assert e.value.status == status


def test_escape_load_ignores_lines_beyond_m(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("4 2\n0 1\n2 3\n1 2\n")
    with R.DeviceGraph.load(str(p), "escape") as g:
        off, nb = g.csr()
    assert off.tolist() == [0, 1, 2, 3, 4] and nb.tolist() == [1, 0, 3, 2]


def _expected_snap(a, b):
    """The conversion's result computed with numpy: self-loops dropped, ids
    ranked among the distinct ids, undirected pairs once, sorted."""
    keep = a != b
    a, b = a[keep], b[keep]
    ids, inv = np.unique(np.concatenate([a, b]), return_inverse=True)
    ra, rb = inv[: len(a)], inv[len(a):]
    lo, hi = np.minimum(ra, rb), np.maximum(ra, rb)
    pairs = np.unique(np.stack([lo, hi], 1), axis=0)
    return len(ids), pairs


def _snap_text(a, b, rng) -> str:
    out = ["# Directed graph (each unordered pair of nodes is saved once)",
           "# Nodes: 0 Edges: 0", "% a percent comment", "   # indented comment", ""]
    for i, (x, y) in enumerate(zip(a.tolist(), b.tolist())):
        r = rng.random()
        sx = f"{x:_}" if (i % 50 == 0 and x >= 1000) else (f"+{x}" if i % 61 == 0 and x > 0 else str(x))
        sep = "\t" if i % 7 == 0 else " "
        out.append(f"{sx}{sep}{y}" + (" 0.5 extra" if r < 0.1 else ""))
        if r > 0.97:
            out.append(f"{y} {x}")            # reversed duplicate
        if 0.93 < r <= 0.95:
            out.append(f"{x} {y}")            # exact duplicate
        if i % 113 == 0:
            out += [f"{x} {x}", "7", "12 abc", "  "]  # self-loop, one token, bad token, blank
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("relabel", ["gaps_monotone", "one_based", "permuted_negative"])
def test_snap_convert_matches_rules(relabel, tmp_path):
    n, u, v = golden_graph("clean_astro")
    u = u.astype(np.int64)
    v = v.astype(np.int64)
    rng = np.random.default_rng(7)
    if relabel == "gaps_monotone":
        lab = 3 * np.arange(n, dtype=np.int64) + 7
    elif relabel == "one_based":
        lab = np.arange(n, dtype=np.int64) + 1
    else:
        lab = rng.permutation(n).astype(np.int64) * 5 - 40_000
    a, b = lab[u], lab[v]
    p = tmp_path / "g.snap"
    p.write_text(_snap_text(a, b, rng))
    nv, pairs = _expected_snap(a, b)
    out = tmp_path / "g.rtr"
    print(f"[gpu] ingest snap {relabel}", flush=True)
    n2, m2 = R.snap_convert(str(p), str(out))
    assert (n2, m2) == (nv, len(pairs))
    want = f"{nv} {len(pairs)}\n" + "".join(f"{x} {y}\n" for x, y in pairs.tolist())
    assert out.read_text() == want
    # the graph built from the SNAP file directly is the converted graph's
    off0, nb0 = O.csr_from_undirected(nv, pairs[:, 0], pairs[:, 1])
    with R.DeviceGraph.load(str(p), "snap") as g:
        off, nb = g.csr()
    assert np.array_equal(off, off0) and np.array_equal(nb, nb0)
    if relabel != "permuted_negative":  # an order-preserving relabelling comes back
        deg = np.bincount(np.concatenate([u, v]), minlength=n)
        assert (deg > 0).all()
        assert np.array_equal(pairs, np.unique(np.stack([np.minimum(u, v), np.maximum(u, v)], 1),
                                               axis=0))


def test_snap_refuses_ids_beyond_64_bits(tmp_path):
    p = tmp_path / "big.snap"
    p.write_text("1 2\n99999999999999999999 3\n")
    with pytest.raises(R.RtrexError) as e:
        R.snap_convert(str(p), str(tmp_path / "o.txt"))
    assert e.value.status == 1
