"""Writes the golden fixtures of tests/golden from the REFERENCE itself.

Run in the build container (needs /root/reference and oracle/_ref/ref_dump,
built by ``make -C oracle``):

    python tests/golden/make_golden.py

For every input it runs the reference's own phases (oracle/_ref/ref_dump:
commonNbr, initialClean, disjointExtract compiled from the reference sources)
and stores:
  graphs/<name>.npz        the input edge list (u, v as listed in the file, n)
  <name>_eps<e>.npz        triangle totals, survivors (canonical u<v pairs),
                           cluster map, per-cluster (size, edges, cut)
  golden.json              SHA-256 of the raw FP64 perEdge / perVertex bytes
                           after count and after the initial clean, plus the
                           reference's phase timings
Inputs: the four graphs shipped with the reference (RTRex_Sequential/Example
and clustering/), cliques K_5..K_9, and the seeded C1 workload
(rtrex_cuda_amd.graphs.config1).  Nothing of the reference's source is
copied; the shipped graph files are data and are stored as edge arrays.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

REF = "/root/reference/RTRex_Sequential"
BUNDLED = {
    "small_test": f"{REF}/Example/small-test.txt",
    "random_graph_rtr": f"{REF}/clustering/random_graph_rtr.txt",
    "clean_email": f"{REF}/clustering/clean_email.txt",
    "clean_astro": f"{REF}/clustering/clean_astro.txt",
}
EPS = (0.3, 0.1)


def read_escape(path):
    rows, n = [], None
    with open(path) as f:
        for ln in f:
            if ln.startswith("#") or not ln.strip():
                continue
            a, b = ln.split()[:2]
            if n is None:
                n, m = int(a), int(b)
                continue
            rows.append((int(a), int(b)))
    e = np.array(rows, np.int64).reshape(-1, 2)
    return n, e[:, 0].astype(np.uint32), e[:, 1].astype(np.uint32)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def survivors(off, nb, status):
    src = np.repeat(np.arange(len(off) - 1), np.diff(off))
    keep = (status[: len(nb)] == ord("Y")) & (nb > src)
    return np.stack([src[keep], nb[keep]], 1).astype(np.uint32)


def main():
    os.makedirs(os.path.join(HERE, "graphs"), exist_ok=True)
    inputs = dict(BUNDLED)
    tmp = tempfile.mkdtemp()
    for k in range(5, 10):
        iu = np.array([(a, b) for a in range(k) for b in range(a + 1, k)], np.uint32)
        p = os.path.join(tmp, f"k{k}.txt")
        O.write_escape(p, k, iu[:, 0], iu[:, 1])
        inputs[f"clique_k{k}"] = p
    from rtrex_cuda_amd import graphs as G
    n1, u1, v1 = G.config1()
    p = os.path.join(tmp, "c1.txt")
    O.write_escape(p, n1, u1, v1)
    inputs["config1"] = p

    meta = {}
    for name, path in inputs.items():
        n, u, v = read_escape(path)
        np.savez_compressed(os.path.join(HERE, "graphs", f"{name}.npz"), n=n, u=u, v=v)
        for eps in EPS:
            r = O.ref_dump(path, eps)
            off, nb = r["offsets"], r["nbors"]
            cl = r["clusters"]
            sizes = np.array([c[0] for c in cl], np.int64)
            stats = np.array([[c[0], c[1], c[2]] for c in cl], np.int64).reshape(-1, 3)
            tag = f"{name}_eps{eps}"
            np.savez_compressed(
                os.path.join(HERE, f"{tag}.npz"),
                count_total=r["count_total"], clean_total=r["clean_total"],
                survivors=survivors(off, nb, r["clean_status"]),
                cluster_map=r["cluster_map"].astype(np.int32), stats=stats)
            meta[tag] = dict(
                n=int(n), m=int(len(u)), eps=eps,
                count_total=r["count_total"], clean_total=r["clean_total"],
                n_clusters=len(cl), largest_cluster=int(sizes.max()) if len(sizes) else 0,
                csr_sha=sha(np.concatenate([off, nb])),
                count_perEdge_sha=sha(r["count_perEdge"]),
                count_perVertex_sha=sha(r["count_perVertex"]),
                clean_perEdge_sha=sha(r["clean_perEdge"]),
                clean_perVertex_sha=sha(r["clean_perVertex"]),
                ref_times=r["times"])
            print(tag, meta[tag]["count_total"], len(cl), flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
