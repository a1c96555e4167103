"""Writes tests/golden/serve_mix.json: hashes of the REFERENCE's own outputs on
one graph of each kind of the serving mix (seed 42, rtrex_cuda_amd.graphs):

  cl       serve_graph("cl", 0)       Chung-Lu, m in [1k, 50k]
  planted  serve_graph("planted", 3)  planted communities, 1.42 M edges
  rmat     serve_graph("rmat", 0)     R-MAT scale 18, 3.8 M edges

Run in the build container (needs oracle/_ref/ref_dump):

    python tests/golden/make_golden_serve.py <workdir>

For each graph and eps 0.3 it runs ref_dump (the reference's commonNbr,
initialClean and disjointExtract compiled from its sources) and records the
fields make_golden_c3.py describes (CSR, count, supports, survivors, clusters
as SHA-256).  The graphs are regenerated from their seeds where needed.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import make_golden_c3 as M  # noqa: E402  (shared helpers; sets up sys.path)
from rtrex_cuda_amd import graphs as G  # noqa: E402
import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "serve_mix.json")
GRAPHS = (("cl", 0), ("planted", 3), ("rmat", 0))
EPS = 0.3


def main(argv):
    work = os.path.abspath(argv[1])
    os.makedirs(work, exist_ok=True)
    meta = {}
    for kind, i in GRAPHS:
        n, u, v = G.serve_graph(kind, i)
        txt = os.path.join(work, f"{kind}{i}.txt")
        M.write_escape_fast(txt, n, u, v)
        d = os.path.join(work, f"{kind}{i}_eps{EPS}")
        os.makedirs(d, exist_ok=True)
        subprocess.run([O.REF_DUMP, txt, repr(EPS), d, "all"], cwd=d, check=True)
        meta[f"{kind}{i}_eps{EPS}"] = dict(M.summarise(d, EPS), kind=kind, index=i)
        print(kind, i, json.dumps(meta[f"{kind}{i}_eps{EPS}"]), flush=True)
    with open(OUT, "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv)
