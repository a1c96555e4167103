"""Writes tests/golden/files.json: SHA-256 and size of every output file the
REFERENCE's own driver writes (oracle/_ref/RTRex, built from
RTRex_Sequential/clustering/RTRex.cpp by ``make -C oracle``), per input and eps.

    python tests/golden/make_golden_files.py

The driver (RTRex.cpp:141-206, mode s) writes, in its working directory:
  RTRex-<name>-{decomposition,stats,cut,freq}.txt            printDecomposition
  connected-components-<name>-{decomposition,stats,cut,freq}.txt
  amazon_cleaned_edges.txt                                   survivor dump
In mode e (expandClusters) at thresholds 0, 2 and 10 it also records the
RTRex-expanded-g-{decomposition,stats,cut,freq,scores}.txt files as
"RTRex-expanded-t<threshold>-<suffix>".
Beside the file hashes it records the canonical partition hash
(rtrex_cuda_amd.validate.canonical_sets_sha) of both decomposition files, so
that a partition computed elsewhere (cluster ids and member order aside) is
compared with the reference's: "RTRex-sets" and "connected-components-sets".
The inputs are the graphs of tests/golden/graphs (the reference's bundled
graphs, cliques K_5..K_9 and the C1 workload), written as Escape files.
"""
from __future__ import annotations

import glob
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from rtrex_cuda_amd import validate as V  # noqa: E402

RTREX = os.path.join(ROOT, "oracle", "_ref", "RTRex")
EPS = (0.3, 0.1)
SUFFIXES = ("decomposition", "stats", "cut", "freq")
THRESHOLDS = (0, 2, 10)


def sha_file(p: str) -> str:
    with open(p, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def main():
    if not os.path.exists(RTREX):
        sys.exit("oracle/_ref/RTRex is not built (make -C oracle)")
    out = {}
    for gp in sorted(glob.glob(os.path.join(HERE, "graphs", "*.npz"))):
        name = os.path.basename(gp)[:-4]
        z = np.load(gp)
        for eps in EPS:
            with tempfile.TemporaryDirectory() as d:
                ef = os.path.join(d, "g.txt")
                O.write_escape(ef, int(z["n"]), z["u"], z["v"])
                subprocess.run([RTREX, ef, "g", str(eps), "s"], cwd=d, check=True,
                               stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
                files = {}
                n = int(z["n"])
                for pre in ("RTRex-g", "connected-components-g"):
                    for suf in SUFFIXES:
                        p = os.path.join(d, f"{pre}-{suf}.txt")
                        ent = [sha_file(p), os.path.getsize(p)]
                        if suf == "freq":
                            # printDecomposition zeroes freq[0..n-1] only
                            # (RTRex.cpp:41-43): with a cluster of all n
                            # vertices its "Size = n" line is uninitialised
                            # memory.  Also keep the hash without that line.
                            with open(p, "rb") as f:
                                lines = f.read().split(b"\n")
                            keep = [ln for ln in lines if not ln.startswith(b"Size = %d," % n)]
                            if len(keep) != len(lines):
                                ent.append(hashlib.sha256(b"\n".join(keep)).hexdigest())
                        files[f"{pre.split('-g')[0]}-{suf}"] = ent
                    sets = V.load_clusters(os.path.join(d, f"{pre}-decomposition.txt"))
                    files[f"{pre.split('-g')[0]}-sets"] = V.canonical_sets_sha(sets)
                p = os.path.join(d, "amazon_cleaned_edges.txt")
                files["survivors"] = [sha_file(p), os.path.getsize(p)]
                # mode e: expandClusters at several thresholds (RTRex.cpp:166-201)
                for thr in THRESHOLDS:
                    subprocess.run([RTREX, ef, "g", str(eps), "e", str(thr)], cwd=d, check=True,
                                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
                    for suf in SUFFIXES + ("scores",):
                        p = os.path.join(d, f"RTRex-expanded-g-{suf}.txt")
                        ent = [sha_file(p), os.path.getsize(p)]
                        if suf == "freq":
                            with open(p, "rb") as f:
                                lines = f.read().split(b"\n")
                            keep = [ln for ln in lines if not ln.startswith(b"Size = %d," % n)]
                            if len(keep) != len(lines):
                                ent.append(hashlib.sha256(b"\n".join(keep)).hexdigest())
                        files[f"RTRex-expanded-t{thr}-{suf}"] = ent
            out[f"{name}_eps{eps}"] = files
            print(name, eps, flush=True)
    with open(os.path.join(HERE, "files.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
