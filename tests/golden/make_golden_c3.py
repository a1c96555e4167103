"""Writes tests/golden/config3.json: hashes of the REFERENCE's own outputs on
the skewed large workload (R-MAT scale 22, edge factor 16, a,b,c = 0.57, 0.19,
0.19, seed 11; 4,194,304 vertices, about 64.2 M undirected edges).

Run in the build container (needs /root/reference and oracle/_ref/ref_dump,
built by ``make -C oracle``).  The reference is single-threaded; one eps takes
hours, so the raw outputs are kept in a work directory and the hashing step can
be re-run on them:

    python tests/golden/make_golden_c3.py <workdir> 0.3 [0.02 ...]

For each eps it runs ref_dump (commonNbr, initialClean, disjointExtract
compiled from the reference sources) into <workdir>/eps<e>/ unless that
directory already holds a finished run, then records in config3.json:
  csr_sha              SHA-256 of offsets ++ nbors (int64, makeCSR layout)
  count_total          triangles before the clean (int)
  count_perEdge_sha    SHA-256 of the FP64 perEdge bytes after commonNbr
  count_perVertex_sha  the same for perVertex
  clean_total          triangles left after initialClean
  survivors            surviving undirected edges after initialClean
  survivors_sha        SHA-256 of the surviving (u, v) pairs, u < v, sorted,
                       as uint32 (the reference's survivor dump, canonical)
  n_clusters           clusters made by disjointExtract
  cluster_map_sha      SHA-256 of cluster_map as int64 (ids in creation order)
  cluster_sets_sha     rtrex_cuda_amd.validate.canonical_cluster_sha of it
  ref_times            the reference's phase wall times (seconds)
The graph itself is regenerated from its seed wherever it is needed
(rtrex_cuda_amd.graphs.config3), so only hashes are committed.
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from rtrex_cuda_amd import graphs as G  # noqa: E402
from rtrex_cuda_amd import validate as V  # noqa: E402

OUT = os.path.join(HERE, "config3.json")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def survivors(off, nb, status):
    src = np.repeat(np.arange(len(off) - 1, dtype=np.int64), np.diff(off))
    keep = (status[: len(nb)] == ord("Y")) & (nb > src)
    return np.stack([src[keep], nb[keep]], 1).astype(np.uint32)


def write_escape_fast(path: str, n: int, u: np.ndarray, v: np.ndarray) -> None:
    """Escape text, 'n m' then 'u v' per line, written in chunks."""
    with open(path, "w") as f:
        f.write(f"{n} {len(u)}\n")
        step = 1 << 22
        for s in range(0, len(u), step):
            a = np.stack([u[s:s + step], v[s:s + step]], 1).astype(np.int64)
            np.savetxt(f, a, fmt="%d")


def run(workdir: str, eps: float) -> str:
    d = os.path.abspath(os.path.join(workdir, f"eps{eps}"))
    if os.path.exists(os.path.join(d, "times.txt")):
        return d
    os.makedirs(d, exist_ok=True)
    txt = os.path.join(workdir, "config3.txt")
    if not os.path.exists(txt):
        n, u, v = G.config3()
        write_escape_fast(txt + ".tmp", n, u, v)
        os.replace(txt + ".tmp", txt)
    t0 = time.time()
    print(f"ref_dump eps {eps} started", flush=True)
    subprocess.run([O.REF_DUMP, os.path.abspath(txt), repr(eps), d, "all"], cwd=d,
                   check=True)
    print(f"ref_dump eps {eps} done in {time.time() - t0:.0f} s", flush=True)
    return d


def summarise(d: str, eps: float) -> dict:
    rd = lambda name, dt: np.fromfile(os.path.join(d, name), dtype=dt)
    off, nb = rd("csr_offsets.i64", np.int64), rd("csr_nbors.i64", np.int64)
    cm = rd("cluster_map.i64", np.int64)
    with open(os.path.join(d, "times.txt")) as f:
        times = {k: float(x) for k, x in (ln.split() for ln in f)}
    surv = survivors(off, nb, rd("clean_status.u8", np.uint8))
    return dict(
        n=int(len(off) - 1), m=int(len(nb) // 2), eps=eps,
        csr_sha=sha(np.concatenate([off, nb])),
        count_total=int(rd("count_total.i64", np.int64)[0]),
        count_perEdge_sha=sha(rd("count_perEdge.f64", np.float64)),
        count_perVertex_sha=sha(rd("count_perVertex.f64", np.float64)),
        clean_total=int(rd("clean_total.i64", np.int64)[0]),
        survivors=int(len(surv)), survivors_sha=sha(surv),
        n_clusters=int(cm.max()) + 1 if len(cm) else 0,
        cluster_map_sha=sha(cm),
        cluster_sets_sha=V.canonical_cluster_sha(cm),
        ref_times=times)


def main(argv):
    workdir = argv[1]
    eps_list = [float(x) for x in argv[2:]] or [0.3]
    meta = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            meta = json.load(f)
    for eps in eps_list:
        d = run(workdir, eps)
        meta[f"config3_eps{eps}"] = summarise(d, eps)
        with open(OUT, "w") as f:
            json.dump(meta, f, indent=1, sort_keys=True)
        print(json.dumps(meta[f"config3_eps{eps}"]), flush=True)


if __name__ == "__main__":
    main(sys.argv)
