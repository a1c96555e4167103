"""Comparison of two decompositions, and canonical hashes of cluster sets.

Counterpart of the reference's tolerance report (RTRex_CUDA/validate.py), which
compares a decomposition file written by one implementation with another's:

  load_clusters      RTRex_CUDA/validate.py:21-35   one cluster per line
  cluster_stats      :37-54   totals, non-trivial clusters, size histogram
  jaccard            :76-84
  match_clusters     :86-106  greedy matching, Jaccard >= threshold
  js_divergence      :108-126 of the two size histograms (log base 2)
  report             :166-221 the clustered-vertex ratio bands (2 % / 5 %)
                              and the JS-divergence bands (0.05 / 0.10)

The matching is the reference's greedy rule: reference clusters in file order,
each takes the still-unmatched candidate of largest Jaccard similarity (the
lowest index on a tie, the order in which the reference iterates its set of
small integers), if that similarity is at least the threshold.  Candidates are
found through an inverted index (vertex -> cluster) instead of a scan of every
candidate: a candidate sharing no vertex has similarity 0, which never beats
the reference's starting best of 0, so the result is the same.

canonical_cluster_sha gives one SHA-256 per partition, independent of cluster
ids and of member order, so that large decompositions are compared by hash:
clusters ordered by their smallest vertex, each written as its size then its
vertices ascending, all as little-endian int64.

Command line (the reference's usage, with the two files named explicitly):

    python -m rtrex_cuda_amd.validate <reference-decomposition.txt> <other.txt>
"""
from __future__ import annotations

import hashlib
import math
import sys
from collections import Counter, defaultdict

import numpy as np


# ----------------------------------------------------------------- loading
def load_clusters(path: str) -> list[set]:
    """Clusters of a -decomposition.txt file, one set per non-blank line."""
    out = []
    with open(path) as f:
        for ln in f:
            ln = ln.strip()
            if ln:
                out.append(set(int(x) for x in ln.split()))
    return out


def clusters_from_map(cluster_map) -> list[set]:
    """Cluster sets from a cluster id per vertex (ids 0..k-1), in id order."""
    cm = np.asarray(cluster_map, np.int64)
    if cm.size == 0:
        return []
    order = np.argsort(cm, kind="stable")
    cuts = np.flatnonzero(np.diff(cm[order])) + 1
    return [set(int(x) for x in grp) for grp in np.split(order, cuts)]


# ----------------------------------------------------------- statistics
def cluster_stats(clusters) -> dict:
    sizes = [len(c) for c in clusters]
    nontrivial = [s for s in sizes if s > 1]
    return {
        "total_clusters": len(clusters),
        "nontrivial_clusters": len(nontrivial),
        "max_size": max(sizes) if sizes else 0,
        "clustered_vertices": sum(nontrivial),
        "total_vertices": sum(sizes),
        "size_distribution": dict(sorted(Counter(sizes).items())),
    }


def jaccard(a: set, b: set) -> float:
    if not a and not b:
        return 1.0
    if not a or not b:
        return 0.0
    inter = len(a & b)
    return inter / (len(a) + len(b) - inter)


def match_clusters(ref, other, threshold: float = 0.5) -> list[tuple]:
    """Greedy matching: list of (ref index, other index, similarity)."""
    owner = {}
    for j, c in enumerate(other):
        for x in c:
            owner.setdefault(x, []).append(j)
    taken = set()
    out = []
    for i, c in enumerate(ref):
        inter = defaultdict(int)
        for x in c:
            for j in owner.get(x, ()):
                if j not in taken:
                    inter[j] += 1
        best, best_j = 0.0, -1
        for j in sorted(inter):
            s = inter[j] / (len(c) + len(other[j]) - inter[j])
            if s > best:
                best, best_j = s, j
        if best >= threshold:
            out.append((i, best_j, best))
            taken.add(best_j)
    return out


def js_divergence(h1: dict, h2: dict) -> float:
    t1, t2 = sum(h1.values()), sum(h2.values())
    jsd = 0.0
    for k in set(h1) | set(h2):
        p = h1.get(k, 0) / t1 if t1 > 0 else 0.0
        q = h2.get(k, 0) / t2 if t2 > 0 else 0.0
        m = (p + q) / 2
        if p > 0 and m > 0:
            jsd += p * math.log2(p / m)
        if q > 0 and m > 0:
            jsd += q * math.log2(q / m)
    return jsd / 2


def report(ref, other, threshold: float = 0.5) -> dict:
    """The reference's comparison as a dict: statistics of both, relative
    deltas, JS divergence of the size histograms, greedy matches and the
    tolerance bands."""
    a, b = cluster_stats(ref), cluster_stats(other)
    keys = ("total_clusters", "nontrivial_clusters", "max_size", "clustered_vertices",
            "total_vertices")
    delta_pct = {k: ((b[k] - a[k]) / a[k] * 100.0) if a[k] else 0.0 for k in keys}
    jsd = js_divergence(a["size_distribution"], b["size_distribution"])
    matches = match_clusters(ref, other, threshold)
    out = dict(ref=a, other=b, delta_pct=delta_pct, jsd=jsd, matches=len(matches),
               ref_matched=len({m[0] for m in matches}),
               other_matched=len({m[1] for m in matches}),
               mean_jaccard=(sum(m[2] for m in matches) / len(matches)) if matches else None)
    if a["clustered_vertices"] > 0:
        r = b["clustered_vertices"] / a["clustered_vertices"]
        out["clustered_ratio"] = r
        out["clustered_band"] = ("within 2%" if r > 0.98 else
                                 "within 5%" if r > 0.95 else "outside 5%")
    out["jsd_band"] = ("excellent" if jsd < 0.05 else "acceptable" if jsd < 0.10
                       else "diverged")
    return out


# ------------------------------------------------------- canonical hashing
def canonical_cluster_sha(cluster_map) -> str:
    """SHA-256 of a partition given as a cluster id per vertex."""
    cm = np.asarray(cluster_map, np.int64)
    if cm.size == 0:
        return hashlib.sha256(b"").hexdigest()
    order = np.lexsort((np.arange(cm.size), cm))  # by cluster, members ascending
    cuts = np.flatnonzero(np.diff(cm[order])) + 1
    starts = np.concatenate([[0], cuts])
    sizes = np.diff(np.concatenate([starts, [cm.size]]))
    rank = np.empty(len(starts), np.int64)  # cluster -> position by smallest vertex
    rank[np.argsort(order[starts], kind="stable")] = np.arange(len(starts))
    head = np.empty(len(starts), np.int64)  # where each cluster's record starts
    sz_by_rank = np.empty(len(starts), np.int64)
    sz_by_rank[rank] = sizes
    head_by_rank = np.concatenate([[0], np.cumsum(sz_by_rank + 1)[:-1]])
    head = head_by_rank[rank]
    flat = np.empty(cm.size + len(starts), np.int64)
    flat[head] = sizes
    cl = np.repeat(np.arange(len(starts)), sizes)  # cluster of each sorted member
    within = np.arange(cm.size) - np.repeat(starts, sizes)
    flat[head[cl] + 1 + within] = order
    return hashlib.sha256(flat.tobytes()).hexdigest()


def canonical_sets_sha(clusters) -> str:
    """The same hash for a list of vertex sets (a decomposition file)."""
    cs = sorted((sorted(c) for c in clusters if c), key=lambda c: c[0])
    flat = []
    for c in cs:
        flat.append(len(c))
        flat.extend(c)
    return hashlib.sha256(np.asarray(flat, np.int64).tobytes()).hexdigest()


def main(argv) -> int:
    if len(argv) < 3:
        print("usage: python -m rtrex_cuda_amd.validate <reference.txt> <other.txt>")
        return 1
    ref, other = load_clusters(argv[1]), load_clusters(argv[2])
    r = report(ref, other)
    print(f"{'Metric':<30} {'reference':>12} {'other':>12} {'Delta':>12}")
    for label, k in (("Total clusters", "total_clusters"),
                     ("Non-trivial clusters", "nontrivial_clusters"),
                     ("Max cluster size", "max_size"),
                     ("Clustered vertices", "clustered_vertices"),
                     ("Total vertices", "total_vertices")):
        print(f"{label:<30} {r['ref'][k]:>12} {r['other'][k]:>12} {r['delta_pct'][k]:>+10.2f}%")
    print(f"Size distribution JS-divergence: {r['jsd']:.6f} ({r['jsd_band']})")
    print(f"Matched (Jaccard >= 0.5): {r['ref_matched']}/{r['ref']['nontrivial_clusters']} "
          f"reference, {r['other_matched']}/{r['other']['nontrivial_clusters']} other")
    if r["mean_jaccard"] is not None:
        print(f"Average Jaccard similarity: {r['mean_jaccard']:.4f}")
    if "clustered_ratio" in r:
        print(f"Clustered vertex ratio: {r['clustered_ratio']:.4f} ({r['clustered_band']})")
    same = canonical_sets_sha(ref) == canonical_sets_sha(other)
    print(f"Identical partitions: {same}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
