"""CPU oracle for the RTRExtractor decomposition path -- TEST INFRASTRUCTURE ONLY.

This package wraps ``liboracle.so`` (a project-written C++ restatement of the
reference's sequential algorithm, ``rtrex_oracle.cpp``) with ctypes.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import it, and only as the checker.  The product (``rtrex_cuda_amd``) never
imports it.

``ref_dump`` helpers run ``oracle/_ref/ref_dump`` (built from the reference
sources by ``oracle/Makefile``) in a scratch directory and read its raw outputs.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile
import time
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DUMP = os.path.join(HERE, "_ref", "ref_dump")

_lib = None


def build() -> None:
    """Compile liboracle.so (and _ref/ref_dump when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        p64, pd, pu8 = C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_uint8)
        L.oracle_make_csr.argtypes = [C.c_int64, C.c_int64, p64, p64, p64, p64]
        L.oracle_count.argtypes = [C.c_int64, p64, p64, p64, pd, pd]
        L.oracle_initial_clean.argtypes = [C.c_int64, p64, p64, C.c_double, pd, pd, p64, pu8]
        L.oracle_delete_and_clean.argtypes = [C.c_int64, p64, p64, C.c_double, p64,
                                              C.c_int64, pd, pd, p64, pu8]
        L.oracle_decompose.argtypes = [C.c_int64, p64, p64, C.c_double, p64, p64, pu8,
                                       p64, C.c_int64, p64, pd, p64]
        L.oracle_connected_components.argtypes = [C.c_int64, p64, p64, p64, p64]
        L.oracle_extract_state.argtypes = [C.c_int64, p64, p64, C.c_double, pd, pd, p64, pu8,
                                           p64, p64]
        L.oracle_extract_inner.argtypes = [C.c_int64, p64, p64, pu8, C.c_int64, C.c_double, p64,
                                           pd, C.c_int64, p64, pd, pd]
        L.oracle_cluster_edges.argtypes = [C.c_int64, p64, p64, pu8, p64, C.c_int64, p64,
                                           C.c_int64, p64]
        for f in (L.oracle_make_csr, L.oracle_count, L.oracle_initial_clean,
                  L.oracle_delete_and_clean, L.oracle_decompose,
                  L.oracle_connected_components, L.oracle_extract_state,
                  L.oracle_extract_inner, L.oracle_cluster_edges):
            f.restype = C.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(C.POINTER(t))


def make_csr(n: int, src: np.ndarray, dst: np.ndarray):
    """Directed COO (both directions present) -> (offsets, nbors), int64."""
    src = np.ascontiguousarray(src, dtype=np.int64)
    dst = np.ascontiguousarray(dst, dtype=np.int64)
    off = np.zeros(n + 1, np.int64)
    nb = np.zeros(len(src), np.int64)
    rc = lib().oracle_make_csr(n, len(src), _p(src, C.c_int64), _p(dst, C.c_int64),
                               _p(off, C.c_int64), _p(nb, C.c_int64))
    if rc:
        raise ValueError(f"oracle_make_csr failed rc={rc}")
    return off, nb


def csr_from_undirected(n: int, u: np.ndarray, v: np.ndarray):
    """Undirected edge list (each edge once) -> CSR, the way the Escape loader
    emits both directions (GraphIO.cpp:54-62) and makeCSR sorts them."""
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    return make_csr(n, np.concatenate([u, v]), np.concatenate([v, u]))


@dataclass
class Count:
    total: int
    perEdge: np.ndarray
    perVertex: np.ndarray


def count(off: np.ndarray, nb: np.ndarray) -> Count:
    n = len(off) - 1
    off = np.ascontiguousarray(off, np.int64)
    nb = np.ascontiguousarray(nb, np.int64)
    pe = np.zeros(len(nb), np.float64)
    pv = np.zeros(n, np.float64)
    tot = np.zeros(1, np.int64)
    rc = lib().oracle_count(n, _p(off, C.c_int64), _p(nb, C.c_int64), _p(tot, C.c_int64),
                            _p(pe, C.c_double), _p(pv, C.c_double))
    if rc:
        raise ValueError(f"oracle_count failed rc={rc}")
    return Count(int(tot[0]), pe, pv)


@dataclass
class Clean:
    total: int
    perEdge: np.ndarray
    perVertex: np.ndarray
    status: np.ndarray  # uint8 'Y'/'N' per slot


def initial_clean(off, nb, eps: float, cnt: Count | None = None) -> Clean:
    n = len(off) - 1
    off = np.ascontiguousarray(off, np.int64)
    nb = np.ascontiguousarray(nb, np.int64)
    cnt = cnt or count(off, nb)
    pe = cnt.perEdge.copy()
    pv = cnt.perVertex.copy()
    tot = np.array([cnt.total], np.int64)
    st = np.zeros(len(nb), np.uint8)
    rc = lib().oracle_initial_clean(n, _p(off, C.c_int64), _p(nb, C.c_int64), eps,
                                    _p(pe, C.c_double), _p(pv, C.c_double),
                                    _p(tot, C.c_int64), _p(st, C.c_uint8))
    if rc:
        raise ValueError(f"oracle_initial_clean failed rc={rc}")
    return Clean(int(tot[0]), pe, pv, st)


def delete_and_clean(off, nb, eps: float, pairs: np.ndarray, state: Clean) -> Clean:
    n = len(off) - 1
    off = np.ascontiguousarray(off, np.int64)
    nb = np.ascontiguousarray(nb, np.int64)
    pairs = np.ascontiguousarray(pairs, np.int64).reshape(-1)
    pe = state.perEdge.copy()
    pv = state.perVertex.copy()
    st = state.status.copy()
    tot = np.array([state.total], np.int64)
    rc = lib().oracle_delete_and_clean(n, _p(off, C.c_int64), _p(nb, C.c_int64), eps,
                                       _p(pairs, C.c_int64), len(pairs) // 2,
                                       _p(pe, C.c_double), _p(pv, C.c_double),
                                       _p(tot, C.c_int64), _p(st, C.c_uint8))
    if rc:
        raise ValueError(f"oracle_delete_and_clean failed rc={rc}")
    return Clean(int(tot[0]), pe, pv, st)


@dataclass
class Decomposition:
    cluster_map: np.ndarray
    n_clusters: int
    status_after_clean: np.ndarray
    clusters: list  # list of (nVertices, nEdges, cut, vertices-in-reference-order)
    times: tuple  # (count, clean, extract) seconds
    member_rank: np.ndarray  # position at which each vertex joined its cluster


def decompose(off, nb, eps: float, with_clusters: bool = True) -> Decomposition:
    n = len(off) - 1
    off = np.ascontiguousarray(off, np.int64)
    nb = np.ascontiguousarray(nb, np.int64)
    cm = np.zeros(n, np.int64)
    nc = np.zeros(1, np.int64)
    st = np.zeros(len(nb), np.uint8)
    cap = 4 * n + 8 if with_clusters else 0
    flat = np.zeros(max(cap, 1), np.int64)
    flen = np.zeros(1, np.int64)
    t = np.zeros(3, np.float64)
    mr = np.zeros(n, np.int64)
    rc = lib().oracle_decompose(n, _p(off, C.c_int64), _p(nb, C.c_int64), eps,
                                _p(cm, C.c_int64), _p(nc, C.c_int64), _p(st, C.c_uint8),
                                _p(flat, C.c_int64) if with_clusters else None, cap,
                                _p(flen, C.c_int64), _p(t, C.c_double), _p(mr, C.c_int64))
    if rc:
        raise ValueError(f"oracle_decompose failed rc={rc}")
    clusters = []
    if with_clusters:
        p = 0
        fl = flat[: int(flen[0])]
        while p < len(fl):
            nv, ne, cut = int(fl[p]), int(fl[p + 1]), int(fl[p + 2])
            clusters.append((nv, ne, cut, fl[p + 3 : p + 3 + nv].copy()))
            p += 3 + nv
    return Decomposition(cm, int(nc[0]), st, clusters, tuple(float(x) for x in t), mr)


def extract_from_state(off, nb, eps: float, state: Clean):
    """disjointExtract's seed loop from a state after initial_clean and any
    delete_and_clean calls.  Returns (cluster_map, n_clusters)."""
    n = len(off) - 1
    off = np.ascontiguousarray(off, np.int64)
    nb = np.ascontiguousarray(nb, np.int64)
    pe = state.perEdge.copy()
    pv = state.perVertex.copy()
    tot = np.array([state.total], np.int64)
    st = np.ascontiguousarray(state.status, np.uint8)
    cm = np.zeros(n, np.int64)
    nc = np.zeros(1, np.int64)
    rc = lib().oracle_extract_state(n, _p(off, C.c_int64), _p(nb, C.c_int64), eps,
                                    _p(pe, C.c_double), _p(pv, C.c_double), _p(tot, C.c_int64),
                                    _p(st, C.c_uint8), _p(cm, C.c_int64), _p(nc, C.c_int64))
    if rc:
        raise ValueError(f"oracle_extract_state failed rc={rc}")
    return cm, int(nc[0])


def extract_inner(off, nb, status, u: int, eps: float):
    """The inner loop of seed u on the given slot statuses: (candidates in
    ascending id, their weights, internal, potential)."""
    n = len(off) - 1
    off = np.ascontiguousarray(off, np.int64)
    nb = np.ascontiguousarray(nb, np.int64)
    st = np.ascontiguousarray(status, np.uint8)
    cap = max(n, 1)
    cand = np.zeros(cap, np.int64)
    wgt = np.zeros(cap, np.float64)
    nc = np.zeros(1, np.int64)
    iw = np.zeros(2, np.float64)
    rc = lib().oracle_extract_inner(n, _p(off, C.c_int64), _p(nb, C.c_int64), _p(st, C.c_uint8),
                                    u, eps, _p(cand, C.c_int64), _p(wgt, C.c_double), cap,
                                    _p(nc, C.c_int64), _p(iw[:1], C.c_double),
                                    _p(iw[1:], C.c_double))
    if rc:
        raise ValueError(f"oracle_extract_inner failed rc={rc}")
    k = int(nc[0])
    return cand[:k].copy(), wgt[:k].copy(), float(iw[0]), float(iw[1])


def cluster_edges(off, nb, status, verts):
    """The deletion pairs of a cluster with members in the given order, on
    the given slot statuses: shape (k, 2)."""
    n = len(off) - 1
    off = np.ascontiguousarray(off, np.int64)
    nb = np.ascontiguousarray(nb, np.int64)
    st = np.ascontiguousarray(status, np.uint8)
    vs = np.ascontiguousarray(verts, np.int64)
    cap = max(len(nb) // 2, 1)
    out = np.zeros(2 * cap, np.int64)
    k = np.zeros(1, np.int64)
    rc = lib().oracle_cluster_edges(n, _p(off, C.c_int64), _p(nb, C.c_int64), _p(st, C.c_uint8),
                                    _p(vs, C.c_int64), len(vs), _p(out, C.c_int64), cap,
                                    _p(k, C.c_int64))
    if rc:
        raise ValueError(f"oracle_cluster_edges failed rc={rc}")
    return out[:2 * int(k[0])].reshape(-1, 2).copy()


def connected_components(off, nb):
    n = len(off) - 1
    off = np.ascontiguousarray(off, np.int64)
    nb = np.ascontiguousarray(nb, np.int64)
    comp = np.zeros(n, np.int64)
    nc = np.zeros(1, np.int64)
    rc = lib().oracle_connected_components(n, _p(off, C.c_int64), _p(nb, C.c_int64),
                                           _p(comp, C.c_int64), _p(nc, C.c_int64))
    if rc:
        raise ValueError("oracle_connected_components failed")
    return comp, int(nc[0])


# ----------------------------------------------------------------------------
# The reference build (oracle/_ref/ref_dump), when it has been compiled here.

def ref_available() -> bool:
    return os.path.exists(REF_DUMP)


def write_escape(path: str, n: int, u: np.ndarray, v: np.ndarray) -> None:
    """Escape text format: header 'n m', then one 'u v' line per edge."""
    with open(path, "w") as f:
        f.write(f"{n} {len(u)}\n")
        np.savetxt(f, np.stack([u, v], 1).astype(np.int64), fmt="%d")


def ref_dump(path: str, eps: float, extract: bool = True) -> dict:
    """Run the reference's own phases on an Escape file; returns raw outputs."""
    if not ref_available():
        raise FileNotFoundError(REF_DUMP)
    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        subprocess.run([REF_DUMP, os.path.abspath(path), repr(eps), d,
                        "all" if extract else "noextract"], cwd=d, check=True)
        wall = time.perf_counter() - t0
        rd = lambda name, dt: np.fromfile(os.path.join(d, name), dtype=dt)
        out = dict(
            offsets=rd("csr_offsets.i64", np.int64),
            nbors=rd("csr_nbors.i64", np.int64),
            count_total=int(rd("count_total.i64", np.int64)[0]),
            count_perEdge=rd("count_perEdge.f64", np.float64),
            count_perVertex=rd("count_perVertex.f64", np.float64),
            clean_status=rd("clean_status.u8", np.uint8),
            clean_perEdge=rd("clean_perEdge.f64", np.float64),
            clean_perVertex=rd("clean_perVertex.f64", np.float64),
            clean_total=int(rd("clean_total.i64", np.int64)[0]),
            wall=wall,
        )
        with open(os.path.join(d, "times.txt")) as f:
            out["times"] = {k: float(x) for k, x in (ln.split() for ln in f)}
        if extract:
            out["cluster_map"] = rd("cluster_map.i64", np.int64)
            flat = rd("clusters.i64", np.int64)
            clusters, p = [], 1
            while p < len(flat):
                nv, ne, cut = int(flat[p]), int(flat[p + 1]), int(flat[p + 2])
                clusters.append((nv, ne, cut, flat[p + 3 : p + 3 + nv].copy()))
                p += 3 + nv
            out["clusters"] = clusters
        return out
