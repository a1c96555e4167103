"""Seeded synthetic workloads (ctypes over librtrex_gen.so).

Each function returns ``(n, u, v)``: a simple undirected graph, one entry per
edge with ``u < v``, uint32 ids, sorted.  Edge weights (float32) are produced
by ``weights`` as a sidecar; the decomposition path never reads them (the
reference's Escape loader keeps only the first two integers of a line,
GraphIO.cpp:40-41).

Workloads (seeds and parameters of the benchmark configurations):
  C1  chung_lu(10_000, 100_000, 2.5, 500, seed=1)
  C2  planted(335_000, 5_000 communities of 10..80, p=0.3, background
      Chung-Lu gamma 2.2 with 335_000 edges, seed=7).  The community
      parameters alone give ~1.8M intra edges, so the total is ~2.1M edges
      rather than ~926k; the stated community parameters are kept.
  C3  rmat(22, 16, 0.57, 0.19, 0.19, seed=11)   (~64.2M edges)
  C4  4,096 planted graphs, n log-uniform in [50, 20000], communities of
      5..40 with p=0.4 covering all vertices, background mean degree 6
      (Chung-Lu gamma 2.5), seed 23 (graph i uses seed 23 * 1_000_003 + i).
  C5  request mix, seed 42: 70% Chung-Lu with m log-uniform in [1k, 50k]
      (n = m / 5, gamma 2.5, max degree 4 sqrt(n)), 25% planted-community
      with m log-uniform in [50k, 2M] (the C4 generator at n = m / 8.3), 5%
      R-MAT of scale 18, 19 or 20 (edge factor 16, a,b,c = 0.57,0.19,0.19).
      Requests draw graphs from a seeded pool generated once (serve_pool).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import GEN_LIB, RtrexError

_g = None
PU32 = C.POINTER(C.c_uint32)


def _gen() -> C.CDLL:
    global _g
    if _g is None:
        if not os.path.exists(GEN_LIB):
            raise RtrexError(2, f"{GEN_LIB} is not built; run python -m rtrex_cuda_amd.build")
        L = C.CDLL(GEN_LIB)
        L.rtrex_gen_chung_lu.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double,
                                         C.c_uint64, PU32, PU32, C.c_int64]
        L.rtrex_gen_planted.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                        C.c_double, C.c_double, C.c_uint64, C.c_double,
                                        C.c_uint64, PU32, PU32, C.c_int64]
        L.rtrex_gen_rmat.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                     C.c_double, C.c_uint64, PU32, PU32, C.c_int64]
        for f in (L.rtrex_gen_chung_lu, L.rtrex_gen_planted, L.rtrex_gen_rmat):
            f.restype = C.c_int64
        L.rtrex_gen_weights.argtypes = [C.c_uint64, C.c_int, C.c_uint64,
                                        C.POINTER(C.c_float)]
        L.rtrex_gen_weights.restype = None
        _g = L
    return _g


def _call(fn, args, cap: int):
    while True:
        u = np.empty(max(cap, 1), np.uint32)
        v = np.empty(max(cap, 1), np.uint32)
        m = fn(*args, u.ctypes.data_as(PU32), v.ctypes.data_as(PU32), cap)
        if m >= 0:
            return u[:m].copy(), v[:m].copy()
        cap = -m


def chung_lu(n: int, m: int, gamma: float, max_deg: float, seed: int):
    u, v = _call(_gen().rtrex_gen_chung_lu, (n, m, gamma, float(max_deg), seed), m)
    return n, u, v


def planted(n: int, n_comm: int, cmin: int, cmax: int, p_in: float, bg_gamma: float,
            m_bg: int, bg_max_deg: float, seed: int):
    est = int(n * (cmax * p_in / 2 + 1)) + m_bg + 16
    u, v = _call(_gen().rtrex_gen_planted,
                 (n, n_comm, cmin, cmax, p_in, bg_gamma, m_bg, float(bg_max_deg), seed), est)
    return n, u, v


def rmat(scale: int, ef: int, a: float, b: float, c: float, seed: int):
    u, v = _call(_gen().rtrex_gen_rmat, (scale, ef, a, b, c, seed), ef << scale)
    return 1 << scale, u, v


def weights(m: int, kind: str, seed: int) -> np.ndarray:
    """float32 sidecar: kind 'uniform' in (0,1] or 'lognormal' (0,1) clipped to (0,10]."""
    w = np.empty(m, np.float32)
    _gen().rtrex_gen_weights(m, 0 if kind == "uniform" else 1, seed,
                             w.ctypes.data_as(C.POINTER(C.c_float)))
    return w


# ------------------------------------------------------------ configurations
def config1():
    return chung_lu(10_000, 100_000, 2.5, 500, seed=1)


def config2():
    return planted(335_000, 5_000, 10, 80, 0.3, 2.2, 335_000, 2_000, seed=7)


def config3():
    return rmat(22, 16, 0.57, 0.19, 0.19, seed=11)


def config4_graph(i: int, seed: int = 23):
    """Graph i of the 4,096-graph batch."""
    rng = np.random.default_rng(seed * 1_000_003 + i)
    n = int(round(np.exp(rng.uniform(np.log(50), np.log(20_000)))))
    return planted(n, 0, 5, 40, 0.4, 2.5, 3 * n, max(8.0, n ** 0.5),
                   seed=seed * 1_000_003 + i)


def config4(count: int = 4096, seed: int = 23, threads: int = 16):
    """The 4,096-graph batch (generated in parallel; each graph depends only on
    its own seed, so the result does not depend on the thread count)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(lambda i: config4_graph(i, seed), range(count)))


def serve_graph(kind: str, i: int, seed: int = 42):
    """Graph i of one kind of the serving mix ('cl', 'planted' or 'rmat')."""
    rng = np.random.default_rng(seed * 7_919 + i * 31 + {"cl": 1, "planted": 2, "rmat": 3}[kind])
    gseed = seed * 1_000_003 + i * 10 + {"cl": 1, "planted": 2, "rmat": 3}[kind]
    if kind == "cl":
        m = int(np.exp(rng.uniform(np.log(1_000), np.log(50_000))))
        n = max(100, m // 5)
        return chung_lu(n, m, 2.5, 4.0 * n ** 0.5, seed=gseed)
    if kind == "planted":
        m = int(np.exp(rng.uniform(np.log(50_000), np.log(2_000_000))))
        n = max(100, int(m / 8.3))
        return planted(n, 0, 5, 40, 0.4, 2.5, 3 * n, max(8.0, n ** 0.5), seed=gseed)
    scale = 18 + i % 3
    return rmat(scale, 16, 0.57, 0.19, 0.19, seed=gseed)


def serve_pool(n_cl: int = 64, n_planted: int = 32, n_rmat: int = 3, seed: int = 42,
               threads: int = 16):
    """The serving pool: {kind: [graphs]} (generated in parallel)."""
    from concurrent.futures import ThreadPoolExecutor
    jobs = [("cl", i) for i in range(n_cl)] + [("planted", i) for i in range(n_planted)] + \
           [("rmat", i) for i in range(n_rmat)]
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        out = list(ex.map(lambda kj: serve_graph(kj[0], kj[1], seed), jobs))
    pool = {"cl": [], "planted": [], "rmat": []}
    for (k, _), g in zip(jobs, out):
        pool[k].append(g)
    return pool


def serve_trace(seconds: float, rate: float = 200.0, seed: int = 42, pool_sizes=(64, 32, 3)):
    """Poisson arrivals (rate per second) over `seconds`: list of
    (arrival_s, kind, pool index), kinds drawn 70/25/5."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            return out
        r = rng.random()
        kind = "cl" if r < 0.70 else "planted" if r < 0.95 else "rmat"
        idx = int(rng.integers(0, pool_sizes[("cl", "planted", "rmat").index(kind)]))
        out.append((t, kind, idx))
