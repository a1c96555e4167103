"""rtrex_cuda_amd -- MI355X (gfx950) RTRExtractor decomposition path.

Python bindings (ctypes) over the C ABI in ``include/rtrex.h``, implemented by
``librtrex_hip.so`` (hand-written HIP kernels).  The method names mirror the
reference's GPU layer (RTRex_CUDA/gpu_*.cuh) and CPU driver
(RTRex_Sequential/Escape/Decomposition.h):

    DeviceGraph.from_csr(offsets, nbors)     transferGraphToDevice + allocateDeviceState
    g.gpu_common_nbr()                       gpuCommonNbr / commonNbr
    g.sync_weights()                         syncWeightsDeviceToHost
    g.gpu_initial_clean(eps)                 gpuInitialClean / initialClean
    g.sync_edge_status()                     syncEdgeStatusDeviceToHost
    g.gpu_delete_and_clean(pairs, eps)       gpuDeleteAndClean / deleteAndClean
    g.disjoint_extract(eps)                  disjointExtractCUDA / disjointExtract
    g.populate_stats()                       populateStats
    g.close()                                freeDeviceGraph

Errors raise ``RtrexError`` carrying the library's status and message.  There
is no CPU fallback: without the built library or a gfx950 device every call
fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(PKG, "librtrex_hip.so")
GEN_LIB = os.path.join(PKG, "librtrex_gen.so")

STATUS = {0: "OK", 1: "EINVAL", 2: "EDEVICE", 3: "ENOMEM", 4: "EIO", 5: "ESTATE", 6: "EBUSY"}

# Every symbol include/rtrex.h declares.
EXPORTS = [
    "rtrex_last_error", "rtrex_version", "rtrex_escape_header", "rtrex_escape_read",
    "rtrex_graph_create_csr", "rtrex_graph_create_edges", "rtrex_graph_destroy",
    "rtrex_graph_info", "rtrex_graph_get_csr", "rtrex_graph_footprint_bytes",
    "rtrex_count", "rtrex_get_supports", "rtrex_initial_clean", "rtrex_get_edge_status",
    "rtrex_get_survivors", "rtrex_host_alloc", "rtrex_host_free",
    "rtrex_delete_and_clean", "rtrex_extract", "rtrex_decompose", "rtrex_cluster_stats",
    "rtrex_cluster_stats_noalloc", "rtrex_connected_components", "rtrex_graph_load",
    "rtrex_snap_convert",
    "rtrex_get_audit", "rtrex_get_timings", "rtrex_batch_create",
    "rtrex_batch_cluster_counts", "rtrex_batch_decompose",
    "rtrex_server_start", "rtrex_server_submit", "rtrex_server_wait", "rtrex_server_clock",
    "rtrex_server_stop", "rtrex_get_member_order", "rtrex_get_cluster_map", "rtrex_write_decomposition",
    "rtrex_write_components", "rtrex_write_survivors", "rtrex_write_expanded", "rtrex_get_count_params",
    "rtrex_set_count_params", "rtrex_autotune_count", "rtrex_extract_inner",
    "rtrex_cluster_edges",
]


class RtrexError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"rtrex {STATUS.get(status, status)}: {msg}")
        self.status = status


class Audit(C.Structure):
    _fields_ = [(k, C.c_int64) for k in (
        "clean_rounds", "edges_removed", "triangles_removed", "ambiguous_clean",
        "ambiguous_ratio", "seeds", "components", "largest_component", "replayed_edges")]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Timings(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "build_ms", "count_ms", "clean_ms", "extract_ms",
        "count_edge_kernel_ms", "count_vertex_kernel_ms")]

    def as_dict(self) -> dict:
        return {k: float(getattr(self, k)) for k, _ in self._fields_}


class CountParams(C.Structure):
    _fields_ = [("hub_min", C.c_int32), ("hub_max", C.c_int32), ("l0_max", C.c_int32),
                ("l1_max", C.c_int32), ("g_thr", C.c_int32 * 4), ("unit_min", C.c_int32),
                ("unit_hash_mult", C.c_int32), ("l2_max", C.c_int32), ("l3_max", C.c_int32)]

    def as_dict(self) -> dict:
        d = {k: int(getattr(self, k)) for k, _ in self._fields_ if k != "g_thr"}
        d["g_thr"] = [int(x) for x in self.g_thr]
        return d

    @classmethod
    def from_dict(cls, d: dict) -> "CountParams":
        p = cls()
        for k, v in d.items():
            if k == "g_thr":
                for i, x in enumerate(v):
                    p.g_thr[i] = int(x)
            else:
                setattr(p, k, int(v))
        return p


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("workers", C.c_int), ("hbm_budget_bytes", C.c_int64),
                ("small_job_edges", C.c_int64), ("batch_max_edges", C.c_int64),
                ("batch_max_graphs", C.c_int64)]


class JobInfo(C.Structure):
    _fields_ = [("submit_s", C.c_double), ("start_s", C.c_double), ("end_s", C.c_double),
                ("batch_graphs", C.c_int64), ("bytes", C.c_int64), ("worker", C.c_int32),
                ("pooled", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: (float if t is C.c_double else int)(getattr(self, k)) for k, t in self._fields_}


_lib = None
_gen = None
P64, PU32, PD, PU8 = C.POINTER(C.c_int64), C.POINTER(C.c_uint32), C.POINTER(C.c_double), C.POINTER(C.c_uint8)


def lib() -> C.CDLL:
    """Loads librtrex_hip.so (build it first: python -m rtrex_cuda_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(HIP_LIB):
            raise RtrexError(2, f"{HIP_LIB} is not built; run python -m rtrex_cuda_amd.build")
        L = C.CDLL(HIP_LIB)
        G = C.c_void_p
        L.rtrex_last_error.restype = C.c_char_p
        L.rtrex_version.restype = C.c_char_p
        sig = {
            "rtrex_escape_header": [C.c_char_p, P64, P64],
            "rtrex_escape_read": [C.c_char_p, C.c_int64, PU32, PU32],
            "rtrex_graph_create_csr": [P64, P64, C.c_int64, C.c_int64, C.c_int, C.POINTER(G)],
            "rtrex_graph_create_edges": [PU32, PU32, C.c_int64, C.c_int64, C.c_int, C.POINTER(G)],
            "rtrex_graph_info": [G, P64, P64],
            "rtrex_graph_get_csr": [G, P64, P64],
            "rtrex_count": [G, P64],
            "rtrex_get_supports": [G, PD, PD, P64],
            "rtrex_initial_clean": [G, C.c_double],
            "rtrex_get_edge_status": [G, PU8],
            "rtrex_get_survivors": [G, PU32, C.c_int64],
            "rtrex_delete_and_clean": [G, P64, C.c_int64, C.c_double],
            "rtrex_extract_inner": [G, C.c_int64, C.c_double, P64, PD, C.c_int64, P64, PD, PD],
            "rtrex_cluster_edges": [G, P64, C.c_int64, P64, C.c_int64, P64],
            "rtrex_extract": [G, C.c_double, P64, P64],
            "rtrex_decompose": [G, C.c_double, P64, P64],
            "rtrex_cluster_stats": [G, P64, P64, P64],
            "rtrex_cluster_stats_noalloc": [G, P64, P64, P64],
            "rtrex_connected_components": [G, P64, P64],
            "rtrex_graph_load": [C.c_char_p, C.c_int, C.c_int, C.POINTER(G)],
            "rtrex_snap_convert": [C.c_char_p, C.c_char_p, C.c_int, P64, P64],
            "rtrex_get_audit": [G, C.POINTER(Audit)],
            "rtrex_get_timings": [G, C.POINTER(Timings)],
            "rtrex_batch_decompose": [C.c_int64, P64, P64, PU32, PU32, C.c_double, C.c_int,
                                      P64, P64, C.POINTER(Timings), C.POINTER(Audit)],
            "rtrex_batch_create": [C.c_int64, P64, P64, PU32, PU32, C.c_int, C.POINTER(G)],
            "rtrex_batch_cluster_counts": [G, P64, C.c_int64],
            "rtrex_server_start": [C.POINTER(ServerConfig), C.POINTER(G)],
            "rtrex_server_submit": [G, PU32, PU32, C.c_int64, C.c_int64, C.c_double, P64],
            "rtrex_server_wait": [G, C.c_int64, P64, P64, C.POINTER(JobInfo)],
            "rtrex_server_stop": [G],
            "rtrex_get_member_order": [G, P64],
            "rtrex_get_cluster_map": [G, P64],
            "rtrex_write_decomposition": [C.c_char_p, C.c_int64, P64, P64, P64, P64, C.c_int64],
            "rtrex_write_components": [C.c_char_p, C.c_int64, P64, P64],
            "rtrex_write_expanded": [C.c_char_p, C.c_int64, P64, P64, P64, P64, C.c_int64,
                                     C.c_int32, P64, PD],
            "rtrex_write_survivors": [C.c_char_p, C.c_int64, P64, P64, PU8],
            "rtrex_get_count_params": [G, C.POINTER(CountParams)],
            "rtrex_set_count_params": [G, C.POINTER(CountParams)],
            "rtrex_autotune_count": [G, C.c_int, C.POINTER(CountParams), PD, PD],
            "rtrex_host_alloc": [C.c_size_t, C.POINTER(C.c_void_p)],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.rtrex_graph_destroy.argtypes = [G]
        L.rtrex_graph_destroy.restype = None
        L.rtrex_graph_footprint_bytes.argtypes = [C.c_int64, C.c_int64]
        L.rtrex_graph_footprint_bytes.restype = C.c_int64
        L.rtrex_server_clock.argtypes = [G]
        L.rtrex_server_clock.restype = C.c_double
        L.rtrex_host_free.argtypes = [C.c_void_p]
        L.rtrex_host_free.restype = None
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise RtrexError(rc, lib().rtrex_last_error().decode())


def host_empty(shape, dtype=np.int64) -> np.ndarray:
    """An uninitialised numpy array in page-locked host memory the device can
    address (rtrex_host_alloc); result arrays there are written by the device
    directly.  The memory is released when the array is garbage-collected."""
    import weakref
    dt = np.dtype(dtype)
    count = int(np.prod(shape))
    nbytes = max(1, count * dt.itemsize)
    ptr = C.c_void_p()
    _check(lib().rtrex_host_alloc(nbytes, C.byref(ptr)))
    buf = (C.c_byte * nbytes).from_address(ptr.value)
    weakref.finalize(buf, lib().rtrex_host_free, ptr.value)
    return np.frombuffer(buf, dtype=dt, count=count).reshape(shape)


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def version() -> str:
    return lib().rtrex_version().decode()


def footprint_bytes(n: int, nnz: int) -> int:
    return int(lib().rtrex_graph_footprint_bytes(n, nnz))


# --------------------------------------------------------------------- loader
def load_escape(path: str):
    """Escape text file -> (n, u, v) with u/v uint32, one entry per line
    (loadGraph_Escape, GraphIO.cpp:19-74)."""
    n = C.c_int64()
    m = C.c_int64()
    b = os.fsencode(path)
    _check(lib().rtrex_escape_header(b, C.byref(n), C.byref(m)))
    u = np.empty(m.value, np.uint32)
    v = np.empty(m.value, np.uint32)
    _check(lib().rtrex_escape_read(b, m.value, _p(u, PU32), _p(v, PU32)))
    return int(n.value), u, v


# ---------------------------------------------------------------------- graph
class DeviceGraph:
    """A graph resident in HBM with its per-edge / per-vertex state."""

    def __init__(self, handle: C.c_void_p, n: int, nnz: int):
        self._h = handle
        self.n = n
        self.nnz = nnz

    @classmethod
    def from_csr(cls, offsets, nbors, device: int = 0) -> "DeviceGraph":
        off = np.ascontiguousarray(offsets, np.int64)
        nb = np.ascontiguousarray(nbors, np.int64)
        h = C.c_void_p()
        _check(lib().rtrex_graph_create_csr(_p(off, P64), _p(nb, P64), len(off) - 1,
                                            len(nb), device, C.byref(h)))
        return cls(h, len(off) - 1, len(nb))

    @classmethod
    def from_edges(cls, n: int, u, v, device: int = 0) -> "DeviceGraph":
        u = np.ascontiguousarray(u, np.uint32)
        v = np.ascontiguousarray(v, np.uint32)
        h = C.c_void_p()
        _check(lib().rtrex_graph_create_edges(_p(u, PU32), _p(v, PU32), n, len(u), device,
                                              C.byref(h)))
        return cls(h, n, 2 * len(u))

    @classmethod
    def load(cls, path: str, fmt: str = "escape", device: int = 0) -> "DeviceGraph":
        """An edge-list text file parsed on the device and built there: the
        reference's Escape loader (fmt "escape") or snap_to_rtr.py's
        conversion (fmt "snap")."""
        code = {"escape": 0, "snap": 1}[fmt]
        h = C.c_void_p()
        _check(lib().rtrex_graph_load(str(path).encode(), code, device, C.byref(h)))
        n, nnz = C.c_int64(), C.c_int64()
        _check(lib().rtrex_graph_info(h, C.byref(n), C.byref(nnz)))
        return cls(h, int(n.value), int(nnz.value))

    def close(self) -> None:  # freeDeviceGraph
        if self._h:
            lib().rtrex_graph_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def csr(self):
        off = np.empty(self.n + 1, np.int64)
        nb = np.empty(self.nnz, np.int64)
        _check(lib().rtrex_graph_get_csr(self._h, _p(off, P64), _p(nb, P64)))
        return off, nb

    def gpu_common_nbr(self) -> int:
        t = C.c_int64()
        _check(lib().rtrex_count(self._h, C.byref(t)))
        return int(t.value)

    def sync_weights(self):
        pe = np.empty(self.nnz, np.float64)
        pv = np.empty(self.n, np.float64)
        t = C.c_int64()
        _check(lib().rtrex_get_supports(self._h, _p(pe, PD), _p(pv, PD), C.byref(t)))
        return pe, pv, int(t.value)

    def gpu_initial_clean(self, eps: float) -> None:
        _check(lib().rtrex_initial_clean(self._h, float(eps)))

    def sync_edge_status(self) -> np.ndarray:
        st = np.empty(self.nnz, np.uint8)
        _check(lib().rtrex_get_edge_status(self._h, _p(st, PU8)))
        return st

    def survivors(self) -> np.ndarray:
        """Bitmap of the edges alive after the initial clean, one bool per
        undirected edge in (u < v) lexicographic order."""
        m = self.nnz // 2
        words = np.zeros((m + 31) // 32, np.uint32)
        _check(lib().rtrex_get_survivors(self._h, _p(words, PU32), len(words)))
        bits = np.unpackbits(words.view(np.uint8), bitorder="little")
        return bits[:m].astype(bool)

    def survivor_words(self, out: np.ndarray | None = None) -> np.ndarray:
        """The packed survivor bitmap as rtrex_get_survivors returns it: bit
        e % 32 of word e // 32 set when undirected edge e is alive."""
        m = self.nnz // 2
        words = out if out is not None else np.empty((m + 31) // 32, np.uint32)
        _check(lib().rtrex_get_survivors(self._h, _p(words, PU32), len(words)))
        return words

    def gpu_delete_and_clean(self, pairs, eps: float) -> None:
        pr = np.ascontiguousarray(pairs, np.int64).reshape(-1)
        _check(lib().rtrex_delete_and_clean(self._h, _p(pr, P64), len(pr) // 2, float(eps)))

    def extract_inner(self, u: int, eps: float):
        """gpuExtractInner: seed u's two-hop candidates (ascending id), their
        weights, internal and potential weight on the current state."""
        cap = self.n
        cand = np.empty(max(cap, 1), np.int64)
        wgt = np.empty(max(cap, 1), np.float64)
        nc = C.c_int64()
        iw = C.c_double()
        pw = C.c_double()
        _check(lib().rtrex_extract_inner(self._h, int(u), float(eps), _p(cand, P64), _p(wgt, PD),
                                         cap, C.byref(nc), C.byref(iw), C.byref(pw)))
        k = int(nc.value)
        return cand[:k].copy(), wgt[:k].copy(), float(iw.value), float(pw.value)

    def cluster_edges(self, verts):
        """gpuMarkClusterEdges without marking: the (member, neighbour) pairs
        deleting the cluster's alive edges, shape (k, 2), for
        gpu_delete_and_clean."""
        vs = np.ascontiguousarray(verts, np.int64).reshape(-1)
        cap = self.nnz // 2
        out = np.empty(2 * max(cap, 1), np.int64)
        k = C.c_int64()
        _check(lib().rtrex_cluster_edges(self._h, _p(vs, P64), len(vs), _p(out, P64), cap,
                                         C.byref(k)))
        return out[:2 * int(k.value)].reshape(-1, 2).copy()

    def disjoint_extract(self, eps: float):
        cm = np.empty(self.n, np.int64)
        k = C.c_int64()
        _check(lib().rtrex_extract(self._h, float(eps), _p(cm, P64), C.byref(k)))
        return cm, int(k.value)

    def decompose(self, eps: float, host_map: bool = True, out: np.ndarray | None = None):
        """count -> clean -> extract.  Returns (cluster map, cluster count);
        with host_map=False the map stays in HBM (returned as None; fetch it
        later with cluster_map()).  out: an int64 array of n entries to write
        the map into (host_empty memory is written by the device directly)."""
        cm = (out if out is not None else np.empty(self.n, np.int64)) if host_map else None
        k = C.c_int64()
        _check(lib().rtrex_decompose(self._h, float(eps), _p(cm, P64) if host_map else None,
                                     C.byref(k)))
        return cm, int(k.value)

    def cluster_map(self) -> np.ndarray:
        """Cluster id per vertex from the last extraction (int64)."""
        cm = np.empty(self.n, np.int64)
        _check(lib().rtrex_get_cluster_map(self._h, _p(cm, P64)))
        return cm

    def count_params(self) -> dict:
        p = CountParams()
        _check(lib().rtrex_get_count_params(self._h, C.byref(p)))
        return p.as_dict()

    def set_count_params(self, params: dict) -> None:
        p = CountParams()
        _check(lib().rtrex_get_count_params(self._h, C.byref(p)))
        cur = p.as_dict()
        cur.update(params)
        _check(lib().rtrex_set_count_params(self._h, C.byref(CountParams.from_dict(cur))))

    def autotune(self, repeats: int = 2):
        """Coordinate search of the count-plan parameters; keeps the fastest.
        Returns (params, best_ms, start_ms)."""
        p = CountParams()
        b = C.c_double()
        s0 = C.c_double()
        _check(lib().rtrex_autotune_count(self._h, int(repeats), C.byref(p), C.byref(b),
                                          C.byref(s0)))
        return p.as_dict(), float(b.value), float(s0.value)

    def member_order(self) -> np.ndarray:
        """Position at which each vertex joined its cluster (last extraction)."""
        r = np.empty(self.n, np.int64)
        _check(lib().rtrex_get_member_order(self._h, _p(r, P64)))
        return r

    def populate_stats(self, n_clusters: int):
        nv = np.empty(n_clusters, np.int64)
        ne = np.empty(n_clusters, np.int64)
        cut = np.empty(n_clusters, np.int64)
        _check(lib().rtrex_cluster_stats(self._h, _p(nv, P64), _p(ne, P64), _p(cut, P64)))
        return nv, ne, cut

    def connected_components(self):
        """connectedComp: (component per vertex, count), components numbered
        by smallest vertex as the reference's BFS finds them."""
        comp = np.empty(self.n, np.int64)
        k = C.c_int64()
        _check(lib().rtrex_connected_components(self._h, _p(comp, P64), C.byref(k)))
        return comp, int(k.value)

    def populate_stats_noalloc(self, n_clusters: int):
        """populate_stats through the allocation-free entry point (at most
        2^20 clusters; calls overlapping on one GPU race)."""
        nv = np.empty(n_clusters, np.int64)
        ne = np.empty(n_clusters, np.int64)
        cut = np.empty(n_clusters, np.int64)
        _check(lib().rtrex_cluster_stats_noalloc(self._h, _p(nv, P64), _p(ne, P64),
                                                 _p(cut, P64)))
        return nv, ne, cut

    def audit(self) -> dict:
        a = Audit()
        _check(lib().rtrex_get_audit(self._h, C.byref(a)))
        return a.as_dict()

    def timings(self) -> dict:
        t = Timings()
        _check(lib().rtrex_get_timings(self._h, C.byref(t)))
        return t.as_dict()


def _batch_arrays(graphs):
    vtx = np.array([g[0] for g in graphs], np.int64)
    sizes = np.array([len(g[1]) for g in graphs], np.int64)
    eoff = np.zeros(len(graphs) + 1, np.int64)
    np.cumsum(sizes, out=eoff[1:])
    u = np.ascontiguousarray(np.concatenate([np.asarray(g[1], np.uint32) for g in graphs]))
    v = np.ascontiguousarray(np.concatenate([np.asarray(g[2], np.uint32) for g in graphs]))
    return vtx, eoff, u, v


class Batch(DeviceGraph):
    """Many independent graphs resident as one disjoint union (rtrex_batch_*).
    decompose() returns one cluster map per graph with per-graph ids."""

    def __init__(self, graphs, device: int = 0):
        vtx, eoff, u, v = _batch_arrays(graphs)
        h = C.c_void_p()
        _check(lib().rtrex_batch_create(len(graphs), _p(vtx, P64), _p(eoff, P64),
                                        _p(u, PU32), _p(v, PU32), device, C.byref(h)))
        super().__init__(h, int(vtx.sum()), 2 * len(u))
        self.vtx = vtx
        self.vbase = np.zeros(len(graphs) + 1, np.int64)
        np.cumsum(vtx, out=self.vbase[1:])

    def decompose(self, eps: float, out: np.ndarray | None = None):
        """out: an int64 array of all graphs' vertices to write the maps into
        (host_empty memory is written by the device directly); the returned
        maps are views of it."""
        cm, _ = super().decompose(eps, out=out)
        nc = np.empty(len(self.vtx), np.int64)
        _check(lib().rtrex_batch_cluster_counts(self._h, _p(nc, P64), len(nc)))
        maps = [cm[self.vbase[i]:self.vbase[i + 1]] for i in range(len(self.vtx))]
        return maps, nc


def batch_decompose(graphs, eps: float, device: int = 0):
    """graphs: list of (n, u, v) with local ids.  Returns (list of cluster maps,
    list of cluster counts, timings, audit)."""
    vtx = np.array([g[0] for g in graphs], np.int64)
    sizes = np.array([len(g[1]) for g in graphs], np.int64)
    eoff = np.zeros(len(graphs) + 1, np.int64)
    np.cumsum(sizes, out=eoff[1:])
    u = np.ascontiguousarray(np.concatenate([np.asarray(g[1], np.uint32) for g in graphs]))
    v = np.ascontiguousarray(np.concatenate([np.asarray(g[2], np.uint32) for g in graphs]))
    cm = np.empty(int(vtx.sum()), np.int64)
    nc = np.empty(len(graphs), np.int64)
    t = Timings()
    a = Audit()
    _check(lib().rtrex_batch_decompose(len(graphs), _p(vtx, P64), _p(eoff, P64), _p(u, PU32),
                                       _p(v, PU32), float(eps), device, _p(cm, P64),
                                       _p(nc, P64), C.byref(t), C.byref(a)))
    vb = np.zeros(len(graphs) + 1, np.int64)
    np.cumsum(vtx, out=vb[1:])
    maps = [cm[vb[i]:vb[i + 1]] for i in range(len(graphs))]
    return maps, [int(x) for x in nc], t.as_dict(), a.as_dict()


class Server:
    """Request loop with HBM-footprint admission and continuous batching
    (rtrex_server_*): submit() queues a job (the edges are copied) and returns
    its id; wait(id) blocks for (cluster_map, n_clusters, info)."""

    def __init__(self, device: int = 0, workers: int = 4, hbm_budget_bytes: int = 16 << 30,
                 small_job_edges: int = 1 << 16, batch_max_edges: int = 1 << 22,
                 batch_max_graphs: int = 1024):
        cfg = ServerConfig(device, workers, hbm_budget_bytes, small_job_edges,
                           batch_max_edges, batch_max_graphs)
        self._h = C.c_void_p()
        _check(lib().rtrex_server_start(C.byref(cfg), C.byref(self._h)))
        self._n = {}

    def clock(self) -> float:
        return float(lib().rtrex_server_clock(self._h))

    def submit(self, n: int, u, v, eps: float) -> int:
        u = np.ascontiguousarray(u, np.uint32)
        v = np.ascontiguousarray(v, np.uint32)
        j = C.c_int64()
        _check(lib().rtrex_server_submit(self._h, _p(u, PU32), _p(v, PU32), int(n), len(u),
                                         float(eps), C.byref(j)))
        self._n[j.value] = int(n)
        return int(j.value)

    def wait(self, job: int):
        cm = np.empty(self._n.pop(job), np.int64)
        k = C.c_int64()
        info = JobInfo()
        _check(lib().rtrex_server_wait(self._h, int(job), _p(cm, P64), C.byref(k), C.byref(info)))
        return cm, int(k.value), info.as_dict()

    def close(self) -> None:
        if self._h:
            lib().rtrex_server_stop(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ---------------------------------------------------------------- output files
def _csr64(offsets, nbors):
    return (np.ascontiguousarray(offsets, np.int64), np.ascontiguousarray(nbors, np.int64))


def write_decomposition(name: str, offsets, nbors, cluster_of, member_rank, n_clusters: int):
    """<name>-decomposition/-stats/-cut/-freq.txt exactly as the reference's
    printDecomposition writes them (RTRex.cpp:19-118)."""
    off, nb = _csr64(offsets, nbors)
    co = np.ascontiguousarray(cluster_of, np.int64)
    mr = np.ascontiguousarray(member_rank, np.int64)
    _check(lib().rtrex_write_decomposition(name.encode(), len(off) - 1, _p(off, P64), _p(nb, P64),
                                           _p(co, P64), _p(mr, P64), int(n_clusters)))


def snap_convert(in_path: str, out_path: str, device: int = 0):
    """snap_to_rtr.py's conversion (RTRex_CUDA/snap_to_rtr.py:47-142) on the
    device: the SNAP edge list in_path written as the Escape file out_path.
    Returns (n, m)."""
    n, m = C.c_int64(), C.c_int64()
    _check(lib().rtrex_snap_convert(str(in_path).encode(), str(out_path).encode(), device,
                                    C.byref(n), C.byref(m)))
    return int(n.value), int(m.value)


def write_expanded(name: str, offsets, nbors, cluster_of, member_rank, n_clusters: int,
                   threshold: int):
    """expandClusters + the four files and <name>-scores.txt, as the
    reference driver's mode e writes them (Decomposition.h:943-1008,
    RTRex.cpp:120-135,166-174,196-201).  Returns (expanded cluster per vertex,
    score per vertex, -1 where the vertex did not move)."""
    off, nb = _csr64(offsets, nbors)
    co = np.ascontiguousarray(cluster_of, np.int64)
    mr = np.ascontiguousarray(member_rank, np.int64)
    n = len(off) - 1
    eo = np.empty(n, np.int64)
    sc = np.empty(n, np.float64)
    _check(lib().rtrex_write_expanded(name.encode(), n, _p(off, P64), _p(nb, P64), _p(co, P64),
                                      _p(mr, P64), int(n_clusters), int(threshold), _p(eo, P64),
                                      _p(sc, PD)))
    return eo, sc


def write_components(name: str, offsets, nbors):
    """The connected-components files of RTRex.cpp:203-204."""
    off, nb = _csr64(offsets, nbors)
    _check(lib().rtrex_write_components(name.encode(), len(off) - 1, _p(off, P64), _p(nb, P64)))


def write_survivors(path: str, offsets, nbors, status):
    """The survivor dump of initialClean (Decomposition.h:419-452)."""
    off, nb = _csr64(offsets, nbors)
    st = np.ascontiguousarray(status, np.uint8)
    _check(lib().rtrex_write_survivors(path.encode(), len(off) - 1, _p(off, P64), _p(nb, P64),
                                       _p(st, PU8)))
