"""asdmatch_amd -- MI355X (gfx950) ASDMatch matching engine, Python side.

A thin ctypes layer over the C ABI of ``libasdm.so`` (declared in ``include/asdm.h``).  It mirrors
the reference's call sequence in main.cpp:11-147:

    Graph.load(path)                 Graph::loadGraphFromTextFile   relation/graph.cpp:68-148
    ctx.filter(q, g) -> CandSpace    CandidateSpace::buildCandCFL   relation/candidate_space.cpp:550-721
    cs.plan(q) -> Plan               simplePlan / optCostPlan       plan/optimizer.cpp
    ctx.count(plan, cs)              adaptiveShareJoin (count)      matching/adaptive.cpp:1179-2279
    ctx.local_count(plan, cs)        ALL_LEVEL + LOCAL_COUNT build  matching/join.cpp:1016-1041

The compute path is the HIP library: if ``libasdm.so`` is missing this module raises on import,
and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libasdm.so")

ASDM_OK = 0
ASDM_E_EMPTY_CANDIDATES = 1
ASDM_PLAN_LOCAL = 1
ASDM_PLAN_NO_FACTOR = 2
ASDM_PLAN_COST = 4
ASDM_PLAN_BAGS = 8
ASDM_PLAN_FRONTIER = 16

_ERRNAMES = {
    -1: "ASDM_E_INVALID", -2: "ASDM_E_IO", -3: "ASDM_E_OOM", -4: "ASDM_E_HIP",
    -5: "ASDM_E_NODEVICE", -6: "ASDM_E_BUDGET", -7: "ASDM_E_UNSUPPORTED", -8: "ASDM_E_TIMEOUT",
}


class AsdmError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(f"libasdm.so not built at {LIB_PATH}; run __graft_entry__.build()")
_lib = C.CDLL(LIB_PATH)

_P = C.c_void_p
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


class Stats(C.Structure):
    _fields_ = [("filter_ms", C.c_double), ("plan_ms", C.c_double), ("exec_ms", C.c_double),
                ("kernel_ms", C.c_double), ("alg_bytes", C.c_uint64), ("rows_total", C.c_uint64),
                ("launches", C.c_uint64)]


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("asdm_version", C.c_char_p)
_sig("asdm_ctx_create", C.c_int, C.c_int, C.POINTER(_P))
_sig("asdm_ctx_destroy", None, _P)
_sig("asdm_last_error", C.c_char_p, _P)
_sig("asdm_ctx_set_profiling", C.c_int, _P, C.c_int)
_sig("asdm_ctx_set_time_limit", C.c_int, _P, C.c_double)
_sig("asdm_ctx_set_row_limit", C.c_int, _P, C.c_uint64)
_sig("asdm_kernel_stat", C.c_int, _P, C.c_uint32, C.c_char_p, C.c_uint32, C.POINTER(C.c_double), _u64p, _u64p)
_sig("asdm_last_stats", C.c_int, _P, C.POINTER(Stats))
_sig("asdm_graph_load_text", C.c_int, C.c_char_p, C.POINTER(_P))
_sig("asdm_graph_from_edges", C.c_int, C.c_uint32, _u32p, C.c_uint64, _u32p, C.POINTER(_P))
_sig("asdm_graph_write_text", C.c_int, _P, C.c_char_p)
_sig("asdm_graph_save_binary", C.c_int, _P, C.c_char_p)
_sig("asdm_graph_free", None, _P)
_sig("asdm_graph_info", C.c_int, _P, _u32p, _u64p, _u32p)
_sig("asdm_graph_csr", C.c_int, _P, _u32p, _u32p, _u32p)
_sig("asdm_graph_edges", C.c_int, _P, _u32p, C.c_uint64, _u64p)
_sig("asdm_graph_upload", C.c_int, _P, _P)
_sig("asdm_gen_chung_lu", C.c_int, C.c_uint32, C.c_uint64, C.c_double, C.c_uint32, C.c_double, C.c_uint64,
     C.POINTER(_P))
_sig("asdm_gen_rmat", C.c_int, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint32,
     C.c_double, C.c_uint64, C.POINTER(_P))
_sig("asdm_gen_query", C.c_int, _P, C.c_uint32, C.c_double, C.c_uint32, C.c_uint64, C.POINTER(_P))
_sig("asdm_filter_cfl", C.c_int, _P, _P, _P, C.POINTER(_P))
_sig("asdm_cs_free", None, _P)
_sig("asdm_cs_root", C.c_uint32, _P)
_sig("asdm_cs_sizes", C.c_int, _P, _u32p)
_sig("asdm_cs_info", C.c_int, _P, _u32p, _u32p)
_sig("asdm_plan_uses_bags", C.c_int, _P)
_sig("asdm_plan_bag_info", C.c_int, _P, _u32p, C.POINTER(C.c_double), C.POINTER(C.c_double))
_sig("asdm_cs_candidates", C.c_int, _P, C.c_uint32, _u32p, C.c_uint32, _u32p)
_sig("asdm_cs_arc_weight", C.c_int, _P, C.c_uint32, _u64p)
_sig("asdm_cs_query_paths", C.c_int, _P, _u64p, _u32p)
_sig("asdm_cs_query_path", C.c_int, _P, C.c_uint32, C.c_uint32, _u32p, C.c_uint32, _u32p)
_sig("asdm_cartesian_candidates", C.c_int, _P, _P, C.c_uint32, _u32p, C.c_uint32, _u32p, C.c_uint32, _u32p)
_sig("asdm_plan_create", C.c_int, _P, _P, _u32p, C.c_uint32, C.POINTER(_P))
_sig("asdm_plan_create_with_cards", C.c_int, _P, _P, _u64p, C.POINTER(C.c_double), C.c_uint32, C.c_uint32,
     C.POINTER(_P))
_sig("asdm_plan_free", None, _P)
_sig("asdm_plan_create_ordered", C.c_int, _P, _P, _u32p, C.c_uint32, C.c_uint32, C.POINTER(_P))
_sig("asdm_plan_order", C.c_int, _P, _u32p, _u32p)
_sig("asdm_plan_estimate", C.c_int, _P, C.POINTER(C.c_double), C.POINTER(C.c_double))
_sig("asdm_count", C.c_int, _P, _P, _P, C.c_uint64, _u64p)
_sig("asdm_local_count", C.c_int, _P, _P, _P, C.c_uint64, _u64p, _u64p)
_sig("asdm_enumerate", C.c_int, _P, _P, _P, C.c_uint64, C.POINTER(C.c_uint32), C.c_uint64, _u64p)
_sig("asdm_intersect_k", C.c_int, _P, C.POINTER(_u32p), _u32p, C.c_uint32, _u32p, C.c_uint32, _u32p)
_sig("asdm_intersect_k_noalloc", C.c_int, _P, C.POINTER(_u32p), _u32p, C.c_uint32, _u32p, C.c_uint32, _u32p)
_sig("asdm_batch_count", C.c_int, _P, _P, C.POINTER(_P), C.c_uint32, C.c_uint64, _u64p, C.POINTER(C.c_float))
_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int)
_sig("asdm_filter_many", C.c_int, _P, _P, C.POINTER(_P), C.c_uint32, C.POINTER(_P), _i32p)
_sig("asdm_batch_count_learned", C.c_int, _P, _P, C.POINTER(_P), C.c_uint32, C.c_uint64, _P, _u64p, _i32p, _f32p)
_sig("asdm_mpnn_create", C.c_int, _P, C.c_uint64, C.POINTER(_P))
_sig("asdm_mpnn_free", None, _P)
_sig("asdm_mpnn_num_params", C.c_uint32)
_sig("asdm_mpnn_get_params", C.c_int, _P, _f32p)
_sig("asdm_mpnn_set_params", C.c_int, _P, _f32p)
_sig("asdm_mpnn_save", C.c_int, _P, C.c_char_p)
_sig("asdm_mpnn_load", C.c_int, _P, C.c_char_p)
_sig("asdm_mpnn_forward", C.c_int, _P, C.c_uint32, _u32p, _f32p, _u32p, _u32p, _f32p)
_sig("asdm_mpnn_grad", C.c_int, _P, C.c_uint32, _u32p, _f32p, _u32p, _u32p, _f32p, _f32p, _f32p)
_sig("asdm_mpnn_train_step", C.c_int, _P, C.c_uint32, _u32p, _f32p, _u32p, _u32p, _f32p, C.c_float, _f32p)
_sig("asdm_mpnn_features", C.c_int, _P, _P, _u64p, C.c_uint32, _u32p, _f32p, _u32p, _u32p, _u32p, _u32p)
_sig("asdm_mpnn_estimate", C.c_int, _P, _P, _P, _u64p, C.c_uint32, C.POINTER(C.c_double))
_sig("asdm_plan_create_learned", C.c_int, _P, _P, _P, _P, C.c_uint32, C.POINTER(_P))
_sig("asdm_plain_estimate", C.c_int, _P, _u64p, C.c_uint32, C.POINTER(C.c_double))
_sig("asdm_ctx_set_batch_tuning", C.c_int, _P, C.c_uint32, C.c_double, C.c_uint32)
_sig("asdm_ctx_get_batch_tuning", C.c_int, _P, _u32p, C.POINTER(C.c_double), _u32p)
_sig("asdm_ctx_set_bag_workers", C.c_int, _P, C.c_uint32)
_sig("asdm_ctx_get_bag_workers", C.c_int, _P, _u32p)
_sig("asdm_ctx_set_table_form", C.c_int, _P, C.c_int)
_sig("asdm_dense_product", C.c_int, _P, _u64p, C.c_int, _u64p, _u64p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
     _u64p, C.c_int)
_sig("asdm_batch_autotune", C.c_int, _P, _P, C.POINTER(_P), C.c_uint32, C.c_uint64, _P, C.c_uint32, _u32p,
     C.POINTER(C.c_double), _u32p, _f32p)

MPNN_FEATURES = 8
MPNN_HIDDEN = 64


def version() -> str:
    return _lib.asdm_version().decode()


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _ptr(a: np.ndarray, t=C.c_uint32):
    return a.ctypes.data_as(C.POINTER(t))


class Graph:
    """Label-sorted CSR graph owned by libasdm (host memory, optionally mirrored in HBM)."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.asdm_graph_free(h)
            self._h = None

    @staticmethod
    def _check(rc, what):
        if rc != ASDM_OK:
            raise AsdmError(rc, what)

    @classmethod
    def load(cls, path: str) -> "Graph":
        h = _P()
        cls._check(_lib.asdm_graph_load_text(os.fsencode(path), C.byref(h)), f"load {path}")
        return cls(h)

    @classmethod
    def from_edges(cls, labels: Sequence[int], edges) -> "Graph":
        lab = _u32(labels)
        e = _u32(edges).reshape(-1)
        h = _P()
        cls._check(_lib.asdm_graph_from_edges(len(lab), _ptr(lab), len(e) // 2, _ptr(e), C.byref(h)), "from_edges")
        return cls(h)

    @classmethod
    def chung_lu(cls, n, ne, beta, nlabels, zipf_s, seed) -> "Graph":
        h = _P()
        cls._check(_lib.asdm_gen_chung_lu(n, ne, beta, nlabels, zipf_s, seed, C.byref(h)), "gen_chung_lu")
        return cls(h)

    @classmethod
    def rmat(cls, scale, edge_factor, a, b, c, nlabels, zipf_s, seed) -> "Graph":
        h = _P()
        cls._check(_lib.asdm_gen_rmat(scale, edge_factor, a, b, c, nlabels, zipf_s, seed, C.byref(h)), "gen_rmat")
        return cls(h)

    def random_walk_query(self, k: int, density: float, max_extra: int, seed: int) -> "Graph":
        h = _P()
        self._check(_lib.asdm_gen_query(self._h, k, density, max_extra, seed, C.byref(h)), "gen_query")
        return Graph(h)

    def info(self):
        n, m, nl = C.c_uint32(), C.c_uint64(), C.c_uint32()
        _lib.asdm_graph_info(self._h, C.byref(n), C.byref(m), C.byref(nl))
        return n.value, m.value, nl.value

    @property
    def n(self) -> int:
        return self.info()[0]

    def csr(self):
        n, m, _ = self.info()
        off = np.zeros(n + 1, np.uint32)
        nbr = np.zeros(max(m, 1), np.uint32)
        lab = np.zeros(max(n, 1), np.uint32)
        _lib.asdm_graph_csr(self._h, _ptr(off), _ptr(nbr), _ptr(lab))
        return off, nbr[:m], lab[:n]

    def edges(self) -> np.ndarray:
        ne = C.c_uint64()
        _lib.asdm_graph_edges(self._h, None, 0, C.byref(ne))
        out = np.zeros(2 * max(ne.value, 1), np.uint32)
        self._check(_lib.asdm_graph_edges(self._h, _ptr(out), ne.value, C.byref(ne)), "edges")
        return out[: 2 * ne.value].reshape(-1, 2)

    def write(self, path: str):
        self._check(_lib.asdm_graph_write_text(self._h, os.fsencode(path)), f"write {path}")

    def save_binary(self, path: str):
        """Binary CSR cache; Graph.load reads it back without parsing or sorting."""
        self._check(_lib.asdm_graph_save_binary(self._h, os.fsencode(path)), f"save_binary {path}")


class Context:
    """One HIP device + stream.  Raises AsdmError(ASDM_E_NODEVICE) without a gfx950 GPU."""

    def __init__(self, device: int = 0):
        h = _P()
        rc = _lib.asdm_ctx_create(device, C.byref(h))
        if rc != ASDM_OK:
            raise AsdmError(rc, "asdm_ctx_create")
        self._h = h

    def close(self):
        """Destroys the context now; candidate spaces made from it must not be used afterwards."""
        if getattr(self, "_h", None):
            _lib.asdm_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc not in (ASDM_OK,):
            raise AsdmError(rc, f"{what}: {_lib.asdm_last_error(self._h).decode()}")

    def upload(self, g: Graph):
        self._check(_lib.asdm_graph_upload(self._h, g._h), "upload")

    def filter(self, q: Graph, g: Graph) -> Optional["CandSpace"]:
        """CFL filter + candidate edges; None when some candidate set is empty."""
        self.upload(g)
        h = _P()
        rc = _lib.asdm_filter_cfl(self._h, q._h, g._h, C.byref(h))
        if rc == ASDM_E_EMPTY_CANDIDATES:
            return None
        self._check(rc, "filter_cfl")
        return CandSpace(h, q.n, self)

    def count(self, plan: "Plan", cs: "CandSpace", budget_bytes: int = 0) -> int:
        out = C.c_uint64()
        self._check(_lib.asdm_count(self._h, plan._h, cs._h, budget_bytes, C.byref(out)), "count")
        return out.value

    def local_count(self, plan: "Plan", cs: "CandSpace", nv: Optional[int] = None, budget_bytes: int = 0):
        """(matrix uint64[|Vg|, |Vq|], total count); the matrix holds one row per data vertex of the
        graph cs was filtered against (nv, when given, must equal that vertex count)."""
        rows = cs.data_vertices
        if nv is not None and nv != rows:
            raise ValueError(f"nv = {nv}, but the candidate space is over a graph of {rows} vertices")
        mat = np.zeros((rows, cs.nq), np.uint64)
        out = C.c_uint64()
        self._check(_lib.asdm_local_count(self._h, plan._h, cs._h, budget_bytes, _ptr(mat, C.c_uint64),
                                          C.byref(out)), "local_count")
        return mat, out.value

    def enumerate(self, plan: "Plan", cs: "CandSpace", max_rows: int, budget_bytes: int = 0) -> np.ndarray:
        """Every embedding as a row of data ids by query vertex (uint32[count, nq]); the plan must be
        made with factor=False.  Raises AsdmError (ASDM_E_BUDGET) past max_rows embeddings."""
        out = np.zeros((max(max_rows, 1), cs.nq), np.uint32)
        n = C.c_uint64()
        self._check(_lib.asdm_enumerate(self._h, plan._h, cs._h, budget_bytes, _ptr(out), max_rows, C.byref(n)),
                    "enumerate")
        return out[: n.value]

    def batch_count(self, g: Graph, qs, budget_bytes: int = 0):
        """Exact counts of many queries against one data graph in one batched pass; returns
        (counts uint64[n], completion time of each query in ms from the call start)."""
        self.upload(g)
        n = len(qs)
        hs = (_P * max(n, 1))(*[q._h for q in qs])
        counts = np.zeros(max(n, 1), np.uint64)
        lat = np.zeros(max(n, 1), np.float32)
        self._check(_lib.asdm_batch_count(self._h, g._h, hs, n, budget_bytes, _ptr(counts, C.c_uint64),
                                          lat.ctypes.data_as(C.POINTER(C.c_float))), "batch_count")
        return counts[:n], lat[:n]

    def filter_many(self, qs, g: Graph):
        """Batched filter: one candidate space per query (None where a candidate set is empty)."""
        self.upload(g)
        n = len(qs)
        hs = (_P * max(n, 1))(*[q._h for q in qs])
        out = (_P * max(n, 1))()
        st = np.zeros(max(n, 1), np.int32)
        self._check(_lib.asdm_filter_many(self._h, g._h, hs, n, out, st.ctypes.data_as(_i32p)), "filter_many")
        res = [CandSpace(_P(out[i]), qs[i].n, self) if out[i] else None for i in range(n)]
        bad = [i for i in range(n) if st[i] not in (ASDM_OK, ASDM_E_EMPTY_CANDIDATES)]
        if bad:
            raise AsdmError(int(st[bad[0]]), f"filter_many query {bad[0]}")
        return res

    def batch_count_learned(self, g: Graph, qs, mpnn: Optional["Mpnn"] = None, budget_bytes: int = 0):
        """batch_count with learned plans (plain plans when mpnn is None) and per-query statuses:
        returns (counts uint64[n], status int32[n], completion ms[n]).  A query over the row limit
        has status ASDM_E_TIMEOUT and count 0."""
        self.upload(g)
        n = len(qs)
        hs = (_P * max(n, 1))(*[q._h for q in qs])
        counts = np.zeros(max(n, 1), np.uint64)
        st = np.zeros(max(n, 1), np.int32)
        lat = np.zeros(max(n, 1), np.float32)
        self._check(_lib.asdm_batch_count_learned(self._h, g._h, hs, n, budget_bytes, mpnn._h if mpnn else None,
                                                  _ptr(counts, C.c_uint64), st.ctypes.data_as(_i32p),
                                                  lat.ctypes.data_as(_f32p)), "batch_count_learned")
        return counts[:n], st[:n], lat[:n]

    def batch_tuning(self):
        """(lanes per leaf-depth row, lane-class threshold factor, rows per lane group in a tile)."""
        l, t = C.c_uint32(), C.c_uint32()
        sc = C.c_double()
        self._check(_lib.asdm_ctx_get_batch_tuning(self._h, C.byref(l), C.byref(sc), C.byref(t)), "get_batch_tuning")
        return l.value, sc.value, t.value

    def bag_workers(self) -> int:
        """Worker contexts (streams) that count a batch's bag-engine queries concurrently."""
        n = C.c_uint32()
        self._check(_lib.asdm_ctx_get_bag_workers(self._h, C.byref(n)), "get_bag_workers")
        return n.value

    def set_bag_workers(self, n: int):
        self._check(_lib.asdm_ctx_set_bag_workers(self._h, n), "set_bag_workers")

    def set_table_form(self, form: int):
        """Form of the bag engine's relation products: 0 per product from its term count, 1 LDS row
        accumulation, 2 zero-fill + atomic scatter.  Counts do not depend on it."""
        self._check(_lib.asdm_ctx_set_table_form(self._h, form), "set_table_form")

    def set_batch_tuning(self, leaf_lanes: int, class_scale: float, tile_iters: int):
        self._check(_lib.asdm_ctx_set_batch_tuning(self._h, leaf_lanes, float(class_scale), tile_iters),
                    "set_batch_tuning")

    def batch_autotune(self, g: Graph, qs, mpnn: Optional["Mpnn"] = None, reps: int = 2, budget_bytes: int = 0):
        """Times the batch under each launch shape of the tuner's grid and keeps the fastest;
        returns (leaf lanes, class factor, tile rows, best median ms)."""
        self.upload(g)
        n = len(qs)
        hs = (_P * max(n, 1))(*[q._h for q in qs])
        l, t = C.c_uint32(), C.c_uint32()
        sc = C.c_double()
        ms = C.c_float()
        self._check(_lib.asdm_batch_autotune(self._h, g._h, hs, n, budget_bytes, mpnn._h if mpnn else None, reps,
                                             C.byref(l), C.byref(sc), C.byref(t), C.byref(ms)), "batch_autotune")
        return l.value, sc.value, t.value, ms.value

    def stats(self) -> Stats:
        s = Stats()
        _lib.asdm_last_stats(self._h, C.byref(s))
        return s

    def set_time_limit(self, ms: float):
        """Per-query limit for count/local_count (0 = none); exceeding it raises ASDM_E_TIMEOUT."""
        self._check(_lib.asdm_ctx_set_time_limit(self._h, float(ms)), "set_time_limit")

    def set_row_limit(self, rows: int):
        """Limit on partial embeddings materialized per count (0 = none)."""
        self._check(_lib.asdm_ctx_set_row_limit(self._h, int(rows)), "set_row_limit")

    def set_profiling(self, on: bool):
        _lib.asdm_ctx_set_profiling(self._h, 1 if on else 0)

    def kernel_stats(self):
        out = []
        i = 0
        name = C.create_string_buffer(64)
        while True:
            ms, la, by = C.c_double(), C.c_uint64(), C.c_uint64()
            if _lib.asdm_kernel_stat(self._h, i, name, 64, C.byref(ms), C.byref(la), C.byref(by)) != ASDM_OK:
                break
            out.append({"name": name.value.decode(), "ms": ms.value, "launches": la.value, "bytes": by.value})
            i += 1
        return out

    def intersect(self, lists) -> np.ndarray:
        arrs = [_u32(a) for a in lists]
        k = len(arrs)
        ptrs = (_u32p * k)(*[_ptr(a) for a in arrs])
        sizes = _u32([len(a) for a in arrs])
        cap = int(min(len(a) for a in arrs)) if arrs else 0
        out = np.zeros(max(cap, 1), np.uint32)
        n = C.c_uint32()
        self._check(_lib.asdm_intersect_k(self._h, ptrs, _ptr(sizes), k, _ptr(out), cap, C.byref(n)), "intersect")
        return out[: n.value]

    def dense_product(self, A: np.ndarray, u, B: np.ndarray, method: int = 2) -> np.ndarray:
        """G = A diag(u) B mod 2^64 (uint64; u None for ones) by the bag engine's dense x dense
        kernels: method 1 the LDS-tiled 64-bit kernel, 2 the int8-digit matrix-core form.  A may be
        given as a transposed view (A.T of an nv x nx array) and B likewise: the kernel reads them
        in the layout they are stored in."""
        nx, nv = A.shape
        nv2, ny = B.shape
        if nv2 != nv:
            raise ValueError("inner dimensions differ")
        a_kmajor = not A.flags.c_contiguous and A.T.flags.c_contiguous
        b_kmajor = B.flags.c_contiguous or not B.T.flags.c_contiguous
        Ah = np.ascontiguousarray(A.T if a_kmajor else A, dtype=np.uint64)
        Bh = np.ascontiguousarray(B if b_kmajor else B.T, dtype=np.uint64)
        uh = None if u is None else np.ascontiguousarray(u, dtype=np.uint64)
        G = np.zeros((nx, ny), np.uint64)
        self._check(_lib.asdm_dense_product(self._h, _ptr(Ah, C.c_uint64), int(a_kmajor),
                                            None if uh is None else _ptr(uh, C.c_uint64), _ptr(Bh, C.c_uint64),
                                            int(b_kmajor), nx, nv, ny, _ptr(G, C.c_uint64), method), "dense_product")
        return G

    def intersect_noalloc(self, lists) -> np.ndarray:
        """intersect() through the allocation-free entry point (fixed per-GPU scratch; at most 64
        lists and 2^21 elements in total, else AsdmError ASDM_E_BUDGET)."""
        arrs = [_u32(a) for a in lists]
        k = len(arrs)
        ptrs = (_u32p * max(k, 1))(*[_ptr(a) for a in arrs])
        sizes = _u32([len(a) for a in arrs])
        cap = int(min(len(a) for a in arrs)) if arrs else 0
        out = np.zeros(max(cap, 1), np.uint32)
        n = C.c_uint32()
        self._check(_lib.asdm_intersect_k_noalloc(self._h, ptrs, _ptr(sizes), k, _ptr(out), cap, C.byref(n)),
                    "intersect_noalloc")
        return out[: n.value]


class CandSpace:
    """Device candidate space; keeps its Context alive (the context owns the stream it frees on)."""

    def __init__(self, handle, nq, ctx):
        self._h = handle
        self.nq = nq
        self._ctx = ctx

    def __del__(self):
        # a candidate space is released on its context's stream, so only while that context lives
        if getattr(self, "_h", None) and getattr(self._ctx, "_h", None):
            _lib.asdm_cs_free(self._h)
        self._h = None

    @property
    def root(self) -> int:
        return int(_lib.asdm_cs_root(self._h))

    @property
    def data_vertices(self) -> int:
        """Vertex count of the data graph this candidate space was filtered against."""
        nv = C.c_uint32()
        rc = _lib.asdm_cs_info(self._h, None, C.byref(nv))
        if rc != ASDM_OK:
            raise AsdmError(rc, "cs_info")
        return nv.value

    def sizes(self) -> np.ndarray:
        s = np.zeros(self.nq, np.uint32)
        _lib.asdm_cs_sizes(self._h, _ptr(s))
        return s

    def candidates(self, u: int) -> np.ndarray:
        n = C.c_uint32()
        _lib.asdm_cs_candidates(self._h, u, None, 0, C.byref(n))
        out = np.zeros(max(n.value, 1), np.uint32)
        rc = _lib.asdm_cs_candidates(self._h, u, _ptr(out), n.value, C.byref(n))
        if rc != ASDM_OK:
            raise AsdmError(rc, "cs_candidates")
        return out[: n.value]

    def arc_weight(self, a: int) -> int:
        w = C.c_uint64()
        _lib.asdm_cs_arc_weight(self._h, a, C.byref(w))
        return w.value

    def query_paths(self):
        """(dist uint64[nq, nq], next uint32[nq, nq]) as the reference's setQueryGraphWeights leaves
        them (products of 32-bit arc weights, wrapping mod 2^64)."""
        dist = np.zeros((self.nq, self.nq), np.uint64)
        nxt = np.zeros((self.nq, self.nq), np.uint32)
        rc = _lib.asdm_cs_query_paths(self._h, _ptr(dist, C.c_uint64), _ptr(nxt))
        if rc != ASDM_OK:
            raise AsdmError(rc, "query_paths")
        return dist, nxt

    def query_path(self, i: int, j: int):
        """reconstructPath(i, j): the list of query vertices ([] when there is no path).  Raises
        AsdmError (ASDM_E_UNSUPPORTED) where the reference's reconstruction would not end."""
        out = np.zeros(self.nq + 1, np.uint32)
        n = C.c_uint32()
        rc = _lib.asdm_cs_query_path(self._h, i, j, _ptr(out), len(out), C.byref(n))
        if rc != ASDM_OK:
            raise AsdmError(rc, "query_path")
        return out[: n.value].tolist()

    def cartesian_candidates(self, v: int, path) -> np.ndarray:
        """optimizedCartesianProduct: ascending candidates of path[-1] reached from data vertex v."""
        p = _u32(path)
        cap = int(self.sizes()[int(p[-1])]) if len(p) else 0
        out = np.zeros(max(cap, 1), np.uint32)
        n = C.c_uint32()
        rc = _lib.asdm_cartesian_candidates(self._ctx._h, self._h, int(v), _ptr(p), len(p), _ptr(out), cap, C.byref(n))
        if rc != ASDM_OK:
            raise AsdmError(rc, f"cartesian_candidates: {_lib.asdm_last_error(self._ctx._h).decode()}")
        return out[: n.value]

    def plan(self, q: Graph, order=None, local: bool = False, cards=None, factor: bool = True,
             num_enumerated: Optional[int] = None, cost: bool = False, bags: Optional[bool] = None) -> "Plan":
        """cost: multi-start cost-based order search with the independence model's cardinalities.
        bags: True counts through bag tables (tree decomposition), False enumerates partial
        embeddings, None lets the planner choose."""
        h = _P()
        flags = ((ASDM_PLAN_LOCAL if local else 0) | (0 if factor else ASDM_PLAN_NO_FACTOR) |
                 (ASDM_PLAN_COST if cost else 0) | (ASDM_PLAN_BAGS if bags else 0) |
                 (ASDM_PLAN_FRONTIER if bags is False else 0))
        if num_enumerated is not None:
            rc = _lib.asdm_plan_create_ordered(q._h, self._h, _ptr(_u32(order)), num_enumerated, flags, C.byref(h))
            if rc != ASDM_OK:
                raise AsdmError(rc, "plan_create_ordered")
            return Plan(h, self.nq)
        if cards:
            masks = np.array(list(cards.keys()), np.uint64)
            vals = np.array(list(cards.values()), np.float64)
            rc = _lib.asdm_plan_create_with_cards(q._h, self._h, _ptr(masks, C.c_uint64),
                                                  vals.ctypes.data_as(C.POINTER(C.c_double)), len(masks), flags,
                                                  C.byref(h))
        else:
            o = None if order is None else _ptr(_u32(order))
            rc = _lib.asdm_plan_create(q._h, self._h, o, flags, C.byref(h))
        if rc != ASDM_OK:
            raise AsdmError(rc, f"plan_create: {_lib.asdm_last_error(self._ctx._h).decode()}")
        return Plan(h, self.nq)


    def plan_learned(self, q: Graph, g: Graph, mpnn: "Mpnn", local: bool = False, cost: bool = False) -> "Plan":
        """Plan ordered greedily by the learned estimator's cardinalities (cost: the best greedy
        order over all start vertices by summed estimated prefix cardinalities)."""
        h = _P()
        flags = (ASDM_PLAN_LOCAL if local else 0) | (ASDM_PLAN_COST if cost else 0)
        rc = _lib.asdm_plan_create_learned(mpnn._h, g._h, q._h, self._h, flags, C.byref(h))
        if rc != ASDM_OK:
            raise AsdmError(rc, "plan_create_learned")
        return Plan(h, self.nq)


class GraphBatch:
    """A batch of small graphs for the estimator: node_off[B+1], x[N, F] float32, adj_off[N+1],
    adj[E] (global node ids, symmetric adjacency)."""

    def __init__(self, node_off, x, adj_off, adj):
        self.node_off = _u32(node_off)
        self.x = np.ascontiguousarray(x, np.float32).reshape(-1, MPNN_FEATURES)
        self.adj_off = _u32(adj_off)
        self.adj = _u32(adj)

    @property
    def size(self) -> int:
        return len(self.node_off) - 1

    def _args(self):
        adj = self.adj if len(self.adj) else np.zeros(1, np.uint32)
        return (self.size, _ptr(self.node_off), self.x.ctypes.data_as(_f32p), _ptr(self.adj_off), _ptr(adj))

    @staticmethod
    def concat(batches) -> "GraphBatch":
        node_off, x, adj_off, adj = [np.zeros(1, np.uint32)], [], [np.zeros(1, np.uint32)], []
        nbase = ebase = 0
        for b in batches:
            node_off.append(b.node_off[1:] + nbase)
            adj_off.append(b.adj_off[1:] + ebase)
            adj.append(b.adj + nbase)
            x.append(b.x)
            nbase += int(b.node_off[-1])
            ebase += int(b.adj_off[-1])
        return GraphBatch(np.concatenate(node_off), np.concatenate(x) if x else np.zeros((0, MPNN_FEATURES)),
                          np.concatenate(adj_off), np.concatenate(adj) if adj else np.zeros(0, np.uint32))

    def subset(self, idx) -> "GraphBatch":
        return GraphBatch.concat([self._one(int(i)) for i in idx])

    def _one(self, i: int) -> "GraphBatch":
        n0, n1 = int(self.node_off[i]), int(self.node_off[i + 1])
        e0, e1 = int(self.adj_off[n0]), int(self.adj_off[n1])
        return GraphBatch(np.array([0, n1 - n0], np.uint32), self.x[n0:n1], self.adj_off[n0:n1 + 1] - e0,
                          self.adj[e0:e1] - n0)


class Mpnn:
    """Learned cardinality estimator on a context's device (bf16 weights, fp32 masters and
    accumulation).  Predicts log2(1 + #embeddings) of each graph of a GraphBatch."""

    def __init__(self, ctx: Context, seed: int = 1):
        h = _P()
        ctx._check(_lib.asdm_mpnn_create(ctx._h, seed, C.byref(h)), "mpnn_create")
        self._h = h
        self._ctx = ctx
        self.num_params = int(_lib.asdm_mpnn_num_params())

    def __del__(self):
        if getattr(self, "_h", None) and getattr(self._ctx, "_h", None):
            _lib.asdm_mpnn_free(self._h)
        self._h = None

    def _check(self, rc, what):
        self._ctx._check(rc, what)

    def params(self) -> np.ndarray:
        p = np.zeros(self.num_params, np.float32)
        self._check(_lib.asdm_mpnn_get_params(self._h, p.ctypes.data_as(_f32p)), "mpnn_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        self._check(_lib.asdm_mpnn_set_params(self._h, p.ctypes.data_as(_f32p)), "mpnn_set_params")

    def save(self, path: str):
        self._check(_lib.asdm_mpnn_save(self._h, os.fsencode(path)), f"mpnn_save {path}")

    def load(self, path: str):
        self._check(_lib.asdm_mpnn_load(self._h, os.fsencode(path)), f"mpnn_load {path}")

    def forward(self, b: GraphBatch) -> np.ndarray:
        y = np.zeros(max(b.size, 1), np.float32)
        self._check(_lib.asdm_mpnn_forward(self._h, *b._args(), y.ctypes.data_as(_f32p)), "mpnn_forward")
        return y[: b.size]

    def grad(self, b: GraphBatch, targets):
        t = np.ascontiguousarray(targets, np.float32)
        g = np.zeros(self.num_params, np.float32)
        y = np.zeros(max(b.size, 1), np.float32)
        self._check(_lib.asdm_mpnn_grad(self._h, *b._args(), t.ctypes.data_as(_f32p), g.ctypes.data_as(_f32p),
                                        y.ctypes.data_as(_f32p)), "mpnn_grad")
        return g, y[: b.size]

    def train_step(self, b: GraphBatch, targets, lr: float) -> float:
        t = np.ascontiguousarray(targets, np.float32)
        loss = C.c_float()
        self._check(_lib.asdm_mpnn_train_step(self._h, *b._args(), t.ctypes.data_as(_f32p), float(lr),
                                              C.byref(loss)), "mpnn_train_step")
        return loss.value

    def estimate(self, g: Graph, cs: "CandSpace", masks) -> np.ndarray:
        m = np.ascontiguousarray(masks, np.uint64)
        out = np.zeros(max(len(m), 1), np.float64)
        self._check(_lib.asdm_mpnn_estimate(self._h, g._h, cs._h, _ptr(m, C.c_uint64), len(m),
                                            out.ctypes.data_as(C.POINTER(C.c_double))), "mpnn_estimate")
        return out[: len(m)]


def mpnn_features(g: Graph, cs: "CandSpace", masks) -> GraphBatch:
    """Estimator inputs of the sub-queries induced by masks on a candidate space."""
    m = np.ascontiguousarray(masks, np.uint64)
    nn, na = C.c_uint32(), C.c_uint32()
    rc = _lib.asdm_mpnn_features(g._h, cs._h, _ptr(m, C.c_uint64), len(m), None, None, None, None,
                                 C.byref(nn), C.byref(na))
    if rc != ASDM_OK:
        raise AsdmError(rc, "mpnn_features")
    node_off = np.zeros(len(m) + 1, np.uint32)
    x = np.zeros((max(nn.value, 1), MPNN_FEATURES), np.float32)
    adj_off = np.zeros(nn.value + 1, np.uint32)
    adj = np.zeros(max(na.value, 1), np.uint32)
    rc = _lib.asdm_mpnn_features(g._h, cs._h, _ptr(m, C.c_uint64), len(m), _ptr(node_off),
                                 x.ctypes.data_as(_f32p), _ptr(adj_off), _ptr(adj), C.byref(nn), C.byref(na))
    if rc != ASDM_OK:
        raise AsdmError(rc, "mpnn_features")
    return GraphBatch(node_off, x[: nn.value], adj_off, adj[: na.value])


def plain_estimate(cs: "CandSpace", masks) -> np.ndarray:
    """The planner's own independence-model cardinalities of the sub-queries induced by masks."""
    m = np.ascontiguousarray(masks, np.uint64)
    out = np.zeros(max(len(m), 1), np.float64)
    rc = _lib.asdm_plain_estimate(cs._h, _ptr(m, C.c_uint64), len(m), out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != ASDM_OK:
        raise AsdmError(rc, "plain_estimate")
    return out[: len(m)]


class Plan:
    def __init__(self, handle, nq):
        self._h = handle
        self.nq = nq

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.asdm_plan_free(self._h)
            self._h = None

    def estimate(self):
        """(estimated rows at the leaf depth, estimated rows over all enumerated depths)."""
        a, b = C.c_double(), C.c_double()
        _lib.asdm_plan_estimate(self._h, C.byref(a), C.byref(b))
        return a.value, b.value

    @property
    def uses_bags(self) -> bool:
        """True when the plan counts through bag tables rather than partial embeddings."""
        return bool(_lib.asdm_plan_uses_bags(self._h))

    def bag_info(self):
        """(partition terms, estimated table-element operations, entries of the largest table)."""
        t, c, pk = C.c_uint32(), C.c_double(), C.c_double()
        _lib.asdm_plan_bag_info(self._h, C.byref(t), C.byref(c), C.byref(pk))
        return t.value, c.value, pk.value

    def order(self):
        o = np.zeros(self.nq, np.uint32)
        d = C.c_uint32()
        _lib.asdm_plan_order(self._h, _ptr(o), C.byref(d))
        return o.tolist(), d.value


def count_embeddings(ctx: Context, q: Graph, g: Graph, budget_bytes: int = 0, order=None,
                     bags: Optional[bool] = None) -> int:
    """Filter, plan and count: the reference's `Number of matches` (0 when a candidate set is empty).
    bags: None lets the planner pick the engine, True/False forces bag tables / enumeration."""
    cs = ctx.filter(q, g)
    if cs is None:
        return 0
    plan = cs.plan(q, order=order, bags=bags)
    return ctx.count(plan, cs, budget_bytes)


def local_count(ctx: Context, q: Graph, g: Graph, budget_bytes: int = 0, bags: Optional[bool] = None):
    cs = ctx.filter(q, g)
    if cs is None:
        return np.zeros((g.n, q.n), np.uint64), 0
    plan = cs.plan(q, local=True, bags=bags)
    return ctx.local_count(plan, cs, g.n, budget_bytes)
