"""ctypes wrapper of oracle/liboracle.so -- the CPU restatement of the reference's matching path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg as the checker.  The product (asdmatch_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_DIR = os.path.join(_HERE, "_ref")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


if not os.path.exists(LIB_PATH):
    build()
_lib = C.CDLL(LIB_PATH)
_P = C.c_void_p
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)


_sig("orc_graph_load", _P, C.c_char_p)
_sig("orc_graph_from_edges", _P, C.c_uint32, _u32p, C.c_uint32, _u32p)
_sig("orc_graph_free", None, _P)
_sig("orc_graph_info", None, _P, _u32p, _u32p, _u32p)
_sig("orc_graph_csr", None, _P, _u32p, _u32p, _u32p)
_sig("orc_cfl", _P, _P, _P)
_sig("orc_cs_free", None, _P)
_sig("orc_cs_empty", C.c_int, _P)
_sig("orc_cs_root", C.c_uint32, _P)
_sig("orc_cs_size", C.c_uint32, _P, C.c_uint32)
_sig("orc_cs_cand", None, _P, C.c_uint32, _u32p)
_sig("orc_cs_arc_weight", C.c_uint64, _P, C.c_uint32)
_sig("orc_count", C.c_int, _P, _P, _P, _u32p, _u64p)
_sig("orc_local_count", C.c_int, _P, _P, _P, _u32p, _u64p, _u64p)
_sig("orc_bruteforce_count", C.c_int, _P, _P, _u64p)
_sig("orc_unique_combine", C.c_uint64, _u32p, _u32p, C.c_int)
_sig("orc_bf_unique_combine", C.c_uint64, _u32p, _u32p, C.c_int)
_sig("orc_paths_from_weights", None, C.c_uint32, _u32p, _u32p, _u32p, _u64p, _u32p)
_sig("orc_query_paths", None, _P, _P, _u64p, _u32p)
_sig("orc_reconstruct_path", C.c_int, C.c_uint32, _u32p, C.c_uint32, C.c_uint32, _u32p)
_sig("orc_cartesian", C.c_uint32, _P, _P, C.c_uint32, _u32p, C.c_uint32, _u32p)


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _ptr(a, t=C.c_uint32):
    return a.ctypes.data_as(C.POINTER(t))


class OGraph:
    def __init__(self, h):
        if not h:
            raise ValueError("oracle graph construction failed")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.orc_graph_free(self._h)
            self._h = None

    @classmethod
    def load(cls, path):
        return cls(_lib.orc_graph_load(os.fsencode(path)))

    @classmethod
    def from_edges(cls, labels, edges):
        lab = _u32(labels)
        e = _u32(edges).reshape(-1)
        return cls(_lib.orc_graph_from_edges(len(lab), _ptr(lab), len(e) // 2, _ptr(e)))

    def info(self):
        n, m, nl = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _lib.orc_graph_info(self._h, C.byref(n), C.byref(m), C.byref(nl))
        return n.value, m.value, nl.value

    def csr(self):
        n, m, _ = self.info()
        off = np.zeros(n + 1, np.uint32)
        nbr = np.zeros(max(m, 1), np.uint32)
        lab = np.zeros(max(n, 1), np.uint32)
        _lib.orc_graph_csr(self._h, _ptr(off), _ptr(nbr), _ptr(lab))
        return off, nbr[:m], lab[:n]


class OCandSpace:
    def __init__(self, q: OGraph, g: OGraph):
        self._q, self._g = q, g
        self._h = _lib.orc_cfl(q._h, g._h)
        self.nq = q.info()[0]

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.orc_cs_free(self._h)
            self._h = None

    @property
    def empty(self):
        return bool(_lib.orc_cs_empty(self._h))

    @property
    def root(self):
        return int(_lib.orc_cs_root(self._h))

    def candidates(self, u):
        n = _lib.orc_cs_size(self._h, u)
        out = np.zeros(max(n, 1), np.uint32)
        _lib.orc_cs_cand(self._h, u, _ptr(out))
        return out[:n]

    def arc_weight(self, a):
        return int(_lib.orc_cs_arc_weight(self._h, a))

    def count(self, order=None):
        c = C.c_uint64()
        o = None if order is None else _ptr(_u32(order))
        rc = _lib.orc_count(self._q._h, self._g._h, self._h, o, C.byref(c))
        if rc < 0:
            raise RuntimeError(f"orc_count failed: {rc}")
        return c.value

    def query_paths(self):
        """(dist uint64[nq, nq], next uint32[nq, nq]) as setQueryGraphWeights leaves them."""
        dist = np.zeros((self.nq, self.nq), np.uint64)
        nxt = np.zeros((self.nq, self.nq), np.uint32)
        _lib.orc_query_paths(self._q._h, self._h, _ptr(dist, C.c_uint64), _ptr(nxt))
        return dist, nxt

    def cartesian(self, v, path):
        """optimizedCartesianProduct: ascending candidates of path[-1] reachable from v."""
        p = _u32(path)
        out = np.zeros(max(1, len(self.candidates(int(p[-1])))), np.uint32)
        n = _lib.orc_cartesian(self._q._h, self._h, int(v), _ptr(p), len(p), _ptr(out))
        return out[:n]

    def local_count(self, order=None):
        nv = self._g.info()[0]
        mat = np.zeros((nv, self.nq), np.uint64)
        c = C.c_uint64()
        o = None if order is None else _ptr(_u32(order))
        rc = _lib.orc_local_count(self._q._h, self._g._h, self._h, o, _ptr(mat, C.c_uint64), C.byref(c))
        if rc < 0:
            raise RuntimeError(f"orc_local_count failed: {rc}")
        return mat, c.value


def count(q: OGraph, g: OGraph, order=None) -> int:
    cs = OCandSpace(q, g)
    return 0 if cs.empty else cs.count(order)


def bruteforce_count(q: OGraph, g: OGraph) -> int:
    c = C.c_uint64()
    rc = _lib.orc_bruteforce_count(q._h, g._h, C.byref(c))
    if rc < 0:
        raise RuntimeError(f"orc_bruteforce_count failed: {rc}")
    return c.value


def unique_combine(sets):
    flat = _u32(np.concatenate([np.asarray(s, np.uint32) for s in sets]) if sets else [])
    sizes = _u32([len(s) for s in sets])
    return int(_lib.orc_unique_combine(_ptr(flat) if len(flat) else None, _ptr(sizes), len(sets)))


def bf_unique_combine(sets):
    flat = _u32(np.concatenate([np.asarray(s, np.uint32) for s in sets]) if sets else [])
    sizes = _u32([len(s) for s in sets])
    return int(_lib.orc_bf_unique_combine(_ptr(flat) if len(flat) else None, _ptr(sizes), len(sets)))


def paths_from_weights(nq, qoff, qnbr, w32):
    """allSourcesShortestPaths over 32-bit arc weights (query CSR order): (dist, next)."""
    dist = np.zeros((nq, nq), np.uint64)
    nxt = np.zeros((nq, nq), np.uint32)
    _lib.orc_paths_from_weights(nq, _ptr(_u32(qoff)), _ptr(_u32(qnbr)), _ptr(_u32(w32)), _ptr(dist, C.c_uint64),
                                _ptr(nxt))
    return dist, nxt


def reconstruct_path(nxt, i, j):
    """reconstructPath: the list of vertices, [] when there is none, None when following next
    never reaches j (the reference's loop would not end)."""
    nq = nxt.shape[0]
    out = np.zeros(nq + 1, np.uint32)
    n = _lib.orc_reconstruct_path(nq, _ptr(np.ascontiguousarray(nxt, np.uint32)), i, j, _ptr(out))
    return None if n < 0 else out[:n].tolist()
