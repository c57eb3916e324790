"""Host-side parts of libasdm that need no GPU: the C ABI surface, the t/v/e loader against the
oracle's restatement of relation/graph.cpp:68-148, the seeded generators, and loud failure when
there is no gfx950 device."""
import ctypes
import os
import re
import tempfile

import numpy as np
import pytest

import asdmatch_amd as A
from oracle import oracle as O
from tests import common as T

HEADER = os.path.join(T.GOLDEN, "..", "..", "include", "asdm.h")


def test_every_declared_symbol_is_exported():
    text = open(HEADER).read()
    names = set(re.findall(r"\b(asdm_[a-z0-9_]+)\s*\(", text))
    assert len(names) > 30
    lib = ctypes.CDLL(A.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_loader_matches_oracle_on_yeast():
    paths = [T.yeast_data_path()] + [T.yeast_query_path(n) for n in list(T.yeast_golden()["queries"])[:40]]
    for p in paths:
        a = A.Graph.load(p).csr()
        b = O.OGraph.load(p).csr()
        for x, y in zip(a, b):
            assert np.array_equal(x, y), p


def test_loader_rejects_bad_input():
    with pytest.raises(A.AsdmError) as e:
        A.Graph.load("/nonexistent/x.graph")
    assert e.value.code == -2
    with tempfile.NamedTemporaryFile("w", suffix=".graph", delete=False) as f:
        f.write("t 2 1\nv 0 0 2\nv 1 0 1\ne 0 1\n")  # declared degree 2 but one edge
        bad = f.name
    try:
        with pytest.raises(A.AsdmError):
            A.Graph.load(bad)
    finally:
        os.remove(bad)


def test_extra_tokens_on_edge_lines_are_ignored():
    with tempfile.NamedTemporaryFile("w", suffix=".graph", delete=False) as f:
        f.write("t 3 2\nv 0 1 1\nv 1 0 2\nv 2 1 1\ne 0 1 0\ne 1 2 0\n")
        p = f.name
    try:
        off, nbr, lab = A.Graph.load(p).csr()
        assert off.tolist() == [0, 1, 3, 4]
        assert nbr.tolist() == [1, 0, 2, 1]
        assert lab.tolist() == [1, 0, 1]
    finally:
        os.remove(p)


def _same_csr(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a.csr(), b.csr()))


def test_parallel_parse_of_large_files_matches_sequential_and_oracle():
    # 3M undirected edges: about 50 MB of text, above the parallel-parse threshold (32 MB)
    g = A.Graph.chung_lu(200_000, 3_000_000, 2.5, 16, 1.0, 0x51)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "big.graph")
        g.write(p)
        assert os.path.getsize(p) > 32 << 20
        h = A.Graph.load(p)
        assert _same_csr(g, h)
        assert all(np.array_equal(x, y) for x, y in zip(h.csr(), O.OGraph.load(p).csr()))
        # the same records with the vertex lines after the first edge line: the sequential scan decides
        lines = open(p).read().split("\n")
        nv = 1 + g.n
        moved = [lines[0], lines[nv]] + lines[1:nv] + lines[nv + 1:]
        q = os.path.join(d, "moved.graph")
        open(q, "w").write("\n".join(moved))
        with pytest.raises(A.AsdmError):  # the reference's loader needs offsets before edges
            A.Graph.load(q)
        # vertex and edge lines interleaved, each edge after both endpoints' vertex lines: valid for
        # the reference's loader, so the sequential scan (taken when an edge precedes a vertex)
        # must load it
        off, nbr, lab = g.csr()
        out = [lines[0]]
        for v in range(g.n):
            out.append(lines[1 + v])
            out.extend(f"e {int(w)} {v}" for w in nbr[off[v]:off[v + 1]] if int(w) < v)
        r2 = os.path.join(d, "interleaved.graph")
        open(r2, "w").write("\n".join(out) + "\n")
        assert _same_csr(g, A.Graph.load(r2))
        # a malformed record deep inside a chunk fails as the sequential scan does
        bad = lines[:]
        bad[len(bad) // 2] = "e 1"
        r = os.path.join(d, "bad.graph")
        open(r, "w").write("\n".join(bad))
        with pytest.raises(A.AsdmError):
            A.Graph.load(r)
        # declared degrees beyond the header's edge count, and an out-of-range label, in the
        # parallel path: rejected before any adjacency entry is placed
        for v, line in ((3, f"v 3 {int(lab[3])} 1000000"), (7, f"v 7 4294967295 {int(off[8] - off[7])}")):
            bad = lines[:]
            bad[1 + v] = line
            open(r, "w").write("\n".join(bad))
            with pytest.raises(A.AsdmError) as e:
                A.Graph.load(r)
            assert e.value.code == (-2 if v == 3 else -7)


def test_malformed_degrees_labels_and_headers_are_rejected():
    cases = [
        # degrees declared beyond the header's edge count, then more edge records than fit
        ("t 2 1\nv 0 0 0\nv 1 0 5\ne 1 1\ne 1 1\n", -2),
        # a label whose per-label tables cannot exist (max label + 1 overflows)
        ("t 2 1\nv 0 4294967295 1\nv 1 0 1\ne 0 1\n", -7),
        ("t 2 1\nv 0 16777216 1\nv 1 0 1\ne 0 1\n", -7),
        # a header edge count that the file cannot hold (each edge needs a record)
        ("t 2 1000000\nv 0 0 1\nv 1 0 1\ne 0 1\n", -2),
    ]
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "m.graph")
        for text, code in cases:
            open(p, "w").write(text)
            with pytest.raises(A.AsdmError) as e:
                A.Graph.load(p)
            assert e.value.code == code, text
        # the largest label value is accepted
        open(p, "w").write("t 2 1\nv 0 16777215 1\nv 1 0 1\ne 0 1\n")
        assert A.Graph.load(p).info()[2] == 1 << 24
        with pytest.raises(A.AsdmError) as e:
            A.Graph.from_edges([0, 1 << 24], [(0, 1)])
        assert e.value.code == -7
        # a binary cache whose label count field is out of range
        g = A.Graph.from_edges([0, 1, 1], [(0, 1), (1, 2)])
        b = os.path.join(d, "g.csr")
        g.save_binary(b)
        raw = bytearray(open(b, "rb").read())
        raw[12:16] = np.uint32(0xFFFFFFFF).tobytes()
        open(b, "wb").write(bytes(raw))
        with pytest.raises(A.AsdmError) as e:
            A.Graph.load(b)
        assert e.value.code == -7


def test_binary_cache_round_trip():
    g = A.Graph.chung_lu(5000, 40_000, 2.5, 7, 1.0, 0x52)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "g.csr")
        g.save_binary(p)
        h = A.Graph.load(p)
        assert _same_csr(g, h)
        assert h.info() == g.info()
        # truncated or corrupted files are rejected
        raw = open(p, "rb").read()
        open(p, "wb").write(raw[:-4])
        with pytest.raises(A.AsdmError):
            A.Graph.load(p)
        off, nbr, lab = g.csr()
        body = bytearray(raw)
        i = 24 + 4 * g.n + 4 * (g.n + 1)  # first adjacency entry
        body[i:i + 4] = np.uint32(g.n + 5).tobytes()
        open(p, "wb").write(bytes(body))
        with pytest.raises(A.AsdmError):
            A.Graph.load(p)


def test_adjacency_sorted_by_label_then_id():
    g = A.Graph.chung_lu(500, 3000, 2.5, 5, 1.0, 1)
    off, nbr, lab = g.csr()
    for v in range(g.n):
        s = nbr[off[v]:off[v + 1]]
        keys = list(zip(lab[s].tolist(), s.tolist()))
        assert keys == sorted(keys)


def test_generators_are_deterministic_and_well_formed():
    a = A.Graph.chung_lu(3000, 12500, 2.5, 70, 1.0, 0xA5D00001)
    b = A.Graph.chung_lu(3000, 12500, 2.5, 70, 1.0, 0xA5D00001)
    assert np.array_equal(a.edges(), b.edges())
    e = a.edges()
    assert len(e) == 12500
    assert np.all(e[:, 0] < e[:, 1])  # no self-loops, canonical order
    assert len(np.unique(e[:, 0].astype(np.uint64) << 32 | e[:, 1])) == len(e)  # no duplicates
    c = A.Graph.chung_lu(3000, 12500, 2.5, 70, 1.0, 0xA5D00002)
    assert not np.array_equal(a.edges(), c.edges())
    r = A.Graph.rmat(10, 8, 0.57, 0.19, 0.19, 16, 0.0, 9)
    er = r.edges()
    assert np.all(er[:, 0] < er[:, 1]) and len(er) <= 8 * 1024
    assert r.info()[2] <= 16


def test_random_walk_query_is_connected_induced_subgraph():
    g = A.Graph.chung_lu(2000, 10000, 2.5, 8, 1.0, 4)
    _, _, glab = g.csr()
    ge = set(map(tuple, g.edges().tolist()))
    for dens, cap in ((1.0, 0xFFFFFFFF), (0.0, 0xFFFFFFFF), (1.0, 2), (0.5, 0xFFFFFFFF)):
        q = g.random_walk_query(10, dens, cap, 17)
        qe = q.edges()
        assert q.n == 10
        assert len(qe) >= 9  # spanning tree at least
        if dens == 0.0:
            assert len(qe) == 9
        if cap == 2:
            assert len(qe) <= 11
        # connected
        adj = {i: set() for i in range(10)}
        for a, b in qe.tolist():
            adj[a].add(b)
            adj[b].add(a)
        seen, st = {0}, [0]
        while st:
            for w in adj[st.pop()]:
                if w not in seen:
                    seen.add(w)
                    st.append(w)
        assert len(seen) == 10


def test_context_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(A.AsdmError) as e:
        A.Context(0)
    assert e.value.code == -5


def test_failed_host_allocation_returns_oom_instead_of_aborting(tmp_path):
    # A header of 4e9 vertices asks for tens of GB of host arrays.  Under an 8 GiB address-space
    # limit the allocation throws inside the library, and the entry point must turn that into
    # ASDM_E_OOM rather than let the exception cross the C ABI (which would abort the process).
    import subprocess
    import sys

    p = tmp_path / "huge.graph"
    p.write_text("t 4000000000 0\n")
    code = (
        "import resource, ctypes\n"
        "resource.setrlimit(resource.RLIMIT_AS, (8 << 30, 8 << 30))\n"
        "import asdmatch_amd as A\n"
        "lib = ctypes.CDLL(A.LIB_PATH)\n"
        "g = ctypes.c_void_p()\n"
        f"print('rc', lib.asdm_graph_load_text({str(p)!r}.encode(), ctypes.byref(g)), flush=True)\n"
    )
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240,
                       cwd=os.path.join(os.path.dirname(__file__), ".."))
    assert r.returncode == 0, r.stderr[-2000:]
    assert "rc -3" in r.stdout, r.stdout + r.stderr[-2000:]


def test_bench_refuses_more_than_one_gpu():
    # The bench measures one GPU; asked for several it says so and exits before touching a device.
    import subprocess
    import sys

    root = os.path.join(os.path.dirname(__file__), "..")
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], capture_output=True, text=True, timeout=300,
                       cwd=root)
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr


def test_bench_traffic_summaries_and_device_aggregate(tmp_path, monkeypatch):
    """The bench's roofline plumbing: each committed traffic summary parses into one block per kernel,
    a block is used only for its own kernel and for launches of its algorithmic size (within 5 %),
    and the device aggregate sums the timed kernels' algorithmic bytes (host phases and the loops'
    touched-byte counters excluded) over the step's wall time."""
    import sys

    root = os.path.join(os.path.dirname(__file__), "..")
    sys.path.insert(0, root)
    import bench

    for w, path in bench.TRAFFIC_FILES.items():
        if not os.path.exists(path):
            continue
        kernels = [line.split(":", 1)[1].strip() for line in open(path) if line.startswith("kernel:")]
        assert kernels, w
        for line in open(path):
            if line.startswith("alg_bytes_per_launch:"):
                assert float(line.split(":", 1)[1]) > 0, w
    f = tmp_path / "t.txt"
    f.write_text("# header\n\nkernel: a\ntraffic_bytes_per_launch: 200\nalg_bytes_per_launch: 100\nsource: x\n\n"
                 "kernel: b\ntraffic_bytes_per_launch: 900\nalg_bytes_per_launch: 1000\nsource: y\n")
    monkeypatch.setitem(bench.TRAFFIC_FILES, "t", str(f))
    assert bench.read_traffic("a", "t", 103) == (200.0, "x")
    assert bench.read_traffic("b", "t", 960) == (900.0, "y")
    assert bench.read_traffic("b", "t", 1100)[0] is None  # launches of another size
    assert bench.read_traffic("c", "t", 100)[0] is None   # no block for this kernel
    ks = [{"name": "k1", "ms": 2.0, "launches": 4, "bytes": 8e9},
          {"name": "k2", "ms": 1.0, "launches": 1, "bytes": 2e9},
          {"name": "host: wait", "ms": 5.0, "launches": 1, "bytes": 7e9},
          {"name": "k1_touched", "ms": 0.0, "launches": 0, "bytes": 9e9},
          {"name": "counter only", "ms": 0.0, "launches": 3, "bytes": 4e9}]
    agg = bench.device_aggregate(ks, steps=2, wall_s=0.5)
    assert agg["alg_bytes_per_step"] == 5e9
    assert agg["achieved"] == 10.0 and agg["frac"] == round(10.0 / bench.HBM_PEAK_GBS, 5)
    r = bench.roofline_of(ks[:2], "t")
    assert r["kernel"] == "k1" and r["launches"] == 4 and r["avg_launch_ms"] == 0.5
    assert r["achieved"] == 4000.0  # GB/s: 2e9 bytes per launch in 0.5 ms
