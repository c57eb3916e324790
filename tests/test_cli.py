"""The drop-in command line (ASDMatch.out / prepare.out) and the .card reader/writer."""
import ctypes as C
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

import asdmatch_amd as A
from tests import common as T

BIN = os.path.join(os.path.dirname(A.LIB_PATH), "bin")
CARDS = os.path.join(T.YEAST, "cards")
lib = C.CDLL(A.LIB_PATH)


def read_card(path):
    n, t = C.c_uint32(), C.c_double()
    assert lib.asdm_card_read(path.encode(), None, None, 0, C.byref(n), C.byref(t)) == 0
    m, v = (C.c_uint64 * n.value)(), (C.c_double * n.value)()
    assert lib.asdm_card_read(path.encode(), m, v, n.value, C.byref(n), C.byref(t)) == 0
    return list(m), list(v), t.value


def run(args, **kw):
    return subprocess.run([os.path.join(BIN, args[0])] + args[1:], capture_output=True, text=True, timeout=600, **kw)


def test_reference_card_files_parse():
    for q in ("query_8_1", "query_10_7", "query_12_26", "query_16_52"):
        masks, vals, t = read_card(os.path.join(CARDS, q + ".card"))
        nq = int(q.split("_")[1])
        assert len(masks) > nq  # every single vertex plus larger subsets
        assert all(0 < m < (1 << nq) for m in masks if m)
        assert all(v >= 0 for v in vals)
        assert 0 <= t < 100


def test_card_write_read_roundtrip():
    masks = (C.c_uint64 * 3)(1, 3, 7)
    vals = (C.c_double * 3)(10.0, 123.456, 7e9)
    with tempfile.NamedTemporaryFile(suffix=".card", delete=False) as f:
        p = f.name
    try:
        assert lib.asdm_card_write(p.encode(), masks, vals, 3, C.c_double(0.25)) == 0
        m, v, t = read_card(p)
        assert m == [1, 3, 7] and v == [10.0, 123.45, 7e9] and abs(t - 0.25) < 1e-4
    finally:
        os.remove(p)


def test_cli_reports_unopenable_files_like_the_reference():
    r = run(["ASDMatch.out", "-q", "/nonexistent.graph", "-d", T.yeast_data_path()])
    assert "Can not open file /nonexistent.graph ." in r.stdout
    assert r.returncode == 255


def test_cli_rejects_a_negative_memory_budget():
    r = run(["ASDMatch.out", "-q", "/nonexistent.graph", "-d", T.yeast_data_path(), "-m", "-5"])
    assert r.returncode == 1 and "-m takes a positive size" in r.stderr


@pytest.mark.gpu
def test_cli_counts_and_lines(tmp_path):
    gold = T.yeast_golden()["queries"]
    names = list(gold)[:12]
    for name in names:
        r = run(["ASDMatch.out", "-q", T.yeast_query_path(name), "-d", T.yeast_data_path()])
        assert r.returncode == 0, r.stderr
        if gold[name]["empty"]:
            assert r.stdout == ""
            continue
        lines = r.stdout.strip().split("\n")
        heads = [ln.split(":")[0] for ln in lines]
        assert heads == ["Filtering time (s)", "FHD time (s)", "Planning Time", "cost", "Execution Time",
                         "Number of matches", "Number of fixed attributes", "Number of intersections",
                         "Number of fix cases"]
        assert int(lines[5].split(":")[1]) == gold[name]["count"], name
    out = tmp_path / "res.txt"
    r = run(["ASDMatch.out", "-q", T.yeast_query_path(names[0]), "-d", T.yeast_data_path(), "-r", str(out)])
    assert r.stdout == "" and "Number of matches" in out.read_text()


@pytest.mark.gpu
def test_cli_accepts_a_binary_data_graph(tmp_path):
    # the binary CSR cache of the data graph gives the same counts as its text file
    gold = T.yeast_golden()["queries"]
    binpath = tmp_path / "yeast.csr"
    A.Graph.load(T.yeast_data_path()).save_binary(str(binpath))
    for name in [n for n in gold if not gold[n]["empty"]][:6]:
        r = run(["ASDMatch.out", "-q", T.yeast_query_path(name), "-d", str(binpath)])
        assert r.returncode == 0, r.stderr
        assert int(r.stdout.strip().split("\n")[5].split(":")[1]) == gold[name]["count"], name


@pytest.mark.gpu
def test_cli_reads_reference_card_files():
    want = json.load(open(os.path.join(CARDS, "golden.json")))["counts"]
    for q, c in want.items():
        r = run(["ASDMatch.out", "-q", os.path.join(CARDS, q + ".graph"), "-d", T.yeast_data_path(), "-s",
                 os.path.join(CARDS, q + ".card")])
        assert r.returncode == 0, r.stderr
        assert "dp time: " in r.stdout
        assert f"Number of matches: {c}\n" in r.stdout, q


@pytest.mark.gpu
def test_prepare_then_match(tmp_path):
    q = os.path.join(CARDS, "query_12_26.graph")
    card = tmp_path / "q.card"
    r = run(["prepare.out", "-q", q, "-d", T.yeast_data_path(), "-s", str(card)])
    assert r.returncode == 0, r.stderr
    masks, vals, _ = read_card(str(card))
    assert len(masks) >= 1
    r = run(["ASDMatch.out", "-q", q, "-d", T.yeast_data_path(), "-s", str(card)])
    want = json.load(open(os.path.join(CARDS, "golden.json")))["counts"]["query_12_26"]
    assert f"Number of matches: {want}\n" in r.stdout


@pytest.mark.gpu
def test_cli_local_count_csv(tmp_path):
    case = [c for c in T.synth_golden() if c["name"] == "c1"][0]
    g = T.make_graph(case)
    rec = case["queries"][1]
    q = g.random_walk_query(rec["k"], rec["density"], rec["max_extra"], rec["seed"])
    dpath, qpath, csv = tmp_path / "d.graph", tmp_path / "q.graph", tmp_path / "local.csv"
    g.write(str(dpath))
    q.write(str(qpath))
    r = run(["ASDMatch.out", "-q", str(qpath), "-d", str(dpath), "-c", str(csv)])
    assert r.returncode == 0, r.stderr
    mat = np.loadtxt(csv, delimiter=",", dtype=np.uint64, ndmin=2)
    ref = np.load(f"{T.GOLDEN}/{rec['local_npz']}")["local"]
    assert np.array_equal(mat, ref)
