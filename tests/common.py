"""Shared helpers for the tests: golden fixtures and seeded inputs built identically for the HIP
library and the oracle."""
import hashlib
import json
import os
import tempfile

import numpy as np

import asdmatch_amd as A
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
YEAST = os.path.join(GOLDEN, "yeast")


def yeast_golden():
    with open(os.path.join(YEAST, "golden.json")) as f:
        return json.load(f)


def yeast_query_path(name):
    return os.path.join(YEAST, "queries", name + ".graph")


def yeast_data_path():
    return os.path.join(YEAST, "yeast.graph")


def synth_golden():
    with open(os.path.join(GOLDEN, "synth.json")) as f:
        return json.load(f)


def make_graph(case):
    if case["kind"] == "chung_lu":
        return A.Graph.chung_lu(case["n"], case["ne"], case["beta"], case["nl"], case["zipf"], case["seed"])
    return A.Graph.rmat(case["scale"], case["ef"], case["a"], case["b"], case["c"], case["nl"], case["zipf"],
                        case["seed"])


def text_sha256(g: "A.Graph") -> str:
    with tempfile.NamedTemporaryFile(suffix=".graph", delete=False) as f:
        path = f.name
    try:
        g.write(path)
        with open(path, "rb") as f:
            return hashlib.sha256(f.read()).hexdigest()
    finally:
        os.remove(path)


def to_oracle(g: "A.Graph") -> "O.OGraph":
    """Same vertices, labels and edges as g, built by the oracle's own loader logic."""
    _, _, lab = g.csr()
    return O.OGraph.from_edges(lab, g.edges())


def graph_from_lists(labels, edges):
    return A.Graph.from_edges(labels, np.asarray(edges, np.uint32).reshape(-1, 2)), \
        O.OGraph.from_edges(labels, np.asarray(edges, np.uint32).reshape(-1, 2))
