"""Host-side pieces of the learned estimator's pipeline (no GPU): sub-query sampling, induced
sub-queries, batching of estimator inputs, q-error."""
import numpy as np

import asdmatch_amd as A
from asdmatch_amd import learned as LR
from oracle import mpnn_oracle as M


def _connected(off, nbr, mask):
    ids = [u for u in range(len(off) - 1) if mask >> u & 1]
    seen, stack = {ids[0]}, [ids[0]]
    while stack:
        u = stack.pop()
        for w in nbr[off[u]:off[u + 1]]:
            w = int(w)
            if mask >> w & 1 and w not in seen:
                seen.add(w)
                stack.append(w)
    return len(seen) == len(ids)


def test_connected_masks_are_connected_and_include_the_query():
    g = A.Graph.chung_lu(3000, 12500, 2.5, 70, 1.0, 0xA5D00001)
    rng = np.random.default_rng(4)
    for seed in range(20):
        q = g.random_walk_query(9, 0.5, 0xFFFFFFFF, 1000 + seed)
        off, nbr, _ = q.csr()
        masks = LR.connected_masks(off, nbr, rng, 12)
        assert (1 << q.n) - 1 in masks
        assert len(set(masks)) == len(masks)
        for m in masks:
            assert bin(m).count("1") >= 2
            assert _connected(off, nbr, m)


def test_induced_subquery_keeps_labels_and_edges():
    g = A.Graph.chung_lu(3000, 12500, 2.5, 70, 1.0, 0xA5D00001)
    q = g.random_walk_query(8, 1.0, 0xFFFFFFFF, 77)
    off, nbr, lab = q.csr()
    qe = {(min(a, b), max(a, b)) for a, b in q.edges().tolist()}
    mask = 0b10110111
    sub = LR.induced(q, mask)
    ids = [u for u in range(q.n) if mask >> u & 1]
    _, _, slab = sub.csr()
    assert slab.tolist() == [int(lab[u]) for u in ids]
    want = {(i, j) for i in range(len(ids)) for j in range(i + 1, len(ids))
            if (min(ids[i], ids[j]), max(ids[i], ids[j])) in qe}
    got = {(min(a, b), max(a, b)) for a, b in sub.edges().tolist()}
    assert got == want


def test_graph_batch_concat_and_subset_roundtrip():
    rng = np.random.default_rng(9)
    parts = [A.GraphBatch(*M.random_batch(rng, n, 10)) for n in (3, 1, 5)]
    b = A.GraphBatch.concat(parts)
    assert b.size == 9
    # subset then re-concat reproduces the same graphs in the new order
    idx = [8, 0, 4, 4]
    s = b.subset(idx)
    assert s.size == 4
    for k, i in enumerate(idx):
        one, ref = s._one(k), b._one(i)
        assert np.array_equal(one.x, ref.x)
        assert np.array_equal(one.adj_off, ref.adj_off)
        assert np.array_equal(one.adj, ref.adj)
    # the oracle sees the same graphs in both layouts
    y1, _ = M.forward(M.init_params(1), b.node_off, b.x, b.adj_off, b.adj)
    y2, _ = M.forward(M.init_params(1), s.node_off, s.x, s.adj_off, s.adj)
    assert np.allclose(y2, y1[idx])


def test_q_errors_match_oracle_definition():
    est = np.array([0.0, 10.0, 3.0, 1e12])
    true = np.array([0, 5, 12, 1e9])
    assert np.allclose(LR.q_errors(est, true), M.q_error(est, true))
