"""GPU parity of the bag-table engine (counting through a tree decomposition of the query with
Moebius inversion over same-label partitions): counts and full local-count matrices against brute
force, the oracle's local counts, the reference's golden counts and matrices, and the frontier
engine.
Every test here runs kernels on the MI355X."""
import random
import sys

import numpy as np
import pytest

import asdmatch_amd as A
from oracle import oracle as O
from tests import common as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True, file=sys.stderr)


def bag_plan(cs, q, local=False):
    """The bag plan of q, or None when some quotient needs a table over three variables."""
    try:
        p = cs.plan(q, local=local, bags=True)
    except A.AsdmError as e:
        if e.code == -7:  # ASDM_E_UNSUPPORTED
            return None
        raise
    assert p.uses_bags
    return p


def random_case(rng, n_lo, n_hi, k_lo, k_hi, nlab):
    n = rng.randint(n_lo, n_hi)
    p = rng.uniform(0.15, 0.5)
    edges = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < p]
    if not edges:
        edges = [(0, 1)]
    glab = [rng.randrange(nlab) for _ in range(n)]
    k = rng.randint(k_lo, k_hi)
    qe = set((rng.randrange(i), i) for i in range(1, k))
    for a in range(k):
        for b in range(a + 1, k):
            if rng.random() < 0.15:
                qe.add((a, b))
    qlab = [rng.randrange(nlab) for _ in range(k)]
    return glab, edges, qlab, sorted(qe)


def test_bag_counts_and_local_counts_against_bruteforce(ctx):
    """Random small graphs with few labels (many same-label pairs, hence many partition terms):
    the count equals brute force and the local-count matrix equals the oracle's enumeration."""
    rng = random.Random(5)
    done = 0
    for trial in range(200):
        glab, ge, qlab, qe = random_case(rng, 6, 14, 2, 7, rng.randint(1, 3))
        g, og = T.graph_from_lists(glab, ge)
        q, oq = T.graph_from_lists(qlab, qe)
        cs = ctx.filter(q, g)
        want = O.bruteforce_count(oq, og)
        if cs is None:
            assert want == 0
            continue
        p = bag_plan(cs, q, local=True)
        if p is None:
            continue
        say(f"bags trial {trial}: k={len(qlab)} edges={qe} labels={qlab}")
        mat, c = ctx.local_count(p, cs)
        assert c == want, (trial, qlab, qe)
        ocs = O.OCandSpace(oq, og)
        omat, oc = ocs.local_count()
        assert oc == want
        assert np.array_equal(mat, omat), (trial, qlab, qe)
        assert ctx.count(bag_plan(cs, q), cs) == want
        done += 1
    assert done >= 80


def test_bag_counts_and_local_counts_match_reference(ctx):
    """The synthetic golden cases (counts and LOCAL_COUNT matrices of the reference's ALL_LEVEL
    build) whose quotients all have treewidth <= 2: equal counts, bit-exact matrices."""
    n = nloc = 0
    for case in T.synth_golden():
        g = T.make_graph(case)
        for rec in case["queries"]:
            if rec["count"] is None:
                continue
            q = g.random_walk_query(rec["k"], rec["density"], rec["max_extra"], rec["seed"])
            cs = ctx.filter(q, g)
            local = bool(rec.get("local_npz"))
            p = bag_plan(cs, q, local=local)
            if p is None:
                continue
            say(f"bags reference {case['name']} {rec['seed']} local={local}")
            if local:
                mat, c = ctx.local_count(p, cs)
                ref = np.load(f"{T.GOLDEN}/{rec['local_npz']}")["local"]
                assert np.array_equal(mat, ref), (case["name"], rec["seed"])
                nloc += 1
            else:
                c = ctx.count(p, cs)
            assert c == rec["count"], (case["name"], rec["seed"])
            n += 1
    say(f"bags reference: {n} counts, {nloc} matrices")
    assert n >= 20 and nloc >= 12


def test_bag_engine_matches_frontier_on_power_law_graphs(ctx):
    """Chung-Lu graphs with 4-16 labels and random-walk queries of 5-9 vertices: the bag engine's
    count and local counts equal the frontier engine's (both exact; independent methods)."""
    checked = 0
    for seed, nl in ((11, 4), (12, 8), (13, 16)):
        g = A.Graph.chung_lu(3000, 15000, 2.5, nl, 0.0, seed)
        ctx.upload(g)
        for j in range(12):
            k = 5 + j % 5
            try:
                q = g.random_walk_query(k, 0.5, 0xFFFFFFFF, seed * 100 + j)
            except A.AsdmError:
                continue
            cs = ctx.filter(q, g)
            if cs is None:
                continue
            p = bag_plan(cs, q, local=True)
            if p is None:
                continue
            say(f"bags vs frontier: graph {seed} query {j} k={k}")
            pf = cs.plan(q, local=True, bags=False)
            assert not pf.uses_bags
            mf, cf = ctx.local_count(pf, cs)
            mb, cb = ctx.local_count(p, cs)
            assert cb == cf, (seed, j)
            assert np.array_equal(mb, mf), (seed, j)
            assert ctx.count(bag_plan(cs, q), cs) == ctx.count(cs.plan(q, bags=False), cs)
            checked += 1
    assert checked >= 15


def test_bag_count_wraps_like_size_t(ctx):
    """A star with many same-label leaves on a complete graph has more than 2^64 embeddings: the
    bag engine's count is the true count mod 2^64 (falling factorial of the hub degree)."""
    n = 300
    glab = [0] * n
    ge = [(a, b) for a in range(n) for b in range(a + 1, n)]
    g, _ = T.graph_from_lists(glab, ge)
    k = 8  # hub + 7 leaves (Bell(7) = 877 partition terms): n (n-1) ... (n-7) > 2^64
    q, _ = T.graph_from_lists([0] * k, [(0, i) for i in range(1, k)])
    cs = ctx.filter(q, g)
    p = bag_plan(cs, q)
    assert p is not None
    want = 1
    for i in range(k):
        want *= n - i
    assert want > 2 ** 64
    assert ctx.count(p, cs) == want % 2 ** 64


def _c3_small_cases():
    import json
    with open(f"{T.GOLDEN}/c3_small12.json") as f:
        return json.load(f)


def test_twelve_vertex_local_counts_match_reference(ctx):
    """12-vertex random-walk queries (config 3's size) on small 16-label power-law graphs: the
    LOCAL_COUNT matrices of the reference's ALL_LEVEL build, bit-exact, through the bag engine and
    through the frontier engine."""
    nb = 0
    for case in _c3_small_cases():
        g = T.make_graph(case)
        ctx.upload(g)
        for rec in case["queries"]:
            q = g.random_walk_query(rec["k"], rec["density"], rec["max_extra"], rec["seed"])
            assert T.text_sha256(q) == rec["query_sha256"], (case["name"], rec["seed"])
            cs = ctx.filter(q, g)
            ref = np.load(f"{T.GOLDEN}/{rec['local_npz']}")["local"]
            say(f"12-vertex local counts: {case['name']} seed {rec['seed']} count {rec['count']}")
            mf, cf = ctx.local_count(cs.plan(q, local=True, bags=False), cs)
            assert cf == rec["count"] and np.array_equal(mf, ref), (case["name"], rec["seed"])
            p = bag_plan(cs, q, local=True)
            if p is not None:
                mb, cb = ctx.local_count(p, cs)
                assert cb == rec["count"] and np.array_equal(mb, ref), (case["name"], rec["seed"])
                nb += 1
    assert nb >= 4


def test_config3_query_at_full_size(ctx):
    """The admitted 12-vertex query of config 3 (200,000 vertices, 4,000,000 edges, 16 labels),
    checked through properties that hold at any size: the local-count matrix's column sums equal
    the count; a cell is zero when its vertex is outside the query vertex's candidate set; the count
    is unchanged, and the columns are permuted, when the query's vertices are renumbered."""
    from asdmatch_amd import workloads as W

    (slot, seed, count), = [tuple(x) for x in W.load_manifest()["c3"]["queries"]]
    g = W.graph("c3")
    ctx.upload(g)
    q = W.query(g, W.slots("c3")[slot], seed)
    cs = ctx.filter(q, g)
    p = cs.plan(q, local=True)
    say(f"config 3: slot {slot} seed {seed}, bag engine: {p.uses_bags}")
    mat, c = ctx.local_count(p, cs)
    assert c == count
    assert all(int(s) == c for s in mat.sum(axis=0, dtype=np.uint64))
    for u in range(q.n):
        inside = np.zeros(g.n, bool)
        inside[cs.candidates(u)] = True
        assert not mat[~inside, u].any(), u
    assert ctx.count(cs.plan(q), cs) == c
    # renumber the query vertices
    rng = np.random.default_rng(3)
    perm = rng.permutation(q.n)  # new id of old vertex i is perm[i]
    _, _, qlab = q.csr()
    lab2 = np.zeros(q.n, np.uint32)
    lab2[perm] = qlab
    q2 = A.Graph.from_edges(lab2, perm[q.edges().astype(np.int64)].astype(np.uint32))
    cs2 = ctx.filter(q2, g)
    mat2, c2 = ctx.local_count(cs2.plan(q2, local=True), cs2)
    assert c2 == c
    assert np.array_equal(mat2[:, perm], mat)


def test_relation_product_forms_agree(ctx):
    """The two forms of the product of two edge relations into a dense table (row accumulation in
    LDS; zero-fill + atomic scatter of the terms), each forced for every product: the same counts
    and local-count matrices as brute force and the reference's golden matrices, and the same
    config-3 matrix and config-4 batch counts at full size.  Out-of-range forms are rejected."""
    from asdmatch_amd import workloads as W

    try:
        rng = random.Random(21)
        cases = []
        for trial in range(120):
            glab, ge, qlab, qe = random_case(rng, 6, 14, 2, 7, rng.randint(1, 3))
            cases.append((glab, ge, qlab, qe))
        for form in (1, 2):
            ctx.set_table_form(form)
            done = 0
            for trial, (glab, ge, qlab, qe) in enumerate(cases):
                g, og = T.graph_from_lists(glab, ge)
                q, oq = T.graph_from_lists(qlab, qe)
                cs = ctx.filter(q, g)
                if cs is None:
                    continue
                p = bag_plan(cs, q, local=True)
                if p is None:
                    continue
                mat, c = ctx.local_count(p, cs)
                assert c == O.bruteforce_count(oq, og), (form, trial)
                assert np.array_equal(mat, O.OCandSpace(oq, og).local_count()[0]), (form, trial)
                done += 1
            assert done >= 40, form
            nloc = 0
            for case in T.synth_golden():
                g = T.make_graph(case)
                for rec in case["queries"]:
                    if rec["count"] is None or not rec.get("local_npz"):
                        continue
                    q = g.random_walk_query(rec["k"], rec["density"], rec["max_extra"], rec["seed"])
                    cs = ctx.filter(q, g)
                    p = bag_plan(cs, q, local=True)
                    if p is None:
                        continue
                    mat, c = ctx.local_count(p, cs)
                    ref = np.load(f"{T.GOLDEN}/{rec['local_npz']}")["local"]
                    assert c == rec["count"] and np.array_equal(mat, ref), (form, case["name"], rec["seed"])
                    nloc += 1
            assert nloc >= 12, form
            say(f"table form {form}: {done} brute-force cases, {nloc} reference matrices")
        # config 3 at full size: both forms give the same matrix, with the manifest's count
        (slot, seed, count), = [tuple(x) for x in W.load_manifest()["c3"]["queries"]]
        g = W.graph("c3")
        ctx.upload(g)
        q = W.query(g, W.slots("c3")[slot], seed)
        cs = ctx.filter(q, g)
        mats = []
        for form in (1, 2):
            ctx.set_table_form(form)
            mat, c = ctx.local_count(cs.plan(q, local=True), cs)
            assert c == count, form
            mats.append(mat)
        assert np.array_equal(mats[0], mats[1])
        # a share of config 4's bag-engine queries through the batched call
        man = W.load_manifest()["c4"]
        g4 = W.graph("c4")
        sl = W.slots("c4")
        bag = [(i, s, c) for (i, s, c), e in zip(man["queries"], man["engines"]) if e == "bags"][1::8]
        qs = [W.query(g4, sl[i], s) for i, s, _ in bag]
        want = [c for _, _, c in bag]
        for form in (1, 2):
            ctx.set_table_form(form)
            counts, st, _ = ctx.batch_count_learned(g4, qs)
            assert [int(c) for c in counts] == want and not any(st), form
            say(f"table form {form}: {len(qs)} config-4 bag queries")
        for bad in (-1, 3):
            with pytest.raises(A.AsdmError):
                ctx.set_table_form(bad)
    finally:
        ctx.set_table_form(0)


def _dense_ref(A, u, B):
    """A diag(u) B mod 2^64 in numpy (uint64 products and sums wrap like the kernels')."""
    Au = A * (np.ones(A.shape[1], np.uint64) if u is None else u)[None, :]
    G = np.zeros((A.shape[0], B.shape[1]), np.uint64)
    for k in range(A.shape[1]):
        G += np.outer(Au[:, k], B[k])
    return G


def test_dense_product_forms_match_numpy(ctx):
    """The dense x dense table product, G = A diag(u) B mod 2^64: the LDS-tiled 64-bit kernel and the
    int8-digit matrix-core form against numpy's wrapping uint64 arithmetic, on shapes that are not
    multiples of the tiles, full-range values (every balanced digit, carries through all eight
    bytes), all four storage orientations of the matrix-core form, and an inner size past one
    int32 accumulation span (8,192)."""
    rng = np.random.default_rng(17)
    full = lambda *s: rng.integers(0, 2**64, size=s, dtype=np.uint64)
    cases = [(1, 1, 1), (31, 63, 33), (64, 64, 64), (70, 200, 97), (129, 65, 130), (40, 300, 50)]
    for nx, nv, ny in cases:
        A, B, u = full(nx, nv), full(nv, ny), full(nv)
        want = _dense_ref(A, u, B)
        assert np.array_equal(ctx.dense_product(A, u, B, method=1), want), (nx, nv, ny)
        assert np.array_equal(ctx.dense_product(A, u, B, method=2), want), (nx, nv, ny)
        At = np.ascontiguousarray(A.T).T  # stored nv x nx
        Bt = np.ascontiguousarray(B.T).T  # stored ny x nv
        for a, b in ((At, B), (A, Bt), (At, Bt)):
            assert np.array_equal(ctx.dense_product(a, u, b, method=2), want), (nx, nv, ny)
        say(f"dense product {nx}x{nv}x{ny}: both forms equal numpy")
    # digits at the extremes: all-ones bytes (0xff.. -> -1 digits with carries), 0x80 bytes (the
    # carry edge), and u = None
    A = np.full((33, 100), 0xFFFFFFFFFFFFFFFF, np.uint64)
    A[::2] = 0x8080808080808080
    B = np.full((100, 35), 0x7F7F7F7F7F7F7F7F, np.uint64)
    B[:, ::3] = 0x8000000000000001
    assert np.array_equal(ctx.dense_product(A, None, B, method=2), _dense_ref(A, None, B))
    # an inner size beyond one int32 accumulation span
    A, B, u = full(40, 9000), full(9000, 36), full(9000)
    want = _dense_ref(A, u, B)
    assert np.array_equal(ctx.dense_product(A, u, B, method=2), want)
    assert np.array_equal(ctx.dense_product(A, u, B, method=1), want)
