"""Pins the CPU float64 estimator oracle: finite differences, closed forms, bf16 rounding."""
import numpy as np

from oracle import mpnn_oracle as M


def test_bf16_round_nearest_even():
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties go to the even one (1.0)
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0, np.inf], np.float32)
    r = M.bf16_round(v)
    assert r[0] == 1.0
    assert r[1] == 1.0
    assert r[2] == np.float32(1.0 + 2 * 2.0 ** -7)
    assert r[3] == -2.5 and r[4] == 0.0 and np.isinf(r[5])
    x = np.random.default_rng(0).standard_normal(1000).astype(np.float32)
    rx = M.bf16_round(x)
    assert np.all(rx.view(np.uint32) & 0xFFFF == 0)
    assert np.all(np.abs(rx - x) <= np.abs(x) * 2.0 ** -8)


def test_zero_weights_give_output_bias():
    p = np.zeros(M.NPARAMS, np.float32)
    p[M.B_OUT] = 3.25
    no, x, ao, adj = M.random_batch(np.random.default_rng(1), 5)
    y, _ = M.forward(p, no, x, ao, adj)
    assert np.allclose(y, 3.25)


def test_single_node_chain_closed_form():
    # One node, feature x0 = 2, W_in[0][0] = 0.5, each Ws_l[0][0] = 2, w_out[0] = 1: no neighbours,
    # so h0 = 1, h1 = 2, h2 = 4, h3 = 8 in unit 0 and y = 8 + b_out.
    p = np.zeros(M.NPARAMS, np.float32)
    p[M.W_IN + 0 * M.H + 0] = 0.5
    for l in range(M.L):
        p[M.LAYER0 + l * M.LAYER + 0] = 2.0
    p[M.W_OUT] = 1.0
    p[M.B_OUT] = -1.0
    x = np.zeros((1, M.F), np.float32)
    x[0, 0] = 2.0
    y, _ = M.forward(p, np.array([0, 1]), x, np.array([0, 0]), np.array([], np.uint32))
    assert y[0] == 7.0


def test_neighbour_aggregation_closed_form():
    # Path a - b: with only Wn_0[0][0] = 1 (and W_in passing x0), layer 0 gives each node its
    # neighbour's h0; later layers pass unit 0 through with Ws = 1.
    p = np.zeros(M.NPARAMS, np.float32)
    p[M.W_IN] = 1.0
    p[M.LAYER0 + M.H * M.H] = 1.0  # Wn_0[0][0]
    for l in (1, 2):
        p[M.LAYER0 + l * M.LAYER] = 1.0  # Ws_l[0][0]
    p[M.W_OUT] = 1.0
    x = np.zeros((2, M.F), np.float32)
    x[:, 0] = [3.0, 5.0]
    y, _ = M.forward(p, np.array([0, 2]), x, np.array([0, 1, 2]), np.array([1, 0]))
    assert y[0] == 8.0  # h1 = [5, 3] -> sum 8


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    p = M.init_params(3).astype(np.float64)
    p[M.B_IN:M.B_IN + M.H] = 0.05  # keep units active
    no, x, ao, adj = M.random_batch(rng, 6, max_nodes=6)
    t = rng.uniform(0, 5, 6)
    g, _ = M.grad(p, no, x, ao, adj, t, quantize=False)

    def loss(q):
        y, _ = M.forward(q, no, x, ao, adj, quantize=False)
        return 0.5 * np.mean((y - t) ** 2)

    idx = np.concatenate([rng.choice(M.NPARAMS, 60, replace=False), [M.B_OUT, M.W_OUT, M.W_IN, M.B_IN,
                                                                      M.LAYER0 + 2 * M.H * M.H]])
    eps = 1e-6
    for i in idx:
        e = np.zeros(M.NPARAMS)
        e[i] = eps
        fd = (loss(p + e) - loss(p - e)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_gradient_planted_mistake_is_caught():
    # A gradient with the neighbour term's transpose dropped (A instead of A^T) differs from
    # finite differences on a directed (asymmetric) adjacency -- the check above would see it.
    rng = np.random.default_rng(11)
    p = M.init_params(5).astype(np.float64)
    p[M.B_IN:M.B_IN + M.H] = 0.05
    no = np.array([0, 3])
    x = rng.uniform(0, 1, (3, M.F))
    ao = np.array([0, 1, 2, 2])
    adj = np.array([1, 2])  # 0 -> 1, 1 -> 2 only
    t = np.array([1.0])
    g, _ = M.grad(p, no, x, ao, adj, t, quantize=False)
    A = M.adjacency(no, ao, adj)
    assert not np.allclose(A, A.T)

    def loss(q):
        y, _ = M.forward(q, no, x, ao, adj, quantize=False)
        return 0.5 * np.mean((y - t) ** 2)

    fd = np.zeros(M.F * M.H)
    for i in range(M.F * M.H):
        e = np.zeros(M.NPARAMS)
        e[M.W_IN + i] = 1e-6
        fd[i] = (loss(p + e) - loss(p - e)) / 2e-6
    assert np.allclose(g[M.W_IN:M.W_IN + M.F * M.H], fd, rtol=1e-5, atol=1e-6)
    # the planted variant: propagate with A rather than A^T
    w = M.unpack(p)
    _, c = M.forward(p, no, x, ao, adj, quantize=False)
    hs = c["hs"]
    dy = (c["r"] @ w["w_out"] + w["b_out"] - t)
    dh = c["seg"].T @ np.outer(dy, w["w_out"])
    for l in reversed(range(M.L)):
        dz = dh * (hs[l + 1] > 0)
        dh = dz @ w[f"ws{l}"].T + A @ (dz @ w[f"wn{l}"].T)
    dz = dh * (hs[0] > 0)
    wrong = (c["x"].T @ dz).ravel()
    assert np.max(np.abs(wrong - fd)) > 1e-3 * max(1e-3, np.max(np.abs(fd)))


def test_q_error():
    assert np.allclose(M.q_error([10, 1, 0, 50], [5, 4, 0, 50]), [2, 4, 1, 1])
