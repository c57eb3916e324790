"""CPU float64 restatement of the learned cardinality estimator -- TEST INFRASTRUCTURE ONLY.

Only tests, the smoke check and the bench's CPU baseline use this module, as the checker of the
GPU kernels; the product never calls it.

The reference has no network: it fills subsetToCard (plan/estimator.cpp:7) with sampling
estimates (plan/estimator.cpp:173-291).  This file restates the network that takes that role
here, from its definition:

    h0      = relu(x W_in + b_in)
    h_{l+1} = relu(h_l Ws_l + (A h_l) Wn_l + b_l)            l = 0, 1, 2
    y_g     = (sum_{n in g} h3[n]) . w_out + b_out
    loss    = 0.5 * mean_g (y_g - t_g)^2

with the weights taken at bf16 precision (round to nearest even from the fp32 masters), every
product in float64.  Parameter layout (floats): w_in [F][H], b_in [H], then per layer Ws [H][H],
Wn [H][H], b [H], then w_out [H], b_out [1]; F = 8, H = 64, 3 layers.

Pinning: the analytic gradient is checked against central finite differences of the forward
pass, and the forward pass against closed forms (zero weights give y = b_out; a single-node graph
with identity-like weights reduces to a hand-computable chain) in tests/test_mpnn_oracle.py.
"""
from __future__ import annotations

import numpy as np

F, H, L = 8, 64, 3
LAYER = 2 * H * H + H
W_IN, B_IN = 0, F * H
LAYER0 = B_IN + H
W_OUT = LAYER0 + L * LAYER
B_OUT = W_OUT + H
NPARAMS = B_OUT + 1


def bf16_round(p: np.ndarray) -> np.ndarray:
    """fp32 -> bf16 -> fp32 with round-to-nearest-even (inf/nan kept)."""
    u = np.ascontiguousarray(p, np.float32).view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(special, u & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32)


def unpack(p: np.ndarray) -> dict:
    p = np.asarray(p, np.float64)
    assert p.size == NPARAMS
    w = {"w_in": p[W_IN:W_IN + F * H].reshape(F, H), "b_in": p[B_IN:B_IN + H]}
    for l in range(L):
        b = LAYER0 + l * LAYER
        w[f"ws{l}"] = p[b:b + H * H].reshape(H, H)
        w[f"wn{l}"] = p[b + H * H:b + 2 * H * H].reshape(H, H)
        w[f"b{l}"] = p[b + 2 * H * H:b + LAYER]
    w["w_out"] = p[W_OUT:W_OUT + H]
    w["b_out"] = p[B_OUT]
    return w


def adjacency(node_off, adj_off, adj) -> np.ndarray:
    """Dense N x N matrix A with A[n, m] = multiplicity of m in n's list (the aggregation adds the
    neighbours' states)."""
    N = int(node_off[-1])
    A = np.zeros((N, N))
    for n in range(N):
        for e in range(int(adj_off[n]), int(adj_off[n + 1])):
            A[n, int(adj[e])] += 1.0
    return A


def forward(params, node_off, x, adj_off, adj, quantize: bool = True):
    """Returns (y[B], cache) in float64."""
    w = unpack(bf16_round(params) if quantize else params)
    x = np.asarray(x, np.float64).reshape(-1, F)
    A = adjacency(node_off, adj_off, adj)
    z0 = x @ w["w_in"] + w["b_in"]
    hs = [np.maximum(z0, 0.0)]
    ags = []
    for l in range(L):
        a = A @ hs[l]
        ags.append(a)
        z = hs[l] @ w[f"ws{l}"] + a @ w[f"wn{l}"] + w[f"b{l}"]
        hs.append(np.maximum(z, 0.0))
    B = len(node_off) - 1
    seg = np.zeros((B, int(node_off[-1])))
    for g in range(B):
        seg[g, int(node_off[g]):int(node_off[g + 1])] = 1.0
    r = seg @ hs[L]
    y = r @ w["w_out"] + w["b_out"]
    return y, {"w": w, "x": x, "A": A, "hs": hs, "ags": ags, "seg": seg, "r": r}


def grad(params, node_off, x, adj_off, adj, targets, quantize: bool = True):
    """Gradient of 0.5 * mean (y - t)^2 with respect to the (bf16) weights, float64; returns
    (grad[NPARAMS], y)."""
    y, c = forward(params, node_off, x, adj_off, adj, quantize)
    w, hs, ags, A = c["w"], c["hs"], c["ags"], c["A"]
    B = len(y)
    dy = (y - np.asarray(targets, np.float64)) / B
    g = np.zeros(NPARAMS)
    g[W_OUT:W_OUT + H] = dy @ c["r"]
    g[B_OUT] = dy.sum()
    dh = c["seg"].T @ np.outer(dy, w["w_out"])  # [N, H]
    for l in reversed(range(L)):
        dz = dh * (hs[l + 1] > 0)
        b = LAYER0 + l * LAYER
        g[b:b + H * H] = (hs[l].T @ dz).ravel()
        g[b + H * H:b + 2 * H * H] = (ags[l].T @ dz).ravel()
        g[b + 2 * H * H:b + LAYER] = dz.sum(0)
        dh = dz @ w[f"ws{l}"].T + A.T @ (dz @ w[f"wn{l}"].T)
    dz = dh * (hs[0] > 0)
    g[W_IN:W_IN + F * H] = (c["x"].T @ dz).ravel()
    g[B_IN:B_IN + H] = dz.sum(0)
    return g, y


def init_params(seed: int) -> np.ndarray:
    """A random parameter vector with the same scale as the library's initialisation (uniform
    +-sqrt(6 / (fan_in + H)) on the matrices, zero biases); not bit-identical to it."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, np.float32)

    def fill(off, fan_in, cnt):
        a = np.sqrt(6.0 / (fan_in + H))
        p[off:off + cnt] = rng.uniform(-a, a, cnt)

    fill(W_IN, F, F * H)
    for l in range(L):
        fill(LAYER0 + l * LAYER, H, H * H)
        fill(LAYER0 + l * LAYER + H * H, H, H * H)
    fill(W_OUT, H, H)
    return p


def random_batch(rng: np.random.Generator, B: int, max_nodes: int = 12):
    """Random small connected graphs (a random tree plus extra edges) with random features."""
    node_off, xs, adj_off, adj = [0], [], [0], []
    for _ in range(B):
        k = int(rng.integers(1, max_nodes + 1))
        edges = set()
        for v in range(1, k):
            edges.add((int(rng.integers(0, v)), v))
        for _ in range(int(rng.integers(0, k))):
            a, b = sorted(rng.choice(k, 2, replace=False)) if k > 1 else (0, 0)
            if a != b:
                edges.add((int(a), int(b)))
        nb = [[] for _ in range(k)]
        for a, b in edges:
            nb[a].append(b)
            nb[b].append(a)
        base = node_off[-1]
        for v in range(k):
            adj.extend(base + u for u in sorted(nb[v]))
            adj_off.append(len(adj))
        xs.append(rng.uniform(0.0, 1.0, (k, F)))
        node_off.append(base + k)
    return (np.array(node_off, np.uint32), np.concatenate(xs).astype(np.float32), np.array(adj_off, np.uint32),
            np.array(adj, np.uint32))


def q_error(est, true) -> np.ndarray:
    """max(est/true, true/est) with both sides floored at 1 (counts of 0 compare as 1)."""
    e = np.maximum(np.asarray(est, np.float64), 1.0)
    t = np.maximum(np.asarray(true, np.float64), 1.0)
    return np.maximum(e / t, t / e)
