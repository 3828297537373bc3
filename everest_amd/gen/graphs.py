"""Deterministic synthetic temporal graphs for the benchmark workloads.

All draws come from ``numpy.random.Generator(PCG64(seed)).random`` (uniform
doubles) and are transformed here by inverse CDFs, Box-Muller and -log(1-u),
so a seed fixes the graph independently of numpy's higher-level samplers.
Edges are returned in generation order, which is also the edge-file line
order and therefore fixes the order of equal timestamps after the stable time
sort.

zipf_uniform_graph
    10k nodes / 100k edges; both endpoints Zipf(alpha=2.0) over ranks 1..N
    (P(r) ~ r^-alpha, node id = r-1); v redrawn while v == u; times uniform
    integers over 30 days.
powerlaw_bursty_graph
    Chung-Lu power-law graph: independent out/in weights w ~ Zipf(alpha)
    (discretised Pareto, P(W >= k) = k^-(alpha-1)) over permuted node ids;
    u ~ w_out, v ~ w_in, self-loops redrawn, repeated (u,v) pairs kept.
    Times: every source u runs a Poisson process of rate
    lambda_u ~ LogNormal(0, 1.5) events/day from a uniform random phase,
    wrapped at the span T; its out-edges take the process's event times in
    generation order.
"""
import hashlib

import numpy as np

DAY = 86400


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _inv_cdf_sample(rng, cum, n):
    """Sample indices i with P(i) = (cum[i]-cum[i-1]) / cum[-1]."""
    u = rng.random(n) * cum[-1]
    idx = np.searchsorted(cum, u, side="right")
    return np.minimum(idx, cum.size - 1)


def zipf_uniform_graph(num_nodes=10_000, num_edges=100_000, alpha=2.0, days=30, seed=1):
    rng = _rng(seed)
    w = np.arange(1, num_nodes + 1, dtype=np.float64) ** (-alpha)
    cum = np.cumsum(w)
    u = _inv_cdf_sample(rng, cum, num_edges)
    v = _inv_cdf_sample(rng, cum, num_edges)
    bad = np.nonzero(u == v)[0]
    while bad.size:
        v[bad] = _inv_cdf_sample(rng, cum, bad.size)
        bad = bad[u[bad] == v[bad]]
    span = days * DAY
    t = np.floor(rng.random(num_edges) * span).astype(np.int64)
    t = np.minimum(t, span - 1)
    return u.astype(np.int32), v.astype(np.int32), t.astype(np.int32)


def _zipf_weights(rng, n, alpha, cap):
    u = rng.random(n)
    w = np.floor((1.0 - u) ** (-1.0 / (alpha - 1.0)))
    return np.minimum(w, cap)


def powerlaw_bursty_graph(num_nodes, num_edges, years, alpha=2.1, sigma=1.5, seed=2):
    rng = _rng(seed)
    span = int(round(years * 365 * DAY))
    cap = float(num_nodes)
    w_out = _zipf_weights(rng, num_nodes, alpha, cap)[rng.permutation(num_nodes)]
    w_in = _zipf_weights(rng, num_nodes, alpha, cap)[rng.permutation(num_nodes)]
    cum_out = np.cumsum(w_out)
    cum_in = np.cumsum(w_in)
    u = _inv_cdf_sample(rng, cum_out, num_edges)
    v = _inv_cdf_sample(rng, cum_in, num_edges)
    bad = np.nonzero(u == v)[0]
    while bad.size:
        v[bad] = _inv_cdf_sample(rng, cum_in, bad.size)
        bad = bad[u[bad] == v[bad]]
    # per-source rate (events/day): LogNormal(0, sigma) via Box-Muller
    u1 = rng.random(num_nodes)
    u2 = rng.random(num_nodes)
    z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    lam = np.exp(sigma * z)
    phase = rng.random(num_nodes) * span
    gaps = -np.log1p(-rng.random(num_edges)) * (DAY / lam[u])
    order = np.argsort(u, kind="stable")
    g_sorted = gaps[order]
    cs = np.cumsum(g_sorted)
    u_sorted = u[order]
    starts = np.flatnonzero(np.r_[True, u_sorted[1:] != u_sorted[:-1]])
    seg_len = np.diff(np.r_[starts, num_edges])
    base = np.repeat(cs[starts] - g_sorted[starts], seg_len)
    t_sorted = np.fmod(phase[u_sorted] + (cs - base), float(span))
    t = np.empty(num_edges, dtype=np.float64)
    t[order] = t_sorted
    t = np.minimum(np.floor(t), span - 1).astype(np.int32)
    return u.astype(np.int32), v.astype(np.int32), t


CONFIGS = {
    # name: (factory, kwargs)
    "zipf100k": (zipf_uniform_graph, dict(num_nodes=10_000, num_edges=100_000, alpha=2.0, days=30, seed=1)),
    "pl8m": (powerlaw_bursty_graph, dict(num_nodes=1_000_000, num_edges=8_000_000, years=6.0, seed=2)),
    "pl64m": (powerlaw_bursty_graph, dict(num_nodes=2_600_000, num_edges=64_000_000, years=7.5, seed=3)),
}


def make(name, **over):
    fn, kw = CONFIGS[name]
    kw = dict(kw)
    kw.update(over)
    return fn(**kw)


def edge_text(u, v, t):
    """Edge-file text (``u v t`` per line), the reference input format."""
    a = np.stack([u, v, t], axis=1)
    return "\n".join(" ".join(map(str, r)) for r in a.tolist()) + "\n"


def write_edge_file(path, u, v, t):
    a = np.stack([u, v, t], axis=1).astype(np.int64)
    np.savetxt(path, a, fmt="%d", delimiter=" ")


def sha256_arrays(u, v, t):
    h = hashlib.sha256()
    for a in (u, v, t):
        h.update(np.ascontiguousarray(a, dtype=np.int32).tobytes())
    return h.hexdigest()
