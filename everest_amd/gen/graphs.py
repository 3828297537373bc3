"""Deterministic synthetic temporal graphs for the benchmark workloads.

Thin wrapper over the C++ generators (include/everest/gen.h, exposed as
corelib.gen): counter-based random numbers, so a seed fixes the graph for any
thread count.  Edges come back in generation order (= edge-file line order).

zipf100k  10k nodes, 100k edges; endpoints Zipf(2.0) over ranks, no
          self-loops; times uniform over 30 days.
pl8m      1M nodes, 8M edges; Chung-Lu power law, Zipf(2.1) out/in weights
          capped at the natural cutoff N^(1/(alpha-1));
          bursty per-source Poisson times (rate LogNormal(0, 1.5)/day) over
          6 years.
pl64m     2.6M nodes, 64M edges; same generator over 7.5 years.
zipf8m    1M nodes, 8M edges; endpoints Zipf(1.0) over ranks (hubs of about
          550k edges), times uniform over one year: dense in short temporal
          cycles, like the reference's social-network inputs, for the
          constraint-feature queries.
"""
import hashlib

import numpy as np

DAY = 86400
YEAR = 365 * DAY

CONFIGS = {
    "zipf100k": dict(kind="zipf", N=10_000, E=100_000, alpha=2.0, span=30 * DAY, seed=1),
    "pl8m": dict(kind="pl", N=1_000_000, E=8_000_000, alpha=2.1, sigma=1.5, span=6 * YEAR, seed=2),
    "pl64m": dict(kind="pl", N=2_600_000, E=64_000_000, alpha=2.1, sigma=1.5, span=int(7.5 * YEAR), seed=3),
    "zipf8m": dict(kind="zipf", N=1_000_000, E=8_000_000, alpha=1.0, span=YEAR, seed=7),
}


def _gen():
    from everest_amd import core
    return core.corelib.gen


def make(name, **over):
    """Return (U, V, T) int32 arrays of a named configuration."""
    kw = dict(CONFIGS[name])
    kw.update(over)
    g = _gen()
    if kw["kind"] == "zipf":
        return g.zipf_uniform(kw["N"], kw["E"], kw["alpha"], kw["span"], kw["seed"])
    return g.powerlaw_bursty(kw["N"], kw["E"], kw["alpha"], kw["sigma"], kw["span"], kw["seed"])


def write_edge_file(path, u, v, t):
    """Write the reference input format: one "u v t" line per edge."""
    a = np.stack([u, v, t], axis=1).astype(np.int64)
    np.savetxt(path, a, fmt="%d", delimiter=" ")


def sha256_arrays(u, v, t):
    h = hashlib.sha256()
    for a in (u, v, t):
        h.update(np.ascontiguousarray(a, dtype=np.int32).tobytes())
    return h.hexdigest()
