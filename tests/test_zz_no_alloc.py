"""The entry point that allocates no device memory (tm_batch_count_noalloc,
buffers in a static device array) against tm_batch_count: two calls in a
row, the second smaller, bit-exact; a third call with a forced-donation
launch shape that fills the smaller ring several times over; sizes beyond
the array's maximum are refused before any GPU work."""
import sys

import numpy as np
import pytest

from oracle import oracle, prep
from helpers import random_motif

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_noalloc_batch_equals_batch():
    from everest_amd import abi
    from everest_amd.gen import graphs
    U, V, T = graphs.make("zipf100k")
    g = prep.preprocess(U, V, T)
    dg = abi.DeviceGraph.from_dict(g)
    rng = np.random.default_rng(71)
    # 2-5 edges on at most 5 nodes, windows of 5-50 minutes: about 2.5 s of
    # 8-thread oracle work for the whole set (long windows on this graph's hubs
    # cost minutes of GPU time per query)
    qs = [(random_motif(rng, int(rng.integers(2, 6)), max_nodes=5), int(rng.integers(300, 3000))) for _ in range(60)]
    say("noalloc: call 1, %d queries" % len(qs))
    want = dg.batch_count(qs)
    assert dg.batch_count_noalloc(qs) == want
    small = qs[10:17]
    say("noalloc: call 2, %d queries" % len(small))
    assert dg.batch_count_noalloc(small) == want[10:17]
    say("noalloc: call 3, forced donation")
    force = abi.tuning(steal_after_us=1, steal_every=1, steal_min=1, steal_max=4)
    assert dg.batch_count_noalloc(qs[:20], tune=force) == want[:20]
    say("noalloc: the oracle on five of them")
    for i in range(5):
        assert want[i] == oracle.count(g, qs[i][0], qs[i][1], nthreads=8)
    say("noalloc: beyond the maximum")
    with pytest.raises(abi.TMError) as ei:
        dg.batch_count_noalloc(qs * 5)
    assert ei.value.status == 1  # TM_ERR_ARG
    assert dg.batch_count_noalloc(qs[:3]) == want[:3]
    dg.close()
