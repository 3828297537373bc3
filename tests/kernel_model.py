"""CPU model of the gfx950 mining kernel's wave-step algorithm.

Mirrors everest_amd/csrc/hip/tm_kernel.cuh step by step for one wavefront
(64 lanes simulated with Python lists): the device-side indexes built at
upload (packed times, neighbour arrays, pair index, time buckets), root
dispatch, the LDS frame stack with its ballot/prefix-scan batching, child
ranges and last-edge range arithmetic.  Every array access is bounds-checked
and the frame stack is checked against its capacity, so an indexing mistake in
the design shows up here on the CPU before the kernel runs on a GPU.

It is a model of the product's algorithm, checked against the oracle; it is not
the oracle.
"""
import numpy as np

WAVE = 64
K_NEW_DST, K_NEW_SRC, K_CLOSE = 0, 1, 2


class Arr:
    """Bounds-checked read-only array."""

    def __init__(self, a, name):
        self.a = np.asarray(a)
        self.name = name

    def __getitem__(self, i):
        if not (0 <= i < self.a.size):
            raise IndexError("%s[%d] out of [0,%d)" % (self.name, i, self.a.size))
        return int(self.a[i])


def build_view(g):
    """Device-side indexes as tm_graph_upload builds them."""
    E, V = g["E"], g["V"]
    Eg = g["Eg"]
    T = Eg[:, 2].copy()
    outPos, inPos = g["outV"], g["inV"]
    outNbr = Eg[outPos, 1]
    inNbr = Eg[inPos, 0]
    okey = Eg[outPos, 0].astype(np.int64) * (1 << 32) + Eg[outPos, 1].astype(np.int64)
    o = np.argsort(okey, kind="stable")
    ikey = Eg[inPos, 1].astype(np.int64) * (1 << 32) + Eg[inPos, 0].astype(np.int64)
    i = np.argsort(ikey, kind="stable")
    tmin, tmax = int(T[0]), int(T[-1])
    span = tmax - tmin + 1
    x = max(1, (span * 16) // max(1, E))
    bshift = 0
    while (1 << bshift) < x:
        bshift += 1
    bshift = min(bshift, 30)
    nb = (span + (1 << bshift) - 1) >> bshift
    keys = tmin + (np.arange(nb + 1, dtype=np.int64) << bshift)
    bucket = np.searchsorted(T.astype(np.int64), keys, side="left")
    outPairPos = outPos[o]
    outSelf = np.empty(E, dtype=np.int64); outSelf[outPos] = np.arange(E)
    inSelf = np.empty(E, dtype=np.int64); inSelf[inPos] = np.arange(E)
    outPairSelf = np.empty(E, dtype=np.int64); outPairSelf[outPairPos] = np.arange(E)
    outR, inR = g["outR"], g["inR"]
    outNext = np.array([outR[Eg[p, 1]] + np.searchsorted(outPos[outR[Eg[p, 1]]:outR[Eg[p, 1] + 1]], p, side="right")
                        for p in range(E)], dtype=np.int64)
    inNext = np.array([inR[Eg[p, 0]] + np.searchsorted(inPos[inR[Eg[p, 0]]:inR[Eg[p, 0] + 1]], p, side="right")
                       for p in range(E)], dtype=np.int64)
    NONE = 2**31 - 1
    tf = np.full((6, E), NONE, dtype=np.int64)
    outPairNbr = outNbr[o]
    for p in range(E):
        u_, v_ = int(Eg[p, 0]), int(Eg[p, 1])
        s_ = outSelf[p] + 1
        if s_ < outR[u_ + 1]:
            tf[0, p] = T[outPos[s_]]
        s_ = outNext[p]
        if s_ < outR[v_ + 1]:
            tf[1, p] = T[outPos[s_]]
        s_ = inSelf[p] + 1
        if s_ < inR[v_ + 1]:
            tf[2, p] = T[inPos[s_]]
        s_ = inNext[p]
        if s_ < inR[u_ + 1]:
            tf[3, p] = T[inPos[s_]]
        s_ = outPairSelf[p] + 1
        if s_ < outR[u_ + 1] and outPairNbr[s_] == v_:
            tf[4, p] = T[outPairPos[s_]]
        rec = [outPairPos[k] for k in range(outR[v_], outR[v_ + 1]) if outPairNbr[k] == u_ and outPairPos[k] > p]
        if rec:
            tf[5, p] = T[min(rec)]
    # first-candidate times in the slot order of each candidate list
    # (tm_graph_view.stf, lists in TM_L_* order: out, in, out-pair, in-pair)
    lists = [outPos, inPos, outPos[o], inPos[i]]
    stf = [[Arr(tf[k][np.asarray(lst)], "stf%d_%d" % (l, k)) for k in range(6)] for l, lst in enumerate(lists)]
    # root gaps (tm_graph_view.rgap): first-candidate time minus the root's
    # time, clamped to INT_MAX; 0xffffffff for self-loops and no candidate;
    # padded to a multiple of 256 entries plus 256 with 0xffffffff
    epad = (E + 255) // 256 * 256 + 256
    rg = np.full((6, epad), 0xFFFFFFFF, dtype=np.int64)
    ok = (Eg[:, 0] != Eg[:, 1])[None, :] & (tf != NONE)
    d = np.minimum(tf - T[None, :].astype(np.int64), 2**31 - 1)
    rg[:, :E][ok] = d[ok]
    return dict(E=E, V=V, Eg=Eg, T=Arr(T, "T"), tfirst=[Arr(tf[k], "tfirst%d" % k) for k in range(6)],
                rgap=[Arr(rg[k], "rgap%d" % k) for k in range(6)], epad=epad, stf=stf,
                outSelf=Arr(outSelf, "outSelf"), inSelf=Arr(inSelf, "inSelf"),
                outNext=Arr(outNext, "outNext"), inNext=Arr(inNext, "inNext"),
                outPairSelf=Arr(outPairSelf, "outPairSelf"),
                outR=Arr(g["outR"], "outR"), outPos=Arr(outPos, "outPos"), outNbr=Arr(outNbr, "outNbr"),
                inR=Arr(g["inR"], "inR"), inPos=Arr(inPos, "inPos"), inNbr=Arr(inNbr, "inNbr"),
                outPairNbr=Arr(outNbr[o], "outPairNbr"), outPairPos=Arr(outPos[o], "outPairPos"),
                inPairNbr=Arr(inNbr[i], "inPairNbr"), inPairPos=Arr(inPos[i], "inPairPos"),
                bucket=Arr(bucket, "bucket"), tmin=tmin, bshift=bshift, nbuckets=nb,
                Euid=Arr(g["Euid"], "Euid"))


S_SEARCH, S_OUT_SELF, S_OUT_NEXT, S_IN_SELF, S_IN_NEXT, S_PAIR_SELF = range(6)
TM_BLOCK = 4 * WAVE
SB_FLUSH = 32  # survivors that trigger a full-width expansion (counting kernel; 0: none buffered)


L_OUT, L_IN, L_OUTPAIR, L_INPAIR = range(4)


def next_first_time(g, md, j, slot, direction, M):
    """First-candidate time of motif edge j + 1 for the candidate at `slot`
    of edge j's list, read from the slot-ordered copy (None: no rule)."""
    tm = md["tmode"][j + 1] if j + 1 < M else -1
    if tm < 0:
        return None
    kind = md["kind"][j]
    lst = L_OUT if kind == K_NEW_DST else (L_IN if kind == K_NEW_SRC else (L_INPAIR if direction else L_OUTPAIR))
    return g["stf"][lst][tm][slot]


def root_passes(g, md, delta, r):
    """Root filter of the dispatch: the root's gap under edge 1's start rule
    is within delta (unsigned compare; self-loops carry 0xffffffff)."""
    ok = delta >= 0 and g["rgap"][md["tmode"][1]][r] <= delta
    # the root's further rules (later edges at the root's nodes)
    return ok and all(g["rgap"][x][r] <= delta for x in range(6) if (md["xmask"][0] >> x) & 1)


def second_ok(g, md, j, slot, direction, tl):
    """The candidate of edge j at `slot` passes edge j's further rules: for
    each later edge k >= j + 2 meeting a node of edge j, that node's next edge
    of the right direction after the candidate's position is in the window."""
    xm = md["xmask"][j] if j < len(md["xmask"]) else 0
    kind = md["kind"][j]
    lst = L_OUT if kind == K_NEW_DST else (L_IN if kind == K_NEW_SRC else (L_INPAIR if direction else L_OUTPAIR))
    return all(g["stf"][lst][x][slot] <= tl for x in range(6) if (xm >> x) & 1)


STRIPES = 8


def claim_blocks(total_blocks, nwaves, per_claim, rng):
    """Fixed-size claims over striped counters, as the kernel claims root
    blocks: stripe k holds blocks [k B / S, (k+1) B / S); a wavefront claims
    per_claim blocks from its own stripe first and moves on when it runs
    dry.  Wavefronts take turns in a random order; returns the claims as
    (first, end) block ranges."""
    counters = [0] * STRIPES
    waves = [dict(home=w % STRIPES, off=0) for w in range(nwaves)]
    claims = []
    live = list(range(nwaves))
    while live:
        wi = live[int(rng.integers(0, len(live)))]
        w = waves[wi]
        got = False
        while w["off"] < STRIPES:
            k = (w["home"] + w["off"]) % STRIPES
            sb, se = total_blocks * k // STRIPES, total_blocks * (k + 1) // STRIPES
            o = sb + counters[k]
            counters[k] += per_claim
            if o < se:
                claims.append((o, min(o + per_claim, se)))
                got = True
                break
            w["off"] += 1
        if not got:
            live.remove(wi)
    return claims


def batch_dispatch(g, queries, nwaves=5, per_claim=3, seed=0, instr=False):
    """Model of the root dispatch of one launch over several queries.

    queries: list of (motif descriptor, delta, root_beg, root_end).  Mirrors
    the host plan (one space of 256-root blocks, query after query; query
    q's block i starts at root (root_beg & ~3) + 256 i) and the kernel's
    walk: blocks claimed in fixed-size claims, the block's query found
    by a search over the block prefix, the block read as one aligned 16-byte
    load of gaps per lane, skipped when no gap is within delta, its quarters
    expanded when some lane passes.  Returns the (query, root) pairs
    expanded with that lane passing (instr: every in-range lane)."""
    rng = np.random.default_rng(seed)
    nq = len(queries)
    prefix = [0]
    for md, d, b, e in queries:
        assert 0 <= md["tmode"][1] < 6
        prefix.append(prefix[-1] + ((e - (b & ~3) + TM_BLOCK - 1) // TM_BLOCK if e > b else 0))
    claims = claim_blocks(prefix[-1], nwaves, per_claim, rng)
    seen = sorted(x for c in claims for x in range(*c))
    assert seen == list(range(prefix[-1])), "every block claimed exactly once"
    out = []
    for c0, c1 in claims:
        for bi in range(c0, c1):
            lo, hi = 0, nq
            while hi - lo > 1:
                mid = (lo + hi) >> 1
                if prefix[mid] <= bi:
                    lo = mid
                else:
                    hi = mid
            q = lo
            assert prefix[q] <= bi < prefix[q + 1]
            md, d, b, e = queries[q]
            rule = md["tmode"][1]
            blk = (b & ~3) + (bi - prefix[q]) * TM_BLOCK
            gaps = []
            for lane in range(WAVE):
                r0 = blk + 4 * lane
                assert r0 % 4 == 0 and r0 + 3 < g["epad"]  # one aligned 16-byte load
                xm = md["xmask"][0] if not instr else 0
                gaps.append([max([g["rgap"][rule][r0 + i]] + [g["rgap"][x][r0 + i] for x in range(6) if (xm >> x) & 1])
                             if b <= r0 + i < e else 0xFFFFFFFF for i in range(4)])
            near = any(d >= 0 and min(x) <= d for x in gaps)
            if not near and not instr:
                continue
            for quarter in range(4):
                for lane in range(WAVE):
                    r = blk + 4 * lane + quarter
                    inq = b <= r < e
                    ps = inq and d >= 0 and gaps[lane][quarter] <= d
                    if ps or (instr and inq):
                        out.append((q, r))
    return out


def build_motif(edges):
    """tmh::build_motif: kinds / node slots / start rules per edge."""
    M = len(edges)
    assert edges[0][0] == 0 and edges[0][1] == 1
    kind, A, B, mapped, smode = [0] * M, [0] * M, [0] * M, [0] * M, [0] * M
    tmode = [-1] * M
    n = 2
    for j in range(1, M):
        a, b = edges[j][0], edges[j][1]
        pa, pb = edges[j - 1][0], edges[j - 1][1]
        assert a != b
        A[j], B[j], mapped[j] = a, b, n
        if a < n and b < n:
            kind[j] = K_CLOSE
            smode[j] = S_PAIR_SELF if (a, b) == (pa, pb) else S_SEARCH
            # the pair's own next edge, else a shared node's next edge of the right direction
            tmode[j] = next(r for r, c in ((4, (a, b) == (pa, pb)), (5, (a, b) == (pb, pa)), (0, a == pa),
                                           (1, a == pb), (2, b == pb), (3, b == pa), (-1, True)) if c)
        elif a < n and b == n:
            kind[j] = K_NEW_DST
            smode[j] = S_OUT_SELF if a == pa else (S_OUT_NEXT if a == pb else S_SEARCH)
            tmode[j] = 0 if a == pa else (1 if a == pb else -1)
            n += 1
        elif b < n and a == n:
            kind[j] = K_NEW_SRC
            smode[j] = S_IN_SELF if b == pb else (S_IN_NEXT if b == pa else S_SEARCH)
            tmode[j] = 2 if b == pb else (3 if b == pa else -1)
            n += 1
        else:
            raise ValueError("not prefix-connected")
    # second rule per edge j: a later edge k >= j + 2 at a node of edge j (pair rules preferred)
    def rule_at(pa, pb, a, b):
        return next(r for r, c in ((4, (a, b) == (pa, pb)), (5, (a, b) == (pb, pa)), (0, a == pa), (1, a == pb),
                                   (2, b == pb), (3, b == pa), (-1, True)) if c)
    xmask = [0] * M
    for j in range(M - 2):
        mask = 0
        for k in range(j + 2, M):
            r = rule_at(edges[j][0], edges[j][1], edges[k][0], edges[k][1])
            if r >= 0:
                mask |= 1 << r
        nxt = (1 << tmode[j + 1]) if tmode[j + 1] >= 0 else 0
        if (mask | nxt) & (1 << 4):
            mask &= ~0b101   # pair rule implies the source's next out- and destination's next in-edge
        if (mask | nxt) & (1 << 5):
            mask &= ~0b1010  # reversed pair implies the other two
        xmask[j] = mask & ~nxt
    return dict(M=M, n=n, kind=kind, a=A, b=B, mapped=mapped, smode=smode, tmode=tmode, xmask=xmask)


def lower_bound(a, lo, hi, x):
    while lo < hi:
        mid = (lo + hi) >> 1
        if a[mid] < x:
            lo = mid + 1
        else:
            hi = mid
    return lo


def upper_bound(a, lo, hi, x):
    while lo < hi:
        mid = (lo + hi) >> 1
        if a[mid] <= x:
            lo = mid + 1
        else:
            hi = mid
    return lo


def pend_of(g, tl):
    k = (tl - g["tmin"]) >> g["bshift"]
    if k < 0:
        return 0
    if k >= g["nbuckets"]:
        return g["E"]
    lo, hi = g["bucket"][k], g["bucket"][k + 1]
    return upper_bound(g["T"], lo, hi, tl)


def gallop_lb(a, lo, hi, x):
    if lo >= hi or a[lo] >= x:
        return lo
    prev, step = lo, 1
    while True:
        nxt = prev + step
        if nxt >= hi:
            return lower_bound(a, prev + 1, hi, x)
        if a[nxt] >= x:
            return lower_bound(a, prev + 1, nxt, x)
        prev = nxt
        step <<= 1


def gallop_ub(a, lo, hi, x):
    if lo >= hi or a[lo] > x:
        return lo
    prev, step = lo, 1
    while True:
        nxt = prev + step
        if nxt >= hi:
            return upper_bound(a, prev + 1, hi, x)
        if a[nxt] > x:
            return upper_bound(a, prev + 1, nxt, x)
        prev = nxt
        step <<= 1


def pair_count(pn, pp, lo, hi, key, p, pend):
    s0 = lower_bound(pn, lo, hi, key)
    if s0 >= hi or pn[s0] != key:
        return 0
    s1 = gallop_ub(pn, s0 + 1, hi, key)
    c0 = upper_bound(pp, s0, s1, p)
    c1 = gallop_lb(pp, c0, s1, pend)
    return c1 - c0


def pair_last(g, x, y, p, pend):
    """The last position in (p, pend) of an edge x -> y, or -1."""
    olo, ohi = g["outR"][x], g["outR"][x + 1]
    ilo, ihi = g["inR"][y], g["inR"][y + 1]
    lo, hi, key, pn, pp = olo, ohi, y, g["outPairNbr"], g["outPairPos"]
    if ihi - ilo < ohi - olo:
        lo, hi, key, pn, pp = ilo, ihi, x, g["inPairNbr"], g["inPairPos"]
    s0 = lower_bound(pn, lo, hi, key)
    if s0 >= hi or pn[s0] != key:
        return -1
    s1 = gallop_ub(pn, s0 + 1, hi, key)
    i = lower_bound(pp, s0, s1, pend) - 1
    return pp[i] if i >= s0 and pp[i] > p else -1


def edge_count(g, x, y, p, pend):
    olo, ohi = g["outR"][x], g["outR"][x + 1]
    ilo, ihi = g["inR"][y], g["inR"][y + 1]
    if ohi - olo <= ihi - ilo:
        return pair_count(g["outPairNbr"], g["outPairPos"], olo, ohi, y, p, pend)
    return pair_count(g["inPairNbr"], g["inPairPos"], ilo, ihi, x, p, pend)


def cand_start(g, md, j, m, p):
    """Returns (pos, nbr, lo, hi, dir) as the kernel's CandList."""
    k, mode = md["kind"][j], md["smode"][j]
    if k == K_NEW_DST:
        x = m[md["a"][j]]
        hi = g["outR"][x + 1]
        if mode == S_OUT_SELF:
            lo = g["outSelf"][p] + 1
        elif mode == S_OUT_NEXT:
            lo = g["outNext"][p]
        else:
            lo = upper_bound(g["outPos"], g["outR"][x], hi, p)
        return g["outPos"], g["outNbr"], lo, hi, 0
    if k == K_NEW_SRC:
        y = m[md["b"][j]]
        hi = g["inR"][y + 1]
        if mode == S_IN_SELF:
            lo = g["inSelf"][p] + 1
        elif mode == S_IN_NEXT:
            lo = g["inNext"][p]
        else:
            lo = upper_bound(g["inPos"], g["inR"][y], hi, p)
        return g["inPos"], g["inNbr"], lo, hi, 0
    x, y = m[md["a"][j]], m[md["b"][j]]
    if mode == S_PAIR_SELF:
        lo = g["outPairSelf"][p] + 1
        hi = gallop_ub(g["outPairNbr"], lo, g["outR"][x + 1], y)
        return g["outPairPos"], None, lo, hi, 0
    olo, ohi = g["outR"][x], g["outR"][x + 1]
    ilo, ihi = g["inR"][y], g["inR"][y + 1]
    if ohi - olo <= ihi - ilo:
        s0 = lower_bound(g["outPairNbr"], olo, ohi, y)
        s1 = gallop_ub(g["outPairNbr"], s0, ohi, y)
        pos, d = g["outPairPos"], 0
    else:
        s0 = lower_bound(g["inPairNbr"], ilo, ihi, x)
        s1 = gallop_ub(g["inPairNbr"], s0, ihi, x)
        pos, d = g["inPairPos"], 1
    return pos, None, upper_bound(pos, s0, s1, p), s1, d


def leaf_count(g, md, j, m, L, p, pend, scan_max):
    pos, nbr, lo, hi, _ = L
    b1 = gallop_lb(pos, lo, hi, pend)
    c = b1 - lo
    if c <= 0 or md["kind"][j] == K_CLOSE:
        return max(c, 0)
    nm = md["mapped"][j]
    if c <= scan_max:
        return c - sum(1 for s in range(lo, b1) if nbr[s] in m[:nm])
    out = md["kind"][j] == K_NEW_DST
    base = m[md["a"][j]] if out else m[md["b"][j]]
    for i in range(nm):
        c -= edge_count(g, base, m[i], p, pend) if out else edge_count(g, m[i], base, p, pend)
    return c


def ref_probes(g, md, k, m, p, pend):
    """Reference probes of one resolved list (instrumented kernel variant)."""
    kind = md["kind"][k]
    x, y = m[md["a"][k]], m[md["b"][k]]
    use_out = kind == K_NEW_DST or (kind == K_CLOSE and not (
        g["inR"][y + 1] - g["inR"][y] < g["outR"][x + 1] - g["outR"][x]))
    if use_out:
        pos, lo, hi = g["outPos"], g["outR"][x], g["outR"][x + 1]
    else:
        pos, lo, hi = g["inPos"], g["inR"][y], g["inR"][y + 1]
    lo = upper_bound(pos, lo, hi, p)
    e = gallop_lb(pos, lo, hi, pend)
    return (e - lo) + (1 if e < hi else 0)


SWAPS = [0]  # swapped frames decided so far (tests read it)


def _swap_frame(g, md, f, t2_min, frames, w):
    """The kernel's two-edge-tail decision at a push of a frame of edge
    M-2 (stack index w): returns the frame to push (swapped to the last
    edge's candidates when those are fewer), or None when no match can
    follow.  Checks that the side-table entry w mod 64 is not in use by
    another live swapped frame."""
    nA = f["end"] - f["cur"]
    if t2_min is None or nA < t2_min or not tail2_eligible(md):
        return f
    pos, nbr = _list_arrays(g, md, md["M"] - 2, f["dir"])
    A = (pos, nbr, f["cur"], f["end"], f["dir"])
    bpos, bnbr, blo, bhi, bdir = tail2_lists(g, md, f["m"], A, f["pend"])
    if 2 * (bhi - blo) >= nA:
        return f
    SWAPS[0] += 1
    if bhi <= blo:
        return None
    for i in range(w):
        fi = frames[i]
        if fi is not None and fi.get("swap") and i % WAVE == w % WAVE:
            raise AssertionError("side-table entry %d in use by frame %d" % (w % WAVE, i))
    return dict(f, cur=blo, end=bhi, swap=True, alo=f["cur"], ahi=f["end"], bdir=bdir)


def _list_arrays(g, md, j, direction):
    kind = md["kind"][j]
    if kind == K_NEW_DST:
        return g["outPos"], g["outNbr"]
    if kind == K_NEW_SRC:
        return g["inPos"], g["inNbr"]
    return (g["inPairPos"] if direction else g["outPairPos"]), None


def _swap_item(g, md, f, slot):
    """Count of one dealt candidate of a swapped frame (the kernel's tail2_item)."""
    M = md["M"]
    pos, nbr = _list_arrays(g, md, M - 2, f["dir"])
    A = (pos, nbr, f["alo"], f["ahi"], f["dir"])
    kk = md["kind"][M - 1]
    if tail2_mode(md) == T2_JOIN:
        into_y = md["a"][M - 1] == md["mapped"][M - 2]
        qpos, z = ((g["inPos"][slot], g["inNbr"][slot]) if into_y else (g["outPos"][slot], g["outNbr"][slot]))
    elif kk == K_NEW_DST:
        qpos, z = g["outPos"][slot], g["outNbr"][slot]
    elif kk == K_NEW_SRC:
        qpos, z = g["inPos"][slot], g["inNbr"][slot]
    else:
        qpos, z = (g["inPairPos"][slot] if f["bdir"] else g["outPairPos"][slot]), -1
    return tail2_item(g, md, f["m"], A, qpos, z)


def run_wave(g, md, delta, beg, end, enumerate_=False, scan_max=16, chunk_roots=256, cap=1 << 30, instr=False,
             sb_flush=SB_FLUSH, sb_inline=SB_FLUSH, t2_min=None):
    """Simulate one wavefront mining roots [beg, end) alone; returns
    (count, rows, stats).  With instr, stats["ref"] = [R, X1, X2, P]."""
    M = md["M"]
    MAXN = 8
    lvls = max(1, (M - 1) if enumerate_ else (M - 2))
    FR = WAVE * lvls
    frames = [None] * FR  # each: dict(cur,end,p,pend,lvl,m,e)
    top = 0
    cnt = 0
    rows = []
    stats = dict(steps=0, items=0, max_top=0)
    ref = [max(0, end - beg), 0, 0, 0]
    stats["ref"] = ref
    last = (M - 1) if enumerate_ else (M - 2)
    Eg = g["Eg"]
    nchunks = (end - beg + chunk_roots - 1) // chunk_roots
    for c in range(nchunks):
        first = beg + c * chunk_roots
        groups = min(chunk_roots // WAVE, (end - first + WAVE - 1) // WAVE)
        for grp in range(groups):
            # ---- root dispatch (stack is empty here)
            assert top == 0
            pushes = []
            for lane in range(WAVE):
                r = first + grp * WAVE + lane
                if r >= end:
                    continue
                u, v, t = (int(x) for x in Eg[r])
                if u == v:
                    continue
                m = [0] * MAXN
                m[0], m[1] = u, v
                if not root_passes(g, md, delta, r) and not instr:
                    continue
                L = cand_start(g, md, 1, m, r)
                if instr:
                    ref[2 if md["kind"][1] == K_CLOSE else 1] += 1
                    ref[3] += ref_probes(g, md, 1, m, r, pend_of(g, t + delta))
                if not (L[2] < L[3] and g["T"][L[0][L[2]]] <= t + delta):
                    continue
                pend = pend_of(g, t + delta)
                if pend - r - 1 < M - 1 and not instr:
                    continue
                tl = t + delta
                if (not enumerate_ and not instr and M >= 3 and md["kind"][M - 1] == K_CLOSE
                        and md["a"][M - 1] < 2 and md["b"][M - 1] < 2):
                    # the last edge closes the root's nodes: the window ends at that pair's last edge
                    x = u if md["a"][M - 1] == 0 else v
                    y = u if md["b"][M - 1] == 0 else v
                    pl = pair_last(g, x, y, r, pend)
                    if pl < 0:
                        continue
                    pend, tl = pl + 1, g["T"][pl]
                    if pend - r - 1 < M - 1:
                        continue
                if last == 0:
                    cnt += leaf_count(g, md, 1, m, L, r, pend, scan_max)
                    continue
                cb, d = L[2], L[4]
                ce = gallop_lb(L[0], L[2], L[3], pend)
                if ce > cb:
                    pushes.append(dict(cur=cb, end=ce, p=r, pend=pend, lvl=0, dir=d, m=m[:], e=[r], tl=tl))
            assert top + len(pushes) <= FR
            for f in pushes:
                if not enumerate_ and not instr and M == 3:
                    f = _swap_frame(g, md, f, t2_min, frames, top)
                    if f is None:
                        continue
                frames[top] = f
                top += 1
            # ---- expansion steps until the stack and the survivor buffer drain
            compact = not enumerate_ and not instr and sb_flush > 0  # the counting kernel buffers survivors
            sb, sb_level = [], 0
            while top > 0 or sb:
                stats["max_top"] = max(stats["max_top"], top)
                Ltop = frames[top - 1]["lvl"] if top > 0 else -1
                if compact and sb and (top == 0 or Ltop != sb_level or len(sb) >= sb_flush):
                    # full-width expansion of the buffered survivors
                    assert len(sb) <= WAVE
                    j = sb_level + 1
                    children = []
                    for pos, m, f in sb:
                        Lc = cand_start(g, md, j + 1, m, pos)
                        if j == last:
                            cnt += leaf_count(g, md, j + 1, m, Lc, pos, f["pend"], scan_max)
                            continue
                        ce = gallop_lb(Lc[0], Lc[2], Lc[3], f["pend"])
                        if ce > Lc[2]:
                            children.append(dict(cur=Lc[2], end=ce, p=pos, pend=f["pend"], lvl=j, dir=Lc[4], m=m,
                                                 e=[], tl=f["tl"]))
                    sb = []
                    assert top + len(children) <= FR, "frame stack overflow"
                    for ch in children:
                        if j == M - 3:
                            ch = _swap_frame(g, md, ch, t2_min, frames, top)
                            if ch is None:
                                continue
                        frames[top] = ch
                        top += 1
                    continue
                L = Ltop
                lim = WAVE - len(sb) if compact else WAVE
                fcnt = [0] * WAVE
                for lane in range(WAVE):
                    fi = top - 1 - lane
                    if fi >= 0 and frames[fi]["lvl"] == L:
                        fcnt[lane] = frames[fi]["end"] - frames[fi]["cur"]
                incl = [int(x) for x in np.cumsum(fcnt)]
                excl = [incl[i] - fcnt[i] for i in range(WAVE)]
                total = min(lim, incl[-1])
                part = [fcnt[i] > 0 and excl[i] < lim for i in range(WAVE)]
                nparts = sum(part)
                take = [min(fcnt[i], lim - excl[i]) if part[i] else 0 for i in range(WAVE)]
                last_partial = any(part[i] and take[i] < fcnt[i] for i in range(WAVE))
                sm = 0
                for i in range(WAVE):
                    if part[i]:
                        sm |= 1 << excl[i]
                items = []
                for lane in range(total):
                    rank = bin(sm & ((2 << lane) - 1)).count("1") - 1
                    off = lane - excl[rank]
                    f = frames[top - 1 - rank]
                    slot = f["cur"] + off
                    assert f["cur"] <= slot < f["end"]
                    items.append((slot, f))
                stats["steps"] += 1
                stats["items"] += total
                for i in range(WAVE):
                    if part[i] and take[i] < fcnt[i]:
                        frames[top - 1 - i] = dict(frames[top - 1 - i])
                        frames[top - 1 - i]["cur"] += take[i]
                ntop = top - nparts + (1 if last_partial else 0)
                j = L + 1
                kind = md["kind"][j]
                children = []
                cur = []  # survivors of this step (compaction)
                for slot, f in items:
                    if f.get("swap"):
                        cnt += _swap_item(g, md, f, slot)
                        continue
                    m = list(f["m"])
                    if kind == K_NEW_DST:
                        pos, node = g["outPos"][slot], g["outNbr"][slot]
                        valid = node not in m[:md["mapped"][j]]
                        m[md["b"][j]] = node
                    elif kind == K_NEW_SRC:
                        pos, node = g["inPos"][slot], g["inNbr"][slot]
                        valid = node not in m[:md["mapped"][j]]
                        m[md["a"][j]] = node
                    else:
                        pos = g["inPairPos"][slot] if f["dir"] else g["outPairPos"][slot]
                        valid = True
                    valid = valid and (instr or second_ok(g, md, j, slot, f["dir"], f["tl"]))
                    if not valid:
                        continue
                    if instr and j + 1 < M:
                        ref[2 if md["kind"][j + 1] == K_CLOSE else 1] += 1
                        ref[3] += ref_probes(g, md, j + 1, m, pos, f["pend"])
                    e = f["e"] + [pos]
                    tfn = next_first_time(g, md, j, slot, f["dir"], M)
                    if tfn is not None:
                        assert tfn == g["tfirst"][md["tmode"][j + 1]][pos]
                    empty_next = tfn is not None and tfn > f["tl"]
                    if compact:
                        if not empty_next:
                            cur.append((pos, m, f))
                        continue
                    if j == last:
                        if enumerate_:
                            if cnt < cap:
                                rows.append(tuple(g["Euid"][x] for x in e))
                            cnt += 1
                        elif not empty_next:
                            L = cand_start(g, md, j + 1, m, pos)
                            cnt += leaf_count(g, md, j + 1, m, L, pos, f["pend"], scan_max)
                        continue
                    if empty_next:
                        continue
                    L = cand_start(g, md, j + 1, m, pos)
                    cb, d = L[2], L[4]
                    ce = gallop_lb(L[0], L[2], L[3], f["pend"])
                    if ce > cb:
                        children.append(dict(cur=cb, end=ce, p=pos, pend=f["pend"], lvl=j, dir=d, m=m, e=e, tl=f["tl"]))
                if compact:
                    # buffered while few and more of this level follows, else
                    # expanded in this step (nothing buffered)
                    level_done = ntop == 0 or frames[ntop - 1]["lvl"] != L
                    if sb or (len(cur) < sb_inline and not level_done):
                        sb += cur
                        sb_level = L
                        assert len(sb) <= WAVE
                        top = ntop
                        stats["buffered_steps"] = stats.get("buffered_steps", 0) + 1
                        continue
                    for pos, m, f in cur:
                        Lc = cand_start(g, md, j + 1, m, pos)
                        if j == last:
                            cnt += leaf_count(g, md, j + 1, m, Lc, pos, f["pend"], scan_max)
                            continue
                        ce = gallop_lb(Lc[0], Lc[2], Lc[3], f["pend"])
                        if ce > Lc[2]:
                            children.append(dict(cur=Lc[2], end=ce, p=pos, pend=f["pend"], lvl=j, dir=Lc[4], m=m,
                                                 e=[], tl=f["tl"]))
                assert ntop + len(children) <= FR, "frame stack overflow"
                top = ntop
                for ch in children:
                    if not enumerate_ and not instr and j == M - 3:
                        ch = _swap_frame(g, md, ch, t2_min, frames, top)
                        if ch is None:
                            continue
                    frames[top] = ch
                    top += 1
    return cnt, rows, stats


# ------------------------------------------------------------ many wavefronts

STRIPES = 8


class Ring:
    """The donation ring and its release words, kept across launches as the
    device keeps them: records are tagged with the launch generation, so a
    record or release left by an earlier launch never matches."""

    def __init__(self, pool_cap, gen_bits=18):
        assert pool_cap & (pool_cap - 1) == 0
        self.cap = pool_cap
        self.slots = [None] * pool_cap   # slot -> (generation, record index, frame)
        self.done = [None] * pool_cap    # slot -> (generation, lap) of the last record taken
        self.gen = 0                     # device generation (advanced by each launch's reset)
        self.host_gen = 0                # launches since the ring was last cleared (host count)
        self.gen_bits = gen_bits

    def launch(self):
        """Host side of a launch (count_launch) then the reset kernel's increment."""
        if self.host_gen + 1 >= (1 << self.gen_bits):
            # generations wrap: clear the ring so no stale tag can match
            self.slots = [None] * self.cap
            self.done = [None] * self.cap
            self.gen = 0
            self.host_gen = 0
        self.host_gen += 1
        self.gen = (self.gen + 1) % (1 << self.gen_bits)


class _Shared:
    def __init__(self, nchunks, pool_cap, ring=None):
        self.next_chunk = [0] * STRIPES   # striped chunk counters
        self.nchunks = nchunks
        if ring is None:
            ring = Ring(pool_cap)
            ring.launch()
        self.ringst = ring
        self.cap = ring.cap
        self.head = 0           # tickets drawn
        self.tail = 0           # records reserved
        self.ndonated = 0
        self.taken = set()
        self.waiters = 0
        self.idle_cap = 1 << 30
        self.busy = 0
        self.idle = 0
        self.abort = False      # a protocol wait gave up (error bit 8): every wavefront leaves
        self.error = 0
        self.spin_limit = None  # bounded wait of a donor on its ring slot (None: unbounded)

    def record(self, slot):
        return self.ringst.slots[slot]

    def done_lap(self, slot):
        d = self.ringst.done[slot]
        return d[1] if d is not None and d[0] == self.ringst.gen else -1


class Wave:
    """One wavefront of mine_kernel as a step machine (one loop iteration per
    step), sharing the chunk counters and the donation pool with the others."""
    count = 0

    def __init__(self, g, md, delta, beg, end, sh, scan_max, chunk_roots, steal_every, steal_max, steal_min=1):
        self.g, self.md, self.delta, self.beg, self.end = g, md, delta, beg, end
        self.sh, self.scan_max, self.chunk_roots = sh, scan_max, chunk_roots
        self.steal_every, self.steal_max, self.steal_min = steal_every, steal_max, steal_min
        M = md["M"]
        self.FR = WAVE * max(1, M - 2)
        self.frames = [None] * self.FR
        self.top = 0
        self.cnt = 0
        self.chunk_left = 0
        self.next_root = 0
        self.roots_done = False
        self.is_idle = False
        self.ticket = None
        self.pending = None
        self.waiter = False
        self.steps = 0
        self.done = False
        self.home = Wave.count % STRIPES
        Wave.count += 1
        self.stripe_off = 0
        self.spins = 0
        sh.busy += 1

    def step(self):
        sh = self.sh
        if sh.abort:  # a wait gave up somewhere: leave (the launch reports an error)
            self.done = True
            return
        if self.pending is not None:  # spinning on a ring slot
            if not self.write_record():
                self.spins += 1
                if sh.spin_limit is not None and self.spins > sh.spin_limit:
                    sh.error |= 8
                    sh.abort = True
                    self.done = True
            return
        g, md = self.g, self.md
        M = md["M"]
        last = M - 2
        if self.top == 0 and self.roots_done:
            # ticket protocol: draw a ticket, wait for its record or for
            # busy == 0 (busy counts holders of work plus untaken records)
            if not self.is_idle:
                sh.busy -= 1
                sh.idle += 1
                self.is_idle = True
                self.ticket = None
            if not self.waiter:
                # strict bound on waiting wavefronts: the others leave
                if sh.waiters >= sh.idle_cap:
                    sh.idle -= 1
                    self.done = True
                    return
                sh.waiters += 1
                self.waiter = True
            if self.ticket is None:
                self.ticket = sh.head
                sh.head += 1
                return
            slot, lap = self.ticket % sh.cap, self.ticket // sh.cap
            rec = sh.record(slot)
            if rec is not None and rec[0] == sh.ringst.gen and rec[1] == self.ticket:
                assert self.ticket not in sh.taken
                sh.taken.add(self.ticket)
                f = rec[2]
                assert sh.done_lap(slot) == lap - 1
                sh.ringst.done[slot] = (sh.ringst.gen, lap)  # the slot's next lap may be written
                self.ticket = None
                sh.waiters -= 1
                self.waiter = False
                sh.idle -= 1
                self.is_idle = False  # inherits the record's busy credit
                self.frames[0] = dict(f)
                self.top = 1
                return
            if sh.busy == 0:
                sh.idle -= 1
                sh.waiters -= 1
                self.waiter = False
                self.done = True
            return  # spin
        if self.top == 0:
            if self.chunk_left == 0:
                c = -1
                while self.stripe_off < STRIPES:
                    k = (self.home + self.stripe_off) % STRIPES
                    sb, se = sh.nchunks * k // STRIPES, sh.nchunks * (k + 1) // STRIPES
                    o = sh.next_chunk[k]
                    sh.next_chunk[k] += 1
                    if sb + o < se:
                        c = sb + o
                        break
                    self.stripe_off += 1
                if c < 0:
                    self.roots_done = True
                    return
                first = self.beg + c * self.chunk_roots
                self.next_root = first
                self.chunk_left = min(self.chunk_roots // WAVE, (self.end - first + WAVE - 1) // WAVE)
            pushes = []
            for lane in range(WAVE):
                r = self.next_root + lane
                if r >= self.end:
                    continue
                u, v, t = (int(x) for x in g["Eg"][r])
                if u == v:
                    continue
                m = [0] * 8
                m[0], m[1] = u, v
                if not root_passes(g, md, self.delta, r):
                    continue
                L = cand_start(g, md, 1, m, r)
                if not (L[2] < L[3] and g["T"][L[0][L[2]]] <= t + self.delta):
                    continue
                pend = pend_of(g, t + self.delta)
                if pend - r - 1 < M - 1:
                    continue
                if last == 0:
                    self.cnt += leaf_count(g, md, 1, m, L, r, pend, self.scan_max)
                    continue
                ce = gallop_lb(L[0], L[2], L[3], pend)
                if ce > L[2]:
                    pushes.append(dict(cur=L[2], end=ce, p=r, pend=pend, lvl=0, dir=L[4], m=m[:], tl=t + self.delta))
            self.next_root += WAVE
            self.chunk_left -= 1
            assert self.top + len(pushes) <= self.FR
            for f in pushes:
                self.frames[self.top] = f
                self.top += 1
            return
        # expansion step (same batching as run_wave)
        top, frames = self.top, self.frames
        L_ = frames[top - 1]["lvl"]
        fcnt = [0] * WAVE
        for lane in range(WAVE):
            fi = top - 1 - lane
            if fi >= 0 and frames[fi]["lvl"] == L_:
                fcnt[lane] = frames[fi]["end"] - frames[fi]["cur"]
        assert fcnt[0] > 0
        incl = [int(x) for x in np.cumsum(fcnt)]
        excl = [incl[i] - fcnt[i] for i in range(WAVE)]
        total = min(WAVE, incl[-1])
        part = [fcnt[i] > 0 and excl[i] < WAVE for i in range(WAVE)]
        nparts = sum(part)
        take = [min(fcnt[i], WAVE - excl[i]) if part[i] else 0 for i in range(WAVE)]
        last_partial = any(part[i] and take[i] < fcnt[i] for i in range(WAVE))
        sm = 0
        for i in range(WAVE):
            if part[i]:
                sm |= 1 << excl[i]
        items = []
        for lane in range(total):
            rank = bin(sm & ((2 << lane) - 1)).count("1") - 1
            f = frames[top - 1 - rank]
            slot = f["cur"] + lane - excl[rank]
            assert f["cur"] <= slot < f["end"]
            items.append((slot, f))
        for i in range(WAVE):
            if part[i] and take[i] < fcnt[i]:
                frames[top - 1 - i] = dict(frames[top - 1 - i])
                frames[top - 1 - i]["cur"] += take[i]
        ntop = top - nparts + (1 if last_partial else 0)
        j = L_ + 1
        kind = md["kind"][j]
        children = []
        for slot, f in items:
            m = list(f["m"])
            if kind == K_NEW_DST:
                pos, node = g["outPos"][slot], g["outNbr"][slot]
                valid = node not in m[:md["mapped"][j]]
                m[md["b"][j]] = node
            elif kind == K_NEW_SRC:
                pos, node = g["inPos"][slot], g["inNbr"][slot]
                valid = node not in m[:md["mapped"][j]]
                m[md["a"][j]] = node
            else:
                pos = g["inPairPos"][slot] if f["dir"] else g["outPairPos"][slot]
                valid = True
            valid = valid and second_ok(g, md, j, slot, f["dir"], f["tl"])
            if not valid:
                continue
            tfn = next_first_time(g, md, j, slot, f["dir"], M)
            if tfn is not None and tfn > f["tl"]:
                continue
            if j == last:
                Lc = cand_start(g, md, j + 1, m, pos)
                self.cnt += leaf_count(g, md, j + 1, m, Lc, pos, f["pend"], self.scan_max)
                continue
            Lc = cand_start(g, md, j + 1, m, pos)
            ce = gallop_lb(Lc[0], Lc[2], Lc[3], f["pend"])
            if ce > Lc[2]:
                children.append(dict(cur=Lc[2], end=ce, p=pos, pend=f["pend"], lvl=j, dir=Lc[4], m=m, tl=f["tl"]))
        assert ntop + len(children) <= self.FR
        for k, ch in enumerate(children):
            frames[ntop + k] = ch
        self.top = ntop + len(children)
        self.steps += 1
        # donation
        if self.top > 0 and self.steps % self.steal_every == 0:
            want = min(sh.head - sh.tail, self.steal_max)
            for _ in range(max(0, want)):
                cand = [i for i in range(min(self.top, WAVE)) if frames[i]["end"] - frames[i]["cur"] >= 2 * self.steal_min]
                if not cand:
                    break
                sh.busy += 1  # the record's credit, before it can be seen
                r = sh.tail
                sh.tail += 1
                self.pending = (r, cand[0])
                if not self.write_record():
                    return  # wait for the ring slot's previous lap to be taken

    def write_record(self):
        """Write the pending donation once its ring slot is free."""
        sh, frames = self.sh, self.frames
        r, b = self.pending
        slot, lap = r % sh.cap, r // sh.cap
        if lap > 0 and sh.done_lap(slot) != lap - 1:
            return False
        f = dict(frames[b])
        mid = f["cur"] + (f["end"] - f["cur"]) // 2
        don = dict(f)
        don["cur"] = mid
        sh.ringst.slots[slot] = (sh.ringst.gen, r, don)
        sh.ndonated += 1
        f["end"] = mid
        frames[b] = f
        self.pending = None
        self.spins = 0
        return True


def run_waves(g, md, delta, beg, end, nwaves=4, scan_max=16, chunk_roots=64, steal_every=1, steal_max=4,
              pool_cap=1 << 20, seed=0, idle_cap=1 << 30, ring=None, spin_limit=None, with_error=False):
    """Simulate nwaves wavefronts stepping in a random interleaving; returns
    (count, number of donated frames) [, error bits].  ring: a Ring kept
    from earlier launches (this launch is counted on it first);
    spin_limit: the donor's bounded wait on its ring slot."""
    rng = np.random.default_rng(seed)
    nchunks = max(0, (end - beg + chunk_roots - 1) // chunk_roots)
    if ring is not None:
        ring.launch()
    sh = _Shared(nchunks, pool_cap, ring)
    sh.idle_cap = idle_cap
    sh.spin_limit = spin_limit
    waves = [Wave(g, md, delta, beg, end, sh, scan_max, chunk_roots, steal_every, steal_max)
             for _ in range(nwaves)]
    guard = 0
    while not all(w.done for w in waves):
        live = [w for w in waves if not w.done]
        w = live[int(rng.integers(0, len(live)))]
        w.step()
        guard += 1
        assert guard < 10_000_000, "no progress"
    if not sh.abort:
        assert len(sh.taken) == sh.tail == sh.ndonated, "a donated record was never taken"
        assert sh.busy == 0 and sh.idle == 0 and sh.waiters == 0
    res = (sum(w.cnt for w in waves), sh.ndonated)
    return res + (sh.error,) if with_error else res


# ------------------------------------------------ two-edge tail by the shorter list

T2_NONE, T2_INDEP, T2_JOIN = range(3)


def tail2_mode(md):
    """INDEP: the last edge k = M-1 does not touch the node edge j = M-2
    introduces (if it introduces one): given the partial match through edge
    j - 1, the last edge's candidate list is then fixed, and only the position
    order (and, when both edges bring new nodes, their distinctness) ties the
    two.  JOIN: edge j brings node w and edge k closes w with an earlier node."""
    M = md["M"]
    if M < 3:
        return T2_NONE
    j, k = M - 2, M - 1
    if md["kind"][j] == K_CLOSE:
        return T2_INDEP
    w = md["mapped"][j]  # the node edge j introduces
    if md["a"][k] != w and md["b"][k] != w:
        return T2_INDEP
    return T2_JOIN if md["kind"][k] == K_CLOSE else T2_NONE


def tail2_eligible(md):
    return tail2_mode(md) != T2_NONE


def tail2_lists(g, md, m, A, pend):
    """A = edge j's list (pos, nbr, lo, hi, dir) bounded by the window ([lo, hi)
    all inside (p_{j-1}, pend)); returns edge k's list restricted to positions
    in (pos of A's first slot, pend): (pos, nbr, lo, hi, dir)."""
    M = md["M"]
    k = M - 1
    pa0 = A[0][A[2]]
    kind = md["kind"][k]
    if tail2_mode(md) == T2_JOIN:
        w = md["mapped"][k - 1]
        into_y = md["a"][k] == w
        y = m[md["b"][k]] if into_y else m[md["a"][k]]
        pos, nbr, R = (g["inPos"], g["inNbr"], g["inR"]) if into_y else (g["outPos"], g["outNbr"], g["outR"])
        lo = upper_bound(pos, R[y], R[y + 1], pa0)
        return pos, nbr, lo, gallop_lb(pos, lo, R[y + 1], pend), 0
    if kind == K_NEW_DST:
        x = m[md["a"][k]]
        lo = upper_bound(g["outPos"], g["outR"][x], g["outR"][x + 1], pa0)
        hi = gallop_lb(g["outPos"], lo, g["outR"][x + 1], pend)
        return g["outPos"], g["outNbr"], lo, hi, 0
    if kind == K_NEW_SRC:
        y = m[md["b"][k]]
        lo = upper_bound(g["inPos"], g["inR"][y], g["inR"][y + 1], pa0)
        hi = gallop_lb(g["inPos"], lo, g["inR"][y + 1], pend)
        return g["inPos"], g["inNbr"], lo, hi, 0
    x, y = m[md["a"][k]], m[md["b"][k]]
    olo, ohi = g["outR"][x], g["outR"][x + 1]
    ilo, ihi = g["inR"][y], g["inR"][y + 1]
    if ohi - olo <= ihi - ilo:
        s0 = lower_bound(g["outPairNbr"], olo, ohi, y)
        s1 = gallop_ub(g["outPairNbr"], s0, ohi, y)
        pos, d = g["outPairPos"], 0
    else:
        s0 = lower_bound(g["inPairNbr"], ilo, ihi, x)
        s1 = gallop_ub(g["inPairNbr"], s0, ihi, x)
        pos, d = g["inPairPos"], 1
    lo = upper_bound(pos, s0, s1, pa0)
    return pos, None, lo, gallop_lb(pos, lo, s1, pend), d


def tail2_item(g, md, m, A, q_pos, q_nbr):
    """Matches of the last two edges whose last edge is the candidate at
    position q_pos (new endpoint q_nbr when edge k brings a node): the
    candidates c of edge j before it (position < q_pos) whose new node, if
    any, is not mapped and differs from q_nbr."""
    M = md["M"]
    j, k = M - 2, M - 1
    nm = md["mapped"][j]
    if tail2_mode(md) == T2_JOIN:
        if q_nbr in m[:nm]:
            return 0
        pos = A[0]
        pa0 = pos[A[2]]
        if md["kind"][j] == K_NEW_DST:
            return edge_count(g, m[md["a"][j]], q_nbr, pa0 - 1, q_pos)
        return edge_count(g, q_nbr, m[md["b"][j]], pa0 - 1, q_pos)
    if md["kind"][k] != K_CLOSE and q_nbr in m[:nm]:
        return 0
    pos, nbr, lo, hi, _ = A
    r = lower_bound(pos, lo, hi, q_pos) - lo
    if r <= 0 or md["kind"][j] == K_CLOSE:
        return max(r, 0)
    pa0 = pos[lo]
    if md["kind"][j] == K_NEW_DST:
        base = m[md["a"][j]]
        pn, pp, plo, phi = g["outPairNbr"], g["outPairPos"], g["outR"][base], g["outR"][base + 1]
    else:
        base = m[md["b"][j]]
        pn, pp, plo, phi = g["inPairNbr"], g["inPairPos"], g["inR"][base], g["inR"][base + 1]
    for i in range(nm):
        r -= pair_count(pn, pp, plo, phi, m[i], pa0 - 1, q_pos)
    if md["kind"][k] != K_CLOSE:
        r -= pair_count(pn, pp, plo, phi, q_nbr, pa0 - 1, q_pos)
    return r


def tail2_count(g, md, m, A, pend):
    """Matches of the last two motif edges for one frame of edge j (list A
    window-bounded), summed over the last edge's candidates instead of edge
    j's: the kernel takes this way when edge k's window list is the shorter."""
    B = tail2_lists(g, md, m, A, pend)
    pos, nbr, lo, hi, _ = B
    return sum(tail2_item(g, md, m, A, pos[s], nbr[s] if nbr is not None else -1) for s in range(lo, hi))


def tail2_direct(g, md, m, A, pend, scan_max=16):
    """The same sum the way the kernel takes it otherwise: every candidate c
    of edge j, then the last edge's leaf count."""
    M = md["M"]
    j = M - 2
    pos, nbr, lo, hi, d = A
    total = 0
    for s in range(lo, hi):
        mm = list(m)
        if md["kind"][j] == K_NEW_DST:
            if nbr[s] in m[:md["mapped"][j]]:
                continue
            mm[md["b"][j]] = nbr[s]
        elif md["kind"][j] == K_NEW_SRC:
            if nbr[s] in m[:md["mapped"][j]]:
                continue
            mm[md["a"][j]] = nbr[s]
        Lc = cand_start(g, md, j + 1, mm, pos[s])
        total += leaf_count(g, md, j + 1, mm, Lc, pos[s], pend, scan_max)
    return total


# ----------------------------------- persistent launches sharing the GPU's slots

def run_concurrent_launches(grids, capacity, idle_cap, work, rng, max_ticks=100_000):
    """Several persistent launches on one GPU with `capacity` resident
    wavefront slots.  Launch L has grids[L] wavefronts; its `busy` counter
    starts at grids[L] (a wavefront not yet started counts as holding work,
    as the reset kernel sets it).  Each tick the hardware starts waiting
    wavefronts of any launch (random interleaving) into free slots, and every
    resident wavefront takes a step: it works off its roots (work[L][w]
    steps), then goes idle (busy - 1) and either waits for donations (at most
    idle_cap waiters per launch; a waiter holds its slot until its launch's
    busy reaches 0) or leaves.  Returns "done" or "deadlock" (no wavefront can
    change state and no slot is free for an unstarted one)."""
    K = len(grids)
    unstarted = [list(range(g)) for g in grids]
    busy = list(grids)
    waiters = [0] * K
    resident = []  # [launch, wave, remaining work, state]  state: 0 work, 1 waiting
    for tick in range(max_ticks):
        # admit
        free = capacity - len(resident)
        cands = [L for L in range(K) if unstarted[L]]
        while free > 0 and cands:
            L = cands[int(rng.integers(0, len(cands)))]
            w = unstarted[L].pop()
            resident.append([L, w, work[L][w], 0])
            free -= 1
            cands = [L for L in range(K) if unstarted[L]]
        if not resident and not any(unstarted):
            return "done"
        changed = False
        keep = []
        for r in resident:
            L = r[0]
            if r[3] == 0:
                if r[2] > 0:
                    r[2] -= 1
                    changed = True
                    keep.append(r)
                    continue
                busy[L] -= 1
                changed = True
                if waiters[L] >= idle_cap:
                    continue  # leaves: its slot frees
                waiters[L] += 1
                r[3] = 1
                keep.append(r)
            else:
                if busy[L] == 0:
                    waiters[L] -= 1
                    changed = True
                    continue
                keep.append(r)
        resident = keep
        if not changed and len(resident) >= capacity and any(unstarted):
            return "deadlock"
    return "timeout"


# ------------------------------------- last two edges on one list: closed forms

TP_NONE, TP_STAR, TP_PAIR, TP_SAME = range(4)


def tail_pattern(md):
    M = md["M"]
    if M < 3:
        return TP_NONE
    j, k = M - 2, M - 1
    kj, kk = md["kind"][j], md["kind"][k]
    samepair = md["a"][k] == md["a"][j] and md["b"][k] == md["b"][j]
    if kj == K_CLOSE:
        return TP_SAME if (kk == K_CLOSE and samepair) else TP_NONE
    if kk == K_CLOSE:
        return TP_PAIR if samepair else TP_NONE
    if kk == kj and (md["a"][k] == md["a"][j] if kj == K_NEW_DST else md["b"][k] == md["b"][j]):
        return TP_STAR
    return TP_NONE


def tail_pattern_count(g, md, m, A, p, pend):
    """The kernel's pattern sum for one frame of edge M-2 (list A bounded by
    the window, p = the position matched for edge M-3): the frame's closed-form
    term plus one same-pair term per slot."""
    pat = tail_pattern(md)
    pos, nbr, lo, hi, _ = A
    K = hi - lo
    if pat == TP_SAME:
        return K * (K - 1) // 2
    j = md["M"] - 2
    out = md["kind"][j] == K_NEW_DST
    base = m[md["a"][j]] if out else m[md["b"][j]]
    kp, c2 = K, 0
    for x in m[:md["mapped"][j]]:
        n = edge_count(g, base, x, p, pend) if out else edge_count(g, x, base, p, pend)
        kp -= n
        c2 += n * (n - 1) // 2
    sp = 0
    for s in range(lo, hi):
        w = nbr[s]
        sp += edge_count(g, base, w, pos[s], pend) if out else edge_count(g, w, base, pos[s], pend)
    if pat == TP_PAIR:
        return sp - c2
    return kp * (kp - 1) // 2 + c2 - sp


# ------------------------------------------------------------------ window sweep
# 3-edge motifs whose last two edges take candidates from one list of a root
# node (the anchor): a star (two new nodes on the anchor's out- or in-list)
# or a new node's pair twice.  Per root r the count is
#   star: C(K', 2) + sum_x C(n_x, 2) - P_r      pair: P_r - sum_x C(n_x, 2)
# (window A_r of K slots, x in the root's nodes, n_x = slots of A_r whose
# other end is x, K' = K - sum n_x, P_r = pairs of slots of A_r with the same
# other end).  The roots of one anchor are the anchor's root list in
# position order; their windows' ends only move forward, so
#   sum over valid roots of P_r = sum over slots c of L of
#       W(i_add) * add(c) - W(i_rem) * rem(c)
# where c enters the window at root i_add (with add(c) = same-pair slots
# already in it) and leaves at root i_rem (rem(c) = same-pair slots after it
# still in it), and W(i) = valid roots at or after i.  Every term is O(log).
WS_STAR, WS_PAIR = 1, 2


def window_sweep_plan(md):
    """(pattern, anchor motif node, anchor list is out) or None."""
    if md["M"] != 3:
        return None
    k1, k2 = md["kind"][1], md["kind"][2]
    a1, b1, a2, b2 = md["a"][1], md["b"][1], md["a"][2], md["b"][2]
    if k1 == K_NEW_DST and k2 == K_NEW_DST and a1 == a2:
        return (WS_STAR, a1, True)
    if k1 == K_NEW_SRC and k2 == K_NEW_SRC and b1 == b2:
        return (WS_STAR, b1, False)
    if k1 == K_NEW_DST and k2 == K_CLOSE and (a2, b2) == (a1, b1):
        return (WS_PAIR, a1, True)
    if k1 == K_NEW_SRC and k2 == K_CLOSE and (a2, b2) == (a1, b1):
        return (WS_PAIR, b1, False)
    return None


def _c2(n):
    return n * (n - 1) // 2


def window_sweep_count(g, md, delta, beg, end):
    """The window-sweep kernel's count (mod 2^64), item by item: items are
    edge positions q in [beg, max(end, qmax)); position q is a root (q < end)
    and a slot of the anchor list of its source (out) or destination (in)."""
    pat, anc, out = window_sweep_plan(md)
    E, Eg, T = g["E"], g["Eg"], g["T"]
    LR, LP = (g["outR"], g["outPos"]) if out else (g["inR"], g["inPos"])
    RR, RP = (g["outR"], g["outPos"]) if anc == 0 else (g["inR"], g["inPos"])
    same = (anc == 0) == out  # the anchor's root list is its window list

    def pcnt(x, y, p, pend):  # slots of x's list whose other end is y, positions in (p, pend)
        return edge_count(g, x, y, p, pend) if out else edge_count(g, y, x, p, pend)

    if end <= beg:
        return 0
    qmax = pend_of(g, T[end - 1] + delta)
    tot = 0
    for q in range(beg, max(end, qmax)):
        u, v = int(Eg[q, 0]), int(Eg[q, 1])
        if q < end and u != v:  # root q
            x = u if anc == 0 else v
            pend = pend_of(g, T[q] + delta)
            if same:
                lo = (g["outSelf"] if out else g["inSelf"])[q] + 1
            else:  # the other list of q's own endpoint: the rank arrays
                lo = int((g["inNext"] if anc == 0 else g["outNext"])[q])
                assert lo == upper_bound(LP, LR[x], LR[x + 1], q)
            hi = lower_bound(LP, lo, LR[x + 1], pend)
            nu, nv = pcnt(x, u, q, pend), pcnt(x, v, q, pend)
            c2 = _c2(nu) + _c2(nv)
            tot += _c2(hi - lo - nu - nv) + c2 if pat == WS_STAR else -c2
        if q > beg:  # slot of q in the list of x
            x, w = (u, v) if out else (v, u)
            r0 = RR[x] if beg == 0 else lower_bound(RP, RR[x], RR[x + 1], beg)
            r1 = RR[x + 1] if (RR[x + 1] > r0 and RP[RR[x + 1] - 1] < end) else lower_bound(RP, r0, RR[x + 1], end)
            assert (r0, r1) == (lower_bound(RP, RR[x], RR[x + 1], beg),
                                lower_bound(RP, lower_bound(RP, RR[x], RR[x + 1], beg), RR[x + 1], end))
            if r0 == r1:
                continue
            # first root whose window end reaches q: time >= T[q] - delta
            tq = T[q] - delta
            lo_, hi_ = r0, r1
            while lo_ < hi_:
                mid = (lo_ + hi_) >> 1
                if T[RP[mid]] < tq:
                    lo_ = mid + 1
                else:
                    hi_ = mid
            ia = lo_
            if ia == r1:
                continue
            P = RP[ia - 1] if ia > r0 else RP[r0]
            if q <= P:
                continue  # before the first root's window: in no window
            add = pcnt(x, w, P, q)
            if same:
                c = (g["outSelf"] if out else g["inSelf"])[q]
                ir = c if q < end else r1
            else:
                ir = lower_bound(RP, r0, r1, q)
            rem = 0
            if ir < r1:
                rem = pcnt(x, w, q, pend_of(g, T[RP[ir]] + delta))

            def valid(i):  # roots at slots [i, r1) that are not self-loops
                n = r1 - i
                if n > 0:
                    n -= edge_count(g, x, x, RP[i] - 1, end)
                return n

            s = valid(ia) * add - (valid(ir) * rem if ir < r1 else 0)
            tot += s if pat == WS_PAIR else -s
    return tot % (1 << 64)


# Root-pair sweep (the pair-sweep kernel): 4-edge motifs root 0 -> 1, edge 1
# a new node 2 on a list of node 0 or 1, edge 2 a new node 3 on a list of
# node 2, edge 3 the root's pair 0 -> 1 again.  The roots of one pair (u, v)
# share their expansions: an edge-1 slot a is expanded once, by the root
# whose window piece holds it (pieces: root i's window (p_i, clip_i) past the
# previous root's clip), and weighted per closing pair edge p_j > a by the
# roots before a whose window reaches p_j:
#   sum_j (A - lo_j) * #(valid edge-2 slots in (a, p_j)),  lo_j < A,
# A = roots of the pair before a, lo_j = the first root with time >= t_j - delta.

def pair_sweep_plan(md):
    """(s1, d1, d2): edge 1 on node s1's out-list (d1 = 1) or in-list, edge 2
    on node 2's out-list (d2 = 1) or in-list; None for other motifs."""
    if md["M"] != 4 or md["kind"][3] != K_CLOSE or (md["a"][3], md["b"][3]) != (0, 1):
        return None
    k1, a1, b1, k2, a2, b2 = md["kind"][1], md["a"][1], md["b"][1], md["kind"][2], md["a"][2], md["b"][2]
    if k1 == K_NEW_DST and a1 in (0, 1):
        s1, d1 = a1, 1
    elif k1 == K_NEW_SRC and b1 in (0, 1):
        s1, d1 = b1, 0
    else:
        return None
    if k2 == K_NEW_DST and a2 == 2:
        d2 = 1
    elif k2 == K_NEW_SRC and b2 == 2:
        d2 = 0
    else:
        return None
    return (s1, d1, d2)


def pair_sweep_count(g, md, delta, beg, end, scan_max=32):
    """The pair-sweep kernel's count (mod 2^64), root by root; edge-2 slots
    scanned directly while at most scan_max of them precede p_j, else
    counted by rank minus pair counts (the kernel's kScan)."""
    s1, d1, d2 = pair_sweep_plan(md)
    Eg, T = g["Eg"], g["T"]
    pp, pn, oR = g["outPairPos"], g["outPairNbr"], g["outR"]
    L1R, L1P, L1N = (g["outR"], g["outPos"], g["outNbr"]) if d1 else (g["inR"], g["inPos"], g["inNbr"])
    L2R, L2P, L2N = (g["outR"], g["outPos"], g["outNbr"]) if d2 else (g["inR"], g["inPos"], g["inNbr"])
    tot = 0
    for r in range(beg, end):
        u, v, t = (int(x) for x in Eg[r])
        if u == v:
            continue
        k = int(g["outPairSelf"][r])
        s0 = lower_bound(pn, oR[u], k, v)
        s1_ = upper_bound(pn, k, oR[u + 1], v)
        c = lower_bound(pp, k, s1_, pend_of(g, t + delta)) - 1
        clip = int(pp[c])
        lo = r
        if k > s0 and pp[k - 1] >= beg:  # the previous root's window covers up to its clip
            cp = lower_bound(pp, k - 1, s1_, pend_of(g, int(T[pp[k - 1]]) + delta)) - 1
            lo = max(lo, int(pp[cp]))
        if clip <= lo:
            continue
        x = u if s1 == 0 else v
        i0 = upper_bound(L1P, L1R[x], L1R[x + 1], lo)
        i1 = lower_bound(L1P, i0, L1R[x + 1], clip)
        rb = lower_bound(pp, s0, s1_, beg)
        for sl in range(i0, i1):
            a, w = int(L1P[sl]), int(L1N[sl])
            if w == u or w == v:
                continue
            # edge-2 list of w after a from the rank arrays (a is in in(w) or out(w))
            if d1:
                base = int(g["outNext"][a]) if d2 else int(g["inSelf"][a]) + 1
            else:
                base = int(g["outSelf"][a]) + 1 if d2 else int(g["inNext"][a])
            assert base == upper_bound(L2P, L2R[w], L2R[w + 1], a)
            hi2 = L2R[w + 1]
            if base >= hi2:
                continue
            first2 = int(L2P[base])
            A = lower_bound(pp, rb, s1_, min(a, end))
            tlim = int(T[pp[A - 1]]) + delta
            j = A if a < end else upper_bound(pp, s0, s1_, a)  # a is not a pair edge, and a > r >= beg
            assert j == upper_bound(pp, s0, s1_, a)
            cur, valid, scan = base, 0, True
            while j < s1_:
                pj = int(pp[j])
                if int(T[pj]) > tlim:
                    break
                if pj <= first2:
                    j += 1
                    continue
                lo_j = max(rb, lower_bound(pp, s0, s1_, pend_of(g, int(T[pj]) - delta - 1)))
                assert lo_j < A
                if scan:
                    while cur < hi2 and cur - base < scan_max and L2P[cur] < pj:
                        y = int(L2N[cur])
                        valid += (y != u) and (y != v) and (y != w)
                        cur += 1
                    scan = not (cur < hi2 and L2P[cur] < pj)
                if scan:
                    n2 = valid
                else:
                    n2 = lower_bound(L2P, base, hi2, pj) - base
                    if d2:
                        n2 -= edge_count(g, w, u, a, pj) + edge_count(g, w, v, a, pj) + edge_count(g, w, w, a, pj)
                    else:
                        n2 -= edge_count(g, u, w, a, pj) + edge_count(g, v, w, a, pj) + edge_count(g, w, w, a, pj)
                tot += (A - lo_j) * n2
                j += 1
    return tot % (1 << 64)
