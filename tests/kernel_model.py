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
    return dict(E=E, V=V, Eg=Eg, T=Arr(T, "T"),
                outR=Arr(g["outR"], "outR"), outPos=Arr(outPos, "outPos"), outNbr=Arr(outNbr, "outNbr"),
                inR=Arr(g["inR"], "inR"), inPos=Arr(inPos, "inPos"), inNbr=Arr(inNbr, "inNbr"),
                outPairNbr=Arr(outNbr[o], "outPairNbr"), outPairPos=Arr(outPos[o], "outPairPos"),
                inPairNbr=Arr(inNbr[i], "inPairNbr"), inPairPos=Arr(inPos[i], "inPairPos"),
                bucket=Arr(bucket, "bucket"), tmin=tmin, bshift=bshift, nbuckets=nb,
                Euid=Arr(g["Euid"], "Euid"))


def build_motif(edges):
    """tmh::build_motif: kinds / node slots per edge from renamed (a, b) pairs."""
    M = len(edges)
    assert edges[0][0] == 0 and edges[0][1] == 1
    kind, A, B, mapped = [0] * M, [0] * M, [0] * M, [0] * M
    n = 2
    for j in range(1, M):
        a, b = edges[j][0], edges[j][1]
        assert a != b
        A[j], B[j], mapped[j] = a, b, n
        if a < n and b < n:
            kind[j] = K_CLOSE
        elif a < n and b == n:
            kind[j] = K_NEW_DST
            n += 1
        elif b < n and a == n:
            kind[j] = K_NEW_SRC
            n += 1
        else:
            raise ValueError("not prefix-connected")
    return dict(M=M, n=n, kind=kind, a=A, b=B, mapped=mapped)


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


def pair_count(pn, pp, lo, hi, y, p, pend):
    s0 = lower_bound(pn, lo, hi, y)
    if s0 >= hi or pn[s0] != y:
        return 0
    s1 = upper_bound(pn, s0 + 1, hi, y)
    c0 = upper_bound(pp, s0, s1, p)
    c1 = lower_bound(pp, c0, s1, pend)
    return c1 - c0


def child_range(g, md, j, m, p, pend):
    k = md["kind"][j]
    if k == K_NEW_DST:
        x = m[md["a"][j]]
        lo, hi = g["outR"][x], g["outR"][x + 1]
        cb = upper_bound(g["outPos"], lo, hi, p)
        return cb, lower_bound(g["outPos"], cb, hi, pend), 0
    if k == K_NEW_SRC:
        y = m[md["b"][j]]
        lo, hi = g["inR"][y], g["inR"][y + 1]
        cb = upper_bound(g["inPos"], lo, hi, p)
        return cb, lower_bound(g["inPos"], cb, hi, pend), 0
    x, y = m[md["a"][j]], m[md["b"][j]]
    olo, ohi = g["outR"][x], g["outR"][x + 1]
    ilo, ihi = g["inR"][y], g["inR"][y + 1]
    if ohi - olo <= ihi - ilo:
        s0 = lower_bound(g["outPairNbr"], olo, ohi, y)
        s1 = upper_bound(g["outPairNbr"], s0, ohi, y)
        pp, d = g["outPairPos"], 0
    else:
        s0 = lower_bound(g["inPairNbr"], ilo, ihi, x)
        s1 = upper_bound(g["inPairNbr"], s0, ihi, x)
        pp, d = g["inPairPos"], 1
    cb = upper_bound(pp, s0, s1, p)
    return cb, lower_bound(pp, cb, s1, pend), d


def leaf_count(g, md, j, m, p, pend, scan_max):
    k = md["kind"][j]
    if k == K_CLOSE:
        x, y = m[md["a"][j]], m[md["b"][j]]
        olo, ohi = g["outR"][x], g["outR"][x + 1]
        ilo, ihi = g["inR"][y], g["inR"][y + 1]
        if ohi - olo <= ihi - ilo:
            return pair_count(g["outPairNbr"], g["outPairPos"], olo, ohi, y, p, pend)
        return pair_count(g["inPairNbr"], g["inPairPos"], ilo, ihi, x, p, pend)
    out = k == K_NEW_DST
    base = m[md["a"][j]] if out else m[md["b"][j]]
    R = g["outR"] if out else g["inR"]
    pos = g["outPos"] if out else g["inPos"]
    nbr = g["outNbr"] if out else g["inNbr"]
    lo, hi = R[base], R[base + 1]
    b0 = upper_bound(pos, lo, hi, p)
    b1 = lower_bound(pos, b0, hi, pend)
    c = b1 - b0
    if c <= 0:
        return 0
    nm = md["mapped"][j]
    if c <= scan_max:
        bad = sum(1 for s in range(b0, b1) if nbr[s] in m[:nm])
        return c - bad
    pn = g["outPairNbr"] if out else g["inPairNbr"]
    pq = g["outPairPos"] if out else g["inPairPos"]
    for i in range(nm):
        c -= pair_count(pn, pq, lo, hi, m[i], p, pend)
    return c


def run_wave(g, md, delta, beg, end, enumerate_=False, scan_max=16, chunk_roots=256, cap=1 << 30):
    """Simulate one wavefront mining roots [beg, end) alone; returns
    (count, rows, stats)."""
    M = md["M"]
    MAXN = 8
    lvls = max(1, (M - 1) if enumerate_ else (M - 2))
    FR = WAVE * lvls
    frames = [None] * FR  # each: dict(cur,end,p,pend,lvl,m,e)
    top = 0
    cnt = 0
    rows = []
    stats = dict(steps=0, items=0, max_top=0)
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
                pend = pend_of(g, t + delta)
                if pend - r - 1 < M - 1:
                    continue
                m = [0] * MAXN
                m[0], m[1] = u, v
                if last == 0:
                    cnt += leaf_count(g, md, 1, m, r, pend, scan_max)
                    continue
                cb, ce, d = child_range(g, md, 1, m, r, pend)
                if ce > cb:
                    pushes.append(dict(cur=cb, end=ce, p=r, pend=pend, lvl=0, dir=d, m=m[:], e=[r]))
            assert top + len(pushes) <= FR
            for f in pushes:
                frames[top] = f
                top += 1
            # ---- expansion steps until the stack drains
            while top > 0:
                stats["max_top"] = max(stats["max_top"], top)
                L = frames[top - 1]["lvl"]
                fcnt = [0] * WAVE
                for lane in range(WAVE):
                    fi = top - 1 - lane
                    if fi >= 0 and frames[fi]["lvl"] == L:
                        fcnt[lane] = frames[fi]["end"] - frames[fi]["cur"]
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
                    if not valid:
                        continue
                    e = f["e"] + [pos]
                    if j == last:
                        if enumerate_:
                            if cnt < cap:
                                rows.append(tuple(g["Euid"][x] for x in e))
                            cnt += 1
                        else:
                            cnt += leaf_count(g, md, j + 1, m, pos, f["pend"], scan_max)
                        continue
                    cb, ce, d = child_range(g, md, j + 1, m, pos, f["pend"])
                    if ce > cb:
                        children.append(dict(cur=cb, end=ce, p=pos, pend=f["pend"], lvl=j, dir=d, m=m, e=e))
                assert ntop + len(children) <= FR, "frame stack overflow"
                for k, ch in enumerate(children):
                    frames[ntop + k] = ch
                top = ntop + len(children)
    return cnt, rows, stats
