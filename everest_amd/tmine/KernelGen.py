"""Per-query worker generator (reference: system/tmine/KernelGen.py:1-538).

The reference expands CUDA snippets into a motif-specialised sm_86 worker.
Here the generator fills templates/worker.hip.tmpl: a gfx950 HIP plugin that
instantiates the library's mining kernel template with the motif as
compile-time tables (kinds, node slots, list start rules, first-candidate
time arrays per edge), the per-query options baked into its launch
parameters, and the reference worker entry point getWorker.
"""
import os
import re
from string import Template

from . import locations

K_NEW_DST, K_NEW_SRC, K_CLOSE = 0, 1, 2
S_SEARCH, S_OUT_SELF, S_OUT_NEXT, S_IN_SELF, S_IN_NEXT, S_PAIR_SELF = range(6)
MAXE = 8
MAXN = 8


def InspectMotif(pathToMotif):
    """(mapped node count before each edge >= 1, max stack size, edge count),
    from the motif file in line order (KernelGen.py:13-29)."""
    mnode, mnode_p = [], []
    allE = 0
    with open(pathToMotif, "r") as f:
        for line in f:
            if not line.split():
                continue
            allE += 1
            u, v, _ = line.split()[:3]
            mnode_p.append(len(mnode))
            if u not in mnode:
                mnode.append(u)
            if v not in mnode:
                mnode.append(v)
    return mnode_p[1:], allE - 2, allE


def motif_edges(pathToMotif):
    """Renamed, time-ordered (src, dst) pairs, read by the core library's motif
    reader (the same reader HostMotifData uses)."""
    from everest_amd import core
    m = core.data.HostMotifData(pathToMotif)
    return [(u, v) for (u, v, _) in m.edges()]


def close_rule(a, b, pa, pb):
    """Pruning rule (TM_T_*) of a closing edge a -> b after edge pa -> pb:
    the pair's own next edge (same or reversed pair), else the next edge of
    the right direction at a node the two edges share, else none."""
    if (a, b) == (pa, pb):
        return 4
    if (a, b) == (pb, pa):
        return 5
    if a == pa:
        return 0
    if a == pb:
        return 1
    if b == pb:
        return 2
    if b == pa:
        return 3
    return -1


def further_rules(edges, j, next_rule):
    """Rules (TM_T_* bits) that later edges k >= j + 2 impose at the graph edge
    matched for edge j, without those implied by another or equal to
    next_rule (edge j + 1's own)."""
    pa, pb = edges[j]
    mask = 0
    for k in range(j + 2, len(edges)):
        r = close_rule(edges[k][0], edges[k][1], pa, pb)
        if r >= 0:
            mask |= 1 << r
    nxt = (1 << next_rule) if next_rule >= 0 else 0
    have = mask | nxt
    if have & (1 << 4):
        mask &= ~((1 << 0) | (1 << 2))
    if have & (1 << 5):
        mask &= ~((1 << 1) | (1 << 3))
    return mask & ~nxt


def motif_tables(edges):
    """Per-edge tables of the kernel spec, as the library builds them at run time."""
    M = len(edges)
    if M < 2 or M > MAXE:
        raise ValueError("motifs need 2..%d edges" % MAXE)
    if edges[0] != (0, 1):
        raise ValueError("motif nodes must be renamed by first appearance")
    kind, A, B, mapped, smode, tmode = ([0] * MAXE for _ in range(6))
    tmode = [-1] * MAXE
    n = 2
    for j in range(1, M):
        a, b = edges[j]
        pa, pb = edges[j - 1]
        if a == b:
            raise ValueError("motif self-loop")
        A[j], B[j], mapped[j] = a, b, n
        if a < n and b < n:
            kind[j] = K_CLOSE
            smode[j] = S_PAIR_SELF if (a, b) == (pa, pb) else S_SEARCH
            tmode[j] = close_rule(a, b, pa, pb)
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
            raise ValueError("motif is not prefix-connected")
        if n > MAXN:
            raise ValueError("motif has more than %d nodes" % MAXN)
    xmask = [0] * MAXE
    for j in range(M - 2):
        xmask[j] = further_rules(edges, j, tmode[j + 1])
    return dict(M=M, N=n, kind=kind, a=A, b=B, mapped=mapped, smode=smode, tmode=tmode, xmask=xmask)


def _c_expr(expr):
    """A constraint expression as the reference pastes it into device code
    (Query.py:229-280): a C expression over `f` (a label), or for the filter
    over eK / nK; only checked to be one expression without statements."""
    expr = str(expr).strip()
    if not expr or any(c in expr for c in ";{}#\\\n"):
        raise ValueError("constraint expression must be one C expression: %r" % expr)
    return expr


def anti_tables(cons, tables):
    """Anti-edge flags and, per motif edge, the nearest earlier edge that is
    not an anti edge (Query.py:39-77: the reference's antiEdge,
    antiEdgeOffset and antiEdgePrev arrays).  Defined here for anti edges
    whose two endpoints are mapped by earlier edges, not directly after the
    root unless last: for an anti edge with a new endpoint the reference maps
    that node to the last candidate it looked at, and after an anti edge at
    position 1 it starts the next edge's list one past the graph's last edge
    (system/tmine/KernelGen.py:490-497), so neither has a count to match."""
    M = tables["M"]
    c = cons.cons
    if "antiEdge" not in c:
        return [False] * M, list(range(-1, M - 1)) if M else []
    raw = c["antiEdge"]
    if not isinstance(raw, (list, tuple)) or len(raw) != M:
        raise ValueError("antiEdge needs one flag per motif edge (%d)" % M)
    anti = []
    for x in raw:
        if isinstance(x, str):
            if x.strip().lower() not in ("true", "false"):
                raise ValueError("antiEdge flags are true / false: %r" % (x,))
            x = x.strip().lower() == "true"
        anti.append(bool(x))
    if anti[0]:
        raise ValueError("the root edge cannot be an anti edge")
    for j in range(1, M):
        if anti[j] and tables["kind"][j] != 2:
            raise NotImplementedError("anti edge %d introduces a motif node; anti edges between nodes that "
                                      "earlier edges map are supported" % j)
    if M > 2 and anti[1]:
        raise NotImplementedError("an anti edge directly after the root edge is supported as the motif's "
                                  "last edge only")
    prev = [-1] + [max(i for i in range(j) if not anti[i]) for j in range(1, M)]
    return anti, prev


def constraint_bodies(cons, M, N, anti=None):
    """Device-function bodies of the constrained worker (templates/worker_cons.hip.tmpl):
    label predicates per motif node / edge (a `node` or `edge` key applies to
    all, `nodeK` / `edgeK` to motif node / edge K and takes precedence), the
    window of each edge after the previous one (`temporal`: gaps between
    consecutive motif edges, KernelGen.py:467-477) and the filter over a
    complete match (`filter`, with eK / nK the labels of matched edge / node
    K, KernelGen.py:530-534 and Query.py:84-102)."""
    c = cons.cons
    def per(prefix, count):
        out = {}
        for i in range(count):
            if "%s%d" % (prefix, i) in c:
                out[i] = _c_expr(c["%s%d" % (prefix, i)])
            elif prefix in c:
                out[i] = _c_expr(c[prefix])
        return out
    nodes, edges = per("node", N), per("edge", M)
    def body(preds, arr, idx, var):
        if not preds:
            return "    (void)%s; (void)%s; (void)%s;\n    return true;" % (var, arr, idx)
        lines = ["    const int f = %s[%s];" % (arr, idx), "    switch (%s) {" % var]
        for i, e in sorted(preds.items()):
            lines.append("      case %d: return (%s);" % (i, e))
        lines += ["      default: return true;", "    }"]
        return "\n".join(lines)
    node_body = body(nodes, "VF", "node", "i")
    edge_body = body(edges, "EF", "p", "j")
    if "temporal" in c:
        gaps = [int(x) for x in c["temporal"]]
        if len(gaps) < M - 1:
            raise ValueError("temporal needs %d gaps (one per pair of consecutive motif edges)" % (M - 1))
        window_body = ("    constexpr int tw[TM_MAXE] = {%s};\n"
                       "    (void)t0;\n"
                       "    const long long w = tmk::pick(tw, j);\n"
                       "    if (j == 0) return tj + w;\n"
                       "    return tj + w < tmax ? tj + w : tmax;"
                       % ", ".join(str(x) for x in (gaps + [0] * MAXE)[:MAXE]))
    else:
        window_body = "    (void)j; (void)tj; (void)t0;\n    return tmax;"
    filt = _c_expr(c["filter"]) if "filter" in c else None
    uses_e = uses_n = False
    if filt:
        uses_e = re.search(r"\be\d+\b", filt) is not None
        uses_n = re.search(r"\bn\d+\b", filt) is not None
        for k in [int(x) for x in re.findall(r"\be(\d+)\b", filt)]:
            if k >= M:
                raise ValueError("filter names e%d of a %d-edge motif" % (k, M))
            if anti and anti[k]:
                # the reference reads the label one past the graph's last edge here
                raise ValueError("filter names e%d, an anti edge, which matches no graph edge" % k)
        for k in [int(x) for x in re.findall(r"\bn(\d+)\b", filt)]:
            if k >= N:
                raise ValueError("filter names n%d of a %d-node motif" % (k, N))
        expr = re.sub(r"\be(\d+)\b", r"EF[pos[\1]]", filt)
        expr = re.sub(r"\bn(\d+)\b", r"VF[m[\1]]", expr)
        filter_body = "    (void)EF; (void)VF; (void)pos; (void)m;\n    return (%s);" % expr
    else:
        filter_body = "    (void)EF; (void)VF; (void)pos; (void)m;\n    return true;"
    return dict(node_body=node_body, edge_body=edge_body, window_body=window_body, filter_body=filter_body,
                need_ef="true" if (edges or uses_e) else "false", need_vf="true" if (nodes or uses_n) else "false")


class KernelGen:
    def __init__(self, pathToMotif, query, tuning=None):
        self.query = query
        self.options = query.options()
        self.cons = query.constraints()
        feats = self.cons.features()
        self.constrained = bool(feats)
        if "antiEdge" in feats and self.options.enumerate():
            # the reference writes the id one past the graph's last edge for an anti edge's slot
            raise NotImplementedError("enumeration with anti edges has no rows to match (an anti edge matches "
                                      "no graph edge); counting with anti edges, and enumeration under the "
                                      "other constraint features, are supported")
        if self.options.multigpu():
            raise NotImplementedError("this build drives one GPU (options.gpu = %d)" % self.options.gpu())
        self.pathToMotif = pathToMotif
        self.edges = motif_edges(pathToMotif)
        self.tables = motif_tables(self.edges)
        self.anti, self.prev = anti_tables(self.cons, self.tables)
        self.enumerate = self.options.enumerate()
        self.enum_NUM = self.options.enumNum() if self.enumerate else -1
        t = dict(waves_per_cu=0, claim_blocks=0, scan_max=0, steal_every=0, steal_max=0, steal_min=0,
                 steal_after_us=0, waves_per_simd_hint=0, sb_flush=0, sb_inline=0)
        if tuning:
            t.update({k: int(v) for k, v in tuning.items() if k in t})
        # register cap of the counting kernel: 80 VGPRs (6 wavefronts per SIMD)
        # when the LDS stack of this motif leaves room for them (at most two
        # stacked levels), else 96 VGPRs (5 per SIMD)
        if t["waves_per_simd_hint"] <= 0:
            t["waves_per_simd_hint"] = 6 if len(self.edges) - 2 <= 2 else 5
        h = t["waves_per_simd_hint"]
        t["wpe"] = 6 if h >= 6 else (5 if h == 5 else 1)
        t["waves_per_simd_hint"] = t["wpe"]
        if not self.options.opts.get("warpBalancing", True):
            t["steal_every"] = 1 << 30  # no work donation between wavefronts
        self.tuning = t

    def generate(self):
        tb = self.tables
        name = "worker_cons.hip.tmpl" if self.constrained else "worker.hip.tmpl"
        with open(os.path.join(locations.templateRoot(), name)) as f:
            tmpl = Template(f.read())
        join = lambda xs: ", ".join(str(int(x)) for x in xs)
        if self.constrained:
            subs = dict(motif_name=os.path.basename(self.pathToMotif), M=tb["M"], N=tb["N"],
                        options=", ".join("%s=%s" % kv for kv in sorted(self.options.opts.items())),
                        cons_text=" ".join(str(self.cons.cons).split()), enum_cap=self.enum_NUM,
                        src=join(a for a, _ in self.edges), dst=join(b for _, b in self.edges),
                        kind=join(tb["kind"]), a=join(tb["a"]), b=join(tb["b"]), mapped=join(tb["mapped"]))
            subs.update(constraint_bodies(self.cons, tb["M"], tb["N"], self.anti))
            subs.update(anti=join((self.anti + [False] * MAXE)[:MAXE]), prev=join((self.prev + [0] * MAXE)[:MAXE]))
            return tmpl.substitute(subs)
        subs = dict(motif_name=os.path.basename(self.pathToMotif), M=tb["M"], N=tb["N"],
                    options=", ".join("%s=%s" % kv for kv in sorted(self.options.opts.items())),
                    src=join(a for a, _ in self.edges), dst=join(b for _, b in self.edges),
                    enum_cap=self.enum_NUM, kind=join(tb["kind"]), a=join(tb["a"]), b=join(tb["b"]),
                    mapped=join(tb["mapped"]), smode=join(tb["smode"]), tmode=join(tb["tmode"]), xmask=join(tb["xmask"]),
                    nmap=4 if tb["N"] <= 4 else 8)
        subs.update(self.tuning)
        return tmpl.substitute(subs)
